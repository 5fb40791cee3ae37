// The SpeciesEqTests case (tests/Cartesian/SpeciesEqTests) in the plugin API
// of mf_plugin_api.h: two species in a frozen unit flow (mflo.solve_navier_stokes
// = 0), S1 advected and diffused (D = 0.1) between Dirichlet values 0 and 1,
// S2 not advected, diffused (D = 1) with a unit source and zero at both ends.
// Steady states (verify_spec.py:26-28, 42): S1 = (e^{x/0.1} - 1) / (e^{10} - 1),
// S2 = (x - x^2) / 2.
#pragma once
#include "mf_plugin_api.h"

namespace species_eq_tests {
using namespace mfp;

struct Case {
    static constexpr int PLUGIN_ID = 101;
    static constexpr int NUM_SPECIES = 2;      // species.H:11
    static constexpr int NUM_GAS_SPECIES = 2;  // the thermo ignores the species
    static constexpr int S1_ID = 0, S2_ID = 1;
    static constexpr bool HAS_CHEMISTRY = true;
    static constexpr bool HAS_FLUID_SOURCE = false;
    static constexpr bool GAMMA_IGNORES_PRESSURE = true;
    static constexpr bool HAS_BG_GAS = false;

    // species.cpp:10-24
    static void init_species(Params &p) {
        p.nspec = NUM_SPECIES;
        p.ngas = NUM_GAS_SPECIES;
        p.molwt[S1_ID] = 0.032;  // O2
        p.molwt[S2_ID] = 0.028;  // N2
        p.advect_flag[S1_ID] = one;
        p.advect_flag[S2_ID] = zeroval;
        p.gamma[S1_ID] = p.gamma[S2_ID] = 1.4;
    }

    // thermo.H: a calorically perfect gas with R = 1, gamma = 1.4
    MFP_HD static Real get_solid_t_from_rhoe(Real, const Params &) { return 300.0; }
    MFP_HD static Real get_r_from_c(const Real *spec, Real, const Params &p) {
        Real rho = zeroval;
        for (int sp = 0; sp < NUM_SPECIES; sp++) rho += spec[sp] * p.molwt[sp];
        return rho;
    }
    MFP_HD static Real get_bgasconc_from_rc(Real, const Real *spec, const Params &) { return spec[S1_ID]; }
    MFP_HD static Real get_t_from_rpc(Real rho, Real pres, const Real *, const Params &) {
        const Real Rgas = one;
        return pres / rho / Rgas;
    }
    MFP_HD static Real get_p_from_rec(Real rhoe, Real, const Real *, const Params &) {
        const Real gamma = 1.4;
        return rhoe * (gamma - one);
    }
    MFP_HD static Real get_gama_from_rpc(Real, Real, const Real *, const Params &) { return 1.4; }
    MFP_HD static Real get_e_from_rpc(Real rho, Real pres, const Real *, const Params &) {
        const Real gamma = 1.4;
        return pres / (gamma - one) / rho;
    }

    // userfuncs.H:20-46
    MFP_HD static void compute_fluid_transport(int i, int j, int k, Array4<Real> const &, Array4<Real> const &tc,
                                               GpuArray<Real, 3>, GpuArray<Real, 3>, const Real, const Params &) {
        tc(i, j, k, VISC_INDX) = 0.01;
        tc(i, j, k, THCOND_INDX) = 0.0;
    }
    MFP_HD static void compute_fluid_source(int i, int j, int k, Array4<Real> const &, Array4<Real> const &source,
                                            GpuArray<Real, 3>, GpuArray<Real, 3>, const Real, const Params &) {
        for (int nc = 0; nc < FLO_NVARS; nc++) source(i, j, k, nc) = 0.0;
    }
    // transport.H:12-21
    MFP_HD static void compute_spec_dcoeff(int i, int j, int k, Array4<Real> const &, Array4<Real> const &dcoeff,
                                           GpuArray<Real, 3>, GpuArray<Real, 3>, const Real, const Params &) {
        dcoeff(i, j, k, S1_ID) = 0.1;
        dcoeff(i, j, k, S2_ID) = 1.0;
    }
    // chemistry.H:13-24
    MFP_HD static void compute_spec_source(int i, int j, int k, Array4<Real> const &, Array4<Real> const &specsource,
                                           GpuArray<Real, 3>, GpuArray<Real, 3>, const Real, const Params &) {
        specsource(i, j, k, FLO_NVARS + S1_ID) = zeroval;
        specsource(i, j, k, FLO_NVARS + S2_ID) = one;
    }
    // externbc.H:13-62: flow variables copied, species reflected about the
    // boundary value (0 at the low end; S1 = 1, S2 = 0 at the high end)
    MFP_HD static void externalbc(const Real *, const Real s_int[], Real s_ext[], const int, const int sgn, const Real,
                                  GeometryData const &, const Params &) {
        s_ext[PRES_INDX] = s_int[PRES_INDX];
        s_ext[DENS_INDX] = s_int[DENS_INDX];
        s_ext[TEMP_INDX] = s_int[TEMP_INDX];
        s_ext[VELX_INDX] = s_int[VELX_INDX];
        s_ext[VELY_INDX] = s_int[VELY_INDX];
        s_ext[VELZ_INDX] = s_int[VELZ_INDX];
        s_ext[RHO_INDX] = s_int[RHO_INDX];
        s_ext[RHOU_INDX] = s_int[RHOU_INDX];
        s_ext[RHOV_INDX] = s_int[RHOV_INDX];
        s_ext[RHOW_INDX] = s_int[RHOW_INDX];
        s_ext[RHOE_INDX] = s_int[RHOE_INDX];
        s_ext[VFRAC_INDX] = one;
        Real val;
        if (sgn > 0) {
            val = zeroval;
            s_ext[FLO_NVARS + S1_ID] = two * val - s_int[FLO_NVARS + S1_ID];
            val = zeroval;
            s_ext[FLO_NVARS + S2_ID] = two * val - s_int[FLO_NVARS + S2_ID];
        } else {
            val = one;
            s_ext[FLO_NVARS + S1_ID] = two * val - s_int[FLO_NVARS + S1_ID];
            val = zeroval;
            s_ext[FLO_NVARS + S2_ID] = two * val - s_int[FLO_NVARS + S2_ID];
        }
    }
};

}  // namespace species_eq_tests
