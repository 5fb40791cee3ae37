// The Channel case (tests/Cartesian/Channel) in the plugin API of
// mf_plugin_api.h: air (one species) between no-slip walls, periodic along
// the channel, viscosity 0.01, an initial uniform axial velocity 0.1 that
// develops the wall profile (verify_vel.py:26 compares it with
// 4 y (w - y) / w^2).
#pragma once
#include "mf_plugin_api.h"

namespace channel {
using namespace mfp;

struct Case {
    static constexpr int PLUGIN_ID = 102;
    static constexpr int NUM_SPECIES = 1;
    static constexpr int NUM_GAS_SPECIES = 1;
    static constexpr int AIR_ID = 0;
    static constexpr bool HAS_CHEMISTRY = false;
    static constexpr bool HAS_FLUID_SOURCE = false;
    static constexpr bool GAMMA_IGNORES_PRESSURE = true;
    static constexpr bool HAS_BG_GAS = false;

    // species.cpp:9-17
    static void init_species(Params &p) {
        p.nspec = NUM_SPECIES;
        p.ngas = NUM_GAS_SPECIES;
        p.molwt[AIR_ID] = 0.0289;
        p.advect_flag[AIR_ID] = one;
        p.gamma[AIR_ID] = 1.4;
    }

    // thermo.H: air, R_air = RU / M_air, gamma 1.4
    MFP_HD static Real get_solid_t_from_rhoe(Real, const Params &) { return 300.0; }
    MFP_HD static Real get_r_from_c(const Real *spec, Real, const Params &p) {
        Real rho = zeroval;
        for (int sp = 0; sp < NUM_SPECIES; sp++) rho += spec[sp] * p.molwt[sp];
        return rho;
    }
    MFP_HD static Real get_bgasconc_from_rc(Real, const Real *spec, const Params &) { return spec[AIR_ID]; }
    MFP_HD static Real get_t_from_rpc(Real rho, Real pres, const Real *, const Params &p) {
        const Real R_air = RU / p.molwt[AIR_ID];
        return pres / R_air / rho;
    }
    MFP_HD static Real get_p_from_rec(Real rhoe, Real, const Real *, const Params &) {
        const Real gamma_air = 1.4;
        return rhoe * (gamma_air - one);
    }
    MFP_HD static Real get_gama_from_rpc(Real, Real, const Real *, const Params &) { return 1.4; }
    MFP_HD static Real get_e_from_rpc(Real rho, Real pres, const Real *, const Params &) {
        const Real gamma_air = 1.4;
        return pres / (gamma_air - one) / rho;
    }

    // userfuncs.H: viscosity tenth*tenth, no conduction, no source
    MFP_HD static void compute_fluid_transport(int i, int j, int k, Array4<Real> const &, Array4<Real> const &tc,
                                               GpuArray<Real, 3>, GpuArray<Real, 3>, const Real, const Params &) {
        tc(i, j, k, VISC_INDX) = tenth * tenth;
        tc(i, j, k, THCOND_INDX) = zeroval;
    }
    MFP_HD static void compute_fluid_source(int i, int j, int k, Array4<Real> const &, Array4<Real> const &source,
                                            GpuArray<Real, 3>, GpuArray<Real, 3>, const Real, const Params &) {
        for (int nc = 0; nc < FLO_NVARS; nc++) source(i, j, k, nc) = zeroval;
    }
    // transport.H, chemistry.H: nothing diffuses or reacts
    MFP_HD static void compute_spec_dcoeff(int i, int j, int k, Array4<Real> const &, Array4<Real> const &dcoeff,
                                           GpuArray<Real, 3>, GpuArray<Real, 3>, const Real, const Params &) {
        dcoeff(i, j, k, AIR_ID) = zeroval;
    }
    MFP_HD static void compute_spec_source(int i, int j, int k, Array4<Real> const &, Array4<Real> const &specsource,
                                           GpuArray<Real, 3>, GpuArray<Real, 3>, const Real, const Params &) {
        specsource(i, j, k, FLO_NVARS + AIR_ID) = zeroval;
    }
    // externbc.H: the case has no ext_dir boundary; the function is empty
    MFP_HD static void externalbc(const Real *, const Real *, Real *, const int, const int, const Real,
                                  GeometryData const &, const Params &) {}
};

}  // namespace channel
