// The reacting porous example of py_utils/rxnscheme_writer/example in the
// plugin API of mf_plugin_api.h: hydrogen as the one tracked gas species in a
// background inert gas that is not a species (its concentration is what the
// density leaves over, thermo.H:17-41), an adsorbed species and a surface
// site, two mass-action reactions with constant rates (species.cpp:40-52,
// chemistry.H:14-55), inflow of hydrogen in the background gas
// (externbc.H:13-52). mflo.using_bg_inertgas = 1 uses get_bgasconc_from_rc /
// get_r_from_c(spec, bgasconc) around the chemistry phases of the split mode.
// Table fields: in_temp, in_pres, in_vel, in_spec[H2_ID] are the reference's
// problem.free_stream_temp / _pres / _vel / _primaryvapor; user[0] is
// problem.chemistry_on.
#pragma once
#include <math.h>

#include "mf_plugin_api.h"

namespace bg_inert_gas {
using namespace mfp;

struct Case {
    static constexpr int PLUGIN_ID = 103;
    static constexpr int NUM_SPECIES = 3;      // species.H:13-16
    static constexpr int H2_ID = 0, H2ADS1_ID = 1, S1_ID = 2;
    static constexpr int NUM_GAS_SPECIES = 1;  // species.H:18
    static constexpr int FIRST_SPEC = H2_ID;
    static constexpr int NUM_REACTIONS = 2;
    static constexpr bool HAS_CHEMISTRY = true;
    static constexpr bool HAS_FLUID_SOURCE = false;
    static constexpr bool GAMMA_IGNORES_PRESSURE = true;  // gamma from rho and the species
    static constexpr bool HAS_BG_GAS = true;

    MFP_HD static Real BG_GAS_MWT() { return 0.04; }   // species.H:21
    MFP_HD static Real GAMMA_BG_GAS() { return 1.67; } // species.H:22
    // gamma_spec[NUM_GAS_SPECIES + 1]: the gas species, then the background gas (species.cpp:23-26)
    MFP_HD static Real gamma_spec(int sp, const Params &p) { return sp < NUM_GAS_SPECIES ? p.gamma[sp] : GAMMA_BG_GAS(); }
    // rxnarray and rateconsts of species.cpp:36-52
    MFP_HD static Real rxnarray(int r, int sp) {
        const Real a[NUM_REACTIONS][NUM_SPECIES] = {{-1.0, 1.0, -1.0}, {1.0, -1.0, 1.0}};
        return a[r][sp];
    }
    MFP_HD static Real rateconsts(int r) { return r == 0 ? 165675000000.00000 : 16567452390.00000; }

    // species.cpp:13-32
    static void init_species(Params &p) {
        p.nspec = NUM_SPECIES;
        p.ngas = NUM_GAS_SPECIES;
        p.nreact = NUM_REACTIONS;
        p.molwt[H2_ID] = 0.00200;
        p.molwt[H2ADS1_ID] = 0.04400;
        p.molwt[S1_ID] = 0.00000;
        p.gamma[H2_ID] = 1.4;
        for (int sp = 0; sp < NUM_SPECIES; sp++) p.advect_flag[sp] = one;
        p.advect_flag[H2ADS1_ID] = zeroval;
        p.advect_flag[S1_ID] = zeroval;
    }

    // thermo.H:17-195
    MFP_HD static Real get_solid_t_from_rhoe(Real, const Params &) { return 300.0; }
    MFP_HD static Real get_r_from_c(const Real *spec, Real bgasconc, const Params &p) {
        Real rho = zeroval;
        for (int sp = FIRST_SPEC; sp < NUM_GAS_SPECIES; sp++) rho += spec[sp] * p.molwt[sp];
        rho += bgasconc * BG_GAS_MWT();
        return rho;
    }
    MFP_HD static Real get_bgasconc_from_rc(Real rho, const Real *spec, const Params &p) {
        Real sum_ciMi = 0.0;
        for (int sp = FIRST_SPEC; sp < NUM_GAS_SPECIES; sp++) sum_ciMi += spec[sp] * p.molwt[sp];
        Real c_bg_gas = (rho - sum_ciMi) / BG_GAS_MWT();
        return c_bg_gas;
    }
    MFP_HD static Real get_r_from_tpc(Real temp, Real pres, const Real *spec, const Params &p) {
        Real sum_ci = zeroval;
        Real sum_ciMi = zeroval;
        for (int sp = FIRST_SPEC; sp < NUM_GAS_SPECIES; sp++) {
            sum_ci += spec[sp];
            sum_ciMi += spec[sp] * p.molwt[sp];
        }
        Real c_total = pres / RU / temp;
        Real c_bg_gas = c_total - sum_ci;
        return (sum_ciMi + c_bg_gas * BG_GAS_MWT());
    }
    MFP_HD static Real get_t_from_rpc(Real rho, Real pres, const Real *spec, const Params &p) {
        Real sum_ci = zeroval;
        Real sum_ciMi = zeroval;
        for (int sp = FIRST_SPEC; sp < NUM_GAS_SPECIES; sp++) {
            sum_ci += spec[sp];
            sum_ciMi += spec[sp] * p.molwt[sp];
        }
        Real c_bg_gas = (rho - sum_ciMi) / BG_GAS_MWT();
        sum_ci += c_bg_gas;
        return (pres / (sum_ci * RU));
    }
    MFP_HD static Real get_p_from_rec(Real rhoe, Real rho, const Real *spec, const Params &p) {
        Real cv_spec[NUM_GAS_SPECIES + 1] = {zeroval};
        Real sum_ciMi = zeroval;
        for (int sp = FIRST_SPEC; sp < NUM_GAS_SPECIES; sp++) sum_ciMi += spec[sp] * p.molwt[sp];
        Real c_bg_gas = (rho - sum_ciMi) / BG_GAS_MWT();
        for (int sp = FIRST_SPEC; sp < (NUM_GAS_SPECIES + 1); sp++) cv_spec[sp] = RU / (gamma_spec(sp, p) - one);
        Real cv_times_spec = zeroval;
        for (int sp = FIRST_SPEC; sp < NUM_GAS_SPECIES; sp++) cv_times_spec += cv_spec[sp] * spec[sp];
        cv_times_spec += cv_spec[NUM_GAS_SPECIES] * c_bg_gas;
        Real T = rhoe / cv_times_spec;
        Real pr = zeroval;
        for (int sp = FIRST_SPEC; sp < NUM_GAS_SPECIES; sp++) pr += spec[sp] * RU * T;
        pr += c_bg_gas * RU * T;
        return pr;
    }
    MFP_HD static Real get_gama_from_rpc(Real rho, Real, const Real *spec, const Params &p) {
        Real x_i[NUM_GAS_SPECIES + 1] = {zeroval};
        Real sum_ciMi = zeroval;
        for (int sp = H2_ID; sp < NUM_GAS_SPECIES; sp++) sum_ciMi += spec[sp] * p.molwt[sp];
        Real c_bg_gas = (rho - sum_ciMi) / BG_GAS_MWT();
        Real sum_ci = zeroval;
        for (int sp = FIRST_SPEC; sp < NUM_GAS_SPECIES; sp++) sum_ci += spec[sp];
        sum_ci += c_bg_gas;
        for (int sp = FIRST_SPEC; sp < NUM_GAS_SPECIES; sp++) x_i[sp] = spec[sp] / sum_ci;
        x_i[NUM_GAS_SPECIES] = c_bg_gas / sum_ci;
        Real term2 = zeroval;
        for (int sp = FIRST_SPEC; sp < (NUM_GAS_SPECIES + 1); sp++) term2 += x_i[sp] / (gamma_spec(sp, p) - one);
        return one + one / term2;
    }
    MFP_HD static Real get_e_from_rpc(Real rho, Real pres, const Real *spec, const Params &p) {
        Real cv_spec[NUM_GAS_SPECIES + 1] = {zeroval};
        Real sum_ciMi = zeroval;
        for (int sp = FIRST_SPEC; sp < NUM_GAS_SPECIES; sp++) sum_ciMi += spec[sp] * p.molwt[sp];
        Real c_bg_gas = (rho - sum_ciMi) / BG_GAS_MWT();
        for (int sp = FIRST_SPEC; sp < NUM_GAS_SPECIES; sp++) cv_spec[sp] = RU / (gamma_spec(sp, p) - one);
        cv_spec[NUM_GAS_SPECIES] = RU / (gamma_spec(NUM_GAS_SPECIES, p) - one);
        Real temp = get_t_from_rpc(rho, pres, spec, p);
        Real sum_cvi_ci = zeroval;
        for (int sp = FIRST_SPEC; sp < NUM_GAS_SPECIES; sp++) sum_cvi_ci += cv_spec[sp] * spec[sp];
        sum_cvi_ci += cv_spec[NUM_GAS_SPECIES] * c_bg_gas;
        return (sum_cvi_ci * temp / rho);
    }

    // userfuncs.H:37-61
    MFP_HD static void compute_fluid_transport(int i, int j, int k, Array4<Real> const &, Array4<Real> const &tc,
                                               GpuArray<Real, 3>, GpuArray<Real, 3>, const Real, const Params &) {
        tc(i, j, k, VISC_INDX) = 2.15e-05;
        tc(i, j, k, THCOND_INDX) = 8.25e-2;
    }
    MFP_HD static void compute_fluid_source(int i, int j, int k, Array4<Real> const &, Array4<Real> const &source,
                                            GpuArray<Real, 3>, GpuArray<Real, 3>, const Real, const Params &) {
        for (int nc = 0; nc < FLO_NVARS; nc++) source(i, j, k, nc) = zeroval;
    }
    // transport.H:12-27
    MFP_HD static void compute_spec_dcoeff(int i, int j, int k, Array4<Real> const &phi, Array4<Real> const &dcoeff,
                                           GpuArray<Real, 3>, GpuArray<Real, 3>, const Real, const Params &) {
        Real solid_dif = 1.e-10;
        Real fluid_dif = 4.27e-5;
        Real vfrac = phi(i, j, k, VFRAC_INDX);
        for (int sp = 0; sp < NUM_SPECIES; sp++) dcoeff(i, j, k, sp) = (one - vfrac) * solid_dif + vfrac * fluid_dif;
        dcoeff(i, j, k, H2ADS1_ID) = zeroval;
        dcoeff(i, j, k, S1_ID) = zeroval;
    }
    // chemistry.H:14-55 (chemistry_on = user[0])
    MFP_HD static void compute_spec_source(int i, int j, int k, Array4<Real> const &phi, Array4<Real> const &specsource,
                                           GpuArray<Real, 3>, GpuArray<Real, 3>, const Real, const Params &p) {
        Real wdot[NUM_SPECIES] = {zeroval};
        if (p.user[0] != zeroval) {
            for (int r = 0; r < NUM_REACTIONS; r++) {
                Real concmult = 1.0;
                for (int sp = 0; sp < NUM_SPECIES; sp++) {
                    if (rxnarray(r, sp) < 0.0) {
                        Real stoichcoeff = fabs(rxnarray(r, sp));
                        concmult *= pow(phi(i, j, k, FLO_NVARS + sp), stoichcoeff);
                    }
                }
                Real reactrate = rateconsts(r) * concmult;
                for (int sp = 0; sp < NUM_SPECIES; sp++)
                    if (fabs(rxnarray(r, sp)) > 0.0) wdot[sp] += rxnarray(r, sp) * reactrate;
            }
        }
        for (int sp = 0; sp < NUM_SPECIES; sp++) specsource(i, j, k, FLO_NVARS + sp) += wdot[sp];
    }
    // externbc.H:13-52: every ext_dir face takes the free stream
    MFP_HD static void externalbc(const Real *, const Real s_int[], Real s_ext[], const int, const int, const Real,
                                  GeometryData const &, const Params &p) {
        Real fs_spec[NUM_SPECIES] = {zeroval};
        fs_spec[H2_ID] = p.in_spec[H2_ID];
        s_ext[PRES_INDX] = s_int[PRES_INDX];
        s_ext[TEMP_INDX] = p.in_temp;
        s_ext[DENS_INDX] = get_r_from_tpc(p.in_temp, s_ext[PRES_INDX], fs_spec, p);
        s_ext[VELX_INDX] = p.in_vel;
        s_ext[VELY_INDX] = zeroval;
        s_ext[VELZ_INDX] = zeroval;
        s_ext[VFRAC_INDX] = one;
        s_ext[RHO_INDX] = s_ext[DENS_INDX];
        s_ext[RHOU_INDX] = s_ext[DENS_INDX] * s_ext[VELX_INDX];
        s_ext[RHOV_INDX] = zeroval;
        s_ext[RHOW_INDX] = zeroval;
        s_ext[RHOE_INDX] = s_ext[DENS_INDX] * get_e_from_rpc(s_ext[DENS_INDX], s_ext[PRES_INDX], fs_spec, p);
        s_ext[RHOE_INDX] += half * s_ext[DENS_INDX] * s_ext[VELX_INDX] * s_ext[VELX_INDX];
        for (int sp = 0; sp < NUM_SPECIES; sp++) s_ext[FLO_NVARS + sp] = fs_spec[sp];
    }
};

}  // namespace bg_inert_gas
