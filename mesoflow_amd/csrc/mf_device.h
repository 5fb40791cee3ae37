// mf_device.h -- device numerics of the RHS path for gfx950 (FP64 throughout).
//
// Each function states the reference lines whose arithmetic it reproduces.
// Operation order follows the reference expression by expression (the build
// turns FP contraction off), so that results agree with the reference's CPU
// build to rounding. pow(x,2) is written x*x and pow(x,-1) as 1/x.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "mesoflow_amd.h"
#include "mf_plugin_api.h"

#define MFD __host__ __device__ __forceinline__

namespace mfd {

// Source/mflo_constants.H:4-28 (third/sixth are the products the macros expand to)
constexpr double kZero = 0.0, kOne = 1.0, kTwo = 2.0, kThree = 3.0, kFour = 4.0, kFive = 5.0;
constexpr double kTwo3rd = 0.6666666666666666667, kHalf = 0.5;
constexpr double kThird = kTwo3rd * kHalf;
constexpr double kEighth = 0.125, kSixteenth = 0.0625, kTwelveth = 0.08333333333;
constexpr double kFourth = 0.25, kThreeTenth = 0.3, kThreeFourth = 0.75;
constexpr double kTiny = 1e-20, kTiny2 = 1e-40;

// Source/globalDefines.H:5-32
enum : int { RHO = 0, RHOU = 1, RHOV = 2, RHOW = 3, RHOE = 4, DENS = 5, VELX = 6, VELY = 7,
             VELZ = 8, PRES = 9, TEMP = 10, VFRAC = 11, FLO = 12, NCV = 5 };

// ---------------------------------------------------------------------------
// views
// ---------------------------------------------------------------------------
struct View {  // amrex::Array4 on the device
    double *p;
    int64_t js, ks, ns;
    int b0, b1, b2;
    MFD double &operator()(int i, int j, int k, int n) const {
        return p[(int64_t)(i - b0) + (int64_t)(j - b1) * js + (int64_t)(k - b2) * ks + (int64_t)n * ns];
    }
    MFD int64_t off(int i, int j, int k) const {
        return (int64_t)(i - b0) + (int64_t)(j - b1) * js + (int64_t)(k - b2) * ks;
    }
};

// State accessor of one box's Sborder: the stage source array plus, for cells of
// the box's own valid region that are cut cells, the values update_cutcells
// produced (kept apart so that other boxes and the RK combine still see the
// raw state; Source/mflo_evolve.cpp:509-517 updates only the valid tilebox).
struct Acc {
    View s;
    const int *cutid;      // level valid cells (nx*ny*nz), -1 if not a cut cell; null: no substitution
    const double *cutbuf;  // [nv][ncut]
    int ncut;
    int nx, ny;
    int blo[3], bhi[3];    // owning box
    MFD int cid(int i, int j, int k) const {
        if (!cutid) return -1;
        if (i < blo[0] || i > bhi[0] || j < blo[1] || j > bhi[1] || k < blo[2] || k > bhi[2]) return -1;
        return cutid[(int64_t)i + (int64_t)nx * ((int64_t)j + (int64_t)ny * k)];
    }
    MFD double get(int i, int j, int k, int n, int c) const {
        return c >= 0 ? cutbuf[(int64_t)n * ncut + c] : s(i, j, k, n);
    }
    MFD double operator()(int i, int j, int k, int n) const { return get(i, j, k, n, cid(i, j, k)); }
};

// All components of one cell of a box's Sborder: p[n * stride]
struct CellPtr {
    const double *p;
    int64_t stride;
    MFD double operator()(int n) const { return p[(int64_t)n * stride]; }
};
MFD CellPtr cell_ptr(const Acc &a, int i, int j, int k, int c) {
    CellPtr r;
    if (c >= 0) { r.p = a.cutbuf + c; r.stride = a.ncut; }
    else { r.p = a.s.p + a.s.off(i, j, k); r.stride = a.s.ns; }
    return r;
}

// One cell seen by a pointwise plugin (the reference's plugins read phi(i,j,k,*) only)
struct CellRef {
    const Acc *a;
    int i, j, k, c;
    MFD double operator()(int n) const { return a->get(i, j, k, n, c); }
};
// A cell held in registers
template <int NV>
struct CellArr {
    const double *v;
    MFD double operator()(int n) const { return v[n]; }
};

// ---------------------------------------------------------------------------
// Plugins: compile-time families with runtime tables (mf_plugin_params)
// ---------------------------------------------------------------------------
// std::max / std::min of the reference (a NaN first argument propagates, a
// NaN second argument is dropped -- fmax/fmin drop both)
MFD double std_max(double a, double b) { return (a < b) ? b : a; }
MFD double std_min(double a, double b) { return (b < a) ? b : a; }

// n / d for a positive d far inside the normal range (WENO weight sums,
// densities, heat capacities, sums of concentrations or of transport
// coefficients plus 1e-30): v_rcp_f64 estimate refined by
// two Newton-Raphson steps r += r (1 - d r) and one residual correction
// q += r (n - d q) -- the steps of the full division without the exponent
// scaling that only extreme operands need
MFD double div_pos(double n, double d) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    const double q = n * r;
    return fma(fma(-d, q, n), r, q);
#else
    return n / d;
#endif
}

// Plugin table on the device: the C-ABI table plus per-species constants the
// plugins would otherwise recompute per call (the same IEEE divisions, done once
// on the host: cv_s = R/(gamma_s-1), gm1_s = gamma_s-1, cvm_s = R/(gamma_s-1)/molwt_s),
// and igm1_s = 1/(gamma_s-1) for the mixture gamma.
struct Tab : mf_plugin_params {
    double cv[MF_MAX_SPECIES];
    double gm1[MF_MAX_SPECIES];
    double cvm[MF_MAX_SPECIES];
    double igm1[MF_MAX_SPECIES];
    mfp::GeometryData geo;  // the level's bounds and cell sizes (for reference-style plugins)
};

// tests/Cartesian/ShockTube_N2/{thermo.H:19-133, transport.H, userfuncs.H, chemistry.H, externbc.H}
struct PShockTubeN2 {
    static constexpr int NS = 1;
    static constexpr int NG = 1;
    // thermo.H:39-48
    MFD static double get_r_from_c(const Tab &T, const double *spec) { return kZero + spec[0] * T.molwt[0]; }
    MFD static double get_t_from_rpc(const Tab &T, double rho, double pres, const double *spec) {
        double sum_ci = kZero + spec[0];
        return pres / (sum_ci * T.R_univ);
    }
    MFD static double get_p_from_rec(const Tab &T, double rhoe, double rho, const double *spec) {
        double cv = T.cv[0];
        double Tm = rhoe / (cv * spec[0]);
        return kZero + spec[0] * T.R_univ * Tm;
    }
    MFD static double get_gama_from_rpc(const Tab &T, double rho, double pres, const double *spec) {
        double sum_ci = kZero + spec[0];
        double x0 = spec[0] / sum_ci;
        return kOne + kOne / (x0 / T.gm1[0]);
    }
    MFD static double get_e_from_rpc(const Tab &T, double rho, double pres, const double *spec) {
        double cv = T.cvm[0];
        double sum_ci = kZero + spec[0];
        double x0 = spec[0] / sum_ci;
        double temp = get_t_from_rpc(T, rho, pres, spec);
        return temp * (cv * x0);
    }
    MFD static double get_solid_t_from_rhoe(const Tab &, double) { return 300.0; }
    template <class C> MFD static void fluid_transport(const Tab &, const C &, double &visc, double &thcond) {
        visc = kZero; thcond = kZero;
    }
    template <class C> MFD static void spec_dcoeff(const Tab &, const C &, double *d) { d[0] = kZero; }
    template <class C> MFD static void fluid_source(const Tab &, const C &, double *src) {
        for (int n = 0; n < FLO; ++n) src[n] = kZero;
    }
    static constexpr bool kHasFluidSource = false;
    static constexpr bool kHasChemistry = false;
    static constexpr bool kGammaIgnoresP = true;  // gamma does not depend on the pressure
    static constexpr bool kHasBgGas = false;
    template <class C> MFD static void spec_source(const Tab &, const C &, double *src) { src[0] = kZero; }
    MFD static void rate_constants(const Tab &, double, double *) {}
    template <class C> MFD static void spec_source_k(const Tab &T, const double *, const C &cell, double *src) {
        spec_source(T, cell, src);
    }
    MFD static void externalbc(const Tab &T, const double *x, const double *s_int, double *s_ext, int idir, int sgn,
                               double time) {
        s_ext[PRES] = s_int[PRES];
        s_ext[DENS] = T.left_rho;
        s_ext[TEMP] = T.left_T;
        s_ext[VELX] = kZero; s_ext[VELY] = kZero; s_ext[VELZ] = kZero;
        s_ext[VFRAC] = kOne;
        s_ext[RHO] = s_ext[DENS];
        s_ext[RHOU] = kZero; s_ext[RHOV] = kZero; s_ext[RHOW] = kZero;
        s_ext[RHOE] = s_ext[DENS] * get_e_from_rpc(T, T.left_rho, T.left_p, T.left_spec);
        s_ext[FLO] = T.left_spec[0];
    }
};

// Mixture ideal gas over gas species 0..NG-1: models/catalystCylinder/thermo.H:36-156
template <int NSv, int NGv>
struct PMixture {
    static constexpr int NS = NSv;
    static constexpr int NG = NGv;
    MFD static double get_r_from_c(const Tab &T, const double *spec) {
        double rho = kZero;
#pragma unroll
        for (int s = 0; s < NG; ++s) rho += spec[s] * T.molwt[s];
        return rho;
    }
    MFD static double get_t_from_rpc(const Tab &T, double rho, double pres, const double *spec) {
        double sum_ci = kZero;
#pragma unroll
        for (int s = 0; s < NG; ++s) sum_ci += spec[s];
        return pres / (sum_ci * T.R_univ);
    }
    MFD static double get_p_from_rec(const Tab &T, double rhoe, double rho, const double *spec) {
        double cvs = kZero;
#pragma unroll
        for (int s = 0; s < NG; ++s) cvs += T.cv[s] * spec[s];
        double Tm = div_pos(rhoe, cvs);
        double p = kZero;
#pragma unroll
        for (int s = 0; s < NG; ++s) p += spec[s] * T.R_univ * Tm;
        return p;
    }
    MFD static double get_gama_from_rpc(const Tab &T, double rho, double pres, const double *spec) {
        double sum_ci = kZero;
#pragma unroll
        for (int s = 0; s < NG; ++s) sum_ci += spec[s];
        // x_s / (gamma_s - 1) with one reciprocal of the total and of each gamma_s - 1
        const double isum = div_pos(kOne, sum_ci);
        double term2 = kZero;
#pragma unroll
        for (int s = 0; s < NG; ++s) term2 += (spec[s] * isum) * T.igm1[s];
        return kOne + div_pos(kOne, term2);
    }
    MFD static double get_e_from_rpc(const Tab &T, double rho, double pres, const double *spec) {
        double temp = get_t_from_rpc(T, rho, pres, spec);
        double cvc = kZero;
#pragma unroll
        for (int s = 0; s < NG; ++s) cvc += T.cv[s] * spec[s];
        return cvc * temp / rho;
    }
    MFD static double get_solid_t_from_rhoe(const Tab &, double) { return 300.0; }
    // models/catalystCylinder/userfuncs.H:27-37
    template <class C> MFD static void fluid_transport(const Tab &T, const C &, double &visc, double &thcond) {
        visc = T.visc; thcond = T.thcond;
    }
    // models/catalystCylinder/transport.H:12-27
    template <class C> MFD static void spec_dcoeff(const Tab &T, const C &cell, double *d) {
        double gasfrac = cell(VFRAC);
        double solfrac = kOne - gasfrac;
#pragma unroll
        for (int s = 0; s < NS; ++s) d[s] = T.dcoef_gas[s] * gasfrac + T.dcoef_solid[s] * solfrac;
    }
    template <class C> MFD static void fluid_source(const Tab &, const C &, double *src) {
        for (int n = 0; n < FLO; ++n) src[n] = kZero;
    }
    static constexpr bool kHasFluidSource = false;
    static constexpr bool kGammaIgnoresP = true;  // gamma does not depend on the pressure
    static constexpr bool kHasBgGas = false;
};

// models/catalystCylinder: chemistry.H:13-29, externbc.H:13-73
struct PCatalystCylinder : PMixture<4, 3> {
    static constexpr bool kHasChemistry = true;
    template <class C> MFD static void spec_source(const Tab &T, const C &cell, double *src) {
        for (int s = 0; s < NS; ++s) src[s] = kZero;
        double reactrate = T.kconst * cell(FLO + 3) * cell(FLO + 1);
        src[1] -= reactrate;
        src[2] += reactrate;
    }
    // a constant rate: nothing depends on temperature
    MFD static void rate_constants(const Tab &, double, double *) {}
    template <class C> MFD static void spec_source_k(const Tab &T, const double *, const C &cell, double *src) {
        spec_source(T, cell, src);
    }
    MFD static void externalbc(const Tab &T, const double *x, const double *s_int, double *s_ext, int idir, int sgn,
                               double time) {
        double spec[NS] = {kZero, kZero, kZero, kZero};
        if (idir == 0 && sgn > 0) {
            s_ext[DENS] = T.in_rho;
            s_ext[PRES] = s_int[PRES];
            s_ext[VELX] = T.in_vel;
            s_ext[VELY] = kZero;
            s_ext[VELZ] = kZero;
            spec[0] = (s_ext[DENS] - T.in_spec[1] * T.molwt[1]) / T.molwt[0];
            spec[1] = T.in_spec[1];
            s_ext[TEMP] = get_t_from_rpc(T, s_ext[DENS], s_ext[PRES], spec);
        } else {
            s_ext[DENS] = s_int[DENS];
            s_ext[PRES] = T.out_pres;
            s_ext[VELX] = s_int[VELX];
            s_ext[VELY] = s_int[VELY];
            s_ext[VELZ] = s_int[VELZ];
            s_ext[TEMP] = s_int[TEMP];
            for (int s = 0; s < NS; ++s) spec[s] = s_int[FLO + s];
        }
        s_ext[RHO] = s_ext[DENS];
        s_ext[RHOU] = s_ext[DENS] * s_ext[VELX];
        s_ext[RHOV] = s_ext[DENS] * s_ext[VELY];
        s_ext[RHOW] = s_ext[DENS] * s_ext[VELZ];
        s_ext[RHOE] = s_ext[DENS] * get_e_from_rpc(T, s_ext[DENS], s_ext[PRES], spec);
        s_ext[RHOE] += kHalf * s_ext[DENS] * (s_ext[VELX] * s_ext[VELX] + s_ext[VELY] * s_ext[VELY] + s_ext[VELZ] * s_ext[VELZ]);
        for (int s = 0; s < NS; ++s) s_ext[FLO + s] = spec[s];
        s_ext[VFRAC] = s_int[VFRAC];
    }
};

// Arrhenius mass-action plugin in the reference's plugin API (the loop of
// py_utils/rxnscheme_writer/example/chemistry.H:14-55 with
// k_r = exp(log_A_r - Ea_r/(R T))); inflow x-lo after
// py_utils/rxnscheme_writer/example/externbc.H:16-52, other ext_dir faces as the
// catalystCylinder outflow.
template <int NSv>
struct PPorous : PMixture<NSv, NSv - 1> {
    using Base = PMixture<NSv, NSv - 1>;
    static constexpr int NS = NSv;
    static constexpr int NG = NSv - 1;
    static constexpr bool kHasChemistry = true;
    // c^a for a stoichiometric coefficient other than 1 and 2, kept out of line:
    // inlined at every (reaction, species) site of an unrolled loop it makes the
    // code larger than the instruction cache
    __host__ __device__ __attribute__((noinline)) static double stoich_pow(double c, double a) { return pow(c, a); }
    // mass-action rate of reaction r with rate constant krate, added to wdot
    template <class C> MFD static void add_reaction(const Tab &T, int r, double krate, const C &cell, double *wdot) {
        double concmult = 1.0;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double nu = T.stoich[r][s];
            if (nu < 0.0) {
                double c = cell(FLO + s);
                double a = fabs(nu);
                concmult *= (a == 1.0) ? c : ((a == 2.0) ? c * c : stoich_pow(c, a));
            }
        }
        double reactrate = krate * concmult;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double nu = T.stoich[r][s];
            if (fabs(nu) > 0.0) wdot[s] += nu * reactrate;
        }
    }
    template <class C> MFD static void spec_source(const Tab &T, const C &cell, double *src) {
        double wdot[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) wdot[s] = kZero;
        double Tm = cell(TEMP);
        for (int r = 0; r < T.nreact; ++r) add_reaction(T, r, exp(T.log_A[r] - T.Ea[r] / (T.R_univ * Tm)), cell, wdot);
#pragma unroll
        for (int s = 0; s < NS; ++s) src[s] = kZero + wdot[s];
    }
    // the temperature-only part of the source, k_r = A_r exp(-Ea_r / (R T)),
    // and the source from given rate constants (the same values as spec_source)
    MFD static void rate_constants(const Tab &T, double Tm, double *k) {
#pragma unroll
        for (int r = 0; r < MF_MAX_REACTIONS; ++r) {
            k[r] = kZero;
            if (r < T.nreact) k[r] = exp(T.log_A[r] - T.Ea[r] / (T.R_univ * Tm));
        }
    }
    template <class C> MFD static void spec_source_k(const Tab &T, const double *k, const C &cell, double *src) {
        double wdot[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) wdot[s] = kZero;
#pragma unroll
        for (int r = 0; r < MF_MAX_REACTIONS; ++r) {
            if (r >= T.nreact) break;
            add_reaction(T, r, k[r], cell, wdot);
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) src[s] = kZero + wdot[s];
    }
    MFD static void externalbc(const Tab &T, const double *x, const double *s_int, double *s_ext, int idir, int sgn,
                               double time) {
        double spec[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) spec[s] = kZero;
        if (idir == 0 && sgn > 0) {
            s_ext[PRES] = s_int[PRES];
            s_ext[TEMP] = T.in_temp;
            double ctot = s_ext[PRES] / (T.R_univ * s_ext[TEMP]);
#pragma unroll
            for (int s = 0; s < NG; ++s) spec[s] = T.in_spec[s] * ctot;
            s_ext[DENS] = Base::get_r_from_c(T, spec);
            s_ext[VELX] = T.in_vel;
            s_ext[VELY] = kZero;
            s_ext[VELZ] = kZero;
            s_ext[VFRAC] = kOne;
        } else {
            s_ext[DENS] = s_int[DENS];
            s_ext[PRES] = T.out_pres;
            s_ext[VELX] = s_int[VELX];
            s_ext[VELY] = s_int[VELY];
            s_ext[VELZ] = s_int[VELZ];
            s_ext[TEMP] = s_int[TEMP];
            s_ext[VFRAC] = s_int[VFRAC];
#pragma unroll
            for (int s = 0; s < NS; ++s) spec[s] = s_int[FLO + s];
        }
        s_ext[RHO] = s_ext[DENS];
        s_ext[RHOU] = s_ext[DENS] * s_ext[VELX];
        s_ext[RHOV] = s_ext[DENS] * s_ext[VELY];
        s_ext[RHOW] = s_ext[DENS] * s_ext[VELZ];
        s_ext[RHOE] = s_ext[DENS] * Base::get_e_from_rpc(T, s_ext[DENS], s_ext[PRES], spec);
        s_ext[RHOE] += kHalf * s_ext[DENS] * (s_ext[VELX] * s_ext[VELX] + s_ext[VELY] * s_ext[VELY] + s_ext[VELZ] * s_ext[VELZ]);
#pragma unroll
        for (int s = 0; s < NS; ++s) s_ext[FLO + s] = spec[s];
    }
};

// ---------------------------------------------------------------------------
// A case written against mf_plugin_api.h (the reference's signature style)
// seen through the interface the kernels use: each point-wise function gets a
// view of its own cell, the level's prob_lo and dx, and the context's table.
// ---------------------------------------------------------------------------
template <class U>
struct PRef {
    static constexpr int NS = U::NUM_SPECIES;
    static constexpr int NG = U::NUM_GAS_SPECIES;
    static constexpr int NV = FLO + NS;
    static constexpr bool kHasChemistry = U::HAS_CHEMISTRY;
    static constexpr bool kHasFluidSource = U::HAS_FLUID_SOURCE;
    static constexpr bool kGammaIgnoresP = U::GAMMA_IGNORES_PRESSURE;
    static constexpr bool kHasBgGas = U::HAS_BG_GAS;
    MFD static mfp::GpuArray<double, 3> lo3(const Tab &T) { return {{T.geo.prob_lo[0], T.geo.prob_lo[1], T.geo.prob_lo[2]}}; }
    MFD static mfp::GpuArray<double, 3> dx3(const Tab &T) { return {{T.geo.dx[0], T.geo.dx[1], T.geo.dx[2]}}; }
    MFD static double get_r_from_c(const Tab &T, const double *spec) { return U::get_r_from_c(spec, kZero, T); }
    MFD static double get_r_from_c_bg(const Tab &T, const double *spec, double bg) { return U::get_r_from_c(spec, bg, T); }
    MFD static double get_bgasconc_from_rc(const Tab &T, double rho, const double *spec) {
        return U::get_bgasconc_from_rc(rho, spec, T);
    }
    MFD static double get_t_from_rpc(const Tab &T, double rho, double pres, const double *spec) {
        return U::get_t_from_rpc(rho, pres, spec, T);
    }
    MFD static double get_p_from_rec(const Tab &T, double rhoe, double rho, const double *spec) {
        return U::get_p_from_rec(rhoe, rho, spec, T);
    }
    MFD static double get_gama_from_rpc(const Tab &T, double rho, double pres, const double *spec) {
        return U::get_gama_from_rpc(rho, pres, spec, T);
    }
    MFD static double get_e_from_rpc(const Tab &T, double rho, double pres, const double *spec) {
        return U::get_e_from_rpc(rho, pres, spec, T);
    }
    MFD static double get_solid_t_from_rhoe(const Tab &T, double rhoe) { return U::get_solid_t_from_rhoe(rhoe, T); }
    template <class C> MFD static void fluid_transport(const Tab &T, const C &cell, double &visc, double &thcond) {
        double ph[NV], out[2] = {kZero, kZero};
#pragma unroll
        for (int n = 0; n < NV; ++n) ph[n] = cell(n);
        U::compute_fluid_transport(0, 0, 0, mfp::cell_view(ph, 0, 0, 0), mfp::cell_view(out, 0, 0, 0), lo3(T), dx3(T),
                                   kZero, T);
        visc = out[0];
        thcond = out[1];
    }
    template <class C> MFD static void spec_dcoeff(const Tab &T, const C &cell, double *d) {
        double ph[NV];
#pragma unroll
        for (int n = 0; n < NV; ++n) ph[n] = cell(n);
#pragma unroll
        for (int s = 0; s < NS; ++s) d[s] = kZero;
        U::compute_spec_dcoeff(0, 0, 0, mfp::cell_view(ph, 0, 0, 0), mfp::cell_view(d, 0, 0, 0), lo3(T), dx3(T), kZero,
                               T);
    }
    template <class C> MFD static void fluid_source(const Tab &T, const C &cell, double *src) {
        double ph[NV];
#pragma unroll
        for (int n = 0; n < NV; ++n) ph[n] = cell(n);
        for (int n = 0; n < FLO; ++n) src[n] = kZero;
        U::compute_fluid_source(0, 0, 0, mfp::cell_view(ph, 0, 0, 0), mfp::cell_view(src, 0, 0, 0), lo3(T), dx3(T),
                                kZero, T);
    }
    template <class C> MFD static void spec_source(const Tab &T, const C &cell, double *src) {
        double ph[NV], out[NV];
#pragma unroll
        for (int n = 0; n < NV; ++n) {
            ph[n] = cell(n);
            out[n] = kZero;
        }
        U::compute_spec_source(0, 0, 0, mfp::cell_view(ph, 0, 0, 0), mfp::cell_view(out, 0, 0, 0), lo3(T), dx3(T), kZero,
                               T);
#pragma unroll
        for (int s = 0; s < NS; ++s) src[s] = out[FLO + s];
    }
    MFD static void rate_constants(const Tab &, double, double *) {}
    template <class C> MFD static void spec_source_k(const Tab &T, const double *, const C &cell, double *src) {
        spec_source(T, cell, src);
    }
    MFD static void externalbc(const Tab &T, const double *x, const double *s_int, double *s_ext, int idir, int sgn,
                               double time) {
        U::externalbc(x, s_int, s_ext, idir, sgn, time, T.geo, T);
    }
};

// ---------------------------------------------------------------------------
// Numerics of Source/Src_3d
// ---------------------------------------------------------------------------
template <class P>
struct Num {
    static constexpr int NS = P::NS;
    static constexpr int NC = NCV + NS;
    static constexpr int NV = FLO + NS;

    // consprim_funcs.H:9-23
    MFD static void cons_to_prim(const Tab &T, const double *u, double *p) {
        p[0] = u[RHO];
        p[1] = u[RHOU] / u[RHO];
        p[2] = u[RHOV] / u[RHO];
        p[3] = u[RHOW] / u[RHO];
        double rhoe = u[RHOE] - kHalf * u[RHO] * (p[1] * p[1] + p[2] * p[2] + p[3] * p[3]);
        p[4] = P::get_p_from_rec(T, rhoe, u[RHO], u + NCV);
    }
    // consprim_funcs.H:26-44 (u holds the species on entry)
    MFD static void prim_to_cons(const Tab &T, double *u, const double *p) {
        u[RHO] = p[0];
        u[RHOU] = p[0] * p[1];
        u[RHOV] = p[0] * p[2];
        u[RHOW] = p[0] * p[3];
        u[RHOE] = u[RHO] * P::get_e_from_rpc(T, p[0], p[4], u + NCV);
        u[RHOE] += kHalf * u[RHO] * (p[1] * p[1] + p[2] * p[2] + p[3] * p[3]);
    }
    // pressure and velocities used by geteulerflux (consprim_funcs.H:57-66)
    MFD static void euler_prims(const Tab &T, const double *u, double &vx, double &vy, double &vz, double &p) {
        double rho = u[RHO];
        const double irho = div_pos(kOne, rho);
        vx = u[RHOU] * irho;
        vy = u[RHOV] * irho;
        vz = u[RHOW] * irho;
        double rhoe = u[RHOE] - kHalf * u[RHO] * (vx * vx + vy * vy + vz * vz);
        p = P::get_p_from_rec(T, rhoe, u[RHO], u + NCV);
    }
    // component c of geteulerflux (consprim_funcs.H:68-80) for the unit normal of
    // direction D; the products with zero normal components add exact zeros.
    template <int D>
    MFD static double euler_comp(int c, const double *u, double vx, double vy, double vz, double p) {
        const double v[3] = {vx, vy, vz};
        const double vn = v[D];
        const double rho = u[RHO];
        if (c == RHO) return rho * vn;
        if (c >= RHOU && c <= RHOW) {
            const int a = c - RHOU;
            double m = rho * v[a];
            return (a == D) ? (m * v[a] + p) : (m * vn);
        }
        if (c == RHOE) return (u[RHOE] + p) * vn;
        return u[c] * vn;
    }
    // flux_schemes.H:10-46 (presR evaluated with the left species)
    template <int D>
    MFD static double max_eigenvalue(const Tab &T, const double *ul, const double *ur) {
        double rhoL = ul[RHO], rhoR = ur[RHO];
        double rhoeL = ul[RHOE] - kHalf * (ul[RHOU] * ul[RHOU] + ul[RHOV] * ul[RHOV] + ul[RHOW] * ul[RHOW]) / ul[RHO];
        double rhoeR = ur[RHOE] - kHalf * (ur[RHOU] * ur[RHOU] + ur[RHOV] * ur[RHOV] + ur[RHOW] * ur[RHOW]) / ur[RHO];
        double presL = P::get_p_from_rec(T, rhoeL, ul[RHO], ul + NCV);
        double presR = P::get_p_from_rec(T, rhoeR, ur[RHO], ul + NCV);
        double velnL = ul[RHOU + D] / ul[RHO];
        double velnR = ur[RHOU + D] / ur[RHO];
        double veln = std_max(fabs(velnL), fabs(velnR));
        double gamaL = P::get_gama_from_rpc(T, rhoL, presL, ul + NCV);
        double gamaR = P::get_gama_from_rpc(T, rhoR, presR, ur + NCV);
        double ahalf = std_max(sqrt(gamaL * presL / rhoL), sqrt(gamaR * presR / rhoR));
        return ahalf + veln;
    }
    // flux_schemes.H:105-141
    MFD static double split_mach(double M, double beta, bool minus) {
        if (minus) {
            if (fabs(M) <= kOne) return -kFourth * (M - kOne) * (M - kOne) - beta * (M * M - kOne) * (M * M - kOne);
            return kHalf * (M - fabs(M));
        }
        if (fabs(M) <= 1) return kFourth * (M + kOne) * (M + kOne) + beta * (M * M - kOne) * (M * M - kOne);
        return kHalf * (M + fabs(M));
    }
    // flux_schemes.H:143-181
    MFD static double split_pres(double M, double alpha, bool minus) {
        if (minus) {
            if (fabs(M) <= kOne) return kFourth * (M - kOne) * (M - kOne) * (kTwo + M) - alpha * M * (M * M - kOne) * (M * M - kOne);
            return kHalf * (M - fabs(M)) / M;
        }
        if (fabs(M) <= kOne) return kFourth * (M + kOne) * (M + kOne) * (kTwo - M) + alpha * M * (M * M - kOne) * (M * M - kOne);
        return kHalf * (M + fabs(M)) / M;
    }
    // flux_schemes.H:183-310, Minf = 0.3, normal = unit vector of D
    template <int D>
    MFD static void ausmp_up(const Tab &T, const double *ul, const double *ur, double *fh) {
        const double Minf2 = kThreeTenth * kThreeTenth;
        const double delta = kThreeTenth, beta = kEighth, sigma = kOne, Kp = kFourth, Ku = kThreeFourth;
        double pl[5], pr[5];
        cons_to_prim(T, ul, pl);
        cons_to_prim(T, ur, pr);
        double gamaL = P::get_gama_from_rpc(T, pl[0], pl[4], ul + NCV);
        double gamaR = P::get_gama_from_rpc(T, pr[0], pr[4], ur + NCV);
        // pl[1]*nx + pl[2]*ny + pl[3]*nz with one unit component
        double vL = pl[1 + D], vR = pr[1 + D];
        double HL = gamaL / (gamaL - kOne) * (pl[4] / pl[0]) + kHalf * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]);
        double HR = gamaR / (gamaR - kOne) * (pr[4] / pr[0]) + kHalf * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]);
        double rho_half = kHalf * (pl[0] + pr[0]);
        double astar2 = kTwo * (gamaL - kOne) / (gamaL + kOne) * HL;
        double ahatL = astar2 / std_max(sqrt(astar2), fabs(vL));
        astar2 = kTwo * (gamaR - kOne) / (gamaR + kOne) * HR;
        double ahatR = astar2 / std_max(sqrt(astar2), fabs(vR));
        double a_half = std_min(ahatL, ahatR);
        double ML = vL / a_half, MR = vR / a_half;
        double MLplus = split_mach(ML, beta, false);
        double MRminus = split_mach(MR, beta, true);
        double M_half = MLplus + MRminus;
        double M_bar2 = kHalf * (vL * vL + vR * vR) / (a_half * a_half);
        double Mo2 = std_min(kOne, std_max(M_bar2, Minf2));
        double Mo = sqrt(Mo2);
        double fa_Mo = Mo * (kTwo - Mo);
        M_half = M_half - (Kp / fa_Mo) * std_max(kOne - sigma * M_bar2, kZero) * (pr[4] - pl[4]) / (rho_half * a_half * a_half);
        double mod_M_half = fabs(M_half);
        if (mod_M_half <= delta) mod_M_half = (M_half * M_half + delta * delta) / (kTwo * delta);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            double fL = (c == RHOE) ? ul[RHOE] + pl[4] : ul[c];
            double fR = (c == RHOE) ? ur[RHOE] + pr[4] : ur[c];
            fh[c] = (kHalf * M_half * (fL + fR) - kHalf * mod_M_half * (fR - fL)) * a_half;
        }
        double alpha = kThree * kSixteenth * (-kFour + kFive * fa_Mo * fa_Mo);
        double PLplus = split_pres(ML, alpha, false);
        double PRminus = split_pres(MR, alpha, true);
        double p_half = (PLplus * pl[4] + PRminus * pr[4]);
        p_half = p_half - Ku * PLplus * PRminus * (pl[0] + pr[0]) * (fa_Mo * a_half) * (vR - vL);
        // fhalf[RHOU+a] += p_half*n[a]: only the normal component changes
        // (adding p_half*0 to the others is exact)
        fh[RHOU + D] += p_half;
    }
    // WENO5-JS half of flux_schemes.H:400-421 (left-biased value at the face
    // between s2 and s3) and :423-444 (right-biased value), eps = 1e-6, p = 2,
    // linear weights (0.1, 0.6, 0.3). With q_k = (eps + beta_k)^2 the nonlinear
    // weights alpha_k = d_k / q_k are normalised with one division:
    //   sum_k alpha_k v_k / sum_k alpha_k
    //     = (v2 q1 q0 + 6 v1 q2 q0 + 3 v0 q2 q1) / (q1 q0 + 6 q2 q0 + 3 q2 q1)
    // (the reference forms 1/q_k and 1/sum alpha; the value is the same to
    // rounding). Building with MF_WENO_REFERENCE_DIVISIONS restores the
    // reference's four divisions.
    MFD static double weno_combine(double b2, double b1, double b0, double v2, double v1, double v0) {
        const double eps = 1e-6;
#ifdef MF_WENO_REFERENCE_DIVISIONS
        b2 = 1.0 / ((eps + b2) * (eps + b2));
        b1 = 1.0 / ((eps + b1) * (eps + b1));
        b0 = 1.0 / ((eps + b0) * (eps + b0));
        double a2 = b2, a1 = 6.0 * b1, a0 = 3.0 * b0;
        double alpha1 = 1.0 / (a2 + a1 + a0);
        return (1.0 / 6.0) * alpha1 * (a2 * v2 + a1 * v1 + a0 * v0);
#else
        const double q2 = (eps + b2) * (eps + b2);
        const double q1 = (eps + b1) * (eps + b1);
        const double q0 = (eps + b0) * (eps + b0);
        const double w2 = q1 * q0, w1 = 6.0 * (q2 * q0), w0 = 3.0 * (q2 * q1);
        return div_pos(w2 * v2 + w1 * v1 + w0 * v0, 6.0 * (w2 + w1 + w0));
#endif
    }
    MFD static double weno_plus(double s0, double s1, double s2, double s3, double s4) {
        double t = s0 - 2.0 * s1 + s2;
        double q = s0 - 4.0 * s1 + 3.0 * s2;
        const double b2 = (13.0 / 12.0) * (t * t) + 0.25 * (q * q);
        t = s1 - 2.0 * s2 + s3;
        const double b1 = (13.0 / 12.0) * (t * t) + 0.25 * (s1 - s3) * (s1 - s3);
        t = s2 - 2.0 * s3 + s4;
        q = 3.0 * s2 - 4.0 * s3 + s4;
        const double b0 = (13.0 / 12.0) * (t * t) + 0.25 * (q * q);
        const double v2 = 2.0 * s0 - 7.0 * s1 + 11.0 * s2;
        const double v1 = -s1 + 5.0 * s2 + 2.0 * s3;
        const double v0 = 2.0 * s2 + 5.0 * s3 - s4;
        return weno_combine(b2, b1, b0, v2, v1, v0);
    }
    MFD static double weno_minus(double s1, double s2, double s3, double s4, double s5) {
        double t = s5 - 2.0 * s4 + s3;
        double q = s5 - 4.0 * s4 + 3.0 * s3;
        const double b2 = (13.0 / 12.0) * (t * t) + 0.25 * (q * q);
        t = s4 - 2.0 * s3 + s2;
        const double b1 = (13.0 / 12.0) * (t * t) + 0.25 * (s4 - s2) * (s4 - s2);
        t = s3 - 2.0 * s2 + s1;
        q = 3.0 * s3 - 4.0 * s2 + s1;
        const double b0 = (13.0 / 12.0) * (t * t) + 0.25 * (q * q);
        const double v2 = 11.0 * s3 - 7.0 * s4 + 2.0 * s5;
        const double v1 = -s4 + 5.0 * s3 + 2.0 * s2;
        const double v0 = 2.0 * s3 + 5.0 * s2 - s1;
        return weno_combine(b2, b1, b0, v2, v1, v0);
    }
    // compute_flux_3d.H:16-22
    MFD static double harmonic_mean(double d1, double d2) { return 2.0 * d1 * d2 / (d1 + d2 + 1e-30); }
    // the same for non-negative coefficients, with div_pos
    MFD static double harmonic_mean_pos(double d1, double d2) { return div_pos(2.0 * d1 * d2, d1 + d2 + 1e-30); }
};

}  // namespace mfd
