// mf_kernels.inc.h -- kernel templates of the RHS + RK3 path, instantiated once
// per plugin family by the mf_kernels_<family>.hip translation units.
#pragma once
#include "mf_args.h"

namespace mfd {

MFD double stage_time(const StageArgs &a) {
    if (!a.ts) return a.time_host;
    return a.stage == 1 ? a.ts->t : a.ts->t + a.ts->dt;
}
MFD double stage_dt(const StageArgs &a) { return a.ts ? a.ts->dt : a.dt_host; }
MFD bool step_skipped(const TimeState *ts) { return ts && (!ts->active || ts->err); }

MFD void raise_err(int *err, int code) {
    if (err) atomicCAS(err, 0, code);
}

// Tiles of a launch: with a grid that is a multiple of 8, block b runs on XCD
// b % 8 (round-robin dispatch); each XCD then walks its own contiguous eighth of
// the tile list, so tiles that share halo cells share that XCD's L2.
struct TileIter {
    int t, step, end;
    MFD TileIter(const Work &w) {
        const int G = gridDim.x, b = blockIdx.x, n = w.t1 - w.t0;
        if ((G & 7) == 0 && n >= G) {
            const int g = b & 7, r = b >> 3, per = G >> 3;
            const int chunk = (n + 7) >> 3;
            const int s0 = w.t0 + g * chunk;
            end = min(w.t1, s0 + chunk);
            t = s0 + r;
            step = per;
        } else {
            t = w.t0 + b;
            step = G;
            end = w.t1;
        }
    }
    MFD bool ok() const { return t < end; }
    MFD void next() { t += step; }
};

// tile t of the work list -> its box and inclusive cell range
MFD const BoxDesc &tile_box(const Work &w, int t, int &lt) {
    if (!w.boxes) { lt = t; return w.one; }
    int lo = 0, hi = w.nboxes - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (w.boxes[mid].tile_start[w.cls] <= t) lo = mid; else hi = mid - 1;
    }
    lt = t - w.boxes[lo].tile_start[w.cls];
    return w.boxes[lo];
}
MFD void tile_range(const BoxDesc &b, const Work &w, int lt, int *c0, int *c1) {
    const int *nt = b.ntile[w.cls];
    int ti = lt % nt[0];
    int r = lt / nt[0];
    int tj = r % nt[1];
    int tk = r / nt[1];
    int tt[3] = {ti, tj, tk};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        c0[d] = b.lo[d] + tt[d] * w.tsh[d];
        c1[d] = min(b.hi[d], c0[d] + w.tsh[d] - 1);
    }
}
MFD Acc make_acc(const StageArgs &a, const BoxDesc &b) {
    Acc acc;
    acc.s = a.src;
    acc.cutid = a.cutid;
    acc.cutbuf = a.cutbuf;
    acc.ncut = a.ncut;
    acc.nx = a.nx;
    acc.ny = a.ny;
#pragma unroll
    for (int d = 0; d < 3; ++d) { acc.blo[d] = b.lo[d]; acc.bhi[d] = b.hi[d]; }
    return acc;
}

// flux component c (0..4 conserved, 5.. species) -> component of the flux view
MFD int flux_comp(int c, int full) { return full ? (c < NCV ? c : c + (FLO - NCV)) : c; }

// pass 2 of a face: viscous stress and work (:293-367), conduction (:373-374),
// species diffusion (:425-428), applied to the stored hyperbolic part
template <class P, int D>
MFD void face_pass2(const Tab &T, const Acc &acc, const StageArgs &a, int i, int j, int k, const View &F,
                    const CellPtr &pL, const CellPtr &pR, int cL, int cR, bool hyp) {
    using N = Num<P>;
    constexpr int NS = P::NS;
    constexpr int T1 = (D + 1) % 3, T2 = (D + 2) % 3;
    constexpr int e0 = D == 0, e1 = D == 1, e2 = D == 2;
    const int iL = i - e0, jL = j - e1, kL = k - e2;
    const double dxs = a.dx[D];
    const bool ns = a.nsflag != 0;
    const int full = a.flux_full;
    double *Fp = F.p + F.off(i, j, k);
    const int64_t Fs = F.ns;
    auto fput = [&](int c, double v) { Fp[(int64_t)flux_comp(c, full) * Fs] = v; };
    auto fget = [&](int c) { return Fp[(int64_t)flux_comp(c, full) * Fs]; };
    if (ns) {
        double visc, thcond;
        {
            // harmonic mean of (i,j,k) and (i-1,j,k) for every sweep direction (:179-180)
            double v0, k0, v1, k1;
            CellRef c0{&acc, i, j, k, cR};
            P::fluid_transport(T, c0, v0, k0);
            CellRef c1{&acc, i - 1, j, k, acc.cid(i - 1, j, k)};
            P::fluid_transport(T, c1, v1, k1);
            visc = N::harmonic_mean(v0, v1);
            thcond = N::harmonic_mean(k0, k1);
        }
        if (hyp) {
            double g[3][3];  // g[q][m]: d(vel q) along normal (m=0), t1 (m=1), t2 (m=2)
            constexpr int p10 = T1 == 0, p11 = T1 == 1, p12 = T1 == 2;
            constexpr int p20 = T2 == 0, p21 = T2 == 1, p22 = T2 == 2;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const int V = VELX + q;
                g[q][0] = (pR(V) - pL(V)) / dxs;
                double top = kHalf * (acc(iL + p10, jL + p11, kL + p12, V) + acc(i + p10, j + p11, k + p12, V));
                double bot = kHalf * (acc(iL - p10, jL - p11, kL - p12, V) + acc(i - p10, j - p11, k - p12, V));
                g[q][1] = (top - bot) / (kTwo * a.dx[T1]);
                double frt = kHalf * (acc(iL + p20, jL + p21, kL + p22, V) + acc(i + p20, j + p21, k + p22, V));
                double bck = kHalf * (acc(iL - p20, jL - p21, kL - p22, V) + acc(i - p20, j - p21, k - p22, V));
                g[q][2] = (frt - bck) / (kTwo * a.dx[T2]);
            }
            // axis -> gradient slot (GET_XDIR/GET_YDIR/GET_ZDIR, globalDefines.H:49-51)
            constexpr int sx = D == 0 ? 0 : (T1 == 0 ? 1 : 2);
            constexpr int sy = D == 1 ? 0 : (T1 == 1 ? 1 : 2);
            constexpr int sz = D == 2 ? 0 : (T1 == 2 ? 1 : 2);
            double dudx = g[0][sx], dudy = g[0][sy], dudz = g[0][sz];
            double dvdx = g[1][sx], dvdy = g[1][sy], dvdz = g[1][sz];
            double dwdx = g[2][sx], dwdy = g[2][sy], dwdz = g[2][sz];
            double delu = dudx + dvdy + dwdz;
            double t0, t1, t2;
            if (D == 0) {
                t0 = visc * (kTwo * dudx - kTwo3rd * delu);
                t1 = visc * (dudy + dvdx);
                t2 = visc * (dudz + dwdx);
            } else if (D == 1) {
                t0 = visc * (dvdx + dudy);
                t1 = visc * (kTwo * dvdy - kTwo3rd * delu);
                t2 = visc * (dvdz + dwdy);
            } else {
                t0 = visc * (dwdx + dudz);
                t1 = visc * (dwdy + dvdz);
                t2 = visc * (kTwo * dwdz - kTwo3rd * delu);
            }
            fput(RHOU, fget(RHOU) - t0);
            fput(RHOV, fget(RHOV) - t1);
            fput(RHOW, fget(RHOW) - t2);
            double e = fget(RHOE);
            e -= t0 * (kHalf * (pL(VELX) + pR(VELX)));
            e -= t1 * (kHalf * (pL(VELY) + pR(VELY)));
            e -= t2 * (kHalf * (pL(VELZ) + pR(VELZ)));
            fput(RHOE, e);
        }
        double dTdn = (pR(TEMP) - pL(TEMP)) / dxs;
        fput(RHOE, fget(RHOE) - thcond * dTdn);
    }
    {
        double dl[NS], dr[NS];
        CellRef l{&acc, iL, jL, kL, cL};
        CellRef r{&acc, i, j, k, cR};
        P::spec_dcoeff(T, l, dl);
        P::spec_dcoeff(T, r, dr);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double Ds = N::harmonic_mean(dl[s], dr[s]);
            fput(NCV + s, fget(NCV + s) - Ds * (pR(FLO + s) - pL(FLO + s)) / dxs);
        }
    }
}

// ---------------------------------------------------------------------------
// Face flux of direction D at face (i,j,k) (between (i,j,k)-e_D and (i,j,k)):
// Source/Src_3d/compute_flux_3d.H:86-430. Pass 1 stores the hyperbolic (or
// convective) part of every flux component; pass 2 applies the viscous stress
// and work, conduction and species diffusion to the stored values in the
// reference's order, so the stencil data and the viscous data are never live
// at the same time.
// ---------------------------------------------------------------------------
template <class P, int ORDER, int D>
MFD void face_flux(const Tab &T, const Acc &acc, const StageArgs &a, int i, int j, int k, const View &F, double time) {
    using N = Num<P>;
    constexpr int NS = P::NS, NC = N::NC;
    constexpr int T1 = (D + 1) % 3, T2 = (D + 2) % 3;
    constexpr int e0 = D == 0, e1 = D == 1, e2 = D == 2;
    const int iL = i - e0, jL = j - e1, kL = k - e2;
    const int cL = acc.cid(iL, jL, kL), cR = acc.cid(i, j, k);
    const CellPtr pL = cell_ptr(acc, iL, jL, kL, cL), pR = cell_ptr(acc, i, j, k, cR);
    const double facevfrac = kHalf * (pL(VFRAC) + pR(VFRAC));
    const double dxs = a.dx[D];
    const bool ns = a.nsflag != 0;
    const bool hyp = ns && facevfrac >= 0.5;
    const int full = a.flux_full;
    double *Fp = F.p + F.off(i, j, k);
    const int64_t Fs = F.ns;
    auto fput = [&](int c, double v) { Fp[(int64_t)flux_comp(c, full) * Fs] = v; };
    auto fget = [&](int c) { return Fp[(int64_t)flux_comp(c, full) * Fs]; };

    // ---- pass 1: hyperbolic / convective part ----
    if (!ns) {
        // species convection without the flow solve (:379-422)
        double velface = kHalf * (pL(VELX + D) + pR(VELX + D));
        if (a.spec_in_solid) {
            double w = (pL(VFRAC) < kOne) ? kZero : kOne;  // both weights test the left cell
            velface = (w * pL(VELX + D) + w * pR(VELX + D)) / (w + w + kTiny);
        }
#pragma unroll
        for (int c = 0; c < NCV; ++c) fput(c, kZero);
#pragma unroll
        for (int s = 0; s < NS; ++s)
            fput(NCV + s, T.advect_flag[s] * (kHalf * (velface + fabs(velface)) * pL(FLO + s) +
                                              kHalf * (velface - fabs(velface)) * pR(FLO + s)));
    } else if (!hyp) {
#pragma unroll
        for (int c = 0; c < NC; ++c) fput(c, kZero);
    } else {
        // stencil along D: q = 0..5 is offset q-3; IB first-order fallback (:233-249)
        CellPtr cp[6];
        bool first = false;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            if (q == 2) cp[q] = pL;
            else if (q == 3) cp[q] = pR;
            else {
                const int ii = i + (q - 3) * e0, jj = j + (q - 3) * e1, kk = k + (q - 3) * e2;
                cp[q] = cell_ptr(acc, ii, jj, kk, acc.cid(ii, jj, kk));
            }
            first = first || (cp[q](VFRAC) < kOne);
        }
        if (first) { cp[0] = pL; cp[1] = pL; cp[4] = pR; cp[5] = pR; }
        auto load = [&](const CellPtr &c, double *u) {
#pragma unroll
            for (int n = 0; n < NCV; ++n) u[n] = c(n);
#pragma unroll
            for (int n = 0; n < NS; ++n) u[NCV + n] = c(FLO + n);
        };
        auto adv = [&](int c, double h) { return c < NCV ? h : T.advect_flag[c - NCV] * h; };
        if (ORDER == 5) {
            // WENO5 flux splitting, flux_schemes.H:448-495
            double lam;
            {
                double ul[NC], ur[NC];
                load(pL, ul);
                load(pR, ur);
                lam = N::template max_eigenvalue<D>(T, ul, ur);
            }
            double pr[6][5];  // rho, vx, vy, vz, p of the six stencil cells
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                double u[NC];
                load(cp[q], u);
                pr[q][0] = u[RHO];
                N::euler_prims(T, u, pr[q][1], pr[q][2], pr[q][3], pr[q][4]);
            }
#pragma unroll 1
            for (int c = 0; c < NC; ++c) {
                const int comp = c < NCV ? c : FLO + (c - NCV);
                double uc[6], fc[6];
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    uc[q] = cp[q](comp);
                    const double vn = pr[q][1 + D];
                    if (c == RHO) fc[q] = pr[q][0] * vn;
                    else if (c >= RHOU && c <= RHOW) {
                        const double va = (c == RHOU) ? pr[q][1] : ((c == RHOV) ? pr[q][2] : pr[q][3]);
                        const double m = pr[q][0] * va;
                        fc[q] = (c - RHOU == D) ? (m * va + pr[q][4]) : (m * vn);
                    } else if (c == RHOE) fc[q] = (uc[q] + pr[q][4]) * vn;
                    else fc[q] = uc[q] * vn;
                }
                double fp = N::weno_plus(fc[0], fc[1], fc[2], fc[3], fc[4]);
                double fm = N::weno_minus(fc[1], fc[2], fc[3], fc[4], fc[5]);
                double up = N::weno_plus(uc[0], uc[1], uc[2], uc[3], uc[4]);
                double um = N::weno_minus(uc[1], uc[2], uc[3], uc[4], uc[5]);
                fput(c, adv(c, kHalf * (fp + lam * up) + kHalf * (fm - lam * um)));
            }
        } else if (ORDER == 1 || ORDER == 2) {
            double ul[NC], ur[NC], fh[NC];
            load(pL, ul);
            load(pR, ur);
            if (ORDER == 2) {
                // compute_flux_3d.H:31-65 with uim2 = ulm1, uip1 = urp1
                double ulm1[NC], urp1[NC];
                load(cp[1], ulm1);
                load(cp[4], urp1);
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    double d1 = urp1[c] - ur[c];
                    double d2 = ul[c] - ulm1[c];
                    double r = fabs((d1 - d2) / (fabs(d1) + fabs(d2) + kTiny));
                    double psi = kOne - r * r;
                    double L = kHalf * psi * (d1 + d2);
                    ul[c] = ul[c] + kHalf * L;
                    ur[c] = ur[c] - kHalf * L;
                }
            }
            N::template ausmp_up<D>(T, ul, ur, fh);
#pragma unroll
            for (int c = 0; c < NC; ++c) fput(c, adv(c, fh[c]));
        } else {
            // flux_schemes.H:312-361 with uim1 = ul, ui = ur, uim2 = ulm1, uip1 = urp1
            double um2[NC], um1[NC], u0[NC], up1[NC];
            load(cp[1], um2);
            load(pL, um1);
            load(pR, u0);
            load(cp[4], up1);
            double pv[4][4];
            N::euler_prims(T, um2, pv[0][0], pv[0][1], pv[0][2], pv[0][3]);
            N::euler_prims(T, um1, pv[1][0], pv[1][1], pv[1][2], pv[1][3]);
            N::euler_prims(T, u0, pv[2][0], pv[2][1], pv[2][2], pv[2][3]);
            N::euler_prims(T, up1, pv[3][0], pv[3][1], pv[3][2], pv[3][3]);
            double rhoL = um1[RHO], rhoR = u0[RHO];
            double rhoeL = um1[RHOE] - kHalf * (um1[RHOU] * um1[RHOU] + um1[RHOV] * um1[RHOV] + um1[RHOW] * um1[RHOW]) / um1[RHO];
            double rhoeR = u0[RHOE] - kHalf * (u0[RHOU] * u0[RHOU] + u0[RHOV] * u0[RHOV] + u0[RHOW] * u0[RHOW]) / u0[RHO];
            double presL = P::get_p_from_rec(T, rhoeL, um1[RHO], um1 + NCV);
            double presR = P::get_p_from_rec(T, rhoeR, u0[RHO], u0 + NCV);
            double veln = std_max(fabs(um1[RHOU + D] / um1[RHO]), fabs(u0[RHOU + D] / u0[RHO]));
            double gamaL = P::get_gama_from_rpc(T, rhoL, presL, um1 + NCV);
            double gamaR = P::get_gama_from_rpc(T, rhoR, presR, u0 + NCV);
            double ahalf = std_max(sqrt(gamaL * presL / rhoL), sqrt(gamaR * presR / rhoR));
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                double fm2 = N::template euler_comp<D>(c, um2, pv[0][0], pv[0][1], pv[0][2], pv[0][3]);
                double fm1 = N::template euler_comp<D>(c, um1, pv[1][0], pv[1][1], pv[1][2], pv[1][3]);
                double f0 = N::template euler_comp<D>(c, u0, pv[2][0], pv[2][1], pv[2][2], pv[2][3]);
                double fp1 = N::template euler_comp<D>(c, up1, pv[3][0], pv[3][1], pv[3][2], pv[3][3]);
                double v = kTwelveth * (7 * f0 + 7 * fm1 - fm2 - fp1);
                v -= kTwelveth * (3 * u0[c] - 3 * um1[c] - up1[c] + u0[c]) * (veln + ahalf) * a.dissfactor;
                fput(c, adv(c, v));
            }
        }
    }
    if (full) {
#pragma unroll
        for (int n = NCV; n < FLO; ++n) F(i, j, k, n) = kZero;
    }

    face_pass2<P, D>(T, acc, a, i, j, k, F, pL, pR, cL, cR, hyp);
}

template <class P, int ORDER, int D>
__global__ __launch_bounds__(256) void k_faces(StageArgs a, Work w) {
    if (step_skipped(a.ts)) return;
    const Tab &T = *a.tab;
    const double time = stage_time(a);
    for (TileIter it(w); it.ok(); it.next()) {
        const int t = it.t;
        int lt;
        const BoxDesc &b = tile_box(w, t, lt);
        int c0[3], c1[3];
        tile_range(b, w, lt, c0, c1);
        if (c1[D] == b.hi[D]) c1[D] += 1;  // growHi(bx, D, 1)
        const int ex = c1[0] - c0[0] + 1, ey = c1[1] - c0[1] + 1, ez = c1[2] - c0[2] + 1;
        const int nf = ex * ey * ez;
        Acc acc = make_acc(a, b);
        View F = b.flux[D];
        for (int f = threadIdx.x; f < nf; f += blockDim.x) {
            int fi = f % ex;
            int r = f / ex;
            int fj = r % ey;
            int fk = r / ey;
            face_flux<P, ORDER, D>(T, acc, a, c0[0] + fi, c0[1] + fj, c0[2] + fk, F, time);
        }
    }
}

// ---------------------------------------------------------------------------
// WENO5 face sweep of direction D with the stencil staged in LDS. A tile of
// faces first stages every stencil cell once (conserved state, the velocities
// and pressure of geteulerflux, gamma, vfrac and the cell's cut-cell index);
// each face then reads its six cells from LDS. Per face the viscous stress and
// work, conduction and species diffusion terms are formed first and kept in
// registers, then each flux component is reconstructed and written once with
// those terms subtracted in the reference's order. Same arithmetic as
// face_flux<P,5,D>; the plugin's gamma must not depend on the pressure (then
// gamma(rho_R, presR, spec_R) of the eigenvalue equals the staged gamma).
// ---------------------------------------------------------------------------
// component n of cell (i,j,k) as a box sees it: the array value, loaded
// independently of the cut-index load, replaced for an own-box cut cell
MFD double sel_get(const Acc &acc, int i, int j, int k, int n, int cid) {
    double v = acc.s(i, j, k, n);
    if (cid >= 0) v = acc.cutbuf[(int64_t)n * acc.ncut + cid];
    return v;
}
// convective flux h of component c with the viscous, conduction and species
// diffusion terms of face_pass2 subtracted in the same order
MFD double flux_with_terms(int c, double h, bool hyp, double dm0, double dm1, double dm2, double de0, double de1, double de2,
                           double de3, double dsp) {
    double v = h;
    if (c >= RHOU && c <= RHOW) {
        if (hyp) v = v - (c == RHOU ? dm0 : (c == RHOV ? dm1 : dm2));
    } else if (c == RHOE) {
        if (hyp) {
            v -= de0;
            v -= de1;
            v -= de2;
        }
        v = v - de3;
    } else if (c >= NCV) {
        v = v - dsp;
    }
    return v;
}
// a staged cell seen by a pointwise plugin: conserved components, species and
// vfrac from its LDS record, the primitives from memory
template <int NC>
struct LdsCell {
    const double *rec;
    const Acc *acc;
    int i, j, k, cid;
    MFD double operator()(int n) const {
        if (n < NCV) return rec[n];
        if (n >= FLO) return rec[NCV + n - FLO];
        if (n == VFRAC) return rec[NC + 5];
        return sel_get(*acc, i, j, k, n, cid);
    }
};

// minimum waves per SIMD the compiler must keep room for (register budget);
// blocks of 256 or 512 threads (512 for tiles of at least 512 faces)
#ifndef MF_FACE_MIN_WAVES
#define MF_FACE_MIN_WAVES 2
#endif
// 1: the per-component reconstruction loop is kept rolled (one component's
// WENO stencils live at a time: fewer registers, more waves per SIMD)
#ifndef MF_FACE_CLOOP
#define MF_FACE_CLOOP 0
#endif
// element s of a small register array for a run-time s (a select chain: no
// dynamic indexing, which would move the array to scratch memory)
template <int M>
MFD double pick(const double (&v)[M], int s) {
    double r = 0.0;
#pragma unroll
    for (int q = 0; q < M; ++q) r = (q == s) ? v[q] : r;
    return r;
}
template <class P, int D>
__global__ __launch_bounds__(512, MF_FACE_MIN_WAVES) void k_faces_lds(StageArgs a, Work w) {
    using N = Num<P>;
    constexpr int NS = P::NS, NC = N::NC;
    constexpr int e0 = D == 0, e1 = D == 1, e2 = D == 2;
    constexpr int T1 = (D + 1) % 3, T2 = (D + 2) % 3;
    // LDS tile: one record of FS doubles per stencil cell (an odd stride keeps
    // 64-bit reads of consecutive cells on distinct banks)
    constexpr int FVX = NC, FP = NC + 3, FG = NC + 4, FVF = NC + 5, NF = NC + 6, FS = NF | 1;
    static_assert(P::kGammaIgnoresP, "the staged gamma needs a gamma that does not depend on the pressure");
    extern __shared__ double lds[];
    if (step_skipped(a.ts)) return;
    const Tab &T = *a.tab;
    // reciprocals of the spacings of the normal and transverse differences
    // (x * (1/dx) instead of x / dx: rounding differs in the last bit)
    const double idn = 1.0 / a.dx[D], it1 = 1.0 / (kTwo * a.dx[T1]), it2 = 1.0 / (kTwo * a.dx[T2]);
    const int full = a.flux_full;
    for (TileIter it(w); it.ok(); it.next()) {
        const int t = it.t;
        int lt;
        const BoxDesc &b = tile_box(w, t, lt);
        int c0[3], c1[3];
        tile_range(b, w, lt, c0, c1);
        if (c1[D] == b.hi[D]) c1[D] += 1;  // growHi(bx, D, 1)
        const int fx = c1[0] - c0[0] + 1, fy = c1[1] - c0[1] + 1, fz = c1[2] - c0[2] + 1;
        const int sx = fx + 5 * e0, sy = fy + 5 * e1, sz = fz + 5 * e2;
        const int ncs = sx * sy * sz;
        const int ox = c0[0] - 3 * e0, oy = c0[1] - 3 * e1, oz = c0[2] - 3 * e2;
        int *lcid = reinterpret_cast<int *>(lds + FS * ncs);
        Acc acc = make_acc(a, b);
        // staging: one pointer per cell (its cut values or the state), so that
        // a cut cell loads its values once and the wave does not branch
        for (int q = threadIdx.x; q < ncs; q += blockDim.x) {
            const int ix = q % sx, r = q / sx, iy = r % sy, iz = r / sy;
            const int ci = ox + ix, cj = oy + iy, ck = oz + iz;
            const int cid = acc.cid(ci, cj, ck);
            const CellPtr cp = cell_ptr(acc, ci, cj, ck, cid);
            double u[NC];
#pragma unroll
            for (int n = 0; n < NCV; ++n) u[n] = cp(n);
#pragma unroll
            for (int n = 0; n < NS; ++n) u[NCV + n] = cp(FLO + n);
            double *rec = lds + q * FS;
#pragma unroll
            for (int c = 0; c < NC; ++c) rec[c] = u[c];
            double vx, vy, vz, p;
            N::euler_prims(T, u, vx, vy, vz, p);
            rec[FVX + 0] = vx;
            rec[FVX + 1] = vy;
            rec[FVX + 2] = vz;
            rec[FP] = p;
            rec[FG] = P::get_gama_from_rpc(T, u[RHO], p, u + NCV);
            rec[FVF] = cp(VFRAC);
            lcid[q] = cid;
        }
        __syncthreads();
        const int nf = fx * fy * fz;
        const int sD = D == 0 ? 1 : (D == 1 ? sx : sx * sy);
        for (int f = threadIdx.x; f < nf; f += blockDim.x) {
            const int ix = f % fx, r = f / fx, iy = r % fy, iz = r / fy;
            const int i = c0[0] + ix, j = c0[1] + iy, k = c0[2] + iz;
            const int q0 = ix + sx * (iy + sy * iz);
            const int iL = i - e0, jL = j - e1, kL = k - e2;
            const int qL = q0 + 2 * sD, qR = q0 + 3 * sD;
            const double *recL = lds + qL * FS, *recR = lds + qR * FS;
            const int cL = lcid[qL], cR = lcid[qR];
            const double vfL = recL[FVF], vfR = recR[FVF];
            const double facevfrac = kHalf * (vfL + vfR);
            const bool hyp = facevfrac >= 0.5;
            // The global loads of the viscous terms (velocities and temperatures
            // of L and R, cut ids and velocities of the eight transverse
            // neighbours) are issued first; the hyperbolic fluxes of the flow
            // components, computed from LDS, cover their latency; then the
            // viscous, conduction and diffusion terms, the flow fluxes' stores,
            // and the species components.
            constexpr int p10 = T1 == 0, p11 = T1 == 1, p12 = T1 == 2;
            constexpr int p20 = T2 == 0, p21 = T2 == 1, p22 = T2 == 2;
            int tcid[8];
            if (hyp) {
                tcid[0] = acc.cid(iL + p10, jL + p11, kL + p12); tcid[1] = acc.cid(i + p10, j + p11, k + p12);
                tcid[2] = acc.cid(iL - p10, jL - p11, kL - p12); tcid[3] = acc.cid(i - p10, j - p11, k - p12);
                tcid[4] = acc.cid(iL + p20, jL + p21, kL + p22); tcid[5] = acc.cid(i + p20, j + p21, k + p22);
                tcid[6] = acc.cid(iL - p20, jL - p21, kL - p22); tcid[7] = acc.cid(i - p20, j - p21, k - p22);
            }
            double pvL[3], pvR[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) { pvL[q] = sel_get(acc, iL, jL, kL, VELX + q, cL); pvR[q] = sel_get(acc, i, j, k, VELX + q, cR); }
            const double pTL = sel_get(acc, iL, jL, kL, TEMP, cL), pTR = sel_get(acc, i, j, k, TEMP, cR);
            double hc[NC], tv[8][3], lam = 0.0;
            const double *rs[6] = {recL, recL, recL, recR, recR, recR};
            if (hyp) {
                bool first = false;
#pragma unroll
                for (int m = 0; m < 6; ++m) first = first || (lds[(q0 + m * sD) * FS + FVF] < kOne);
                // stencil records, IB first-order fallback (:233-249)
                const double *rc0 = lds + (first ? qL : q0) * FS;
                const double *rc1 = lds + (first ? qL : q0 + sD) * FS;
                const double *rc4 = lds + (first ? qR : q0 + 4 * sD) * FS;
                const double *rc5 = lds + (first ? qR : q0 + 5 * sD) * FS;
                rs[0] = rc0; rs[1] = rc1; rs[4] = rc4; rs[5] = rc5;
                // max eigenvalue, flux_schemes.H:10-46 (presR with the left species)
                {
                    double ul[NC], ur[NC];
#pragma unroll
                    for (int c = 0; c < NC; ++c) { ul[c] = recL[c]; ur[c] = recR[c]; }
                    double rhoeL = ul[RHOE] - kHalf * (ul[RHOU] * ul[RHOU] + ul[RHOV] * ul[RHOV] + ul[RHOW] * ul[RHOW]) / ul[RHO];
                    double rhoeR = ur[RHOE] - kHalf * (ur[RHOU] * ur[RHOU] + ur[RHOV] * ur[RHOV] + ur[RHOW] * ur[RHOW]) / ur[RHO];
                    double presL = P::get_p_from_rec(T, rhoeL, ul[RHO], ul + NCV);
                    double presR = P::get_p_from_rec(T, rhoeR, ur[RHO], ul + NCV);
                    double veln = std_max(fabs(recL[FVX + D]), fabs(recR[FVX + D]));
                    double gamaL = recL[FG], gamaR = recR[FG];
                    double ahalf = std_max(sqrt(gamaL * presL / ul[RHO]), sqrt(gamaR * presR / ur[RHO]));
                    lam = ahalf + veln;
                }
                {
                    const int ti[8][3] = {{iL + p10, jL + p11, kL + p12}, {i + p10, j + p11, k + p12},
                                          {iL - p10, jL - p11, kL - p12}, {i - p10, j - p11, k - p12},
                                          {iL + p20, jL + p21, kL + p22}, {i + p20, j + p21, k + p22},
                                          {iL - p20, jL - p21, kL - p22}, {i - p20, j - p21, k - p22}};
#pragma unroll
                    for (int m = 0; m < 8; ++m)
#pragma unroll
                        for (int q = 0; q < 3; ++q) tv[m][q] = sel_get(acc, ti[m][0], ti[m][1], ti[m][2], VELX + q, tcid[m]);
                }
            }
            // flow components (the reconstruction loop is written out twice:
            // the species copy below runs after the viscous terms have freed
            // the transverse velocities' registers)
            if (hyp) {
#if MF_FACE_CLOOP
#pragma unroll 1
#else
#pragma unroll
#endif
                for (int c = 0; c < NCV; ++c) {
                    double uc[6], fc[6];
#pragma unroll
                    for (int q = 0; q < 6; ++q) {
                        const double *rq = rs[q];
                        uc[q] = rq[c];
                        const double vn = rq[FVX + D];
                        if (c == RHO) fc[q] = uc[q] * vn;
                        else if (c >= RHOU && c <= RHOW) {
                            const double va = rq[FVX + (c - RHOU)];
                            const double mm = rq[RHO] * va;
                            fc[q] = (c - RHOU == D) ? (mm * va + rq[FP]) : (mm * vn);
                        } else if (c == RHOE) fc[q] = (uc[q] + rq[FP]) * vn;
                        else fc[q] = uc[q] * vn;
                    }
                    double fp = N::weno_plus(fc[0], fc[1], fc[2], fc[3], fc[4]);
                    double fm = N::weno_minus(fc[1], fc[2], fc[3], fc[4], fc[5]);
                    double up = N::weno_plus(uc[0], uc[1], uc[2], uc[3], uc[4]);
                    double um = N::weno_minus(uc[1], uc[2], uc[3], uc[4], uc[5]);
                    const double h = kHalf * (fp + lam * up) + kHalf * (fm - lam * um);
                    hc[c] = c < NCV ? h : T.advect_flag[c - NCV] * h;
                }
            }
            // ---- viscous stress and work (:293-367), conduction (:373-374), species diffusion (:425-428)
            double dm0 = 0.0, dm1 = 0.0, dm2 = 0.0, de0 = 0.0, de1 = 0.0, de2 = 0.0, de3, dsp[NS];
            {
                double visc, thcond;
                {
                    // harmonic mean of (i,j,k) and (i-1,j,k) for every sweep direction (:179-180)
                    double v0, k0, v1, k1;
                    LdsCell<NC> c0c{recR, &acc, i, j, k, cR};
                    P::fluid_transport(T, c0c, v0, k0);
                    if (D == 0) {
                        LdsCell<NC> c1c{recL, &acc, iL, jL, kL, cL};
                        P::fluid_transport(T, c1c, v1, k1);
                    } else {
                        CellRef c1c{&acc, i - 1, j, k, acc.cid(i - 1, j, k)};
                        P::fluid_transport(T, c1c, v1, k1);
                    }
                    visc = N::harmonic_mean_pos(v0, v1);
                    thcond = N::harmonic_mean_pos(k0, k1);
                }
                if (hyp) {
                    double g[3][3];  // g[q][m]: d(vel q) along normal (m=0), t1 (m=1), t2 (m=2)
                    double vL[3], vR[3];
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        vL[q] = pvL[q];
                        vR[q] = pvR[q];
                        g[q][0] = (vR[q] - vL[q]) * idn;
                        double top = kHalf * (tv[0][q] + tv[1][q]);
                        double bot = kHalf * (tv[2][q] + tv[3][q]);
                        g[q][1] = (top - bot) * it1;
                        double frt = kHalf * (tv[4][q] + tv[5][q]);
                        double bck = kHalf * (tv[6][q] + tv[7][q]);
                        g[q][2] = (frt - bck) * it2;
                    }
                    // axis -> gradient slot (GET_XDIR/GET_YDIR/GET_ZDIR, globalDefines.H:49-51)
                    constexpr int sx_ = D == 0 ? 0 : (T1 == 0 ? 1 : 2);
                    constexpr int sy_ = D == 1 ? 0 : (T1 == 1 ? 1 : 2);
                    constexpr int sz_ = D == 2 ? 0 : (T1 == 2 ? 1 : 2);
                    double dudx = g[0][sx_], dudy = g[0][sy_], dudz = g[0][sz_];
                    double dvdx = g[1][sx_], dvdy = g[1][sy_], dvdz = g[1][sz_];
                    double dwdx = g[2][sx_], dwdy = g[2][sy_], dwdz = g[2][sz_];
                    double delu = dudx + dvdy + dwdz;
                    double t0, t1, t2;
                    if (D == 0) {
                        t0 = visc * (kTwo * dudx - kTwo3rd * delu);
                        t1 = visc * (dudy + dvdx);
                        t2 = visc * (dudz + dwdx);
                    } else if (D == 1) {
                        t0 = visc * (dvdx + dudy);
                        t1 = visc * (kTwo * dvdy - kTwo3rd * delu);
                        t2 = visc * (dvdz + dwdy);
                    } else {
                        t0 = visc * (dwdx + dudz);
                        t1 = visc * (dwdy + dvdz);
                        t2 = visc * (kTwo * dwdz - kTwo3rd * delu);
                    }
                    dm0 = t0;
                    dm1 = t1;
                    dm2 = t2;
                    de0 = t0 * (kHalf * (vL[0] + vR[0]));
                    de1 = t1 * (kHalf * (vL[1] + vR[1]));
                    de2 = t2 * (kHalf * (vL[2] + vR[2]));
                }
                const double dTdn = (pTR - pTL) * idn;
                de3 = thcond * dTdn;
                double dl[NS], dr[NS];
                LdsCell<NC> lc{recL, &acc, iL, jL, kL, cL};
                LdsCell<NC> rc{recR, &acc, i, j, k, cR};
                P::spec_dcoeff(T, lc, dl);
                P::spec_dcoeff(T, rc, dr);
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    double Ds = N::harmonic_mean_pos(dl[s], dr[s]);
                    dsp[s] = Ds * (recR[NCV + s] - recL[NCV + s]) * idn;
                }
            }
            const View &F = b.flux[D];
            double *Fp = F.p + F.off(i, j, k);
            const int64_t Fs = F.ns;
#pragma unroll
            for (int c = 0; c < NCV; ++c)
                Fp[(int64_t)flux_comp(c, full) * Fs] = flux_with_terms(c, hyp ? hc[c] : kZero, hyp, dm0, dm1, dm2, de0, de1, de2, de3, 0.0);
            // species components
            if (hyp) {
#if MF_FACE_CLOOP
#pragma unroll 1
#else
#pragma unroll
#endif
                for (int c = NCV; c < NC; ++c) {
                    double uc[6], fc[6];
#pragma unroll
                    for (int q = 0; q < 6; ++q) {
                        const double *rq = rs[q];
                        uc[q] = rq[c];
                        const double vn = rq[FVX + D];
                        if (c == RHO) fc[q] = uc[q] * vn;
                        else if (c >= RHOU && c <= RHOW) {
                            const double va = rq[FVX + (c - RHOU)];
                            const double mm = rq[RHO] * va;
                            fc[q] = (c - RHOU == D) ? (mm * va + rq[FP]) : (mm * vn);
                        } else if (c == RHOE) fc[q] = (uc[q] + rq[FP]) * vn;
                        else fc[q] = uc[q] * vn;
                    }
                    double fp = N::weno_plus(fc[0], fc[1], fc[2], fc[3], fc[4]);
                    double fm = N::weno_minus(fc[1], fc[2], fc[3], fc[4], fc[5]);
                    double up = N::weno_plus(uc[0], uc[1], uc[2], uc[3], uc[4]);
                    double um = N::weno_minus(uc[1], uc[2], uc[3], uc[4], uc[5]);
                    const double h = kHalf * (fp + lam * up) + kHalf * (fm - lam * um);
                    hc[c] = c < NCV ? h : T.advect_flag[c - NCV] * h;
                }
            }
#pragma unroll
            for (int c = NCV; c < NC; ++c)
                Fp[(int64_t)flux_comp(c, full) * Fs] =
                    flux_with_terms(c, hyp ? hc[c] : kZero, hyp, dm0, dm1, dm2, de0, de1, de2, de3, pick(dsp, c - NCV));
            if (full) {
#pragma unroll
                for (int n = NCV; n < FLO; ++n) F(i, j, k, n) = kZero;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Post-step work on one cell (after the stage-3 combine):
// Advance_chemistry_implicit (SSPRK3 of AMReX's TimeIntegrator, rk.type=3),
// update_primitive_vars, clip_neg_speciesconc, residual-norm and dt partials.
// ---------------------------------------------------------------------------
// stage state of the chemistry integrator as seen by the plugin: the species
// of the stage, every other component of the post-flow state (unchanged by
// the integrator: their right-hand side is zero)
template <int NS>
struct ChemCell {
    const double *S;
    const double *sp;
    MFD double operator()(int n) const { return n >= FLO ? sp[n - FLO] : S[n]; }
};

template <class P>
MFD void chemistry_rk3(const Tab &T, double *S, double dt) {
    constexpr int NS = P::NS;
    if (!P::kHasChemistry) return;
    double S0[NS], F[3][NS], st[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) S0[s] = S[FLO + s];
    const double a10 = 1.0, a20 = 0.25, a21 = 0.25;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double v = S0[s];
            if (k == 1) v += (dt * a10) * F[0][s];
            if (k == 2) { v += (dt * a20) * F[0][s]; v += (dt * a21) * F[1][s]; }
            st[s] = v;
        }
        double src[NS];
        P::spec_source(T, ChemCell<NS>{S, st}, src);
#pragma unroll
        for (int s = 0; s < NS; ++s) F[k][s] = 0.0 + src[s];
    }
    const double w0 = 1. / 6., w1 = 1. / 6., w2 = 2. / 3.;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double v = S0[s];
        v += (dt * w0) * F[0][s];
        v += (dt * w1) * F[1][s];
        v += (dt * w2) * F[2][s];
        S[FLO + s] = v;
    }
}

// a x + b y rounded as written: the empty asm makes both products opaque, so
// the device compiler cannot fuse one of them into the add (-ffp-contract=fast
// ignores contraction pragmas)
MFD double lincomb_rn(double a, double x, double b, double y) {
    double p = a * x, q = b * y;
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(p), "+v"(q));
#endif
    return p + q;
}

template <class P>
MFD void dt_partial(const Tab &T, const double *S, int do_ns, double &ws, double &maxd) {
    constexpr int NS = P::NS, NV = Num<P>::NV;
    ws = kZero;
    if (S[VFRAC] > kZero) {
        double fs = sqrt(S[VELX] * S[VELX] + S[VELY] * S[VELY] + S[VELZ] * S[VELZ]);
        double gama = P::get_gama_from_rpc(T, S[DENS], S[PRES], S + FLO);
        double ss = sqrt(gama * S[PRES] / S[DENS]);
        ws = fs + (double)do_ns * ss;
    }
    double d[NS];
    P::spec_dcoeff(T, CellArr<NV>{S}, d);
    maxd = d[0];
#pragma unroll
    for (int s = 0; s < NS; ++s) if (d[s] > maxd) maxd = d[s];
    ws = fabs(ws);
    maxd = fabs(maxd);
}

// block reduction of NR sums and 2 maxima into partials[tile]
template <int NR>
MFD void block_reduce_store(double *sums, double m0, double m1, double *out) {
    __shared__ double red[4][NR + 2];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        double v = sums[q];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        sums[q] = v;
    }
    for (int o = 32; o > 0; o >>= 1) {
        m0 = fmax(m0, __shfl_xor(m0, o, 64));
        m1 = fmax(m1, __shfl_xor(m1, o, 64));
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NR; ++q) red[wid][q] = sums[q];
        red[wid][NR] = m0;
        red[wid][NR + 1] = m1;
    }
    __syncthreads();
    if (threadIdx.x < NR + 2) {
        const int nw = (blockDim.x + 63) >> 6;
        double v = red[0][threadIdx.x];
        for (int q = 1; q < nw; ++q) {
            double x = red[q][threadIdx.x];
            v = (threadIdx.x < NR) ? v + x : fmax(v, x);
        }
        out[threadIdx.x] = v;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// update_residual (div_3d.H:13-76) and the RK3 stage arithmetic
// (mflo_evolve.cpp:283, 296-298, 309-311) per cell of a tile. Each cell forms
// its base pointers once (fluxes of the three directions, stage source, U^n,
// destination), so the loads of all components are independent address
// offsets from them.
// ---------------------------------------------------------------------------
template <class P>
__global__ __launch_bounds__(256) void k_residual(ResArgs r, Work w) {
    using N = Num<P>;
    constexpr int NS = P::NS, NV = N::NV;
    const StageArgs &a = r.st;
    if (step_skipped(a.ts)) return;
    const Tab &T = *a.tab;
    const double dt = stage_dt(a);
    const int nvn = (a.spec_in_solid == 1) ? FLO : FLO + NS;
    const double dx0 = a.dx[0], dx1 = a.dx[1], dx2 = a.dx[2];
    for (TileIter it(w); it.ok(); it.next()) {
        const int t = it.t;
        int lt;
        const BoxDesc &b = tile_box(w, t, lt);
        int c0[3], c1[3];
        tile_range(b, w, lt, c0, c1);
        const int ex = c1[0] - c0[0] + 1, ey = c1[1] - c0[1] + 1, ez = c1[2] - c0[2] + 1;
        const int nc = ex * ey * ez;
        Acc acc = make_acc(a, b);
        const View &fx = b.flux[0], &fy = b.flux[1], &fz = b.flux[2];
        for (int f = threadIdx.x; f < nc; f += blockDim.x) {
            const int i = c0[0] + f % ex;
            const int rr = f / ex;
            const int j = c0[1] + rr % ey;
            const int k = c0[2] + rr / ey;
            const int cid = acc.cid(i, j, k);
            double src[FLO];
            if (P::kHasFluidSource) {
                CellRef cr{&acc, i, j, k, cid};
                P::fluid_source(T, cr, src);
            } else {
#pragma unroll
                for (int n = 0; n < FLO; ++n) src[n] = kZero;
            }
            const double *px = fx.p + fx.off(i, j, k), *py = fy.p + fy.off(i, j, k), *pz = fz.p + fz.off(i, j, k);
            const int64_t xs = fx.ns, ys = fy.ns, zs = fz.ns, yj = fy.js, zk = fz.ks;
            const double *ps = a.src.p + a.src.off(i, j, k);
            const int64_t sn = a.src.ns;
            const double *pc = cid >= 0 ? a.cutbuf + cid : ps;   // the values update_cutcells produced
            const int64_t cn = cid >= 0 ? (int64_t)a.ncut : sn;
            const double vfc = pc[(int64_t)VFRAC * cn];
            const double *po = nullptr;
            int64_t on = 0;
            if (r.combine && a.stage > 1) {
                po = r.old.p + r.old.off(i, j, k);
                on = r.old.ns;
            }
            double *pd;
            int64_t dn;
            if (r.combine) {
                pd = r.dst.p + r.dst.off(i, j, k);
                dn = r.dst.ns;
            } else {
                pd = r.dsdt.p + r.dsdt.off(i, j, k);
                dn = r.dsdt.ns;
            }
            // all loads first, then the stores: a store inside the component loop
            // would keep the next component's loads behind it (the arrays may alias)
            double Sout[NV];
#pragma unroll
            for (int n = 0; n < NV; ++n) {
                double d;
                if (n < NCV || n >= FLO) {
                    const int c = flux_comp(n < NCV ? n : n - (FLO - NCV), a.flux_full);
                    const double *qx = px + (int64_t)c * xs, *qy = py + (int64_t)c * ys, *qz = pz + (int64_t)c * zs;
                    d = (qx[0] - qx[1]) / dx0 + (qy[0] - qy[yj]) / dx1 + (qz[0] - qz[zk]) / dx2 + (n < FLO ? src[n] : kZero);
                } else {
                    d = kZero + src[n];
                }
                if (n != RHOE) {
                    if (n < nvn && vfc < kOne) d = kZero;
                } else {
                    if (!a.conj_ht) { if (vfc < kOne) d = kZero; }
                    else { if (vfc > kZero && vfc < kOne) d = kZero; }
                }
                if (d != d) raise_err(a.err, MF_ERR_NAN_DSDT);
                double S;
                if (!r.combine) {
                    S = d;
                } else if (a.stage == 1) {
                    S = kOne * pc[(int64_t)n * cn] + dt * d;
                } else if (a.stage == 2) {
                    double v = ps[(int64_t)n * sn];
                    v += dt * d;
                    v += 3.0 * po[(int64_t)n * on];
                    S = v * kFourth;
                } else {
                    const double o = po[(int64_t)n * on];
                    double v = ps[(int64_t)n * sn];
                    v += dt * d;
                    v = o + kTwo * v;
                    S = v * kThird;
                }
                Sout[n] = S;
            }
#pragma unroll
            for (int n = 0; n < NV; ++n) pd[(int64_t)n * dn] = Sout[n];
        }
    }
}


// Post-step pass over the level after the stage-3 combine: S3 is read from
// st.src (A with chemistry, B without), U^n from `old` (needed by the residual
// only without chemistry), the final state goes to `dst` (A):
// Advance_chemistry_implicit, update_primitive_vars, clip_neg_speciesconc,
// compute_residual_norms and EstTimeStep partials, point by point.
template <class P, bool CHEM>
__global__ __launch_bounds__(256) void k_post(ResArgs r, Work w) {
    using N = Num<P>;
    constexpr int NS = P::NS, NV = N::NV;
    const StageArgs &a = r.st;
    if (step_skipped(a.ts)) return;
    const Tab &T = *a.tab;
    const double dt = stage_dt(a);
    // partial sums of the block over all its tiles: one slot per block
    double sums[NV];
#pragma unroll
    for (int n = 0; n < NV; ++n) sums[n] = 0.0;
    double mws = 0.0, mD = 0.0;
    for (TileIter it(w); it.ok(); it.next()) {
        const int t = it.t;
        int lt;
        const BoxDesc &b = tile_box(w, t, lt);
        int c0[3], c1[3];
        tile_range(b, w, lt, c0, c1);
        const int ex = c1[0] - c0[0] + 1, ey = c1[1] - c0[1] + 1, ez = c1[2] - c0[2] + 1;
        const int nc = ex * ey * ez;
        for (int f = threadIdx.x; f < nc; f += blockDim.x) {
            const int i = c0[0] + f % ex;
            const int rr = f / ex;
            const int j = c0[1] + rr % ey;
            const int k = c0[2] + rr / ey;
            const double *srcp = a.src.p + a.src.off(i, j, k);
            double *sp = r.dst.p + r.dst.off(i, j, k);
            const int64_t sn = r.dst.ns;
            double S[NV];
#pragma unroll
            for (int n = 0; n < NV; ++n) S[n] = srcp[(int64_t)n * a.src.ns];
            // S_old of compute_residual_norms: the post-flow state with chemistry
            // (second swap, mflo_evolve.cpp:473), else U^n
            double So_p[NV - NCV];  // components DENS.. of S_old
            if (CHEM) {
#pragma unroll
                for (int n = DENS; n < NV; ++n) So_p[n - DENS] = S[n];
                chemistry_rk3<P>(T, S, dt);
            }
            // update_primitive_vars (mflo_evolve.cpp:129-160)
            if (S[VFRAC] == kOne) {
                double u[N::NC], p[NCV];
#pragma unroll
                for (int c = 0; c < NCV; ++c) u[c] = S[c];
#pragma unroll
                for (int s = 0; s < NS; ++s) u[NCV + s] = S[FLO + s];
                N::cons_to_prim(T, u, p);
#pragma unroll
                for (int c = 0; c < NCV; ++c) S[DENS + c] = p[c];
                S[TEMP] = P::get_t_from_rpc(T, S[DENS], S[PRES], u + NCV);
            } else if (a.conj_ht && S[VFRAC] == kZero) {
                S[TEMP] = P::get_solid_t_from_rhoe(T, S[RHOE]);
            }
            if (r.do_ns && S[PRES] < 0) raise_err(a.err, MF_ERR_NEG_PRESSURE);
            if (r.clip) {
#pragma unroll
                for (int s = 0; s < NS; ++s) if (S[FLO + s] < 0.0) S[FLO + s] = 0.0;
            }
            if (r.track_res) {
                // LinComb(a, S_new, b, S_old) rounded as written (no fused
                // multiply-add), so that an unchanged component gives exactly 0
                const double ia = 1.0 / dt, ib = -1.0 / dt;
                if (CHEM) {
                    // conserved components and vfrac: S_new == S_old, exactly zero
                    const double vfo = S[VFRAC];
#pragma unroll
                    for (int n = DENS; n < NV; ++n) {
                        if (n == VFRAC) continue;
                        double df = lincomb_rn(ia, S[n], ib, So_p[n - DENS]);
                        if (n < VFRAC) df *= vfo;
                        sums[n] += df * df;
                    }
                } else {
                    const View &o = r.old;
                    const double *op = o.p + o.off(i, j, k);
                    const double vfo = op[(int64_t)VFRAC * o.ns];
#pragma unroll
                    for (int n = 0; n < NV; ++n) {
                        double df = lincomb_rn(ia, S[n], ib, op[(int64_t)n * o.ns]);
                        if (n < VFRAC) df *= vfo;
                        sums[n] += df * df;
                    }
                }
            }
            double ws, md;
            dt_partial<P>(T, S, r.do_ns, ws, md);
            mws = fmax(mws, ws);
            mD = fmax(mD, md);
#pragma unroll
            for (int n = 0; n < NV; ++n) sp[(int64_t)n * sn] = S[n];
        }
    }
    block_reduce_store<NV>(sums, mws, mD, r.partials + (size_t)(w.t0 + blockIdx.x) * (NV + 2));
}

// mf_box_post: the per-cell work of k_post without the residual and dt
// partials, on one box view in place (the per-box form of
// Advance_chemistry_implicit, update_primitive_vars, clip_neg_speciesconc)
template <class P>
__global__ __launch_bounds__(256) void k_box_post(BoxPostArgs a) {
    using N = Num<P>;
    constexpr int NS = P::NS, NV = N::NV;
    const Tab &T = *a.tab;
    const int64_t nc = (int64_t)a.e[0] * a.e[1] * a.e[2];
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nc; q += (int64_t)gridDim.x * blockDim.x) {
        const int i = a.lo[0] + (int)(q % a.e[0]);
        const int64_t rr = q / a.e[0];
        const int j = a.lo[1] + (int)(rr % a.e[1]);
        const int k = a.lo[2] + (int)(rr / a.e[1]);
        double S[NV];
#pragma unroll
        for (int n = 0; n < NV; ++n) S[n] = a.s(i, j, k, n);
        if (a.chem) chemistry_rk3<P>(T, S, a.dt);
        if (a.prims) {
            if (S[VFRAC] == kOne) {
                double u[N::NC], p[NCV];
#pragma unroll
                for (int c = 0; c < NCV; ++c) u[c] = S[c];
#pragma unroll
                for (int s = 0; s < NS; ++s) u[NCV + s] = S[FLO + s];
                N::cons_to_prim(T, u, p);
#pragma unroll
                for (int c = 0; c < NCV; ++c) S[DENS + c] = p[c];
                S[TEMP] = P::get_t_from_rpc(T, S[DENS], S[PRES], u + NCV);
            } else if (a.conj_ht && S[VFRAC] == kZero) {
                S[TEMP] = P::get_solid_t_from_rhoe(T, S[RHOE]);
            }
            if (a.do_ns && S[PRES] < 0) raise_err(a.err, MF_ERR_NEG_PRESSURE);
        }
        if (a.clip) {
#pragma unroll
            for (int s = 0; s < NS; ++s) if (S[FLO + s] < 0.0) S[FLO + s] = 0.0;
        }
#pragma unroll
        for (int n = 0; n < NV; ++n) a.s(i, j, k, n) = S[n];
    }
}

// Advance_chemistry_implicit alone over the valid cells of a level's tile
// list (a level of a hierarchy, whose primitive update waits for the end of
// the coarse step): st.src -> dst (may alias), dt of st.ts
template <class P>
__global__ __launch_bounds__(256) void k_chem_tiles(ResArgs r, Work w) {
    constexpr int NV = Num<P>::NV;
    const StageArgs &a = r.st;
    if (step_skipped(a.ts)) return;
    const Tab &T = *a.tab;
    const double dt = stage_dt(a);
    for (TileIter it(w); it.ok(); it.next()) {
        int lt;
        const BoxDesc &b = tile_box(w, it.t, lt);
        int c0[3], c1[3];
        tile_range(b, w, lt, c0, c1);
        const int ex = c1[0] - c0[0] + 1, ey = c1[1] - c0[1] + 1, ez = c1[2] - c0[2] + 1;
        const int nc = ex * ey * ez;
        for (int f = threadIdx.x; f < nc; f += blockDim.x) {
            const int i = c0[0] + f % ex;
            const int rr = f / ex;
            const int j = c0[1] + rr % ey;
            const int k = c0[2] + rr / ey;
            const double *sp = a.src.p + a.src.off(i, j, k);
            double *dp = r.dst.p + r.dst.off(i, j, k);
            double S[NV];
#pragma unroll
            for (int n = 0; n < NV; ++n) S[n] = sp[(int64_t)n * a.src.ns];
            chemistry_rk3<P>(T, S, dt);
#pragma unroll
            for (int n = 0; n < NV; ++n) dp[(int64_t)n * r.dst.ns] = S[n];
        }
    }
}

// EstTimeStep partials on the current state (first step of a run)
template <class P>
__global__ __launch_bounds__(256) void k_est(ResArgs r, Work w) {
    constexpr int NV = Num<P>::NV;
    const StageArgs &a = r.st;
    const Tab &T = *a.tab;
    // partial sums of the block over all its tiles: one slot per block
    double sums[NV];
#pragma unroll
    for (int n = 0; n < NV; ++n) sums[n] = 0.0;
    double mws = 0.0, mD = 0.0;
    for (TileIter it(w); it.ok(); it.next()) {
        const int t = it.t;
        int lt;
        const BoxDesc &b = tile_box(w, t, lt);
        int c0[3], c1[3];
        tile_range(b, w, lt, c0, c1);
        const int ex = c1[0] - c0[0] + 1, ey = c1[1] - c0[1] + 1, ez = c1[2] - c0[2] + 1;
        for (int f = threadIdx.x; f < ex * ey * ez; f += blockDim.x) {
            const int i = c0[0] + f % ex, rr = f / ex, j = c0[1] + rr % ey, k = c0[2] + rr / ey;
            double S[NV];
#pragma unroll
            for (int n = 0; n < NV; ++n) S[n] = a.src(i, j, k, n);
            double ws, md;
            dt_partial<P>(T, S, r.do_ns, ws, md);
            mws = fmax(mws, ws);
            mD = fmax(mD, md);
        }
    }
    block_reduce_store<NV>(sums, mws, mD, r.partials + (size_t)(w.t0 + blockIdx.x) * (NV + 2));
}

// ---------------------------------------------------------------------------
// Ghost cells of the padded level array (FillPatchSingleLevel): phase 0 fills
// cells outside only in periodic directions from their periodic images; phase
// p >= 1 fills cells outside in p non-periodic directions with the filcc rules
// (foextrap, reflect_even/odd for walls: Source/mflo.cpp:70-143) in x,y,z order,
// then the ext_dir functor of Source/mflo_bcfill.H:16-128 in x,y,z order.
// ---------------------------------------------------------------------------
// the rules of one ghost cell iv of a.s for the given phase (cells outside the
// domain in phase non-periodic directions; phase 0: only in periodic ones)
template <class P>
MFD void fill_cell_rules(const FillArgs &a, const int *iv, int phase, const Tab &T, double time) {
    constexpr int NV = Num<P>::NV;
    int out_np = 0, out_p = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        bool o = iv[d] < 0 || iv[d] >= a.n[d];
        if (o) { if (a.periodic[d]) ++out_p; else ++out_np; }
    }
    if (phase == 0) {
        if (out_np != 0 || out_p == 0) return;
        int s[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            s[d] = iv[d];
            if (a.periodic[d]) { if (s[d] < 0) s[d] += a.n[d]; else if (s[d] >= a.n[d]) s[d] -= a.n[d]; }
        }
        for (int c = 0; c < NV; ++c) a.s(iv[0], iv[1], iv[2], c) = a.s(s[0], s[1], s[2], c);
        return;
    }
    if (out_np != phase) return;
    for (int d = 0; d < 3; ++d) {
        if (a.periodic[d]) continue;
        int side = iv[d] < 0 ? 0 : (iv[d] >= a.n[d] ? 1 : -1);
        if (side < 0) continue;
        int bc = side == 0 ? a.lo_bc[d] : a.hi_bc[d];
        if (bc == MF_BC_EXT_DIR) continue;
        int s[3] = {iv[0], iv[1], iv[2]};
        bool refl = (bc == MF_BC_WALL);
        if (!refl) s[d] = side == 0 ? 0 : a.n[d] - 1;
        else s[d] = side == 0 ? (-iv[d] - 1) : (2 * (a.n[d] - 1) - iv[d] + 1);
        for (int c = 0; c < NV; ++c) {
            double v = a.s(s[0], s[1], s[2], c);
            bool odd = refl && (c == RHOU || c == RHOV || c == RHOW || c == VELX || c == VELY || c == VELZ);
            a.s(iv[0], iv[1], iv[2], c) = odd ? -v : v;
        }
    }
    double x[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = a.prob_lo[d] + (iv[d] + 0.5) * a.dx[d];
    for (int d = 0; d < 3; ++d) {
        if (a.periodic[d]) continue;
        int sgn = 0, bnd = 0;
        if (a.lo_bc[d] == MF_BC_EXT_DIR && iv[d] < 0) { sgn = 1; bnd = 0; }
        else if (a.hi_bc[d] == MF_BC_EXT_DIR && iv[d] >= a.n[d]) { sgn = -1; bnd = a.n[d] - 1; }
        if (!sgn) continue;
        double s_int[NV], s_ext[NV];
        int loc[3] = {iv[0], iv[1], iv[2]};
        loc[d] = bnd;
        for (int c = 0; c < NV; ++c) { s_int[c] = a.s(loc[0], loc[1], loc[2], c); s_ext[c] = 0.0; }
        P::externalbc(T, x, s_int, s_ext, d, sgn, time);
        for (int c = 0; c < NV; ++c) a.s(iv[0], iv[1], iv[2], c) = s_ext[c];
    }
}

template <class P>
__global__ __launch_bounds__(256) void k_fill(FillArgs a, int phase) {
    if (step_skipped(a.ts)) return;
    const Tab &T = *a.tab;
    const double time = !a.ts ? a.time_host : (a.stage == 1 ? a.ts->t : a.ts->t + a.ts->dt);
    const int G = MF_NGROW;
    const int nx = a.n[0], ny = a.n[1], nz = a.n[2];
    const int PX = nx + 2 * G, PY = ny + 2 * G;
    const int64_t nzs = (int64_t)2 * G * PX * PY;
    const int64_t nys = (int64_t)nz * 2 * G * PX;
    const int64_t nxs = (int64_t)nz * ny * 2 * G;
    const int64_t total = nzs + nys + nxs;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        int iv[3];
        if (q < nzs) {
            int64_t r = q;
            iv[0] = (int)(r % PX) - G; r /= PX;
            iv[1] = (int)(r % PY) - G; r /= PY;
            int g = (int)r;
            iv[2] = g < G ? g - G : nz + (g - G);
        } else if (q < nzs + nys) {
            int64_t r = q - nzs;
            iv[0] = (int)(r % PX) - G; r /= PX;
            int g = (int)(r % (2 * G)); r /= (2 * G);
            iv[1] = g < G ? g - G : ny + (g - G);
            iv[2] = (int)r;
        } else {
            int64_t r = q - nzs - nys;
            int g = (int)(r % (2 * G)); r /= (2 * G);
            iv[0] = g < G ? g - G : nx + (g - G);
            iv[1] = (int)(r % ny); r /= ny;
            iv[2] = (int)r;
        }
        fill_cell_rules<P>(a, iv, phase, T, time);
    }
}

// The physical-boundary rules (phases >= 1) over the cells of the region
// [rlo, rhi] of a.s that lie outside the domain: the ghost region of a box
// whose other cells are already filled (FillPatchTwoLevels' physical boundary
// fill of a fine box). Each rule reads only cells of the region: the source
// planes 0..2 and n-3..n-1 lie inside any box grown by 3 that reaches beyond
// the domain.
template <class P>
__global__ __launch_bounds__(256) void k_fill_region(FillArgs a, int phase, int3 rlo, int3 rhi) {
    const Tab &T = *a.tab;
    const double time = a.time_host;
    const int ex = rhi.x - rlo.x + 1, ey = rhi.y - rlo.y + 1, ez = rhi.z - rlo.z + 1;
    const int64_t total = (int64_t)ex * ey * ez;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        int iv[3];
        int64_t r = q;
        iv[0] = rlo.x + (int)(r % ex); r /= ex;
        iv[1] = rlo.y + (int)(r % ey); r /= ey;
        iv[2] = rlo.z + (int)r;
        fill_cell_rules<P>(a, iv, phase, T, time);
    }
}

// The rules of one phase over a list of cells of a.s ([3][n]: i, j, k), the
// ghost cells of a partly covered level that lie outside the domain
// (FillPatchTwoLevels' physical boundary fill of the fine boxes); the sources
// of a phase are cells of earlier phases or interior cells, so the cells of a
// phase are independent.
template <class P>
__global__ __launch_bounds__(256) void k_fill_list(FillArgs a, int phase, const int *list, int n) {
    if (step_skipped(a.ts)) return;
    const Tab &T = *a.tab;
    const double time = !a.ts ? a.time_host : (a.stage == 1 ? a.ts->t : a.ts->t + a.ts->dt);
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const int iv[3] = {list[q], list[n + q], list[2 * n + q]};
        fill_cell_rules<P>(a, iv, phase, T, time);
    }
}

// ---------------------------------------------------------------------------
// update_cutcells (ib_utils.H:271-436), one thread per cut cell of the list;
// the result goes to cutbuf, the state array is not modified.
// ---------------------------------------------------------------------------
MFD double ib_basis1d(int l, double z) {
    if (l == -1) return kHalf * z * (z - kOne);
    if (l == 0) return (kOne - z * z);
    if (l == 1) return kHalf * z * (z + kOne);
    return 0.0;
}
MFD double ib_basisder1d(int l, double z) {
    if (l == -1) return kHalf * (kTwo * z - kOne);
    if (l == 0) return -kTwo * z;
    if (l == 1) return kHalf * (kTwo * z + kOne);
    return 0.0;
}

template <class P>
MFD void cutcell_values(const Tab &T, const CutArgs &a, int i, int j, int k, double *out) {
    using N = Num<P>;
    constexpr int NS = P::NS, NV = N::NV;
    const View &s = a.s;
#pragma unroll
    for (int n = 0; n < NV; ++n) out[n] = s(i, j, k, n);
    const double vf = out[VFRAC];
    if (!(vf > kZero && vf < kOne)) return;
    const double *dx = a.dx;
    // level-set gradient from the 27-point quadratic basis at the centre:
    // only the (+-1,0,0)-type terms are non-zero, in loop order n,m,l
    double lsg[3] = {0.0, 0.0, 0.0};
    const double zero = 0.0;
    for (int n = -1; n <= 1; n++)
        for (int m = -1; m <= 1; m++)
            for (int l = -1; l <= 1; l++) {
                double v = s(i + l, j + m, k + n, VFRAC);
                double bx = ib_basisder1d(l, zero / dx[0]) / dx[0] * ib_basis1d(m, zero / dx[1]) * ib_basis1d(n, zero / dx[2]);
                double by = ib_basis1d(l, zero / dx[0]) * (ib_basisder1d(m, zero / dx[1]) / dx[1]) * ib_basis1d(n, zero / dx[2]);
                double bz = ib_basis1d(l, zero / dx[0]) * ib_basis1d(m, zero / dx[1]) * (ib_basisder1d(n, zero / dx[2]) / dx[2]);
                lsg[0] += v * bx;
                lsg[1] += v * by;
                lsg[2] += v * bz;
            }
    double mag = sqrt(lsg[0] * lsg[0] + lsg[1] * lsg[1] + lsg[2] * lsg[2]);
    double dist = kZero, xp[3] = {kZero, kZero, kZero};
    if (mag > kTiny) {
        dist = (vf - kHalf) / mag;
#pragma unroll
        for (int d = 0; d < 3; d++) xp[d] = -dist * (lsg[d] / mag);
    }
    const double dxmin = fmin(dx[0], fmin(dx[1], dx[2]));
    const double dnd[3] = {dx[0] / dxmin, dx[1] / dxmin, dx[2] / dxmin};
    const double xnd[3] = {xp[0] / dxmin, xp[1] / dxmin, xp[2] / dxmin};
    const double cutoff = dxmin * 1e-5;
    // Dirichlet (u = 0) IDW of the velocities, :98-152
    if (fabs(dist) < cutoff) {
        out[VELX] = 0.0; out[VELY] = 0.0; out[VELZ] = 0.0;
    } else {
        double su[3] = {0.0, 0.0, 0.0}, sw = 0.0;
        for (int n = -2; n <= 2; n++)
            for (int m = -2; m <= 2; m++)
                for (int l = -2; l <= 2; l++) {
                    if (l == 0 && m == 0 && n == 0) continue;
                    if (s(i + l, j + m, k + n, VFRAC) == 1.0) {
                        double a0 = l * dnd[0], a1 = m * dnd[1], a2 = n * dnd[2];
                        double w2 = a0 * a0 + a1 * a1 + a2 * a2 + kTiny2;
                        su[0] += s(i + l, j + m, k + n, VELX) / w2;
                        su[1] += s(i + l, j + m, k + n, VELY) / w2;
                        su[2] += s(i + l, j + m, k + n, VELZ) / w2;
                        sw += 1.0 / w2;
                    }
                }
        double dn = fabs(dist) / dxmin;
        double wd = 1.0 / (dn * dn);
#pragma unroll
        for (int q3 = 0; q3 < 3; ++q3) out[VELX + q3] = (su[q3] + 0.0 / (dn * dn)) / (sw + wd);
    }
    // homogeneous Neumann IDW of p, T, rho (and species without spec_in_solid), :154-198
    {
        const bool doT = !a.conj_ht, doS = !a.spec_in_solid;
        double sp_ = 0.0, sT = 0.0, sR = 0.0, sw = 0.0;
        double sS[NS];
#pragma unroll
        for (int q = 0; q < NS; ++q) sS[q] = 0.0;
        for (int n = -2; n <= 2; n++)
            for (int m = -2; m <= 2; m++)
                for (int l = -2; l <= 2; l++) {
                    if (s(i + l, j + m, k + n, VFRAC) == 1.0) {
                        double a0 = l * dnd[0] - xnd[0], a1 = m * dnd[1] - xnd[1], a2 = n * dnd[2] - xnd[2];
                        double w2 = a0 * a0 + a1 * a1 + a2 * a2 + kTiny2;
                        sp_ += s(i + l, j + m, k + n, PRES) / w2;
                        if (doT) sT += s(i + l, j + m, k + n, TEMP) / w2;
                        sR += s(i + l, j + m, k + n, DENS) / w2;
                        if (doS) {
#pragma unroll
                            for (int q = 0; q < NS; ++q) sS[q] += s(i + l, j + m, k + n, FLO + q) / w2;
                        }
                        sw += 1.0 / w2;
                    }
                }
        const bool any = sw > 0.0;
        out[PRES] = any ? sp_ / sw : s(i, j, k, PRES);
        if (doT) out[TEMP] = any ? sT / sw : s(i, j, k, TEMP);
        out[DENS] = any ? sR / sw : s(i, j, k, DENS);
        if (doS) {
#pragma unroll
            for (int q = 0; q < NS; ++q) out[FLO + q] = any ? sS[q] / sw : s(i, j, k, FLO + q);
        }
    }
    if (a.conj_ht) {
        double st = 0.0;
        int cnt = 0;
        for (int n = -1; n <= 1; n++)
            for (int m = -1; m <= 1; m++)
                for (int l = -1; l <= 1; l++)
                    if (!(l == 0 && m == 0 && n == 0)) { st += s(i + l, j + m, k + n, TEMP); cnt++; }
        out[TEMP] = st / (double)cnt;
    }
    // prim_to_cons of the updated primitives (:421-433)
    double u[N::NC], pv[NCV];
#pragma unroll
    for (int sp = 0; sp < NS; ++sp) u[NCV + sp] = out[FLO + sp];
    pv[0] = out[DENS]; pv[1] = out[VELX]; pv[2] = out[VELY]; pv[3] = out[VELZ]; pv[4] = out[PRES];
    N::prim_to_cons(T, u, pv);
#pragma unroll
    for (int c = 0; c < NCV; ++c) out[c] = u[c];
}

// block -> chunk of the cut list such that each XCD (block b runs on XCD b % 8
// when the grid is a multiple of 8) walks a contiguous eighth of the list,
// whose IDW neighbourhoods then share that XCD's L2
MFD int xcd_chunk(int b, int G) {
    if (G & 7) return b;
    return (b & 7) * (G >> 3) + (b >> 3);
}

// static geometry of one cut cell (the vfrac-only part of cutcell_values):
// g[0..2] the wall point / dxmin, g[3] the signed distance, g[4] the
// denominator of the velocity IDW (fluid-neighbour weights plus the wall
// term), g[5] the sum of the scalar IDW weights; mb[4] the fluid-neighbour
// mask, bit (n+2)*25 + (m+2)*5 + (l+2) for neighbour (l, m, n)
MFD void cut_geom_cell(const CutArgs &a, int i, int j, int k, double *g, unsigned *mb) {
    const View &s = a.s;
    const double *dx = a.dx;
    const double vf = s(i, j, k, VFRAC);
    double lsg[3] = {0.0, 0.0, 0.0};
    const double zero = 0.0;
    for (int n = -1; n <= 1; n++)
        for (int m = -1; m <= 1; m++)
            for (int l = -1; l <= 1; l++) {
                double v = s(i + l, j + m, k + n, VFRAC);
                double bx = ib_basisder1d(l, zero / dx[0]) / dx[0] * ib_basis1d(m, zero / dx[1]) * ib_basis1d(n, zero / dx[2]);
                double by = ib_basis1d(l, zero / dx[0]) * (ib_basisder1d(m, zero / dx[1]) / dx[1]) * ib_basis1d(n, zero / dx[2]);
                double bz = ib_basis1d(l, zero / dx[0]) * ib_basis1d(m, zero / dx[1]) * (ib_basisder1d(n, zero / dx[2]) / dx[2]);
                lsg[0] += v * bx;
                lsg[1] += v * by;
                lsg[2] += v * bz;
            }
    double mag = sqrt(lsg[0] * lsg[0] + lsg[1] * lsg[1] + lsg[2] * lsg[2]);
    double dist = kZero, xp[3] = {kZero, kZero, kZero};
    if (mag > kTiny) {
        dist = (vf - kHalf) / mag;
#pragma unroll
        for (int d = 0; d < 3; d++) xp[d] = -dist * (lsg[d] / mag);
    }
    const double dxmin = fmin(dx[0], fmin(dx[1], dx[2]));
    const double dnd[3] = {dx[0] / dxmin, dx[1] / dxmin, dx[2] / dxmin};
#pragma unroll
    for (int d = 0; d < 3; ++d) g[d] = xp[d] / dxmin;
    g[3] = dist;
    mb[0] = mb[1] = mb[2] = mb[3] = 0u;
    double swv = 0.0, sws = 0.0;
    for (int nb = 0; nb < 125; ++nb) {
        const int l = nb % 5 - 2, m = (nb / 5) % 5 - 2, n = nb / 25 - 2;
        if (s(i + l, j + m, k + n, VFRAC) == 1.0) {
            mb[nb >> 5] |= 1u << (nb & 31);
            const double a0 = l * dnd[0], a1 = m * dnd[1], a2 = n * dnd[2];
            swv += 1.0 / (a0 * a0 + a1 * a1 + a2 * a2 + kTiny2);
            const double b0 = a0 - g[0], b1 = a1 - g[1], b2 = a2 - g[2];
            sws += 1.0 / (b0 * b0 + b1 * b1 + b2 * b2 + kTiny2);
        }
    }
    const double dn = fabs(dist) / dxmin;
    g[4] = swv + 1.0 / (dn * dn);
    g[5] = sws;
}

// static geometry of the listed cut cells: geom[6][ncut], mask[4][ncut]
template <class P>
__global__ __launch_bounds__(128) void k_cut_geom(CutArgs a, double *geom, unsigned *mask) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.ncut) return;
    double g[6];
    unsigned mb[4];
    cut_geom_cell(a, a.list[q], a.list[a.ncut + q], a.list[2 * a.ncut + q], g, mb);
#pragma unroll
    for (int d = 0; d < 6; ++d) geom[(int64_t)d * a.ncut + q] = g[d];
#pragma unroll
    for (int w = 0; w < 4; ++w) mask[(int64_t)w * a.ncut + q] = mb[w];
}

// the static geometry of the cut cells sub[0..nsub) of the list
template <class P>
__global__ __launch_bounds__(128) void k_cut_geom_sub(CutArgs a, double *geom, unsigned *mask, const int *sub, int nsub) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsub) return;
    const int q = sub[t];
    double g[6];
    unsigned mb[4];
    cut_geom_cell(a, a.list[q], a.list[a.ncut + q], a.list[2 * a.ncut + q], g, mb);
#pragma unroll
    for (int d = 0; d < 6; ++d) geom[(int64_t)d * a.ncut + q] = g[d];
#pragma unroll
    for (int w = 0; w < 4; ++w) mask[(int64_t)w * a.ncut + q] = mb[w];
}

// update_cutcells of the level's cut-cell list (ib_utils.H:271-436), one thread
// per cut cell, with the static geometry of k_cut_geom: the IDW sums run over
// the fluid neighbours of the mask in the reference's loop order.
template <class P>
__global__ __launch_bounds__(128) void k_cutcells(CutArgs a) {
    using N = Num<P>;
    constexpr int NS = P::NS, NV = N::NV;
    if (step_skipped(a.ts)) return;
    const int q = xcd_chunk(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (q >= a.ncut) return;
    const Tab &T = *a.tab;
    const View &s = a.s;
    const int i = a.list[q], j = a.list[a.ncut + q], k = a.list[2 * a.ncut + q];
    double out[NV];
#pragma unroll
    for (int n = 0; n < NV; ++n) out[n] = s(i, j, k, n);
    const double *dx = a.dx;
    const double dxmin = fmin(dx[0], fmin(dx[1], dx[2]));
    const double dnd[3] = {dx[0] / dxmin, dx[1] / dxmin, dx[2] / dxmin};
    const double xnd[3] = {a.geom[q], a.geom[(int64_t)a.ncut + q], a.geom[(int64_t)2 * a.ncut + q]};
    const double dist = a.geom[(int64_t)3 * a.ncut + q];
    const double cutoff = dxmin * 1e-5;
    const bool doT = !a.conj_ht, doS = !a.spec_in_solid;
    const bool vel0 = fabs(dist) < cutoff;
    double su0 = 0.0, su1 = 0.0, su2 = 0.0, swv = 0.0;
    double sp = 0.0, sT = 0.0, sR = 0.0, swn = 0.0, sS[NS];
#pragma unroll
    for (int q2 = 0; q2 < NS; ++q2) sS[q2] = 0.0;
    for (int w = 0; w < 4; ++w) {
        unsigned bits = a.mask[(int64_t)w * a.ncut + q];
        while (bits) {
            const int nb = w * 32 + __ffs(bits) - 1;
            bits &= bits - 1;
            const int l = nb % 5 - 2, m = (nb / 5) % 5 - 2, n = nb / 25 - 2;
            const double *c = s.p + s.off(i + l, j + m, k + n);
            const int64_t ns = s.ns;
            // one reciprocal per weight, multiplied into each value (the
            // reference divides every value by the squared distance)
            if (!vel0 && !(l == 0 && m == 0 && n == 0)) {
                double a0 = l * dnd[0], a1 = m * dnd[1], a2 = n * dnd[2];
                const double iw = div_pos(1.0, a0 * a0 + a1 * a1 + a2 * a2 + kTiny2);
                su0 += c[(int64_t)VELX * ns] * iw;
                su1 += c[(int64_t)VELY * ns] * iw;
                su2 += c[(int64_t)VELZ * ns] * iw;
                swv += iw;
            }
            double a0 = l * dnd[0] - xnd[0], a1 = m * dnd[1] - xnd[1], a2 = n * dnd[2] - xnd[2];
            const double iw = div_pos(1.0, a0 * a0 + a1 * a1 + a2 * a2 + kTiny2);
            sp += c[(int64_t)PRES * ns] * iw;
            if (doT) sT += c[(int64_t)TEMP * ns] * iw;
            sR += c[(int64_t)DENS * ns] * iw;
            if (doS) {
#pragma unroll
                for (int q2 = 0; q2 < NS; ++q2) sS[q2] += c[(int64_t)(FLO + q2) * ns] * iw;
            }
            swn += iw;
        }
    }
    // Dirichlet (u = 0) IDW of the velocities, :98-152
    if (vel0) {
        out[VELX] = 0.0; out[VELY] = 0.0; out[VELZ] = 0.0;
    } else {
        double dn = fabs(dist) / dxmin;
        double wd = 1.0 / (dn * dn);
        out[VELX] = (su0 + 0.0 / (dn * dn)) / (swv + wd);
        out[VELY] = (su1 + 0.0 / (dn * dn)) / (swv + wd);
        out[VELZ] = (su2 + 0.0 / (dn * dn)) / (swv + wd);
    }
    // homogeneous Neumann IDW of p, T, rho (and species without spec_in_solid), :154-198
    if (swn > 0.0) {
        out[PRES] = sp / swn;
        if (doT) out[TEMP] = sT / swn;
        out[DENS] = sR / swn;
        if (doS) {
#pragma unroll
            for (int q2 = 0; q2 < NS; ++q2) out[FLO + q2] = sS[q2] / swn;
        }
    }
    if (a.conj_ht) {
        double st = 0.0;
        int cnt = 0;
        for (int n = -1; n <= 1; n++)
            for (int m = -1; m <= 1; m++)
                for (int l = -1; l <= 1; l++)
                    if (!(l == 0 && m == 0 && n == 0)) { st += s(i + l, j + m, k + n, TEMP); cnt++; }
        out[TEMP] = st / (double)cnt;
    }
    // prim_to_cons of the updated primitives (:421-433)
    double u[N::NC], pv[NCV];
#pragma unroll
    for (int sp2 = 0; sp2 < NS; ++sp2) u[NCV + sp2] = out[FLO + sp2];
    pv[0] = out[DENS]; pv[1] = out[VELX]; pv[2] = out[VELY]; pv[3] = out[VELZ]; pv[4] = out[PRES];
    N::prim_to_cons(T, u, pv);
#pragma unroll
    for (int c = 0; c < NCV; ++c) out[c] = u[c];
#pragma unroll
    for (int n = 0; n < NV; ++n) a.cutbuf[(int64_t)n * a.ncut + q] = out[n];
}

// dense form over one box for the per-box C-ABI call: values to tmp, flags set
template <class P>
__global__ __launch_bounds__(128) void k_cutcells_box(CutArgs a) {
    constexpr int NV = Num<P>::NV;
    const int ex = a.hi[0] - a.lo[0] + 1, ey = a.hi[1] - a.lo[1] + 1, ez = a.hi[2] - a.lo[2] + 1;
    const int64_t ncell = (int64_t)ex * ey * ez;
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= ncell) return;
    const int i = a.lo[0] + (int)(q % ex), j = a.lo[1] + (int)((q / ex) % ey), k = a.lo[2] + (int)(q / ((int64_t)ex * ey));
    const double vf = a.s(i, j, k, VFRAC);
    const int cut = (vf > kZero && vf < kOne) ? 1 : 0;
    a.flag[q] = cut;
    if (!cut) return;
    double out[NV];
    cutcell_values<P>(*a.tab, a, i, j, k, out);
#pragma unroll
    for (int n = 0; n < NV; ++n) a.tmp[(int64_t)n * ncell + q] = out[n];
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Chemistry phase of the split mode (mflo_splitchem.cpp:40-62), one thread per
// valid cell: nsub substeps of Advance_chemistry (:172-195), an explicit
// midpoint rule whose first stage the reference evaluates on phi_new after
// its swap -- the state before the previous step or substep, held here in the
// lag buffer -- and its second on the half-step state; then, with finish,
// update_vars_after_chemsolve (:74-140; no background inert gas). Only the
// species have a right-hand side, so LinComb leaves every other component
// equal to the current state; temperature is constant over the phase.
// Consecutive launches continue the substep sequence exactly (the cell's
// state is its current and its lag buffer).
// ---------------------------------------------------------------------------
template <class P>
__global__ __launch_bounds__(256) void k_chem_split(ChemSplitArgs c) {
    using N = Num<P>;
    constexpr int NS = P::NS, NV = N::NV;
    const Tab &T = *c.tab;
    const int nx = c.n[0], ny = c.n[1];
    const int64_t ncell = (int64_t)nx * ny * c.n[2];
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < ncell; q += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((int64_t)nx * ny));
        double *sp = c.s.p + c.s.off(i, j, k);
        double *lp = c.lag.p + c.lag.off(i, j, k);
        const int64_t sn = c.s.ns, ln = c.lag.ns;
        double S[NV];
#pragma unroll
        for (int n = 0; n < NV; ++n) S[n] = sp[(int64_t)n * sn];
        if constexpr (P::kHasBgGas) {
            // store_inertgas_conc (mflo_splitchem.cpp:142-170): the background
            // gas concentration in the density slot while the chemistry runs
            if (c.bg && c.store_bg && S[VFRAC] >= kOne) S[DENS] = P::get_bgasconc_from_rc(T, S[RHO], S + FLO);
        }
        if (P::kHasChemistry && c.nsub > 0) {
            double L[NV];
#pragma unroll
            for (int n = 0; n < NV; ++n) L[n] = lp[(int64_t)n * ln];
            double u[NS], lsp[NS], src[NS], st[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) u[s] = S[FLO + s];
            const double hdt = kHalf * c.dt;
            // rate constants depend on temperature only: the lag state's for the
            // first stage of the first substep, the current state's otherwise
            // (the same values the source would recompute at every stage)
            double kS[MF_MAX_REACTIONS], kL[MF_MAX_REACTIONS];
            P::rate_constants(T, S[TEMP], kS);
            P::rate_constants(T, L[TEMP], kL);
            // first substep: stage 1 on the whole lag state
            P::spec_source_k(T, kL, ChemCell<NS>{L, L + FLO}, src);
#pragma unroll
            for (int s = 0; s < NS; ++s) st[s] = kOne * u[s] + hdt * (kZero + src[s]);
            P::spec_source_k(T, kS, ChemCell<NS>{S, st}, src);
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                lsp[s] = u[s];
                u[s] = kOne * u[s] + c.dt * (kZero + src[s]);
            }
            // later substeps: the lag state is the current one with the species
            // of the previous substep
            for (int m = 1; m < c.nsub; ++m) {
                P::spec_source_k(T, kS, ChemCell<NS>{S, lsp}, src);
#pragma unroll
                for (int s = 0; s < NS; ++s) st[s] = kOne * u[s] + hdt * (kZero + src[s]);
                P::spec_source_k(T, kS, ChemCell<NS>{S, st}, src);
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    lsp[s] = u[s];
                    u[s] = kOne * u[s] + c.dt * (kZero + src[s]);
                }
            }
            // the state before the last substep becomes the lag state
#pragma unroll
            for (int n = 0; n < NV; ++n) lp[(int64_t)n * ln] = (n >= FLO) ? lsp[n - FLO] : S[n];
#pragma unroll
            for (int s = 0; s < NS; ++s) S[FLO + s] = u[s];
        }
        if (c.finish && S[VFRAC] >= kOne) {
            double rho;
            if constexpr (P::kHasBgGas) {
                // update_vars_after_chemsolve with the background gas (mflo_splitchem.cpp:94-97)
                rho = c.bg ? P::get_r_from_c_bg(T, S + FLO, S[DENS]) : P::get_r_from_c(T, S + FLO);
            } else {
                rho = P::get_r_from_c(T, S + FLO);
            }
            S[RHO] = rho;
            S[DENS] = rho;
            S[RHOU] = rho * S[VELX];
            S[RHOV] = rho * S[VELY];
            S[RHOW] = rho * S[VELZ];
            double uu[N::NC], p[NCV];
#pragma unroll
            for (int n = 0; n < NCV; ++n) uu[n] = S[n];
#pragma unroll
            for (int s = 0; s < NS; ++s) uu[NCV + s] = S[FLO + s];
            N::cons_to_prim(T, uu, p);
#pragma unroll
            for (int n = 0; n < NCV; ++n) S[DENS + n] = p[n];
            S[TEMP] = P::get_t_from_rpc(T, S[DENS], S[PRES], uu + NCV);
        }
        if (c.finish && c.do_ns && S[PRES] < 0) raise_err(c.err, MF_ERR_NEG_PRESSURE);
#pragma unroll
        for (int n = 0; n < NV; ++n) sp[(int64_t)n * sn] = S[n];
    }
}

template <class P>
struct Launch {
    static void fill(const FillArgs &a, int phase, hipStream_t st) {
        const int G = MF_NGROW;
        int64_t PX = a.n[0] + 2 * G, PY = a.n[1] + 2 * G;
        int64_t total = 2 * G * PX * PY + (int64_t)a.n[2] * 2 * G * PX + (int64_t)a.n[2] * a.n[1] * 2 * G;
        int64_t grid = (total + 255) / 256;
        if (grid > 65535) grid = 65535;
        if (grid < 1) grid = 1;
        hipLaunchKernelGGL(k_fill<P>, dim3((unsigned)grid), dim3(256), 0, st, a, phase);
    }
    static void fill_region(const FillArgs &a, int phase, const int *rlo, const int *rhi, hipStream_t st) {
        const int64_t total = (int64_t)(rhi[0] - rlo[0] + 1) * (rhi[1] - rlo[1] + 1) * (rhi[2] - rlo[2] + 1);
        if (total <= 0) return;
        int64_t grid = (total + 255) / 256;
        if (grid > 65535) grid = 65535;
        hipLaunchKernelGGL(k_fill_region<P>, dim3((unsigned)grid), dim3(256), 0, st, a, phase,
                           make_int3(rlo[0], rlo[1], rlo[2]), make_int3(rhi[0], rhi[1], rhi[2]));
    }
    static void fill_list(const FillArgs &a, int phase, const int *list, int n, hipStream_t st) {
        if (n <= 0) return;
        const int grid = std::min((n + 255) / 256, 65535);
        hipLaunchKernelGGL(k_fill_list<P>, dim3((unsigned)grid), dim3(256), 0, st, a, phase, list, n);
    }
    static void chem(const ResArgs &r, const Work &w, int grid, int block, hipStream_t st) {
        hipLaunchKernelGGL(k_chem_tiles<P>, dim3(grid), dim3(block), 0, st, r, w);
    }
    static void cut(const CutArgs &a, hipStream_t st) {
        if (a.ncut <= 0) return;
        int g = (a.ncut + 127) / 128;
        if (g >= 8) g = (g + 7) & ~7;
        hipLaunchKernelGGL(k_cutcells<P>, dim3(g), dim3(128), 0, st, a);
    }
    static void cut_geom(const CutArgs &a, double *geom, unsigned *mask, hipStream_t st) {
        if (a.ncut <= 0) return;
        hipLaunchKernelGGL(k_cut_geom<P>, dim3((a.ncut + 127) / 128), dim3(128), 0, st, a, geom, mask);
    }
    static void cut_geom_sub(const CutArgs &a, double *geom, unsigned *mask, const int *sub, int nsub, hipStream_t st) {
        if (nsub <= 0) return;
        hipLaunchKernelGGL(k_cut_geom_sub<P>, dim3((nsub + 127) / 128), dim3(128), 0, st, a, geom, mask, sub, nsub);
    }
    static void cut_box(const CutArgs &a, hipStream_t st) {
        int64_t n = (int64_t)(a.hi[0] - a.lo[0] + 1) * (a.hi[1] - a.lo[1] + 1) * (a.hi[2] - a.lo[2] + 1);
        if (n <= 0) return;
        hipLaunchKernelGGL(k_cutcells_box<P>, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, a);
    }
    template <int ORDER>
    static void faces_o(const StageArgs &a, const Work &w, int dir, int grid, int block, hipStream_t st) {
        if (dir == 0) hipLaunchKernelGGL((k_faces<P, ORDER, 0>), dim3(grid), dim3(block), 0, st, a, w);
        else if (dir == 1) hipLaunchKernelGGL((k_faces<P, ORDER, 1>), dim3(grid), dim3(block), 0, st, a, w);
        else hipLaunchKernelGGL((k_faces<P, ORDER, 2>), dim3(grid), dim3(block), 0, st, a, w);
    }
    static void faces(const StageArgs &a, const Work &w, int dir, int grid, int block, hipStream_t st) {
        switch (a.order) {
        case 1: faces_o<1>(a, w, dir, grid, block, st); break;
        case 2: faces_o<2>(a, w, dir, grid, block, st); break;
        case 4: faces_o<4>(a, w, dir, grid, block, st); break;
        default: faces_o<5>(a, w, dir, grid, block, st); break;
        }
    }
    static size_t lds_bytes(const int *tsh, int dir) {
        int e[3] = {tsh[0], tsh[1], tsh[2]};
        e[dir] += 6;
        const size_t ncs = (size_t)e[0] * e[1] * e[2];
        return ncs * ((Num<P>::NC + 6) | 1) * sizeof(double) + ((ncs * sizeof(int) + 7) & ~(size_t)7);
    }
    static void faces_lds(const StageArgs &a, const Work &w, int dir, int grid, int block, size_t lds, hipStream_t st) {
        static bool attr = false;
        if (!attr) {
            hipFuncSetAttribute((const void *)k_faces_lds<P, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            hipFuncSetAttribute((const void *)k_faces_lds<P, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            hipFuncSetAttribute((const void *)k_faces_lds<P, 2>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            attr = true;
        }
        if (dir == 0) hipLaunchKernelGGL((k_faces_lds<P, 0>), dim3(grid), dim3(block), lds, st, a, w);
        else if (dir == 1) hipLaunchKernelGGL((k_faces_lds<P, 1>), dim3(grid), dim3(block), lds, st, a, w);
        else hipLaunchKernelGGL((k_faces_lds<P, 2>), dim3(grid), dim3(block), lds, st, a, w);
    }
    static void residual(const ResArgs &r, const Work &w, int grid, int block, hipStream_t st) {
        hipLaunchKernelGGL(k_residual<P>, dim3(grid), dim3(block), 0, st, r, w);
    }
    static void post(const ResArgs &r, const Work &w, int grid, int block, hipStream_t st) {
        if (r.do_chem) hipLaunchKernelGGL((k_post<P, true>), dim3(grid), dim3(block), 0, st, r, w);
        else hipLaunchKernelGGL((k_post<P, false>), dim3(grid), dim3(block), 0, st, r, w);
    }
    static void est(const ResArgs &r, const Work &w, int grid, int block, hipStream_t st) {
        hipLaunchKernelGGL(k_est<P>, dim3(grid), dim3(block), 0, st, r, w);
    }
    static void chem_split(const ChemSplitArgs &a, hipStream_t st) {
        const int64_t ncell = (int64_t)a.n[0] * a.n[1] * a.n[2];
        const int64_t g = std::min<int64_t>((ncell + 255) / 256, 8192);
        hipLaunchKernelGGL(k_chem_split<P>, dim3((unsigned)std::max<int64_t>(g, 1)), dim3(256), 0, st, a);
    }
    static void box_post(const BoxPostArgs &a, hipStream_t st) {
        const int64_t nc = (int64_t)a.e[0] * a.e[1] * a.e[2];
        const int64_t g = std::min<int64_t>((nc + 255) / 256, 8192);
        hipLaunchKernelGGL(k_box_post<P>, dim3((unsigned)std::max<int64_t>(g, 1)), dim3(256), 0, st, a);
    }
    static KernelSet set(const AdjSet *adj = nullptr) {
        KernelSet k;
        k.ns = P::NS;
        k.has_bg_gas = P::kHasBgGas;
        k.adj = adj;
        k.fill = &fill;
        k.fill_region = &fill_region;
        k.fill_list = &fill_list;
        k.chem = &chem;
        k.cut = &cut;
        k.cut_geom = &cut_geom;
        k.cut_geom_sub = &cut_geom_sub;
        k.cut_box = &cut_box;
        k.faces = &faces;
        k.faces_lds = &faces_lds;
        k.lds_bytes = &lds_bytes;
        k.residual = &residual;
        k.est = &est;
        k.post = &post;
        k.chem_split = &chem_split;
        k.box_post = &box_post;
        return k;
    }
};

}  // namespace mfd
