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

// tile t of the work list -> its box and inclusive cell range
MFD const BoxDesc &tile_box(const Work &w, int t, int &lt) {
    if (!w.boxes) { lt = t; return w.one; }
    int lo = 0, hi = w.nboxes - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (w.boxes[mid].tile_start <= t) lo = mid; else hi = mid - 1;
    }
    lt = t - w.boxes[lo].tile_start;
    return w.boxes[lo];
}
MFD void tile_range(const BoxDesc &b, const Work &w, int lt, int *c0, int *c1) {
    int ti = lt % b.ntile[0];
    int r = lt / b.ntile[0];
    int tj = r % b.ntile[1];
    int tk = r / b.ntile[1];
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

// ---------------------------------------------------------------------------
// Face flux of direction D at face (i,j,k) (between (i,j,k)-e_D and (i,j,k)):
// Source/Src_3d/compute_flux_3d.H:86-430.
// ---------------------------------------------------------------------------
template <class P, int ORDER, int D>
MFD void face_flux(const Tab &T, const Acc &acc, const StageArgs &a, int i, int j, int k, const View &F, double time) {
    using N = Num<P>;
    constexpr int NS = P::NS, NC = N::NC;
    constexpr int T1 = (D + 1) % 3, T2 = (D + 2) % 3;
    int e[3] = {0, 0, 0};
    e[D] = 1;
    // stencil cell s (s = -3..2) along D; -1 is `left`, 0 is `right`
    auto ci = [&](int s) { return i + s * e[0]; };
    auto cj = [&](int s) { return j + s * e[1]; };
    auto ck = [&](int s) { return k + s * e[2]; };
    int cid[6];
    double vf[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        cid[s] = acc.cid(ci(s - 3), cj(s - 3), ck(s - 3));
        vf[s] = acc.get(ci(s - 3), cj(s - 3), ck(s - 3), VFRAC, cid[s]);
    }
    const int cL = cid[2], cR = cid[3];
    const int iL = ci(-1), jL = cj(-1), kL = ck(-1);
    // viscosity / conductivity: harmonic mean of (i,j,k) and (i-1,j,k) for every
    // sweep direction (compute_flux_3d.H:179-180)
    double visc, thcond;
    {
        double v0, k0, v1, k1;
        CellRef c0{&acc, i, j, k, cR};
        P::fluid_transport(T, c0, v0, k0);
        CellRef c1{&acc, i - 1, j, k, acc.cid(i - 1, j, k)};
        P::fluid_transport(T, c1, v1, k1);
        visc = N::harmonic_mean(v0, v1);
        thcond = N::harmonic_mean(k0, k1);
    }
    double dl[NS], dr[NS];
    {
        CellRef l{&acc, iL, jL, kL, cL};
        CellRef r{&acc, i, j, k, cR};
        P::spec_dcoeff(T, l, dl);
        P::spec_dcoeff(T, r, dr);
    }
    double facevel[3];
    facevel[0] = kHalf * (acc.get(iL, jL, kL, VELX, cL) + acc.get(i, j, k, VELX, cR));
    facevel[1] = kHalf * (acc.get(iL, jL, kL, VELY, cL) + acc.get(i, j, k, VELY, cR));
    facevel[2] = kHalf * (acc.get(iL, jL, kL, VELZ, cL) + acc.get(i, j, k, VELZ, cR));
    const double facevfrac = kHalf * (vf[2] + vf[3]);
    const double dxs = a.dx[D];
    double Fo[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) Fo[c] = kZero;

    if (a.nsflag) {
        const bool first = (vf[0] < kOne || vf[1] < kOne || vf[2] < kOne || vf[3] < kOne || vf[4] < kOne || vf[5] < kOne);
        if (facevfrac >= 0.5) {
            double fh[NC];
            // stencil positions after the IB first-order fallback (:233-249)
            const int sidx[6] = {first ? -1 : -3, first ? -1 : -2, -1, 0, first ? 0 : 1, first ? 0 : 2};
            auto load = [&](int s, double *u) {
                int q = s + 3;
                int ii = ci(s), jj = cj(s), kk = ck(s);
                int c = (s == -1) ? cL : ((s == 0) ? cR : cid[q]);
#pragma unroll
                for (int n = 0; n < NCV; ++n) u[n] = acc.get(ii, jj, kk, n, c);
#pragma unroll
                for (int n = 0; n < NS; ++n) u[NCV + n] = acc.get(ii, jj, kk, FLO + n, c);
            };
            if (ORDER == 1 || ORDER == 2) {
                double ul[NC], ur[NC];
                load(-1, ul);
                load(0, ur);
                if (ORDER == 2) {
                    // compute_flux_3d.H:31-65 with uim2 = ulm1, uip1 = urp1
                    double ulm1[NC], urp1[NC];
                    load(sidx[1], ulm1);
                    load(sidx[4], urp1);
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
            } else if (ORDER == 4) {
                // flux_schemes.H:312-361 with uim1 = ul, ui = ur, uim2 = ulm1, uip1 = urp1
                double um2[NC], um1[NC], u0[NC], up1[NC];
                load(sidx[1], um2);
                load(-1, um1);
                load(0, u0);
                load(sidx[4], up1);
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
                double veln = fmax(fabs(um1[RHOU + D] / um1[RHO]), fabs(u0[RHOU + D] / u0[RHO]));
                double gamaL = P::get_gama_from_rpc(T, rhoL, presL, um1 + NCV);
                double gamaR = P::get_gama_from_rpc(T, rhoR, presR, u0 + NCV);
                double ahalf = fmax(sqrt(gamaL * presL / rhoL), sqrt(gamaR * presR / rhoR));
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    double fm2 = N::template euler_comp<D>(c, um2, pv[0][0], pv[0][1], pv[0][2], pv[0][3]);
                    double fm1 = N::template euler_comp<D>(c, um1, pv[1][0], pv[1][1], pv[1][2], pv[1][3]);
                    double f0 = N::template euler_comp<D>(c, u0, pv[2][0], pv[2][1], pv[2][2], pv[2][3]);
                    double fp1 = N::template euler_comp<D>(c, up1, pv[3][0], pv[3][1], pv[3][2], pv[3][3]);
                    double v = kTwelveth * (7 * f0 + 7 * fm1 - fm2 - fp1);
                    v -= kTwelveth * (3 * u0[c] - 3 * um1[c] - up1[c] + u0[c]) * (veln + ahalf) * a.dissfactor;
                    fh[c] = v;
                }
            } else {
                // WENO5 flux splitting, flux_schemes.H:448-495
                double lam;
                {
                    double ul[NC], ur[NC];
                    load(-1, ul);
                    load(0, ur);
                    lam = N::template max_eigenvalue<D>(T, ul, ur);
                }
                double pr[6][5];  // rho, vx, vy, vz, p of the six stencil cells
#pragma unroll
                for (int s = 0; s < 6; ++s) {
                    double u[NC];
                    load(sidx[s], u);
                    pr[s][0] = u[RHO];
                    N::euler_prims(T, u, pr[s][1], pr[s][2], pr[s][3], pr[s][4]);
                }
                int sc[6];
#pragma unroll
                for (int s = 0; s < 6; ++s) sc[s] = (sidx[s] == -1) ? cL : ((sidx[s] == 0) ? cR : cid[sidx[s] + 3]);
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    const int comp = c < NCV ? c : FLO + (c - NCV);
                    double uc[6], fc[6];
#pragma unroll
                    for (int s = 0; s < 6; ++s) {
                        int ss = sidx[s];
                        uc[s] = acc.get(ci(ss), cj(ss), ck(ss), comp, sc[s]);
                        double vn = pr[s][1 + D];
                        if (c == RHO) fc[s] = pr[s][0] * vn;
                        else if (c >= RHOU && c <= RHOW) {
                            double m = pr[s][0] * pr[s][c];
                            fc[s] = (c - RHOU == D) ? (m * pr[s][c] + pr[s][4]) : (m * vn);
                        } else if (c == RHOE) fc[s] = (uc[s] + pr[s][4]) * vn;
                        else fc[s] = uc[s] * vn;
                    }
                    double fp = N::weno_plus(fc[0], fc[1], fc[2], fc[3], fc[4]);
                    double fm = N::weno_minus(fc[1], fc[2], fc[3], fc[4], fc[5]);
                    double up = N::weno_plus(uc[0], uc[1], uc[2], uc[3], uc[4]);
                    double um = N::weno_minus(uc[1], uc[2], uc[3], uc[4], uc[5]);
                    fh[c] = kHalf * (fp + lam * up) + kHalf * (fm - lam * um);
                }
            }
#pragma unroll
            for (int c = 0; c < NCV; ++c) Fo[c] = fh[c];
#pragma unroll
            for (int s = 0; s < NS; ++s) Fo[NCV + s] = T.advect_flag[s] * fh[NCV + s];

            // viscous stress and work (:293-367); the unit normal keeps row D of tau
            double g[3][3];  // g[q][m]: derivative of velocity q along normal (m=0), t1 (m=1), t2 (m=2)
            {
                int tl[3] = {iL, jL, kL}, tr[3] = {i, j, k};
                int p1[3] = {0, 0, 0}, p2[3] = {0, 0, 0};
                p1[T1] = 1;
                p2[T2] = 1;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    int V = VELX + q;
                    double vr = acc.get(i, j, k, V, cR), vl = acc.get(iL, jL, kL, V, cL);
                    g[q][0] = (vr - vl) / dxs;
                    double top = kHalf * (acc(tl[0] + p1[0], tl[1] + p1[1], tl[2] + p1[2], V) + acc(tr[0] + p1[0], tr[1] + p1[1], tr[2] + p1[2], V));
                    double bot = kHalf * (acc(tl[0] - p1[0], tl[1] - p1[1], tl[2] - p1[2], V) + acc(tr[0] - p1[0], tr[1] - p1[1], tr[2] - p1[2], V));
                    g[q][1] = (top - bot) / (kTwo * a.dx[T1]);
                    double frt = kHalf * (acc(tl[0] + p2[0], tl[1] + p2[1], tl[2] + p2[2], V) + acc(tr[0] + p2[0], tr[1] + p2[1], tr[2] + p2[2], V));
                    double bck = kHalf * (acc(tl[0] - p2[0], tl[1] - p2[1], tl[2] - p2[2], V) + acc(tr[0] - p2[0], tr[1] - p2[1], tr[2] - p2[2], V));
                    g[q][2] = (frt - bck) / (kTwo * a.dx[T2]);
                }
            }
            // axis a -> slot of the gradient array (GET_XDIR/GET_YDIR/GET_ZDIR)
            auto slot = [&](int ax) { return ax == D ? 0 : (ax == T1 ? 1 : 2); };
            double dudx = g[0][slot(0)], dudy = g[0][slot(1)], dudz = g[0][slot(2)];
            double dvdx = g[1][slot(0)], dvdy = g[1][slot(1)], dvdz = g[1][slot(2)];
            double dwdx = g[2][slot(0)], dwdy = g[2][slot(1)], dwdz = g[2][slot(2)];
            double delu = dudx + dvdy + dwdz;
            double tau0, tau1, tau2;  // row D of tau
            if (D == 0) {
                tau0 = visc * (kTwo * dudx - kTwo3rd * delu);
                tau1 = visc * (dudy + dvdx);
                tau2 = visc * (dudz + dwdx);
            } else if (D == 1) {
                tau0 = visc * (dvdx + dudy);
                tau1 = visc * (kTwo * dvdy - kTwo3rd * delu);
                tau2 = visc * (dvdz + dwdy);
            } else {
                tau0 = visc * (dwdx + dudz);
                tau1 = visc * (dwdy + dvdz);
                tau2 = visc * (kTwo * dwdz - kTwo3rd * delu);
            }
            Fo[RHOU] -= tau0;
            Fo[RHOV] -= tau1;
            Fo[RHOW] -= tau2;
            Fo[RHOE] -= tau0 * facevel[0];
            Fo[RHOE] -= tau1 * facevel[1];
            Fo[RHOE] -= tau2 * facevel[2];
        }
        double dTdn = (acc.get(i, j, k, TEMP, cR) - acc.get(iL, jL, kL, TEMP, cL)) / dxs;
        Fo[RHOE] -= thcond * dTdn;
    } else {
        // species convection without the flow solve (:379-422)
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double phiL = acc.get(iL, jL, kL, FLO + s, cL);
            double phiR = acc.get(i, j, k, FLO + s, cR);
            double velface = facevel[D];
            if (a.spec_in_solid) {
                double w = (vf[2] < kOne) ? kZero : kOne;  // both weights test the left cell
                velface = (w * acc.get(iL, jL, kL, VELX + D, cL) + w * acc.get(i, j, k, VELX + D, cR)) / (w + w + kTiny);
            }
            Fo[NCV + s] = T.advect_flag[s] * (kHalf * (velface + fabs(velface)) * phiL + kHalf * (velface - fabs(velface)) * phiR);
        }
    }
    // species diffusion (:425-428)
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double D_s = N::harmonic_mean(dl[s], dr[s]);
        double cr = acc.get(i, j, k, FLO + s, cR), cl = acc.get(iL, jL, kL, FLO + s, cL);
        Fo[NCV + s] -= D_s * (cr - cl) / dxs;
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) F(i, j, k, flux_comp(c, a.flux_full)) = Fo[c];
    if (a.flux_full) {
#pragma unroll
        for (int n = NCV; n < FLO; ++n) F(i, j, k, n) = kZero;
    }
}

template <class P, int ORDER, int D>
__global__ __launch_bounds__(256) void k_faces(StageArgs a, Work w) {
    if (step_skipped(a.ts)) return;
    const Tab &T = *a.tab;
    const double time = stage_time(a);
    for (int t = w.t0 + blockIdx.x; t < w.t1; t += gridDim.x) {
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
// Post-step work on one cell (fused epilogue of stage 3):
// Advance_chemistry_implicit (SSPRK3 of AMReX's TimeIntegrator, rk.type=3),
// update_primitive_vars, clip_neg_speciesconc, residual-norm and dt partials.
// ---------------------------------------------------------------------------
template <class P>
MFD void chemistry_rk3(const Tab &T, double *S, double dt) {
    constexpr int NS = P::NS, NV = Num<P>::NV;
    if (!P::kHasChemistry) return;
    double S0[NS], F[3][NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) S0[s] = S[FLO + s];
    const double a10 = 1.0, a20 = 0.25, a21 = 0.25;
    double cell[NV];
#pragma unroll
    for (int n = 0; n < NV; ++n) cell[n] = S[n];
#pragma unroll
    for (int st = 0; st < 3; ++st) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double v = S0[s];
            if (st == 1) v += (dt * a10) * F[0][s];
            if (st == 2) { v += (dt * a20) * F[0][s]; v += (dt * a21) * F[1][s]; }
            cell[FLO + s] = v;
        }
        double src[NS];
        P::spec_source(T, CellArr<NV>{cell}, src);
#pragma unroll
        for (int s = 0; s < NS; ++s) F[st][s] = 0.0 + src[s];
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
// (mflo_evolve.cpp:283, 296-298, 309-311) per cell of a tile.
// ---------------------------------------------------------------------------
template <class P>
__global__ __launch_bounds__(256) void k_residual(ResArgs r, Work w) {
    using N = Num<P>;
    constexpr int NS = P::NS, NV = N::NV;
    const StageArgs &a = r.st;
    if (step_skipped(a.ts)) return;
    const Tab &T = *a.tab;
    const double time = stage_time(a);
    const double dt = stage_dt(a);
    (void)time;
    const int nvn = (a.spec_in_solid == 1) ? FLO : FLO + NS;
    for (int t = w.t0 + blockIdx.x; t < w.t1; t += gridDim.x) {
        int lt;
        const BoxDesc &b = tile_box(w, t, lt);
        int c0[3], c1[3];
        tile_range(b, w, lt, c0, c1);
        const int ex = c1[0] - c0[0] + 1, ey = c1[1] - c0[1] + 1, ez = c1[2] - c0[2] + 1;
        const int nc = ex * ey * ez;
        Acc acc = make_acc(a, b);
        double sums[NV];
#pragma unroll
        for (int n = 0; n < NV; ++n) sums[n] = 0.0;
        double mws = 0.0, mD = 0.0;
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
            const double vfc = acc.get(i, j, k, VFRAC, cid);
            double S[NV], So[NV];
#pragma unroll
            for (int n = 0; n < NV; ++n) {
                double d;
                if (n < NCV || n >= FLO) {
                    const int c = flux_comp(n < NCV ? n : n - (FLO - NCV), a.flux_full);
                    const View &fx = b.flux[0], &fy = b.flux[1], &fz = b.flux[2];
                    d = (fx(i, j, k, c) - fx(i + 1, j, k, c)) / a.dx[0] + (fy(i, j, k, c) - fy(i, j + 1, k, c)) / a.dx[1] +
                        (fz(i, j, k, c) - fz(i, j, k + 1, c)) / a.dx[2] + (n < FLO ? src[n] : kZero);
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
                if (!r.combine) {
                    r.dsdt(i, j, k, n) = d;
                    continue;
                }
                if (a.stage == 1) {
                    S[n] = kOne * acc.get(i, j, k, n, cid) + dt * d;
                } else if (a.stage == 2) {
                    double v = a.src(i, j, k, n);
                    v += dt * d;
                    v += 3.0 * r.old(i, j, k, n);
                    S[n] = v * kFourth;
                } else {
                    double o = r.old(i, j, k, n);
                    So[n] = o;
                    double v = a.src(i, j, k, n);
                    v += dt * d;
                    v = o + kTwo * v;
                    S[n] = v * kThird;
                }
            }
            if (!r.combine) continue;
            if (r.epilogue) {
                if (r.do_chem) {
#pragma unroll
                    for (int n = 0; n < NV; ++n) So[n] = S[n];
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
                    const double ia = 1.0 / dt, ib = -1.0 / dt;
#pragma unroll
                    for (int n = 0; n < NV; ++n) {
                        double df = ia * S[n] + ib * So[n];
                        if (n < VFRAC) df *= So[VFRAC];
                        sums[n] += df * df;
                    }
                }
                double ws, md;
                dt_partial<P>(T, S, r.do_ns, ws, md);
                mws = fmax(mws, ws);
                mD = fmax(mD, md);
            }
#pragma unroll
            for (int n = 0; n < NV; ++n) r.dst(i, j, k, n) = S[n];
        }
        if (r.combine && r.epilogue) block_reduce_store<NV>(sums, mws, mD, r.partials + (size_t)t * (NV + 2));
    }
}

// EstTimeStep partials on the current state (first step of a run)
template <class P>
__global__ __launch_bounds__(256) void k_est(ResArgs r, Work w) {
    constexpr int NV = Num<P>::NV;
    const StageArgs &a = r.st;
    const Tab &T = *a.tab;
    for (int t = w.t0 + blockIdx.x; t < w.t1; t += gridDim.x) {
        int lt;
        const BoxDesc &b = tile_box(w, t, lt);
        int c0[3], c1[3];
        tile_range(b, w, lt, c0, c1);
        const int ex = c1[0] - c0[0] + 1, ey = c1[1] - c0[1] + 1, ez = c1[2] - c0[2] + 1;
        double sums[NV];
#pragma unroll
        for (int n = 0; n < NV; ++n) sums[n] = 0.0;
        double mws = 0.0, mD = 0.0;
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
        block_reduce_store<NV>(sums, mws, mD, r.partials + (size_t)t * (NV + 2));
    }
}

// ---------------------------------------------------------------------------
// Ghost cells of the padded level array (FillPatchSingleLevel): phase 0 fills
// cells outside only in periodic directions from their periodic images; phase
// p >= 1 fills cells outside in p non-periodic directions with the filcc rules
// (foextrap, reflect_even/odd for walls: Source/mflo.cpp:70-143) in x,y,z order,
// then the ext_dir functor of Source/mflo_bcfill.H:16-128 in x,y,z order.
// ---------------------------------------------------------------------------
template <class P>
__global__ __launch_bounds__(256) void k_fill(FillArgs a, int phase) {
    constexpr int NV = Num<P>::NV;
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
        int out_np = 0, out_p = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            bool o = iv[d] < 0 || iv[d] >= a.n[d];
            if (o) { if (a.periodic[d]) ++out_p; else ++out_np; }
        }
        if (phase == 0) {
            if (out_np != 0 || out_p == 0) continue;
            int s[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                s[d] = iv[d];
                if (a.periodic[d]) { if (s[d] < 0) s[d] += a.n[d]; else if (s[d] >= a.n[d]) s[d] -= a.n[d]; }
            }
            for (int c = 0; c < NV; ++c) a.s(iv[0], iv[1], iv[2], c) = a.s(s[0], s[1], s[2], c);
            continue;
        }
        if (out_np != phase) continue;
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
    constexpr int NN = 3 + NS;
    int comps[NN];
    int nn = 0;
    comps[nn++] = PRES;
    if (!a.conj_ht) comps[nn++] = TEMP;
    comps[nn++] = DENS;
    if (!a.spec_in_solid) for (int sp = 0; sp < NS; ++sp) comps[nn++] = FLO + sp;
    double sv[NN], sw = 0.0;
    for (int c = 0; c < NN; ++c) sv[c] = 0.0;
    for (int n = -2; n <= 2; n++)
        for (int m = -2; m <= 2; m++)
            for (int l = -2; l <= 2; l++) {
                if (s(i + l, j + m, k + n, VFRAC) == 1.0) {
                    double a0 = l * dnd[0] - xnd[0], a1 = m * dnd[1] - xnd[1], a2 = n * dnd[2] - xnd[2];
                    double w2 = a0 * a0 + a1 * a1 + a2 * a2 + kTiny2;
                    for (int c = 0; c < nn; ++c) sv[c] += s(i + l, j + m, k + n, comps[c]) / w2;
                    sw += 1.0 / w2;
                }
            }
    for (int c = 0; c < nn; ++c) out[comps[c]] = (sw > 0.0) ? sv[c] / sw : s(i, j, k, comps[c]);
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

template <class P>
__global__ __launch_bounds__(128) void k_cutcells(CutArgs a) {
    constexpr int NV = Num<P>::NV;
    if (step_skipped(a.ts)) return;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.ncut) return;
    double out[NV];
    cutcell_values<P>(*a.tab, a, a.list[q], a.list[a.ncut + q], a.list[2 * a.ncut + q], out);
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
    static void cut(const CutArgs &a, hipStream_t st) {
        if (a.ncut <= 0) return;
        hipLaunchKernelGGL(k_cutcells<P>, dim3((a.ncut + 127) / 128), dim3(128), 0, st, a);
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
    static void residual(const ResArgs &r, const Work &w, int grid, int block, hipStream_t st) {
        hipLaunchKernelGGL(k_residual<P>, dim3(grid), dim3(block), 0, st, r, w);
    }
    static void est(const ResArgs &r, const Work &w, int grid, int block, hipStream_t st) {
        hipLaunchKernelGGL(k_est<P>, dim3(grid), dim3(block), 0, st, r, w);
    }
    static KernelSet set() {
        KernelSet k;
        k.ns = P::NS;
        k.fill = &fill;
        k.cut = &cut;
        k.cut_box = &cut_box;
        k.faces = &faces;
        k.residual = &residual;
        k.est = &est;
        return k;
    }
};

}  // namespace mfd
