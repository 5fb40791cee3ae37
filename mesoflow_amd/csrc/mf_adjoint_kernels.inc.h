// mf_adjoint_kernels.inc.h -- discrete-adjoint kernels of the RHS + RK3 path,
// instantiated per plugin family by the mf_kernels_<family>.hip translation units
// (the porous-reactor families). The reverse sweep of one RK3 step runs, per
// stage in reverse order:
//   k_faces_adj (x, y, z)  adjoint of the face fluxes and update_residual
//   k_cut_adj              adjoint of update_cutcells
//   k_fill_adj (3, 2, 1, 0) adjoint of the ghost fill
// after k_post_adj (adjoint of chemistry, update_primitive_vars, clip and of
// the objective's outlet term of the step).
#pragma once
#include "mf_kernels.inc.h"
#include "mf_adjoint.h"

#define MFK __device__ __forceinline__

namespace mfd {

// global accumulation of an input adjoint: own-box cut cells were read through
// the stage's cut values, every other cell from the array
struct AdjSink {
    View lam;
    double *lcut;
    int ncut;
    MFK void add(int i, int j, int k, int n, int cid, double v) const {
        if (v == 0.0) return;
        if (cid >= 0) unsafeAtomicAdd(lcut + (int64_t)n * ncut + cid, v);
        else unsafeAtomicAdd(&lam(i, j, k, n), v);
    }
};

// update_residual's masking of dsdt component n (div_3d.H:35-62)
MFK bool dsdt_kept(int n, double vfc, int nvn, int conj_ht) {
    if (n != RHOE) return !(n < nvn && vfc < kOne);
    if (!conj_ht) return !(vfc < kOne);
    return !(vfc > kZero && vfc < kOne);
}

// LDS tile of the adjoint sweep: one record of AS = 2 NC + 9 doubles (odd) per
// stencil cell: forward values u[NC], vx, vy, vz, p, vfrac, then the adjoint
// accumulators ub[NC], vb[3], pb
template <int NC>
struct AdjRec {
    static constexpr int FVX = NC, FP = NC + 3, FVF = NC + 4, AU = NC + 5, AV = 2 * NC + 5, AP = 2 * NC + 8;
    static constexpr int AS = 2 * NC + 9;
};
// stencil of one face in the LDS tile; positions q = 0..5 map to records
// q0 + ms[q]*sD. fbs is the face's flux adjoint, a register array of the
// calling thread: the reverse sweep indexes it with compile-time components
// only.
template <int NC>
struct LdsStencil {
    using R_ = AdjRec<NC>;
    double *lds;
    const double *fbs;
    int q0, sD;
    int ms[6];
    MFK double *rec(int q) const { return lds + (q0 + ms[q] * sD) * R_::AS; }
    MFK double u(int q, int c) const { return rec(q)[c]; }
    MFK double vel(int q, int a) const { return rec(q)[R_::FVX + a]; }
    MFK double pres(int q) const { return rec(q)[R_::FP]; }
    MFK double fb(int c) const { return fbs[c]; }
    MFK void add_u(int q, int c, double x) const { if (x != 0.0) atomicAdd(rec(q) + R_::AU + c, x); }
    MFK void add_v(int q, int a, double x) const { if (x != 0.0) atomicAdd(rec(q) + R_::AV + a, x); }
    MFK void add_p(int q, double x) const { if (x != 0.0) atomicAdd(rec(q) + R_::AP, x); }
};

// ---------------------------------------------------------------------------
// Adjoint of the three face sweeps of direction D and of update_residual for
// one stage (do_ns = 1, order_hyp = 5). The flux adjoint of a face follows from
// the adjoint of the stage output: Fbar = (ld(R) - ld(L)) / dx with
// ld = coef * mask * lam_out for the cells of the face's box. The tile stages
// the forward values like k_faces_lds and accumulates the adjoints of the
// stencil cells in LDS; the stored primitives read by pass 2 (velocities,
// temperature) go to global memory directly.
// ---------------------------------------------------------------------------
template <class P, int D>
__global__ __launch_bounds__(128) void k_faces_adj(AdjFaceArgs aa, Work w) {
    using N = Num<P>;
    using R = Rev<P>;
    constexpr int NS = P::NS, NC = N::NC;
    constexpr int e0 = D == 0, e1 = D == 1, e2 = D == 2;
    constexpr int T1 = (D + 1) % 3, T2 = (D + 2) % 3;
    using St = LdsStencil<NC>;
    using RR = AdjRec<NC>;
    constexpr int FVX = RR::FVX, FP = RR::FP, FVF = RR::FVF, AU = RR::AU, AV = RR::AV, AP = RR::AP, AS = RR::AS;
    extern __shared__ double lds[];
    const StageArgs &a = aa.st;
    if (step_skipped(a.ts)) return;
    const Tab &T = *a.tab;
    const double dxs = a.dx[D];
    const int nvn = (a.spec_in_solid == 1) ? FLO : FLO + NS;
    const AdjSink sink{aa.lam, aa.lcut, a.ncut};
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
        Acc acc = make_acc(a, b);
        for (int q = threadIdx.x; q < ncs; q += blockDim.x) {
            const int ix = q % sx, r = q / sx, iy = r % sy, iz = r / sy;
            const int ci = ox + ix, cj = oy + iy, ck = oz + iz;
            const CellPtr cp = cell_ptr(acc, ci, cj, ck, acc.cid(ci, cj, ck));
            double u[NC];
#pragma unroll
            for (int n = 0; n < NCV; ++n) u[n] = cp(n);
#pragma unroll
            for (int n = 0; n < NS; ++n) u[NCV + n] = cp(FLO + n);
            double *rec = lds + q * AS;
#pragma unroll
            for (int c = 0; c < NC; ++c) rec[c] = u[c];
            double vx, vy, vz, p;
            N::euler_prims(T, u, vx, vy, vz, p);
            rec[FVX + 0] = vx;
            rec[FVX + 1] = vy;
            rec[FVX + 2] = vz;
            rec[FP] = p;
            rec[FVF] = cp(VFRAC);
#pragma unroll
            for (int f = AU; f < AS; ++f) rec[f] = 0.0;
        }
        __syncthreads();
        const int nf = fx * fy * fz;
        const int sD = D == 0 ? 1 : (D == 1 ? sx : sx * sy);
        for (int f = threadIdx.x; f < nf; f += blockDim.x) {
            const int ix = f % fx, r = f / fx, iy = r % fy, iz = r / fy;
            const int i = c0[0] + ix, j = c0[1] + iy, k = c0[2] + iz;
            const int q0 = ix + sx * (iy + sy * iz);
            const int iL = i - e0, jL = j - e1, kL = k - e2;
            const int fD = D == 0 ? i : (D == 1 ? j : k);
            const bool Rin = fD <= b.hi[D], Lin = fD - 1 >= b.lo[D];
            double *recL = lds + (q0 + 2 * sD) * AS, *recR = lds + (q0 + 3 * sD) * AS;
            const double vfL = recL[FVF], vfR = recR[FVF];
            double Fb[NC];
            bool any = false;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const int n = c < NCV ? c : FLO + (c - NCV);
                const double lr = (Rin && dsdt_kept(n, vfR, nvn, a.conj_ht)) ? aa.coef * aa.lam_out(i, j, k, n) : 0.0;
                const double ll = (Lin && dsdt_kept(n, vfL, nvn, a.conj_ht)) ? aa.coef * aa.lam_out(iL, jL, kL, n) : 0.0;
                Fb[c] = lr / dxs - ll / dxs;
                any = any || Fb[c] != 0.0;
            }
            if (!any) continue;
            const double facevfrac = kHalf * (vfL + vfR);
            const bool hyp = facevfrac >= 0.5;
            // pass 2's global loads first (cut ids, L/R velocities, transverse
            // cut ids); the LLF reverse covers their latency
            constexpr int p10 = T1 == 0, p11 = T1 == 1, p12 = T1 == 2;
            constexpr int p20 = T2 == 0, p21 = T2 == 1, p22 = T2 == 2;
            const int cL = acc.cid(iL, jL, kL), cR = acc.cid(i, j, k);
            const int ti[8][3] = {{iL + p10, jL + p11, kL + p12}, {i + p10, j + p11, k + p12},
                                  {iL - p10, jL - p11, kL - p12}, {i - p10, j - p11, k - p12},
                                  {iL + p20, jL + p21, kL + p22}, {i + p20, j + p21, k + p22},
                                  {iL - p20, jL - p21, kL - p22}, {i - p20, j - p21, k - p22}};
            int tc[8];
            if (hyp) {
#pragma unroll
                for (int m = 0; m < 8; ++m) tc[m] = acc.cid(ti[m][0], ti[m][1], ti[m][2]);
            }
            const CellPtr pL = cell_ptr(acc, iL, jL, kL, cL), pR = cell_ptr(acc, i, j, k, cR);
            const double vL[3] = {pL(VELX), pL(VELY), pL(VELZ)}, vR[3] = {pR(VELX), pR(VELY), pR(VELZ)};
            if (hyp) {
                bool first = false;
#pragma unroll
                for (int m = 0; m < 6; ++m) first = first || (lds[(q0 + m * sD) * AS + FVF] < kOne);
                St st;
                st.lds = lds;
                st.fbs = Fb;
                st.q0 = q0;
                st.sD = sD;
                st.ms[0] = first ? 2 : 0;
                st.ms[1] = first ? 2 : 1;
                st.ms[2] = 2;
                st.ms[3] = 3;
                st.ms[4] = first ? 3 : 4;
                st.ms[5] = first ? 3 : 5;
                R::template face_hyp_rev<D>(T, st);
            }
            // pass 2: viscous stress and work, conduction, species diffusion
            double visc, thcond;
            {
                double v0, k0, v1, k1;
                CellRef cc0{&acc, i, j, k, cR};
                P::fluid_transport(T, cc0, v0, k0);
                CellRef cc1{&acc, i - 1, j, k, acc.cid(i - 1, j, k)};
                P::fluid_transport(T, cc1, v1, k1);
                visc = N::harmonic_mean(v0, v1);
                thcond = N::harmonic_mean(k0, k1);
            }
            double Ds[NS];
            {
                double dl[NS], dr[NS];
                CellRef l{&acc, iL, jL, kL, cL};
                CellRef rr{&acc, i, j, k, cR};
                P::spec_dcoeff(T, l, dl);
                P::spec_dcoeff(T, rr, dr);
#pragma unroll
                for (int s = 0; s < NS; ++s) Ds[s] = N::harmonic_mean(dl[s], dr[s]);
            }
            double G[3][3];
#pragma unroll
            for (int q = 0; q < 3; ++q) G[q][0] = G[q][1] = G[q][2] = 0.0;
            if (hyp) {
                double tv[8][3];
#pragma unroll
                for (int m = 0; m < 8; ++m) {
#pragma unroll
                    for (int q = 0; q < 3; ++q) tv[m][q] = acc.get(ti[m][0], ti[m][1], ti[m][2], VELX + q, tc[m]);
                }
                R::template grad_tensor<D>(a.dx, vL, vR, tv, G);
            }
            double vLb[3] = {0.0, 0.0, 0.0}, vRb[3] = {0.0, 0.0, 0.0}, TLb = 0.0, TRb = 0.0, cLb[NS], cRb[NS], h1[3], h2[3];
#pragma unroll
            for (int s = 0; s < NS; ++s) { cLb[s] = 0.0; cRb[s] = 0.0; }
            R::template pass2_rev<D>(Fb, true, hyp, visc, thcond, Ds, a.dx, vL, vR, G, vLb, vRb, h1, h2, TLb, TRb, cLb, cRb);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                sink.add(iL, jL, kL, VELX + q, cL, vLb[q]);
                sink.add(i, j, k, VELX + q, cR, vRb[q]);
            }
            sink.add(iL, jL, kL, TEMP, cL, TLb);
            sink.add(i, j, k, TEMP, cR, TRb);
            if (hyp) {
#pragma unroll
                for (int m = 0; m < 8; ++m)
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        const double h = (m < 4 ? h1[q] : h2[q]) * ((m & 2) ? -kHalf : kHalf);
                        sink.add(ti[m][0], ti[m][1], ti[m][2], VELX + q, tc[m], h);
                    }
            }
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                if (cLb[s] != 0.0) atomicAdd(recL + AU + NCV + s, cLb[s]);
                if (cRb[s] != 0.0) atomicAdd(recR + AU + NCV + s, cRb[s]);
            }
        }
        __syncthreads();
        // primitive adjoints -> conserved, then the tile's accumulators to global
        for (int q = threadIdx.x; q < ncs; q += blockDim.x) {
            const int ix = q % sx, r = q / sx, iy = r % sy, iz = r / sy;
            const int ci = ox + ix, cj = oy + iy, ck = oz + iz;
            const double *rec = lds + q * AS;
            double u[NC], ub[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) { u[c] = rec[c]; ub[c] = rec[AU + c]; }
            const double vb[3] = {rec[AV], rec[AV + 1], rec[AV + 2]};
            const double pb = rec[AP];
            if (vb[0] != 0.0 || vb[1] != 0.0 || vb[2] != 0.0 || pb != 0.0) R::prims_rev(T, u, vb, pb, ub);
            const int cid = acc.cid(ci, cj, ck);
#pragma unroll
            for (int c = 0; c < NC; ++c) sink.add(ci, cj, ck, c < NCV ? c : FLO + (c - NCV), cid, ub[c]);
        }
        __syncthreads();
    }
}

// adjoint of update_cutcells in gather form: k_cut_coef, one thread per cut
// cell, writes the record of terms for the cell's fluid neighbours and adds
// the cell's own terms; k_cut_gather, one thread per receiving fluid cell,
// sums the records of the cut cells around it. Every adjoint value has one
// writer: no atomics.
template <class P>
__global__ __launch_bounds__(128) void k_cut_coef(AdjCutArgs aa) {
    constexpr int NV = Num<P>::NV;
    using R = Rev<P>;
    const CutArgs &a = aa.c;
    if (step_skipped(a.ts)) return;
    const int q = xcd_chunk(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (q >= a.ncut) return;
    double ob[NV], out[NV], g[6];
#pragma unroll
    for (int n = 0; n < NV; ++n) {
        ob[n] = aa.lcut[(int64_t)n * a.ncut + q];
        out[n] = a.cutbuf[(int64_t)n * a.ncut + q];
    }
#pragma unroll
    for (int d = 0; d < 6; ++d) g[d] = a.geom[(int64_t)d * a.ncut + q];
    const int i = a.list[q], j = a.list[a.ncut + q], k = a.list[2 * a.ncut + q];
    const View lam = aa.lam;
    auto self = [&](int n, double v) {
        if (v != 0.0) lam(i, j, k, n) += v;
    };
    double rec[R::CUTREC];
    R::cut_coef(*a.tab, a, g, ob, out, rec, self);
    double *dst = aa.coef + (int64_t)q * R::CUTREC;
#pragma unroll
    for (int c = 0; c < R::CUTREC; ++c) dst[c] = rec[c];
}

template <class P>
__global__ __launch_bounds__(128) void k_cut_gather(AdjCutArgs aa) {
    constexpr int NS = P::NS;
    using R = Rev<P>;
    const CutArgs &a = aa.c;
    if (step_skipped(a.ts)) return;
    const int q = xcd_chunk(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (q >= aa.nnbr) return;
    const int i = aa.nbr[q], j = aa.nbr[aa.nnbr + q], k = aa.nbr[2 * aa.nnbr + q];
    unsigned mb[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) mb[w] = aa.nbrmask[(int64_t)w * aa.nnbr + q];
    const int nx = aa.n[0], ny = aa.n[1];
    auto rec_of = [&](int l, int m, int n) -> const double * {
        const int cid = aa.cutid[(int64_t)(i - l) + (int64_t)nx * ((int64_t)(j - m) + (int64_t)ny * (k - n))];
        return aa.coef + (int64_t)cid * R::CUTREC;
    };
    double res[6 + NS];
    R::cut_gather(a.dx, mb, rec_of, res);
    const View lam = aa.lam;
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (res[d] != 0.0) lam(i, j, k, VELX + d) += res[d];
    if (res[3] != 0.0) lam(i, j, k, PRES) += res[3];
    if (res[4] != 0.0) lam(i, j, k, TEMP) += res[4];
    if (res[5] != 0.0) lam(i, j, k, DENS) += res[5];
#pragma unroll
    for (int s = 0; s < NS; ++s)
        if (res[6 + s] != 0.0) lam(i, j, k, FLO + s) += res[6 + s];
}

// adjoint of one phase of k_fill: ghost adjoints go to the cells the ghost was
// filled from (the last rule that wrote it) and are cleared
template <class P>
__global__ __launch_bounds__(256) void k_fill_adj(AdjFillArgs aa, int phase) {
    constexpr int NV = Num<P>::NV;
    const FillArgs &a = aa.f;
    if (step_skipped(a.ts)) return;
    const Tab &T = *a.tab;
    const View lam = aa.lam;
    const int G = MF_NGROW;
    const int nx = a.n[0], ny = a.n[1], nz = a.n[2];
    const int PX = nx + 2 * G, PY = ny + 2 * G;
    const int64_t nzs = (int64_t)2 * G * PX * PY;
    const int64_t nys = (int64_t)nz * 2 * G * PX;
    const int64_t nxs = (int64_t)nz * ny * 2 * G;
    const int64_t total = nzs + nys + nxs;
    double invelb = 0.0;
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
            for (int c = 0; c < NV; ++c) {
                const double v = lam(iv[0], iv[1], iv[2], c);
                if (v != 0.0) unsafeAtomicAdd(&lam(s[0], s[1], s[2], c), v);
                lam(iv[0], iv[1], iv[2], c) = 0.0;
            }
            continue;
        }
        if (out_np != phase) continue;
        int ext = -1, sgn = 0, bnd = 0, fil = -1;
        for (int d = 0; d < 3; ++d) {
            if (a.periodic[d]) continue;
            if (a.lo_bc[d] == MF_BC_EXT_DIR && iv[d] < 0) { ext = d; sgn = 1; bnd = 0; }
            else if (a.hi_bc[d] == MF_BC_EXT_DIR && iv[d] >= a.n[d]) { ext = d; sgn = -1; bnd = a.n[d] - 1; }
            const int side = iv[d] < 0 ? 0 : (iv[d] >= a.n[d] ? 1 : -1);
            if (side >= 0 && (side == 0 ? a.lo_bc[d] : a.hi_bc[d]) != MF_BC_EXT_DIR) fil = d;
        }
        if (ext >= 0) {
            int loc[3] = {iv[0], iv[1], iv[2]};
            loc[ext] = bnd;
            double s_int[NV], lext[NV], lint[NV];
            for (int c = 0; c < NV; ++c) {
                s_int[c] = a.s(loc[0], loc[1], loc[2], c);
                lext[c] = lam(iv[0], iv[1], iv[2], c);
                lint[c] = 0.0;
            }
            Rev<P>::externalbc_rev(T, s_int, lext, ext, sgn, lint, invelb);
            for (int c = 0; c < NV; ++c) {
                if (lint[c] != 0.0) unsafeAtomicAdd(&lam(loc[0], loc[1], loc[2], c), lint[c]);
                lam(iv[0], iv[1], iv[2], c) = 0.0;
            }
        } else if (fil >= 0) {
            const int d = fil;
            const int side = iv[d] < 0 ? 0 : 1;
            const int bc = side == 0 ? a.lo_bc[d] : a.hi_bc[d];
            int s[3] = {iv[0], iv[1], iv[2]};
            const bool refl = (bc == MF_BC_WALL);
            if (!refl) s[d] = side == 0 ? 0 : a.n[d] - 1;
            else s[d] = side == 0 ? (-iv[d] - 1) : (2 * (a.n[d] - 1) - iv[d] + 1);
            for (int c = 0; c < NV; ++c) {
                const double v = lam(iv[0], iv[1], iv[2], c);
                const bool odd = refl && (c == RHOU || c == RHOV || c == RHOW || c == VELX || c == VELY || c == VELZ);
                if (v != 0.0) unsafeAtomicAdd(&lam(s[0], s[1], s[2], c), odd ? -v : v);
                lam(iv[0], iv[1], iv[2], c) = 0.0;
            }
        }
    }
    if (invelb != 0.0) unsafeAtomicAdd(aa.grad + T.nreact, invelb);
}

// adjoint of the post-step pass (chemistry, update_primitive_vars, clip) with
// the outlet term of the objective of this step added to the adjoint of U^{n+1}
template <class P>
__global__ __launch_bounds__(256, 2) void k_post_adj(AdjPostArgs pa, Work w) {
    constexpr int NV = Num<P>::NV;
    using R = Rev<P>;
    const StageArgs &a = pa.st;
    if (step_skipped(a.ts)) return;
    const Tab &T = *a.tab;
    const double dt = stage_dt(a);
    double gl[MF_MAX_REACTIONS];
#pragma unroll
    for (int r = 0; r < MF_MAX_REACTIONS; ++r) gl[r] = 0.0;
    for (TileIter it(w); it.ok(); it.next()) {
        const int t = it.t;
        int lt;
        const BoxDesc &b = tile_box(w, t, lt);
        int c0[3], c1[3];
        tile_range(b, w, lt, c0, c1);
        const int ex = c1[0] - c0[0] + 1, ey = c1[1] - c0[1] + 1, ez = c1[2] - c0[2] + 1;
        for (int f = threadIdx.x; f < ex * ey * ez; f += blockDim.x) {
            const int i = c0[0] + f % ex, rr = f / ex, j = c0[1] + rr % ey, k = c0[2] + rr / ey;
            double S3[NV], lb[NV];
#pragma unroll
            for (int n = 0; n < NV; ++n) { S3[n] = a.src(i, j, k, n); lb[n] = pa.lu(i, j, k, n); }
            if (i == pa.outlet_i) {
                double So[NV];
#pragma unroll
                for (int n = 0; n < NV; ++n) So[n] = S3[n];
                R::post_cell(T, So, dt, pa.clip);
                const double fct = dt * pa.dA * T.molwt[pa.spec];
                double c = 0.0;
#pragma unroll
                for (int s = 0; s < P::NS; ++s) c = (s == pa.spec) ? So[FLO + s] : c;
                const double rho = So[RHO], ru = So[RHOU];
#pragma unroll
                for (int s = 0; s < P::NS; ++s)
                    if (s == pa.spec) lb[FLO + s] += fct * ru / rho;
                lb[RHOU] += fct * c / rho;
                lb[RHO] += -fct * c * ru / (rho * rho);
            }
            R::post_rev(T, S3, dt, pa.clip, lb, gl);
#pragma unroll
            for (int n = 0; n < NV; ++n) pa.ld(i, j, k, n) = lb[n];
        }
    }
#pragma unroll
    for (int r = 0; r < MF_MAX_REACTIONS; ++r) {
        if (r >= T.nreact) break;
        double v = gl[r];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if ((threadIdx.x & 63) == 0 && v != 0.0) unsafeAtomicAdd(pa.grad + r, v);
    }
}

// objective of one step: J += dt * dA * sum over the outlet layer of M_s c_s (rho u)/rho
template <class P>
__global__ __launch_bounds__(1024) void k_objective(ObjArgs o) {
    if (o.ts->err) return;
    __shared__ double red[1024];
    const Tab &T = *o.tab;
    double acc = 0.0;
    const int m = o.ny * o.nz;
    for (int q = threadIdx.x; q < m; q += blockDim.x) {
        const int j = q % o.ny, k = q / o.ny;
        const double c = o.s(o.outlet_i, j, k, FLO + o.spec), rho = o.s(o.outlet_i, j, k, RHO), ru = o.s(o.outlet_i, j, k, RHOU);
        acc += T.molwt[o.spec] * c * ru / rho;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) o.J[0] += o.ts->dt * o.dA * red[0];
}

template <class P>
struct AdjLaunch {
    static size_t lds_bytes(const int *tsh, int dir) {
        int e[3] = {tsh[0], tsh[1], tsh[2]};
        e[dir] += 6;
        return (size_t)e[0] * e[1] * e[2] * AdjRec<Num<P>::NC>::AS * sizeof(double);
    }
    static void faces(const AdjFaceArgs &a, const Work &w, int dir, int grid, int block, size_t lds, hipStream_t st) {
        static bool attr = false;
        if (!attr) {
            hipFuncSetAttribute((const void *)k_faces_adj<P, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            hipFuncSetAttribute((const void *)k_faces_adj<P, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            hipFuncSetAttribute((const void *)k_faces_adj<P, 2>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            attr = true;
        }
        if (dir == 0) hipLaunchKernelGGL((k_faces_adj<P, 0>), dim3(grid), dim3(block), lds, st, a, w);
        else if (dir == 1) hipLaunchKernelGGL((k_faces_adj<P, 1>), dim3(grid), dim3(block), lds, st, a, w);
        else hipLaunchKernelGGL((k_faces_adj<P, 2>), dim3(grid), dim3(block), lds, st, a, w);
    }
    // grids rounded up to a multiple of 8 blocks for the XCD-chunked mapping
    static int cut_grid(int n) {
        int g = (n + 127) / 128;
        if (g >= 8) g = (g + 7) & ~7;
        return g;
    }
    static void cut(const AdjCutArgs &a, hipStream_t st) {
        if (a.c.ncut <= 0) return;
        hipLaunchKernelGGL(k_cut_coef<P>, dim3(cut_grid(a.c.ncut)), dim3(128), 0, st, a);
        if (a.nnbr > 0) hipLaunchKernelGGL(k_cut_gather<P>, dim3(cut_grid(a.nnbr)), dim3(128), 0, st, a);
    }
    static void fill(const AdjFillArgs &a, int phase, hipStream_t st) {
        const int G = MF_NGROW;
        int64_t PX = a.f.n[0] + 2 * G, PY = a.f.n[1] + 2 * G;
        int64_t total = 2 * G * PX * PY + (int64_t)a.f.n[2] * 2 * G * PX + (int64_t)a.f.n[2] * a.f.n[1] * 2 * G;
        int64_t grid = (total + 255) / 256;
        if (grid > 65535) grid = 65535;
        if (grid < 1) grid = 1;
        hipLaunchKernelGGL(k_fill_adj<P>, dim3((unsigned)grid), dim3(256), 0, st, a, phase);
    }
    static void post(const AdjPostArgs &a, const Work &w, int grid, int block, hipStream_t st) {
        hipLaunchKernelGGL(k_post_adj<P>, dim3(grid), dim3(block), 0, st, a, w);
    }
    static void objective(const ObjArgs &o, hipStream_t st) {
        hipLaunchKernelGGL(k_objective<P>, dim3(1), dim3(1024), 0, st, o);
    }
    static AdjSet set() {
        AdjSet s;
        s.faces = &faces;
        s.lds_bytes = &lds_bytes;
        s.cut = &cut;
        s.fill = &fill;
        s.post = &post;
        s.objective = &objective;
        s.cut_rec = Rev<P>::CUTREC;
        return s;
    }
};

}  // namespace mfd
