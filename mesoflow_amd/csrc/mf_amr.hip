// mf_amr.hip -- two-level coupling operators on device views.
//
// The AMReX calls with which mesoflow couples a fine level to its coarse
// level, as per-box device operations of a context (all nv components):
//   average_down            AverageDownTo (Source/mflo_gridwork.cpp:125-130)
//   time interpolation      GetData + FillPatchSingleLevel (Source/mflo.cpp:310-330,
//                           Source/mflo_gridwork.cpp:147-165)
//   cell_cons_interp        the mapper of FillPatchTwoLevels / FillCoarsePatch
//                           (Source/mflo_gridwork.cpp:151-165, 170-193)
//   FluxRegister            CrseInit / FineAdd (Source/mflo_evolve.cpp:655-681),
//                           Reflux (Source/mflo_evolve.cpp:243-246)
//   tagging                 ErrorEst (Source/mflo.cpp:183-260, Source/mflo_tagging.H)
// One thread per output value (cell or face, component); every product that
// feeds an add is rounded on its own (no FMA), so the results equal the CPU
// restatement's bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "mf_args.h"

using namespace mfd;

int mf_ctx_stream(void *ctx, hipStream_t *stream, int *nv);  // mf_capi.hip
int mf_ctx_geom(void *ctx, mf_geom *geom);                   // mf_capi.hip
int mf_ctx_aux(void *ctx, size_t bytes, double **p);         // mf_capi.hip
int mf_ctx_fill_physbc(void *ctx, mf_array4 view, mf_box region, double time);  // mf_capi.hip

namespace {

constexpr int kBlock = 256;
constexpr int kMaxGrid = 8192;

// the product rounded on its own: the empty asm hides it from FMA contraction
__device__ __forceinline__ double rn(double p) {
    asm volatile("" : "+v"(p));
    return p;
}

__device__ __forceinline__ int floor_div(int a, int r) { return (a >= 0) ? a / r : -((-a + r - 1) / r); }

__device__ __forceinline__ double mc_slope(double um, double u0, double up) {
    const double dc = 0.5 * (up - um);
    const double df = 2.0 * (up - u0);
    const double db = 2.0 * (u0 - um);
    const double lim = (df * db >= 0.0) ? fmin(fabs(df), fabs(db)) : 0.0;
    return copysign(1.0, dc) * fmin(lim, fabs(dc));
}

struct Ext {  // a box as origin + extents, and a flat index -> (i, j, k, n)
    int lo[3], e[3];
    int64_t cells;
    __device__ void at(int64_t q, int &i, int &j, int &k, int &n) const {
        int64_t r = q;
        i = lo[0] + (int)(r % e[0]);
        r /= e[0];
        j = lo[1] + (int)(r % e[1]);
        r /= e[1];
        k = lo[2] + (int)(r % e[2]);
        n = (int)(r / e[2]);
    }
};

Ext ext_of(const mf_box &b) {
    Ext x;
    x.cells = 1;
    for (int d = 0; d < 3; ++d) {
        x.lo[d] = b.lo[d];
        x.e[d] = b.hi[d] - b.lo[d] + 1;
        x.cells *= x.e[d];
    }
    return x;
}

View view_of(const mf_array4 &a) {
    View v;
    v.p = a.p;
    v.js = a.jstride;
    v.ks = a.kstride;
    v.ns = a.nstride;
    v.b0 = a.begin[0];
    v.b1 = a.begin[1];
    v.b2 = a.begin[2];
    return v;
}

bool holds(const mf_array4 &a, const int lo[3], const int hi[3], int nc) {
    if (!a.p || a.ncomp < nc) return false;
    for (int d = 0; d < 3; ++d)
        if (hi[d] < lo[d] || lo[d] < a.begin[d] || hi[d] >= a.end[d]) return false;
    return true;
}

bool valid_box(const mf_box &b) {
    for (int d = 0; d < 3; ++d)
        if (b.hi[d] < b.lo[d]) return false;
    return true;
}

int grid_for(int64_t n) { return (int)std::min<int64_t>((n + kBlock - 1) / kBlock, kMaxGrid); }

__global__ __launch_bounds__(kBlock) void k_average_down(Ext cb, int nv, View fine, View crse, int r, double volfrac) {
    const int64_t tot = cb.cells * nv;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < tot; q += (int64_t)gridDim.x * kBlock) {
        int i, j, k, n;
        cb.at(q, i, j, k, n);
        double s = 0.0;
        for (int kr = 0; kr < r; ++kr)
            for (int jr = 0; jr < r; ++jr)
                for (int ir = 0; ir < r; ++ir) s += fine(i * r + ir, j * r + jr, k * r + kr, n);
        crse(i, j, k, n) = s * volfrac;
    }
}

__global__ __launch_bounds__(kBlock) void k_time_interp(Ext b, int nv, View dst, View so, View sn, int mode, double alpha,
                                                        double beta) {
    const int64_t tot = b.cells * nv;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < tot; q += (int64_t)gridDim.x * kBlock) {
        int i, j, k, n;
        b.at(q, i, j, k, n);
        double v;
        if (mode == 1) v = sn(i, j, k, n);
        else if (mode == 0) v = so(i, j, k, n);
        else v = rn(alpha * so(i, j, k, n)) + rn(beta * sn(i, j, k, n));
        dst(i, j, k, n) = v;
    }
}

// one thread per coarse cell and component: the limited slopes once, then the
// r^3 children that lie in the fine box fb
__global__ __launch_bounds__(kBlock) void k_interp_cons(Ext cb, int nv, View c, View f, int r, double dfac, int flo0,
                                                        int flo1, int flo2, int fhi0, int fhi1, int fhi2) {
    const int64_t tot = cb.cells * nv;
    const double rr = (double)r;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < tot; q += (int64_t)gridDim.x * kBlock) {
        int ic, jc, kc, n;
        cb.at(q, ic, jc, kc, n);
        const double u = c(ic, jc, kc, n);
        double sx = mc_slope(c(ic - 1, jc, kc, n), u, c(ic + 1, jc, kc, n));
        double sy = mc_slope(c(ic, jc - 1, kc, n), u, c(ic, jc + 1, kc, n));
        double sz = mc_slope(c(ic, jc, kc - 1, n), u, c(ic, jc, kc + 1, n));
        double alpha = 1.0;
        if (sx != 0.0 || sy != 0.0 || sz != 0.0) {
            const double dumax = (rn(fabs(sx) * dfac) + rn(fabs(sy) * dfac)) + rn(fabs(sz) * dfac);
            double umax = u, umin = u;
            for (int ko = -1; ko <= 1; ++ko)
                for (int jo = -1; jo <= 1; ++jo)
                    for (int io = -1; io <= 1; ++io) {
                        const double v = c(ic + io, jc + jo, kc + ko, n);
                        umax = fmax(umax, v);
                        umin = fmin(umin, v);
                    }
            if (rn(dumax * alpha) > (umax - u)) alpha = (umax - u) / dumax;
            if (rn(dumax * alpha) > (u - umin)) alpha = (u - umin) / dumax;
        }
        sx *= alpha;
        sy *= alpha;
        sz *= alpha;
        for (int kr = 0; kr < r; ++kr) {
            const int k = kc * r + kr;
            if (k < flo2 || k > fhi2) continue;
            const double zt = rn((((double)kr + 0.5) / rr - 0.5) * sz);
            for (int jr = 0; jr < r; ++jr) {
                const int j = jc * r + jr;
                if (j < flo1 || j > fhi1) continue;
                const double yt = rn((((double)jr + 0.5) / rr - 0.5) * sy);
                for (int ir = 0; ir < r; ++ir) {
                    const int i = ic * r + ir;
                    if (i < flo0 || i > fhi0) continue;
                    const double xt = rn((((double)ir + 0.5) / rr - 0.5) * sx);
                    f(i, j, k, n) = ((u + xt) + yt) + zt;
                }
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_crse_add(Ext fb, int nv, View reg, View fl, double scale) {
    const int64_t tot = fb.cells * nv;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < tot; q += (int64_t)gridDim.x * kBlock) {
        int i, j, k, n;
        fb.at(q, i, j, k, n);
        reg(i, j, k, n) = reg(i, j, k, n) + rn(scale * fl(i, j, k, n));
    }
}

__global__ __launch_bounds__(kBlock) void k_fine_add(Ext fb, int nv, View reg, View fl, int dir, int r, double scale) {
    const int64_t tot = fb.cells * nv;
    const int t1 = (dir == 0) ? 1 : 0, t2 = (dir == 2) ? 1 : 2;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < tot; q += (int64_t)gridDim.x * kBlock) {
        int i, j, k, n;
        fb.at(q, i, j, k, n);
        const int c[3] = {i, j, k};
        double acc = reg(i, j, k, n);
        for (int b = 0; b < r; ++b)
            for (int a = 0; a < r; ++a) {
                int p[3];
                p[dir] = c[dir] * r;
                p[t1] = c[t1] * r + a;
                p[t2] = c[t2] * r + b;
                acc = acc + rn(scale * fl(p[0], p[1], p[2], n));
            }
        reg(i, j, k, n) = acc;
    }
}

__global__ __launch_bounds__(kBlock) void k_reflux(Ext fb, int nv, View s, View reg, int dir, int side, double mult,
                                                   double vol) {
    const int64_t tot = fb.cells * nv;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < tot; q += (int64_t)gridDim.x * kBlock) {
        int i, j, k, n;
        fb.at(q, i, j, k, n);
        const int sh = side == 0 ? 1 : 0;
        const int ci = i - (dir == 0 ? sh : 0), cj = j - (dir == 1 ? sh : 0), ck = k - (dir == 2 ? sh : 0);
        s(ci, cj, ck, n) = s(ci, cj, ck, n) + (mult * reg(i, j, k, n)) / vol;
    }
}

// dst = src shifted by sh (a periodic image) over the region
__global__ __launch_bounds__(kBlock) void k_copy_shift(Ext b, int nv, View dst, View src, int sh0, int sh1, int sh2) {
    const int64_t tot = b.cells * nv;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < tot; q += (int64_t)gridDim.x * kBlock) {
        int i, j, k, n;
        b.at(q, i, j, k, n);
        dst(i, j, k, n) = src(i + sh0, j + sh1, k + sh2, n);
    }
}

constexpr int kMaxTagVars = 16;
struct TagVars {
    int n;
    int comp[kMaxTagVars];
    double phi[kMaxTagVars], grad[kMaxTagVars];
};

// ErrorEst's two criteria per cell (Source/mflo_tagging.H:6-70); tags dense over bx
__global__ __launch_bounds__(kBlock) void k_tag(Ext b, View s, unsigned char *tags, TagVars tv, unsigned char tagval) {
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < b.cells; q += (int64_t)gridDim.x * kBlock) {
        int i, j, k, n0;
        b.at(q, i, j, k, n0);
        bool set = false;
        for (int c = 0; c < tv.n; ++c)
            if (fabs(s(i, j, k, tv.comp[c])) > tv.phi[c]) set = true;
        for (int c = 0; c < tv.n; ++c) {
            const int n = tv.comp[c];
            const double u = s(i, j, k, n);
            // amrex::max is std::max ((a < b) ? b : a: a NaN first argument
            // propagates, a NaN second argument is dropped), and
            // amrex::max(gx, gy, gz) = max(max(gx, gy), gz)
            const double gx = std_max(fabs(s(i + 1, j, k, n) - u), fabs(u - s(i - 1, j, k, n)));
            const double gy = std_max(fabs(s(i, j + 1, k, n) - u), fabs(u - s(i, j - 1, k, n)));
            const double gz = std_max(fabs(s(i, j, k + 1, n) - u), fabs(u - s(i, j, k - 1, n)));
            if (std_max(std_max(gx, gy), gz) > tv.grad[c]) set = true;
        }
        if (set) tags[q] = tagval;
    }
}

int tag_vars(const double *phi, const double *grad, const int *comps, int ntag, int nv, TagVars &tv) {
    if (ntag < 0 || ntag > kMaxTagVars || (ntag > 0 && (!phi || !grad || !comps))) return MF_ERR_BAD_ARG;
    tv.n = ntag;
    for (int c = 0; c < ntag; ++c) {
        if (comps[c] < 0 || comps[c] >= nv) return MF_ERR_BAD_ARG;
        tv.comp[c] = comps[c];
        tv.phi[c] = phi[c];
        tv.grad[c] = grad[c];
    }
    return MF_OK;
}

}  // namespace

int mf_level_ctx(void *level, void **ctx);  // mf_capi.hip

extern "C" {

int mf_fill_patch_two_levels(void *ctx, mf_box fbx, mf_array4 sborder, const mf_box *fboxes, const mf_array4 *fdata,
                             int nfine, mf_array4 crse_old, mf_array4 crse_new, double t_old, double t_new, double time,
                             int ratio) {
    hipStream_t st;
    int nv;
    if (!ctx || ratio < 1 || ratio > 8 || !valid_box(fbx) || nfine < 0 || (nfine > 0 && (!fboxes || !fdata)))
        return MF_ERR_BAD_ARG;
    int rc = mf_ctx_stream(ctx, &st, &nv);
    if (rc) return rc;
    mf_geom g;
    rc = mf_ctx_geom(ctx, &g);
    if (rc) return rc;
    mf_box gb;  // the region filled: fbx grown by the ghost width
    for (int d = 0; d < 3; ++d) {
        gb.lo[d] = fbx.lo[d] - MF_NGROW;
        gb.hi[d] = fbx.hi[d] + MF_NGROW;
        if (fbx.lo[d] < 0 || fbx.hi[d] >= g.n_cell[d]) return MF_ERR_BAD_ARG;
        if (g.n_cell[d] % ratio) return MF_ERR_BAD_ARG;
    }
    if (!holds(sborder, gb.lo, gb.hi, nv)) return MF_ERR_BAD_ARG;
    for (int b = 0; b < nfine; ++b)
        if (!valid_box(fboxes[b]) || !holds(fdata[b], fboxes[b].lo, fboxes[b].hi, nv)) return MF_ERR_BAD_ARG;
    // 1. coarse data at the fill time on coarsen(gb) grown by one
    mf_box cb;
    for (int d = 0; d < 3; ++d) {
        cb.lo[d] = ((gb.lo[d] >= 0) ? gb.lo[d] / ratio : -((-gb.lo[d] + ratio - 1) / ratio)) - 1;
        cb.hi[d] = ((gb.hi[d] >= 0) ? gb.hi[d] / ratio : -((-gb.hi[d] + ratio - 1) / ratio)) + 1;
    }
    if (!holds(crse_old, cb.lo, cb.hi, nv) || !holds(crse_new, cb.lo, cb.hi, nv)) return MF_ERR_BAD_ARG;
    const Ext ce = ext_of(cb);
    double *tmp = nullptr;
    rc = mf_ctx_aux(ctx, (size_t)ce.cells * nv * sizeof(double), &tmp);
    if (rc) return rc;
    mf_array4 ct{};
    ct.p = tmp;
    ct.jstride = ce.e[0];
    ct.kstride = (int64_t)ce.e[0] * ce.e[1];
    ct.nstride = ce.cells;
    for (int d = 0; d < 3; ++d) { ct.begin[d] = cb.lo[d]; ct.end[d] = cb.hi[d] + 1; }
    ct.ncomp = nv;
    rc = mf_box_time_interp(ctx, cb, ct, crse_old, crse_new, t_old, t_new, time);
    // 2. the conservative interpolation over the whole region
    if (!rc) rc = mf_box_interp_cons(ctx, gb, ct, sborder, ratio);
    // 3. fine data wherever a fine box (or its periodic image) covers the region
    for (int b = 0; !rc && b < nfine; ++b) {
        int span[3][2];
        for (int d = 0; d < 3; ++d) {
            span[d][0] = g.is_periodic[d] ? -1 : 0;
            span[d][1] = g.is_periodic[d] ? 1 : 0;
        }
        for (int pk = span[2][0]; pk <= span[2][1]; ++pk)
            for (int pj = span[1][0]; pj <= span[1][1]; ++pj)
                for (int pi = span[0][0]; pi <= span[0][1]; ++pi) {
                    const int sh[3] = {pi * g.n_cell[0], pj * g.n_cell[1], pk * g.n_cell[2]};
                    mf_box x;
                    bool empty = false;
                    for (int d = 0; d < 3; ++d) {  // region cells whose shifted position lies in fine box b
                        x.lo[d] = std::max(gb.lo[d], fboxes[b].lo[d] + sh[d]);
                        x.hi[d] = std::min(gb.hi[d], fboxes[b].hi[d] + sh[d]);
                        if (x.hi[d] < x.lo[d]) empty = true;
                    }
                    if (empty) continue;
                    const Ext xe = ext_of(x);
                    hipLaunchKernelGGL(k_copy_shift, dim3(grid_for(xe.cells * nv)), dim3(kBlock), 0, st, xe, nv,
                                       view_of(sborder), view_of(fdata[b]), -sh[0], -sh[1], -sh[2]);
                    if (hipGetLastError() != hipSuccess) rc = MF_ERR_HIP;
                }
    }
    // 4. the fine level's physical boundary conditions where the region leaves
    // the domain in a non-periodic direction
    if (!rc) rc = mf_ctx_fill_physbc(ctx, sborder, gb, time);
    return rc;
}

int mf_box_tag(void *ctx, mf_box bx, mf_array4 state, unsigned char *tags, const double *refine_phi,
               const double *refine_phigrad, const int *comps, int ntag, int tagval) {
    hipStream_t st;
    int nv;
    if (!ctx || !tags || !valid_box(bx)) return MF_ERR_BAD_ARG;
    int rc = mf_ctx_stream(ctx, &st, &nv);
    if (rc) return rc;
    TagVars tv;
    rc = tag_vars(refine_phi, refine_phigrad, comps, ntag, nv, tv);
    if (rc) return rc;
    int glo[3], ghi[3];
    for (int d = 0; d < 3; ++d) { glo[d] = bx.lo[d] - 1; ghi[d] = bx.hi[d] + 1; }
    if (!holds(state, glo, ghi, nv)) return MF_ERR_BAD_ARG;
    if (ntag == 0) return MF_OK;
    const Ext e = ext_of(bx);
    hipLaunchKernelGGL(k_tag, dim3(grid_for(e.cells)), dim3(kBlock), 0, st, e, view_of(state), tags, tv,
                       (unsigned char)tagval);
    return hipGetLastError() == hipSuccess ? MF_OK : MF_ERR_HIP;
}

int mf_level_tag(void *level, double time, const double *refine_phi, const double *refine_phigrad, const int *comps,
                 int ntag, unsigned char *tags, int64_t *ntagged) {
    if (!level || !tags) return MF_ERR_BAD_ARG;
    mf_array4 v;
    mf_geom g;
    int rc = mf_level_state_view(level, &v);
    if (!rc) rc = mf_level_boxes(level, nullptr, 0, nullptr, &g, nullptr);
    if (!rc) rc = mf_fill_ghosts_single_level(level, time);
    if (rc) return rc;
    mf_box dom;
    for (int d = 0; d < 3; ++d) { dom.lo[d] = 0; dom.hi[d] = g.n_cell[d] - 1; }
    const int64_t ncell = (int64_t)g.n_cell[0] * g.n_cell[1] * g.n_cell[2];
    hipStream_t st = nullptr;
    int nv = 0;
    void *ctx = nullptr;
    rc = mf_level_ctx(level, &ctx);
    if (!rc) rc = mf_ctx_stream(ctx, &st, &nv);
    if (rc) return rc;
    unsigned char *dt = nullptr;
    if (hipMalloc(&dt, ncell) != hipSuccess) return MF_ERR_HIP;
    if (hipMemsetAsync(dt, 0, ncell, st) != hipSuccess) rc = MF_ERR_HIP;
    if (!rc) rc = mf_box_tag(ctx, dom, v, dt, refine_phi, refine_phigrad, comps, ntag, 1);
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = MF_ERR_HIP;
    if (!rc && hipMemcpy(tags, dt, ncell, hipMemcpyDeviceToHost) != hipSuccess) rc = MF_ERR_HIP;
    hipFree(dt);
    if (rc) return rc;
    if (ntagged) {
        int64_t cnt = 0;
        for (int64_t q = 0; q < ncell; ++q) cnt += tags[q] != 0;
        *ntagged = cnt;
    }
    return MF_OK;
}

int mf_box_average_down(void *ctx, mf_box cbx, mf_array4 fine, mf_array4 crse, int ratio) {
    hipStream_t st;
    int nv;
    if (!ctx || ratio < 1 || ratio > 8 || !valid_box(cbx)) return MF_ERR_BAD_ARG;
    int rc = mf_ctx_stream(ctx, &st, &nv);
    if (rc) return rc;
    int flo[3], fhi[3];
    for (int d = 0; d < 3; ++d) { flo[d] = cbx.lo[d] * ratio; fhi[d] = cbx.hi[d] * ratio + ratio - 1; }
    if (!holds(crse, cbx.lo, cbx.hi, nv) || !holds(fine, flo, fhi, nv)) return MF_ERR_BAD_ARG;
    const Ext e = ext_of(cbx);
    hipLaunchKernelGGL(k_average_down, dim3(grid_for(e.cells * nv)), dim3(kBlock), 0, st, e, nv, view_of(fine),
                       view_of(crse), ratio, 1.0 / (double)(ratio * ratio * ratio));
    return hipGetLastError() == hipSuccess ? MF_OK : MF_ERR_HIP;
}

int mf_box_time_interp(void *ctx, mf_box bx, mf_array4 dst, mf_array4 s_old, mf_array4 s_new, double t_old,
                       double t_new, double time) {
    hipStream_t st;
    int nv;
    if (!ctx || !valid_box(bx)) return MF_ERR_BAD_ARG;
    int rc = mf_ctx_stream(ctx, &st, &nv);
    if (rc) return rc;
    if (!holds(dst, bx.lo, bx.hi, nv) || !holds(s_old, bx.lo, bx.hi, nv) || !holds(s_new, bx.lo, bx.hi, nv))
        return MF_ERR_BAD_ARG;
    const double teps = (t_new - t_old) * 1.e-3;
    int mode = 2;
    if (time > t_new - teps && time < t_new + teps) mode = 1;
    else if (time > t_old - teps && time < t_old + teps) mode = 0;
    const double alpha = (mode == 2) ? (t_new - time) / (t_new - t_old) : 0.0;
    const double beta = (mode == 2) ? (time - t_old) / (t_new - t_old) : 0.0;
    const Ext e = ext_of(bx);
    hipLaunchKernelGGL(k_time_interp, dim3(grid_for(e.cells * nv)), dim3(kBlock), 0, st, e, nv, view_of(dst),
                       view_of(s_old), view_of(s_new), mode, alpha, beta);
    return hipGetLastError() == hipSuccess ? MF_OK : MF_ERR_HIP;
}

int mf_box_interp_cons(void *ctx, mf_box fbx, mf_array4 crse, mf_array4 fine, int ratio) {
    hipStream_t st;
    int nv;
    if (!ctx || ratio < 1 || ratio > 8 || !valid_box(fbx)) return MF_ERR_BAD_ARG;
    int rc = mf_ctx_stream(ctx, &st, &nv);
    if (rc) return rc;
    int clo[3], chi[3];
    for (int d = 0; d < 3; ++d) {
        const int l = fbx.lo[d], h = fbx.hi[d];
        clo[d] = ((l >= 0) ? l / ratio : -((-l + ratio - 1) / ratio)) - 1;
        chi[d] = ((h >= 0) ? h / ratio : -((-h + ratio - 1) / ratio)) + 1;
    }
    if (!holds(fine, fbx.lo, fbx.hi, nv) || !holds(crse, clo, chi, nv)) return MF_ERR_BAD_ARG;
    mf_box cbx;  // coarsen(fbx)
    for (int d = 0; d < 3; ++d) { cbx.lo[d] = clo[d] + 1; cbx.hi[d] = chi[d] - 1; }
    const Ext e = ext_of(cbx);
    hipLaunchKernelGGL(k_interp_cons, dim3(grid_for(e.cells * nv)), dim3(kBlock), 0, st, e, nv, view_of(crse),
                       view_of(fine), ratio, (double)(ratio - 1) / (double)(2 * ratio), fbx.lo[0], fbx.lo[1],
                       fbx.lo[2], fbx.hi[0], fbx.hi[1], fbx.hi[2]);
    return hipGetLastError() == hipSuccess ? MF_OK : MF_ERR_HIP;
}

int mf_fluxreg_crse_add(void *ctx, int dir, mf_box face, mf_array4 reg, mf_array4 cflux, double scale) {
    hipStream_t st;
    int nv;
    if (!ctx || dir < 0 || dir > 2 || !valid_box(face) || face.lo[dir] != face.hi[dir]) return MF_ERR_BAD_ARG;
    int rc = mf_ctx_stream(ctx, &st, &nv);
    if (rc) return rc;
    if (!holds(reg, face.lo, face.hi, nv) || !holds(cflux, face.lo, face.hi, nv)) return MF_ERR_BAD_ARG;
    const Ext e = ext_of(face);
    hipLaunchKernelGGL(k_crse_add, dim3(grid_for(e.cells * nv)), dim3(kBlock), 0, st, e, nv, view_of(reg),
                       view_of(cflux), scale);
    return hipGetLastError() == hipSuccess ? MF_OK : MF_ERR_HIP;
}

int mf_fluxreg_fine_add(void *ctx, int dir, mf_box face, mf_array4 reg, mf_array4 fflux, int ratio, double scale) {
    hipStream_t st;
    int nv;
    if (!ctx || dir < 0 || dir > 2 || ratio < 1 || ratio > 8 || !valid_box(face) || face.lo[dir] != face.hi[dir])
        return MF_ERR_BAD_ARG;
    int rc = mf_ctx_stream(ctx, &st, &nv);
    if (rc) return rc;
    const int t1 = (dir == 0) ? 1 : 0, t2 = (dir == 2) ? 1 : 2;
    int flo[3], fhi[3];
    flo[dir] = fhi[dir] = face.lo[dir] * ratio;
    flo[t1] = face.lo[t1] * ratio;
    fhi[t1] = face.hi[t1] * ratio + ratio - 1;
    flo[t2] = face.lo[t2] * ratio;
    fhi[t2] = face.hi[t2] * ratio + ratio - 1;
    if (!holds(reg, face.lo, face.hi, nv) || !holds(fflux, flo, fhi, nv)) return MF_ERR_BAD_ARG;
    const Ext e = ext_of(face);
    hipLaunchKernelGGL(k_fine_add, dim3(grid_for(e.cells * nv)), dim3(kBlock), 0, st, e, nv, view_of(reg),
                       view_of(fflux), dir, ratio, scale);
    return hipGetLastError() == hipSuccess ? MF_OK : MF_ERR_HIP;
}

int mf_fluxreg_reflux(void *ctx, int dir, int side, mf_box face, mf_array4 state, mf_array4 reg, double scale,
                      double vol) {
    hipStream_t st;
    int nv;
    if (!ctx || dir < 0 || dir > 2 || (side != 0 && side != 1) || !(vol > 0.0) || !valid_box(face) ||
        face.lo[dir] != face.hi[dir])
        return MF_ERR_BAD_ARG;
    int rc = mf_ctx_stream(ctx, &st, &nv);
    if (rc) return rc;
    mf_box cells = face;
    if (side == 0) { cells.lo[dir] -= 1; cells.hi[dir] -= 1; }
    if (!holds(reg, face.lo, face.hi, nv) || !holds(state, cells.lo, cells.hi, nv)) return MF_ERR_BAD_ARG;
    const Ext e = ext_of(face);
    hipLaunchKernelGGL(k_reflux, dim3(grid_for(e.cells * nv)), dim3(kBlock), 0, st, e, nv, view_of(state),
                       view_of(reg), dir, side, side == 0 ? -scale : scale, vol);
    return hipGetLastError() == hipSuccess ? MF_OK : MF_ERR_HIP;
}

}  // extern "C"
