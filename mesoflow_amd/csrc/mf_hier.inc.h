// mf_hier.inc.h -- the two-level driver; part of mf_capi.hip's translation
// unit (it drives the Level objects defined there).
//
// mflo::EvolveAMR over levels 0 and 1 with subcycling (Source/mflo_evolve.cpp:
// 18-89), timeStep's recursion (:175-249), Advance_coupled_strang per level
// (:251-318) with the flux-register updates of compute_dsdt_flow (:655-681),
// FillPatch of level 1 = FillPatchTwoLevels (Source/mflo_gridwork.cpp:147-165)
// with GetData's time selection (Source/mflo.cpp:310-330), ComputeDt over the
// levels (Source/mflo_timestep.cpp:18-50), Reflux and AverageDownTo
// (Source/mflo_evolve.cpp:243-248, Source/mflo_gridwork.cpp:125-130), and the
// regrid every regrid_int coarse steps (Source/mflo_evolve.cpp:177-209) with
// RemakeLevel / MakeNewLevelFromCoarse / ClearLevel (Source/mflo_gridwork.cpp:
// 21-89).
//
// Layout: each level keeps its three RK stage arrays padded over its whole
// domain; the fine level's boxes cover part of it. A fine ghost cell has one
// value whichever fine box reads it, so the fine ghost cells are filled once
// per stage, in place, from cell lists built when the boxes change:
//   coarse patch   the coarse cells that the interpolation reads,
//                  time-interpolated into a coarse-shaped array whose ghost
//                  layers then get the periodic images and the physical
//                  boundary rules at the fill time (FillPatchSingleLevel);
//   interpolation  cell_cons_interp for the fine ghost cells no fine box
//                  covers (periodic images included);
//   images         copies of fine valid cells into the ghost cells that are
//                  their periodic images;
//   boundaries     the fine physical boundary rules on the ghost cells outside
//                  the domain, phase by phase.
// The flux register holds the coarse faces of the coarsened fine boxes whose
// outside coarse cell is valid and uncovered; CrseInit reads the coarse flux
// of the box owning that cell; a refluxed cell takes its faces' corrections
// in the order x-lo, x-hi, y-lo, y-hi, z-lo, z-hi. Times, dt and flags stay in
// device memory: a whole coarse step is one launch sequence (one persistent
// launch per level, substep and stage for the face sweeps, the residual and
// the passes), which a hipGraph replays. Every product that feeds an add in
// the coupling arithmetic is rounded on its own, as on the CPU.

#include "mf_io.h"

namespace {

constexpr int kHierMaxLev = 2;

// t_old / t_new of each level (GetData), the time argument of the level's
// advance in flight and dt[lev] of the coarse step in flight
struct HierTime {
    double t_old[kHierMaxLev], t_new[kHierMaxLev];
    double time[kHierMaxLev];
    double dt[kHierMaxLev];
};

// one coarse face of the flux register
struct RegDev {
    int dir, side;            // side 0: the fine region lies above the face (reflux with -1)
    int64_t cofs, cns;        // coarse flux: element offset of component 0, component stride
    int64_t fofs, fns;        // first of the r^2 fine faces: element offset, component stride
    int64_t fs1, fs2;         // fine-face strides along the two transverse directions
};
struct TargetDev {            // a refluxed coarse cell and its register faces (-1: none)
    int c[3];
    int e[6];
};

__device__ __forceinline__ double hrn(double p) {
    asm volatile("" : "+v"(p));
    return p;
}
__device__ __forceinline__ int hfdiv(int a, int r) { return (a >= 0) ? a / r : -((-a + r - 1) / r); }
__device__ __forceinline__ double hmc_slope(double um, double u0, double up) {
    const double dc = 0.5 * (up - um);
    const double df = 2.0 * (up - u0);
    const double db = 2.0 * (u0 - um);
    const double lim = (df * db >= 0.0) ? fmin(fabs(df), fabs(db)) : 0.0;
    return copysign(1.0, dc) * fmin(lim, fabs(dc));
}
// a step that EvolveAMR's stop conditions or an error ended (ts0 null: a
// regrid's fill, never skipped)
__device__ __forceinline__ bool hier_skip(const TimeState *ts0) { return ts0 && (!ts0->active || ts0->err); }

// ComputeDt over the two levels (Source/mflo_timestep.cpp:18-50) with
// EstTimeStep's formula (:132-152) on each level's wave speed and diffusivity
struct HierDtArgs {
    TimeState *ts[kHierMaxLev];
    HierTime *ht;
    int nlev;
    int nsub[kHierMaxLev];
    double dx[kHierMaxLev][3];
    double cfl, dt_fixed;
    int fixed;
};
__global__ void k_hier_dt(HierDtArgs a) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    TimeState *t0 = a.ts[0];
    if (!t0->active || t0->err) {
        for (int l = 1; l < a.nlev; ++l) a.ts[l]->active = 0;
        return;
    }
    if (!(t0->t < t0->stop_time)) {
        for (int l = 0; l < a.nlev; ++l) a.ts[l]->active = 0;
        return;
    }
    const double change_max = 1.1;
    double dt0 = 0.0;
    int n_factor = 1;
    for (int l = 0; l < a.nlev; ++l) {
        const TimeState *ts = a.ts[l];
        double est = DBL_MAX;
        if (ts->max_ws > 0)
            for (int d = 0; d < 3; ++d) est = fmin(est, a.dx[l][d] / ts->max_ws);
        if (ts->max_D > 0)
            for (int d = 0; d < 3; ++d) est = fmin(est, kHalf * a.dx[l][d] * a.dx[l][d] / ts->max_D / 3);
        est *= a.cfl;
        if (a.fixed) est = a.dt_fixed;
        if (l == 0) dt0 = est;
        est = fmin(est, change_max * ts->dt_prev);
        n_factor *= a.nsub[l];
        dt0 = fmin(dt0, n_factor * est);
    }
    if (a.fixed) dt0 = a.dt_fixed;
    const double eps = 1.e-3 * dt0;
    if (t0->t + dt0 > t0->stop_time - eps) dt0 = t0->stop_time - t0->t;
    t0->dt = dt0;
    a.ht->dt[0] = dt0;
    for (int l = 1; l < a.nlev; ++l) {
        const double d = a.ht->dt[l - 1] / a.nsub[l];
        a.ht->dt[l] = d;
        a.ts[l]->dt = d;
        a.ts[l]->active = 1;
        a.ts[l]->err = 0;
    }
}

// the start of a level's advance: its time argument (timeStep: time of level
// 0 = cur_time, substep i of level 1: time + (i-1) dt[1]) and t_old/t_new
// (Advance_coupled_strang :256-257)
__global__ void k_hier_begin(HierTime *ht, TimeState *ts_l, const TimeState *ts0, int lev, int i) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (hier_skip(ts0)) return;
    const double time = lev == 0 ? ts0->t : ht->time[0] + hrn((double)(i - 1) * ht->dt[lev]);
    ht->time[lev] = time;
    ht->t_old[lev] = ht->t_new[lev];
    ht->t_new[lev] = ht->t_new[lev] + ht->dt[lev];
    ts_l->t = time;
    if (lev > 0) ts_l->istep += 1;  // ++istep[lev] of each substep (level 0: k_finalize)
}

__global__ void k_hier_zero(double *p, int64_t n, const TimeState *ts0) {
    if (hier_skip(ts0)) return;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) p[q] = 0.0;
}

// the coarse patch: GetData of level 0 at the fill time and
// FillPatchSingleLevel's time interpolation of the listed (interior) coarse
// cells; mode >= 0 forces old (0) or new (1) data
__global__ void k_hier_ct(View ct, View so, View sn, const int *list, int n, int nv, const HierTime *ht,
                          const TimeState *tsf, int stage, int mode_forced, const TimeState *ts0) {
    if (hier_skip(ts0)) return;
    int mode = mode_forced;
    double alpha = 0.0, beta = 0.0;
    if (mode < 0) {
        const double t_old = ht->t_old[0], t_new = ht->t_new[0];
        const double time = stage == 1 ? tsf->t : tsf->t + tsf->dt;
        const double teps = hrn((t_new - t_old) * 1.e-3);
        mode = 2;
        if (time > t_new - teps && time < t_new + teps) mode = 1;
        else if (time > t_old - teps && time < t_old + teps) mode = 0;
        if (mode == 2) {
            alpha = (t_new - time) / (t_new - t_old);
            beta = (time - t_old) / (t_new - t_old);
        }
    }
    const int64_t tot = (int64_t)n * nv;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int e = (int)(q % n), c = (int)(q / n);
        const int i = list[e], j = list[n + e], k = list[2 * n + e];
        double v;
        if (mode == 1) v = sn(i, j, k, c);
        else if (mode == 0) v = so(i, j, k, c);
        else v = hrn(alpha * so(i, j, k, c)) + hrn(beta * sn(i, j, k, c));
        ct(i, j, k, c) = v;
    }
}

// cell_cons_interp of the listed fine cells from the coarse patch: one thread
// per listed parent coarse cell (list [3][n]) and component forms the limited
// slopes once and writes the children marked in its mask (bit ir + r (jr + r kr))
__global__ void k_hier_cf(View f, View ct, const int *list, const uint64_t *mask, int n, int nv, int r, double dfac,
                          const TimeState *ts0) {
    if (hier_skip(ts0)) return;
    const int64_t tot = (int64_t)n * nv;
    const double rr = (double)r;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int e = (int)(q % n), c = (int)(q / n);
        const int ic = list[e], jc = list[n + e], kc = list[2 * n + e];
        const uint64_t bits = mask[e];
        const double u = ct(ic, jc, kc, c);
        double sx = hmc_slope(ct(ic - 1, jc, kc, c), u, ct(ic + 1, jc, kc, c));
        double sy = hmc_slope(ct(ic, jc - 1, kc, c), u, ct(ic, jc + 1, kc, c));
        double sz = hmc_slope(ct(ic, jc, kc - 1, c), u, ct(ic, jc, kc + 1, c));
        double alpha = 1.0;
        if (sx != 0.0 || sy != 0.0 || sz != 0.0) {
            const double dumax = (hrn(fabs(sx) * dfac) + hrn(fabs(sy) * dfac)) + hrn(fabs(sz) * dfac);
            double umax = u, umin = u;
            for (int ko = -1; ko <= 1; ++ko)
                for (int jo = -1; jo <= 1; ++jo)
                    for (int io = -1; io <= 1; ++io) {
                        const double v = ct(ic + io, jc + jo, kc + ko, c);
                        umax = fmax(umax, v);
                        umin = fmin(umin, v);
                    }
            if (hrn(dumax * alpha) > (umax - u)) alpha = (umax - u) / dumax;
            if (hrn(dumax * alpha) > (u - umin)) alpha = (u - umin) / dumax;
        }
        sx *= alpha;
        sy *= alpha;
        sz *= alpha;
        for (int kr = 0; kr < r; ++kr) {
            const double zt = hrn((((double)kr + 0.5) / rr - 0.5) * sz);
            for (int jr = 0; jr < r; ++jr) {
                const double yt = hrn((((double)jr + 0.5) / rr - 0.5) * sy);
                for (int ir = 0; ir < r; ++ir) {
                    if (!((bits >> (ir + r * (jr + r * kr))) & 1)) continue;
                    const double xt = hrn((((double)ir + 0.5) / rr - 0.5) * sx);
                    f(ic * r + ir, jc * r + jr, kc * r + kr, c) = ((u + xt) + yt) + zt;
                }
            }
        }
    }
}

// fine ghost cells that are periodic images of fine valid cells
__global__ void k_hier_per(View f, const int *list, int n, int nv, int nx, int ny, int nz, const TimeState *ts0) {
    if (hier_skip(ts0)) return;
    const int64_t tot = (int64_t)n * nv;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int e = (int)(q % n), c = (int)(q / n);
        const int i = list[e], j = list[n + e], k = list[2 * n + e];
        const int si = ((i % nx) + nx) % nx, sj = ((j % ny) + ny) % ny, sk = ((k % nz) + nz) % nz;
        f(i, j, k, c) = f(si, sj, sk, c);
    }
}

// FluxRegister::CrseInit(flux, d, 0, 0, ncomp, -dt dA w, ADD) over the
// register faces (components: the 5 + NS flux components)
__global__ void k_hier_crse_init(double *regv, const RegDev *reg, int nreg, int nc, const double *cflux, double dA0,
                                 double dA1, double dA2, double w, const TimeState *tsc, const TimeState *ts0) {
    if (hier_skip(ts0)) return;
    const double dt = tsc->dt;
    const int64_t tot = (int64_t)nreg * nc;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int e = (int)(q % nreg), c = (int)(q / nreg);
        const RegDev g = reg[e];
        const double dA = g.dir == 0 ? dA0 : (g.dir == 1 ? dA1 : dA2);
        const double scale = hrn(hrn(-dt * dA) * w);
        double *rp = regv + (int64_t)c * nreg + e;
        *rp = *rp + hrn(scale * cflux[g.cofs + (int64_t)c * g.cns]);
    }
}

// FluxRegister::FineAdd(flux, d, 0, 0, ncomp, dt dA w): the r^2 fine faces of
// each register face, second transverse index outer
__global__ void k_hier_fine_add(double *regv, const RegDev *reg, int nreg, int nc, const double *fflux, int r,
                                double dA0, double dA1, double dA2, double w, const TimeState *tsf,
                                const TimeState *ts0) {
    if (hier_skip(ts0)) return;
    const double dt = tsf->dt;
    const int64_t tot = (int64_t)nreg * nc;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int e = (int)(q % nreg), c = (int)(q / nreg);
        const RegDev g = reg[e];
        const double dA = g.dir == 0 ? dA0 : (g.dir == 1 ? dA1 : dA2);
        const double scale = hrn(hrn(dt * dA) * w);
        double *rp = regv + (int64_t)c * nreg + e;
        double acc = *rp;
        const double *fp = fflux + g.fofs + (int64_t)c * g.fns;
        for (int b = 0; b < r; ++b)
            for (int a = 0; a < r; ++a) acc = acc + hrn(scale * fp[(int64_t)a * g.fs1 + (int64_t)b * g.fs2]);
        *rp = acc;
    }
}

// FluxRegister::Reflux(phi_new[0], 1.0, 0, 0, ncomp, geom[0]) on the
// refluxed coarse cells
__global__ void k_hier_reflux(View s, const TargetDev *tg, int ntg, const RegDev *reg, const double *regv, int nreg,
                              int nc, double vol, const TimeState *ts0) {
    if (hier_skip(ts0)) return;
    const int64_t tot = (int64_t)ntg * nc;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int t = (int)(q % ntg), c = (int)(q / ntg);
        const TargetDev g = tg[t];
        const int n = c < NCV ? c : c + (FLO - NCV);
        double v = s(g.c[0], g.c[1], g.c[2], n);
        for (int sl = 0; sl < 6; ++sl) {
            const int e = g.e[sl];
            if (e < 0) continue;
            const double mult = reg[e].side == 0 ? -1.0 : 1.0;
            v = v + (mult * regv[(int64_t)c * nreg + e]) / vol;
        }
        s(g.c[0], g.c[1], g.c[2], n) = v;
    }
}

// amrex::average_down over the listed coarse cells (k outer, i inner)
__global__ void k_hier_avg(View crse, View fine, const int *list, int n, int nv, int r, double volfrac,
                           const TimeState *ts0) {
    if (hier_skip(ts0)) return;
    const int64_t tot = (int64_t)n * nv;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int e = (int)(q % n), c = (int)(q / n);
        const int i = list[e], j = list[n + e], k = list[2 * n + e];
        double s = 0.0;
        for (int kr = 0; kr < r; ++kr)
            for (int jr = 0; jr < r; ++jr)
                for (int ir = 0; ir < r; ++ir) s += fine(i * r + ir, j * r + jr, k * r + kr, c);
        crse(i, j, k, c) = s * volfrac;
    }
}

// vfrac of the valid cells of a padded level array, dense (x fastest)
__global__ void k_hier_vfrac(View s, double *out, int nx, int ny, int nz) {
    const int64_t n = (int64_t)nx * ny * nz;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((int64_t)nx * ny));
        out[q] = s(i, j, k, VFRAC);
    }
}

// the end of a coarse step (after k_finalize of level 0): every level's t_new
// is cur_time + dt[0]; the fine level keeps its dt as the next step's growth
// cap; an error of the fine context stops the run
__global__ void k_hier_end(HierTime *ht, TimeState *ts0, TimeState *ts1, const int *err1, int nlev) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (err1 && *err1 && !ts0->err) ts0->err = *err1;
    for (int l = 0; l < nlev; ++l) ht->t_new[l] = ts0->t;
    if (nlev > 1 && ts1) {
        ts1->t = ts0->t;
        ts1->dt_prev = ts1->dt;
        ts1->active = ts0->active;
    }
}

// times of a level made by a regrid at `time` (t_old = time - 1e200)
__global__ void k_hier_newlevel(HierTime *ht, TimeState *ts1, int lev, double time, double dt, int set_dt) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    ht->t_new[lev] = time;
    ht->t_old[lev] = time - 1.e200;
    ts1->t = time;
    if (set_dt) {
        ht->dt[lev] = dt;
        ts1->dt = dt;
        ts1->dt_prev = dt;
    }
}

unsigned hgrid(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384)); }

struct Hier {
    Level *L[kHierMaxLev] = {nullptr, nullptr};
    Ctx *cf = nullptr;            // context of level 1 (owned)
    int nlev = 1, ratio = 2;
    HierTime *ht = nullptr;
    // fill lists of level 1, [3][n] each
    int *cneed = nullptr, *fcf = nullptr, *fper = nullptr, *fphys[3] = {nullptr, nullptr, nullptr}, *avgc = nullptr;
    int ncneed = 0, nfcf = 0, nfper = 0, nfphys[3] = {0, 0, 0}, navg = 0;
    uint64_t *fcfm = nullptr;     // child masks of the parents listed in fcf
    double *ct = nullptr;         // coarse patch, padded coarse shape
    RegDev *reg = nullptr;
    double *regv = nullptr;
    TargetDev *tgt = nullptr;
    int nreg = 0, ntgt = 0;
    int do_reflux = 1;
    int only_flow = 0;
    // regrid every regrid_int coarse steps (0: never) with ErrorEst's criteria
    int regrid_int = 0, n_error_buf = 1, blocking_factor = 8, max_grid_size = 32;
    double grid_eff = 0.7;
    std::vector<double> tag_phi, tag_grad;
    std::vector<int> tag_comp;
    int last_regrid_step = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_only_flow = 0;
};

void hier_drop_graph(Hier *H) {
    if (H->gexec) { hipGraphExecDestroy(H->gexec); H->gexec = nullptr; }
    if (H->graph) { hipGraphDestroy(H->graph); H->graph = nullptr; }
}

void hier_free_lists(Hier *H) {
    int **lists[] = {&H->cneed, &H->fcf, &H->fper, &H->fphys[0], &H->fphys[1], &H->fphys[2], &H->avgc};
    for (int **p : lists)
        if (*p) { hipFree(*p); *p = nullptr; }
    if (H->fcfm) { hipFree(H->fcfm); H->fcfm = nullptr; }
    if (H->reg) { hipFree(H->reg); H->reg = nullptr; }
    if (H->regv) { hipFree(H->regv); H->regv = nullptr; }
    if (H->tgt) { hipFree(H->tgt); H->tgt = nullptr; }
    H->ncneed = H->nfcf = H->nfper = H->navg = H->nreg = H->ntgt = 0;
    for (int p = 0; p < 3; ++p) H->nfphys[p] = 0;
}

template <class T>
int upload_list(const std::vector<T> &v, T **dst) {
    if (v.empty()) return MF_OK;
    HIPCHK(hipMalloc(dst, sizeof(T) * v.size()));
    HIPCHK(hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return MF_OK;
}

// cell sets of a two-level fill of the fine cells in `regions` given fine
// data over `fboxes` (host): interpolated cells, periodic images, cells
// outside the domain by phase, and the coarse cells the interpolation reads
struct FillSets {
    std::vector<int> cf, per, phys[3], cneed;   // [3][n] after pack()
    std::vector<int> cfp;                       // parents of the cf cells, [3][n] after pack()
    std::vector<uint64_t> cfm;                  // their children among the cf cells
};
void pack3(std::vector<int> &v) {  // (i,j,k) triples -> [3][n]
    const size_t n = v.size() / 3;
    std::vector<int> o(v.size());
    for (size_t q = 0; q < n; ++q) {
        o[q] = v[3 * q];
        o[n + q] = v[3 * q + 1];
        o[2 * n + q] = v[3 * q + 2];
    }
    v.swap(o);
}
void build_fill_sets(const int *nc, const int *per, int r, const std::vector<mf_box> &fboxes,
                     const std::vector<mf_box> &regions, FillSets &S) {
    const int G = MF_NGROW;
    int nf[3];
    for (int d = 0; d < 3; ++d) nf[d] = nc[d] * r;
    std::vector<unsigned char> cov((size_t)nf[0] * nf[1] * nf[2], 0);
    for (const mf_box &B : fboxes)
        for (int k = B.lo[2]; k <= B.hi[2]; ++k)
            for (int j = B.lo[1]; j <= B.hi[1]; ++j)
                std::memset(cov.data() + (size_t)nf[0] * (j + (size_t)nf[1] * k) + B.lo[0], 1, (size_t)(B.hi[0] - B.lo[0] + 1));
    const int64_t px = nf[0] + 2 * G, py = nf[1] + 2 * G, pz = nf[2] + 2 * G;
    std::vector<unsigned char> seen((size_t)(px * py * pz), 0);
    std::vector<unsigned char> cseen((size_t)nc[0] * nc[1] * nc[2], 0);
    std::map<int64_t, size_t> parent;  // parent coarse cell (padded index) -> its entry
    for (const mf_box &R : regions)
        for (int k = R.lo[2]; k <= R.hi[2]; ++k)
            for (int j = R.lo[1]; j <= R.hi[1]; ++j)
                for (int i = R.lo[0]; i <= R.hi[0]; ++i) {
                    unsigned char &m = seen[(size_t)((i + G) + px * ((j + G) + py * (int64_t)(k + G)))];
                    if (m) continue;
                    m = 1;
                    const int iv[3] = {i, j, k};
                    int out_np = 0, w[3];
                    bool inside = true;
                    for (int d = 0; d < 3; ++d) {
                        const bool o = iv[d] < 0 || iv[d] >= nf[d];
                        if (o) { inside = false; if (!per[d]) ++out_np; }
                        w[d] = per[d] ? ((iv[d] % nf[d]) + nf[d]) % nf[d] : iv[d];
                    }
                    if (out_np > 0) {
                        std::vector<int> &P = S.phys[out_np - 1];
                        P.push_back(i); P.push_back(j); P.push_back(k);
                        continue;
                    }
                    const bool covered = cov[(size_t)w[0] + (size_t)nf[0] * (w[1] + (size_t)nf[1] * w[2])] != 0;
                    if (covered) {
                        if (!inside) { S.per.push_back(i); S.per.push_back(j); S.per.push_back(k); }
                        continue;
                    }
                    S.cf.push_back(i); S.cf.push_back(j); S.cf.push_back(k);
                    int p[3];
                    for (int d = 0; d < 3; ++d) p[d] = (iv[d] >= 0) ? iv[d] / r : -((-iv[d] + r - 1) / r);
                    {
                        const int64_t key = (int64_t)(p[0] + G) + (int64_t)(nc[0] + 2 * G) *
                                            ((int64_t)(p[1] + G) + (int64_t)(nc[1] + 2 * G) * (p[2] + G));
                        auto it = parent.find(key);
                        if (it == parent.end()) {
                            S.cfp.push_back(p[0]); S.cfp.push_back(p[1]); S.cfp.push_back(p[2]);
                            S.cfm.push_back(0);
                            it = parent.emplace(key, S.cfm.size() - 1).first;
                        }
                        const int bit = (iv[0] - p[0] * r) + r * ((iv[1] - p[1] * r) + r * (iv[2] - p[2] * r));
                        S.cfm[it->second] |= (uint64_t)1 << bit;
                    }
                    for (int dk = -1; dk <= 1; ++dk)
                        for (int dj = -1; dj <= 1; ++dj)
                            for (int di = -1; di <= 1; ++di) {
                                int q[3] = {p[0] + di, p[1] + dj, p[2] + dk};
                                bool ok = true;
                                for (int d = 0; d < 3; ++d) {
                                    if (per[d]) q[d] = ((q[d] % nc[d]) + nc[d]) % nc[d];
                                    else if (q[d] < 0 || q[d] >= nc[d]) ok = false;
                                }
                                if (!ok) continue;
                                unsigned char &cm = cseen[(size_t)q[0] + (size_t)nc[0] * (q[1] + (size_t)nc[1] * q[2])];
                                if (cm) continue;
                                cm = 1;
                                S.cneed.push_back(q[0]); S.cneed.push_back(q[1]); S.cneed.push_back(q[2]);
                            }
                }
    pack3(S.cf);
    pack3(S.cfp);
    pack3(S.per);
    pack3(S.cneed);
    for (int p = 0; p < 3; ++p) pack3(S.phys[p]);
}

// element offset of cell/face (i,j,k) component 0 in a view relative to base
int64_t view_ofs(const View &v, const double *base, int i, int j, int k) {
    return (int64_t)(v.p - base) + v.off(i, j, k);
}

// the fill lists, the register faces, the reflux targets and the
// average-down cells of level 1's boxes
int hier_build(Hier *H) {
    hier_free_lists(H);
    hier_drop_graph(H);
    if (H->nlev < 2) return MF_OK;
    Level *Lc = H->L[0], *Lf = H->L[1];
    const Ctx *c = Lc->c;
    const int r = H->ratio;
    const int *nc = Lc->n;
    const int *per = c->geom.is_periodic;
    std::vector<mf_box> grown;
    for (const mf_box &B : Lf->boxes) {
        mf_box g;
        for (int d = 0; d < 3; ++d) { g.lo[d] = B.lo[d] - MF_NGROW; g.hi[d] = B.hi[d] + MF_NGROW; }
        grown.push_back(g);
    }
    FillSets S;
    build_fill_sets(nc, per, r, Lf->boxes, grown, S);
    int st = upload_list(S.cneed, &H->cneed);
    if (!st) st = upload_list(S.cfp, &H->fcf);
    if (!st) st = upload_list(S.cfm, &H->fcfm);
    if (!st) st = upload_list(S.per, &H->fper);
    for (int p = 0; p < 3 && !st; ++p) st = upload_list(S.phys[p], &H->fphys[p]);
    if (st) return st;
    H->ncneed = (int)(S.cneed.size() / 3);
    H->nfcf = (int)S.cfm.size();
    H->nfper = (int)(S.per.size() / 3);
    for (int p = 0; p < 3; ++p) H->nfphys[p] = (int)(S.phys[p].size() / 3);
    // coarse cells under the fine level, owners of the coarse cells
    const int64_t ncc = (int64_t)nc[0] * nc[1] * nc[2];
    std::vector<unsigned char> ccov((size_t)ncc, 0);
    std::vector<int> avg;
    for (const mf_box &B : Lf->boxes)
        for (int k = B.lo[2] / r; k <= B.hi[2] / r; ++k)
            for (int j = B.lo[1] / r; j <= B.hi[1] / r; ++j)
                for (int i = B.lo[0] / r; i <= B.hi[0] / r; ++i) {
                    ccov[(size_t)i + (size_t)nc[0] * (j + (size_t)nc[1] * k)] = 1;
                    avg.push_back(i); avg.push_back(j); avg.push_back(k);
                }
    pack3(avg);
    H->navg = (int)(avg.size() / 3);
    if ((st = upload_list(avg, &H->avgc))) return st;
    std::vector<int> owner((size_t)ncc, -1);
    for (size_t b = 0; b < Lc->boxes.size(); ++b) {
        const mf_box &B = Lc->boxes[b];
        for (int k = B.lo[2]; k <= B.hi[2]; ++k)
            for (int j = B.lo[1]; j <= B.hi[1]; ++j)
                for (int i = B.lo[0]; i <= B.hi[0]; ++i) owner[(size_t)i + (size_t)nc[0] * (j + (size_t)nc[1] * k)] = (int)b;
    }
    // register faces: the faces of the coarsened fine boxes whose outside cell is valid and uncovered
    std::vector<RegDev> regs;
    std::vector<TargetDev> tg;
    std::map<int64_t, int> tmap;
    for (size_t fb = 0; fb < Lf->boxes.size(); ++fb) {
        const mf_box &F = Lf->boxes[fb];
        const BoxDesc &fd = Lf->hbox[fb];
        int clo[3], chi[3];
        for (int d = 0; d < 3; ++d) { clo[d] = F.lo[d] / r; chi[d] = F.hi[d] / r; }
        for (int d = 0; d < 3; ++d)
            for (int side = 0; side < 2; ++side) {
                const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
                const int fidx = side == 0 ? clo[d] : chi[d] + 1;
                for (int b2 = clo[t2]; b2 <= chi[t2]; ++b2)
                    for (int a1 = clo[t1]; a1 <= chi[t1]; ++a1) {
                        int face[3];
                        face[d] = fidx;
                        face[t1] = a1;
                        face[t2] = b2;
                        int o[3] = {face[0], face[1], face[2]};
                        if (side == 0) o[d] -= 1;
                        int w[3] = {o[0], o[1], o[2]};
                        if (per[d]) w[d] = ((o[d] % nc[d]) + nc[d]) % nc[d];
                        else if (o[d] < 0 || o[d] >= nc[d]) continue;
                        const int64_t cell = (int64_t)w[0] + (int64_t)nc[0] * (w[1] + (int64_t)nc[1] * w[2]);
                        if (ccov[(size_t)cell]) continue;
                        const int ob = owner[(size_t)cell];
                        int cface[3] = {face[0], face[1], face[2]};
                        cface[d] += w[d] - o[d];
                        RegDev g;
                        g.dir = d;
                        g.side = side;
                        const View &cv = Lc->hbox[(size_t)ob].flux[d];
                        g.cofs = view_ofs(cv, Lc->flux, cface[0], cface[1], cface[2]);
                        g.cns = cv.ns;
                        const View &fv = fd.flux[d];
                        int p[3];
                        p[d] = face[d] * r;
                        p[t1] = face[t1] * r;
                        p[t2] = face[t2] * r;
                        g.fofs = view_ofs(fv, Lf->flux, p[0], p[1], p[2]);
                        g.fns = fv.ns;
                        const int64_t str[3] = {1, fv.js, fv.ks};
                        g.fs1 = str[t1];
                        g.fs2 = str[t2];
                        auto it = tmap.find(cell);
                        if (it == tmap.end()) {
                            TargetDev t;
                            for (int q = 0; q < 3; ++q) t.c[q] = w[q];
                            for (int q = 0; q < 6; ++q) t.e[q] = -1;
                            tg.push_back(t);
                            it = tmap.emplace(cell, (int)tg.size() - 1).first;
                        }
                        tg[(size_t)it->second].e[2 * d + (side == 1 ? 0 : 1)] = (int)regs.size();
                        regs.push_back(g);
                    }
            }
    }
    H->nreg = (int)regs.size();
    H->ntgt = (int)tg.size();
    const int NC = NCV + c->ns;
    if (H->nreg) {
        HIPCHK(hipMalloc(&H->reg, sizeof(RegDev) * regs.size()));
        HIPCHK(hipMemcpy(H->reg, regs.data(), sizeof(RegDev) * regs.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMalloc(&H->regv, sizeof(double) * (size_t)NC * H->nreg));
        HIPCHK(hipMalloc(&H->tgt, sizeof(TargetDev) * tg.size()));
        HIPCHK(hipMemcpy(H->tgt, tg.data(), sizeof(TargetDev) * tg.size(), hipMemcpyHostToDevice));
    }
    if (!H->ct) {
        HIPCHK(hipMalloc(&H->ct, Lc->padded * sizeof(double)));
        HIPCHK(hipMemset(H->ct, 0, Lc->padded * sizeof(double)));
    }
    return MF_OK;
}

// FillPatch of level 1 into its stage array `dst` at the stage's time: the
// coarse patch from level 0's old/new data (mode_forced >= 0: that data
// alone, at time_host), then the interpolation, the images and the boundaries
void hier_fill_fine(Hier *H, double *dst, double *c_old, double *c_new, int stage, int mode_forced, double time_host,
                    const int *cneed, int ncneed, const int *fcf, const uint64_t *fcfm, int nfcf, bool ghosts) {
    Level *Lc = H->L[0], *Lf = H->L[1];
    Ctx *c = Lc->c, *cf = Lf->c;
    const int nv = c->nv;
    hipStream_t st = c->stream;
    const View ctv = padded_view(H->ct, Lc->n, nv);
    const TimeState *ts0 = mode_forced >= 0 ? nullptr : Lc->ts;
    if (ncneed)
        hipLaunchKernelGGL(k_hier_ct, dim3(hgrid((int64_t)ncneed * nv)), dim3(256), 0, st, ctv,
                           padded_view(c_old, Lc->n, nv), padded_view(c_new, Lc->n, nv), cneed, ncneed, nv, H->ht,
                           Lf->ts, stage, mode_forced, ts0);
    FillArgs fa = fill_args(Lc, H->ct, stage);
    if (mode_forced >= 0) { fa.ts = nullptr; fa.time_host = time_host; }
    else fa.ts = Lf->ts;
    for (int ph = 0; ph <= 3; ++ph) c->ks->fill(fa, ph, st);
    const View fv = padded_view(dst, Lf->n, nv);
    if (nfcf)
        hipLaunchKernelGGL(k_hier_cf, dim3(hgrid((int64_t)nfcf * nv)), dim3(256), 0, st, fv, ctv, fcf, fcfm, nfcf, nv,
                           H->ratio, (double)(H->ratio - 1) / (double)(2 * H->ratio), ts0);
    if (!ghosts) return;
    if (H->nfper)
        hipLaunchKernelGGL(k_hier_per, dim3(hgrid((int64_t)H->nfper * nv)), dim3(256), 0, st, fv, H->fper, H->nfper, nv,
                           Lf->n[0], Lf->n[1], Lf->n[2], ts0);
    FillArgs ff = fill_args(Lf, dst, stage);
    ff.ts = Lf->ts;
    for (int ph = 1; ph <= 3; ++ph) cf->ks->fill_list(ff, ph, H->fphys[ph - 1], H->nfphys[ph - 1], st);
}

// one RK3 stage of a level of the hierarchy (Advance_coupled_strang
// :268-317): fill, cut cells, face sweeps, residual with the stage combine,
// register updates. Arrays: level 0 A -> B -> C -> B; level 1 A -> B -> C -> A.
void hier_stage(Hier *H, int lev, int stage, int mode, double *c_old, double *c_new) {
    Level *L = H->L[lev];
    Ctx *c = L->c;
    hipStream_t st = c->stream;
    double *src = L->arr[stage == 1 ? 0 : (stage == 2 ? 1 : 2)];
    const double w = stage == 3 ? kTwo3rd : kTwo3rd * kHalf * kHalf;
    if (lev == 0) {
        FillArgs f = fill_args(L, src, stage);
        for (int ph = 0; ph <= 3; ++ph) c->ks->fill(f, ph, st);
    } else {
        hier_fill_fine(H, src, c_old, c_new, stage, -1, 0.0, H->cneed, H->ncneed, H->fcf, H->fcfm, H->nfcf, true);
    }
    if (L->ncut) {
        CutArgs ca = cut_args(L, src);
        if (L->partial) c->ks->cut_geom_sub(ca, L->cutgeom, L->cutmask, L->geomsub, L->ngeomsub, st);
        c->ks->cut(ca, st);
    }
    StageArgs a = stage_args(L, stage);
    for (int d = 0; d < 3; ++d) launch_faces(L, a, d, mode);
    ResArgs r;
    std::memset(&r, 0, sizeof(r));
    r.st = a;
    r.combine = 1;
    r.old = L->view[0];
    r.dst = L->view[stage == 1 ? 1 : (stage == 2 ? 2 : (lev == 0 ? 1 : 0))];
    r.do_ns = c->prm.do_ns;
    r.partials = L->partials;
    launch_residual(L, r, mode);
    const int NC = NCV + c->ns;
    const double *dx = c->dx;
    if (H->nreg) {
        const TimeState *ts0 = H->L[0]->ts;
        if (lev == 0)
            hipLaunchKernelGGL(k_hier_crse_init, dim3(hgrid((int64_t)H->nreg * NC)), dim3(256), 0, st, H->regv, H->reg,
                               H->nreg, NC, (const double *)L->flux, dx[1] * dx[2], dx[0] * dx[2], dx[0] * dx[1], w, L->ts,
                               ts0);
        else
            hipLaunchKernelGGL(k_hier_fine_add, dim3(hgrid((int64_t)H->nreg * NC)), dim3(256), 0, st, H->regv, H->reg,
                               H->nreg, NC, (const double *)L->flux, H->ratio, dx[1] * dx[2], dx[0] * dx[2],
                               dx[0] * dx[1], w, L->ts, ts0);
    }
}

void hier_chem(Level *L, View src, View dst, int mode) {
    Ctx *c = L->c;
    ResArgs r;
    std::memset(&r, 0, sizeof(r));
    r.st = stage_args(L, 3);
    r.st.src = src;
    r.dst = dst;
    Work w = level_work(L, kCellClass);
    (void)mode;
    c->ks->chem(r, w, grid_for(L->ntiles[kCellClass], L->max_grid), 256, c->stream);
}

// the end-of-step passes of a level: update_primitive_vars, clip and (level
// 0) the residual norms against `old`, EstTimeStep partials; src -> A
void hier_post(Level *L, View src, View old, int track, int mode) {
    Ctx *c = L->c;
    ResArgs r;
    std::memset(&r, 0, sizeof(r));
    r.st = stage_args(L, 3);
    r.st.src = src;
    r.combine = 1;
    r.old = old;
    r.dst = L->view[0];
    r.do_chem = 0;
    r.clip = c->prm.clip_species;
    r.track_res = track;
    r.do_ns = c->prm.do_ns;
    r.partials = L->partials;
    Work w = level_work(L, kCellClass);
    (void)mode;
    L->post_slots = grid_for(L->ntiles[kCellClass], L->max_grid);
    c->ks->post(r, w, L->post_slots, 256, c->stream);
}

// ComputeDt of the hierarchy
void hier_enqueue_dt(Hier *H) {
    Level *L0 = H->L[0];
    Ctx *c = L0->c;
    HierDtArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nlev = H->nlev;
    a.ht = H->ht;
    for (int l = 0; l < H->nlev; ++l) {
        a.ts[l] = H->L[l]->ts;
        a.nsub[l] = l == 0 ? 1 : H->ratio;
        for (int d = 0; d < 3; ++d) a.dx[l][d] = H->L[l]->c->dx[d];
    }
    a.cfl = c->prm.cfl;
    a.fixed = c->prm.fixed_dt;
    a.dt_fixed = c->prm.dt_fixed;
    hipLaunchKernelGGL(k_hier_dt, dim3(1), dim3(1), 0, c->stream, a);
}

// timeStep(0) and the rest of EvolveAMR's step after ComputeDt
void hier_enqueue_advance(Hier *H, int mode) {
    Level *L0 = H->L[0], *L1 = H->nlev > 1 ? H->L[1] : nullptr;
    Ctx *c = L0->c;
    hipStream_t st = c->stream;
    const TimeState *ts0 = L0->ts;
    const int NC = NCV + c->ns;
    const bool chem0 = c->prm.do_chemistry && !H->only_flow;
    hipLaunchKernelGGL(k_hier_begin, dim3(1), dim3(1), 0, st, H->ht, L0->ts, ts0, 0, 1);
    if (H->nreg)
        hipLaunchKernelGGL(k_hier_zero, dim3(hgrid((int64_t)H->nreg * NC)), dim3(256), 0, st, H->regv,
                           (int64_t)H->nreg * NC, ts0);
    for (int s = 1; s <= 3; ++s) hier_stage(H, 0, s, mode, nullptr, nullptr);
    // Advance_chemistry_implicit (:313-316): phi_old = S3 (B), phi_new = C
    if (chem0) hier_chem(L0, L0->view[1], L0->view[2], mode);
    double *c_old = chem0 ? L0->arr[1] : L0->arr[0];
    double *c_new = chem0 ? L0->arr[2] : L0->arr[1];
    const View v_new = padded_view(c_new, L0->n, c->nv);
    if (L1) {
        // the recursion passes no only_flow (:237): the fine level runs its chemistry
        const bool chem1 = L1->c->prm.do_chemistry != 0;
        for (int i = 1; i <= H->ratio; ++i) {
            hipLaunchKernelGGL(k_hier_begin, dim3(1), dim3(1), 0, st, H->ht, L1->ts, ts0, 1, i);
            for (int s = 1; s <= 3; ++s) hier_stage(H, 1, s, mode, c_old, c_new);
            if (chem1) hier_chem(L1, L1->view[0], L1->view[0], mode);
        }
        if (H->do_reflux && H->ntgt) {
            const double vol = c->dx[0] * c->dx[1] * c->dx[2];
            hipLaunchKernelGGL(k_hier_reflux, dim3(hgrid((int64_t)H->ntgt * NC)), dim3(256), 0, st, v_new, H->tgt,
                               H->ntgt, H->reg, H->regv, H->nreg, NC, vol, ts0);
        }
        if (H->navg)
            hipLaunchKernelGGL(k_hier_avg, dim3(hgrid((int64_t)H->navg * c->nv)), dim3(256), 0, st, v_new,
                               L1->view[0], H->avgc, H->navg, c->nv, H->ratio,
                               1.0 / (double)(H->ratio * H->ratio * H->ratio), ts0);
    }
    // update_primitive_vars and clip on every level, residual norms of level 0 (:32-57)
    hier_post(L0, v_new, padded_view(c_old, L0->n, c->nv), c->prm.track_residual_norms, mode);
    if (L1) hier_post(L1, L1->view[0], L1->view[0], 0, mode);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, st, L0->partials, L0->post_slots, c->nv, L0->ts, c->d_err, 0,
                       c->prm.track_residual_norms);
    if (L1)
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, st, L1->partials, L1->post_slots, c->nv, L1->ts,
                           L1->c->d_err, 1, 0);
    hipLaunchKernelGGL(k_hier_end, dim3(1), dim3(1), 0, st, H->ht, L0->ts, L1 ? L1->ts : nullptr,
                       L1 ? (const int *)L1->c->d_err : nullptr, H->nlev);
}

void hier_sync_ctx(Hier *H) {
    if (H->nlev < 2) return;
    Ctx *c = H->L[0]->c, *cf = H->L[1]->c;
    cf->stream = c->stream;
    if (std::memcmp(&cf->tab, &c->tab, sizeof(c->tab)) != 0) mf_ctx_set_plugin_params(cf, &c->tab);
}

int hier_est(Hier *H) {
    for (int l = 0; l < H->nlev; ++l)
        if (H->L[l]->need_est) {
            enqueue_est(H->L[l]);
            H->L[l]->need_est = false;
        }
    HIPCHK(hipGetLastError());
    return MF_OK;
}

// a level-1 Level over `boxes` (fine indices) in a context with the refined geometry
int hier_new_fine(Hier *H, const mf_box *boxes, int nb, Level **out) {
    Level *Lc = H->L[0];
    Ctx *c = Lc->c;
    if (!H->cf) {
        mf_geom g = c->geom;
        for (int d = 0; d < 3; ++d) g.n_cell[d] = c->geom.n_cell[d] * H->ratio;
        void *cf = nullptr;
        int st = mf_ctx_create(&c->tab, &c->prm, &g, &cf);
        if (st) return st;
        H->cf = (Ctx *)cf;
    }
    H->cf->stream = c->stream;
    Level *L = nullptr;
    int st = level_alloc(H->cf, boxes, nb, true, &L);
    if (st) return st;
    for (int q = 0; q < kTileClasses; ++q)
        for (int d = 0; d < 3; ++d) L->tsh[q][d] = Lc->tsh[q][d];
    L->max_grid = Lc->max_grid;
    if ((st = rebuild_tiles(L))) { level_free(L); return st; }
    L->hier = H;
    *out = L;
    return MF_OK;
}

// fine boxes: inside the fine domain, aligned to the ratio, disjoint, at least one
int hier_check_boxes(const Hier *H, const mf_box *boxes, int nb) {
    if (nb < 1 || !boxes) return MF_ERR_BAD_ARG;
    const int r = H->ratio;
    int nf[3];
    for (int d = 0; d < 3; ++d) nf[d] = H->L[0]->n[d] * r;
    std::vector<unsigned char> m((size_t)nf[0] * nf[1] * nf[2], 0);
    for (int b = 0; b < nb; ++b) {
        for (int d = 0; d < 3; ++d) {
            const int lo = boxes[b].lo[d], hi = boxes[b].hi[d];
            if (lo < 0 || hi >= nf[d] || hi < lo || lo % r || (hi + 1) % r) return MF_ERR_BAD_ARG;
        }
        for (int k = boxes[b].lo[2]; k <= boxes[b].hi[2]; ++k)
            for (int j = boxes[b].lo[1]; j <= boxes[b].hi[1]; ++j) {
                unsigned char *row = m.data() + (size_t)nf[0] * (j + (size_t)nf[1] * k);
                for (int i = boxes[b].lo[0]; i <= boxes[b].hi[0]; ++i) {
                    if (row[i]) return MF_ERR_BAD_ARG;
                    row[i] = 1;
                }
            }
    }
    return MF_OK;
}

// the cut list of level 1 from its current data
int hier_fine_cutlist(Level *Lf) {
    const int64_t ncell = (int64_t)Lf->n[0] * Lf->n[1] * Lf->n[2];
    double *dense = nullptr;
    HIPCHK(hipMalloc(&dense, sizeof(double) * ncell));
    hipLaunchKernelGGL(k_hier_vfrac, dim3(hgrid(ncell)), dim3(256), 0, Lf->c->stream, Lf->view[0], dense, Lf->n[0],
                       Lf->n[1], Lf->n[2]);
    std::vector<double> vf((size_t)ncell);
    hipError_t e = hipStreamSynchronize(Lf->c->stream);
    if (e == hipSuccess) e = hipMemcpy(vf.data(), dense, sizeof(double) * ncell, hipMemcpyDeviceToHost);
    hipFree(dense);
    if (e != hipSuccess) { last_hip_error = e; return MF_ERR_HIP; }
    return build_cutlist(Lf, vf.data());
}

// regrid of level 1 at `time` with new boxes: RemakeLevel (level 1 exists:
// its data where the old boxes covered, the coarse level interpolated
// elsewhere), MakeNewLevelFromCoarse (it does not), ClearLevel (no boxes)
// (Source/mflo_gridwork.cpp:21-89); `time` is t_new of level 0, whose data is
// A. in_step: the regrid inside timeStep, after ComputeDt (a new level gets
// this step's dt[0] / ratio); else between steps (dt[0] of the last step, or
// 1e100 before the first)
int hier_regrid_to(Hier *H, const mf_box *boxes, int nb, double time, bool in_step) {
    Level *Lc = H->L[0];
    hipStream_t st = Lc->c->stream;
    HIPCHK(hipStreamSynchronize(st));
    hier_drop_graph(H);
    if (nb == 0) {
        if (H->nlev > 1) {
            H->L[1]->hier = nullptr;
            level_free(H->L[1]);
            H->L[1] = nullptr;
            H->nlev = 1;
        }
        return hier_build(H);
    }
    int rc = hier_check_boxes(H, boxes, nb);
    if (rc) return rc;
    std::vector<mf_box> nbx(boxes, boxes + nb);
    std::vector<mf_box> old;
    Level *Lf = nullptr;
    const bool had = H->nlev > 1;
    if (had) {
        old = H->L[1]->boxes;
        if (old.size() == nbx.size() && std::memcmp(old.data(), nbx.data(), sizeof(mf_box) * nbx.size()) == 0) return MF_OK;
        // the same arrays (old data in place) with the new boxes
        Level *O = H->L[1];
        if ((rc = hier_new_fine(H, boxes, nb, &Lf))) return rc;
        for (int q = 0; q < 3; ++q) std::swap(Lf->arr[q], O->arr[q]);
        for (int q = 0; q < 3; ++q) {
            Lf->view[q] = padded_view(Lf->arr[q], Lf->n, Lf->c->nv);
            O->view[q] = padded_view(O->arr[q], O->n, O->c->nv);
        }
        std::swap(Lf->ts, O->ts);
        O->hier = nullptr;
        level_free(O);
    } else {
        if ((rc = hier_new_fine(H, boxes, nb, &Lf))) return rc;
    }
    H->L[1] = Lf;
    H->nlev = 2;
    // the new cells: interpolated from level 0's current data (GetData at t_new)
    FillSets S;
    build_fill_sets(Lc->n, Lc->c->geom.is_periodic, H->ratio, old, nbx, S);
    int *cneed = nullptr, *fcf = nullptr;
    uint64_t *fcfm = nullptr;
    rc = upload_list(S.cneed, &cneed);
    if (!rc) rc = upload_list(S.cfp, &fcf);
    if (!rc) rc = upload_list(S.cfm, &fcfm);
    if (!rc && !H->ct) {
        if (hipMalloc(&H->ct, Lc->padded * sizeof(double)) != hipSuccess ||
            hipMemset(H->ct, 0, Lc->padded * sizeof(double)) != hipSuccess)
            rc = MF_ERR_HIP;
    }
    if (!rc) {
        TimeState h0;
        rc = read_ts(Lc, &h0);
        if (!rc) {
            // t_new = time, t_old = time - 1e200 (:34-35, :60-61); a new level gets dt[0] / ratio (:203-206)
            const double dt0 = in_step ? h0.dt : h0.dt_prev;
            hipLaunchKernelGGL(k_hier_newlevel, dim3(1), dim3(1), 0, st, H->ht, Lf->ts, 1, time, dt0 / H->ratio,
                               had ? 0 : 1);
            hier_fill_fine(H, Lf->arr[0], Lc->arr[0], Lc->arr[0], 1, 1, time, cneed, (int)(S.cneed.size() / 3), fcf,
                           fcfm, (int)S.cfm.size(), false);
            if (hipStreamSynchronize(st) != hipSuccess) rc = MF_ERR_HIP;
        }
    }
    if (cneed) hipFree(cneed);
    if (fcf) hipFree(fcf);
    if (fcfm) hipFree(fcfm);
    if (rc) return rc;
    if ((rc = hier_fine_cutlist(Lf))) return rc;
    // the next ComputeDt estimates over the new boxes (inside a step the
    // step's post-pass does it)
    Lf->need_est = !in_step;
    return hier_build(H);
}

// EvolveAMR's regrid check at the start of a coarse step (timeStep :177-209):
// at istep[0] > last regrid and istep[0] % regrid_int == 0, after ComputeDt
int hier_maybe_regrid(Hier *H, int istep0, double time, bool *did) {
    *did = false;
    if (H->regrid_int <= 0 || istep0 <= H->last_regrid_step || istep0 % H->regrid_int) return MF_OK;
    Level *Lc = H->L[0];
    const int64_t ncell = (int64_t)Lc->n[0] * Lc->n[1] * Lc->n[2];
    std::vector<unsigned char> tags((size_t)ncell);
    int rc = mf_level_tag(Lc, time, H->tag_phi.data(), H->tag_grad.data(), H->tag_comp.data(), (int)H->tag_comp.size(),
                          tags.data(), nullptr);
    if (rc) return rc;
    int32_t n[3], per[3];
    for (int d = 0; d < 3; ++d) { n[d] = Lc->n[d]; per[d] = Lc->c->geom.is_periodic[d]; }
    int nb = 0;
    rc = mf_tags_to_boxes(tags.data(), n, per, H->n_error_buf, H->blocking_factor, H->max_grid_size, H->grid_eff,
                          H->ratio, nullptr, 0, &nb);
    if (rc) return rc;
    std::vector<mf_box> bx((size_t)std::max(nb, 1));
    if (nb > 0) rc = mf_tags_to_boxes(tags.data(), n, per, H->n_error_buf, H->blocking_factor, H->max_grid_size,
                                      H->grid_eff, H->ratio, bx.data(), nb, &nb);
    if (rc) return rc;
    rc = hier_regrid_to(H, bx.data(), nb, time, true);
    if (rc) return rc;
    H->last_regrid_step = istep0;
    *did = true;
    return MF_OK;
}

int hier_steps(Hier *H, int nsteps, int mode) {
    Level *L0 = H->L[0];
    Ctx *c = L0->c;
    hier_sync_ctx(H);
    int rc = hier_est(H);
    if (rc) return rc;
    for (int s = 0; s < nsteps; ++s) {
        if (H->regrid_int > 0) {
            TimeState h;
            if ((rc = read_ts(L0, &h))) return rc;
            if (!h.active || h.err) break;
            if (h.istep > H->last_regrid_step && h.istep % H->regrid_int == 0) {
                // ComputeDt, then the regrid inside timeStep, then the advance
                hier_enqueue_dt(H);
                bool did = false;
                if ((rc = hier_maybe_regrid(H, h.istep, h.t, &did))) return rc;
                hier_sync_ctx(H);
                hier_enqueue_advance(H, mode == MF_MODE_GRAPH ? MF_MODE_BATCHED : mode);
                continue;
            }
        }
        if (mode == MF_MODE_GRAPH) {
            if (H->gexec && H->graph_only_flow != H->only_flow) hier_drop_graph(H);
            if (!H->gexec) {
                H->graph_only_flow = H->only_flow;
                hipStream_t cap = c->stream;
                bool own = false;
                if (!cap) {
                    HIPCHK(hipStreamCreate(&cap));
                    own = true;
                }
                HIPCHK(hipStreamSynchronize(c->stream));
                hipStream_t saved = c->stream;
                c->stream = cap;
                hier_sync_ctx(H);
                HIPCHK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
                hier_enqueue_dt(H);
                hier_enqueue_advance(H, MF_MODE_BATCHED);
                hipError_t e = hipStreamEndCapture(cap, &H->graph);
                c->stream = saved;
                hier_sync_ctx(H);
                if (own) hipStreamDestroy(cap);
                if (e != hipSuccess) { last_hip_error = e; return MF_ERR_HIP; }
                HIPCHK(hipGraphInstantiate(&H->gexec, H->graph, nullptr, nullptr, 0));
            }
            HIPCHK(hipGraphLaunch(H->gexec, c->stream));
        } else {
            hier_enqueue_dt(H);
            hier_enqueue_advance(H, mode);
        }
    }
    HIPCHK(hipGetLastError());
    return MF_OK;
}

int hier_status(Hier *H) {
    TimeState h;
    int rc = read_ts(H->L[0], &h);
    if (rc) return rc;
    return h.err;
}

}  // namespace

extern "C" {

int mf_hier_create(void *level0, int ratio, void **hier_out) {
    Level *L0 = (Level *)level0;
    if (!L0 || !hier_out || L0->partial || L0->hier || ratio < 2 || ratio > 4) return MF_ERR_BAD_ARG;
    *hier_out = nullptr;
    for (int d = 0; d < 3; ++d)
        if ((int64_t)L0->n[d] * ratio > (1 << 24)) return MF_ERR_BAD_ARG;
    Hier *H = new (std::nothrow) Hier();
    if (!H) return MF_ERR_BAD_ARG;
    H->L[0] = L0;
    H->ratio = ratio;
    if (hipMalloc(&H->ht, sizeof(HierTime)) != hipSuccess) { delete H; return MF_ERR_HIP; }
    TimeState h;
    int rc = read_ts(L0, &h);
    if (rc) { hipFree(H->ht); delete H; return rc; }
    HierTime t;
    std::memset(&t, 0, sizeof(t));
    for (int l = 0; l < kHierMaxLev; ++l) { t.t_new[l] = h.t; t.t_old[l] = -1.e100; t.dt[l] = h.dt; }
    hipMemcpy(H->ht, &t, sizeof(t), hipMemcpyHostToDevice);
    H->last_regrid_step = 0;
    L0->hier = H;
    *hier_out = H;
    return MF_OK;
}

int mf_hier_destroy(void *hier) {
    Hier *H = (Hier *)hier;
    if (!H) return MF_ERR_BAD_ARG;
    hipStreamSynchronize(H->L[0]->c->stream);
    hier_drop_graph(H);
    hier_free_lists(H);
    if (H->ct) hipFree(H->ct);
    if (H->nlev > 1) {
        H->L[1]->hier = nullptr;
        level_free(H->L[1]);
    }
    if (H->cf) mf_ctx_destroy(H->cf);
    if (H->ht) hipFree(H->ht);
    H->L[0]->hier = nullptr;
    delete H;
    return MF_OK;
}

int mf_hier_add_level(void *hier, const mf_box *boxes, int nboxes, const double *state, int is_device, void **level_out) {
    Hier *H = (Hier *)hier;
    if (!H || !state || H->nlev != 1) return MF_ERR_BAD_ARG;
    int rc = hier_check_boxes(H, boxes, nboxes);
    if (rc) return rc;
    Level *Lf = nullptr;
    if ((rc = hier_new_fine(H, boxes, nboxes, &Lf))) return rc;
    H->L[1] = Lf;
    H->nlev = 2;
    rc = mf_level_set_state(Lf, state, is_device);
    if (!rc) {
        TimeState h0;
        rc = read_ts(H->L[0], &h0);
        // MakeNewLevelFromScratch: t_new = the coarse time, t_old = -1e100, dt = 1e100 until ComputeDt
        if (!rc) hipLaunchKernelGGL(k_hier_newlevel, dim3(1), dim3(1), 0, H->L[0]->c->stream, H->ht, Lf->ts, 1, h0.t,
                                    1.e100, 1);
    }
    if (!rc) rc = hier_build(H);
    if (rc) {
        H->L[1] = nullptr;
        H->nlev = 1;
        Lf->hier = nullptr;
        level_free(Lf);
        return rc;
    }
    if (level_out) *level_out = Lf;
    return MF_OK;
}

int mf_hier_level(void *hier, int lev, void **level_out) {
    Hier *H = (Hier *)hier;
    if (!H || !level_out || lev < 0) return MF_ERR_BAD_ARG;
    *level_out = lev < H->nlev ? H->L[lev] : nullptr;
    return lev < H->nlev ? MF_OK : MF_ERR_BAD_ARG;
}

int mf_hier_nlevels(void *hier, int *nlev) {
    Hier *H = (Hier *)hier;
    if (!H || !nlev) return MF_ERR_BAD_ARG;
    *nlev = H->nlev;
    return MF_OK;
}

int mf_hier_autotune(void *hier) {
    Hier *H = (Hier *)hier;
    if (!H) return MF_ERR_BAD_ARG;
    hier_sync_ctx(H);
    for (int l = 0; l < H->nlev; ++l) {
        int rc = level_autotune(H->L[l]);
        if (rc) return rc;
    }
    return hier_build(H);
}

int mf_hier_set_reflux(void *hier, int do_reflux) {
    Hier *H = (Hier *)hier;
    if (!H) return MF_ERR_BAD_ARG;
    if (H->do_reflux != (do_reflux ? 1 : 0)) hier_drop_graph(H);
    H->do_reflux = do_reflux ? 1 : 0;
    return MF_OK;
}

int mf_hier_set_regrid(void *hier, int regrid_int, const double *refine_phi, const double *refine_phigrad,
                       const int *comps, int ntag, int n_error_buf, int blocking_factor, int max_grid_size,
                       double grid_eff) {
    Hier *H = (Hier *)hier;
    if (!H || regrid_int < 0 || ntag < 0 || ntag > 16 || n_error_buf < 0 || blocking_factor < H->ratio ||
        blocking_factor % H->ratio || max_grid_size < blocking_factor || !(grid_eff > 0.0 && grid_eff <= 1.0))
        return MF_ERR_BAD_ARG;
    if (ntag > 0 && (!refine_phi || !refine_phigrad || !comps)) return MF_ERR_BAD_ARG;
    H->regrid_int = regrid_int;
    H->tag_phi.assign(refine_phi, refine_phi + ntag);
    H->tag_grad.assign(refine_phigrad, refine_phigrad + ntag);
    H->tag_comp.assign(comps, comps + ntag);
    H->n_error_buf = n_error_buf;
    H->blocking_factor = blocking_factor;
    H->max_grid_size = max_grid_size;
    H->grid_eff = grid_eff;
    return MF_OK;
}

int mf_hier_regrid(void *hier, const mf_box *boxes, int nboxes) {
    Hier *H = (Hier *)hier;
    if (!H || nboxes < 0 || (nboxes > 0 && !boxes)) return MF_ERR_BAD_ARG;
    TimeState h;
    int rc = read_ts(H->L[0], &h);
    if (rc) return rc;
    return hier_regrid_to(H, boxes, nboxes, h.t, false);
}

int mf_hier_step(void *hier, int nsteps, int mode) {
    Hier *H = (Hier *)hier;
    if (!H || nsteps < 0 || mode < 0 || mode > 2) return MF_ERR_BAD_ARG;
    int rc = hier_steps(H, nsteps, mode);
    if (rc) return rc;
    return hier_status(H);
}

int mf_hier_evolve(void *hier, double final_time, int max_step, int only_flow, int mode, int *nsteps_out) {
    Hier *H = (Hier *)hier;
    if (!H || max_step < 0 || mode < 0 || mode > 2) return MF_ERR_BAD_ARG;
    int n = 0, rc = MF_OK;
    H->only_flow = only_flow ? 1 : 0;
    TimeState h;
    rc = read_ts(H->L[0], &h);
    while (!rc && !h.err && h.istep < max_step && h.t < final_time && h.active) {
        rc = hier_steps(H, 1, mode);
        if (rc) break;
        ++n;
        rc = read_ts(H->L[0], &h);
    }
    H->only_flow = 0;
    if (nsteps_out) *nsteps_out = n;
    if (rc) return rc;
    return h.err;
}

int mf_hier_info(void *hier, double *time, double *dt, int *istep, double *resnorms) {
    Hier *H = (Hier *)hier;
    if (!H) return MF_ERR_BAD_ARG;
    return mf_level_info(H->L[0], time, dt, istep, resnorms);
}

}  // extern "C"

namespace {
// a level's part of the output files (its state copied to the host)
int hier_level_data(Hier *H, int l, std::vector<double> &buf, mfio::LevelData &D) {
    Level *L = H->L[l];
    D.geom = L->c->geom;
    D.boxes = L->boxes;
    D.nv = L->c->nv;
    buf.resize((size_t)D.nv * L->n[0] * L->n[1] * L->n[2]);
    int rc = mf_level_get_state(L, buf.data(), 0);
    D.state = buf.data();
    return rc;
}
}  // namespace

extern "C" {

int mf_hier_write_plotfile(void *hier, const char *dir, const char *const *varnames) {
    Hier *H = (Hier *)hier;
    if (!H || !dir) return MF_ERR_BAD_ARG;
    std::vector<mfio::LevelData> D((size_t)H->nlev);
    std::vector<std::vector<double>> buf((size_t)H->nlev);
    std::vector<int> isteps;
    double t = 0.0;
    for (int l = 0; l < H->nlev; ++l) {
        TimeState h;
        int rc = read_ts(H->L[l], &h);
        if (!rc) rc = hier_level_data(H, l, buf[(size_t)l], D[(size_t)l]);
        if (rc) return rc;
        isteps.push_back(h.istep);
        if (l == 0) t = h.t;
    }
    return mfio::write_plotfile_levels(D, t, isteps, H->ratio, dir, varnames);
}

int mf_hier_write_checkpoint(void *hier, const char *dir, const char *const *varnames) {
    Hier *H = (Hier *)hier;
    if (!H || !dir) return MF_ERR_BAD_ARG;
    std::vector<mfio::LevelData> D((size_t)H->nlev);
    std::vector<std::vector<double>> buf((size_t)H->nlev);
    std::vector<int> isteps(kHierMaxLev, 0);
    std::vector<double> dt(kHierMaxLev, 0.0), t_new(kHierMaxLev, 0.0);
    for (int l = 0; l < kHierMaxLev; ++l) {
        TimeState h;
        int rc = read_ts(H->L[l < H->nlev ? l : 0], &h);
        if (rc) return rc;
        isteps[(size_t)l] = l < H->nlev ? h.istep : 0;
        dt[(size_t)l] = l < H->nlev ? h.dt_prev : h.dt_prev / H->ratio;
        t_new[(size_t)l] = h.t;
        if (l < H->nlev && (rc = hier_level_data(H, l, buf[(size_t)l], D[(size_t)l]))) return rc;
    }
    return mfio::write_checkpoint_levels(D, t_new, dt, isteps, dir, varnames);
}

int mf_hier_read_checkpoint(void *hier, const char *dir, const char *const *varnames) {
    Hier *H = (Hier *)hier;
    if (!H || !dir) return MF_ERR_BAD_ARG;
    int finest = -1;
    std::vector<int> isteps;
    std::vector<double> dt, t_new;
    std::vector<std::vector<mf_box>> bx;
    int rc = mfio::read_checkpoint_header(dir, kHierMaxLev - 1, &finest, isteps, dt, t_new, bx);
    if (rc) return rc;
    Level *L0 = H->L[0];
    if (bx[0].size() != L0->boxes.size() ||
        std::memcmp(bx[0].data(), L0->boxes.data(), sizeof(mf_box) * bx[0].size()) != 0)
        return MF_ERR_BAD_ARG;
    if (finest == 1 && (rc = hier_check_boxes(H, bx[1].data(), (int)bx[1].size()))) return rc;
    HIPCHK(hipStreamSynchronize(L0->c->stream));
    hier_drop_graph(H);
    {
        std::vector<double> s0((size_t)L0->c->nv * L0->n[0] * L0->n[1] * L0->n[2]);
        if ((rc = mf_level_get_state(L0, s0.data(), 0))) return rc;
        if ((rc = mfio::read_checkpoint_level(dir, 0, L0->c->geom, bx[0], L0->c->nv, s0.data(), varnames))) return rc;
        if ((rc = mf_level_set_state(L0, s0.data(), 0))) return rc;
        if ((rc = mf_level_set_time(L0, t_new[0], dt[0], isteps[0]))) return rc;
    }
    const bool same = finest == 1 && H->nlev > 1 && H->L[1]->boxes.size() == bx[1].size() &&
                      std::memcmp(H->L[1]->boxes.data(), bx[1].data(), sizeof(mf_box) * bx[1].size()) == 0;
    if (H->nlev > 1 && !same) {
        H->L[1]->hier = nullptr;
        level_free(H->L[1]);
        H->L[1] = nullptr;
        H->nlev = 1;
    }
    HierTime t;
    std::memset(&t, 0, sizeof(t));
    for (int l = 0; l < kHierMaxLev; ++l) {
        t.t_new[l] = t_new[(size_t)std::min(l, (int)t_new.size() - 1)];
        t.t_old[l] = -1.e100;
        t.dt[l] = dt[(size_t)std::min(l, (int)dt.size() - 1)];
    }
    if (finest == 1) {
        if (H->nlev < 2) {
            Level *Lf = nullptr;
            if ((rc = hier_new_fine(H, bx[1].data(), (int)bx[1].size(), &Lf))) return rc;
            H->L[1] = Lf;
            H->nlev = 2;
        }
        Level *Lf = H->L[1];
        std::vector<double> s1((size_t)Lf->c->nv * Lf->n[0] * Lf->n[1] * Lf->n[2], 0.0);
        if ((rc = mfio::read_checkpoint_level(dir, 1, Lf->c->geom, bx[1], Lf->c->nv, s1.data(), varnames))) return rc;
        if ((rc = mf_level_set_state(Lf, s1.data(), 0))) return rc;
        if ((rc = mf_level_set_time(Lf, t_new[1], dt[1], isteps[1]))) return rc;
    }
    HIPCHK(hipMemcpy(H->ht, &t, sizeof(t), hipMemcpyHostToDevice));
    H->last_regrid_step = 0;  // the static last_regrid_step of a fresh run (:185)
    return hier_build(H);
}

}  // extern "C"
