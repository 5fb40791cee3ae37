// mf_capi.hip -- host driver and C-ABI of the MI355X RHS + RK3 path.
//
// The level driver keeps three padded state arrays A, B, C per level:
//   stage 1: A (U^n, ghost-filled)  -> B = S1
//   stage 2: B (S1, ghost-filled)   -> C = S2
//   stage 3: C (S2, ghost-filled)   -> A = S3, chemistry, primitives, clip,
//            residual and dt partials fused (A is only read point-wise there)
// so every kernel argument of a step is fixed and a step can be replayed from a
// hipGraph. dt, time and step count live in a device TimeState.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <mutex>
#include <vector>
#include "mf_args.h"

using namespace mfd;

namespace mfd {
// the families compiled from case headers (filled while the library loads)
static std::vector<CaseFamily> &case_families() {
    static std::vector<CaseFamily> v;
    return v;
}
int register_case_family(const CaseFamily &f) {
    case_families().push_back(f);
    return (int)case_families().size();
}
const CaseFamily *find_case_family(int id) {
    for (const CaseFamily &f : case_families())
        if (f.id == id) return &f;
    return nullptr;
}
}  // namespace mfd

// Scratch of mf_box_rhs_noalloc: the face fluxes of one box, in a file-scope
// device array that each GPU holds from the library's load to the end of the
// process. Host code uses only the address hipGetSymbolAddress returns for
// the current device (noalloc_ptr below), never the array's name.
static constexpr size_t kNoAllocDoubles = (size_t)8 << 20;  // 64 MiB
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {

#define HIPCHK(x)                                      \
    do {                                               \
        hipError_t e_ = (x);                           \
        if (e_ != hipSuccess) {                        \
            last_hip_error = e_;                       \
            return MF_ERR_HIP;                         \
        }                                              \
    } while (0)

thread_local hipError_t last_hip_error = hipSuccess;

constexpr int kMaxDevices = 64;
std::mutex noalloc_mu;
double *noalloc_ptr[kMaxDevices] = {};
int noalloc_state[kMaxDevices] = {};  // 0 not resolved, 1 resolved, -1 failed

// resolves the scratch array's address on the current device once
int resolve_noalloc(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return MF_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (noalloc_state[dev] == 0) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            noalloc_state[dev] = -1;
        } else {
            noalloc_ptr[dev] = (double *)p;
            noalloc_state[dev] = 1;
        }
    }
    return noalloc_state[dev] == 1 ? MF_OK : MF_ERR_HIP;
}

struct Ctx {
    mf_plugin_params tab;
    mf_params prm;
    mf_geom geom;
    double dx[3];
    int nv = 0, ns = 0;
    Tab *d_tab = nullptr;
    int *d_err = nullptr;
    hipStream_t stream = nullptr;
    const KernelSet *ks = nullptr;
    int device = 0;
    // scratch of the per-box C-ABI calls
    double *scratch = nullptr;
    size_t scratch_bytes = 0;
    // work buffer of the two-level operators (coarse data of a fill)
    double *aux = nullptr;
    size_t aux_bytes = 0;
    int tsh[3] = {32, 4, 2};
    // tiles of the LDS-staged WENO5 face sweeps of the per-box calls (x, y, z)
    int ftsh[3][3] = {{64, 4, 1}, {32, 8, 1}, {32, 1, 8}};
};

// HIP-event timing of the kernel classes of a step (mf_level_step_timed)
enum { KC_DT = 0, KC_FILL = 1, KC_CUT = 2, KC_FACES = 3, KC_RESID = 4, KC_FINAL = 5, KC_POST = 6, KC_N = 7 };
struct Timer {
    bool on = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
    void begin(int cls, hipStream_t st) {
        if (!on) return;
        hipEvent_t a, b;
        hipEventCreate(&a);
        hipEventCreate(&b);
        hipEventRecord(a, st);
        ev.push_back({cls, {a, b}});
    }
    void end(hipStream_t st) {
        if (!on) return;
        hipEventRecord(ev.back().second.second, st);
    }
};

struct Level {
    Ctx *c = nullptr;
    Timer timer;
    int n[3];
    std::vector<mf_box> boxes;
    std::vector<BoxDesc> hbox;
    BoxDesc *d_box = nullptr;
    int ntiles[kTileClasses] = {0, 0, 0, 0, 0, 0, 0};
    // adjoint sweeps (classes 4..6): shapes measured fastest on the 128^3 porous case
    int tsh[kTileClasses][3] = {{64, 4, 1}, {32, 8, 1}, {32, 1, 8}, {64, 2, 2}, {16, 2, 4}, {8, 16, 1}, {8, 1, 16}};
    int max_grid = 2048;
    double *arr[3] = {nullptr, nullptr, nullptr};  // A, B, C
    size_t padded = 0;
    View view[3];
    double *flux = nullptr;
    size_t flux_elems = 0;
    int *cutid = nullptr, *cutlist = nullptr;
    double *cutbuf = nullptr;
    double *cutgeom = nullptr;      // static geometry of the cut cells ([4][ncut])
    unsigned *cutmask = nullptr;    // fluid-neighbour masks ([4][ncut])
    int ncut = 0;
    // level above 0 of a hierarchy: the cut cells (list indices) whose 5^3
    // neighbourhood reaches a cell its boxes do not cover; only their
    // geometry changes with the two-level fill
    int *geomsub = nullptr;
    int ngeomsub = 0;
    int geom_gen = 0;  // incremented whenever the cut-cell list and geometry are rebuilt
    double *partials = nullptr;
    size_t partials_cap = 0;
    TimeState *ts = nullptr;
    bool need_est = true;
    // split mode (mf_level_evolve / mf_level_advance_chemistry): only_flow of
    // the steps being enqueued, the graph's value of it, and the other state
    // buffer of the reference (phi_old) that the chemistry phase's first
    // stage reads: lag_mode 0 unknown (after a coupled step), 1 equal to the
    // current state (after set_state), 2 held in lag
    int only_flow = 0, graph_only_flow = 0;
    double *lag = nullptr;
    int lag_mode = 1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    struct AdjState *adj = nullptr;
    int post_slots = 1;  // partial-sum slots written by the post-step pass (one per block)
    // a level of a hierarchy (mf_hier_*): `partial` for a level above 0 whose
    // boxes cover part of its domain (its ghost cells come from the two-level
    // fill; its cut list holds the cut cells of its boxes only); `hier` the
    // owning hierarchy, whose driver alone steps the level
    bool partial = false;
    void *hier = nullptr;
};

// Discrete adjoint of a fixed-dt run of nsteps steps (mf_adjoint_forward /
// mf_adjoint_backward): binomial checkpoints of the level state, the stage
// arrays of the step being reversed and the adjoint arrays.
struct AdjState {
    int nsteps = 0, ncheck = 0, spec = 0, ncut = -1;
    bool ready = false;
    double t0 = 0.0, dt = 0.0;
    double *D = nullptr;                         // S3 of the recomputed step
    double *lam[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // U^{n+1}, A (U^n), B, C, D
    double *cutv[3] = {nullptr, nullptr, nullptr};                   // cut-cell values per stage
    double *lcut[3] = {nullptr, nullptr, nullptr};                   // their adjoints
    double *grad = nullptr;                      // [MF_MAX_REACTIONS + 1]
    double *J = nullptr;
    std::vector<double *> pool;                  // free snapshot buffers
    std::vector<double *> bufs;                  // all snapshot buffers
    std::map<int, double *> snaps;               // step -> snapshot of U^step
    // gather form of the cut-cell adjoint: per-cut-cell records and the fluid
    // cells that receive them (rebuilt when the level's geometry changes)
    double *cutcoef = nullptr;
    int *nbr = nullptr;
    unsigned *nbrmask = nullptr;
    int nnbr = 0, geom_gen = -1;
};

int level_alloc(Ctx *c, const mf_box *boxes, int nboxes, bool partial, Level **out);
int level_free(Level *L);
int level_autotune(Level *L);

const KernelSet *pick_kernels(const mf_plugin_params &t) {
    if (const CaseFamily *f = find_case_family(t.kind)) return t.nspec == f->nspec ? f->ks : nullptr;
    switch (t.kind) {
    case MF_PLUGIN_SHOCKTUBE_N2: return (t.nspec == 1 && t.ngas == 1) ? &ks_shocktube_n2 : nullptr;
    case MF_PLUGIN_CATALYST_CYLINDER: return (t.nspec == 4 && t.ngas == 3) ? &ks_catalyst_cylinder : nullptr;
    case MF_PLUGIN_POROUS_ARRHENIUS:
        if (t.nspec == 5 && t.ngas == 4) return &ks_porous5;
        if (t.nspec == 10 && t.ngas == 9) return &ks_porous10;
        return nullptr;
    }
    return nullptr;
}

View padded_view(double *p, const int *n, int nv) {
    View v;
    v.p = p;
    v.b0 = v.b1 = v.b2 = -MF_NGROW;
    v.js = n[0] + 2 * MF_NGROW;
    v.ks = v.js * (int64_t)(n[1] + 2 * MF_NGROW);
    v.ns = v.ks * (int64_t)(n[2] + 2 * MF_NGROW);
    (void)nv;
    return v;
}

View from_array4(const mf_array4 &a) {
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

int div_up(int a, int b) { return (a + b - 1) / b; }

// the view holds bx (grown by one cell at the high end of direction hi_dir, if 0..2) and nv components
bool covers(const mf_array4 &a, const mf_box &bx, int nv, int hi_dir = -1) {
    if (!a.p || a.ncomp < nv) return false;
    for (int d = 0; d < 3; ++d) {
        const int hi = bx.hi[d] + (d == hi_dir ? 1 : 0);
        if (bx.lo[d] < a.begin[d] || hi >= a.end[d]) return false;
    }
    return true;
}

// ---------------------------------------------------------------------------
// plugin-independent kernels
// ---------------------------------------------------------------------------
__global__ void k_pack(const double *__restrict__ dense, View pad, int nx, int ny, int nz, int nv, int to_pad) {
    const int64_t ncell = (int64_t)nx * ny * nz;
    const int64_t total = ncell * nv;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        int64_t c = q / ncell, r = q % ncell;
        int i = (int)(r % nx), j = (int)((r / nx) % ny), k = (int)(r / ((int64_t)nx * ny));
        if (to_pad) pad(i, j, k, (int)c) = dense[q];
        else const_cast<double *>(dense)[q] = pad(i, j, k, (int)c);
    }
}

__global__ void k_combine(int stage, View snew, View src, View dsdt, int lo0, int lo1, int lo2, int ex, int ey, int ez, int nv,
                          double dt) {
    const int64_t ncell = (int64_t)ex * ey * ez;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < ncell * nv; q += (int64_t)gridDim.x * blockDim.x) {
        int n = (int)(q / ncell);
        int64_t r = q % ncell;
        int i = lo0 + (int)(r % ex), j = lo1 + (int)((r / ex) % ey), k = lo2 + (int)(r / ((int64_t)ex * ey));
        double d = dsdt(i, j, k, n);
        if (stage == 1) snew(i, j, k, n) = kOne * src(i, j, k, n) + dt * d;
        else if (stage == 2) {
            double v = snew(i, j, k, n);
            v += dt * d;
            v += 3.0 * src(i, j, k, n);
            snew(i, j, k, n) = v * kFourth;
        } else {
            double v = snew(i, j, k, n);
            v += dt * d;
            v = src(i, j, k, n) + kTwo * v;
            snew(i, j, k, n) = v * kThird;
        }
    }
}

__global__ void k_scatter_cut(View sb, const double *tmp, const int *flag, int lo0, int lo1, int lo2, int ex, int ey, int ez,
                              int nv) {
    const int64_t ncell = (int64_t)ex * ey * ez;
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= ncell || !flag[q]) return;
    int i = lo0 + (int)(q % ex), j = lo1 + (int)((q / ex) % ey), k = lo2 + (int)(q / ((int64_t)ex * ey));
    for (int n = 0; n < nv; ++n) sb(i, j, k, n) = tmp[(int64_t)n * ncell + q];
}

// ComputeDt (Source/mflo_timestep.cpp:18-50) + EstTimeStep's final formula (:132-152)
__global__ void k_compute_dt(TimeState *ts, double dx0, double dx1, double dx2, double cfl, int fixed, double dt_fixed) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (!ts->active || ts->err) return;
    if (!(ts->t < ts->stop_time)) { ts->active = 0; return; }
    const double dx[3] = {dx0, dx1, dx2};
    double est = DBL_MAX;
    if (ts->max_ws > 0)
        for (int d = 0; d < 3; ++d) est = fmin(est, dx[d] / ts->max_ws);
    if (ts->max_D > 0)
        for (int d = 0; d < 3; ++d) est = fmin(est, kHalf * dx[d] * dx[d] / ts->max_D / 3);
    est *= cfl;
    if (fixed) est = dt_fixed;
    const double change_max = 1.1;
    double dt0 = est;
    double dtl = fmin(est, change_max * ts->dt_prev);
    dt0 = fmin(dt0, 1 * dtl);
    if (fixed) dt0 = dt_fixed;
    const double eps = 1.e-3 * dt0;
    if (ts->t + dt0 > ts->stop_time - eps) dt0 = ts->stop_time - ts->t;
    ts->dt = dt0;
}

// reduces the per-tile partials; mode 0 closes a step, mode 1 only refreshes
// the dt estimate inputs
__global__ __launch_bounds__(1024) void k_finalize(const double *partials, int ntiles, int nv, TimeState *ts, int *err,
                                                  int mode, int track) {
    __shared__ double red[1024];
    const int np = nv + 2;
    if (mode == 0 && (!ts->active || ts->err)) {
        if (threadIdx.x == 0 && *err && !ts->err) ts->err = *err;
        return;
    }
    // thread folds whole tiles (contiguous reads), then one block reduction
    // per value; the arrays are indexed with constants only (no scratch)
    double acc[kMaxNV + 2];
#pragma unroll
    for (int q = 0; q < kMaxNV + 2; ++q) acc[q] = 0.0;
    for (int t = threadIdx.x; t < ntiles; t += blockDim.x) {
        const double *pt = partials + (size_t)t * np;
#pragma unroll
        for (int q = 0; q < kMaxNV + 2; ++q) {
            if (q < np) {
                const double x = pt[q];
                acc[q] = (q >= nv) ? fmax(acc[q], x) : acc[q] + x;
            }
        }
    }
    double vals[kMaxNV + 2];
#pragma unroll
    for (int q = 0; q < kMaxNV + 2; ++q) {
        vals[q] = 0.0;
        if (q >= np) continue;
        const bool is_max = q >= nv;
        red[threadIdx.x] = acc[q];
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) {
                double a = red[threadIdx.x], b = red[threadIdx.x + s];
                red[threadIdx.x] = is_max ? fmax(a, b) : a + b;
            }
            __syncthreads();
        }
        vals[q] = red[0];
        __syncthreads();
    }
    double mws = 0.0, mD = 0.0;
#pragma unroll
    for (int q = 0; q < kMaxNV + 2; ++q) {
        if (q == nv) mws = vals[q];
        if (q == nv + 1) mD = vals[q];
    }
    if (threadIdx.x != 0) return;
    ts->max_ws = mws;
    ts->max_D = mD;
    ts->have_est = 1;
    if (mode == 1) return;
    if (track) {
#pragma unroll
        for (int q = 0; q < kMaxNV; ++q)
            if (q < nv) ts->resnorm[q] = sqrt(vals[q]);
    }
    ts->t = ts->t + ts->dt;
    ts->dt_prev = ts->dt;
    ts->istep += 1;
    if (*err) ts->err = *err;
    if (ts->t >= ts->stop_time - 1.e-6 * ts->dt) ts->active = 0;
}

__global__ void k_copy(double *__restrict__ dst, const double *__restrict__ src, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// level helpers
// ---------------------------------------------------------------------------
StageArgs stage_args(Level *L, int stage) {
    Ctx *c = L->c;
    StageArgs a;
    a.tab = c->d_tab;
    a.src = L->view[stage == 1 ? 0 : (stage == 2 ? 1 : 2)];
    a.cutid = L->ncut ? L->cutid : nullptr;
    a.cutbuf = L->cutbuf;
    a.ncut = L->ncut;
    a.nx = L->n[0];
    a.ny = L->n[1];
    for (int d = 0; d < 3; ++d) a.dx[d] = c->dx[d];
    a.ts = L->ts;
    a.time_host = 0.0;
    a.dt_host = 0.0;
    a.stage = stage;
    a.order = c->prm.order_hyp;
    a.dissfactor = c->prm.dissfactor;
    a.nsflag = c->prm.do_ns;
    a.spec_in_solid = c->prm.species_in_solid;
    a.conj_ht = c->prm.conj_ht;
    a.flux_full = 0;
    a.err = c->d_err;
    return a;
}

Work level_work(Level *L, int cls) {
    Work w;
    w.boxes = L->d_box;
    w.nboxes = (int)L->boxes.size();
    w.t0 = 0;
    w.t1 = L->ntiles[cls];
    w.cls = cls;
    for (int d = 0; d < 3; ++d) w.tsh[d] = L->tsh[cls][d];
    std::memset(&w.one, 0, sizeof(w.one));
    return w;
}
int box_tiles(const BoxDesc &bd, int cls) { return bd.ntile[cls][0] * bd.ntile[cls][1] * bd.ntile[cls][2]; }

FillArgs fill_args(Level *L, double *arr, int stage) {
    Ctx *c = L->c;
    FillArgs f;
    f.tab = c->d_tab;
    f.s = padded_view(arr, L->n, c->nv);
    for (int d = 0; d < 3; ++d) {
        f.n[d] = L->n[d];
        f.periodic[d] = c->geom.is_periodic[d];
        f.lo_bc[d] = c->prm.lo_bc[d];
        f.hi_bc[d] = c->prm.hi_bc[d];
        f.prob_lo[d] = c->geom.prob_lo[d];
        f.dx[d] = c->dx[d];
    }
    f.ts = L->ts;
    f.time_host = 0.0;
    f.stage = stage;
    return f;
}

CutArgs cut_args(Level *L, double *arr) {
    Ctx *c = L->c;
    CutArgs a;
    std::memset(&a, 0, sizeof(a));
    a.tab = c->d_tab;
    a.s = padded_view(arr, L->n, c->nv);
    a.list = L->cutlist;
    a.cutbuf = L->cutbuf;
    a.ncut = L->ncut;
    a.geom = L->cutgeom;
    a.mask = L->cutmask;
    for (int d = 0; d < 3; ++d) a.dx[d] = c->dx[d];
    a.spec_in_solid = c->prm.species_in_solid;
    a.conj_ht = c->prm.conj_ht;
    a.ts = L->ts;
    return a;
}

int rebuild_tiles(Level *L) {
    Ctx *c = L->c;
    for (int q = 0; q < kTileClasses; ++q)
        for (int d = 0; d < 3; ++d) L->tsh[q][d] = std::max(1, L->tsh[q][d]);
    const int NC = NCV + c->ns;
    int tiles[kTileClasses] = {0, 0, 0, 0, 0, 0, 0};
    size_t fofs = 0;
    L->hbox.resize(L->boxes.size());
    for (size_t b = 0; b < L->boxes.size(); ++b) {
        BoxDesc &bd = L->hbox[b];
        std::memset(&bd, 0, sizeof(bd));
        const mf_box &B = L->boxes[b];
        int len[3];
        for (int d = 0; d < 3; ++d) {
            bd.lo[d] = B.lo[d];
            bd.hi[d] = B.hi[d];
            len[d] = B.hi[d] - B.lo[d] + 1;
        }
        for (int q = 0; q < kTileClasses; ++q) {
            for (int d = 0; d < 3; ++d) bd.ntile[q][d] = div_up(len[d], L->tsh[q][d]);
            bd.tile_start[q] = tiles[q];
            tiles[q] += box_tiles(bd, q);
        }
        for (int d = 0; d < 3; ++d) {
            int fl[3] = {len[0], len[1], len[2]};
            fl[d] += 1;
            View &v = bd.flux[d];
            v.p = L->flux + fofs;
            v.b0 = B.lo[0];
            v.b1 = B.lo[1];
            v.b2 = B.lo[2];
            v.js = fl[0];
            v.ks = (int64_t)fl[0] * fl[1];
            v.ns = v.ks * fl[2];
            fofs += (size_t)v.ns * NC;
        }
    }
    for (int q = 0; q < kTileClasses; ++q) L->ntiles[q] = tiles[q];
    if (L->d_box) hipFree(L->d_box);
    HIPCHK(hipMalloc(&L->d_box, sizeof(BoxDesc) * L->hbox.size()));
    HIPCHK(hipMemcpy(L->d_box, L->hbox.data(), sizeof(BoxDesc) * L->hbox.size(), hipMemcpyHostToDevice));
    size_t need = (size_t)tiles[kCellClass] * (c->nv + 2);
    if (need > L->partials_cap) {
        if (L->partials) hipFree(L->partials);
        HIPCHK(hipMalloc(&L->partials, need * sizeof(double)));
        L->partials_cap = need;
    }
    if (L->gexec) { hipGraphExecDestroy(L->gexec); L->gexec = nullptr; }
    if (L->graph) { hipGraphDestroy(L->graph); L->graph = nullptr; }
    return MF_OK;
}

int build_cutlist(Level *L, const double *vfrac_host) {
    const int nx = L->n[0], ny = L->n[1], nz = L->n[2];
    const int64_t ncell = (int64_t)nx * ny * nz;
    std::vector<int> id(ncell, -1), li, lj, lk;
    std::vector<unsigned char> cov;
    if (L->partial) {
        cov.assign((size_t)ncell, 0);
        for (const mf_box &B : L->boxes)
            for (int k = B.lo[2]; k <= B.hi[2]; ++k)
                for (int j = B.lo[1]; j <= B.hi[1]; ++j)
                    std::memset(cov.data() + (size_t)nx * (j + (size_t)ny * k) + B.lo[0], 1, (size_t)(B.hi[0] - B.lo[0] + 1));
    }
    int nc = 0;
    for (int64_t q = 0; q < ncell; ++q) {
        double v = vfrac_host[q];
        if (v > 0.0 && v < 1.0 && (cov.empty() || cov[(size_t)q])) {
            id[q] = nc++;
            li.push_back((int)(q % nx));
            lj.push_back((int)((q / nx) % ny));
            lk.push_back((int)(q / ((int64_t)nx * ny)));
        }
    }
    // cut cells whose neighbourhood (radius 2) leaves the covered region or the domain
    std::vector<int> sub;
    if (L->partial) {
        for (int q = 0; q < nc; ++q) {
            bool edge = false;
            for (int n = -2; n <= 2 && !edge; ++n)
                for (int m = -2; m <= 2 && !edge; ++m)
                    for (int l = -2; l <= 2 && !edge; ++l) {
                        const int i = li[q] + l, j = lj[q] + m, k = lk[q] + n;
                        if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz ||
                            !cov[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)])
                            edge = true;
                    }
            if (edge) sub.push_back(q);
        }
    }
    HIPCHK(hipStreamSynchronize(L->c->stream));
    if (L->cutlist) { hipFree(L->cutlist); L->cutlist = nullptr; }
    if (L->cutbuf) { hipFree(L->cutbuf); L->cutbuf = nullptr; }
    if (L->cutgeom) { hipFree(L->cutgeom); L->cutgeom = nullptr; }
    if (L->cutmask) { hipFree(L->cutmask); L->cutmask = nullptr; }
    if (L->geomsub) { hipFree(L->geomsub); L->geomsub = nullptr; }
    L->ngeomsub = 0;
    L->ncut = nc;
    ++L->geom_gen;
    HIPCHK(hipMemcpy(L->cutid, id.data(), sizeof(int) * ncell, hipMemcpyHostToDevice));
    if (nc) {
        std::vector<int> lst;
        lst.reserve(3 * (size_t)nc);
        lst.insert(lst.end(), li.begin(), li.end());
        lst.insert(lst.end(), lj.begin(), lj.end());
        lst.insert(lst.end(), lk.begin(), lk.end());
        HIPCHK(hipMalloc(&L->cutlist, sizeof(int) * lst.size()));
        HIPCHK(hipMemcpy(L->cutlist, lst.data(), sizeof(int) * lst.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMalloc(&L->cutbuf, sizeof(double) * (size_t)nc * L->c->nv));
        HIPCHK(hipMalloc(&L->cutgeom, sizeof(double) * 6 * (size_t)nc));
        HIPCHK(hipMalloc(&L->cutmask, sizeof(unsigned) * 4 * (size_t)nc));
        // the geometry reads vfrac of the 5^3 neighbourhoods: fill the ghost
        // layers first. A partly covered level computes it here from its valid
        // data for the cut cells whose neighbourhoods its boxes cover, and
        // after each of its two-level fills for the listed others (their ghost
        // vfrac comes from the coarse level)
        if (!L->partial) {
            FillArgs f = fill_args(L, L->arr[0], 1);
            f.ts = nullptr;
            f.time_host = 0.0;
            for (int ph = 0; ph <= 3; ++ph) L->c->ks->fill(f, ph, L->c->stream);
            L->c->ks->cut_geom(cut_args(L, L->arr[0]), L->cutgeom, L->cutmask, L->c->stream);
        } else {
            L->c->ks->cut_geom(cut_args(L, L->arr[0]), L->cutgeom, L->cutmask, L->c->stream);
            if (!sub.empty()) {
                HIPCHK(hipMalloc(&L->geomsub, sizeof(int) * sub.size()));
                HIPCHK(hipMemcpy(L->geomsub, sub.data(), sizeof(int) * sub.size(), hipMemcpyHostToDevice));
                L->ngeomsub = (int)sub.size();
            }
        }
        HIPCHK(hipGetLastError());
    }
    if (L->gexec) { hipGraphExecDestroy(L->gexec); L->gexec = nullptr; }
    if (L->graph) { hipGraphDestroy(L->graph); L->graph = nullptr; }
    return MF_OK;
}

// launch of the face and residual kernels, one launch per box (mode 0) or one
// persistent launch over the whole tile list
// persistent grid: at most max_grid blocks, a multiple of 8 (one share per XCD)
int grid_for(int ntiles, int max_grid) {
    int g = std::min(ntiles, max_grid);
    if (g >= 8) g &= ~7;
    return std::max(g, 1);
}

bool use_lds(Level *L) { return L->c->prm.order_hyp == 5 && L->c->prm.do_ns; }

// threads of an LDS face-sweep block: 512 for tiles of at least 512 faces (a
// box-end tile has one more face per row along the sweep, which at 256 faces
// per block leaves a mostly idle extra wave pass), else 256
int face_block(const int *t) { return t[0] * t[1] * t[2] >= 512 ? 512 : 256; }

void launch_faces(Level *L, const StageArgs &a, int dir, int mode) {
    Ctx *c = L->c;
    const bool lds = use_lds(L);
    const int cls = lds ? dir : kCellClass;
    Work w = level_work(L, cls);
    const size_t shm = lds ? c->ks->lds_bytes(L->tsh[cls], dir) : 0;
    auto go = [&](int grid) {
        if (lds) c->ks->faces_lds(a, w, dir, grid, face_block(L->tsh[cls]), shm, c->stream);
        else c->ks->faces(a, w, dir, grid, 256, c->stream);
    };
    if (mode == MF_MODE_PER_BOX) {
        for (size_t b = 0; b < L->hbox.size(); ++b) {
            const BoxDesc &bd = L->hbox[b];
            w.t0 = bd.tile_start[cls];
            w.t1 = w.t0 + box_tiles(bd, cls);
            go(w.t1 - w.t0);
        }
    } else {
        go(grid_for(L->ntiles[cls], L->max_grid));
    }
}
void launch_residual(Level *L, const ResArgs &r, int mode) {
    Ctx *c = L->c;
    Work w = level_work(L, kCellClass);
    if (mode == MF_MODE_PER_BOX) {
        for (size_t b = 0; b < L->hbox.size(); ++b) {
            const BoxDesc &bd = L->hbox[b];
            w.t0 = bd.tile_start[kCellClass];
            w.t1 = w.t0 + box_tiles(bd, kCellClass);
            c->ks->residual(r, w, w.t1 - w.t0, 256, c->stream);
        }
    } else {
        c->ks->residual(r, w, grid_for(L->ntiles[kCellClass], L->max_grid), 256, c->stream);
    }
}

void enqueue_stage(Level *L, int stage, int mode) {
    Ctx *c = L->c;
    double *src = L->arr[stage == 1 ? 0 : (stage == 2 ? 1 : 2)];
    FillArgs f = fill_args(L, src, stage);
    L->timer.begin(KC_FILL, c->stream);
    for (int ph = 0; ph <= 3; ++ph) c->ks->fill(f, ph, c->stream);
    L->timer.end(c->stream);
    if (L->ncut) {
        L->timer.begin(KC_CUT, c->stream);
        c->ks->cut(cut_args(L, src), c->stream);
        L->timer.end(c->stream);
    }
    StageArgs a = stage_args(L, stage);
    for (int d = 0; d < 3; ++d) {
        L->timer.begin(KC_FACES, c->stream);
        launch_faces(L, a, d, mode);
        L->timer.end(c->stream);
    }
    ResArgs r;
    r.st = a;
    r.combine = 1;
    std::memset(&r.dsdt, 0, sizeof(r.dsdt));
    r.old = L->view[0];
    // stage 3 writes A in place with chemistry; without, B keeps S3 so that the
    // post pass still sees U^n in A for the residual
    const int chem = c->prm.do_chemistry && !L->only_flow;
    r.dst = L->view[stage == 1 ? 1 : (stage == 2 ? 2 : (chem ? 0 : 1))];
    r.do_chem = chem;
    r.clip = c->prm.clip_species;
    r.track_res = c->prm.track_residual_norms;
    r.do_ns = c->prm.do_ns;
    r.partials = L->partials;
    L->timer.begin(KC_RESID, c->stream);
    launch_residual(L, r, mode);
    L->timer.end(c->stream);
    if (stage == 3) {
        r.st.src = r.dst;
        r.dst = L->view[0];
        L->timer.begin(KC_POST, c->stream);
        Work w = level_work(L, kCellClass);
        if (mode == MF_MODE_PER_BOX) {
            for (size_t b = 0; b < L->hbox.size(); ++b) {
                const BoxDesc &bd = L->hbox[b];
                w.t0 = bd.tile_start[kCellClass];
                w.t1 = w.t0 + box_tiles(bd, kCellClass);
                c->ks->post(r, w, w.t1 - w.t0, 256, c->stream);
            }
            L->post_slots = L->ntiles[kCellClass];
        } else {
            L->post_slots = grid_for(L->ntiles[kCellClass], L->max_grid);
            c->ks->post(r, w, L->post_slots, 256, c->stream);
        }
        L->timer.end(c->stream);
    }
}

void enqueue_step(Level *L, int mode) {
    Ctx *c = L->c;
    L->timer.begin(KC_DT, c->stream);
    hipLaunchKernelGGL(k_compute_dt, dim3(1), dim3(1), 0, c->stream, L->ts, c->dx[0], c->dx[1], c->dx[2], c->prm.cfl,
                       c->prm.fixed_dt, c->prm.dt_fixed);
    L->timer.end(c->stream);
    for (int s = 1; s <= 3; ++s) enqueue_stage(L, s, mode);
    L->timer.begin(KC_FINAL, c->stream);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, c->stream, L->partials, L->post_slots, c->nv, L->ts, c->d_err, 0,
                       c->prm.track_residual_norms);
    L->timer.end(c->stream);
}

void enqueue_est(Level *L) {
    Ctx *c = L->c;
    ResArgs r;
    std::memset(&r, 0, sizeof(r));
    r.st = stage_args(L, 1);
    r.st.ts = nullptr;
    r.do_ns = c->prm.do_ns;
    r.partials = L->partials;
    Work w = level_work(L, kCellClass);
    const int grid = std::min(L->ntiles[kCellClass], L->max_grid);
    c->ks->est(r, w, grid, 256, c->stream);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, c->stream, L->partials, grid, c->nv, L->ts, c->d_err, 1, 0);
}

// ---------------------------------------------------------------------------
// discrete adjoint: host side
// ---------------------------------------------------------------------------
__global__ void k_set_ts(TimeState *ts, double t, double dt) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    ts->t = t;
    ts->dt = dt;
    ts->dt_prev = dt;
    ts->active = 1;
    ts->stop_time = 1.e300;
}

__global__ void k_zero(double *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

// Adjoint of the RK3 stage arithmetic over the valid cells:
//   stage 3: S3 = (U^n + 2 (S2 + dt d)) third      -> la = third lo, lx = 2 third lo
//   stage 2: S2 = ((S1 + dt d) + 3 U^n) / 4        -> lx = lo / 4, la += 3/4 lo
//   stage 1: S1 = Sb + dt d (Sb: cut values at cut cells) -> lcut = lo (cut cells), la += lo
// (the dsdt part goes through k_faces_adj)
__global__ void k_adj_comb(int stage, View lo, View la, View lx, const int *cutid, double *lcut, int ncut, int nx, int ny,
                           int nz, int nv) {
    const int64_t ncell = (int64_t)nx * ny * nz;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < ncell; q += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((int64_t)nx * ny));
        const int cid = stage == 1 ? cutid[q] : -1;
        for (int n = 0; n < nv; ++n) {
            const double v = lo(i, j, k, n);
            if (stage == 3) {
                la(i, j, k, n) = kThird * v;
                lx(i, j, k, n) = kTwo * kThird * v;
            } else if (stage == 2) {
                lx(i, j, k, n) = kFourth * v;
                la(i, j, k, n) += 3.0 * kFourth * v;
            } else {
                if (cid >= 0) lcut[(int64_t)n * ncut + cid] = v;
                else la(i, j, k, n) += v;
            }
        }
    }
}

void adj_free(AdjState *A) {
    if (!A) return;
    if (A->D) hipFree(A->D);
    for (int q = 0; q < 5; ++q) if (A->lam[q]) hipFree(A->lam[q]);
    for (int q = 0; q < 3; ++q) {
        if (A->cutv[q]) hipFree(A->cutv[q]);
        if (A->lcut[q]) hipFree(A->lcut[q]);
    }
    if (A->grad) hipFree(A->grad);
    if (A->J) hipFree(A->J);
    if (A->cutcoef) hipFree(A->cutcoef);
    if (A->nbr) hipFree(A->nbr);
    if (A->nbrmask) hipFree(A->nbrmask);
    for (double *b : A->bufs) hipFree(b);
    delete A;
}

// fluid cells that receive cut-cell adjoint terms. k_nbr_flag marks every
// fluid neighbour of a listed cut cell (its mask bits) in a padded byte array;
// k_nbr_list, per flagged cell, sets bit (n+2)*25 + (m+2)*5 + (l+2) of its
// mask for each listed cut cell at offset -(l, m, n), counts the cells
// (pass 0) or writes them (pass 1, one slot per cell claimed per wave).
__global__ void k_nbr_flag(const int *list, const unsigned *mask, int ncut, unsigned char *flag, int nx, int ny) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ncut) return;
    const int G = MF_NGROW;
    const int64_t PX = nx + 2 * G, PY = ny + 2 * G;
    const int i = list[q], j = list[ncut + q], k = list[2 * ncut + q];
    for (int w = 0; w < 4; ++w) {
        unsigned bits = mask[(int64_t)w * ncut + q];
        while (bits) {
            const int nb = w * 32 + __builtin_ctz(bits);
            bits &= bits - 1u;
            const int l = nb % 5 - 2, m = (nb / 5) % 5 - 2, n = nb / 25 - 2;
            flag[(i + l + G) + PX * ((j + m + G) + PY * (int64_t)(k + n + G))] = 1;
        }
    }
}

__global__ __launch_bounds__(256) void k_nbr_list(int pass, const unsigned char *flag, const int *cutid, int nx, int ny, int nz,
                                                  int *count, int *nbr, unsigned *nbrmask, int cap) {
    const int G = MF_NGROW;
    const int PX = nx + 2 * G, PY = ny + 2 * G, PZ = nz + 2 * G;
    const int64_t npad = (int64_t)PX * PY * PZ;
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    unsigned mb[4] = {0u, 0u, 0u, 0u};
    int i = 0, j = 0, k = 0;
    if (p < npad && flag[p]) {
        i = (int)(p % PX) - G;
        j = (int)((p / PX) % PY) - G;
        k = (int)(p / ((int64_t)PX * PY)) - G;
        for (int nb = 0; nb < 125; ++nb) {
            const int l = nb % 5 - 2, m = (nb / 5) % 5 - 2, n = nb / 25 - 2;
            const int ci = i - l, cj = j - m, ck = k - n;
            if (ci < 0 || ci >= nx || cj < 0 || cj >= ny || ck < 0 || ck >= nz) continue;
            if (cutid[ci + (int64_t)nx * (cj + (int64_t)ny * ck)] >= 0) mb[nb >> 5] |= 1u << (nb & 31);
        }
    }
    const bool has = (mb[0] | mb[1] | mb[2] | mb[3]) != 0u;
    const unsigned long long bal = __ballot(has);
    if (bal == 0ull) return;
    const int lane = threadIdx.x & 63, lead = __builtin_ctzll(bal);
    int base = 0;
    if (lane == lead) base = atomicAdd(count, __popcll(bal));
    base = __shfl(base, lead, 64);
    if (pass == 1 && has) {
        const int slot = base + __popcll(bal & ((1ull << lane) - 1ull));
        if (slot < cap) {
            nbr[slot] = i;
            nbr[cap + slot] = j;
            nbr[2 * cap + slot] = k;
            for (int w = 0; w < 4; ++w) nbrmask[(int64_t)w * cap + slot] = mb[w];
        }
    }
}

int adj_build_nbr(Level *L) {
    Ctx *c = L->c;
    AdjState *A = L->adj;
    if (A->nbr) { hipFree(A->nbr); A->nbr = nullptr; }
    if (A->nbrmask) { hipFree(A->nbrmask); A->nbrmask = nullptr; }
    A->nnbr = 0;
    A->geom_gen = L->geom_gen;
    if (L->ncut == 0) return MF_OK;
    const int nx = L->n[0], ny = L->n[1], nz = L->n[2], G = MF_NGROW;
    const int64_t npad = (int64_t)(nx + 2 * G) * (ny + 2 * G) * (nz + 2 * G);
    unsigned char *flag = nullptr;
    int *cnt = nullptr;
    HIPCHK(hipMalloc(&flag, npad));
    HIPCHK(hipMalloc(&cnt, sizeof(int)));
    HIPCHK(hipMemsetAsync(flag, 0, npad, c->stream));
    HIPCHK(hipMemsetAsync(cnt, 0, sizeof(int), c->stream));
    hipLaunchKernelGGL(k_nbr_flag, dim3((L->ncut + 127) / 128), dim3(128), 0, c->stream, L->cutlist, L->cutmask, L->ncut, flag, nx, ny);
    const unsigned nblk = (unsigned)((npad + 255) / 256);
    hipLaunchKernelGGL(k_nbr_list, dim3(nblk), dim3(256), 0, c->stream, 0, flag, L->cutid, nx, ny, nz, cnt, nullptr, nullptr, 0);
    int n = 0;
    HIPCHK(hipMemcpyAsync(&n, cnt, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (n > 0) {
        HIPCHK(hipMalloc(&A->nbr, sizeof(int) * 3 * (size_t)n));
        HIPCHK(hipMalloc(&A->nbrmask, sizeof(unsigned) * 4 * (size_t)n));
        HIPCHK(hipMemsetAsync(cnt, 0, sizeof(int), c->stream));
        hipLaunchKernelGGL(k_nbr_list, dim3(nblk), dim3(256), 0, c->stream, 1, flag, L->cutid, nx, ny, nz, cnt, A->nbr, A->nbrmask, n);
        int n2 = 0;
        HIPCHK(hipMemcpyAsync(&n2, cnt, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (n2 != n) {
            hipFree(flag);
            hipFree(cnt);
            return MF_ERR_HIP;
        }
    }
    A->nnbr = n;
    hipFree(flag);
    hipFree(cnt);
    HIPCHK(hipGetLastError());
    return MF_OK;
}

int adj_alloc(Level *L, int ncheck) {
    Ctx *c = L->c;
    AdjState *A = L->adj;
    if (A && A->ncheck == ncheck && A->ncut == L->ncut)
        return A->geom_gen == L->geom_gen ? MF_OK : adj_build_nbr(L);
    if (A) { adj_free(A); L->adj = nullptr; }
    A = new (std::nothrow) AdjState();
    if (!A) return MF_ERR_BAD_ARG;
    L->adj = A;
    A->ncheck = ncheck;
    A->ncut = L->ncut;
    const size_t pb = L->padded * sizeof(double);
    const size_t cb = (size_t)std::max(L->ncut, 1) * c->nv * sizeof(double);
    HIPCHK(hipMalloc(&A->D, pb));
    for (int q = 0; q < 5; ++q) {
        HIPCHK(hipMalloc(&A->lam[q], pb));
        HIPCHK(hipMemset(A->lam[q], 0, pb));
    }
    for (int q = 0; q < 3; ++q) {
        HIPCHK(hipMalloc(&A->cutv[q], cb));
        HIPCHK(hipMalloc(&A->lcut[q], cb));
    }
    HIPCHK(hipMalloc(&A->grad, (MF_MAX_REACTIONS + 1) * sizeof(double)));
    HIPCHK(hipMalloc(&A->J, sizeof(double)));
    HIPCHK(hipMalloc(&A->cutcoef, (size_t)std::max(L->ncut, 1) * c->ks->adj->cut_rec * sizeof(double)));
    for (int q = 0; q < ncheck; ++q) {
        double *b = nullptr;
        HIPCHK(hipMalloc(&b, pb));
        A->bufs.push_back(b);
    }
    return adj_build_nbr(L);
}

// binomial split of the steps [a, b) with s free snapshots: the first
// checkpoint after a (Griewank's beta(s, t) = C(s + t, s) rule)
int adj_split(int a, int b, int s) {
    const int l = b - a;
    auto beta = [](int s_, int t_) -> double {
        double v = 1.0;
        for (int q = 1; q <= s_; ++q) v = v * (double)(t_ + q) / (double)q;
        return v;
    };
    int t = 1;
    while (beta(s, t) < (double)l) ++t;
    const double right = beta(s - 1, t);
    int left = (double)l > right ? (int)((double)l - right) : 1;
    left = std::max(1, std::min(left, l - 1));
    return a + left;
}

void adj_objective(Level *L) {
    Ctx *c = L->c;
    ObjArgs o;
    o.tab = c->d_tab;
    o.s = L->view[0];
    o.ts = L->ts;
    o.spec = L->adj->spec;
    o.outlet_i = L->n[0] - 1;
    o.ny = L->n[1];
    o.nz = L->n[2];
    o.dA = c->dx[1] * c->dx[2];
    o.J = L->adj->J;
    c->ks->adj->objective(o, c->stream);
}

void adj_copy(Level *L, double *dst, const double *src) {
    const int64_t n = (int64_t)L->padded;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 16384)), dim3(256), 0, L->c->stream, dst, src, n);
}
void adj_zero(Level *L, double *p, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 16384)), dim3(256), 0, L->c->stream, p, n);
}
// U^step from its snapshot into A, time t0 + step dt
void adj_restore(Level *L, int step) {
    AdjState *A = L->adj;
    adj_copy(L, L->arr[0], A->snaps[step]);
    hipLaunchKernelGGL(k_set_ts, dim3(1), dim3(1), 0, L->c->stream, L->ts, A->t0 + step * A->dt, A->dt);
}
void adj_store(Level *L, int step) {
    AdjState *A = L->adj;
    double *b = A->pool.back();
    A->pool.pop_back();
    adj_copy(L, b, L->arr[0]);
    A->snaps[step] = b;
}
void adj_release(Level *L, int step) {
    AdjState *A = L->adj;
    auto it = A->snaps.find(step);
    if (it == A->snaps.end()) return;
    A->pool.push_back(it->second);
    A->snaps.erase(it);
}

// forward of one step keeping its stage arrays: A = U^n -> B = S1, C = S2,
// D = S3 (post-flow, before the post-step pass), stage cut values in cutv[]
void adj_forward_keep(Level *L) {
    Ctx *c = L->c;
    AdjState *A = L->adj;
    for (int stage = 1; stage <= 3; ++stage) {
        double *src = L->arr[stage - 1];
        FillArgs f = fill_args(L, src, stage);
        for (int ph = 0; ph <= 3; ++ph) c->ks->fill(f, ph, c->stream);
        StageArgs a = stage_args(L, stage);
        a.cutbuf = A->cutv[stage - 1];
        if (L->ncut) {
            CutArgs ca = cut_args(L, src);
            ca.cutbuf = A->cutv[stage - 1];
            c->ks->cut(ca, c->stream);
        }
        for (int d = 0; d < 3; ++d) launch_faces(L, a, d, MF_MODE_BATCHED);
        ResArgs r;
        std::memset(&r, 0, sizeof(r));
        r.st = a;
        r.combine = 1;
        r.old = L->view[0];
        r.dst = stage == 1 ? L->view[1] : (stage == 2 ? L->view[2] : padded_view(A->D, L->n, c->nv));
        r.do_chem = c->prm.do_chemistry;
        r.do_ns = c->prm.do_ns;
        r.partials = L->partials;
        launch_residual(L, r, MF_MODE_BATCHED);
    }
}

// reverse of one stage: faces (3 sweeps), cut cells, ghost fill
void adj_stage_reverse(Level *L, int stage, View lam_out, double coef, double *lam_in, bool zero_lcut) {
    Ctx *c = L->c;
    AdjState *A = L->adj;
    const AdjSet *S = c->ks->adj;
    double *src = L->arr[stage - 1];
    View lin = padded_view(lam_in, L->n, c->nv);
    if (zero_lcut && L->ncut) adj_zero(L, A->lcut[stage - 1], (int64_t)L->ncut * c->nv);
    AdjFaceArgs fa;
    fa.st = stage_args(L, stage);
    fa.st.cutbuf = A->cutv[stage - 1];
    fa.lam_out = lam_out;
    fa.coef = coef;
    fa.lam = lin;
    fa.lcut = A->lcut[stage - 1];
    for (int d = 0; d < 3; ++d) {
        Work w = level_work(L, kAdjFaceClass + d);
        const size_t shm = S->lds_bytes(L->tsh[kAdjFaceClass + d], d);
        S->faces(fa, w, d, grid_for(L->ntiles[kAdjFaceClass + d], L->max_grid), 128, shm, c->stream);
    }
    if (L->ncut) {
        AdjCutArgs ca;
        ca.c = cut_args(L, src);
        ca.c.cutbuf = A->cutv[stage - 1];
        ca.lam = lin;
        ca.lcut = A->lcut[stage - 1];
        ca.coef = A->cutcoef;
        ca.nbr = A->nbr;
        ca.nbrmask = A->nbrmask;
        ca.nnbr = A->nnbr;
        ca.cutid = L->cutid;
        for (int d = 0; d < 3; ++d) ca.n[d] = L->n[d];
        S->cut(ca, c->stream);
    }
    AdjFillArgs fl;
    fl.f = fill_args(L, src, stage);
    fl.lam = lin;
    fl.grad = A->grad;
    for (int ph = 3; ph >= 0; --ph) S->fill(fl, ph, c->stream);
}

// adjoint of step n: A holds U^n; lam[0] holds the adjoint of U^{n+1} on entry
// and of U^n on exit
void adj_step(Level *L, int n) {
    Ctx *c = L->c;
    AdjState *A = L->adj;
    const AdjSet *S = c->ks->adj;
    const double dt = A->dt;
    hipLaunchKernelGGL(k_set_ts, dim3(1), dim3(1), 0, c->stream, L->ts, A->t0 + n * dt, dt);
    adj_forward_keep(L);
    View lu = padded_view(A->lam[0], L->n, c->nv), la = padded_view(A->lam[1], L->n, c->nv);
    View lb = padded_view(A->lam[2], L->n, c->nv), lc = padded_view(A->lam[3], L->n, c->nv);
    View ld = padded_view(A->lam[4], L->n, c->nv);
    {
        AdjPostArgs pa;
        pa.st = stage_args(L, 3);
        pa.st.src = padded_view(A->D, L->n, c->nv);
        pa.lu = lu;
        pa.ld = ld;
        pa.clip = c->prm.clip_species;
        pa.spec = A->spec;
        pa.dA = c->dx[1] * c->dx[2];
        pa.outlet_i = L->n[0] - 1;
        pa.grad = A->grad;
        Work w = level_work(L, kCellClass);
        S->post(pa, w, grid_for(L->ntiles[kCellClass], L->max_grid), 256, c->stream);
    }
    const unsigned cg = (unsigned)std::min<int64_t>(((int64_t)L->n[0] * L->n[1] * L->n[2] + 255) / 256, 8192);
    hipLaunchKernelGGL(k_adj_comb, dim3(cg), dim3(256), 0, c->stream, 3, ld, la, lc, L->cutid, A->lcut[0], L->ncut, L->n[0],
                       L->n[1], L->n[2], c->nv);
    adj_stage_reverse(L, 3, ld, kTwo * dt * kThird, A->lam[3], true);
    hipLaunchKernelGGL(k_adj_comb, dim3(cg), dim3(256), 0, c->stream, 2, lc, la, lb, L->cutid, A->lcut[0], L->ncut, L->n[0],
                       L->n[1], L->n[2], c->nv);
    adj_stage_reverse(L, 2, lc, dt * kFourth, A->lam[2], true);
    hipLaunchKernelGGL(k_adj_comb, dim3(cg), dim3(256), 0, c->stream, 1, lb, la, lb, L->cutid, A->lcut[0], L->ncut, L->n[0],
                       L->n[1], L->n[2], c->nv);
    adj_stage_reverse(L, 1, lb, dt, A->lam[1], false);
    std::swap(A->lam[0], A->lam[1]);
}

void adj_advance(Level *L, int from, int count, bool objective) {
    for (int q = 0; q < count; ++q) {
        enqueue_step(L, MF_MODE_BATCHED);
        if (objective) adj_objective(L);
    }
    (void)from;
}

// reversal of the steps [a, b) with s free snapshot buffers; U^a is stored
void adj_reverse(Level *L, int a, int b, int s) {
    AdjState *A = L->adj;
    if (b - a == 1) {
        adj_restore(L, a);
        adj_step(L, a);
        return;
    }
    if (s <= 0) {
        for (int st = b - 1; st >= a; --st) {
            adj_restore(L, a);
            adj_advance(L, a, st - a, false);
            adj_step(L, st);
        }
        return;
    }
    const int m = adj_split(a, b, s);
    if (!A->snaps.count(m)) {
        adj_restore(L, a);
        adj_advance(L, a, m - a, false);
        adj_store(L, m);
    }
    adj_reverse(L, m, b, s - 1);
    adj_release(L, m);
    adj_reverse(L, a, m, s);
}

int read_ts(Level *L, TimeState *h) {
    HIPCHK(hipStreamSynchronize(L->c->stream));
    HIPCHK(hipMemcpy(h, L->ts, sizeof(TimeState), hipMemcpyDeviceToHost));
    return MF_OK;
}

// grows the context's scratch buffer; after a failed allocation the context
// holds no buffer (scratch null, scratch_bytes 0), so a later smaller request
// allocates again instead of using a freed pointer
int ensure_scratch(Ctx *c, size_t bytes) {
    if (c->scratch && bytes <= c->scratch_bytes) return MF_OK;
    if (c->scratch) {
        HIPCHK(hipStreamSynchronize(c->stream));
        hipFree(c->scratch);
        c->scratch = nullptr;
    }
    c->scratch_bytes = 0;
    double *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess || !p) {
        (void)hipGetLastError();
        return MF_ERR_HIP;
    }
    c->scratch = p;
    c->scratch_bytes = bytes;
    return MF_OK;
}

// the device table: the C-ABI table, per-species constants and the geometry
void make_tab(const mf_plugin_params *plugin, const mf_geom &g, Tab &ht) {
    std::memset(&ht, 0, sizeof(ht));
    std::memcpy(static_cast<mf_plugin_params *>(&ht), plugin, sizeof(mf_plugin_params));
    for (int q = 0; q < MF_MAX_SPECIES; ++q) {
        ht.gm1[q] = plugin->gamma[q] - 1.0;
        ht.igm1[q] = 1.0 / (plugin->gamma[q] - 1.0);
        ht.cv[q] = plugin->R_univ / (plugin->gamma[q] - 1.0);
        ht.cvm[q] = plugin->R_univ / (plugin->gamma[q] - 1.0) / plugin->molwt[q];
    }
    for (int d = 0; d < 3; ++d) {
        ht.geo.prob_lo[d] = g.prob_lo[d];
        ht.geo.prob_hi[d] = g.prob_hi[d];
        ht.geo.dx[d] = (g.prob_hi[d] - g.prob_lo[d]) / g.n_cell[d];
    }
}

int upload_tab(Ctx *c, const mf_plugin_params *plugin) {
    Tab ht;
    make_tab(plugin, c->geom, ht);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(c->d_tab, &ht, sizeof(Tab), hipMemcpyHostToDevice));
    c->tab = *plugin;
    return MF_OK;
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

const char *mf_status_string(int st) {
    switch (st) {
    case MF_OK: return "ok";
    case MF_ERR_NAN_DSDT: return "NaN detected in dsdt";
    case MF_ERR_NEG_PRESSURE: return "Pressure has gone negative";
    case MF_ERR_UNKNOWN_SCHEME: return "unknown hyperbolics scheme";
    case MF_ERR_BAD_BC: return "invalid boundary condition";
    case MF_ERR_BAD_ARG: return "invalid argument";
    case MF_ERR_HIP: return hipGetErrorString(last_hip_error);
    case MF_ERR_UNSUPPORTED: return "plugin family / species count not compiled";
    case MF_ERR_NO_DEVICE: return "no HIP device";
    case MF_ERR_TOO_LARGE: return "size above the supported maximum";
    }
    return "unknown status";
}

int mf_device_count(int *count) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    if (count) *count = n;
    return n > 0 ? MF_OK : MF_ERR_NO_DEVICE;
}

int mf_ctx_create(const mf_plugin_params *plugin, const mf_params *params, const mf_geom *geom, void **ctx_out) {
    if (!plugin || !params || !geom || !ctx_out) return MF_ERR_BAD_ARG;
    *ctx_out = nullptr;
    const KernelSet *ks = pick_kernels(*plugin);
    if (!ks) return MF_ERR_UNSUPPORTED;
    int o = params->order_hyp;
    if (!(o == 1 || o == 2 || o == 4 || o == 5)) return MF_ERR_UNKNOWN_SCHEME;
    for (int d = 0; d < 3; ++d) {
        if (geom->n_cell[d] <= 0 || !(geom->prob_hi[d] > geom->prob_lo[d])) return MF_ERR_BAD_ARG;
        if (geom->is_periodic[d]) continue;
        const int bcs[2] = {params->lo_bc[d], params->hi_bc[d]};
        for (int s = 0; s < 2; ++s)
            if (!(bcs[s] == MF_BC_FOEXTRAP || bcs[s] == MF_BC_EXT_DIR || bcs[s] == MF_BC_WALL)) return MF_ERR_BAD_BC;
    }
    if (plugin->nreact < 0 || plugin->nreact > MF_MAX_REACTIONS) return MF_ERR_BAD_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MF_ERR_NO_DEVICE;
    Ctx *c = new (std::nothrow) Ctx();
    if (!c) return MF_ERR_BAD_ARG;
    c->tab = *plugin;
    c->prm = *params;
    c->geom = *geom;
    c->ks = ks;
    c->ns = plugin->nspec;
    c->nv = FLO + c->ns;
    for (int d = 0; d < 3; ++d) c->dx[d] = (geom->prob_hi[d] - geom->prob_lo[d]) / geom->n_cell[d];
    hipGetDevice(&c->device);
    Tab ht;
    make_tab(plugin, *geom, ht);
    if (hipMalloc(&c->d_tab, sizeof(Tab)) != hipSuccess || hipMalloc(&c->d_err, sizeof(int)) != hipSuccess ||
        hipMemcpy(c->d_tab, &ht, sizeof(Tab), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(c->d_err, 0, sizeof(int)) != hipSuccess) {
        delete c;
        return MF_ERR_HIP;
    }
    // the no-allocation entry point's scratch on this device (a failure here
    // only disables mf_box_rhs_noalloc, which then returns MF_ERR_HIP)
    (void)resolve_noalloc(c->device);
    *ctx_out = c;
    return MF_OK;
}

int mf_plugin_defaults(int kind, mf_plugin_params *tab) {
    if (!tab) return MF_ERR_BAD_ARG;
    const CaseFamily *f = find_case_family(kind);
    if (!f) return MF_ERR_UNSUPPORTED;
    std::memset(tab, 0, sizeof(*tab));
    tab->kind = kind;
    tab->R_univ = mfp::RU;
    f->init_species(tab);
    return MF_OK;
}

int mf_plugin_name(int kind, const char **name) {
    if (!name) return MF_ERR_BAD_ARG;
    const CaseFamily *f = find_case_family(kind);
    if (!f) return MF_ERR_UNSUPPORTED;
    *name = f->name;
    return MF_OK;
}

int mf_ctx_destroy(void *ctx) {
    Ctx *c = (Ctx *)ctx;
    if (!c) return MF_ERR_BAD_ARG;
    hipStreamSynchronize(c->stream);
    if (c->d_tab) hipFree(c->d_tab);
    if (c->d_err) hipFree(c->d_err);
    if (c->scratch) hipFree(c->scratch);
    if (c->aux) hipFree(c->aux);
    delete c;
    return MF_OK;
}

int mf_ctx_set_plugin_params(void *ctx, const mf_plugin_params *plugin) {
    Ctx *c = (Ctx *)ctx;
    if (!c || !plugin) return MF_ERR_BAD_ARG;
    if (plugin->kind != c->tab.kind || plugin->nspec != c->tab.nspec || plugin->ngas != c->tab.ngas) return MF_ERR_BAD_ARG;
    if (plugin->nreact < 0 || plugin->nreact > MF_MAX_REACTIONS) return MF_ERR_BAD_ARG;
    return upload_tab(c, plugin);
}

int mf_ctx_set_stream(void *ctx, void *stream) {
    Ctx *c = (Ctx *)ctx;
    if (!c) return MF_ERR_BAD_ARG;
    c->stream = (hipStream_t)stream;
    return MF_OK;
}

int mf_ctx_nvars(void *ctx, int *nv) {
    Ctx *c = (Ctx *)ctx;
    if (!c || !nv) return MF_ERR_BAD_ARG;
    *nv = c->nv;
    return MF_OK;
}

int mf_ctx_sync(void *ctx) {
    Ctx *c = (Ctx *)ctx;
    if (!c) return MF_ERR_BAD_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    int e = 0;
    HIPCHK(hipMemcpy(&e, c->d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (e) HIPCHK(hipMemset(c->d_err, 0, sizeof(int)));
    return e;
}

// ----- per-box plain forms ---------------------------------------------------
int mf_box_cutcells(void *ctx, mf_box bx, mf_array4 sborder, double time) {
    Ctx *c = (Ctx *)ctx;
    if (!c || !sborder.p || sborder.ncomp < c->nv) return MF_ERR_BAD_ARG;
    (void)time;
    const int ex = bx.hi[0] - bx.lo[0] + 1, ey = bx.hi[1] - bx.lo[1] + 1, ez = bx.hi[2] - bx.lo[2] + 1;
    if (ex <= 0 || ey <= 0 || ez <= 0) return MF_ERR_BAD_ARG;
    for (int d = 0; d < 3; ++d)
        if (bx.lo[d] - MF_NGROW < sborder.begin[d] || bx.hi[d] + MF_NGROW >= sborder.end[d]) return MF_ERR_BAD_ARG;
    const int64_t ncell = (int64_t)ex * ey * ez;
    int st = ensure_scratch(c, (size_t)ncell * (c->nv * sizeof(double) + sizeof(int)));
    if (st) return st;
    CutArgs a;
    std::memset(&a, 0, sizeof(a));
    a.tab = c->d_tab;
    a.s = from_array4(sborder);
    for (int d = 0; d < 3; ++d) { a.dx[d] = c->dx[d]; a.lo[d] = bx.lo[d]; a.hi[d] = bx.hi[d]; }
    a.spec_in_solid = c->prm.species_in_solid;
    a.conj_ht = c->prm.conj_ht;
    a.tmp = c->scratch;
    a.flag = (int *)(c->scratch + (size_t)ncell * c->nv);
    c->ks->cut_box(a, c->stream);
    hipLaunchKernelGGL(k_scatter_cut, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, c->stream, a.s, a.tmp, a.flag,
                       bx.lo[0], bx.lo[1], bx.lo[2], ex, ey, ez, c->nv);
    HIPCHK(hipGetLastError());
    return MF_OK;
}

// mf_box_rhs with the face fluxes either in flux3 or, with flux3 null, in
// fluxbuf (cap doubles; null: the context's scratch buffer, allocated here)
int box_rhs_impl(Ctx *c, mf_box bx, mf_array4 sborder, mf_array4 dsdt, const mf_array4 *flux3, double time,
                 double *fluxbuf, size_t cap) {
    if (!c || !sborder.p || !dsdt.p || sborder.ncomp < c->nv || dsdt.ncomp < c->nv) return MF_ERR_BAD_ARG;
    int len[3];
    for (int d = 0; d < 3; ++d) {
        len[d] = bx.hi[d] - bx.lo[d] + 1;
        if (len[d] <= 0) return MF_ERR_BAD_ARG;
        if (bx.lo[d] - MF_NGROW < sborder.begin[d] || bx.hi[d] + MF_NGROW >= sborder.end[d]) return MF_ERR_BAD_ARG;
        if (bx.lo[d] < dsdt.begin[d] || bx.hi[d] >= dsdt.end[d]) return MF_ERR_BAD_ARG;
    }
    Work w;
    std::memset(&w, 0, sizeof(w));
    w.boxes = nullptr;
    w.cls = kCellClass;
    for (int d = 0; d < 3; ++d) {
        w.tsh[d] = c->tsh[d];
        w.one.lo[d] = bx.lo[d];
        w.one.hi[d] = bx.hi[d];
        w.one.ntile[kCellClass][d] = div_up(len[d], w.tsh[d]);
    }
    w.t0 = 0;
    w.t1 = box_tiles(w.one, kCellClass);
    const int NC = NCV + c->ns;
    int full = 0;
    if (flux3) {
        full = 1;
        for (int d = 0; d < 3; ++d) {
            if (!covers(flux3[d], bx, c->nv, d)) return MF_ERR_BAD_ARG;
            w.one.flux[d] = from_array4(flux3[d]);
        }
    } else {
        size_t tot = 0;
        for (int d = 0; d < 3; ++d) {
            int fl[3] = {len[0], len[1], len[2]};
            fl[d] += 1;
            tot += (size_t)fl[0] * fl[1] * fl[2] * NC;
        }
        double *base = fluxbuf;
        if (base) {
            if (tot > cap) return MF_ERR_TOO_LARGE;
        } else {
            int st = ensure_scratch(c, tot * sizeof(double));
            if (st) return st;
            base = c->scratch;
        }
        size_t ofs = 0;
        for (int d = 0; d < 3; ++d) {
            int fl[3] = {len[0], len[1], len[2]};
            fl[d] += 1;
            View &v = w.one.flux[d];
            v.p = base + ofs;
            v.b0 = bx.lo[0]; v.b1 = bx.lo[1]; v.b2 = bx.lo[2];
            v.js = fl[0];
            v.ks = (int64_t)fl[0] * fl[1];
            v.ns = v.ks * fl[2];
            ofs += (size_t)v.ns * NC;
        }
    }
    StageArgs a;
    std::memset(&a, 0, sizeof(a));
    a.tab = c->d_tab;
    a.src = from_array4(sborder);
    a.cutid = nullptr;
    for (int d = 0; d < 3; ++d) a.dx[d] = c->dx[d];
    a.ts = nullptr;
    a.time_host = time;
    a.stage = 1;
    a.order = c->prm.order_hyp;
    a.dissfactor = c->prm.dissfactor;
    a.nsflag = c->prm.do_ns;
    a.spec_in_solid = c->prm.species_in_solid;
    a.conj_ht = c->prm.conj_ht;
    a.flux_full = full;
    a.err = c->d_err;
    // WENO5 with the flow solve: the LDS-staged sweeps over face tiles of the
    // box (the level path's kernels); the other schemes: one face per thread
    const bool lds = a.order == 5 && a.nsflag;
    for (int d = 0; d < 3; ++d) {
        if (!lds) {
            c->ks->faces(a, w, d, w.t1, 256, c->stream);
            continue;
        }
        Work wf = w;
        wf.cls = d;
        for (int q = 0; q < 3; ++q) {
            wf.tsh[q] = c->ftsh[d][q];
            wf.one.ntile[d][q] = div_up(len[q], wf.tsh[q]);
        }
        wf.t0 = 0;
        wf.t1 = box_tiles(wf.one, d);
        c->ks->faces_lds(a, wf, d, std::min(wf.t1, 2048), face_block(c->ftsh[d]), c->ks->lds_bytes(c->ftsh[d], d),
                         c->stream);
    }
    ResArgs r;
    std::memset(&r, 0, sizeof(r));
    r.st = a;
    r.combine = 0;
    r.dsdt = from_array4(dsdt);
    c->ks->residual(r, w, w.t1, 256, c->stream);
    HIPCHK(hipGetLastError());
    return MF_OK;
}

int mf_box_rhs(void *ctx, mf_box bx, mf_array4 sborder, mf_array4 dsdt, const mf_array4 *flux3, double time) {
    return box_rhs_impl((Ctx *)ctx, bx, sborder, dsdt, flux3, time, nullptr, 0);
}

int mf_box_rhs_noalloc(void *ctx, mf_box bx, mf_array4 sborder, mf_array4 dsdt, double time) {
    Ctx *c = (Ctx *)ctx;
    if (!c) return MF_ERR_BAD_ARG;
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev != c->device) return MF_ERR_BAD_ARG;  // the context's device must be current
    if (dev < 0 || dev >= kMaxDevices || noalloc_state[dev] != 1) return MF_ERR_HIP;
    return box_rhs_impl(c, bx, sborder, dsdt, nullptr, time, noalloc_ptr[dev], kNoAllocDoubles);
}

int mf_box_rk3_combine(void *ctx, int stage, mf_box bx, mf_array4 s_new, mf_array4 src, mf_array4 dsdt, double dt) {
    Ctx *c = (Ctx *)ctx;
    if (!c || stage < 1 || stage > 3 || !s_new.p || !src.p || !dsdt.p) return MF_ERR_BAD_ARG;
    const int ex = bx.hi[0] - bx.lo[0] + 1, ey = bx.hi[1] - bx.lo[1] + 1, ez = bx.hi[2] - bx.lo[2] + 1;
    if (ex <= 0 || ey <= 0 || ez <= 0) return MF_ERR_BAD_ARG;
    if (!covers(s_new, bx, c->nv) || !covers(src, bx, c->nv) || !covers(dsdt, bx, c->nv)) return MF_ERR_BAD_ARG;
    const int64_t tot = (int64_t)ex * ey * ez * c->nv;
    int grid = (int)std::min<int64_t>((tot + 255) / 256, 65535);
    hipLaunchKernelGGL(k_combine, dim3(grid), dim3(256), 0, c->stream, stage, from_array4(s_new), from_array4(src),
                       from_array4(dsdt), bx.lo[0], bx.lo[1], bx.lo[2], ex, ey, ez, c->nv, dt);
    HIPCHK(hipGetLastError());
    return MF_OK;
}

int mf_box_post(void *ctx, mf_box bx, mf_array4 state, double dt, int flags) {
    Ctx *c = (Ctx *)ctx;
    if (!c || (flags & ~(MF_POST_CHEMISTRY | MF_POST_PRIMITIVES | MF_POST_CLIP))) return MF_ERR_BAD_ARG;
    for (int d = 0; d < 3; ++d)
        if (bx.hi[d] < bx.lo[d]) return MF_ERR_BAD_ARG;
    if (!covers(state, bx, c->nv)) return MF_ERR_BAD_ARG;
    if (!flags) return MF_OK;
    BoxPostArgs a;
    a.tab = c->d_tab;
    a.s = from_array4(state);
    for (int d = 0; d < 3; ++d) { a.lo[d] = bx.lo[d]; a.e[d] = bx.hi[d] - bx.lo[d] + 1; }
    a.dt = dt;
    a.chem = (flags & MF_POST_CHEMISTRY) ? 1 : 0;
    a.prims = (flags & MF_POST_PRIMITIVES) ? 1 : 0;
    a.clip = (flags & MF_POST_CLIP) ? 1 : 0;
    a.conj_ht = c->prm.conj_ht;
    a.do_ns = c->prm.do_ns;
    a.err = c->d_err;
    c->ks->box_post(a, c->stream);
    HIPCHK(hipGetLastError());
    return MF_OK;
}

// ----- level forms -------------------------------------------------------------
int mf_level_create(void *ctx, const mf_box *boxes, int nboxes, void **level_out) {
    Ctx *c = (Ctx *)ctx;
    if (!c || !level_out || nboxes < 0 || (nboxes > 0 && !boxes)) return MF_ERR_BAD_ARG;
    *level_out = nullptr;
    Level *L = nullptr;
    int st = level_alloc(c, boxes, nboxes, false, &L);
    if (st) return st;
    *level_out = L;
    return MF_OK;
}

}  // extern "C"

namespace {
// a level over the context's domain: boxes that tile it (partial: disjoint
// boxes inside it, a level above 0 of a hierarchy)
int level_alloc(Ctx *c, const mf_box *boxes, int nboxes, bool partial, Level **out) {
    *out = nullptr;
    Level *L = new (std::nothrow) Level();
    if (!L) return MF_ERR_BAD_ARG;
    L->c = c;
    L->partial = partial;
    for (int d = 0; d < 3; ++d) L->n[d] = c->geom.n_cell[d];
    const int64_t ncell = (int64_t)L->n[0] * L->n[1] * L->n[2];
    if (nboxes == 0) {
        mf_box b;
        for (int d = 0; d < 3; ++d) { b.lo[d] = 0; b.hi[d] = L->n[d] - 1; }
        L->boxes.push_back(b);
    } else {
        int64_t cov = 0;
        for (int b = 0; b < nboxes; ++b) {
            for (int d = 0; d < 3; ++d)
                if (boxes[b].lo[d] < 0 || boxes[b].hi[d] >= L->n[d] || boxes[b].hi[d] < boxes[b].lo[d]) { delete L; return MF_ERR_BAD_ARG; }
            cov += (int64_t)(boxes[b].hi[0] - boxes[b].lo[0] + 1) * (boxes[b].hi[1] - boxes[b].lo[1] + 1) *
                   (boxes[b].hi[2] - boxes[b].lo[2] + 1);
            L->boxes.push_back(boxes[b]);
        }
        if (!partial && cov != ncell) { delete L; return MF_ERR_BAD_ARG; }  // boxes must tile the domain
        if (nboxes > 1 || partial) {
            // disjoint (with the volume check: they cover the domain); overlapping
            // boxes would have tiles writing the same cells and leave a gap unupdated
            std::vector<unsigned char> mark((size_t)ncell, 0);
            for (int b = 0; b < nboxes; ++b)
                for (int k = boxes[b].lo[2]; k <= boxes[b].hi[2]; ++k)
                    for (int j = boxes[b].lo[1]; j <= boxes[b].hi[1]; ++j) {
                        unsigned char *row = mark.data() + (size_t)L->n[0] * (j + (size_t)L->n[1] * k);
                        for (int i = boxes[b].lo[0]; i <= boxes[b].hi[0]; ++i) {
                            if (row[i]) { delete L; return MF_ERR_BAD_ARG; }
                            row[i] = 1;
                        }
                    }
        }
    }
    const int NC = NCV + c->ns;
    L->padded = (size_t)(L->n[0] + 6) * (L->n[1] + 6) * (L->n[2] + 6) * c->nv;
    size_t fl = 0;
    for (const mf_box &B : L->boxes) {
        size_t e[3];
        for (int d = 0; d < 3; ++d) e[d] = B.hi[d] - B.lo[d] + 1;
        fl += ((e[0] + 1) * e[1] * e[2] + e[0] * (e[1] + 1) * e[2] + e[0] * e[1] * (e[2] + 1)) * NC;
    }
    L->flux_elems = fl;
    bool ok = true;
    for (int q = 0; q < 3 && ok; ++q) ok = hipMalloc(&L->arr[q], L->padded * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&L->flux, fl * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&L->cutid, sizeof(int) * ncell) == hipSuccess;
    ok = ok && hipMalloc(&L->ts, sizeof(TimeState)) == hipSuccess;
    if (!ok) {
        level_free(L);
        return MF_ERR_HIP;
    }
    for (int q = 0; q < 3; ++q) {
        hipMemset(L->arr[q], 0, L->padded * sizeof(double));
        L->view[q] = padded_view(L->arr[q], L->n, c->nv);
    }
    hipMemset(L->cutid, 0xff, sizeof(int) * ncell);
    TimeState h;
    std::memset(&h, 0, sizeof(h));
    h.t = 0.0;
    h.dt = 0.0;
    h.dt_prev = 1.e100;
    h.stop_time = c->prm.stop_time;
    h.active = 1;
    hipMemcpy(L->ts, &h, sizeof(h), hipMemcpyHostToDevice);
    int st = rebuild_tiles(L);
    if (st) { level_free(L); return st; }
    *out = L;
    return MF_OK;
}
}  // namespace

extern "C" {

int mf_level_destroy(void *level) {
    Level *L = (Level *)level;
    if (!L || L->hier) return MF_ERR_BAD_ARG;  // a level of a hierarchy goes with the hierarchy
    return level_free(L);
}

}  // extern "C"

namespace {
int level_free(Level *L) {
    if (L->c) hipStreamSynchronize(L->c->stream);
    if (L->gexec) hipGraphExecDestroy(L->gexec);
    if (L->graph) hipGraphDestroy(L->graph);
    for (int q = 0; q < 3; ++q) if (L->arr[q]) hipFree(L->arr[q]);
    if (L->flux) hipFree(L->flux);
    if (L->cutid) hipFree(L->cutid);
    if (L->cutlist) hipFree(L->cutlist);
    if (L->cutbuf) hipFree(L->cutbuf);
    if (L->cutgeom) hipFree(L->cutgeom);
    if (L->cutmask) hipFree(L->cutmask);
    if (L->geomsub) hipFree(L->geomsub);
    if (L->partials) hipFree(L->partials);
    if (L->d_box) hipFree(L->d_box);
    if (L->ts) hipFree(L->ts);
    if (L->lag) hipFree(L->lag);
    adj_free(L->adj);
    delete L;
    return MF_OK;
}
}  // namespace

extern "C" {

int mf_level_set_state(void *level, const double *state, int is_device) {
    Level *L = (Level *)level;
    if (!L || !state) return MF_ERR_BAD_ARG;
    Ctx *c = L->c;
    const int64_t ncell = (int64_t)L->n[0] * L->n[1] * L->n[2];
    const size_t bytes = (size_t)ncell * c->nv * sizeof(double);
    const double *dsrc = state;
    double *tmp = nullptr;
    if (!is_device) {
        HIPCHK(hipMalloc(&tmp, bytes));
        HIPCHK(hipMemcpy(tmp, state, bytes, hipMemcpyHostToDevice));
        dsrc = tmp;
    }
    hipLaunchKernelGGL(k_pack, dim3(4096), dim3(256), 0, c->stream, dsrc, L->view[0], L->n[0], L->n[1], L->n[2], c->nv, 1);
    std::vector<double> vf(ncell);
    if (!is_device) std::memcpy(vf.data(), state + (size_t)VFRAC * ncell, ncell * sizeof(double));
    else HIPCHK(hipMemcpy(vf.data(), state + (size_t)VFRAC * ncell, ncell * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (tmp) hipFree(tmp);
    int st = build_cutlist(L, vf.data());
    if (st) return st;
    L->need_est = true;
    L->lag_mode = 1;
    return MF_OK;
}

int mf_level_get_state(void *level, double *state, int is_device) {
    Level *L = (Level *)level;
    if (!L || !state) return MF_ERR_BAD_ARG;
    Ctx *c = L->c;
    const int64_t ncell = (int64_t)L->n[0] * L->n[1] * L->n[2];
    const size_t bytes = (size_t)ncell * c->nv * sizeof(double);
    double *dst = state;
    double *tmp = nullptr;
    if (!is_device) {
        HIPCHK(hipMalloc(&tmp, bytes));
        dst = tmp;
    }
    hipLaunchKernelGGL(k_pack, dim3(4096), dim3(256), 0, c->stream, dst, L->view[0], L->n[0], L->n[1], L->n[2], c->nv, 0);
    HIPCHK(hipStreamSynchronize(c->stream));
    if (tmp) {
        HIPCHK(hipMemcpy(state, tmp, bytes, hipMemcpyDeviceToHost));
        hipFree(tmp);
    }
    return MF_OK;
}

int mf_level_state_view(void *level, mf_array4 *v) {
    Level *L = (Level *)level;
    if (!L || !v) return MF_ERR_BAD_ARG;
    View w = L->view[0];
    v->p = w.p;
    v->jstride = w.js;
    v->kstride = w.ks;
    v->nstride = w.ns;
    for (int d = 0; d < 3; ++d) {
        v->begin[d] = -MF_NGROW;
        v->end[d] = L->n[d] + MF_NGROW;
    }
    v->ncomp = L->c->nv;
    v->pad_ = 0;
    return MF_OK;
}

int mf_fill_ghosts_single_level(void *level, double time) {
    Level *L = (Level *)level;
    if (!L) return MF_ERR_BAD_ARG;
    if (L->partial) return MF_ERR_UNSUPPORTED;  // filled from its coarse level
    Ctx *c = L->c;
    FillArgs f = fill_args(L, L->arr[0], 1);
    f.ts = nullptr;
    f.time_host = time;
    for (int ph = 0; ph <= 3; ++ph) c->ks->fill(f, ph, c->stream);
    HIPCHK(hipGetLastError());
    return MF_OK;
}

int mf_est_dt(void *level, double *dt_out) {
    Level *L = (Level *)level;
    if (!L || !dt_out) return MF_ERR_BAD_ARG;
    Ctx *c = L->c;
    enqueue_est(L);
    HIPCHK(hipGetLastError());
    TimeState h;
    int st = read_ts(L, &h);
    if (st) return st;
    L->need_est = false;
    double est = DBL_MAX;
    if (h.max_ws > 0)
        for (int d = 0; d < 3; ++d) est = std::min(est, c->dx[d] / h.max_ws);
    if (h.max_D > 0)
        for (int d = 0; d < 3; ++d) est = std::min(est, 0.5 * c->dx[d] * c->dx[d] / h.max_D / 3);
    *dt_out = est * c->prm.cfl;
    return MF_OK;
}

// nsteps steps of the level (one graph replay per step in graph mode)
int level_steps(Level *L, int nsteps, int mode) {
    Ctx *c = L->c;
    if (L->need_est) {
        enqueue_est(L);
        L->need_est = false;
    }
    if (mode == MF_MODE_GRAPH) {
        if (L->gexec && L->graph_only_flow != L->only_flow) {
            HIPCHK(hipStreamSynchronize(c->stream));
            hipGraphExecDestroy(L->gexec);
            hipGraphDestroy(L->graph);
            L->gexec = nullptr;
            L->graph = nullptr;
        }
        if (!L->gexec) {
            L->graph_only_flow = L->only_flow;
            hipStream_t cap = c->stream;
            bool own = false;
            if (!cap) {
                HIPCHK(hipStreamCreate(&cap));
                own = true;
            }
            HIPCHK(hipStreamSynchronize(c->stream));
            hipStream_t saved = c->stream;
            c->stream = cap;
            HIPCHK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
            enqueue_step(L, MF_MODE_BATCHED);
            hipError_t e = hipStreamEndCapture(cap, &L->graph);
            c->stream = saved;
            if (own) hipStreamDestroy(cap);
            if (e != hipSuccess) { last_hip_error = e; return MF_ERR_HIP; }
            HIPCHK(hipGraphInstantiate(&L->gexec, L->graph, nullptr, nullptr, 0));
        }
        for (int s = 0; s < nsteps; ++s) HIPCHK(hipGraphLaunch(L->gexec, c->stream));
    } else {
        for (int s = 0; s < nsteps; ++s) enqueue_step(L, mode);
    }
    HIPCHK(hipGetLastError());
    return MF_OK;
}

int mf_level_step(void *level, int nsteps, int mode) {
    Level *L = (Level *)level;
    if (!L || nsteps < 0 || mode < 0 || mode > 2) return MF_ERR_BAD_ARG;
    if (L->hier) return MF_ERR_UNSUPPORTED;  // stepped by its hierarchy
    int st = level_steps(L, nsteps, mode);
    if (st) return st;
    if (nsteps > 0) L->lag_mode = 0;
    TimeState h;
    st = read_ts(L, &h);
    if (st) return st;
    return h.err;
}

// ---------------------------------------------------------------------------
// split-chemistry mode (Source/mflo_splitchem.cpp; EvolveAMR with only_flow)
// ---------------------------------------------------------------------------
int lag_alloc(Level *L) {
    if (!L->lag) HIPCHK(hipMalloc(&L->lag, L->padded * sizeof(double)));
    return MF_OK;
}

int level_evolve(Level *L, double final_time, int max_step, int only_flow, int mode, int *taken) {
    Ctx *c = L->c;
    int n = 0, st = MF_OK;
    TimeState h;
    st = read_ts(L, &h);
    if (st) return st;
    if (h.err) return h.err;
    if (only_flow && (st = lag_alloc(L))) return st;
    L->only_flow = only_flow;
    while (h.istep < max_step && h.t < final_time && h.active) {
        // the reference's phi_old after a flow-only step is the state the step started from
        if (only_flow) {
            const int64_t np = (int64_t)L->padded;
            hipLaunchKernelGGL(k_copy, dim3((unsigned)std::min<int64_t>((np + 255) / 256, 65536)), dim3(256), 0, c->stream, L->lag,
                               L->arr[0], np);
        }
        st = level_steps(L, 1, mode);
        if (st) break;
        ++n;
        st = read_ts(L, &h);
        if (st) break;
        if (h.err) { st = h.err; break; }
    }
    L->only_flow = 0;
    if (n > 0) L->lag_mode = only_flow ? 2 : 0;
    if (taken) *taken = n;
    return st;
}

int level_chem_phase(Level *L, double react_time, double dt_react, double dt_rk) {
    (void)react_time;
    Ctx *c = L->c;
    if (c->prm.using_bg_inertgas && !c->ks->has_bg_gas) return MF_ERR_UNSUPPORTED;
    if (!(dt_rk > 0.0) || !(dt_react < 1e300)) return MF_ERR_BAD_ARG;
    if (L->lag_mode == 0) return MF_ERR_UNSUPPORTED;
    int st = lag_alloc(L);
    if (st) return st;
    if (L->lag_mode == 1) {
        const int64_t np = (int64_t)L->padded;
        hipLaunchKernelGGL(k_copy, dim3((unsigned)std::min<int64_t>((np + 255) / 256, 65536)), dim3(256), 0, c->stream, L->lag,
                           L->arr[0], np);
    }
    // the substeps of `while (tchem < dt_react) { ...; tchem += dt_rk; }`
    int64_t nsub = 0;
    for (double tchem = 0.0; tchem < dt_react; tchem += dt_rk) {
        if (++nsub > (int64_t)1 << 40) return MF_ERR_BAD_ARG;
    }
    ChemSplitArgs a;
    a.tab = c->d_tab;
    a.s = L->view[0];
    a.lag = padded_view(L->lag, L->n, c->nv);
    for (int d = 0; d < 3; ++d) a.n[d] = L->n[d];
    a.dt = dt_rk;
    a.do_ns = c->prm.do_ns;
    a.err = c->d_err;
    a.bg = c->prm.using_bg_inertgas ? 1 : 0;
    // launches of at most kChunk substeps (the kernel continues the sequence exactly)
    const int64_t kChunk = 1024;
    int64_t done = 0;
    do {
        const int64_t m = std::min(kChunk, nsub - done);
        a.nsub = (int)m;
        a.store_bg = done == 0;
        done += m;
        a.finish = done == nsub;
        c->ks->chem_split(a, c->stream);
    } while (done < nsub);
    HIPCHK(hipGetLastError());
    L->lag_mode = 2;
    L->need_est = true;  // the next ComputeDt sees the state after the chemistry
    TimeState h;
    st = read_ts(L, &h);
    if (st) return st;
    return h.err;
}

int mf_level_step_timed(void *level, int nsteps, int mode, double *class_ms, int *class_launches) {
    Level *L = (Level *)level;
    if (!L || nsteps < 0 || mode < 0 || mode > 1 || !class_ms) return MF_ERR_BAD_ARG;
    if (L->hier) return MF_ERR_UNSUPPORTED;  // stepped by its hierarchy
    Ctx *c = L->c;
    if (L->need_est) {
        enqueue_est(L);
        L->need_est = false;
    }
    L->timer.on = true;
    L->timer.ev.clear();
    for (int s = 0; s < nsteps; ++s) enqueue_step(L, mode);
    L->timer.on = false;
    HIPCHK(hipGetLastError());
    TimeState h;
    int st = read_ts(L, &h);
    for (int q = 0; q < KC_N; ++q) {
        class_ms[q] = 0.0;
        if (class_launches) class_launches[q] = 0;
    }
    for (auto &e : L->timer.ev) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, e.second.first, e.second.second);
        class_ms[e.first] += ms;
        if (class_launches) class_launches[e.first] += 1;
        hipEventDestroy(e.second.first);
        hipEventDestroy(e.second.second);
    }
    L->timer.ev.clear();
    if (st) return st;
    return h.err;
}

int mf_level_evolve(void *level, double final_time, int max_step, int only_flow, int mode, int *nsteps_out) {
    Level *L = (Level *)level;
    if (!L || mode < 0 || mode > 2 || max_step < 0) return MF_ERR_BAD_ARG;
    if (L->hier) return MF_ERR_UNSUPPORTED;  // stepped by its hierarchy
    return level_evolve(L, final_time, max_step, only_flow ? 1 : 0, mode, nsteps_out);
}

int mf_level_advance_chemistry(void *level, double react_time, double dt_react, double dt_react_rk) {
    Level *L = (Level *)level;
    if (!L) return MF_ERR_BAD_ARG;
    if (L->hier) return MF_ERR_UNSUPPORTED;  // stepped by its hierarchy
    return level_chem_phase(L, react_time, dt_react, dt_react_rk);
}

int mf_level_evolve_split(void *level, double t_ss, double t_react, double dt_react, double dt_react_rk, double dt_ss,
                          int max_step, int mode) {
    Level *L = (Level *)level;
    if (!L || mode < 0 || mode > 2 || max_step < 0) return MF_ERR_BAD_ARG;
    if (L->hier) return MF_ERR_UNSUPPORTED;  // stepped by its hierarchy
    if (L->c->prm.using_bg_inertgas && !L->c->ks->has_bg_gas) return MF_ERR_UNSUPPORTED;
    double react_time = 0.0, flow_time;
    int st = level_evolve(L, t_ss, max_step, 1, mode, nullptr);
    if (st) return st;
    flow_time = t_ss;
    while (react_time < t_react) {
        st = level_chem_phase(L, react_time, dt_react, dt_react_rk);
        if (st) return st;
        st = level_evolve(L, flow_time + dt_ss, max_step, 1, mode, nullptr);
        if (st) return st;
        react_time += dt_react;
        flow_time += dt_ss;
    }
    return MF_OK;
}

int mf_level_info(void *level, double *time, double *dt, int *istep, double *resnorms) {
    Level *L = (Level *)level;
    if (!L) return MF_ERR_BAD_ARG;
    TimeState h;
    int st = read_ts(L, &h);
    if (st) return st;
    if (time) *time = h.t;
    if (dt) *dt = h.dt;
    if (istep) *istep = h.istep;
    if (resnorms) for (int q = 0; q < L->c->nv; ++q) resnorms[q] = h.resnorm[q];
    return MF_OK;
}

int mf_level_boxes(void *level, mf_box *boxes, int cap, int *nboxes, mf_geom *geom, int *nv) {
    Level *L = (Level *)level;
    if (!L) return MF_ERR_BAD_ARG;
    const int nb = (int)L->boxes.size();
    if (nboxes) *nboxes = nb;
    if (geom) *geom = L->c->geom;
    if (nv) *nv = L->c->nv;
    if (boxes) {
        if (cap < nb) return MF_ERR_BAD_ARG;
        std::memcpy(boxes, L->boxes.data(), sizeof(mf_box) * nb);
    }
    return MF_OK;
}

int mf_level_set_time(void *level, double time, double dt_prev, int istep) {
    Level *L = (Level *)level;
    if (!L) return MF_ERR_BAD_ARG;
    TimeState h;
    int st = read_ts(L, &h);
    if (st) return st;
    h.t = time;
    h.dt_prev = dt_prev;
    h.istep = istep;
    h.active = 1;
    h.err = 0;
    HIPCHK(hipMemcpy(L->ts, &h, sizeof(h), hipMemcpyHostToDevice));
    return MF_OK;
}

int mf_level_set_tile(void *level, int cls, int tx, int ty, int tz) {
    Level *L = (Level *)level;
    if (!L || cls < 0 || cls >= kTileClasses) return MF_ERR_BAD_ARG;
    if (L->hier) return MF_ERR_UNSUPPORTED;  // stepped by its hierarchy
    if (tx < 1 || ty < 1 || tz < 1 || (int64_t)tx * ty * tz > 4096) return MF_ERR_BAD_ARG;
    int t[3] = {tx, ty, tz};
    if (cls < 3 && L->c->ks->lds_bytes(t, cls) > 160 * 1024) return MF_ERR_TOO_LARGE;
    if (cls >= kAdjFaceClass && L->c->ks->adj && L->c->ks->adj->lds_bytes(t, cls - kAdjFaceClass) > 160 * 1024)
        return MF_ERR_TOO_LARGE;
    HIPCHK(hipStreamSynchronize(L->c->stream));
    L->tsh[cls][0] = tx;
    L->tsh[cls][1] = ty;
    L->tsh[cls][2] = tz;
    return rebuild_tiles(L);
}

int mf_level_get_tile(void *level, int cls, int *tx, int *ty, int *tz) {
    Level *L = (Level *)level;
    if (!L || cls < 0 || cls >= kTileClasses) return MF_ERR_BAD_ARG;
    if (tx) *tx = L->tsh[cls][0];
    if (ty) *ty = L->tsh[cls][1];
    if (tz) *tz = L->tsh[cls][2];
    return MF_OK;
}

// Autotuner: for each tile class, times the kernels of that class (the face
// sweep of one direction, or the stage-1 residual/combine) on the current state
// for every candidate shape and keeps the fastest. The state in A is only read
// (stage 1 writes B, the fluxes and the cut-cell buffer).
int mf_level_autotune(void *level) {
    Level *L = (Level *)level;
    if (!L) return MF_ERR_BAD_ARG;
    if (L->hier) return MF_ERR_UNSUPPORTED;  // tuned through its hierarchy (mf_hier_autotune)
    return level_autotune(L);
}

}  // extern "C"

namespace {
// times every kernel class on the level's current state for candidate tile
// shapes (the last ones narrow, for the small boxes of a refined level) and
// keeps the fastest; the valid cells of the state are only read
int level_autotune(Level *L) {
    Ctx *c = L->c;
    constexpr int kFaceCand = 13, kCellCand = 11;
    static const int face_cand[3][kFaceCand][3] = {
        {{64, 4, 1}, {64, 2, 2}, {32, 16, 1}, {32, 4, 2}, {32, 8, 1}, {64, 1, 4}, {32, 8, 2}, {32, 2, 4}, {32, 4, 4}, {32, 2, 8},
         {16, 8, 2}, {16, 4, 4}, {8, 8, 4}},
        {{32, 8, 1}, {16, 16, 1}, {16, 16, 2}, {32, 4, 2}, {16, 8, 2}, {64, 8, 1}, {32, 16, 1}, {16, 12, 2}, {8, 32, 2}, {16, 32, 1},
         {16, 8, 1}, {8, 16, 2}, {8, 8, 4}},
        {{32, 1, 8}, {16, 2, 8}, {32, 2, 4}, {16, 2, 16}, {16, 1, 16}, {64, 1, 8}, {32, 1, 16}, {16, 2, 12}, {8, 2, 32}, {16, 1, 32},
         {16, 1, 8}, {8, 2, 16}, {8, 4, 8}}};
    static const int cell_cand[kCellCand][3] = {{64, 2, 2}, {32, 4, 2}, {64, 4, 1}, {128, 2, 1}, {32, 8, 1}, {16, 4, 4},
                                                {32, 2, 4}, {64, 1, 4}, {16, 8, 2}, {8, 8, 4}, {16, 16, 1}};
    HIPCHK(hipStreamSynchronize(c->stream));
    TimeState saved;
    int st = read_ts(L, &saved);
    if (st) return st;
    TimeState probe = saved;
    probe.active = 1;
    probe.err = 0;
    if (!(probe.dt > 0)) probe.dt = 1e-30;
    HIPCHK(hipMemcpy(L->ts, &probe, sizeof(probe), hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    // ghost cells and cut cells of stage 1 once, so the timed kernels see real
    // data (a level above 0 of a hierarchy: the single-level fill of its
    // padded array stands in for its two-level fill, whose next stage
    // overwrites these ghost cells; its cut geometry is made first)
    {
        FillArgs f = fill_args(L, L->arr[0], 1);
        for (int ph = 0; ph <= 3; ++ph) c->ks->fill(f, ph, c->stream);
        if (L->ncut) {
            if (L->partial) c->ks->cut_geom(cut_args(L, L->arr[0]), L->cutgeom, L->cutmask, c->stream);
            c->ks->cut(cut_args(L, L->arr[0]), c->stream);
        }
    }
    StageArgs a = stage_args(L, 1);
    ResArgs r;
    std::memset(&r, 0, sizeof(r));
    r.st = a;
    r.combine = 1;
    r.old = L->view[0];
    r.dst = L->view[1];
    r.do_ns = c->prm.do_ns;
    r.partials = L->partials;
    // the fastest of three rounds of three launches (after one untimed): a
    // single round let run-to-run noise pick between near-equal candidates
    auto time_it = [&](auto &&launch) -> double {
        launch();
        double best = 1e300;
        for (int round = 0; round < 3; ++round) {
            hipEventRecord(e0, c->stream);
            for (int q = 0; q < 3; ++q) launch();
            hipEventRecord(e1, c->stream);
            hipEventSynchronize(e1);
            float ms = 0.f;
            hipEventElapsedTime(&ms, e0, e1);
            best = std::min(best, (double)ms);
        }
        return best;
    };
    const bool lds = use_lds(L);
    for (int cls = 0; cls <= kCellClass && st == MF_OK; ++cls) {
        if (cls < 3 && !lds) continue;
        const int ncand = cls < 3 ? kFaceCand : kCellCand;
        double best = 1e300;
        int bi[3] = {L->tsh[cls][0], L->tsh[cls][1], L->tsh[cls][2]};
        for (int q = 0; q < ncand; ++q) {
            const int *cand = cls < 3 ? face_cand[cls][q] : cell_cand[q];
            bool fits = true;
            for (int d = 0; d < 3; ++d) if (cand[d] > 2 * L->n[d] && cand[d] > 1) fits = false;
            if (cls < 3 && c->ks->lds_bytes(cand, cls) > 96 * 1024) fits = false;
            if (!fits) continue;
            for (int d = 0; d < 3; ++d) L->tsh[cls][d] = cand[d];
            st = rebuild_tiles(L);
            if (st) break;
            double ms;
            if (cls < 3) ms = time_it([&] { launch_faces(L, a, cls, MF_MODE_BATCHED); });
            else {
                // the cell tiles also drive the post-step pass (chemistry,
                // primitives, norms): it is timed with the residual, reading B
                // and writing C so that the state in A stays as it was
                ResArgs rp = r;
                rp.st.src = L->view[1];
                rp.dst = L->view[2];
                rp.do_chem = c->prm.do_chemistry;
                rp.clip = c->prm.clip_species;
                rp.track_res = c->prm.track_residual_norms;
                auto post = [&] {
                    c->ks->post(rp, level_work(L, kCellClass), grid_for(L->ntiles[kCellClass], L->max_grid), 256,
                                c->stream);
                };
                if (!lds) ms = time_it([&] {
                    for (int d = 0; d < 3; ++d) launch_faces(L, a, d, MF_MODE_BATCHED);
                    launch_residual(L, r, MF_MODE_BATCHED);
                    post();
                });
                else ms = time_it([&] {
                    launch_residual(L, r, MF_MODE_BATCHED);
                    post();
                });
            }
            if (ms < best) { best = ms; bi[0] = cand[0]; bi[1] = cand[1]; bi[2] = cand[2]; }
        }
        for (int d = 0; d < 3; ++d) L->tsh[cls][d] = bi[d];
        if (st == MF_OK) st = rebuild_tiles(L);
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    HIPCHK(hipMemcpy(L->ts, &saved, sizeof(saved), hipMemcpyHostToDevice));
    int ok = 0;
    HIPCHK(hipMemcpy(c->d_err, &ok, sizeof(int), hipMemcpyHostToDevice));
    return st;
}
}  // namespace

extern "C" {

int mf_util_copy(void *ctx, double *dst, const double *src, int64_t n) {
    Ctx *c = (Ctx *)ctx;
    if (!c || (n > 0 && (!dst || !src))) return MF_ERR_BAD_ARG;
    if (n <= 0) return MF_OK;
    int grid = (int)std::min<int64_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, c->stream, dst, src, n);
    HIPCHK(hipGetLastError());
    return MF_OK;
}

// div_pos against the IEEE division over n seeded operand pairs: numerators
// of either sign with exponents in [-40, 40], positive denominators with
// exponents in [lo_exp, hi_exp]; counts[0] equal, [1] one ulp apart, [2] more
__global__ void k_div_check(int64_t n, int lo_exp, int hi_exp, unsigned long long *counts) {
    unsigned long long c0 = 0, c1 = 0, c2 = 0;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        // splitmix64 of the index
        uint64_t z = (uint64_t)q * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
        auto next = [&]() {
            z += 0x9E3779B97F4A7C15ull;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
            x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
            return x ^ (x >> 31);
        };
        const uint64_t r1 = next(), r2 = next(), r3 = next(), r4 = next();
        const double m1 = 1.0 + (double)(r1 >> 11) * 0x1.0p-53, m2 = 1.0 + (double)(r2 >> 11) * 0x1.0p-53;
        const int e1 = (int)(r3 % 81) - 40, e2 = lo_exp + (int)(r4 % (uint64_t)(hi_exp - lo_exp + 1));
        const double num = ((r3 >> 32) & 1 ? -1.0 : 1.0) * ldexp(m1, e1);
        const double den = ldexp(m2, e2);
        const double a = num / den, b = div_pos(num, den);
        const int64_t ia = __double_as_longlong(a), ib = __double_as_longlong(b);
        const int64_t d = ia > ib ? ia - ib : ib - ia;
        if (d == 0) ++c0;
        else if (d == 1) ++c1;
        else ++c2;
    }
    atomicAdd(counts + 0, c0);
    atomicAdd(counts + 1, c1);
    atomicAdd(counts + 2, c2);
}

int mf_util_div_check(void *ctx, int64_t n, int lo_exp, int hi_exp, int64_t *counts) {
    Ctx *c = (Ctx *)ctx;
    if (!c || n < 0 || !counts || lo_exp > hi_exp || lo_exp < -900 || hi_exp > 900) return MF_ERR_BAD_ARG;
    unsigned long long *d = nullptr;
    HIPCHK(hipMalloc(&d, 3 * sizeof(unsigned long long)));
    HIPCHK(hipMemsetAsync(d, 0, 3 * sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(k_div_check, dim3(2048), dim3(256), 0, c->stream, n, lo_exp, hi_exp, d);
    unsigned long long h[3];
    HIPCHK(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    hipFree(d);
    for (int q = 0; q < 3; ++q) counts[q] = (int64_t)h[q];
    return MF_OK;
}

int mf_adjoint_forward(void *level, int nsteps, int ncheck, int spec, double *J_out) {
    Level *L = (Level *)level;
    if (!L || nsteps < 1 || ncheck < 1) return MF_ERR_BAD_ARG;
    if (L->hier) return MF_ERR_UNSUPPORTED;  // stepped by its hierarchy
    Ctx *c = L->c;
    if (spec < 0 || spec >= c->ns) return MF_ERR_BAD_ARG;
    if (!c->ks->adj || !c->prm.fixed_dt || !(c->prm.dt_fixed > 0) || c->prm.do_ns != 1 || c->prm.order_hyp != 5 ||
        c->prm.conj_ht != 0 || c->prm.do_chemistry != 1)
        return MF_ERR_UNSUPPORTED;
    ncheck = std::min(ncheck, nsteps);
    int st = adj_alloc(L, ncheck);
    if (st) return st;
    AdjState *A = L->adj;
    A->nsteps = nsteps;
    A->spec = spec;
    A->ready = false;
    A->snaps.clear();
    A->pool = A->bufs;
    TimeState h;
    st = read_ts(L, &h);
    if (st) return st;
    A->t0 = h.t;
    A->dt = c->prm.dt_fixed;
    HIPCHK(hipMemset(A->J, 0, sizeof(double)));
    hipLaunchKernelGGL(k_set_ts, dim3(1), dim3(1), 0, c->stream, L->ts, A->t0, A->dt);
    adj_store(L, 0);
    // the right spine of the binomial reversal is stored on the way
    int pos = 0, s = ncheck - 1;
    while (nsteps - pos > 1 && s > 0) {
        const int m = adj_split(pos, nsteps, s);
        adj_advance(L, pos, m - pos, true);
        adj_store(L, m);
        pos = m;
        --s;
    }
    adj_advance(L, pos, nsteps - pos, true);
    HIPCHK(hipGetLastError());
    st = read_ts(L, &h);
    if (st) return st;
    if (h.err) return h.err;
    double J = 0.0;
    HIPCHK(hipMemcpy(&J, A->J, sizeof(double), hipMemcpyDeviceToHost));
    if (J_out) *J_out = J;
    A->ready = true;
    return MF_OK;
}

int mf_adjoint_backward(void *level, double *grad_logA, double *grad_invel) {
    Level *L = (Level *)level;
    if (!L || !L->adj || !L->adj->ready) return MF_ERR_BAD_ARG;
    Ctx *c = L->c;
    AdjState *A = L->adj;
    HIPCHK(hipMemset(A->grad, 0, (MF_MAX_REACTIONS + 1) * sizeof(double)));
    for (int q = 0; q < 5; ++q) HIPCHK(hipMemset(A->lam[q], 0, L->padded * sizeof(double)));
    adj_reverse(L, 0, A->nsteps, A->ncheck - 1);
    hipLaunchKernelGGL(k_set_ts, dim3(1), dim3(1), 0, c->stream, L->ts, A->t0, A->dt);
    HIPCHK(hipGetLastError());
    TimeState h;
    int st = read_ts(L, &h);
    if (st) return st;
    double g[MF_MAX_REACTIONS + 1];
    HIPCHK(hipMemcpy(g, A->grad, sizeof(g), hipMemcpyDeviceToHost));
    if (grad_logA) for (int r = 0; r < c->tab.nreact; ++r) grad_logA[r] = g[r];
    if (grad_invel) *grad_invel = g[c->tab.nreact];
    A->ready = false;
    if (h.err) return h.err;
    return MF_OK;
}

int mf_adjoint_state_grad(void *level, double *out, int is_device) {
    Level *L = (Level *)level;
    if (!L || !out || !L->adj) return MF_ERR_BAD_ARG;
    Ctx *c = L->c;
    const int64_t ncell = (int64_t)L->n[0] * L->n[1] * L->n[2];
    const size_t bytes = (size_t)ncell * c->nv * sizeof(double);
    double *dst = out, *tmp = nullptr;
    if (!is_device) {
        HIPCHK(hipMalloc(&tmp, bytes));
        dst = tmp;
    }
    hipLaunchKernelGGL(k_pack, dim3(4096), dim3(256), 0, c->stream, dst, padded_view(L->adj->lam[0], L->n, c->nv), L->n[0], L->n[1],
                       L->n[2], c->nv, 0);
    HIPCHK(hipStreamSynchronize(c->stream));
    if (tmp) {
        HIPCHK(hipMemcpy(out, tmp, bytes, hipMemcpyDeviceToHost));
        hipFree(tmp);
    }
    return MF_OK;
}

}  // extern "C"

// dt of the level's last step (the dt[lev] that a checkpoint records and the
// next ComputeDt caps against; 1e100 before the first step), for mf_io.hip
int mf_level_dt_prev(void *level, double *dt_prev) {
    Level *L = (Level *)level;
    if (!L || !dt_prev) return MF_ERR_BAD_ARG;
    TimeState h;
    int st = read_ts(L, &h);
    if (st) return st;
    *dt_prev = h.dt_prev;
    return MF_OK;
}

// stream and component count of a context, for mf_amr.hip
int mf_ctx_stream(void *ctx, hipStream_t *stream, int *nv) {
    Ctx *c = (Ctx *)ctx;
    if (!c || !stream || !nv) return MF_ERR_BAD_ARG;
    *stream = c->stream;
    *nv = c->nv;
    return MF_OK;
}

// context of a level, for mf_amr.hip
int mf_level_ctx(void *level, void **ctx) {
    Level *L = (Level *)level;
    if (!L || !ctx) return MF_ERR_BAD_ARG;
    *ctx = L->c;
    return MF_OK;
}

// geometry of a context, for mf_amr.hip
int mf_ctx_geom(void *ctx, mf_geom *geom) {
    Ctx *c = (Ctx *)ctx;
    if (!c || !geom) return MF_ERR_BAD_ARG;
    *geom = c->geom;
    return MF_OK;
}

// a device buffer of at least `bytes` owned by the context, for mf_amr.hip:
// grown when needed (after the stream's work that may use the old one),
// reused by later calls in stream order
int mf_ctx_aux(void *ctx, size_t bytes, double **p) {
    Ctx *c = (Ctx *)ctx;
    if (!c || !p) return MF_ERR_BAD_ARG;
    if (!c->aux || bytes > c->aux_bytes) {
        if (c->aux) {
            HIPCHK(hipStreamSynchronize(c->stream));
            hipFree(c->aux);
            c->aux = nullptr;
        }
        c->aux_bytes = 0;
        double *q = nullptr;
        if (hipMalloc(&q, bytes) != hipSuccess || !q) {
            (void)hipGetLastError();
            return MF_ERR_HIP;
        }
        c->aux = q;
        c->aux_bytes = bytes;
    }
    *p = c->aux;
    return MF_OK;
}

// the physical-boundary rules of the single-level fill (foextrap, wall
// reflection, the plugin's externalbc; Source/mflo_bcfill.H:16-128) on the
// cells of `region` of `view` outside the domain in non-periodic directions,
// in phases of 1, 2, 3 such directions; every other cell of the region must
// be filled already. For mf_amr.hip (FillPatchTwoLevels of a fine box).
int mf_ctx_fill_physbc(void *ctx, mf_array4 view, mf_box region, double time) {
    Ctx *c = (Ctx *)ctx;
    if (!c || !view.p || view.ncomp < c->nv) return MF_ERR_BAD_ARG;
    bool outside = false;
    for (int d = 0; d < 3; ++d) {
        if (region.lo[d] < view.begin[d] || region.hi[d] >= view.end[d] || region.hi[d] < region.lo[d])
            return MF_ERR_BAD_ARG;
        if (!c->geom.is_periodic[d] && (region.lo[d] < 0 || region.hi[d] >= c->geom.n_cell[d])) outside = true;
    }
    if (!outside) return MF_OK;
    FillArgs f;
    std::memset(&f, 0, sizeof(f));
    f.tab = c->d_tab;
    f.s = from_array4(view);
    for (int d = 0; d < 3; ++d) {
        f.n[d] = c->geom.n_cell[d];
        f.periodic[d] = c->geom.is_periodic[d];
        f.lo_bc[d] = c->prm.lo_bc[d];
        f.hi_bc[d] = c->prm.hi_bc[d];
        f.prob_lo[d] = c->geom.prob_lo[d];
        f.dx[d] = c->dx[d];
    }
    f.ts = nullptr;
    f.time_host = time;
    f.stage = 1;
    for (int ph = 1; ph <= 3; ++ph) c->ks->fill_region(f, ph, region.lo, region.hi, c->stream);
    HIPCHK(hipGetLastError());
    return MF_OK;
}

// the two-level driver (mf_hier_*)
#include "mf_hier.inc.h"
