// mf_args.h -- argument blocks shared by the kernels and the host driver.
#pragma once
#include "mf_device.h"

namespace mfd {

constexpr int kMaxNV = FLO + MF_MAX_SPECIES;

// Device-resident step state: everything a step needs from the previous one,
// so that a whole step can be captured in a hipGraph and replayed.
struct TimeState {
    double t;          // t_new of the level (cur_time of EvolveAMR)
    double dt;         // dt of the step in flight
    double dt_prev;    // dt of the previous step (ComputeDt growth cap); 1e100 at start
    double max_ws;     // norm0 of the wave speed (EstTimeStep)
    double max_D;      // norm0 of the largest species diffusivity
    double stop_time;
    int istep;
    int active;        // 0 once EvolveAMR's stop conditions are met
    int err;           // first MF_ERR_* raised on the device
    int have_est;      // max_ws/max_D are current
    double resnorm[kMaxNV];
};

// Tile classes: 0..2 the LDS-staged face sweeps of direction 0..2, 3 cells
// (residual/combine, dt estimate and the face sweeps without LDS staging),
// 4..6 the adjoint face sweeps of direction 0..2.
constexpr int kTileClasses = 7;
constexpr int kCellClass = 3;
constexpr int kAdjFaceClass = 4;

// One box of a level (or the single box of a per-box call).
struct BoxDesc {
    int lo[3], hi[3];
    int ntile[kTileClasses][3];
    int tile_start[kTileClasses];
    View flux[3];      // face-centred flux of each direction; begin = lo
};

// Tiles [t0, t1) of class cls of the box/tile list (boxes == nullptr: the single box `one`).
struct Work {
    const BoxDesc *boxes;
    int nboxes;
    int t0, t1;
    int cls;
    int tsh[3];
    BoxDesc one;
};

struct StageArgs {
    const Tab *tab;
    View src;                  // stage source (padded level array or caller's Sborder)
    const int *cutid;          // substitution of own-box cut cells (null: none)
    const double *cutbuf;
    int ncut;
    int nx, ny;
    double dx[3];
    const TimeState *ts;       // null: use time_host / dt_host
    double time_host, dt_host;
    int stage;                 // 1..3: RHS time is t (stage 1) or t+dt (stages 2, 3)
    int order;
    double dissfactor;
    int nsflag, spec_in_solid, conj_ht;
    int flux_full;             // flux views hold all NV comps (else the 5+NS flux comps)
    int *err;
};

struct ResArgs {
    StageArgs st;
    int combine;               // 0: write dsdt; 1: RK3 stage combine into dst
    View dsdt;                 // combine == 0 (single box)
    View old;                  // U^n (S_old)
    View dst;                  // output state of the stage
    int do_chem, clip, track_res, do_ns;
    double *partials;          // [tile][nv + 2]
};

struct FillArgs {
    const Tab *tab;
    View s;                    // padded level array
    int n[3];
    int periodic[3];
    int lo_bc[3], hi_bc[3];
    double prob_lo[3], dx[3];
    const TimeState *ts;       // null: use time_host
    double time_host;
    int stage;
};

struct CutArgs {
    const Tab *tab;
    View s;
    const int *list;           // [3][ncut] i,j,k of the cut cells
    double *cutbuf;            // [nv][ncut]
    int ncut;
    double dx[3];
    int spec_in_solid, conj_ht;
    const TimeState *ts;
    // static geometry of the listed cut cells (vfrac does not change within a
    // run): geom[6][ncut] = the wall point xnd[0..2], the signed distance and
    // the two IDW weight sums (velocity denominator, scalar sum),
    // mask[4][ncut] = bit (n+2)*25 + (m+2)*5 + (l+2) set for a fluid
    // (vfrac == 1) neighbour (l, m, n) of the 5^3 neighbourhood
    const double *geom;
    const unsigned *mask;
    // dense per-box form (list == nullptr): cells of the box [lo, hi], result
    // to tmp[nv][cells] and flag[cells]
    int lo[3], hi[3];
    double *tmp;
    int *flag;
};

// ---------------------------------------------------------------------------
// discrete adjoint
// ---------------------------------------------------------------------------
struct AdjFaceArgs {
    StageArgs st;              // forward stage: input array, its cut values, dx, flags
    View lam_out;              // adjoint of the stage output (valid cells)
    double coef;               // d(stage output)/d(dsdt): dt, dt/4 or 2 dt third
    View lam;                  // adjoint of the stage input array (padded, accumulated)
    double *lcut;              // adjoint of the stage's cut-cell values [nv][ncut] (accumulated)
};
struct AdjCutArgs {
    CutArgs c;                 // forward stage input array and cut list
    View lam;                  // adjoint of the input array (accumulated)
    const double *lcut;        // adjoint of the cut-cell values
    double *coef;              // [ncut][record] terms for the fluid neighbours
    // fluid cells that receive cut-cell terms: nbr[3][nnbr] (i, j, k),
    // nbrmask[4][nnbr] bit (n+2)*25 + (m+2)*5 + (l+2) set when the cell at
    // offset -(l, m, n) is a listed cut cell
    const int *nbr;
    const unsigned *nbrmask;
    int nnbr;
    const int *cutid;          // cut index of valid cell i + nx (j + ny k), or -1
    int n[3];                  // valid cells of the level
};
struct AdjFillArgs {
    FillArgs f;                // forward array (its interior values feed externalbc)
    View lam;                  // adjoint of the array: ghosts are folded into their sources and cleared
    double *grad;              // grad[nreact] += d/d(in_vel)
};
struct AdjPostArgs {
    StageArgs st;              // src = S3 (post-flow state); ts gives dt
    View lu;                   // adjoint of U^{n+1}
    View ld;                   // out: adjoint of S3
    int clip;
    int spec;                  // objective species
    double dA;                 // outlet face area of a cell
    int outlet_i;              // i of the outlet cell layer
    double *grad;              // grad[r] += d/d(log_A[r])
};
struct ObjArgs {
    const Tab *tab;
    View s;                    // post-step state
    const TimeState *ts;
    int spec, outlet_i, ny, nz;
    double dA;
    double *J;                 // J += dt * outlet flux
};
struct AdjSet {
    void (*faces)(const AdjFaceArgs &, const Work &, int dir, int grid, int block, size_t lds, hipStream_t);
    size_t (*lds_bytes)(const int *tsh, int dir);
    void (*cut)(const AdjCutArgs &, hipStream_t);
    void (*fill)(const AdjFillArgs &, int phase, hipStream_t);
    void (*post)(const AdjPostArgs &, const Work &, int grid, int block, hipStream_t);
    void (*objective)(const ObjArgs &, hipStream_t);
    int cut_rec;               // doubles per cut-cell record of AdjCutArgs::coef
};

// chemistry phase of the split mode over the valid cells of a level
struct ChemSplitArgs {
    const Tab *tab;
    View s;                    // current state (phi_new), updated in place
    View lag;                  // the other buffer (phi_old): first-stage state of the next substep
    int n[3];                  // valid cells
    int nsub;                  // substeps of this launch
    double dt;                 // react_time_step
    int finish;                // 1: update_vars_after_chemsolve after the substeps
    int bg;                    // mflo.using_bg_inertgas (a family with a background gas)
    int store_bg;              // 1: first launch of a phase: store_inertgas_conc first
    int do_ns;
    int *err;
};

// mf_box_post: chemistry, primitive update and clip on one box, in place
struct BoxPostArgs {
    const Tab *tab;
    View s;
    int lo[3], e[3];
    double dt;
    int chem, prims, clip, conj_ht, do_ns;
    int *err;
};

// Host launchers of one compiled plugin family.
struct KernelSet {
    int ns;
    bool has_bg_gas;           // the family's thermo has a background gas (using_bg_inertgas)
    void (*fill)(const FillArgs &, int phase, hipStream_t);
    // physical-boundary phase of the fill over the region [rlo, rhi] of a.s
    void (*fill_region)(const FillArgs &, int phase, const int *rlo, const int *rhi, hipStream_t);
    // physical-boundary phase of the fill over a list of cells ([3][n])
    void (*fill_list)(const FillArgs &, int phase, const int *list, int n, hipStream_t);
    // chemistry alone over the tile list (st.src -> dst)
    void (*chem)(const ResArgs &, const Work &, int grid, int block, hipStream_t);
    void (*cut)(const CutArgs &, hipStream_t);
    void (*cut_geom)(const CutArgs &, double *geom, unsigned *mask, hipStream_t);
    // the same for the listed cut cells only (indices into the cut list)
    void (*cut_geom_sub)(const CutArgs &, double *geom, unsigned *mask, const int *sub, int nsub, hipStream_t);
    void (*cut_box)(const CutArgs &, hipStream_t);
    void (*faces)(const StageArgs &, const Work &, int dir, int grid, int block, hipStream_t);
    // LDS-staged WENO5 sweeps (order_hyp = 5); lds_bytes(tile) gives the shared memory of a tile
    void (*faces_lds)(const StageArgs &, const Work &, int dir, int grid, int block, size_t lds, hipStream_t);
    size_t (*lds_bytes)(const int *tsh, int dir);
    void (*residual)(const ResArgs &, const Work &, int grid, int block, hipStream_t);
    void (*est)(const ResArgs &, const Work &, int grid, int block, hipStream_t);
    void (*post)(const ResArgs &, const Work &, int grid, int block, hipStream_t);
    void (*chem_split)(const ChemSplitArgs &, hipStream_t);
    void (*box_post)(const BoxPostArgs &, hipStream_t);
    const AdjSet *adj;         // discrete adjoint (null: not compiled for this family)
};

// A plugin family compiled from a case header written against
// mf_plugin_api.h (mf_kernels_case.hip), registered at library load.
struct CaseFamily {
    int id;                                  // mf_plugin_params::kind
    int nspec;
    bool has_bg_gas;
    const KernelSet *ks;
    void (*init_species)(mf_plugin_params *);
    const char *name;
};
int register_case_family(const CaseFamily &f);   // mf_capi.hip
const CaseFamily *find_case_family(int id);

extern const KernelSet ks_shocktube_n2;
extern const KernelSet ks_catalyst_cylinder;
extern const KernelSet ks_porous5;
extern const KernelSet ks_porous10;

}  // namespace mfd
