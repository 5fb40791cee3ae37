/*
 * mesoflow_amd.h -- C-ABI of the MI355X-native RHS + RK3 path of mesoflow.
 *
 * Every entry point has C linkage, takes plain pointers and sizes, and
 * returns an int status (MF_OK == 0).  No exception and no abort crosses
 * this boundary: where the reference calls amrex::Abort the device sets an
 * error word that the host turns into one of the MF_ERR_* codes below.
 *
 * Layout of every state view (mf_array4) is the reference's amrex::Array4:
 * component-major, x fastest, FP64,
 *     p[(i-begin.x) + (j-begin.y)*jstride + (k-begin.z)*kstride + n*nstride]
 * with the component order of Source/globalDefines.H:19-32
 * (0-4 rho,rhou,rhov,rhow,rhoE; 5-10 density,velx,vely,velz,pressure,temperature;
 *  11 volume fraction; 12.. species concentrations in mol/m^3).
 *
 * A context (mf_ctx) is not thread-safe: use one per host thread or stream.
 */
#ifndef MESOFLOW_AMD_H
#define MESOFLOW_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MF_MAX_SPECIES   16
#define MF_MAX_REACTIONS 8
#define MF_NGROW         3      /* num_grow, Source/mflo_evolve.cpp:253 */

/* component indices, Source/globalDefines.H:19-32 */
enum {
    MF_RHO = 0, MF_RHOU = 1, MF_RHOV = 2, MF_RHOW = 3, MF_RHOE = 4,
    MF_DENS = 5, MF_VELX = 6, MF_VELY = 7, MF_VELZ = 8, MF_PRES = 9,
    MF_TEMP = 10, MF_VFRAC = 11, MF_FLO_NVARS = 12
};

/* status codes (replace the amrex::Abort sites of the reference) */
enum {
    MF_OK = 0,
    MF_ERR_NAN_DSDT = 1,        /* Source/Src_3d/div_3d.H:72-75            */
    MF_ERR_NEG_PRESSURE = 2,    /* Source/mflo_evolve.cpp:162-167          */
    MF_ERR_UNKNOWN_SCHEME = 3,  /* Source/Src_3d/compute_flux_3d.H:273-276 */
    MF_ERR_BAD_BC = 4,          /* Source/mflo.cpp:113,141                 */
    MF_ERR_BAD_ARG = 5,
    MF_ERR_HIP = 6,
    MF_ERR_UNSUPPORTED = 7,
    MF_ERR_NO_DEVICE = 8,
    MF_ERR_TOO_LARGE = 9
};

/* boundary codes of mflo.lo_bc / mflo.hi_bc (legend in every inputs file,
 * e.g. tests/Cartesian/ShockTube_N2/inputs_x:30-36; mapping mflo.cpp:70-143) */
enum { MF_BC_PERIODIC = 0, MF_BC_FOEXTRAP = 2, MF_BC_EXT_DIR = 3, MF_BC_WALL = 6 };

/* compile-time plugin families built into the library (the reference selects
 * its plugin headers per case through Make.package, e.g.
 * tests/Cartesian/ShockTube_N2/Make.package:1-2) */
enum {
    MF_PLUGIN_SHOCKTUBE_N2 = 1,      /* tests/Cartesian/ShockTube_N2 headers, NUM_SPECIES=1 */
    MF_PLUGIN_CATALYST_CYLINDER = 2, /* models/catalystCylinder headers, NUM_SPECIES=4  */
    MF_PLUGIN_POROUS_ARRHENIUS = 3   /* mixture thermo of models/catalystCylinder/thermo.H
                                        with Arrhenius mass-action kinetics, NS = 5 or 10 */
};

/* amrex::Array4<Real> (a view; the caller owns the memory). begin is the
 * inclusive low corner (ghost cells included), end the exclusive high one. */
typedef struct mf_array4 {
    double *p;
    int64_t jstride, kstride, nstride;
    int32_t begin[3];
    int32_t end[3];
    int32_t ncomp;
    int32_t pad_;
} mf_array4;

/* amrex::Box (cell-centred, inclusive bounds) */
typedef struct mf_box {
    int32_t lo[3];
    int32_t hi[3];
} mf_box;

/* geometry.* keys: prob_lo, prob_hi, amr.n_cell, is_periodic */
typedef struct mf_geom {
    double prob_lo[3];
    double prob_hi[3];
    int32_t n_cell[3];
    int32_t is_periodic[3];
} mf_geom;

/* mflo.* runtime parameters (Source/mflo.cpp:263-307, defaults mflo.H:226-259) */
typedef struct mf_params {
    int32_t order_hyp;          /* 1, 2 (default), 4, 5                         */
    int32_t do_ns;              /* solve_navier_stokes                          */
    int32_t species_in_solid;
    int32_t conj_ht;            /* conjugate_heat_transfer                       */
    int32_t using_bg_inertgas;
    int32_t clip_species;
    int32_t track_residual_norms;
    int32_t do_chemistry;       /* 1: Advance_chemistry_implicit after the flow
                                   RK3 (EvolveAMR default); 0: only_flow         */
    double dissfactor;
    double cfl;
    double stop_time;           /* stop_time clamp of ComputeDt                  */
    int32_t lo_bc[3];
    int32_t hi_bc[3];
    int32_t fixed_dt;           /* 1: use dt_fixed instead of the CFL estimate   */
    int32_t pad_;
    double dt_fixed;
} mf_params;

/* Runtime parameter tables of the compile-time plugins. In the reference
 * these live in managed globals (e.g. tests/Cartesian/ShockTube_N2/species.H:18-19,
 * userfuncs.cpp:3-14); here they are passed by value into the kernels. */
typedef struct mf_plugin_params {
    int32_t kind;                /* MF_PLUGIN_*                                 */
    int32_t nspec;               /* NUM_SPECIES                                 */
    int32_t ngas;                /* NUM_GAS_SPECIES: species 0..ngas-1 are gas  */
    int32_t nreact;              /* NUM_REACTIONS                               */
    double R_univ;               /* gas constant used by the thermo plugin      */
    double molwt[MF_MAX_SPECIES];        /* kg/mol (species.cpp)               */
    double gamma[MF_MAX_SPECIES];        /* gas species ratio of heats         */
    double advect_flag[MF_MAX_SPECIES];  /* mflo_species::advect_flags         */
    double visc, thcond;                 /* compute_fluid_transport            */
    double dcoef_gas[MF_MAX_SPECIES];    /* compute_spec_dcoeff, vfrac == 1    */
    double dcoef_solid[MF_MAX_SPECIES];  /* compute_spec_dcoeff, vfrac == 0    */
    double stoich[MF_MAX_REACTIONS][MF_MAX_SPECIES]; /* rxnarray              */
    double log_A[MF_MAX_REACTIONS];      /* ln of the pre-exponential          */
    double Ea[MF_MAX_REACTIONS];         /* activation energy, J/mol           */
    double kconst;                       /* constant rate of catalystCylinder  */
    /* problem state read by externalbc / initdata */
    double left_p, left_rho, left_T;     /* ShockTube p_l, rho_l, T_l          */
    double right_p, right_rho, right_T;  /* ShockTube p_r, rho_r, T_r          */
    double left_spec[MF_MAX_SPECIES];
    double right_spec[MF_MAX_SPECIES];
    double in_vel, in_temp, in_rho, in_pres;  /* inflow state (x-lo ext_dir)   */
    double in_spec[MF_MAX_SPECIES];
    double out_pres;                     /* catalystCylinder outflow p0        */
    double siteconc;
    double user[8];                      /* parameters of a case written against
                                            the plugin API (its problem.* inputs) */
} mf_plugin_params;

/* ------------------------------------------------------------------------ */
/* Context                                                                    */
/* ------------------------------------------------------------------------ */
const char *mf_status_string(int status);
int mf_device_count(int *count);

/* Creates a context bound to the current HIP device. Checks the plugin/table
 * combination (MF_ERR_UNSUPPORTED), the scheme (MF_ERR_UNKNOWN_SCHEME) and the
 * boundary codes (MF_ERR_BAD_BC). */
int mf_ctx_create(const mf_plugin_params *plugin, const mf_params *params,
                  const mf_geom *geom, void **ctx_out);
int mf_ctx_destroy(void *ctx);
/* Plugin families compiled from case headers written against the plugin API
 * (mesoflow_amd/csrc/mf_plugin_api.h; kinds >= 100): the table a context of
 * `kind` starts from -- zeroed, kind, R_univ = RU, then the case's species
 * tables (its species.cpp init(): nspec, ngas, molwt, advect_flag, ...);
 * MF_ERR_UNSUPPORTED for a kind that is not such a family. The reference has
 * these tables as compile-time globals (e.g.
 * tests/Cartesian/SpeciesEqTests/species.cpp:10-24). Host only. */
int mf_plugin_defaults(int kind, mf_plugin_params *tab);
/* The case (namespace) name of such a family, a static string. */
int mf_plugin_name(int kind, const char **name);
/* Replaces the plugin parameter table of a context (same kind, nspec, ngas;
 * MF_ERR_BAD_ARG otherwise), e.g. new rate constants or inflow state. Work
 * enqueued afterwards uses the new table. No counterpart in the reference
 * (its tables are compile-time globals); used by parameter sweeps and the
 * finite-difference checks of the adjoint. */
int mf_ctx_set_plugin_params(void *ctx, const mf_plugin_params *plugin);
/* All later work of the context is enqueued on this stream (NULL = default). */
int mf_ctx_set_stream(void *ctx, void *hip_stream);
int mf_ctx_nvars(void *ctx, int *nv);          /* TOTAL_NVARS = 12 + NUM_SPECIES */
/* Waits for the context's stream and returns (and clears) the device error word. */
int mf_ctx_sync(void *ctx);

/* ------------------------------------------------------------------------ */
/* Plain per-box forms: the body of the MFIter loop of                        */
/* mflo::compute_dsdt_flow (Source/mflo_evolve.cpp:579-652) and its neighbours */
/* ------------------------------------------------------------------------ */

/* update_cutcells over the valid box bx, in place on sborder
 * (Source/mflo_evolve.cpp:509-517 -> Source/Src_3d/ib_utils.H:271-436). */
int mf_box_cutcells(void *ctx, mf_box bx, mf_array4 sborder, double time);

/* Per-box right-hand side: transport, species diffusion coefficients, fluid
 * source, the three face-flux sweeps and update_residual
 * (Source/mflo_evolve.cpp:602-651; Source/Src_3d/compute_flux_3d.H:86-430,
 * Source/Src_3d/div_3d.H:13-76). sborder must hold bx grown by MF_NGROW
 * (already ghost-filled and cut-cell updated). dsdt covers bx. flux may be
 * NULL or point to three face-centred views (growHi(bx,d,1)). */
int mf_box_rhs(void *ctx, mf_box bx, mf_array4 sborder, mf_array4 dsdt,
               const mf_array4 *flux3, double time);

/* mf_box_rhs without flux views that allocates nothing. No counterpart in the
 * reference: for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation). The face
 * fluxes go to a scratch array of 8M doubles (64 MiB) in the library's device
 * memory, one per GPU, held until the process ends; a box needs
 * 3 (n_x+1) n_y n_z (5 + NUM_SPECIES) doubles (a 64^3 box with up to 5
 * species, a 32^3 box with any), a larger one returns MF_ERR_TOO_LARGE. Each
 * call writes every scratch value it reads, so nothing carries over between
 * calls; but every call on a GPU shares that GPU's one array, so calls that
 * overlap on one GPU (other streams or host threads) race. The context's
 * device must be the current one (MF_ERR_BAD_ARG otherwise). Same results as
 * mf_box_rhs(ctx, bx, sborder, dsdt, NULL, time). */
int mf_box_rhs_noalloc(void *ctx, mf_box bx, mf_array4 sborder, mf_array4 dsdt, double time);

/* The MultiFab arithmetic of one SSP-RK3 stage on the valid box
 * (Source/mflo_evolve.cpp:283, 296-298, 309-311), same operation order:
 *   stage 1: s_new = 1*src + dt*dsdt                    (src = Sborder)
 *   stage 2: s_new = ((s_new + dt*dsdt) + 3*src) * 1/4  (src = S_old)
 *   stage 3: s_new = (src + 2*(s_new + dt*dsdt)) * third (src = S_old) */
int mf_box_rk3_combine(void *ctx, int stage, mf_box bx, mf_array4 s_new,
                       mf_array4 src, mf_array4 dsdt, double dt);

/* The per-cell passes after a level's RK3 flow step, on one box in place
 * (flags, applied in this order): MF_POST_CHEMISTRY Advance_chemistry_implicit
 * over dt (Source/mflo_evolve.cpp:470-493, 684-720: SSPRK3 on the species
 * source), MF_POST_PRIMITIVES update_primitive_vars (:114-171; a negative
 * pressure with do_ns is reported as MF_ERR_NEG_PRESSURE by mf_ctx_sync),
 * MF_POST_CLIP clip_neg_speciesconc (:521-540). state must hold bx. */
enum { MF_POST_CHEMISTRY = 1, MF_POST_PRIMITIVES = 2, MF_POST_CLIP = 4 };
int mf_box_post(void *ctx, mf_box bx, mf_array4 state, double dt, int flags);

/* ------------------------------------------------------------------------ */
/* Two-level coupling: the AMReX operations with which mesoflow couples a     */
/* fine level (refinement ratio r) to its coarse level, as per-box device     */
/* operations over all nv components of the context's state.                 */
/* ------------------------------------------------------------------------ */

/* amrex::average_down (AverageDownTo, Source/mflo_gridwork.cpp:125-130) over
 * the coarse box cbx: crse = (sum of the r^3 fine cells) * 1/r^3. fine must
 * hold refine(cbx). */
int mf_box_average_down(void *ctx, mf_box cbx, mf_array4 fine, mf_array4 crse, int ratio);

/* The data FillPatch reads at a time (GetData, Source/mflo.cpp:310-330, with
 * FillPatchSingleLevel's time interpolation): within 1e-3 (t_new - t_old) of
 * t_new the new state, of t_old the old state, else
 * ((t_new-time)/(t_new-t_old)) s_old + ((time-t_old)/(t_new-t_old)) s_new. */
int mf_box_time_interp(void *ctx, mf_box bx, mf_array4 dst, mf_array4 s_old, mf_array4 s_new, double t_old,
                       double t_new, double time);

/* cell_cons_interp, the mapper of FillPatchTwoLevels and FillCoarsePatch
 * (Source/mflo_gridwork.cpp:151-165, 170-193), over the fine box fbx:
 * MC-limited slopes of each coarse cell, scaled so that no fine value leaves
 * the range of the 27 coarse neighbours; conservative (the mean of the r^3
 * children is the coarse value). crse must hold coarsen(fbx) grown by 1. */
int mf_box_interp_cons(void *ctx, mf_box fbx, mf_array4 crse, mf_array4 fine, int ratio);

/* FillPatchTwoLevels (Source/mflo_gridwork.cpp:147-165) for the fine box fbx:
 * sborder over fbx grown by MF_NGROW gets, in this order, the coarse data at
 * `time` (mf_box_time_interp of crse_old/crse_new, which must hold the coarse
 * cells under the region grown by one: ghost-filled coarse arrays) mapped by
 * mf_box_interp_cons, then the fine data of every fine box (fboxes/fdata, the
 * fine level's valid data at `time`, fbx's own included) and of its periodic
 * images where they cover the region, then, where the region leaves the
 * domain in a non-periodic direction, the fine level's physical boundary
 * conditions (the rules of mf_fill_ghosts_single_level: foextrap, wall
 * reflection, the plugin's externalbc at `time`). ctx is the fine level's
 * context (its geometry gives the domain, periodicity and boundary codes);
 * fbx must lie inside the domain. The coarse data goes to a device buffer the
 * context keeps and grows; the call is asynchronous on the context's stream. */
int mf_fill_patch_two_levels(void *ctx, mf_box fbx, mf_array4 sborder, const mf_box *fboxes, const mf_array4 *fdata,
                             int nfine, mf_array4 crse_old, mf_array4 crse_new, double t_old, double t_new, double time,
                             int ratio);

/* FluxRegister on one face box (coarse face indices; along dir lo == hi is the
 * face index) of the boundary of a coarsened fine box; reg is a face-centred
 * view of it.
 * CrseInit(flux, dir, .., scale, ADD) (Source/mflo_evolve.cpp:655-667):
 *   reg += scale * cflux on the face box. */
int mf_fluxreg_crse_add(void *ctx, int dir, mf_box face, mf_array4 reg, mf_array4 cflux, double scale);
/* FineAdd(flux, dir, .., scale) (:668-679): reg += scale * f for each of the
 * r^2 fine faces on every coarse face. */
int mf_fluxreg_fine_add(void *ctx, int dir, mf_box face, mf_array4 reg, mf_array4 fflux, int ratio, double scale);
/* Reflux(S, scale, ..) (:243-246): the coarse cell outside the fine region
 * next to each face (below it for side 0, the fine box's low side; above it
 * for side 1) gets S += (-scale for side 0, +scale for side 1) * reg / vol. */
int mf_fluxreg_reflux(void *ctx, int dir, int side, mf_box face, mf_array4 state, mf_array4 reg, double scale,
                      double vol);

/* ErrorEst (Source/mflo.cpp:183-260) on one box: state holds bx grown by one
 * (ghost-filled); tags is device memory, one byte per cell of bx (x fastest).
 * For each of the ntag (<= 16) variables, state_based_refinement and then
 * stategrad_based_refinement (Source/mflo_tagging.H:6-70) set the byte to
 * tagval where |state(comps[c])| > refine_phi[c] or the largest one-sided
 * absolute difference in x, y, z exceeds refine_phigrad[c]. Tags are only
 * set, never cleared. The parameter arrays are host memory. */
int mf_box_tag(void *ctx, mf_box bx, mf_array4 state, unsigned char *tags, const double *refine_phi,
               const double *refine_phigrad, const int *comps, int ntag, int tagval);

/* ------------------------------------------------------------------------ */
/* Level forms: one AMR level of boxes stored as one domain-sized array with  */
/* MF_NGROW ghost layers. Boxes must tile the domain (level 0).               */
/* ------------------------------------------------------------------------ */
int mf_level_create(void *ctx, const mf_box *boxes, int nboxes, void **level_out);
int mf_level_destroy(void *level);
/* State of the valid domain, [nv][nz][ny][nx] doubles, host or device memory. */
int mf_level_set_state(void *level, const double *state, int is_device);
int mf_level_get_state(void *level, double *state, int is_device);
/* Device pointer of the current state (padded by MF_NGROW; strides returned). */
int mf_level_state_view(void *level, mf_array4 *view);

/* FillPatchSingleLevel of the current state at `time` (Source/mflo_gridwork.cpp:135-146,
 * Source/mflo_bcfill.H:16-128): periodic images, foextrap, wall reflection and
 * the plugin's externalbc, written into the ghost layers of the level array. */
int mf_fill_ghosts_single_level(void *level, double time);

/* EstTimeStep + ComputeDt (Source/mflo_timestep.cpp:18-153) on the current state. */
int mf_est_dt(void *level, double *dt_out);

/* EvolveAMR single-level step loop (Source/mflo_evolve.cpp:18-89): nsteps
 * coarse steps, each ComputeDt -> Advance_coupled_strang (flow RK3 + explicit
 * chemistry) -> update_primitive_vars -> clip_neg_speciesconc ->
 * compute_residual_norms.  mode selects how the same kernels are launched:
 *   MF_MODE_PER_BOX   one launch per box and kernel (the reference's shape),
 *   MF_MODE_BATCHED   one persistent launch per kernel over the box/tile list,
 *   MF_MODE_GRAPH     MF_MODE_BATCHED with each step replayed from a hipGraph. */
enum { MF_MODE_PER_BOX = 0, MF_MODE_BATCHED = 1, MF_MODE_GRAPH = 2 };
int mf_level_step(void *level, int nsteps, int mode);

/* mf_level_step (modes MF_MODE_PER_BOX / MF_MODE_BATCHED) with a HIP event pair
 * around every launch group on the context stream. class_ms[7] receives the
 * summed event time of: 0 ComputeDt, 1 ghost fill, 2 cut cells, 3 face-flux
 * sweeps, 4 update_residual + RK3 combine, 5 step close, 6 post-step pass
 * (chemistry, primitives, clip, residual and dt partials); class_launches[7]
 * (may be NULL) the number of groups. Instrumentation of this
 * build; no counterpart in the reference. */
int mf_level_step_timed(void *level, int nsteps, int mode, double *class_ms, int *class_launches);

/* EvolveAMR(final_time, only_flow) (Source/mflo_evolve.cpp:18-89): steps while
 * istep < max_step and t < final_time (ComputeDt per step, the stop_time
 * break); only_flow = 1 leaves out the chemistry of each step
 * (Advance_coupled_strang :314-317). *nsteps_out (may be NULL) receives the
 * number of steps taken. */
int mf_level_evolve(void *level, double final_time, int max_step, int only_flow, int mode, int *nsteps_out);

/* Chemistry phase of the split mode (Source/mflo_splitchem.cpp:40-62):
 * Advance_chemistry (:172-195, explicit midpoint) substeps of dt_react_rk from
 * react_time while tchem < dt_react, then update_vars_after_chemsolve
 * (:74-140). The first stage of a substep reads the reference's other state
 * buffer (phi_old after its swap): the state before the previous flow step or
 * substep; after mf_level_set_state, the state itself. With
 * using_bg_inertgas = 1 the phase starts with store_inertgas_conc (:142-170:
 * the background gas concentration into the density slot of every fluid cell)
 * and update_vars_after_chemsolve forms the density with it (:94-97); only a
 * family whose thermo has a background gas accepts that flag. Returns
 * MF_ERR_UNSUPPORTED after coupled steps (mf_level_step, the adjoint), whose
 * pre-chemistry state is not kept, and for using_bg_inertgas = 1 with a
 * family without a background gas. All substeps of a cell run in one thread. */
int mf_level_advance_chemistry(void *level, double react_time, double dt_react, double dt_react_rk);

/* Evolve_split (Source/mflo_splitchem.cpp:24-72; main.cpp:36-56 reads its
 * five times from mflo.steady_flow_time, react_final_time,
 * react_increment_time, react_time_step, flow_coupling_time): flow-only
 * steps to t_ss, then until react_time >= t_react a chemistry phase of
 * dt_react followed by flow-only steps to the next coupling time (+dt_ss). */
int mf_level_evolve_split(void *level, double t_ss, double t_react, double dt_react, double dt_react_rk, double dt_ss,
                          int max_step, int mode);

/* Time, last dt, step count and residual norms (TOTAL_NVARS doubles, may be NULL). */
int mf_level_info(void *level, double *time, double *dt, int *istep, double *resnorms);
/* Overrides the level time/dt history (restart). */
int mf_level_set_time(void *level, double time, double dt_prev, int istep);

/* ErrorEst over a whole level: ghost cells filled at time (FillPatch), tags
 * (host memory, one byte per domain cell, x fastest) cleared and set to 1
 * where a criterion of mf_box_tag holds; *ntagged (may be NULL) counts them. */
int mf_level_tag(void *level, double time, const double *refine_phi, const double *refine_phigrad, const int *comps,
                 int ntag, unsigned char *tags, int64_t *ntagged);

/* The boxes of the next finer level from tags of a level (the regrid of
 * Source/mflo_evolve.cpp:177-209 -> MakeNewGrids with amr.n_error_buf,
 * blocking_factor, max_grid_size, grid_eff and the refinement ratio). tags:
 * host bytes over the domain n_cell (x fastest); every tag grows by
 * n_error_buf (wrapping around the domain where is_periodic[d] is set;
 * is_periodic may be NULL: no periodic direction), tags are coarsened by blocking_factor/ratio, clustered
 * (Berger-Rigoutsos: accept at tags/volume >= grid_eff, else cut at a hole,
 * an inflection of the signature, or in half), refined to the fine level and
 * chopped to max_grid_size. Boxes are in fine-level indices, disjoint,
 * aligned to blocking_factor, and cover every buffered tag. With boxes NULL
 * and cap 0 only *nboxes is returned. Host code: needs no device. */
int mf_tags_to_boxes(const unsigned char *tags, const int32_t n_cell[3], const int32_t *is_periodic, int n_error_buf,
                     int blocking_factor, int max_grid_size, double grid_eff, int ratio, mf_box *boxes, int cap,
                     int *nboxes);

/* Box list, geometry and component count of a level (boxes may be NULL to
 * query the count; cap is the capacity of boxes). */
int mf_level_boxes(void *level, mf_box *boxes, int cap, int *nboxes, mf_geom *geom, int *nv);

/* ------------------------------------------------------------------------ */
/* Two-level hierarchy: mflo's AmrCore levels 0 and 1                        */
/* ------------------------------------------------------------------------ */
/* A hierarchy over a level 0 (created by mf_level_create, which keeps its
 * state and time): level 1 refines it by ratio (2..4), with subcycling
 * (nsubsteps = ratio, Source/mflo.cpp:56-61). While the hierarchy exists the
 * level-0 handle is stepped only through it (its single-level step and
 * evolve calls return MF_ERR_UNSUPPORTED) and mf_level_destroy refuses it;
 * mf_hier_destroy releases level 1 and returns level 0 to its caller. */
int mf_hier_create(void *level0, int ratio, void **hier_out);
int mf_hier_destroy(void *hier);
/* MakeNewLevelFromScratch (Source/mflo_gridwork.cpp:91-123) with the
 * caller's initial data: level 1 over the boxes (fine indices, inside the
 * refined domain, aligned to the ratio, disjoint), state dense over the whole
 * fine domain [nv][nz*r][ny*r][nx*r] (only the boxes' cells are used), host or
 * device memory. *level_out (may be NULL) receives the level's handle: its
 * state, boxes, plotfile and checkpoint calls work as for level 0, its state
 * view holds the valid cells of its boxes. */
int mf_hier_add_level(void *hier, const mf_box *boxes, int nboxes, const double *state, int is_device,
                      void **level_out);
/* Handle and count of the hierarchy's levels. */
int mf_hier_level(void *hier, int lev, void **level_out);
int mf_hier_nlevels(void *hier, int *nlev);
/* mf_level_autotune for each level of the hierarchy on its own boxes (the
 * fine level's small boxes take narrower tiles). */
int mf_hier_autotune(void *hier);
/* mflo.do_reflux (default 1): Reflux of level 0 after the fine substeps. */
int mf_hier_set_reflux(void *hier, int do_reflux);
/* Regrid every regrid_int coarse steps (0: never; Source/mflo_evolve.cpp:
 * 177-209): at the start of a step with istep > the last regrid and istep %
 * regrid_int == 0, after ComputeDt, the tags of level 0 (mf_level_tag with
 * the ErrorEst criteria refine_phi / refine_phigrad of the components comps)
 * give the new level-1 boxes (mf_tags_to_boxes with amr.n_error_buf,
 * blocking_factor (a multiple of the ratio), max_grid_size, grid_eff); level
 * 1 is then remade (RemakeLevel: its data where the old boxes covered, the
 * coarse level interpolated elsewhere), made (MakeNewLevelFromCoarse, dt[1] =
 * dt[0] / ratio) or cleared (no tags). */
int mf_hier_set_regrid(void *hier, int regrid_int, const double *refine_phi, const double *refine_phigrad,
                       const int *comps, int ntag, int n_error_buf, int blocking_factor, int max_grid_size,
                       double grid_eff);
/* The same regrid now, with the caller's boxes for level 1 (none: clear it). */
int mf_hier_regrid(void *hier, const mf_box *boxes, int nboxes);
/* EvolveAMR's step over the hierarchy (Source/mflo_evolve.cpp:20-82, no I/O),
 * nsteps coarse steps: ComputeDt over the levels (Source/mflo_timestep.cpp:
 * 18-50), timeStep's recursion (:175-249): Advance_coupled_strang of level 0
 * (with CrseInit of its fluxes at each stage), ratio substeps of level 1
 * (FillPatchTwoLevels at each stage, FineAdd, its chemistry at every substep:
 * the recursion passes no only_flow), Reflux and AverageDownTo; then
 * update_primitive_vars and clip on both levels and the residual norms of
 * level 0. mode as mf_level_step (per box, one persistent launch per level,
 * substep and stage, or the whole coarse step replayed from a hipGraph). */
int mf_hier_step(void *hier, int nsteps, int mode);
/* EvolveAMR(final_time, only_flow) over the hierarchy (only_flow applies to
 * level 0, as in the reference). */
int mf_hier_evolve(void *hier, double final_time, int max_step, int only_flow, int mode, int *nsteps_out);
/* Time, dt[0], coarse step count and residual norms (level 0). */
int mf_hier_info(void *hier, double *time, double *dt, int *istep, double *resnorms);
/* Output of the hierarchy (Source/mflo_plot.cpp): WritePlotFile with
 * amrex::WriteMultiLevelPlotfile (:45-56; Header with finest_level, the
 * refinement ratio, each level's domain, steps and cell sizes, and
 * Level_0/Cell, Level_1/Cell), WriteCheckpointFile (:57-145; finest_level,
 * istep, dt and t_new of both levels, each level's BoxArray and
 * Level_<l>/<varnames[0]>) and ReadCheckpointFile (:147-239) into a hierarchy
 * over a level 0 with the same boxes (level 1 is made or remade with the
 * file's boxes, or cleared). varnames as mf_level_write_plotfile. */
int mf_hier_write_plotfile(void *hier, const char *dir, const char *const *varnames);
int mf_hier_write_checkpoint(void *hier, const char *dir, const char *const *varnames);
int mf_hier_read_checkpoint(void *hier, const char *dir, const char *const *varnames);

/* Output of a level (Source/mflo_plot.cpp). varnames: nv names
 * (Source/mflo.cpp:34-53 order), or NULL for rho, rhou, rhov, rhow, rhoe,
 * density, velx, vely, velz, pressure, temperature, volfrac, S0, S1, ..
 * WritePlotFile (:45-56, amrex::WriteMultiLevelPlotfile): <dir>/Header
 * (HyperCLaw-V1.1) and <dir>/Level_0/Cell_H, Cell_D_00000, one FAB per box
 * of the valid cells, time and step from the level. */
int mf_level_write_plotfile(void *level, const char *dir, const char *const *varnames);
/* WriteCheckpointFile (:57-145): <dir>/Header ("Checkpoint file for mflo",
 * finest_level 0, istep, dt of the last step, t_new at precision 17, the
 * BoxArray) and the state as <dir>/Level_0/<varnames[0]>_H, _D_00000. */
int mf_level_write_checkpoint(void *level, const char *dir, const char *const *varnames);
/* ReadCheckpointFile (:147-239) into a level created with the same boxes
 * (MF_ERR_BAD_ARG if the Header's BoxArray or the FABs differ,
 * MF_ERR_UNSUPPORTED for finest_level > 0: mf_hier_read_checkpoint reads
 * those): the state (mf_level_set_state) and the time, dt and step
 * (mf_level_set_time). */
int mf_level_read_checkpoint(void *level, const char *dir, const char *const *varnames);

/* The same files from host data (no device work): state is
 * [nv][nz][ny][nx] over geom's domain, only the cells of the boxes are
 * written or read; boxes must lie inside the domain (MF_ERR_BAD_ARG). For a
 * host that holds the state itself, e.g. after mf_level_get_state. */
int mf_write_plotfile(const char *dir, const mf_geom *geom, const mf_box *boxes, int nboxes, int nv,
                      const double *state, double time, int istep, const char *const *varnames);
int mf_write_checkpoint(const char *dir, const mf_geom *geom, const mf_box *boxes, int nboxes, int nv,
                        const double *state, double time, double dt, int istep, const char *const *varnames);
int mf_read_checkpoint(const char *dir, const mf_geom *geom, const mf_box *boxes, int nboxes, int nv, double *state,
                       double *time, double *dt, int *istep, const char *const *varnames);

/* Tile shape of a kernel class: 0, 1, 2 the LDS-staged WENO5 face sweeps of
 * direction x, y, z; 3 the cell tiles (residual + RK3 combine, dt estimate,
 * face sweeps of the other schemes). */
int mf_level_set_tile(void *level, int cls, int tx, int ty, int tz);
int mf_level_get_tile(void *level, int cls, int *tx, int *ty, int *tz);
/* Autotuner: times every kernel class on the current state for a set of
 * candidate tile shapes and keeps the fastest (the state is only read). */
int mf_level_autotune(void *level);

/* ------------------------------------------------------------------------ */
/* Discrete adjoint (no counterpart in the reference)                         */
/* ------------------------------------------------------------------------ */
/* Objective J = sum over the nsteps steps of dt * dy * dz * sum_{j,k} of
 * molwt[spec] * c_spec * (rho u) / rho in the outlet cell layer i = nx-1 of
 * the state after the step (the species mass flux leaving through x-hi).
 * mf_adjoint_forward runs nsteps steps of dt_fixed from the current state and
 * time, returns J and keeps binomial checkpoints (ncheck state snapshots,
 * 1 <= ncheck; ncheck >= nsteps stores every step). mf_adjoint_backward then
 * reverses the run with the discrete adjoint kernels (recomputing forward
 * segments from the checkpoints) and returns dJ/d(log_A[r]) for r < nreact and
 * dJ/d(in_vel); afterwards the level holds the initial state of the run again.
 * Supported: the Arrhenius porous-reactor families with fixed_dt = 1,
 * do_ns = 1, order_hyp = 5, conj_ht = 0, do_chemistry = 1
 * (MF_ERR_UNSUPPORTED otherwise). */
int mf_adjoint_forward(void *level, int nsteps, int ncheck, int spec, double *J_out);
int mf_adjoint_backward(void *level, double *grad_logA, double *grad_invel);
/* dJ/d(initial state) of the last backward sweep, [nv][nz][ny][nx]. */
int mf_adjoint_state_grad(void *level, double *out, int is_device);

/* Plain HBM copy of n doubles on the context stream (bandwidth calibration). */
int mf_util_copy(void *ctx, double *dst, const double *src, int64_t n);
/* Check of this build's positive-denominator division (no counterpart in the
 * reference): n seeded operand pairs, numerators of either sign with binary
 * exponents in [-40, 40], denominators in [2^lo_exp, 2^(hi_exp+1)); counts[3]
 * receives how many quotients equal the IEEE division, differ by one ulp, or
 * by more. */
int mf_util_div_check(void *ctx, int64_t n, int lo_exp, int hi_exp, int64_t *counts);

#ifdef __cplusplus
}
#endif
#endif /* MESOFLOW_AMD_H */
