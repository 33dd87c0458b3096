/*
 * letkf_gpu.h -- C ABI of the MI355X local-analysis library.
 *
 * One handle holds one analysis on one GPU: the background ensemble of the
 * local gridpoints, the observations, the localizer and the solver
 * parameters.  letkf_gpu_run() performs the work of the reference's
 * letkf_solver_run() (src/letkf_solver.F90:150-389): local-observation
 * search, Gaspari-Cohn localization, the LETKF ensemble transform
 * (src/letkf_core.F90:69-146), the application of the transform to every
 * state slab and the post-processing (recentering, increment clipping,
 * RTPS / RTPP / multiplicative inflation, member bounds, spread).
 *
 * Conventions
 *  - Every function returns 0 (LETKF_OK) or a LETKF_E* status;
 *    letkf_gpu_last_error() returns the message of the last failure.
 *  - All arrays are Fortran column-major with the same index order as the
 *    reference's module arrays; sizes are int64_t.
 *  - The caller owns every host buffer.  The set_* calls copy host data to
 *    device memory owned by the handle; the get_* calls copy results back.
 *    Nothing the caller passes is kept or freed.
 *  - Calls are synchronous.  A handle owns two HIP streams (default
 *    priority: the solve, and the search of the next chunk of gridpoints)
 *    and is not thread safe; use one handle per host thread.
 *  - Fortran LOGICAL arrays are passed as int32 (non-zero = .TRUE.).
 *  - Gridpoint, slab and observation indices are 0-based at this ABI.
 */
#ifndef LETKF_GPU_H
#define LETKF_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define LETKF_OK            0
#define LETKF_EINVAL        1  /* invalid argument or parameter value       */
#define LETKF_EUNSUPPORTED  2  /* localizer class / option not on the GPU   */
#define LETKF_ENOMEM        3  /* device allocation failed                  */
#define LETKF_EHIP          4  /* HIP runtime error                         */
#define LETKF_ECAPACITY     5  /* a gridpoint's search set exceeds max_obs  */
#define LETKF_ESTATE        6  /* call order violated (e.g. run before set) */

/* ---- arithmetic of the local analysis ---------------------------------- */
#define LETKF_PREC_FP64     0  /* everything in FP64                        */
#define LETKF_PREC_FP32     1  /* products in FP32, eigensolve in FP64      */
#define LETKF_PREC_BF16X3   2  /* products split-BF16 (3 MFMA terms), FP32
                                  accumulation; eigensolve in FP64          */
#define LETKF_PREC_BF16     3  /* products plain BF16, FP32 accumulation     */
#define LETKF_PREC_BF16X4   4  /* products split-BF16 with all 4 MFMA terms,
                                  compensated FP32 accumulation; eigensolve
                                  in FP64 (FP32 products are compensated
                                  the same way)                             */

/* ---- execution path ---------------------------------------------------- */
#define LETKF_PATH_PLAIN    0  /* one gridpoint per workgroup, unbucketed   */
#define LETKF_PATH_BATCHED  1  /* gridpoint tiles bucketed by local obs count */

/* ---- localizer classes (YAML localization.class) ------------------------ */
#define LETKF_LOC_NOVRT     1  /* subclass.letkf_localizer/letkf_loc_novrt.F90 */
#define LETKF_LOC_OCEAN     2  /* subclass.letkf_localizer/letkf_loc_ocean.F90 */
#define LETKF_LOC_VGAUSS    3  /* horizontal GC x vertical Gaussian, one group
                                  per level (new class, no reference file)   */

#define LETKF_MAX_INTERP    32
#define LETKF_MAX_IDLIST    64

/* Problem sizes, fixed for the life of a handle.
 * Replaces: mpi.ens_size (src/letkf_mpi.F90:134), grid_ns and ij_count
 * (src/letkf_state.F90:182-192), max_obs (src/letkf_solver.F90:165). */
typedef struct letkf_gpu_dims {
  int64_t ens_size;   /* k                                                  */
  int64_t grid_ns;    /* slabs per gridpoint (sum of levels over variables) */
  int64_t ij_count;   /* gridpoints owned by this handle                    */
  int64_t max_obs;    /* search-set capacity per gridpoint; 0 -> 100000     */
  int32_t precision;  /* LETKF_PREC_*                                       */
  int32_t path;       /* LETKF_PATH_*                                       */
} letkf_gpu_dims;

/* One `linearinterp_lat` table exactly as written in the YAML `value:` list
 * (src/letkf_loc.F90:251-294).  The library adds the implicit 0 and 90
 * degree end points the same way the reference does. */
typedef struct letkf_linearinterp_lat {
  int32_t npoints;
  float   lat[LETKF_MAX_INTERP];     /* degrees, ascending */
  float   radius[LETKF_MAX_INTERP];  /* metres             */
} letkf_linearinterp_lat;

/* Localizer descriptor.  Replaces the letkf_localizer object selected by
 * localization.class (src/letkf_loc.F90:54-110,164-210). */
typedef struct letkf_loc_desc {
  int32_t cls;                        /* LETKF_LOC_*                         */
  /* LOC_NOVRT and LOC_VGAUSS: hzloc (letkf_loc_novrt.F90:98-105)          */
  letkf_linearinterp_lat hzloc;
  /* LOC_OCEAN (letkf_loc_ocean.F90:120-232)                                */
  letkf_linearinterp_lat hzloc_prof;
  letkf_linearinterp_lat hzloc_sat;
  float   tloc_prof;                  /* <= 0 disables (default -1)          */
  float   tloc_sat;
  int32_t vtloc_surf;                 /* 0 = none, 1 = bkg_t                 */
  float   bkg_t_delta;
  int32_t bkg_t_var_idx;              /* first slab of bkg_t_var (0-based)   */
  int32_t bkg_t_var_len;              /* levels of bkg_t_var                 */
  int32_t n_sat_obs;                  /* obsid list, 0 = empty list          */
  int32_t sat_obs[LETKF_MAX_IDLIST];
  int32_t n_sat_plats;                /* platid list, 0 = empty list         */
  int32_t sat_plats[LETKF_MAX_IDLIST];
  /* LOC_VGAUSS                                                             */
  float   vt_scale;                   /* Gaussian scale in zdim units        */
  int32_t nlev;                       /* number of levels                    */
  const float* lev_depth;             /* nlev level depths (copied)          */
} letkf_loc_desc;

/* Per-gridpoint diagnostics (the fields of diag.solver.nc,
 * src/letkf_solver.F90:59-62,212,220,265-268).  Any pointer may be NULL.
 * lg_* arrays are (ij_count, maxgroups) column-major like the reference. */
typedef struct letkf_gpu_diag {
  float*    maxhz;        /* ij_count                                  */
  float*    obscount;     /* ij_count: size of the search set          */
  float*    lg_obscount;  /* kept observations per group               */
  float*    lg_obsloc;    /* sum of localization weights per group     */
  uint64_t* lg_selhash;   /* order-free hash of the kept index set     */
} letkf_gpu_diag;

/* Phase timings of the last letkf_gpu_run (HIP events on the handle's
 * stream), in milliseconds. */
typedef struct letkf_gpu_timing {
  double search_ms;    /* local-obs search + localization          */
  double product_ms;   /* Yb^T R^-1 Yb and Yb^T R^-1 d              */
  double eigen_ms;     /* eigensolve, transform, application        */
  double finalize_ms;  /* recenter / clip / inflation / spread      */
  double total_ms;
  int64_t n_problems;  /* (gridpoint, group) solves with m > 0       */
  int64_t n_pairs;     /* kept (gridpoint, group, obs) triples        */
  int64_t n_cand;      /* candidate (gridpoint, obs) pairs tested     */
  int64_t n_launch;    /* launches per phase (gridpoint chunks)       */
  double order_ms;     /* batched path: problem order (keys + sort); the
                          product_ms above is then the product kernel alone */
} letkf_gpu_timing;

/* Launch choices of the analysis (no counterpart in the reference, which
 * has no device kernels to tune).  order: problem order of the batched
 * product, 0 gridpoint order, 1 longest first, 2 buckets of the kept
 * count largest first, 3 buckets with a Morton walk inside (the default);
 * bucket_bits:
 * the bucket edges are the powers of two split into 2^bucket_bits parts
 * (0..2); tridiag_reg: the Householder reduction for 32 <= k < 64, 1 one
 * problem per wave with the matrix in registers, 2 two problems per wave
 * (k <= 50; 1 above; the default), 0 the LDS kernel; overlap: 1 searches the next chunk of
 * gridpoints on a second stream while the current one is solved;
 * chunk_log2: at most 2^chunk_log2 (gridpoint, group) problems per chunk
 * (17..20); xcd: how the product kernel's workgroups meet the 8 XCDs (each
 * with its own L2), 0 in launch order, 1..5 each XCD runs of 32, 64, 128,
 * 256 or 512 consecutive problems, the runs of the XCDs interleaved
 * (default 4); tile:
 * the batched product's MFMA tiles, 0 32x32 blocks over rows padded to a
 * multiple of 32, 1 16x16 tiles over rows padded to a multiple of 16
 * (BF16 and split-BF16 products, k <= 111; 0 elsewhere; the default 1).  None of them
 * changes the analysis, except that the tile shape changes the order of
 * the FP32 accumulation (the products agree to rounding). */
typedef struct letkf_gpu_tuning {
  int32_t order;
  int32_t bucket_bits;
  int32_t tridiag_reg;
  int32_t n_tried;     /* autotune: candidates timed                  */
  double best_ms;      /* autotune: analysis time of the chosen one   */
  double default_ms;   /* autotune: analysis time of the defaults     */
  int32_t overlap;
  int32_t chunk_log2;
  int32_t xcd;
  int32_t tile;
} letkf_gpu_tuning;

typedef struct letkf_gpu_t letkf_gpu_t;

/* Library version string. */
const char* letkf_gpu_version(void);

/* Create a handle on the current HIP device.
 * Replaces letkf_core_init (src/letkf_core.F90:56-61) and the buffer
 * set-up of letkf_solver_init (src/letkf_solver.F90:127-140). */
int letkf_gpu_create(const letkf_gpu_dims* dims, letkf_gpu_t** handle);

/* Release every device resource of the handle (letkf_solver_final's
 * deallocation, src/letkf_solver.F90:397-527).  NULL is accepted. */
int letkf_gpu_destroy(letkf_gpu_t* handle);

/* Message of the last failed call on this handle ("" if none).  With a
 * NULL handle, the message of the last failed letkf_gpu_create. */
const char* letkf_gpu_last_error(const letkf_gpu_t* handle);

/* Background ensemble of the local gridpoints (src/letkf_state.F90:182-192):
 *   state_ij (k, ns, ij_count)  background perturbations (mean removed)
 *   mean_ij  (ns, ij_count)     background mean
 *   sprd_ij  (ns, ij_count)     background spread (k-1 denominator)
 *   lat_ij, lon_ij (ij_count)   degrees
 *   mask_ij  (ij_count)         non-zero = masked, skipped as in :207 */
int letkf_gpu_set_state(letkf_gpu_t* h, const float* state_ij,
                        const float* mean_ij, const float* sprd_ij,
                        const float* lat_ij, const float* lon_ij,
                        const int32_t* mask_ij);

/* Observations after letkf_obs_read (src/letkf_obs.F90:396-507): the
 * QC-compacted obs_def fields as separate arrays, obs_hx perturbations
 * (k, n_obs) and obs_hx_mean (n_obs).  The search structure that replaces
 * the kd-tree of :491-503 is built on the device by this call. */
int letkf_gpu_set_obs(letkf_gpu_t* h, int64_t n_obs,
                      const float* lat, const float* lon, const float* zdim,
                      const float* time, const float* val, const float* err,
                      const int32_t* obsid, const int32_t* platid,
                      const float* hx_pert, const float* hx_mean);

/* Background ensemble as the members' full values in file order, members
 * (ij_count, grid_ns, ens_size) with the gridpoint fastest: the state that
 * letkf_state_read_ens scatters (src/letkf_state.F90:312-314) before
 * letkf_state_init removes the mean.  The mean SUM(.,1)/ens_size, the
 * perturbations and the spread SQRT(DOT_PRODUCT(x',x')/(ens_size-1)) are
 * formed on the device in FP32 with the reference's order of operations
 * (src/letkf_state.F90:326-334).  Replaces letkf_gpu_set_state's host-side
 * preparation; ens_size must be at least 2. */
int letkf_gpu_set_state_members(letkf_gpu_t* h, const float* members,
                                const float* lat_ij, const float* lon_ij,
                                const int32_t* mask_ij);

/* Observations before letkf_obs_read's own processing
 * (src/letkf_obs.F90:396-507): every obs_def field including qc, and the
 * members' hx (ens_size, n_obs) as read by read_hx.  On the device:
 * obs_hx_mean = SUM(hx,1)/ens_size and obs_hx = hx - obs_hx_mean in FP32
 * (:459-466), then the obs with qc > 0 are dropped keeping the order
 * (:479-489), then the search structure is built as in letkf_gpu_set_obs. */
int letkf_gpu_set_obs_members(letkf_gpu_t* h, int64_t n_obs,
                              const float* lat, const float* lon,
                              const float* zdim, const float* time,
                              const float* val, const float* err,
                              const int32_t* obsid, const int32_t* platid,
                              const int32_t* qc, const float* hx_members);

/* State variables (statedef, src/letkf_state.F90:83-98): slab start
 * (0-based grid_s_idx), levels, ana_inc_max and ana_bounds (2 per var). */
int letkf_gpu_set_statevars(letkf_gpu_t* h, int32_t nvars,
                            const int32_t* grid_s_idx, const int32_t* levels,
                            const float* ana_inc_max, const float* ana_bounds);

/* Localizer (letkf_loc_init + class%init).  Classes other than the ones
 * above return LETKF_EUNSUPPORTED so the caller keeps its CPU loop. */
int letkf_gpu_set_localizer(letkf_gpu_t* h, const letkf_loc_desc* desc);

/* Inflation (solver.inflation.{rtps,rtpp,mul}); validated exactly like
 * letkf_solver_init (src/letkf_solver.F90:114-125). */
int letkf_gpu_set_solver(letkf_gpu_t* h, float rtps, float rtpp, float mul);

/* Run the local analysis over every gridpoint of the handle.  Inputs stay
 * unchanged on the device, so run may be repeated.  A run that fails once
 * its work has started leaves no analysis (LETKF_ESTATE from the getters
 * until a run succeeds).  LETKF_ECAPACITY (a search set above max_obs)
 * comes from the chunk's count pass when the localizer has several groups,
 * and after the run's work with one group. */
int letkf_gpu_run(letkf_gpu_t* h);

/* letkf_gpu_run with the whole analysis captured as one HIP graph and
 * replayed.  The first call (and the first after a buffer of the handle is
 * reallocated or a setting changes: set_statevars, set_localizer,
 * set_solver, set_tuning, more observations than the handle has room for)
 * runs eagerly, which sizes the kept lists, then captures the launch
 * sequence of the analysis on the handle's stream (search, list-space
 * check, problem order, products, transform, post-processing: no host
 * synchronisation inside); later calls replay it on the current inputs.
 * New backgrounds (set_state*) and new observation sets that fit the
 * handle's observation buffers (which keep a quarter of headroom over the
 * largest set so far) are replayed without a new capture, which is how a
 * cycled analysis runs.  If a replayed analysis needs more list space than
 * the capture was sized for, it is redone eagerly with larger lists (and
 * captured again on the next call).  The analysis is identical to
 * letkf_gpu_run's; timings of a replay hold total_ms only.  No counterpart
 * in the reference (it has no device work to capture). */
int letkf_gpu_run_graph(letkf_gpu_t* h);

/* Analysis in the reference's single-precision layout:
 * state_ij (k, ns, ij_count), mean_ij and sprd_ij (ns, ij_count).
 * Masked points hold the background perturbations, mean and spread. */
int letkf_gpu_get_state(letkf_gpu_t* h, float* state_ij, float* mean_ij,
                        float* sprd_ij);

/* Analysis in file order, the layout letkf_state_write_ens writes from
 * (src/letkf_state.F90:877-955): members (ij_count, grid_ns, ens_size),
 * mean and sprd (ij_count, grid_ns) for letkf_state_write_meansprd
 * (:357-451); FP32.  Any pointer may be NULL. */
int letkf_gpu_get_state_members(letkf_gpu_t* h, float* members, float* mean_ij,
                                float* sprd_ij);

/* letkf_gpu_get_state_members without any device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (tight serving or cycling loops, harnesses that forbid
 * hipMalloc).  The file-order conversion is staged through a 64 MiB device
 * array owned by the library (one per GPU, held until the process ends) in
 * slices of whole gridpoints, each copied down before the next; any
 * ij_count works, but one gridpoint's column (4 (k ns + 2 ns) bytes) must
 * fit in the array (LETKF_ECAPACITY otherwise).  Results are identical to
 * letkf_gpu_get_state_members.  Every handle on a GPU shares that GPU's
 * array: calls that overlap in time on one GPU (from several host threads)
 * race.  Any pointer may be NULL. */
int letkf_gpu_get_state_members_noalloc(letkf_gpu_t* h, float* members, float* mean_ij,
                                        float* sprd_ij);

/* The same analysis at the precision it was computed in (FP64 for
 * LETKF_PREC_FP64; the FP32 results widened otherwise). */
int letkf_gpu_get_state_f64(letkf_gpu_t* h, double* state_ij,
                            double* mean_ij, double* sprd_ij);

/* Diagnostics of the last run; maxgroups = letkf_gpu_maxgroups(h). */
int letkf_gpu_get_diag(letkf_gpu_t* h, const letkf_gpu_diag* diag);

/* localizer_class%maxgroups of the configured localizer. */
int letkf_gpu_maxgroups(const letkf_gpu_t* h, int32_t* maxgroups);

/* Kept observations of one (gridpoint, group) of the last run: original
 * obs indices and localization weights, in the library's order.  m_out is
 * set to the full count even when it exceeds cap. */
int letkf_gpu_get_selection(letkf_gpu_t* h, int64_t ij, int32_t group,
                            int32_t* idx_out, float* rloc_out, int64_t cap,
                            int64_t* m_out);

/* Phase timings and work counts of the last run. */
int letkf_gpu_get_timing(const letkf_gpu_t* h, letkf_gpu_timing* t);

/* Autotuner (no counterpart in the reference): times the analysis of the
 * handle's current inputs under each launch choice (coordinate search:
 * order and bucket edges, the tridiagonalization kernel, then the chunk
 * size with the chunk overlap on and off), `repeats`
 * runs each (best taken, 1..16), keeps the fastest (a candidate must beat
 * the current choice by 0.5%) in the handle and
 * returns it.  Needs everything letkf_gpu_run needs; leaves the analysis
 * of the last timed run, identical under every choice. */
int letkf_gpu_autotune(letkf_gpu_t* h, int32_t repeats, letkf_gpu_tuning* out);

/* Read or set the launch choices (for example a tuning saved from an
 * earlier autotune of the same configuration). */
int letkf_gpu_get_tuning(const letkf_gpu_t* h, letkf_gpu_tuning* out);
int letkf_gpu_set_tuning(letkf_gpu_t* h, const letkf_gpu_tuning* t);

#ifdef __cplusplus
}
#endif

#endif /* LETKF_GPU_H */
