// letkf_internal.h -- device-side descriptors shared by the kernels and the
// host orchestration of libletkf_gpu.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/letkf_gpu.h"

namespace lk {

// search cells: 1 x 1 degree, band-major, lon cells within a band
constexpr int kBands = 180;
constexpr int kLonCells = 360;
constexpr int kCells = kBands * kLonCells;
constexpr int kTab = LETKF_MAX_INTERP + 2;
constexpr int kMaxGroups = 128;       // largest maxgroups handled on the GPU
constexpr int kMaxK = 128;            // largest ensemble size

// completed linearinterp_lat tables and the localizer switches
struct LocDev {
  int cls;
  int maxgroups;
  int n_hz, n_prof, n_sat;
  float hz_lat[kTab], hz_dist[kTab];
  float prof_lat[kTab], prof_dist[kTab];
  float sat_lat[kTab], sat_dist[kTab];
  float tloc_prof, tloc_sat;
  int vtloc_surf;
  float bkg_t_delta;
  int bkg_t_var_idx, bkg_t_var_len;
  int max_var_lev;                    // largest levels over variables
  float vt_scale;
  int nlev;
};

// sorted observations (search order) and their per-localizer factors
struct ObsDev {
  int64_t n;
  const double* x; const double* y; const double* z;
  const float* zdim;
  const float* err;
  const float* dep;          // val - hx_mean in FP32
  const int32_t* oidx;       // original index
  const uint8_t* surf;       // LOC_OCEAN surface flag
  const float* tw_prof;      // GC(time, tloc_prof) or 1
  const float* tw_sat;
  const float* hx;           // (k, n) perturbations, sorted columns
  const int32_t* cell_start; // kCells + 1
  const float* hxp;          // (kp, n): hx rows, departure in row k, zero rows
  int kp;                    // 32-multiple >= k + 1
};

// list offsets are multiples of kListAlign entries (aligned vector loads)
constexpr int kListAlign = 8;
struct PadCount {
  __host__ __device__ int64_t operator()(int64_t c) const {
    return (c + kListAlign - 1) & ~(int64_t)(kListAlign - 1);
  }
};

// background state
struct StateDev {
  int64_t k, ns, nij;
  const float* state;  // (k, ns, nij)
  const float* mean;   // (ns, nij)
  const float* sprd;
  const float* lat;
  const float* lon;
  const int32_t* mask;
  const double* gx; const double* gy; const double* gz;
};

struct VarsDev {
  int nvars;
  const int32_t* slab_var;   // (ns) variable of each slab, -1 if none
  const int32_t* slab_lev;   // (ns) 0-based level within the variable
  const float* inc_max;      // (nvars)
  const float* bounds;       // (2, nvars)
  const float* lev_depth;    // LOC_VGAUSS level depths
};

// per-chunk work arrays; pair = (ij - i0) * G + g
struct ChunkDev {
  int64_t i0, ngp;
  int G;
  int64_t* cnt;        // kept count per pair
  int64_t* off;        // list offset per pair
  uint8_t* solve;      // pair needs a transform
  int32_t* list_idx;   // sorted obs index
  float* list_w;       // localization weight, or sqrt(weight / err) (list_mode 1)
  int list_mode;       // 0: rloc (FP64 products), 1: sqrt(rloc / err) (mixed)
  int xcd;             // product kernels' block order: 0 as launched, 1..3 runs of
                       // 32 .. 512 problems per XCD, interleaved (xcd_order)
};

// XCD-interleaved runs: the dispatcher deals workgroups round-robin over
// the 8 XCDs, each with its own L2.  Remap so that the workgroups of one XCD
// (b % 8) take runs of R = 16 << mode consecutive problems (mode 1..5: 32
// to 512) and the runs of the 8 XCDs are interleaved: the problems in
// flight on one XCD are neighbours in the launch order (sharing observation
// columns in that XCD's L2) while the 8 XCDs still move through the order
// together (a largest-first order stays balanced; giving each XCD one
// contiguous range would not).  A bijection on [0, nb): the tail past the
// last whole group of 8 runs keeps its place.  Mode 0: launch order.
constexpr int kXcdModes = 6;   // xcd_order modes 0..5
__device__ __forceinline__ int64_t xcd_order(int mode, int64_t b, int64_t nb) {
  constexpr int kXcd = 8;
  if (mode <= 0) return b;
  const int64_t run = 16ll << mode;
  const int64_t full = nb / (kXcd * run) * (kXcd * run);
  if (b >= full) return b;
  const int64_t x = b % kXcd, j = b / kXcd;
  return ((j / run) * kXcd + x) * run + j % run;
}

struct DiagDev {
  float* maxhz;        // (nij)
  int32_t* obscount;   // (nij)
  int32_t* lg_m;       // (G, nij)
  double* lg_obsloc;   // (G, nij)
  uint64_t* lg_hash;   // (G, nij)
  int32_t* mld;        // (nij) mixed-layer level (LOC_OCEAN bkg_t)
  // counters of the search, spread over kCtrSlots slots (slot = block index
  // mod kCtrSlots) so that the waves' atomics do not queue on one line:
  // slot s holds solves at ctr[s * kCtrStride + kCtrSolves], candidates,
  // kept pairs of solves and the largest search set at the other offsets
  unsigned long long* ctr;
};

constexpr int kCtrSlots = 64;
constexpr int kCtrStride = 64;        // 512 B per slot
constexpr int kCtrSolves = 0, kCtrCand = 16, kCtrPairs = 32, kCtrMax = 48;
constexpr int kCtrOverflow = 8;       // slot 0 only: pairs refused list space

// group of slab s at a gridpoint (letkf_loc_ocean.F90:360-375 for bkg_t)
__device__ __forceinline__ int slab_group(const LocDev& L, const VarsDev& V,
                                          int s, int mld) {
  if (L.cls == LETKF_LOC_VGAUSS) return V.slab_lev[s];
  if (L.cls == LETKF_LOC_OCEAN && L.vtloc_surf == 1) {
    if (V.slab_var[s] < 0) return -1;
    return (V.slab_lev[s] + 1 <= mld) ? 0 : 1;
  }
  return 0;
}

// buffer resource over n bytes at p: loads past the end read 0 (and touch
// no memory), so a load can be issued for every lane with an offset past
// the end where its value is not wanted
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
constexpr uint32_t kBufOut = 0x7ffffff0u;   // an offset past every buffer
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, int64_t n) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)n, 0x00020000);
}
// FP64 / FP32 value at byte offset vo + so of a buffer (0 past its end)
__device__ __forceinline__ double buf_f64(__amdgpu_buffer_rsrc_t r, uint32_t vo, int so) {
  const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0);
  return __hiloint2double((int)v[1], (int)v[0]);
}
__device__ __forceinline__ float buf_f32(__amdgpu_buffer_rsrc_t r, uint32_t vo, int so) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
}

// host launchers (kernels_*.hip)
void launch_obs_prep(hipStream_t st, int64_t n, int k, const float* lat,
                     const float* lon, int32_t* cell, int32_t* iota);
void launch_obs_gather(hipStream_t st, int64_t n, int k, const int32_t* perm,
                       const float* lat, const float* lon, const float* zdim,
                       const float* time, const float* val, const float* err,
                       const int32_t* obsid, const int32_t* platid,
                       const float* hx, const float* hxmean,
                       const int32_t* sorted_cell, double* x, double* y,
                       double* z, float* ozdim, float* oerr, float* odep,
                       int32_t* oidx, int32_t* oobsid, int32_t* oplatid,
                       float* otime, float* ohx, int32_t* cell_start);
void launch_obs_flags(hipStream_t st, const LocDev& L, int64_t n,
                      const int32_t* obsid, const int32_t* platid,
                      const float* time, const int32_t* sat_obs, int n_sat_obs,
                      const int32_t* sat_plats, int n_sat_plats, uint8_t* surf,
                      float* tw_prof, float* tw_sat);
void launch_obs_pad_hx(hipStream_t st, int64_t n, int k, int kp, const float* ohx,
                       const float* odep, float* hxp);
// observation arrays of set_obs in input order: lat, lon, zdim, time, val,
// err, hx_mean (f); obsid, platid (i); hx perturbations (k, n)
struct ObsFields {
  static constexpr int kF = 7, kI = 2;
  float* f[kF];
  int32_t* i[kI];
  float* hx;
};
void launch_f64_to_f32(hipStream_t st, int64_t n, const double* in, float* out);
void launch_state_from_members(hipStream_t st, int64_t nij, int64_t ns, int k,
                               const float* members, float* state, float* mean, float* sprd);
void launch_members_from_state(hipStream_t st, int64_t nij, int64_t ns, int k,
                               const double* ana, const double* amean, const double* asprd,
                               float* members, float* mean, float* sprd);
void launch_obs_hx_split(hipStream_t st, int64_t n, int k, const float* hx, const int32_t* qc,
                         float* pert, float* hmean, uint8_t* keep);
void launch_obs_compact(hipStream_t st, int64_t nsel, int k, const int32_t* sel,
                        const ObsFields& in, const ObsFields& out);
void launch_grid_xyz(hipStream_t st, int64_t n, const float* lat,
                     const float* lon, double* x, double* y, double* z);
void launch_clear_u64(hipStream_t st, int64_t n, unsigned long long* p);
void launch_list_guard(hipStream_t st, int64_t np, int64_t* cnt, const int64_t* off,
                       uint8_t* solve, int64_t cap, unsigned long long* ctr);
void launch_search(hipStream_t st, bool fill, const LocDev& L,
                   const ObsDev& O, const StateDev& S, const VarsDev& V,
                   const ChunkDev& C, const DiagDev& D);
void launch_product_f64(hipStream_t st, const ObsDev& O, const StateDev& S,
                        const ChunkDev& C, double* B, double* bvec);
hipError_t launch_eigen_apply_f64(hipStream_t st, const LocDev& L, const StateDev& S,
                                  const VarsDev& V, const ChunkDev& C, const DiagDev& D,
                                  const double* B, const double* bvec, const int* only_flagged,
                                  double* ana);
// tile 0: 32x32 MFMA blocks; tile 1: 16x16 tiles where product_t16_supported
void launch_product_mx(hipStream_t st, int prec, const ObsDev& O, const StateDev& S,
                       const ChunkDev& C, const int32_t* perm, double* B, double* bvec,
                       int fold_log2, int tile);
// a 16x16-tile product kernel exists (BF16 and split-BF16, k <= 111)
bool product_t16_supported(int prec, int k);
void launch_bucket_keys(hipStream_t st, int mode, int bbits, int64_t np, int G, int64_t i0,
                        const int64_t* cnt, const uint8_t* solve, const float* lat,
                        const float* lon, uint32_t* keys, int32_t* vals);
// register-resident tridiagonalization for rows of KQ (a multiple of 8,
// k <= KQ <= 64); false if no kernel exists for (KQ, NV)
bool launch_tridiag_reg(hipStream_t st, unsigned nb, int KQ, int NV, int k, int K,
                        const LocDev& L, const StateDev& S, const VarsDev& V,
                        const ChunkDev& C, const DiagDev& D, double* B, const double* bvec,
                        double* tau, double* dd, double* ee, double* U, int* nvec,
                        int* status);
// the same reduction for two problems per wave (rows of KQ <= 50, at most
// 3 carried vectors), the problems taken in the order perm (nullptr:
// 0..np-1).  With rnb > 0 (rational node tables) and carried vectors the
// rational solve runs in the same kernel.  Returns 0 if no kernel exists
// for (KQ, NV), 1 for the reduction alone, 2 with the rational solve.
int launch_tridiag_pair(hipStream_t st, int64_t np, const int32_t* perm, int KQ, int NV, int k,
                        int K, const LocDev& L, const StateDev& S, const VarsDev& V,
                        const ChunkDev& C, const DiagDev& D, double* B, const double* bvec,
                        double* tau, double* dd, double* ee, double* U, int* nvec,
                        int* status, const double* rtab, const int* rtabn, int rnb);
int eigen_pipeline_K(int k);
void eigen_caps(int k, int* cap_rot, int* cap_it);
// vectors carried per problem on the vector path (b + slabs of a group,
// rounded up to an instantiated count), 0 if maxslab is too large
int eigen_vec_nv(int maxslab);
// node tables of the rational approximation of x^(-1/2) on [1, 2^b]:
// tab[b * 64 + j] = s_j, tab[b * 64 + 32 + j] = w_j (j < n[b] <= 31)
struct RatTab {
  const double* tab = nullptr;
  const int* n = nullptr;
  int nb = 0;                 // 0: not used
};
constexpr int kRatBuckets = 48;
// NV == 0: eigenvectors formed (Z, V = QZ) and applied to every slab;
// NV > 0: only the NV carried vectors are transformed (U: NV x K per
// problem), by the QL rotations, or, with a node table (R.nb > 0), by the
// rational approximation of A^(-1/2) without an eigensolve
void launch_eigen_pipeline(hipStream_t st, const LocDev& L, const StateDev& S,
                           const VarsDev& V, const ChunkDev& C, const DiagDev& D,
                           double* B, const double* bvec, double* tau, double* dd,
                           double* ee, double2* rot, int cap_rot, int4* itr,
                           int cap_it, int* nit, int* status, double* Z, int NV, double* U,
                           int* nvec, const RatTab& R, double* ana, int trireg,
                           const int32_t* perm, bool fuse);
// explicit transform matrix and its MFMA application (64 <= k <= 128)
bool transform_mat_supported(int k);
// post-processing parameters (src/letkf_solver.F90:299-379) and the
// analysis mean and spread it writes
struct FinArgs {
  float rtps, rtpp, mul;
  double* amean;
  double* asprd;
};
// fin != nullptr (mixed precisions only): the apply post-processes the slabs
// it transforms, and launch_finalize must then run with skip_solved
bool transform_mat_fuses_finalize(int prec);
void launch_transform_mat(hipStream_t st, int prec, const LocDev& L, const StateDev& S,
                          const VarsDev& V, const ChunkDev& C, const DiagDev& D, double* B,
                          const double* bvec, double* tau, double* dd, double* ee,
                          const RatTab& R, double* ana, bool blk, const FinArgs* fin);
// skip_solved: only masked gridpoints and the slabs of groups without a
// solved problem (the rest were post-processed by the fused apply)
void launch_finalize(hipStream_t st, const LocDev& L, const StateDev& S,
                     const VarsDev& V, const ChunkDev& C, const DiagDev& D,
                     const FinArgs& F, double* ana, bool skip_solved);

}  // namespace lk
