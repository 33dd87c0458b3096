// letkf_gpu.hip -- C ABI (include/letkf_gpu.h) and host orchestration.
//
// letkf_gpu_run() replaces letkf_solver_run (src/letkf_solver.F90:150-389).
// Gridpoints are processed in chunks so that the per-(gridpoint, group)
// work arrays stay bounded; each chunk runs
//   search count -> offsets scan -> search fill -> products -> eigensolve
//   and application -> post-processing
// on the handle's stream.  Inputs stay untouched on the device, so a run can
// be repeated (benchmarks, cycling).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "det_math.h"
#include "letkf_internal.h"
#include "ratnodes.h"

using namespace lk;

namespace lk {
size_t eigen_lds_bytes(int k);
}

struct letkf_gpu_t {
  letkf_gpu_dims dims{};
  int64_t k = 0, ns = 0, nij = 0;
  hipStream_t stream = nullptr;
  int device = 0;
  std::string err;
  bool have_state = false, have_obs = false, have_vars = false, have_loc = false;
  bool have_solver = false, have_run = false;

  // background
  float *d_state = nullptr, *d_mean = nullptr, *d_sprd = nullptr;
  float *d_lat = nullptr, *d_lon = nullptr;
  int32_t* d_mask = nullptr;
  double *d_gx = nullptr, *d_gy = nullptr, *d_gz = nullptr;

  // observations (raw copies, then sorted)
  int64_t nobs = 0;
  double *d_ox = nullptr, *d_oy = nullptr, *d_oz = nullptr;
  float *d_ozdim = nullptr, *d_oerr = nullptr, *d_odep = nullptr, *d_otime = nullptr;
  int32_t *d_oidx = nullptr, *d_oobsid = nullptr, *d_oplatid = nullptr;
  float* d_ohx = nullptr;
  float* d_ohxp = nullptr;   // padded operand columns (kp, n)
  int kp = 0;
  int32_t* d_cell_start = nullptr;
  uint8_t* d_osurf = nullptr;
  float *d_otwp = nullptr, *d_otws = nullptr;

  // variables / localizer / solver
  int nvars = 0;
  std::vector<int32_t> var_s, var_lev;
  int32_t *d_slab_var = nullptr, *d_slab_lev = nullptr;
  float *d_inc_max = nullptr, *d_bounds = nullptr;
  LocDev loc{};
  std::vector<int32_t> sat_obs, sat_plats;
  float* d_lev_depth = nullptr;
  float rtps = 0.0f, rtpp = 0.0f, mul = 1.0f;
  int G = 1;

  // outputs
  double *d_ana = nullptr, *d_amean = nullptr, *d_asprd = nullptr;
  float* d_maxhz = nullptr;
  int32_t *d_obscount = nullptr, *d_lgm = nullptr, *d_mld = nullptr;
  unsigned long long* d_ctr = nullptr;   // search counters, kCtrSlots x kCtrStride
  unsigned long long* h_ctr = nullptr;   // pinned copy
  double* d_lgloc = nullptr;
  uint64_t* d_lghash = nullptr;

  // chunk work
  int64_t chunk_gp = 0;
  int64_t alloc_G = 0;
  int64_t* d_cnt = nullptr;
  int64_t* d_off = nullptr;
  uint8_t* d_solve = nullptr;
  int32_t* d_list_idx = nullptr;
  float* d_list_w = nullptr;
  int64_t list_cap = 0;
  // second set of per-chunk search outputs and a second stream: the search
  // of chunk c + 1 runs while chunk c is solved
  hipStream_t stream2 = nullptr;
  int64_t* d_cnt2 = nullptr;
  int64_t* d_off2 = nullptr;
  uint8_t* d_solve2 = nullptr;
  int32_t* d_list_idx2 = nullptr;
  float* d_list_w2 = nullptr;
  int64_t list_cap2 = 0;
  void* d_tmp2 = nullptr;
  int64_t* h_tot2 = nullptr;   // pinned
  int64_t alloc_chunk = 0;
  double *d_B = nullptr, *d_b = nullptr;
  // eigensolver work (per chunk pair)
  double *d_tau = nullptr, *d_dd = nullptr, *d_ee = nullptr;
  double2* d_rot = nullptr;
  int4* d_itr = nullptr;
  int *d_nit = nullptr, *d_status = nullptr;
  double* d_Z = nullptr;
  // vector path: carried vectors (NV x K per problem) and slab counts
  double* d_U = nullptr;
  int* d_nvec = nullptr;
  int vec_nv = 0, alloc_nv = -1;
  // batched path: node tables of the rational A^(-1/2)
  double* d_rat = nullptr;
  int* d_ratn = nullptr;
  bool use_rat = false;
  // batched path: longest-first order of the problems of a chunk
  uint32_t *d_bkey = nullptr, *d_bkey2 = nullptr;
  int32_t *d_bval = nullptr, *d_perm = nullptr;
  void* d_stmp = nullptr;
  size_t stmp_bytes = 0;
  int cap_rot = 0, cap_it = 0, eigK = 0;
  void* d_tmp = nullptr;
  size_t tmp_bytes = 0;
  int64_t* h_tot = nullptr;    // pinned

  letkf_gpu_timing timing{};
  // device address of this device's no-allocation staging array
  void* d_scratch = nullptr;
  hipError_t scratch_err = hipSuccess;
  // captured analysis (letkf_gpu_run_graph): valid while gen == graph_gen.
  // gen counts every reallocation of a handle buffer and every change of a
  // setting the kernels take by value
  uint64_t gen = 1, graph_gen = 0;
  int graph_cfg = -1;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int64_t obs_cap = 0;     // columns allocated for the observation buffers
  int list_headroom_div = 4;   // lists grow to total + total / div
  // device memory of the handle: live allocations and an optional cap
  std::unordered_map<void*, size_t> live;
  size_t live_bytes = 0;
  size_t mem_limit = 0;   // 0 = none
  int tune_xcd = 4;       // product kernels: runs of 256 problems per XCD
  int tune_tile = 1;      // batched product: 0 32x32 MFMA blocks, 1 16x16 tiles
                          // (where that kernel exists)
  int fold_log2 = 7;      // compensated products: partial sums folded every 2^n obs
  int tune_bucket = 3;    // 0 gridpoint order, 1 longest first, 2 log2 buckets,
                          // 3 log2 buckets with a Morton walk inside
  int tune_bbits = 0;     // bucket edges: powers of two split in 2^bbits parts
  int tune_trireg = 2;    // tridiagonalization (32 <= k < 64): 0 LDS, 1 register,
                          // 2 two problems per wave (k <= 50; 1 above)
  int tune_overlap = 0;   // search of the next chunk concurrent with the solve
  int tune_chunk_log2 = 20;   // at most 2^n problems per chunk (17..20)
  int tune_vec = 1;       // vector path when every group has few slabs
  int tune_fuse = 1;      // rational solve inside the two-problem tridiagonalization
  int tune_matblk = 1;    // 64 <= k <= 128: 2D-block reduction and back-transform
  int tune_fusefin = 1;   // 64 <= k <= 128, mixed: post-processing inside the apply
  int tune_rat = 1;       // vector path: rational A^(-1/2), no eigensolve (1 batched
                          // path only, 2 the plain FP64 path too, 0 off)
  int tune_mat = 1;       // batched, 64 <= k <= 128: explicit T applied on MFMA
  bool use_mat = false;
  int alloc_mat = -1;

  bool batched() const {
    return dims.path == LETKF_PATH_BATCHED && dims.precision != LETKF_PREC_FP64;
  }
};

static std::string g_create_err;

// ---------------------------------------------------------------------------
// Staging memory of letkf_gpu_get_state_members_noalloc: a file-scope device
// array, so that the call allocates nothing.  Every GPU the library runs on
// holds its own copy (64 MiB) until the process ends.  In host code the
// array's name is not a device address; its address on each device is
// looked up once per device, with that device current, by
// hipGetSymbolAddress at the first letkf_gpu_create there, and only that
// address is handed to kernels and copies.  All calls on one device share
// the array: calls that overlap in time on one GPU race.
// ---------------------------------------------------------------------------
static constexpr size_t kNoAllocScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::once_flag g_scratch_once[kMaxDevices];
void* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

// device address of `scratch` on the current device `dev` (looked up once)
void* scratch_address(int dev, hipError_t* err) {
  if (dev < 0 || dev >= kMaxDevices) {
    *err = hipErrorInvalidDevice;
    return nullptr;
  }
  std::call_once(g_scratch_once[dev], [dev] {
    void* p = nullptr;
    g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_ptr[dev] = g_scratch_err[dev] == hipSuccess ? p : nullptr;
  });
  *err = g_scratch_err[dev];
  return g_scratch_ptr[dev];
}
}  // namespace

static int fail(letkf_gpu_t* h, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (h) h->err = buf; else g_create_err = buf;
  return code;
}

#define HIPCHK(h, call)                                                        \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail((h), e_ == hipErrorOutOfMemory ? LETKF_ENOMEM : LETKF_EHIP,  \
                  "%s: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__,    \
                  __LINE__);                                                   \
  } while (0)

// Every device allocation of a handle goes through dalloc / dfree, which
// keep the handle's live bytes.  With LETKF_GPU_MEM_LIMIT (bytes, read at
// create) an allocation that would take them past the limit fails with
// hipErrorOutOfMemory, as a full device does: the LETKF_ENOMEM path can be
// exercised on a GPU with memory to spare.
static hipError_t raw_alloc(letkf_gpu_t* h, void** p, size_t bytes) {
  if (h->mem_limit && h->live_bytes + bytes > h->mem_limit) return hipErrorOutOfMemory;
  const hipError_t e = hipMalloc(p, bytes);
  if (e == hipSuccess) {
    h->live[*p] = bytes;
    h->live_bytes += bytes;
  }
  return e;
}
static void raw_free(letkf_gpu_t* h, void* p) {
  if (!p) return;
  auto it = h->live.find(p);
  if (it != h->live.end()) {
    h->live_bytes -= it->second;
    h->live.erase(it);
  }
  hipFree(p);
}
// a buffer pointer stored in the handle itself (not a local temporary)
static bool handle_member(const letkf_gpu_t* h, const void* p) {
  const char* a = (const char*)p;
  return a >= (const char*)h && a < (const char*)(h + 1);
}
template <typename T>
static hipError_t dalloc(letkf_gpu_t* h, T** p, size_t n) {
  if (handle_member(h, p)) ++h->gen;     // a captured analysis no longer matches
  if (*p) { raw_free(h, *p); *p = nullptr; }
  if (n == 0) n = 1;
  return raw_alloc(h, (void**)p, n * sizeof(T));
}
// allocate a handle buffer of a size fixed for the handle's life once
template <typename T>
static hipError_t dalloc_once(letkf_gpu_t* h, T** p, size_t n) {
  return *p ? hipSuccess : dalloc(h, p, n);
}
template <typename T>
static void dfree(letkf_gpu_t* h, T*& p) {
  raw_free(h, p);
  p = nullptr;
}

extern "C" {

const char* letkf_gpu_version(void) { return "letkf_gpu 0.1 (gfx950)"; }

const char* letkf_gpu_last_error(const letkf_gpu_t* h) {
  return h ? h->err.c_str() : g_create_err.c_str();
}

int letkf_gpu_create(const letkf_gpu_dims* d, letkf_gpu_t** out) {
  if (!d || !out) return fail(nullptr, LETKF_EINVAL, "null argument");
  *out = nullptr;
  if (d->ens_size < 2 || d->ens_size > kMaxK - 1)
    return fail(nullptr, LETKF_EUNSUPPORTED, "ens_size %lld outside [2, %d]",
                (long long)d->ens_size, kMaxK - 1);
  if (d->grid_ns < 1 || d->ij_count < 0)
    return fail(nullptr, LETKF_EINVAL, "invalid grid_ns / ij_count");
  if (d->precision < LETKF_PREC_FP64 || d->precision > LETKF_PREC_BF16X4)
    return fail(nullptr, LETKF_EINVAL, "invalid precision %d", d->precision);
  if (d->path != LETKF_PATH_PLAIN && d->path != LETKF_PATH_BATCHED)
    return fail(nullptr, LETKF_EINVAL, "invalid path %d", d->path);
  if (d->path == LETKF_PATH_PLAIN && d->precision != LETKF_PREC_FP64)
    return fail(nullptr, LETKF_EUNSUPPORTED, "the plain path computes in FP64 only");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(nullptr, LETKF_EHIP, "no HIP device");
  letkf_gpu_t* h = new letkf_gpu_t();
  h->dims = *d;
  if (h->dims.max_obs <= 0) h->dims.max_obs = 100000;   // src/letkf_solver.F90:165
  h->k = d->ens_size;
  h->ns = d->grid_ns;
  h->nij = d->ij_count;
  if (const char* e = getenv("LETKF_GPU_MEM_LIMIT")) h->mem_limit = strtoull(e, nullptr, 10);
  if (const char* e = getenv("LETKF_GPU_XCD")) h->tune_xcd = std::min(kXcdModes - 1, std::max(0, atoi(e)));
  if (const char* e = getenv("LETKF_GPU_FOLD")) h->fold_log2 = std::min(10, std::max(5, atoi(e)));
  if (const char* e = getenv("LETKF_GPU_TILE")) h->tune_tile = atoi(e) != 0;
  if (const char* e = getenv("LETKF_GPU_FUSE")) h->tune_fuse = atoi(e) != 0;
  if (const char* e = getenv("LETKF_GPU_MATBLK")) h->tune_matblk = atoi(e) != 0;
  if (const char* e = getenv("LETKF_GPU_FUSEFIN")) h->tune_fusefin = atoi(e) != 0;
  if (const char* e = getenv("LETKF_GPU_BUCKET")) h->tune_bucket = std::min(3, std::max(0, atoi(e)));
  if (const char* e = getenv("LETKF_GPU_BBITS")) h->tune_bbits = std::min(2, std::max(0, atoi(e)));
  if (const char* e = getenv("LETKF_GPU_TRIREG")) h->tune_trireg = std::min(2, std::max(0, atoi(e)));
  if (const char* e = getenv("LETKF_GPU_VEC")) h->tune_vec = atoi(e);
  if (const char* e = getenv("LETKF_GPU_RAT")) h->tune_rat = atoi(e);
  if (const char* e = getenv("LETKF_GPU_MAT")) h->tune_mat = atoi(e);
  if (const char* e = getenv("LETKF_GPU_OVERLAP")) h->tune_overlap = atoi(e) != 0;
  if (const char* e = getenv("LETKF_GPU_CHUNK_LOG2"))
    h->tune_chunk_log2 = std::min(20, std::max(17, atoi(e)));
  hipGetDevice(&h->device);
  h->d_scratch = scratch_address(h->device, &h->scratch_err);
  // both streams at the default priority
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return fail(nullptr, LETKF_EHIP, "hipStreamCreate failed");
  }
  if (hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking) != hipSuccess) {
    hipStreamDestroy(h->stream);
    delete h;
    return fail(nullptr, LETKF_EHIP, "hipStreamCreate failed");
  }
  if (hipHostMalloc((void**)&h->h_tot, 4 * sizeof(int64_t), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&h->h_tot2, 4 * sizeof(int64_t), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&h->h_ctr, sizeof(unsigned long long) * kCtrSlots * kCtrStride,
                    hipHostMallocDefault) != hipSuccess) {
    if (h->h_tot) hipHostFree(h->h_tot);
    if (h->h_tot2) hipHostFree(h->h_tot2);
    hipStreamDestroy(h->stream);
    hipStreamDestroy(h->stream2);
    delete h;
    return fail(nullptr, LETKF_ENOMEM, "pinned allocation failed");
  }
  *out = h;
  return LETKF_OK;
}

int letkf_gpu_destroy(letkf_gpu_t* h) {
  if (!h) return LETKF_OK;
  hipStreamSynchronize(h->stream);
  hipStreamSynchronize(h->stream2);
  if (h->gexec) hipGraphExecDestroy(h->gexec);
  if (h->graph) hipGraphDestroy(h->graph);
  dfree(h, h->d_state); dfree(h, h->d_mean); dfree(h, h->d_sprd); dfree(h, h->d_lat); dfree(h, h->d_lon);
  dfree(h, h->d_mask); dfree(h, h->d_gx); dfree(h, h->d_gy); dfree(h, h->d_gz);
  dfree(h, h->d_ox); dfree(h, h->d_oy); dfree(h, h->d_oz); dfree(h, h->d_ozdim); dfree(h, h->d_oerr);
  dfree(h, h->d_odep); dfree(h, h->d_otime); dfree(h, h->d_oidx); dfree(h, h->d_oobsid);
  dfree(h, h->d_oplatid); dfree(h, h->d_ohx); dfree(h, h->d_ohxp); dfree(h, h->d_cell_start); dfree(h, h->d_osurf);
  dfree(h, h->d_otwp); dfree(h, h->d_otws);
  dfree(h, h->d_slab_var); dfree(h, h->d_slab_lev); dfree(h, h->d_inc_max); dfree(h, h->d_bounds);
  dfree(h, h->d_lev_depth);
  dfree(h, h->d_ana); dfree(h, h->d_amean); dfree(h, h->d_asprd); dfree(h, h->d_maxhz);
  dfree(h, h->d_obscount); dfree(h, h->d_lgm); dfree(h, h->d_mld); dfree(h, h->d_ctr);
  dfree(h, h->d_lgloc); dfree(h, h->d_lghash);
  dfree(h, h->d_cnt); dfree(h, h->d_off); dfree(h, h->d_solve); dfree(h, h->d_list_idx);
  dfree(h, h->d_list_w); dfree(h, h->d_B); dfree(h, h->d_b); dfree(h, h->d_tmp);
  dfree(h, h->d_cnt2); dfree(h, h->d_off2); dfree(h, h->d_solve2); dfree(h, h->d_list_idx2);
  dfree(h, h->d_list_w2); dfree(h, h->d_tmp2);
  dfree(h, h->d_tau); dfree(h, h->d_dd); dfree(h, h->d_ee); dfree(h, h->d_rot); dfree(h, h->d_itr);
  dfree(h, h->d_nit); dfree(h, h->d_status); dfree(h, h->d_Z); dfree(h, h->d_U); dfree(h, h->d_nvec);
  dfree(h, h->d_rat); dfree(h, h->d_ratn);
  dfree(h, h->d_bkey); dfree(h, h->d_bkey2); dfree(h, h->d_bval); dfree(h, h->d_perm); dfree(h, h->d_stmp);
  if (h->h_tot) hipHostFree(h->h_tot);
  if (h->h_tot2) hipHostFree(h->h_tot2);
  if (h->h_ctr) hipHostFree(h->h_ctr);
  hipStreamDestroy(h->stream);
  hipStreamDestroy(h->stream2);
  delete h;
  return LETKF_OK;
}

}  // extern "C"

// background buffers: sizes fixed by the handle's dims, allocated by the
// first set_state* and reused by later ones (a new cycle's background keeps
// the device addresses a captured analysis was recorded with)
static hipError_t alloc_state(letkf_gpu_t* h) {
  const size_t k = (size_t)h->k, ns = (size_t)h->ns, nij = (size_t)h->nij;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = dalloc_once(h, &h->d_state, k * ns * nij);
  if (e == hipSuccess) e = dalloc_once(h, &h->d_mean, ns * nij);
  if (e == hipSuccess) e = dalloc_once(h, &h->d_sprd, ns * nij);
  if (e == hipSuccess) e = dalloc_once(h, &h->d_lat, nij);
  if (e == hipSuccess) e = dalloc_once(h, &h->d_lon, nij);
  if (e == hipSuccess) e = dalloc_once(h, &h->d_mask, nij);
  if (e == hipSuccess) e = dalloc_once(h, &h->d_gx, nij);
  if (e == hipSuccess) e = dalloc_once(h, &h->d_gy, nij);
  if (e == hipSuccess) e = dalloc_once(h, &h->d_gz, nij);
  return e;
}

extern "C" {

int letkf_gpu_set_state_members(letkf_gpu_t* h, const float* members, const float* lat_ij,
                                const float* lon_ij, const int32_t* mask_ij) {
  if (!h) return LETKF_EINVAL;
  const int64_t nij = h->nij, k = h->k, ns = h->ns;
  if (nij > 0 && (!members || !lat_ij || !lon_ij || !mask_ij))
    return fail(h, LETKF_EINVAL, "null state array");
  if (k < 2) return fail(h, LETKF_EINVAL, "ens_size below 2: no spread");
  h->have_run = false;   // new inputs: the previous analysis is gone
  HIPCHK(h, alloc_state(h));
  if (nij > 0) {
    float* raw = nullptr;
    HIPCHK(h, dalloc(h, &raw, (size_t)(k * ns * nij)));
    hipError_t e = hipMemcpyAsync(raw, members, sizeof(float) * k * ns * nij,
                                  hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
      launch_state_from_members(h->stream, nij, ns, (int)k, raw, h->d_state, h->d_mean,
                                h->d_sprd);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    dfree(h, raw);
    HIPCHK(h, e);
    HIPCHK(h, hipMemcpyAsync(h->d_lat, lat_ij, sizeof(float) * nij, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_lon, lon_ij, sizeof(float) * nij, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_mask, mask_ij, sizeof(int32_t) * nij, hipMemcpyHostToDevice, h->stream));
    launch_grid_xyz(h->stream, nij, h->d_lat, h->d_lon, h->d_gx, h->d_gy, h->d_gz);
    HIPCHK(h, hipGetLastError());
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->have_state = true;
  return LETKF_OK;
}

int letkf_gpu_set_state(letkf_gpu_t* h, const float* state_ij, const float* mean_ij,
                        const float* sprd_ij, const float* lat_ij, const float* lon_ij,
                        const int32_t* mask_ij) {
  if (!h) return LETKF_EINVAL;
  const int64_t nij = h->nij, k = h->k, ns = h->ns;
  if (nij > 0 && (!state_ij || !mean_ij || !sprd_ij || !lat_ij || !lon_ij || !mask_ij))
    return fail(h, LETKF_EINVAL, "null state array");
  h->have_run = false;
  HIPCHK(h, alloc_state(h));
  if (nij > 0) {
    HIPCHK(h, hipMemcpyAsync(h->d_state, state_ij, sizeof(float) * k * ns * nij, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_mean, mean_ij, sizeof(float) * ns * nij, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_sprd, sprd_ij, sizeof(float) * ns * nij, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_lat, lat_ij, sizeof(float) * nij, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_lon, lon_ij, sizeof(float) * nij, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_mask, mask_ij, sizeof(int32_t) * nij, hipMemcpyHostToDevice, h->stream));
    launch_grid_xyz(h->stream, nij, h->d_lat, h->d_lon, h->d_gx, h->d_gy, h->d_gz);
    HIPCHK(h, hipGetLastError());
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->have_state = true;
  return LETKF_OK;
}

static int upload_flags(letkf_gpu_t* h);

// n input observations from host arrays.  With qc and hx_full (the members'
// hx, (k, n)) the mean / perturbation split and the QC compaction of
// letkf_obs_read run on the device first; otherwise hx_pert and hx_mean are
// the reference's obs_hx and obs_hx_mean after letkf_obs_read.
static int set_obs_impl(letkf_gpu_t* h, int64_t n, const float* lat, const float* lon,
                        const float* zdim, const float* time, const float* val,
                        const float* err, const int32_t* obsid, const int32_t* platid,
                        const float* hx_pert, const float* hx_mean, const int32_t* qc,
                        const float* hx_full) {
  if (!h) return LETKF_EINVAL;
  if (n < 0) return fail(h, LETKF_EINVAL, "negative n_obs");
  if (n > 0x7ffffff0) return fail(h, LETKF_EUNSUPPORTED, "n_obs above 2^31");
  const bool members = hx_full != nullptr;
  if (n > 0 && (!lat || !lon || !zdim || !time || !val || !err || !obsid || !platid ||
                (members ? !qc : (!hx_pert || !hx_mean))))
    return fail(h, LETKF_EINVAL, "null observation array");
  h->have_run = false;
  const int64_t k = h->k;
  // raw uploads and work arrays (freed at the end of the call)
  float *r_lat = nullptr, *r_lon = nullptr, *r_zdim = nullptr, *r_time = nullptr;
  float *r_val = nullptr, *r_err = nullptr, *r_hx = nullptr, *r_hxm = nullptr;
  int32_t *r_obsid = nullptr, *r_platid = nullptr, *cell = nullptr, *scell = nullptr;
  int32_t *iota = nullptr, *perm = nullptr, *r_qc = nullptr, *sel = nullptr, *nsel_d = nullptr;
  float* r_full = nullptr;
  uint8_t* keep = nullptr;
  ObsFields cf{};
  void* tmp = nullptr;
  int rc = LETKF_OK;
  auto cleanup = [&]() {
    dfree(h, r_lat); dfree(h, r_lon); dfree(h, r_zdim); dfree(h, r_time); dfree(h, r_val); dfree(h, r_err);
    dfree(h, r_hx); dfree(h, r_hxm); dfree(h, r_obsid); dfree(h, r_platid); dfree(h, cell); dfree(h, scell);
    dfree(h, iota); dfree(h, perm); dfree(h, r_qc); dfree(h, sel); dfree(h, nsel_d); dfree(h, r_full);
    dfree(h, keep);
    for (auto& f : cf.f) dfree(h, f);
    for (auto& f : cf.i) dfree(h, f);
    dfree(h, cf.hx);
    raw_free(h, tmp);
  };
#define OBSCHK(call)                                                           \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      rc = fail(h, e_ == hipErrorOutOfMemory ? LETKF_ENOMEM : LETKF_EHIP,      \
                "%s: %s", #call, hipGetErrorString(e_));                       \
      cleanup();                                                               \
      return rc;                                                               \
    }                                                                          \
  } while (0)
  hipStream_t st = h->stream;
  size_t nn = (size_t)(n > 0 ? n : 1);
  OBSCHK(dalloc(h, &r_lat, nn)); OBSCHK(dalloc(h, &r_lon, nn)); OBSCHK(dalloc(h, &r_zdim, nn));
  OBSCHK(dalloc(h, &r_time, nn)); OBSCHK(dalloc(h, &r_val, nn)); OBSCHK(dalloc(h, &r_err, nn));
  OBSCHK(dalloc(h, &r_hx, nn * k)); OBSCHK(dalloc(h, &r_hxm, nn)); OBSCHK(dalloc(h, &r_obsid, nn));
  OBSCHK(dalloc(h, &r_platid, nn));
  if (n > 0) {
    OBSCHK(hipMemcpyAsync(r_lat, lat, 4 * n, hipMemcpyHostToDevice, st));
    OBSCHK(hipMemcpyAsync(r_lon, lon, 4 * n, hipMemcpyHostToDevice, st));
    OBSCHK(hipMemcpyAsync(r_zdim, zdim, 4 * n, hipMemcpyHostToDevice, st));
    OBSCHK(hipMemcpyAsync(r_time, time, 4 * n, hipMemcpyHostToDevice, st));
    OBSCHK(hipMemcpyAsync(r_val, val, 4 * n, hipMemcpyHostToDevice, st));
    OBSCHK(hipMemcpyAsync(r_err, err, 4 * n, hipMemcpyHostToDevice, st));
    OBSCHK(hipMemcpyAsync(r_obsid, obsid, 4 * n, hipMemcpyHostToDevice, st));
    OBSCHK(hipMemcpyAsync(r_platid, platid, 4 * n, hipMemcpyHostToDevice, st));
    if (!members) {
      OBSCHK(hipMemcpyAsync(r_hx, hx_pert, 4 * n * k, hipMemcpyHostToDevice, st));
      OBSCHK(hipMemcpyAsync(r_hxm, hx_mean, 4 * n, hipMemcpyHostToDevice, st));
    }
  }
  if (members && n > 0) {
    // obs_hx_mean, perturbations and the qc <= 0 flags (src/letkf_obs.F90:459-466)
    OBSCHK(dalloc(h, &r_full, nn * k)); OBSCHK(dalloc(h, &r_qc, nn)); OBSCHK(dalloc(h, &keep, nn));
    OBSCHK(dalloc(h, &sel, nn)); OBSCHK(dalloc(h, &nsel_d, 1));
    OBSCHK(hipMemcpyAsync(r_full, hx_full, 4 * n * k, hipMemcpyHostToDevice, st));
    OBSCHK(hipMemcpyAsync(r_qc, qc, 4 * n, hipMemcpyHostToDevice, st));
    launch_obs_hx_split(st, n, (int)k, r_full, r_qc, r_hx, r_hxm, keep);
    OBSCHK(hipGetLastError());
    // qc > 0 dropped, order kept (:479-489)
    hipcub::CountingInputIterator<int32_t> idx(0);
    size_t tb = 0;
    OBSCHK(hipcub::DeviceSelect::Flagged(nullptr, tb, idx, keep, sel, nsel_d, (int)n, st));
    OBSCHK(raw_alloc(h, &tmp, tb));
    OBSCHK(hipcub::DeviceSelect::Flagged(tmp, tb, idx, keep, sel, nsel_d, (int)n, st));
    int32_t nsel = 0;
    OBSCHK(hipMemcpyAsync(&nsel, nsel_d, 4, hipMemcpyDeviceToHost, st));
    OBSCHK(hipStreamSynchronize(st));
    raw_free(h, tmp);
    tmp = nullptr;
    const size_t ns1 = (size_t)(nsel > 0 ? nsel : 1);
    for (auto& f : cf.f) OBSCHK(dalloc(h, &f, ns1));
    for (auto& f : cf.i) OBSCHK(dalloc(h, &f, ns1));
    OBSCHK(dalloc(h, &cf.hx, ns1 * k));
    const ObsFields in{{r_lat, r_lon, r_zdim, r_time, r_val, r_err, r_hxm},
                       {r_obsid, r_platid}, r_hx};
    launch_obs_compact(st, nsel, (int)k, sel, in, cf);
    OBSCHK(hipGetLastError());
    std::swap(r_lat, cf.f[0]); std::swap(r_lon, cf.f[1]); std::swap(r_zdim, cf.f[2]);
    std::swap(r_time, cf.f[3]); std::swap(r_val, cf.f[4]); std::swap(r_err, cf.f[5]);
    std::swap(r_hxm, cf.f[6]); std::swap(r_obsid, cf.i[0]); std::swap(r_platid, cf.i[1]);
    std::swap(r_hx, cf.hx);
    n = nsel;
    nn = ns1;
  }
  h->nobs = n;
  h->kp = (int)(((k + 1) + 31) / 32 * 32);
  if ((int64_t)nn > h->obs_cap) {
    // grow-only, with a quarter of headroom (kept below the 2^31-byte range
    // of the padded columns), so that later observation sets of similar size
    // keep the device addresses a captured analysis was recorded with
    const int64_t lim = ((1ll << 31) - 1) / (4ll * h->kp);
    const int64_t cap = std::max<int64_t>((int64_t)nn,
                                          std::min<int64_t>((int64_t)nn + (int64_t)nn / 4, lim));
    const size_t c = (size_t)cap;
    OBSCHK(dalloc(h, &h->d_ox, c)); OBSCHK(dalloc(h, &h->d_oy, c)); OBSCHK(dalloc(h, &h->d_oz, c));
    OBSCHK(dalloc(h, &h->d_ozdim, c)); OBSCHK(dalloc(h, &h->d_oerr, c)); OBSCHK(dalloc(h, &h->d_odep, c));
    OBSCHK(dalloc(h, &h->d_otime, c)); OBSCHK(dalloc(h, &h->d_oidx, c)); OBSCHK(dalloc(h, &h->d_oobsid, c));
    OBSCHK(dalloc(h, &h->d_oplatid, c)); OBSCHK(dalloc(h, &h->d_ohx, c * k));
    OBSCHK(dalloc(h, &h->d_osurf, c)); OBSCHK(dalloc(h, &h->d_otwp, c)); OBSCHK(dalloc(h, &h->d_otws, c));
    OBSCHK(dalloc(h, &h->d_ohxp, c * h->kp));
    h->obs_cap = cap;
  }
  OBSCHK(dalloc_once(h, &h->d_cell_start, (size_t)kCells + 1));
  OBSCHK(dalloc(h, &cell, nn)); OBSCHK(dalloc(h, &scell, nn));
  OBSCHK(dalloc(h, &iota, nn)); OBSCHK(dalloc(h, &perm, nn));
  if (n > 0) {
    launch_obs_prep(st, n, (int)k, r_lat, r_lon, cell, iota);
    OBSCHK(hipGetLastError());
    size_t tb = 0;
    OBSCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, cell, scell, iota, perm, (int)n, 0, 17, st));
    OBSCHK(raw_alloc(h, &tmp, tb));
    OBSCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, cell, scell, iota, perm, (int)n, 0, 17, st));
  }
  launch_obs_gather(st, n, (int)k, perm, r_lat, r_lon, r_zdim, r_time, r_val, r_err,
                    r_obsid, r_platid, r_hx, r_hxm, scell, h->d_ox, h->d_oy, h->d_oz,
                    h->d_ozdim, h->d_oerr, h->d_odep, h->d_oidx, h->d_oobsid,
                    h->d_oplatid, h->d_otime, h->d_ohx, h->d_cell_start);
  OBSCHK(hipGetLastError());
  launch_obs_pad_hx(st, n, (int)k, h->kp, h->d_ohx, h->d_odep, h->d_ohxp);
  OBSCHK(hipGetLastError());
  OBSCHK(hipStreamSynchronize(st));
  cleanup();
#undef OBSCHK
  h->have_obs = true;
  if (h->have_loc) return upload_flags(h);
  return LETKF_OK;
}

int letkf_gpu_set_obs(letkf_gpu_t* h, int64_t n, const float* lat, const float* lon,
                      const float* zdim, const float* time, const float* val,
                      const float* err, const int32_t* obsid, const int32_t* platid,
                      const float* hx_pert, const float* hx_mean) {
  return set_obs_impl(h, n, lat, lon, zdim, time, val, err, obsid, platid, hx_pert, hx_mean,
                      nullptr, nullptr);
}

int letkf_gpu_set_obs_members(letkf_gpu_t* h, int64_t n, const float* lat, const float* lon,
                              const float* zdim, const float* time, const float* val,
                              const float* err, const int32_t* obsid, const int32_t* platid,
                              const int32_t* qc, const float* hx_members) {
  if (n > 0 && !hx_members) return h ? fail(h, LETKF_EINVAL, "null hx_members") : LETKF_EINVAL;
  static const float kNone = 0.0f;
  return set_obs_impl(h, n, lat, lon, zdim, time, val, err, obsid, platid, nullptr, nullptr, qc,
                      hx_members ? hx_members : &kNone);
}

static void build_interp(const letkf_linearinterp_lat& c, int* n, float* lat, float* dist) {
  // linearinterp_lat_init, src/letkf_loc.F90:262-293
  int m = 0;
  float l = 0.0f, r = 0.0f;
  for (int i = 0; i < c.npoints; ++i) {
    l = c.lat[i];
    r = c.radius[i];
    if (l > 0.0f && m == 0) { lat[m] = 0.0f; dist[m] = r; ++m; }
    lat[m] = l; dist[m] = r; ++m;
  }
  if (l < 90.0f) { lat[m] = 90.0f; dist[m] = r; ++m; }
  *n = c.npoints > 0 ? m : 0;
}

static int rebuild_slabs(letkf_gpu_t* h) {
  std::vector<int32_t> sv(h->ns, -1), sl(h->ns, 0);
  int maxlev = 1;
  for (int v = 0; v < h->nvars; ++v) {
    maxlev = std::max(maxlev, (int)h->var_lev[v]);
    for (int j = 0; j < h->var_lev[v]; ++j) {
      sv[h->var_s[v] + j] = v;
      sl[h->var_s[v] + j] = j;
    }
  }
  h->loc.max_var_lev = maxlev;
  HIPCHK(h, dalloc(h, &h->d_slab_var, (size_t)h->ns));
  HIPCHK(h, dalloc(h, &h->d_slab_lev, (size_t)h->ns));
  HIPCHK(h, hipMemcpy(h->d_slab_var, sv.data(), 4 * h->ns, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->d_slab_lev, sl.data(), 4 * h->ns, hipMemcpyHostToDevice));
  return LETKF_OK;
}

int letkf_gpu_set_statevars(letkf_gpu_t* h, int32_t nvars, const int32_t* s_idx,
                            const int32_t* levels, const float* inc_max,
                            const float* bounds) {
  if (!h) return LETKF_EINVAL;
  if (nvars < 0 || (nvars > 0 && (!s_idx || !levels || !inc_max || !bounds)))
    return fail(h, LETKF_EINVAL, "invalid statevars");
  for (int v = 0; v < nvars; ++v) {
    if (levels[v] < 1 || s_idx[v] < 0 || s_idx[v] + levels[v] > h->ns)
      return fail(h, LETKF_EINVAL, "variable %d: slabs [%d, %d) outside grid_ns %lld", v,
                  s_idx[v], s_idx[v] + levels[v], (long long)h->ns);
  }
  h->have_run = false;
  ++h->gen;
  h->nvars = nvars;
  h->var_s.assign(s_idx, s_idx + nvars);
  h->var_lev.assign(levels, levels + nvars);
  HIPCHK(h, dalloc(h, &h->d_inc_max, (size_t)std::max(nvars, 1)));
  HIPCHK(h, dalloc(h, &h->d_bounds, (size_t)std::max(2 * nvars, 1)));
  if (nvars > 0) {
    HIPCHK(h, hipMemcpy(h->d_inc_max, inc_max, 4 * nvars, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_bounds, bounds, 8 * nvars, hipMemcpyHostToDevice));
  }
  int rc = rebuild_slabs(h);
  if (rc) return rc;
  h->have_vars = true;
  return LETKF_OK;
}

static int upload_flags(letkf_gpu_t* h) {
  int32_t *so = nullptr, *sp = nullptr;
  HIPCHK(h, dalloc(h, &so, std::max<size_t>(h->sat_obs.size(), 1)));
  HIPCHK(h, dalloc(h, &sp, std::max<size_t>(h->sat_plats.size(), 1)));
  if (!h->sat_obs.empty())
    HIPCHK(h, hipMemcpy(so, h->sat_obs.data(), 4 * h->sat_obs.size(), hipMemcpyHostToDevice));
  if (!h->sat_plats.empty())
    HIPCHK(h, hipMemcpy(sp, h->sat_plats.data(), 4 * h->sat_plats.size(), hipMemcpyHostToDevice));
  launch_obs_flags(h->stream, h->loc, h->nobs, h->d_oobsid, h->d_oplatid, h->d_otime, so,
                   (int)h->sat_obs.size(), sp, (int)h->sat_plats.size(), h->d_osurf,
                   h->d_otwp, h->d_otws);
  hipError_t e = hipGetLastError();
  hipStreamSynchronize(h->stream);
  dfree(h, so);
  dfree(h, sp);
  if (e != hipSuccess) return fail(h, LETKF_EHIP, "obs flags: %s", hipGetErrorString(e));
  return LETKF_OK;
}

int letkf_gpu_set_localizer(letkf_gpu_t* h, const letkf_loc_desc* d) {
  if (!h || !d) return LETKF_EINVAL;
  h->have_run = false;
  ++h->gen;
  LocDev L{};
  L.cls = d->cls;
  switch (d->cls) {
    case LETKF_LOC_NOVRT:
      L.maxgroups = 1;
      break;
    case LETKF_LOC_OCEAN:
      if (d->vtloc_surf != 0 && d->vtloc_surf != 1)
        return fail(h, LETKF_EINVAL, "illegal vtloc_surf type %d", d->vtloc_surf);
      L.maxgroups = d->vtloc_surf == 1 ? 2 : 1;
      break;
    case LETKF_LOC_VGAUSS:
      if (d->nlev < 1 || !d->lev_depth || !(d->vt_scale > 0.0f))
        return fail(h, LETKF_EINVAL, "LOC_VGAUSS needs nlev >= 1, lev_depth and vt_scale > 0");
      L.maxgroups = d->nlev;
      break;
    default:
      return fail(h, LETKF_EUNSUPPORTED, "localizer class %d has no GPU implementation", d->cls);
  }
  if (L.maxgroups > kMaxGroups)
    return fail(h, LETKF_EUNSUPPORTED, "maxgroups %d above %d", L.maxgroups, kMaxGroups);
  auto chk_tab = [&](const letkf_linearinterp_lat& t, const char* name) -> int {
    if (t.npoints < 1 || t.npoints > LETKF_MAX_INTERP)
      return fail(h, LETKF_EINVAL, "%s: linearinterp_lat needs 1..%d points", name, LETKF_MAX_INTERP);
    return LETKF_OK;
  };
  int rc;
  if (d->cls == LETKF_LOC_OCEAN) {
    if ((rc = chk_tab(d->hzloc_prof, "hzloc_prof"))) return rc;
    if ((rc = chk_tab(d->hzloc_sat, "hzloc_sat"))) return rc;
    if (d->vtloc_surf == 1 && (d->bkg_t_var_idx < 0 || d->bkg_t_var_len < 1 ||
                               d->bkg_t_var_idx + d->bkg_t_var_len > h->ns))
      return fail(h, LETKF_EINVAL, "bkg_t_var slabs outside the state");
    if (d->n_sat_obs < 0 || d->n_sat_obs > LETKF_MAX_IDLIST || d->n_sat_plats < 0 ||
        d->n_sat_plats > LETKF_MAX_IDLIST)
      return fail(h, LETKF_EINVAL, "sat_obs / sat_plats list too long");
  } else {
    if ((rc = chk_tab(d->hzloc, "hzloc"))) return rc;
  }
  build_interp(d->hzloc, &L.n_hz, L.hz_lat, L.hz_dist);
  build_interp(d->hzloc_prof, &L.n_prof, L.prof_lat, L.prof_dist);
  build_interp(d->hzloc_sat, &L.n_sat, L.sat_lat, L.sat_dist);
  L.tloc_prof = d->tloc_prof;
  L.tloc_sat = d->tloc_sat;
  L.vtloc_surf = d->vtloc_surf;
  L.bkg_t_delta = d->bkg_t_delta;
  L.bkg_t_var_idx = d->bkg_t_var_idx;
  L.bkg_t_var_len = d->bkg_t_var_len;
  L.vt_scale = d->vt_scale;
  L.nlev = d->cls == LETKF_LOC_VGAUSS ? d->nlev : 0;
  L.max_var_lev = h->loc.max_var_lev > 0 ? h->loc.max_var_lev : 1;
  h->sat_obs.assign(d->sat_obs, d->sat_obs + (d->cls == LETKF_LOC_OCEAN ? d->n_sat_obs : 0));
  h->sat_plats.assign(d->sat_plats, d->sat_plats + (d->cls == LETKF_LOC_OCEAN ? d->n_sat_plats : 0));
  if (d->cls == LETKF_LOC_VGAUSS) {
    HIPCHK(h, dalloc(h, &h->d_lev_depth, (size_t)d->nlev));
    HIPCHK(h, hipMemcpy(h->d_lev_depth, d->lev_depth, 4 * d->nlev, hipMemcpyHostToDevice));
  }
  h->loc = L;
  h->G = L.maxgroups;
  h->have_loc = true;
  if (h->have_vars) { rc = rebuild_slabs(h); if (rc) return rc; }
  if (h->have_obs) return upload_flags(h);
  return LETKF_OK;
}

int letkf_gpu_set_solver(letkf_gpu_t* h, float rtps, float rtpp, float mul) {
  if (!h) return LETKF_EINVAL;
  // src/letkf_solver.F90:114-125
  if (rtps > 1.0f || rtps < 0.0f)
    return fail(h, LETKF_EINVAL, "illegal value for infl_rtps (should be <1.0 and >0.0).");
  if (rtpp > 1.0f || rtpp < 0.0f)
    return fail(h, LETKF_EINVAL, "illegal value for infl_rtpp (should be <1.0 and >0.0).");
  if (rtpp != 0.0f && rtps != 0.0f)
    return fail(h, LETKF_EINVAL, "cannot have both RTPS and RTPP enabled at the same time. ");
  if (mul < 1.0f) return fail(h, LETKF_EINVAL, "infl_mul must be >=1.0.");
  h->have_run = false;
  ++h->gen;
  h->rtps = rtps;
  h->rtpp = rtpp;
  h->mul = mul;
  h->have_solver = true;
  return LETKF_OK;
}

int letkf_gpu_maxgroups(const letkf_gpu_t* h, int32_t* g) {
  if (!h || !g) return LETKF_EINVAL;
  if (!h->have_loc) return LETKF_ESTATE;
  *g = h->G;
  return LETKF_OK;
}

// largest number of slabs one (gridpoint, group) problem can hold
// (slab_group: LOC_VGAUSS groups by level, LOC_OCEAN bkg_t splits the
// variable slabs at the mixed layer, every other case has one group)
static int max_slabs_per_group(const letkf_gpu_t* h) {
  if (h->loc.cls == LETKF_LOC_VGAUSS) {
    std::vector<int> per(h->G, 0);
    int mx = 0;
    for (int64_t s = 0; s < h->ns; ++s) {
      int lev = 0, var = -1;
      for (int v = 0; v < h->nvars; ++v)
        if (s >= h->var_s[v] && s < h->var_s[v] + h->var_lev[v]) { var = v; lev = (int)(s - h->var_s[v]); }
      (void)var;
      if (lev < h->G) mx = std::max(mx, ++per[lev]);
    }
    return mx;
  }
  if (h->loc.cls == LETKF_LOC_OCEAN && h->loc.vtloc_surf == 1) {
    int n = 0;
    for (int v = 0; v < h->nvars; ++v) n += h->var_lev[v];
    return n;
  }
  return (int)h->ns;
}

static hipcub::TransformInputIterator<int64_t, PadCount, const int64_t*> padded_counts(
    const int64_t* cnt) {
  return hipcub::TransformInputIterator<int64_t, PadCount, const int64_t*>(cnt, PadCount());
}

static int ensure_outputs(letkf_gpu_t* h) {
  const int64_t k = h->k, ns = h->ns, nij = h->nij, G = h->G;
  if (!h->d_ana) {
    HIPCHK(h, dalloc(h, &h->d_ana, (size_t)(k * ns * nij)));
    HIPCHK(h, dalloc(h, &h->d_amean, (size_t)(ns * nij)));
    HIPCHK(h, dalloc(h, &h->d_asprd, (size_t)(ns * nij)));
    HIPCHK(h, dalloc(h, &h->d_maxhz, (size_t)nij));
    HIPCHK(h, dalloc(h, &h->d_obscount, (size_t)nij));
    HIPCHK(h, dalloc(h, &h->d_mld, (size_t)nij));
    HIPCHK(h, dalloc(h, &h->d_ctr, (size_t)kCtrSlots * kCtrStride));
  }
  // LETKF_GPU_EIGEN=jacobi selects the parallel Jacobi solver for every
  // problem (reference solver for tests); default: tridiagonal QL pipeline
  const char* ev = getenv("LETKF_GPU_EIGEN");
  const bool force_jacobi = ev && strcmp(ev, "jacobi") == 0;
  // 64 <= k <= 128: explicit transform matrix, applied on MFMA (batched)
  // or in FP64 (plain); LETKF_GPU_MAT=0 keeps the Jacobi solver (k <= 94)
  const bool use_mat = !force_jacobi && h->tune_mat && transform_mat_supported((int)k);
  const int eigK = (force_jacobi || use_mat) ? 0 : eigen_pipeline_K((int)k);
  const int nv = (eigK && h->tune_vec) ? eigen_vec_nv(max_slabs_per_group(h)) : 0;
  if (h->alloc_G == G && h->eigK == eigK && h->alloc_nv == nv && h->alloc_mat == (int)use_mat &&
      h->alloc_chunk == h->tune_chunk_log2)
    return LETKF_OK;
  h->eigK = eigK;
  h->vec_nv = nv;
  h->use_mat = use_mat;
  HIPCHK(h, dalloc(h, &h->d_lgm, (size_t)(G * nij)));
  HIPCHK(h, dalloc(h, &h->d_lgloc, (size_t)(G * nij)));
  HIPCHK(h, dalloc(h, &h->d_lghash, (size_t)(G * nij)));
  // chunk sizing: bound the per-pair work arrays to about 32 GB
  h->use_rat = nv > 0 && (h->batched() ? h->tune_rat != 0 : h->tune_rat == 2) && k >= 2;
  h->cap_rot = h->cap_it = 0;
  if (h->eigK && !h->use_rat) eigen_caps((int)k, &h->cap_rot, &h->cap_it);
  const int64_t per_pair = 8 * k * k + 8 * k + 3 * 8 * (int64_t)(use_mat ? 128 : h->eigK) +
                           (nv ? 8 * (int64_t)nv * h->eigK : 8 * (int64_t)h->eigK * h->eigK) +
                           16 * (int64_t)h->cap_rot + 16 * (int64_t)h->cap_it + 40;
  int64_t pairs = std::max<int64_t>(
      1, std::min<int64_t>(1ll << h->tune_chunk_log2, (32ll << 30) / per_pair));
  int64_t cg = std::max<int64_t>(1, pairs / G);
  cg = std::min<int64_t>(cg, std::max<int64_t>(nij, 1));
  h->chunk_gp = cg;
  const int64_t np = cg * G;
  HIPCHK(h, dalloc(h, &h->d_cnt, (size_t)np));
  HIPCHK(h, dalloc(h, &h->d_off, (size_t)np));
  HIPCHK(h, dalloc(h, &h->d_solve, (size_t)np));
  HIPCHK(h, dalloc(h, &h->d_cnt2, (size_t)np));
  HIPCHK(h, dalloc(h, &h->d_off2, (size_t)np));
  HIPCHK(h, dalloc(h, &h->d_solve2, (size_t)np));
  HIPCHK(h, dalloc(h, &h->d_B, (size_t)(np * k * k)));
  HIPCHK(h, dalloc(h, &h->d_b, (size_t)(np * k)));
  HIPCHK(h, dalloc(h, &h->d_status, (size_t)np));
  if (use_mat) {
    HIPCHK(h, dalloc(h, &h->d_tau, (size_t)(np * 128)));
    HIPCHK(h, dalloc(h, &h->d_dd, (size_t)(np * 128)));
    HIPCHK(h, dalloc(h, &h->d_ee, (size_t)(np * 128)));
  }
  if (h->eigK) {
    HIPCHK(h, dalloc(h, &h->d_tau, (size_t)(np * h->eigK)));
    HIPCHK(h, dalloc(h, &h->d_dd, (size_t)(np * h->eigK)));
    HIPCHK(h, dalloc(h, &h->d_ee, (size_t)(np * h->eigK)));
    HIPCHK(h, dalloc(h, &h->d_rot, (size_t)(np * h->cap_rot)));
    HIPCHK(h, dalloc(h, &h->d_itr, (size_t)(np * h->cap_it)));
    HIPCHK(h, dalloc(h, &h->d_nit, (size_t)np));
    if (nv) {
      dfree(h, h->d_Z);
      HIPCHK(h, dalloc(h, &h->d_U, (size_t)(np * nv * h->eigK)));
      HIPCHK(h, dalloc(h, &h->d_nvec, (size_t)np));
    } else {
      dfree(h, h->d_U);
      HIPCHK(h, dalloc(h, &h->d_Z, (size_t)(np * h->eigK * h->eigK)));
    }
  }
  size_t tb = 0;
  HIPCHK(h, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, padded_counts(h->d_cnt), h->d_off,
                                             (int)np, h->stream));
  HIPCHK(h, dalloc(h, (char**)&h->d_tmp, tb));
  HIPCHK(h, dalloc(h, (char**)&h->d_tmp2, tb));
  h->tmp_bytes = tb;
  if (h->batched()) {
    HIPCHK(h, dalloc(h, &h->d_bkey, (size_t)np));
    HIPCHK(h, dalloc(h, &h->d_bkey2, (size_t)np));
    HIPCHK(h, dalloc(h, &h->d_bval, (size_t)np));
    HIPCHK(h, dalloc(h, &h->d_perm, (size_t)np));
    size_t sb = 0;
    HIPCHK(h, hipcub::DeviceRadixSort::SortPairs(nullptr, sb, h->d_bkey, h->d_bkey2, h->d_bval,
                                                 h->d_perm, (int)np, 0, 32, h->stream));
    HIPCHK(h, dalloc(h, (char**)&h->d_stmp, sb));
    h->stmp_bytes = sb;
  }
  if ((h->use_rat || h->use_mat) && !h->d_rat) {
    // x^(-1/2) on [1, 2^b]: the fewest nodes (<= 31) for 1e-12 relative error,
    // else 31 nodes
    std::vector<double> tab((size_t)kRatBuckets * 64, 0.0);
    std::vector<int> tn(kRatBuckets, 0);
    for (int b = 1; b < kRatBuckets; ++b) {
      double* sb = tab.data() + (size_t)b * 64;
      int n = rat::choose(std::ldexp(1.0, b), 1e-12, 31, sb, sb + 32);
      if (n < 0) { n = 31; rat::nodes(std::ldexp(1.0, b), n, sb, sb + 32); }
      tn[b] = n;
    }
    HIPCHK(h, dalloc(h, &h->d_rat, tab.size()));
    HIPCHK(h, dalloc(h, &h->d_ratn, tn.size()));
    HIPCHK(h, hipMemcpy(h->d_rat, tab.data(), 8 * tab.size(), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_ratn, tn.data(), 4 * tn.size(), hipMemcpyHostToDevice));
  }
  h->alloc_G = G;
  h->alloc_nv = nv;
  h->alloc_mat = (int)use_mat;
  h->alloc_chunk = h->tune_chunk_log2;
  return LETKF_OK;
}

}  // extern "C"

// The checks of every run, before the analysis buffers are touched, and the
// buffers themselves.
static int run_prepare(letkf_gpu_t* h) {
  if (!h->have_state || !h->have_obs || !h->have_loc || !h->have_solver || !h->have_vars)
    return fail(h, LETKF_ESTATE, "set_state, set_obs, set_statevars, set_localizer and "
                                 "set_solver must precede run");
  if (h->loc.cls == LETKF_LOC_OCEAN && h->loc.vtloc_surf == 1 && h->nvars == 0)
    return fail(h, LETKF_EINVAL, "LOC_OCEAN bkg_t needs state variables");
  if (h->loc.cls == LETKF_LOC_VGAUSS && h->loc.max_var_lev > h->loc.nlev)
    return fail(h, LETKF_EINVAL, "LOC_VGAUSS: a variable has %d levels but the localizer "
                "has %d", h->loc.max_var_lev, h->loc.nlev);
  if (h->ns > 2048)
    return fail(h, LETKF_EUNSUPPORTED, "grid_ns %lld above 2048", (long long)h->ns);
  const char* evj = getenv("LETKF_GPU_EIGEN");
  const bool mat = !(evj && strcmp(evj, "jacobi") == 0) && h->tune_mat &&
                   transform_mat_supported((int)h->k);
  if (!mat && eigen_lds_bytes((int)h->k) + 8192 + 64 > 150 * 1024)
    return fail(h, LETKF_EUNSUPPORTED, "ens_size %lld too large for the FP64 eigensolver",
                (long long)h->k);
  if (h->batched() && h->obs_cap * (int64_t)h->kp * 4 >= (1ll << 31))
    return fail(h, LETKF_EUNSUPPORTED, "batched path: n_obs x %d rows above 2^29 values",
                h->kp);
  return ensure_outputs(h);
}

// the device views of the handle's buffers that every kernel takes
struct RunViews {
  StateDev S;
  ObsDev O;
  VarsDev V;
  DiagDev D;
};
static RunViews run_views(const letkf_gpu_t* h) {
  // the observation count the kernels see is the allocated capacity (the
  // padded columns' buffer range); lists only reference the first nobs
  return RunViews{
      StateDev{h->k, h->ns, h->nij, h->d_state, h->d_mean, h->d_sprd, h->d_lat, h->d_lon,
               h->d_mask, h->d_gx, h->d_gy, h->d_gz},
      ObsDev{h->obs_cap, h->d_ox, h->d_oy, h->d_oz, h->d_ozdim, h->d_oerr, h->d_odep,
             h->d_oidx, h->d_osurf, h->d_otwp, h->d_otws, h->d_ohx, h->d_cell_start,
             h->d_ohxp, h->kp},
      VarsDev{h->nvars, h->d_slab_var, h->d_slab_lev, h->d_inc_max, h->d_bounds,
              h->d_lev_depth},
      DiagDev{h->d_maxhz, h->d_obscount, h->d_lgm, h->d_lgloc, h->d_lghash, h->d_mld,
              h->d_ctr}};
}

static FinArgs fin_args(const letkf_gpu_t* h) {
  return FinArgs{h->rtps, h->rtpp, h->mul, h->d_amean, h->d_asprd};
}

// the large-ensemble mixed path post-processes the slabs of solved problems
// inside the apply; k_finalize then does the rest
static bool fin_fused(const letkf_gpu_t* h) {
  return h->use_mat && h->batched() && h->tune_fusefin &&
         transform_mat_fuses_finalize(h->dims.precision);
}

// the transform of one chunk's problems and its post-processing, on st
static hipError_t enqueue_solve(letkf_gpu_t* h, hipStream_t st, const RunViews& W,
                                const ChunkDev& C) {
  if (h->use_mat) {
    const RatTab R{h->d_rat, h->d_ratn, kRatBuckets};
    const FinArgs F = fin_args(h);
    launch_transform_mat(st, h->batched() ? h->dims.precision : LETKF_PREC_FP64, h->loc, W.S,
                         W.V, C, W.D, h->d_B, h->d_b, h->d_tau, h->d_dd, h->d_ee, R, h->d_ana,
                         h->tune_matblk != 0, fin_fused(h) ? &F : nullptr);
  } else if (h->eigK) {
    RatTab R;
    if (h->use_rat) { R.tab = h->d_rat; R.n = h->d_ratn; R.nb = kRatBuckets; }
    launch_eigen_pipeline(st, h->loc, W.S, W.V, C, W.D, h->d_B, h->d_b, h->d_tau, h->d_dd,
                          h->d_ee, h->d_rot, h->cap_rot, h->d_itr, h->cap_it, h->d_nit,
                          h->d_status, h->d_Z, h->vec_nv, h->d_U, h->d_nvec, R, h->d_ana,
                          h->tune_trireg, h->batched() ? h->d_perm : nullptr,
                          h->tune_fuse != 0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // problems the QL could not finish go to the Jacobi solver
    e = launch_eigen_apply_f64(st, h->loc, W.S, W.V, C, W.D, h->d_B, h->d_b, h->d_status,
                               h->d_ana);
    if (e != hipSuccess) return e;
  } else {
    const hipError_t e = launch_eigen_apply_f64(st, h->loc, W.S, W.V, C, W.D, h->d_B, h->d_b,
                                                nullptr, h->d_ana);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

// The search counters of the last analysis (work counts into T), the
// max_obs capacity check and, for a replayed graph, the pairs refused list
// space (*overflow; nullptr: not looked at).
static int read_counters(letkf_gpu_t* h, letkf_gpu_timing* T, int64_t* overflow) {
  const size_t ctr_bytes = sizeof(unsigned long long) * kCtrSlots * kCtrStride;
  HIPCHK(h, hipMemcpyAsync(h->h_ctr, h->d_ctr, ctr_bytes, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  unsigned long long maxc = 0;
  for (int sl = 0; sl < kCtrSlots; ++sl) {
    T->n_problems += (int64_t)h->h_ctr[sl * kCtrStride + kCtrSolves];
    T->n_cand += (int64_t)h->h_ctr[sl * kCtrStride + kCtrCand];
    T->n_pairs += (int64_t)h->h_ctr[sl * kCtrStride + kCtrPairs];
    maxc = std::max(maxc, h->h_ctr[sl * kCtrStride + kCtrMax]);
  }
  if (overflow) *overflow = (int64_t)h->h_ctr[kCtrOverflow];
  if ((long long)maxc > (long long)h->dims.max_obs) {
    return fail(h, LETKF_ECAPACITY, "a gridpoint has %llu observations in its search "
                "radius, above max_obs = %lld", maxc, (long long)h->dims.max_obs);
  }
  return LETKF_OK;
}

static int run_eager(letkf_gpu_t* h) {
  int rc = LETKF_OK;
  hipStream_t st = h->stream;
  const int64_t k = h->k, G = h->G;

  const RunViews W = run_views(h);
  const StateDev& S = W.S;
  const ObsDev& O = W.O;
  const VarsDev& V = W.V;
  const DiagDev& D = W.D;
  const size_t ctr_bytes = sizeof(unsigned long long) * kCtrSlots * kCtrStride;
  HIPCHK(h, hipMemsetAsync(h->d_ctr, 0, ctr_bytes, st));

  // Chunk pipeline.  With overlap, chunk c is searched on stream2 into
  // buffer set c % 2 while chunk c - 1 is ordered, multiplied, transformed
  // and finalized on the main stream; a set is reused only after the main
  // stream has finished the chunk that last used it.
  const bool overlap = h->tune_overlap != 0;
  hipStream_t sst = overlap ? h->stream2 : st;
  std::vector<hipEvent_t> evs;                  // all events, destroyed at the end
  auto mk = [&](hipEvent_t* e) -> hipError_t {
    hipError_t r = hipEventCreate(e);
    if (r == hipSuccess) evs.push_back(*e);
    return r;
  };
  auto cleanup_events = [&]() { for (auto e : evs) hipEventDestroy(e); };
#define RUNCHK(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      hipStreamSynchronize(st);                                                   \
      hipStreamSynchronize(h->stream2);                                           \
      cleanup_events();                                                           \
      return fail(h, e_ == hipErrorOutOfMemory ? LETKF_ENOMEM : LETKF_EHIP,       \
                  "%s: %s", #call, hipGetErrorString(e_));                        \
    }                                                                             \
  } while (0)
  hipEvent_t slot_free[2] = {nullptr, nullptr};
  hipEvent_t e_first = nullptr, e_last = nullptr;
  // per chunk: search start / end (search stream), then order, product,
  // transform, finalize boundaries (main stream)
  std::vector<std::array<hipEvent_t, 7>> cev;
  letkf_gpu_timing T{};
  RUNCHK(mk(&e_first));
  RUNCHK(hipEventRecord(e_first, st));
  if (overlap) RUNCHK(hipStreamWaitEvent(sst, e_first, 0));   // counters cleared first
  int chunk = 0;
  for (int64_t i0 = 0; i0 < h->nij; i0 += h->chunk_gp, ++chunk) {
    const int64_t ngp = std::min(h->chunk_gp, h->nij - i0);
    const int64_t np = ngp * G;
    const int slot = overlap ? (chunk & 1) : 0;
    int64_t* b_cnt = slot ? h->d_cnt2 : h->d_cnt;
    int64_t* b_off = slot ? h->d_off2 : h->d_off;
    uint8_t* b_solve = slot ? h->d_solve2 : h->d_solve;
    int32_t*& b_list_idx = slot ? h->d_list_idx2 : h->d_list_idx;
    float*& b_list_w = slot ? h->d_list_w2 : h->d_list_w;
    int64_t& b_list_cap = slot ? h->list_cap2 : h->list_cap;
    void* b_tmp = slot ? h->d_tmp2 : h->d_tmp;
    int64_t* b_tot = slot ? h->h_tot2 : h->h_tot;
    std::array<hipEvent_t, 7> ce{};
    for (auto& e : ce) RUNCHK(mk(&e));
    cev.push_back(ce);
    if (overlap && slot_free[slot]) RUNCHK(hipStreamWaitEvent(sst, slot_free[slot], 0));
    ChunkDev C{i0, ngp, (int)G, b_cnt, b_off, b_solve, b_list_idx, b_list_w,
               h->batched() ? 1 : 0, h->tune_xcd};
    RUNCHK(hipEventRecord(ce[0], sst));
    launch_search(sst, false, h->loc, O, S, V, C, D);
    RUNCHK(hipGetLastError());
    size_t tb = h->tmp_bytes;
    RUNCHK(hipcub::DeviceScan::ExclusiveSum(b_tmp, tb, padded_counts(b_cnt), b_off, (int)np,
                                            sst));
    RUNCHK(hipMemcpyAsync(b_tot, b_off + (np - 1), 8, hipMemcpyDeviceToHost, sst));
    RUNCHK(hipMemcpyAsync(b_tot + 1, b_cnt + (np - 1), 8, hipMemcpyDeviceToHost, sst));
    RUNCHK(hipMemcpyAsync(h->h_ctr, h->d_ctr, ctr_bytes, hipMemcpyDeviceToHost, sst));
    RUNCHK(hipStreamSynchronize(sst));
    // list space: padded offsets, plus one alignment unit of slack for the
    // vector loads past the last list
    const int64_t total = b_tot[0] + PadCount()(b_tot[1]) + kListAlign;
    // search-set sizes: counted by the count pass with several groups
    // (checked here, before any analysis is written), by the fill pass with
    // one (its count pass only bounds them by the cover; checked once the
    // run's work is done)
    if (G > 1) {
      unsigned long long maxc = 0;
      for (int sl = 0; sl < kCtrSlots; ++sl)
        maxc = std::max(maxc, h->h_ctr[sl * kCtrStride + kCtrMax]);
      if ((long long)maxc > (long long)h->dims.max_obs) {
        hipStreamSynchronize(st);
        hipStreamSynchronize(h->stream2);
        cleanup_events();
        return fail(h, LETKF_ECAPACITY, "a gridpoint has %llu observations in its search "
                    "radius, above max_obs = %lld", maxc, (long long)h->dims.max_obs);
      }
    }
    if (total > b_list_cap) {
      // this set's previous chunk is finished (the search stream waited for it);
      // a quarter of headroom, half for a run that sizes a graph capture
      const int64_t cap = total + total / (h->list_headroom_div) + 1024;
      RUNCHK(dalloc(h, &b_list_idx, (size_t)cap));
      RUNCHK(dalloc(h, &b_list_w, (size_t)cap));
      b_list_cap = cap;
      C.list_idx = b_list_idx;
      C.list_w = b_list_w;
    }
    launch_search(sst, true, h->loc, O, S, V, C, D);
    RUNCHK(hipGetLastError());
    RUNCHK(hipEventRecord(ce[1], sst));
    if (overlap) RUNCHK(hipStreamWaitEvent(st, ce[1], 0));
    RUNCHK(hipEventRecord(ce[2], st));
    if (h->batched()) {
      launch_bucket_keys(st, h->tune_bucket, h->tune_bbits, np, (int)G, i0, b_cnt, b_solve,
                         h->d_lat, h->d_lon, h->d_bkey, h->d_bval);
      size_t sb = h->stmp_bytes;
      RUNCHK(hipcub::DeviceRadixSort::SortPairs(h->d_stmp, sb, h->d_bkey, h->d_bkey2,
                                                h->d_bval, h->d_perm, (int)np, 0, 32, st));
      RUNCHK(hipEventRecord(ce[3], st));
      launch_product_mx(st, h->dims.precision, O, S, C, h->d_perm, h->d_B, h->d_b,
                        h->fold_log2, h->tune_tile);
    } else {
      RUNCHK(hipEventRecord(ce[3], st));
      launch_product_f64(st, O, S, C, h->d_B, h->d_b);
    }
    RUNCHK(hipGetLastError());
    RUNCHK(hipEventRecord(ce[4], st));
    RUNCHK(enqueue_solve(h, st, W, C));
    RUNCHK(hipEventRecord(ce[5], st));
    launch_finalize(st, h->loc, S, V, C, D, fin_args(h), h->d_ana, fin_fused(h));
    RUNCHK(hipGetLastError());
    RUNCHK(hipEventRecord(ce[6], st));
    slot_free[slot] = ce[6];
  }
  RUNCHK(mk(&e_last));
  RUNCHK(hipEventRecord(e_last, st));
  RUNCHK(hipStreamSynchronize(st));
  RUNCHK(hipStreamSynchronize(h->stream2));
  for (const auto& ce : cev) {
    float a = 0, o = 0, p = 0, e = 0, f = 0;
    hipEventElapsedTime(&a, ce[0], ce[1]);
    hipEventElapsedTime(&o, ce[2], ce[3]);
    hipEventElapsedTime(&p, ce[3], ce[4]);
    hipEventElapsedTime(&e, ce[4], ce[5]);
    hipEventElapsedTime(&f, ce[5], ce[6]);
    T.search_ms += a;
    T.order_ms += o;
    T.product_ms += p;
    T.eigen_ms += e;
    T.finalize_ms += f;
    T.n_launch += 1;
  }
  {
    float tt = 0;
    hipEventElapsedTime(&tt, e_first, e_last);
    T.total_ms = tt;
  }
  cleanup_events();
#undef RUNCHK
  rc = read_counters(h, &T, nullptr);
  if (rc) return rc;
  h->timing = T;
  h->have_run = true;
  return LETKF_OK;
}

// The whole analysis as one launch sequence on h->stream with no host
// interaction: counters cleared by a kernel, each chunk's list space
// checked on the device (k_list_guard) against the capacity an earlier
// eager run sized, the search of each chunk, its problem order, products,
// transform and post-processing.  Captured into a graph by run_graph.
static hipError_t enqueue_analysis(letkf_gpu_t* h) {
  hipStream_t st = h->stream;
  const RunViews W = run_views(h);
  const int64_t G = h->G;
  launch_clear_u64(st, (int64_t)kCtrSlots * kCtrStride, h->d_ctr);
  hipError_t e = hipGetLastError();
  for (int64_t i0 = 0; i0 < h->nij && e == hipSuccess; i0 += h->chunk_gp) {
    const int64_t ngp = std::min(h->chunk_gp, h->nij - i0);
    const int64_t np = ngp * G;
    ChunkDev C{i0, ngp, (int)G, h->d_cnt, h->d_off, h->d_solve, h->d_list_idx, h->d_list_w,
               h->batched() ? 1 : 0, h->tune_xcd};
    launch_search(st, false, h->loc, W.O, W.S, W.V, C, W.D);
    e = hipGetLastError();
    if (e != hipSuccess) break;
    size_t tb = h->tmp_bytes;
    e = hipcub::DeviceScan::ExclusiveSum(h->d_tmp, tb, padded_counts(h->d_cnt), h->d_off,
                                         (int)np, st);
    if (e != hipSuccess) break;
    launch_list_guard(st, np, h->d_cnt, h->d_off, h->d_solve, h->list_cap, h->d_ctr);
    launch_search(st, true, h->loc, W.O, W.S, W.V, C, W.D);
    e = hipGetLastError();
    if (e != hipSuccess) break;
    if (h->batched()) {
      launch_bucket_keys(st, h->tune_bucket, h->tune_bbits, np, (int)G, i0, h->d_cnt,
                         h->d_solve, h->d_lat, h->d_lon, h->d_bkey, h->d_bval);
      size_t sb = h->stmp_bytes;
      e = hipcub::DeviceRadixSort::SortPairs(h->d_stmp, sb, h->d_bkey, h->d_bkey2, h->d_bval,
                                             h->d_perm, (int)np, 0, 32, st);
      if (e != hipSuccess) break;
      launch_product_mx(st, h->dims.precision, W.O, W.S, C, h->d_perm, h->d_B, h->d_b,
                        h->fold_log2, h->tune_tile);
    } else {
      launch_product_f64(st, W.O, W.S, C, h->d_B, h->d_b);
    }
    e = hipGetLastError();
    if (e == hipSuccess) e = enqueue_solve(h, st, W, C);
    if (e != hipSuccess) break;
    launch_finalize(st, h->loc, W.S, W.V, C, W.D, fin_args(h), h->d_ana, fin_fused(h));
    e = hipGetLastError();
  }
  return e;
}

// an eager run on the main stream alone: the graph uses the first set of
// chunk buffers only, so their lists are sized by such a run
static int run_eager_single(letkf_gpu_t* h) {
  const int ov = h->tune_overlap;
  h->tune_overlap = 0;
  h->list_headroom_div = 2;
  const int rc = run_eager(h);
  h->tune_overlap = ov;
  h->list_headroom_div = 4;
  return rc;
}

// the launch choices a captured graph was recorded under (besides gen)
static int graph_config(const letkf_gpu_t* h) {
  const char* ev = getenv("LETKF_GPU_EIGEN");
  const int jac = ev && strcmp(ev, "jacobi") == 0;
  return jac | (h->eigK << 1) | ((int)h->use_mat << 9) | ((int)h->use_rat << 10) |
         (h->vec_nv << 11);
}

static void drop_graph(letkf_gpu_t* h) {
  if (h->gexec) hipGraphExecDestroy(h->gexec);
  if (h->graph) hipGraphDestroy(h->graph);
  h->gexec = nullptr;
  h->graph = nullptr;
  h->graph_gen = 0;
}

static int run_graph(letkf_gpu_t* h) {
  hipStream_t st = h->stream;
  if (h->gexec && h->graph_gen == h->gen && h->graph_cfg == graph_config(h)) {
    // replay: the whole analysis in one launch, timed as a whole
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(h, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
      hipEventDestroy(e0);
      return fail(h, LETKF_EHIP, "hipEventCreate failed");
    }
    hipError_t e = hipEventRecord(e0, st);
    if (e == hipSuccess) e = hipGraphLaunch(h->gexec, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    letkf_gpu_timing T{};
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    HIPCHK(h, e);
    T.total_ms = ms;
    T.n_launch = 1;
    int64_t overflow = 0;
    const int rc = read_counters(h, &T, &overflow);
    if (rc) return rc;
    if (overflow > 0) {
      // this analysis needed more list space than the capture had: run it
      // eagerly (which grows the lists, so the next call captures again)
      return run_eager_single(h);
    }
    h->timing = T;
    h->have_run = true;
    return LETKF_OK;
  }
  // first call, or buffers / settings changed: an eager run sizes and
  // allocates everything (and gives this call's analysis), then the launch
  // sequence is captured for the calls that follow
  drop_graph(h);
  int rc = run_eager_single(h);
  if (rc) return rc;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    const hipError_t ee = enqueue_analysis(h);
    e = hipStreamEndCapture(st, &g);
    if (e == hipSuccess) e = ee;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    if (g) hipGraphDestroy(g);
    h->gexec = nullptr;
    h->have_run = false;
    return fail(h, LETKF_EHIP, "capturing the analysis as a graph: %s", hipGetErrorString(e));
  }
  h->graph = g;
  h->graph_gen = h->gen;
  h->graph_cfg = graph_config(h);
  return LETKF_OK;
}

extern "C" {

int letkf_gpu_run(letkf_gpu_t* h) {
  if (!h) return LETKF_EINVAL;
  // whatever happens below, the previous analysis no longer matches the
  // inputs: a run that fails leaves no analysis to fetch
  h->have_run = false;
  const int rc = run_prepare(h);
  if (rc) return rc;
  return run_eager(h);
}

int letkf_gpu_run_graph(letkf_gpu_t* h) {
  if (!h) return LETKF_EINVAL;
  h->have_run = false;
  const int rc = run_prepare(h);
  if (rc) return rc;
  return run_graph(h);
}

int letkf_gpu_get_state_f64(letkf_gpu_t* h, double* state, double* mean, double* sprd) {
  if (!h) return LETKF_EINVAL;
  if (!h->have_run) return fail(h, LETKF_ESTATE, "no analysis: call letkf_gpu_run first");
  const int64_t k = h->k, ns = h->ns, nij = h->nij;
  if (state) HIPCHK(h, hipMemcpy(state, h->d_ana, 8 * k * ns * nij, hipMemcpyDeviceToHost));
  if (mean) HIPCHK(h, hipMemcpy(mean, h->d_amean, 8 * ns * nij, hipMemcpyDeviceToHost));
  if (sprd) HIPCHK(h, hipMemcpy(sprd, h->d_asprd, 8 * ns * nij, hipMemcpyDeviceToHost));
  return LETKF_OK;
}

int letkf_gpu_get_state_members(letkf_gpu_t* h, float* members, float* mean, float* sprd) {
  if (!h) return LETKF_EINVAL;
  if (!h->have_run) return fail(h, LETKF_ESTATE, "no analysis: call letkf_gpu_run first");
  const int64_t k = h->k, ns = h->ns, nij = h->nij;
  if (nij * ns <= 0) return LETKF_OK;
  float *dm = nullptr, *dmu = nullptr, *dsp = nullptr;
  auto done = [&](hipError_t e) {
    dfree(h, dm); dfree(h, dmu); dfree(h, dsp);
    return e;
  };
  hipError_t e = hipSuccess;
  if (members && e == hipSuccess) e = dalloc(h, &dm, (size_t)(k * ns * nij));
  if (mean && e == hipSuccess) e = dalloc(h, &dmu, (size_t)(ns * nij));
  if (sprd && e == hipSuccess) e = dalloc(h, &dsp, (size_t)(ns * nij));
  if (e == hipSuccess) {
    launch_members_from_state(h->stream, nij, ns, (int)k, h->d_ana, h->d_amean, h->d_asprd, dm,
                              dmu, dsp);
    e = hipGetLastError();
  }
  if (e == hipSuccess && members)
    e = hipMemcpyAsync(members, dm, 4 * k * ns * nij, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess && mean)
    e = hipMemcpyAsync(mean, dmu, 4 * ns * nij, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess && sprd)
    e = hipMemcpyAsync(sprd, dsp, 4 * ns * nij, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  HIPCHK(h, done(e));
  return LETKF_OK;
}

int letkf_gpu_get_state_members_noalloc(letkf_gpu_t* h, float* members, float* mean,
                                        float* sprd) {
  if (!h) return LETKF_EINVAL;
  if (!h->have_run) return fail(h, LETKF_ESTATE, "no analysis: call letkf_gpu_run first");
  if (!h->d_scratch)
    return fail(h, LETKF_EHIP, "no-allocation staging array: hipGetSymbolAddress failed (%s)",
                hipGetErrorString(h->scratch_err));
  const int64_t k = h->k, ns = h->ns, nij = h->nij;
  if (nij * ns <= 0) return LETKF_OK;
  // one gridpoint's column (members, mean, spread) must fit the array
  const int64_t col_bytes = 4 * (k * ns + 2 * ns);
  if (col_bytes > (int64_t)kNoAllocScratchBytes)
    return fail(h, LETKF_ECAPACITY, "a gridpoint column of %lld bytes exceeds the %zu-byte "
                "no-allocation staging array", (long long)col_bytes, kNoAllocScratchBytes);
  // gridpoints per slice: as many columns as fit, optionally fewer
  // (LETKF_GPU_NOALLOC_SLICE, gridpoints; used to test ragged slices)
  int64_t slice = (int64_t)kNoAllocScratchBytes / col_bytes;
  if (const char* e = getenv("LETKF_GPU_NOALLOC_SLICE")) {
    const long long v = atoll(e);
    if (v > 0) slice = std::min<int64_t>(slice, v);
  }
  float* sm = (float*)h->d_scratch;   // (c, ns, k) file-order members of the slice
  hipError_t e = hipSuccess;
  for (int64_t i0 = 0; i0 < nij && e == hipSuccess; i0 += slice) {
    const int64_t c = std::min(slice, nij - i0);
    float* smu = sm + c * ns * k;      // (c, ns) mean
    float* ssp = smu + c * ns;         // (c, ns) spread
    // every value of the three staging regions is written before it is
    // copied, so nothing left by an earlier call needs clearing
    launch_members_from_state(h->stream, c, ns, (int)k, h->d_ana + i0 * k * ns,
                              h->d_amean + i0 * ns, h->d_asprd + i0 * ns,
                              members ? sm : nullptr, mean ? smu : nullptr,
                              sprd ? ssp : nullptr);
    e = hipGetLastError();
    // rows of c gridpoints go to column i0 of the caller's (nij, rows) arrays
    if (e == hipSuccess && members)
      e = hipMemcpy2DAsync(members + i0, 4 * nij, sm, 4 * c, 4 * c, (size_t)(ns * k),
                           hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && mean)
      e = hipMemcpy2DAsync(mean + i0, 4 * nij, smu, 4 * c, 4 * c, (size_t)ns,
                           hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && sprd)
      e = hipMemcpy2DAsync(sprd + i0, 4 * nij, ssp, 4 * c, 4 * c, (size_t)ns,
                           hipMemcpyDeviceToHost, h->stream);
    // the next slice overwrites the staging array
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  }
  if (e != hipSuccess)
    return fail(h, LETKF_EHIP, "get_state_members_noalloc: %s", hipGetErrorString(e));
  return LETKF_OK;
}

int letkf_gpu_get_state(letkf_gpu_t* h, float* state, float* mean, float* sprd) {
  if (!h) return LETKF_EINVAL;
  if (!h->have_run) return fail(h, LETKF_ESTATE, "no analysis: call letkf_gpu_run first");
  const int64_t k = h->k, ns = h->ns, nij = h->nij;
  const int64_t n_max = std::max(k * ns * nij, ns * nij);
  if (n_max <= 0) return LETKF_OK;
  // converted on the device, then copied down in slices of at most 16M
  // values through one staging buffer (host memory holds only the caller's
  // FP32 arrays)
  const int64_t slice = std::min<int64_t>(n_max, 1ll << 24);
  float* stage = nullptr;
  HIPCHK(h, dalloc(h, &stage, (size_t)slice));
  hipError_t e = hipSuccess;
  auto conv = [&](const double* src, float* dst, int64_t n) {
    for (int64_t i = 0; i < n && e == hipSuccess; i += slice) {
      const int64_t c = std::min(slice, n - i);
      launch_f64_to_f32(h->stream, c, src + i, stage);
      e = hipGetLastError();
      if (e == hipSuccess)
        e = hipMemcpyAsync(dst + i, stage, 4 * c, hipMemcpyDeviceToHost, h->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    }
  };
  if (state) conv(h->d_ana, state, k * ns * nij);
  if (mean) conv(h->d_amean, mean, ns * nij);
  if (sprd) conv(h->d_asprd, sprd, ns * nij);
  dfree(h, stage);
  HIPCHK(h, e);
  return LETKF_OK;
}

int letkf_gpu_get_diag(letkf_gpu_t* h, const letkf_gpu_diag* d) {
  if (!h || !d) return LETKF_EINVAL;
  if (!h->have_run) return fail(h, LETKF_ESTATE, "no analysis: call letkf_gpu_run first");
  const int64_t nij = h->nij, G = h->G;
  if (d->maxhz) HIPCHK(h, hipMemcpy(d->maxhz, h->d_maxhz, 4 * nij, hipMemcpyDeviceToHost));
  std::vector<int32_t> ib;
  if (d->obscount) {
    ib.resize((size_t)nij);
    HIPCHK(h, hipMemcpy(ib.data(), h->d_obscount, 4 * nij, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < nij; ++i) d->obscount[i] = (float)ib[i];
  }
  // device layout (G, nij) -> reference layout (nij, G)
  if (d->lg_obscount) {
    ib.resize((size_t)(nij * G));
    HIPCHK(h, hipMemcpy(ib.data(), h->d_lgm, 4 * nij * G, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < nij; ++i)
      for (int64_t g = 0; g < G; ++g) d->lg_obscount[i + nij * g] = (float)ib[i * G + g];
  }
  if (d->lg_obsloc) {
    std::vector<double> db((size_t)(nij * G));
    HIPCHK(h, hipMemcpy(db.data(), h->d_lgloc, 8 * nij * G, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < nij; ++i)
      for (int64_t g = 0; g < G; ++g) d->lg_obsloc[i + nij * g] = (float)db[i * G + g];
  }
  if (d->lg_selhash) {
    std::vector<uint64_t> ub((size_t)(nij * G));
    HIPCHK(h, hipMemcpy(ub.data(), h->d_lghash, 8 * nij * G, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < nij; ++i)
      for (int64_t g = 0; g < G; ++g) d->lg_selhash[i + nij * g] = ub[i * G + g];
  }
  return LETKF_OK;
}

int letkf_gpu_get_selection(letkf_gpu_t* h, int64_t ij, int32_t group, int32_t* idx_out,
                            float* rloc_out, int64_t cap, int64_t* m_out) {
  if (!h || !m_out) return LETKF_EINVAL;
  if (!h->have_run) return fail(h, LETKF_ESTATE, "no analysis: call letkf_gpu_run first");
  if (ij < 0 || ij >= h->nij || group < 0 || group >= h->G)
    return fail(h, LETKF_EINVAL, "gridpoint / group out of range");
  // re-run the search of the chunk that holds ij
  const int64_t i0 = ij;
  ChunkDev C{i0, 1, h->G, h->d_cnt, h->d_off, h->d_solve, h->d_list_idx, h->d_list_w, 0, 0};
  StateDev S{h->k, h->ns, h->nij, h->d_state, h->d_mean, h->d_sprd, h->d_lat, h->d_lon,
             h->d_mask, h->d_gx, h->d_gy, h->d_gz};
  ObsDev O{h->nobs, h->d_ox, h->d_oy, h->d_oz, h->d_ozdim, h->d_oerr, h->d_odep,
           h->d_oidx, h->d_osurf, h->d_otwp, h->d_otws, h->d_ohx, h->d_cell_start,
           h->d_ohxp, h->kp};
  VarsDev V{h->nvars, h->d_slab_var, h->d_slab_lev, h->d_inc_max, h->d_bounds, h->d_lev_depth};
  DiagDev D{h->d_maxhz, h->d_obscount, h->d_lgm, h->d_lgloc, h->d_lghash, h->d_mld,
            h->d_ctr};
  hipStream_t st = h->stream;
  launch_search(st, false, h->loc, O, S, V, C, D);
  size_t tb = h->tmp_bytes;
  HIPCHK(h, hipcub::DeviceScan::ExclusiveSum(h->d_tmp, tb, padded_counts(h->d_cnt), h->d_off,
                                             (int)h->G, st));
  HIPCHK(h, hipMemcpyAsync(h->h_tot, h->d_off + group, 8, hipMemcpyDeviceToHost, st));
  HIPCHK(h, hipMemcpyAsync(h->h_tot + 1, h->d_cnt + group, 8, hipMemcpyDeviceToHost, st));
  HIPCHK(h, hipMemcpyAsync(h->h_tot + 2, h->d_off + (h->G - 1), 8, hipMemcpyDeviceToHost, st));
  HIPCHK(h, hipMemcpyAsync(h->h_tot + 3, h->d_cnt + (h->G - 1), 8, hipMemcpyDeviceToHost, st));
  HIPCHK(h, hipStreamSynchronize(st));
  const int64_t off = h->h_tot[0];
  const int64_t total = h->h_tot[2] + PadCount()(h->h_tot[3]) + kListAlign;
  if (total > h->list_cap) {
    const int64_t c2 = total + 1024;
    HIPCHK(h, dalloc(h, &h->d_list_idx, (size_t)c2));
    HIPCHK(h, dalloc(h, &h->d_list_w, (size_t)c2));
    h->list_cap = c2;
    C.list_idx = h->d_list_idx;
    C.list_w = h->d_list_w;
  }
  launch_search(st, true, h->loc, O, S, V, C, D);
  HIPCHK(h, hipGetLastError());
  // kept count (with one group the count pass gives only a bound)
  HIPCHK(h, hipMemcpyAsync(h->h_tot + 1, h->d_cnt + group, 8, hipMemcpyDeviceToHost, st));
  HIPCHK(h, hipStreamSynchronize(st));
  const int64_t m = h->h_tot[1];
  *m_out = m;
  const int64_t nc = std::min(m, cap);
  if (nc > 0) {
    std::vector<int32_t> sidx((size_t)nc);
    HIPCHK(h, hipMemcpy(sidx.data(), h->d_list_idx + off, 4 * nc, hipMemcpyDeviceToHost));
    if (rloc_out) HIPCHK(h, hipMemcpy(rloc_out, h->d_list_w + off, 4 * nc, hipMemcpyDeviceToHost));
    if (idx_out) {
      std::vector<int32_t> oidx((size_t)h->nobs);
      HIPCHK(h, hipMemcpy(oidx.data(), h->d_oidx, 4 * h->nobs, hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < nc; ++i) idx_out[i] = oidx[sidx[i]];
    }
  }
  return LETKF_OK;
}

int letkf_gpu_get_timing(const letkf_gpu_t* h, letkf_gpu_timing* t) {
  if (!h || !t) return LETKF_EINVAL;
  *t = h->timing;
  return LETKF_OK;
}

int letkf_gpu_get_tuning(const letkf_gpu_t* h, letkf_gpu_tuning* out) {
  if (!h || !out) return LETKF_EINVAL;
  *out = letkf_gpu_tuning{h->tune_bucket, h->tune_bbits, h->tune_trireg, 0, 0.0, 0.0,
                          h->tune_overlap, h->tune_chunk_log2, h->tune_xcd, h->tune_tile};
  return LETKF_OK;
}

int letkf_gpu_set_tuning(letkf_gpu_t* h, const letkf_gpu_tuning* t) {
  if (!h) return LETKF_EINVAL;
  if (!t) return fail(h, LETKF_EINVAL, "null tuning");
  if (t->order < 0 || t->order > 3) return fail(h, LETKF_EINVAL, "tuning order %d outside 0..3", t->order);
  if (t->bucket_bits < 0 || t->bucket_bits > 2)
    return fail(h, LETKF_EINVAL, "tuning bucket_bits %d outside 0..2", t->bucket_bits);
  if (t->tridiag_reg < 0 || t->tridiag_reg > 2)
    return fail(h, LETKF_EINVAL, "tuning tridiag_reg %d outside 0..2", t->tridiag_reg);
  if (t->overlap != 0 && t->overlap != 1)
    return fail(h, LETKF_EINVAL, "tuning overlap %d not 0 or 1", t->overlap);
  if (t->chunk_log2 < 17 || t->chunk_log2 > 20)
    return fail(h, LETKF_EINVAL, "tuning chunk_log2 %d outside 17..20", t->chunk_log2);
  if (t->xcd < 0 || t->xcd >= kXcdModes)
    return fail(h, LETKF_EINVAL, "tuning xcd %d outside 0..3", t->xcd);
  if (t->tile != 0 && t->tile != 1)
    return fail(h, LETKF_EINVAL, "tuning tile %d not 0 or 1", t->tile);
  ++h->gen;
  h->tune_bucket = t->order;
  h->tune_bbits = t->bucket_bits;
  h->tune_trireg = t->tridiag_reg;
  h->tune_overlap = t->overlap;
  h->tune_chunk_log2 = t->chunk_log2;
  h->tune_xcd = t->xcd;
  h->tune_tile = t->tile;
  return LETKF_OK;
}

int letkf_gpu_autotune(letkf_gpu_t* h, int32_t repeats, letkf_gpu_tuning* out) {
  if (!h) return LETKF_EINVAL;
  if (repeats < 1 || repeats > 16)
    return fail(h, LETKF_EINVAL, "autotune repeats %d outside 1..16", repeats);
  letkf_gpu_tuning best{};
  int rc = letkf_gpu_get_tuning(h, &best);
  if (rc) return rc;
  rc = letkf_gpu_run(h);   // allocations and list capacity outside the timings
  if (rc) return rc;
  int tried = 0;
  auto time_of = [&](const letkf_gpu_tuning& t, double* ms) -> int {
    int r = letkf_gpu_set_tuning(h, &t);
    if (r) return r;
    double b = 1e300;
    for (int i = 0; i < repeats; ++i) {
      const auto t0 = std::chrono::steady_clock::now();
      r = letkf_gpu_run(h);
      if (r) return r;
      const std::chrono::duration<double, std::milli> dt = std::chrono::steady_clock::now() - t0;
      b = std::min(b, dt.count());
    }
    ++tried;
    *ms = b;
    return LETKF_OK;
  };
  double best_ms = 0.0;
  rc = time_of(best, &best_ms);
  if (rc) return rc;
  const double default_ms = best_ms;
  // a candidate replaces the current choice only if it is at least 0.5%
  // faster (timing noise does not move the choice)
  auto consider = [&](letkf_gpu_tuning t) -> int {
    double ms = 0.0;
    const int r = time_of(t, &ms);
    if (r) return r;
    if (ms < 0.995 * best_ms) { best = t; best_ms = ms; }
    return LETKF_OK;
  };
  // order and bucket edges (the batched product only), then the
  // tridiagonalization kernel (32 <= k < 64 only)
  if (h->batched()) {
    const letkf_gpu_tuning base = best;
    for (int order = 0; order <= 3; ++order)
      for (int bb = 0; bb <= (order >= 2 ? 2 : 0); ++bb) {
        if (order == base.order && bb == base.bucket_bits) continue;
        letkf_gpu_tuning t = base;
        t.order = order;
        t.bucket_bits = bb;
        if ((rc = consider(t))) return rc;
      }
  }
  // the product kernel's tile shape (32x32 blocks or 16x16 tiles, where the
  // 16x16 kernel exists), then its XCD placement (batched path: that kernel
  // reads the observation columns neighbouring problems share)
  if (h->batched() && product_t16_supported(h->dims.precision, (int)h->k)) {
    letkf_gpu_tuning t = best;
    t.tile = 1 - t.tile;
    if ((rc = consider(t))) return rc;
  }
  if (h->batched()) {
    const letkf_gpu_tuning base = best;
    for (int x = 0; x < kXcdModes; ++x) {
      if (x == base.xcd) continue;
      letkf_gpu_tuning t = base;
      t.xcd = x;
      if ((rc = consider(t))) return rc;
    }
  }
  if (h->k >= 32 && h->k < 64 && h->eigK == 64) {
    const letkf_gpu_tuning base = best;
    for (int tr = 0; tr <= (h->k <= 50 ? 2 : 1); ++tr) {
      if (tr == base.tridiag_reg) continue;
      letkf_gpu_tuning t = base;
      t.tridiag_reg = tr;
      if ((rc = consider(t))) return rc;
    }
  }
  // overlap and chunk size matter only when the smallest chunk size would
  // split the analysis; both overlap settings for every chunk size (the
  // overlap only pays with several chunks, so it is not tried alone)
  if (h->nij * h->G > (1ll << 17)) {
    const letkf_gpu_tuning base = best;
    for (int cl = 17; cl <= 20; ++cl)
      for (int ov = 0; ov <= 1; ++ov) {
        if (cl == base.chunk_log2 && ov == base.overlap) continue;
        letkf_gpu_tuning u = base;
        u.chunk_log2 = cl;
        u.overlap = ov;
        if ((rc = consider(u))) return rc;
      }
  }
  // leave the chosen settings and their analysis in the handle
  rc = letkf_gpu_set_tuning(h, &best);
  if (rc) return rc;
  rc = letkf_gpu_run(h);
  if (rc) return rc;
  best.n_tried = tried;
  best.best_ms = best_ms;
  best.default_ms = default_ms;
  if (out) *out = best;
  return LETKF_OK;
}

}  // extern "C"
