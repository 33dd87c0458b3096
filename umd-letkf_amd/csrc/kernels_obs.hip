// kernels_obs.hip -- observation preparation: the device-side replacement of
// the kd-tree build (src/letkf_obs.F90:491-503).
//
// Observations are binned into 1 x 1 degree lat/lon cells, sorted by cell
// (stable radix sort, so the order inside a cell is the original order) and
// stored as structure-of-arrays in that order together with their unit
// vectors and the hx perturbation columns (k floats, contiguous per obs).
// A gridpoint's radius search then reads a handful of contiguous obs ranges.
#include <hip/hip_runtime.h>

#include "det_math.h"
#include "letkf_internal.h"

namespace lk {

__device__ __forceinline__ int cell_of(float lat, float lon) {
  double la = (double)lat + 90.0;
  int b = (int)floor(la);
  b = b < 0 ? 0 : (b >= kBands ? kBands - 1 : b);
  double lo = (double)lon;
  lo = lo - 360.0 * floor(lo / 360.0);
  int c = (int)floor(lo);
  c = c < 0 ? 0 : (c >= kLonCells ? kLonCells - 1 : c);
  return b * kLonCells + c;
}

__global__ void k_obs_cell(int64_t n, const float* __restrict__ lat,
                           const float* __restrict__ lon, int32_t* cell,
                           int32_t* iota) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cell[i] = cell_of(lat[i], lon[i]);
  iota[i] = (int32_t)i;
}

void launch_obs_prep(hipStream_t st, int64_t n, int k, const float* lat,
                     const float* lon, int32_t* cell, int32_t* iota) {
  if (n <= 0) return;
  k_obs_cell<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, lat, lon, cell, iota);
}

// scalar fields of sorted obs s <- original obs perm[s]
__global__ void k_obs_gather_fields(int64_t n, const int32_t* __restrict__ perm,
                                    const float* __restrict__ lat,
                                    const float* __restrict__ lon,
                                    const float* __restrict__ zdim,
                                    const float* __restrict__ time,
                                    const float* __restrict__ val,
                                    const float* __restrict__ err,
                                    const int32_t* __restrict__ obsid,
                                    const int32_t* __restrict__ platid,
                                    const float* __restrict__ hxmean,
                                    double* x, double* y, double* z,
                                    float* ozdim, float* oerr, float* odep,
                                    int32_t* oidx, int32_t* oobsid,
                                    int32_t* oplatid, float* otime) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int32_t o = perm[s];
  double a, b, c;
  dm::xyz(lat[o], lon[o], &a, &b, &c);
  x[s] = a;
  y[s] = b;
  z[s] = c;
  ozdim[s] = zdim[o];
  oerr[s] = err[o];
  {
#pragma clang fp contract(off)
    odep[s] = val[o] - hxmean[o];     // REAL subtraction, letkf_solver.F90:228
  }
  oidx[s] = o;
  oobsid[s] = obsid[o];
  oplatid[s] = platid[o];
  otime[s] = time[o];
}

// hx columns: one block per 64 sorted obs, k floats each
__global__ void k_obs_gather_hx(int64_t n, int k, const int32_t* __restrict__ perm,
                                const float* __restrict__ hx, float* ohx) {
  const int64_t s0 = (int64_t)blockIdx.x * 64;
  for (int e = threadIdx.x; e < 64 * k; e += blockDim.x) {
    const int64_t s = s0 + e / k;
    const int a = e % k;
    if (s < n) ohx[s * k + a] = hx[(int64_t)perm[s] * k + a];
  }
}

// operand columns of the mixed-precision products: hxp(kp, n) in search
// order, rows 0..k-1 the hx perturbations, row k the departure, rows above
// zero, so a column can be read without row tests
__global__ void k_obs_pad_hx(int64_t n, int k, int kp, const float* __restrict__ ohx,
                             const float* __restrict__ odep, float* hxp) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * kp) return;
  const int64_t s = e / kp;
  const int a = (int)(e - s * kp);
  hxp[e] = a < k ? ohx[s * k + a] : (a == k ? odep[s] : 0.0f);
}

void launch_obs_pad_hx(hipStream_t st, int64_t n, int k, int kp, const float* ohx,
                       const float* odep, float* hxp) {
  if (n <= 0) return;
  const int64_t tot = n * kp;
  k_obs_pad_hx<<<(unsigned)((tot + 255) / 256), 256, 0, st>>>(n, k, kp, ohx, odep, hxp);
}

// cell_start[c] = first sorted obs with cell >= c (c = 0..kCells)
__global__ void k_cell_start(int64_t n, const int32_t* __restrict__ scell,
                             int32_t* cell_start) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > kCells) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (scell[mid] < c) lo = mid + 1; else hi = mid;
  }
  cell_start[c] = (int32_t)lo;
}

void launch_obs_gather(hipStream_t st, int64_t n, int k, const int32_t* perm,
                       const float* lat, const float* lon, const float* zdim,
                       const float* time, const float* val, const float* err,
                       const int32_t* obsid, const int32_t* platid,
                       const float* hx, const float* hxmean,
                       const int32_t* sorted_cell, double* x, double* y,
                       double* z, float* ozdim, float* oerr, float* odep,
                       int32_t* oidx, int32_t* oobsid, int32_t* oplatid,
                       float* otime, float* ohx, int32_t* cell_start) {
  if (n > 0) {
    unsigned nb = (unsigned)((n + 255) / 256);
    k_obs_gather_fields<<<nb, 256, 0, st>>>(n, perm, lat, lon, zdim, time, val,
                                            err, obsid, platid, hxmean, x, y, z,
                                            ozdim, oerr, odep, oidx, oobsid,
                                            oplatid, otime);
    k_obs_gather_hx<<<(unsigned)((n + 63) / 64), 256, 0, st>>>(n, k, perm, hx, ohx);
  }
  k_cell_start<<<(kCells + 1 + 255) / 256, 256, 0, st>>>(n, sorted_cell, cell_start);
}

// localizer factors that depend on the observation only:
// LOC_OCEAN surface flag (letkf_loc_ocean.F90:467-479) and the temporal
// GC factors (:491-499); 1 when the temporal localization is off.
__global__ void k_obs_flags(LocDev L, int64_t n, const int32_t* __restrict__ obsid,
                            const int32_t* __restrict__ platid,
                            const float* __restrict__ time,
                            const int32_t* __restrict__ sat_obs, int n_sat_obs,
                            const int32_t* __restrict__ sat_plats, int n_sat_plats,
                            uint8_t* surf, float* tw_prof, float* tw_sat) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int sf = 1;
  if (n_sat_obs > 0) {
    int hit = 0;
    for (int i = 0; i < n_sat_obs; ++i) hit |= (obsid[s] == sat_obs[i]);
    if (!hit) sf = 0;
  }
  if (n_sat_plats > 0) {
    int hit = 0;
    for (int i = 0; i < n_sat_plats; ++i) hit |= (platid[s] == sat_plats[i]);
    if (!hit) sf = 0;
  }
  surf[s] = (uint8_t)sf;
  tw_prof[s] = L.tloc_prof > 0.0f ? dm::gc(time[s], L.tloc_prof) : 1.0f;
  tw_sat[s] = L.tloc_sat > 0.0f ? dm::gc(time[s], L.tloc_sat) : 1.0f;
}

void launch_obs_flags(hipStream_t st, const LocDev& L, int64_t n,
                      const int32_t* obsid, const int32_t* platid,
                      const float* time, const int32_t* sat_obs, int n_sat_obs,
                      const int32_t* sat_plats, int n_sat_plats, uint8_t* surf,
                      float* tw_prof, float* tw_sat) {
  if (n <= 0) return;
  k_obs_flags<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(
      L, n, obsid, platid, time, sat_obs, n_sat_obs, sat_plats, n_sat_plats,
      surf, tw_prof, tw_sat);
}

__global__ void k_grid_xyz(int64_t n, const float* __restrict__ lat,
                           const float* __restrict__ lon, double* x, double* y,
                           double* z) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a, b, c;
  dm::xyz(lat[i], lon[i], &a, &b, &c);
  x[i] = a;
  y[i] = b;
  z[i] = c;
}

void launch_grid_xyz(hipStream_t st, int64_t n, const float* lat,
                     const float* lon, double* x, double* y, double* z) {
  if (n <= 0) return;
  k_grid_xyz<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, lat, lon, x, y, z);
}

}  // namespace lk
