// kernels_eigen.hip -- batched FP64 symmetric eigensolve and transform.
//
// For every (gridpoint, group) with kept observations, A = (k-1)I + Yb^T R^-1 Yb
// (k x k, FP64, from the product kernel) is decomposed and the LETKF
// transform is applied straight to the group's slabs
// (src/letkf_core.F90:100-141, src/letkf_solver.F90:284-290):
//
//   x_a(j) = (W x_b)(j) + wbar . x_b,   W = V diag(sqrt((k-1)/lam)) V^T,
//   wbar = V diag(1/lam) V^T (Yb^T R^-1 d)
//
// which equals x_b^T T with T = W + wbar 1^T, W being symmetric.
//
// Three kernels, each shaped for what it does:
//  k_tridiag  one wave per problem: Householder reduction A = Q Tri Q^T
//             (LAPACK dsytd2, lower), A in LDS, lane r owns row r;
//  k_tql      one LANE per problem: implicit QL with Wilkinson shift on the
//             tridiagonal (Numerical Recipes tqli), d and e held in registers
//             through a statically unrolled, predicated sweep; every rotation
//             (c, s) and every sweep header (l, m, count) is recorded;
//  k_replay   one wave per problem: replays the recorded rotations into Z
//             (lane r owns row r, registers; records prefetched two sweeps
//             ahead and broadcast with readlane) and stores Z;
//  k_vec_apply  one wave per problem: back-transforms V = Q Z (reflectors
//             bulk-loaded into LDS) and applies the transform to the group's
//             slabs with two matrix-vector products per slab (the k x k W is
//             never formed).
// Problems whose record space overflows or whose QL does not converge are
// flagged and solved by the Jacobi kernel instead.
#include <hip/hip_runtime.h>

#include "letkf_internal.h"

namespace lk {

constexpr double kEps = 2.220446049250313e-16;

// sum over the K lanes of a K-thread block (K = 32 or 64)
template <int K>
__device__ __forceinline__ double wsumK(double v) {
#pragma unroll
  for (int o = K / 2; o >= 1; o >>= 1) v += __shfl_xor(v, o, K);
  return v;
}

// ---------------------------------------------------------------------------
// Householder tridiagonalization, one K-lane block per problem (k < K <= 64)
// ---------------------------------------------------------------------------
template <int K>
__global__ __launch_bounds__(64) void k_tridiag(int k, ChunkDev C, double* __restrict__ B,
                                                const double* __restrict__ bvec,
                                                double* __restrict__ tau,
                                                double* __restrict__ dd,
                                                double* __restrict__ ee) {
  extern __shared__ double As[];      // k x (k+1), column-major
  __shared__ double sv[K];
  __shared__ double sw[K];
  const int LD = k + 1;
  const int64_t pair = blockIdx.x;
  if (!C.solve[pair]) return;
  const int r = threadIdx.x;
  double* Bp = B + pair * (int64_t)k * k;
  for (int c = 0; c < k; ++c) {
    if (r < k) {
      double v = Bp[r + (int64_t)k * c];
      if (r == c) v += (double)(k - 1);
      As[c * LD + r] = v;
    }
  }
  __syncthreads();
  double* taup = tau + pair * K;
  double* ep = ee + pair * K;
  for (int j = 0; j <= k - 2; ++j) {
    const double x = (r > j && r < k) ? As[j * LD + r] : 0.0;
    const double alpha = As[j * LD + j + 1];
    const double sig = wsumK<K>((r > j + 1 && r < k) ? x * x : 0.0);
    if (sig == 0.0) {
      if (r == 0) { taup[j] = 0.0; ep[j] = alpha; }
      continue;
    }
    const double mu = sqrt(alpha * alpha + sig);
    const double beta = alpha >= 0.0 ? -mu : mu;
    const double t = (beta - alpha) / beta;
    const double scal = 1.0 / (alpha - beta);
    const double v = (r == j + 1) ? 1.0 : ((r > j + 1 && r < k) ? x * scal : 0.0);
    sv[r] = v;
    __syncthreads();
    double p0 = 0.0, p1 = 0.0;
    if (r > j && r < k) {
      int c = j + 1;
      for (; c + 1 < k; c += 2) {
        p0 += As[c * LD + r] * sv[c];
        p1 += As[(c + 1) * LD + r] * sv[c + 1];
      }
      if (c < k) p0 += As[c * LD + r] * sv[c];
    }
    const double p = (r > j && r < k) ? t * (p0 + p1) : 0.0;
    const double kc = 0.5 * t * wsumK<K>(p * v);
    const double w = p - kc * v;
    sw[r] = w;
    __syncthreads();
    if (r > j && r < k) {
      for (int c = j + 1; c < k; ++c) As[c * LD + r] -= v * sw[c] + w * sv[c];
      Bp[(int64_t)j * k + r] = v;              // reflector j, rows j+1..k-1
    }
    if (r == 0) { taup[j] = t; ep[j] = beta; }
    __syncthreads();
  }
  if (r < K) {
    dd[pair * K + r] = r < k ? As[r * LD + r] : 0.0;
    if (r >= k - 1) { ep[r] = 0.0; taup[r] = 0.0; }
  }
  (void)bvec;
}

// ---------------------------------------------------------------------------
// implicit QL, one lane per problem; d, e in registers (static indices)
// ---------------------------------------------------------------------------
template <int K, int KS>
__global__ __launch_bounds__(64) void k_tql(int k, int64_t np, ChunkDev C,
                                            double* __restrict__ dd,
                                            const double* __restrict__ ee,
                                            double2* __restrict__ rot, int cap_rot,
                                            int4* __restrict__ itr, int cap_it,
                                            int* __restrict__ nit_out,
                                            int* __restrict__ status) {
  const int64_t p = (int64_t)blockIdx.x * 64 + threadIdx.x;
  bool active = p < np && C.solve[p];
  double d[K], e[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    d[i] = active ? dd[p * KS + i] : 0.0;
    e[i] = (active && i < k - 1) ? ee[p * KS + i] : 0.0;
  }
  double2* rp = rot + (active ? p : 0) * (int64_t)cap_rot;
  int4* hp = itr + (active ? p : 0) * (int64_t)cap_it;
  int l = 0, iters = 0, t = 0, it = 0, st = 0;
  const int maxit = 30 * k;
  while (__any(active)) {
    if (active) {
      int m = k - 1;
      bool found = false;
#pragma unroll
      for (int i = 0; i < K - 1; ++i) {
        if (!found && i >= l && i < k - 1) {
          const double sd = fabs(d[i]) + fabs(d[i + 1]);
          if (fabs(e[i]) <= kEps * sd) { m = i; found = true; }
        }
      }
      if (m == l) {
        ++l;
        if (l >= k) active = false;
      } else if (++iters > maxit) {
        st = 2;
        active = false;
      } else {
        double dl = 0.0, dl1 = 0.0, el = 0.0, dm = 0.0;
#pragma unroll
        for (int i = 0; i < K; ++i) {
          dl = (i == l) ? d[i] : dl;
          el = (i == l) ? e[i] : el;
          dm = (i == m) ? d[i] : dm;
        }
#pragma unroll
        for (int i = 1; i < K; ++i) dl1 = (i == l + 1) ? d[i] : dl1;
        double g = (dl1 - dl) / (2.0 * el);
        double r = sqrt(g * g + 1.0);
        g = dm - dl + el / (g + (g >= 0.0 ? r : -r));
        double s = 1.0, c = 1.0, pp = 0.0;
        int cnt = 0;
        bool broke = false;
#pragma unroll
        for (int i = K - 2; i >= 0; --i) {
          if (!broke && i >= l && i < m) {
            const double f = s * e[i], b = c * e[i];
            const double r2s = f * f + g * g;
            if (r2s == 0.0) {
              e[i + 1] = 0.0;
              d[i + 1] -= pp;
              broke = true;
            } else {
              // 1/sqrt(f^2+g^2): hardware estimate + two Newton steps
              double y = __builtin_amdgcn_rsq(r2s);
              y = y * (1.5 - 0.5 * r2s * y * y);
              y = y * (1.5 - 0.5 * r2s * y * y);
              e[i + 1] = r2s * y;
              s = f * y;
              c = g * y;
              g = d[i + 1] - pp;
              const double r2 = (d[i] - g) * s + 2.0 * c * b;
              pp = s * r2;
              d[i + 1] = g + pp;
              g = c * r2 - b;
              if (t < cap_rot) rp[t] = make_double2(c, s);
              ++t;
              ++cnt;
            }
          }
        }
#pragma unroll
        for (int i = 0; i < K; ++i) {
          if (!broke && i == l) { d[i] -= pp; e[i] = g; }
          if (i == m) e[i] = 0.0;
        }
        if (it < cap_it) hp[it] = make_int4(l, m, cnt, 0);
        ++it;
        if (t > cap_rot || it > cap_it) { st = 1; active = false; }
      }
    }
  }
  if (p < np && C.solve[p]) {
#pragma unroll
    for (int i = 0; i < K; ++i) dd[p * KS + i] = d[i];    // eigenvalues
    nit_out[p] = it;
    status[p] = st;
  }
}

// ---------------------------------------------------------------------------
// eigenvectors and application
// ---------------------------------------------------------------------------
__device__ __forceinline__ double bcast(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

template <int K>
__global__ __launch_bounds__(64) void k_replay(ChunkDev C, const double2* __restrict__ rot,
                                               int cap_rot, const int4* __restrict__ itr,
                                               int cap_it, const int* __restrict__ nit_in,
                                               const int* __restrict__ status,
                                               double* __restrict__ Zout) {
  __shared__ int2 sh[8 * K];
  const int64_t pair = blockIdx.x;
  if (!C.solve[pair] || status[pair] != 0) return;
  const int lane = threadIdx.x;
  const int nit = nit_in[pair];
  const double2* rp = rot + pair * (int64_t)cap_rot;
  const int4* hp = itr + pair * (int64_t)cap_it;
  for (int q = lane; q < nit; q += K) {
    const int4 h = hp[q];
    sh[q] = make_int2(h.y, h.z);               // (m, count)
  }
  __syncthreads();
  double z[K];
#pragma unroll
  for (int c = 0; c < K; ++c) z[c] = (c == lane) ? 1.0 : 0.0;
  auto fetch = [&](int64_t t0) -> double2 {
    const int64_t q = t0 + lane;
    return q < cap_rot ? rp[q] : make_double2(0.0, 0.0);
  };
  double2 cs0 = fetch(0);
  int64_t t1 = nit > 0 ? sh[0].y : 0;
  double2 cs1 = fetch(t1);
  for (int it = 0; it < nit; ++it) {
    const int m = sh[it].x, cnt = sh[it].y;
    const int64_t t2 = t1 + (it + 1 < nit ? sh[it + 1].y : 0);
    const double2 cs2 = fetch(t2);
    const int lo = m - cnt;
#pragma unroll
    for (int i = K - 2; i >= 0; --i) {
      if (i < m && i >= lo) {
        const int q = m - 1 - i;
        const double c = bcast(cs0.x, q), sn = bcast(cs0.y, q);
        const double f = z[i + 1];
        z[i + 1] = sn * z[i] + c * f;
        z[i] = c * z[i] - sn * f;
      }
    }
    cs0 = cs1;
    cs1 = cs2;
    t1 = t2;
  }
  double* Zp = Zout + pair * (int64_t)K * K;
#pragma unroll
  for (int c = 0; c < K; ++c) Zp[lane * K + c] = z[c];
}

template <int K>
__global__ __launch_bounds__(64) void k_vec_apply(LocDev L, StateDev S, VarsDev V,
                                                  ChunkDev C, DiagDev D,
                                                  const double* __restrict__ B,
                                                  const double* __restrict__ bvec,
                                                  const double* __restrict__ tau,
                                                  const double* __restrict__ lam,
                                                  const double* __restrict__ Zin,
                                                  const int* __restrict__ status,
                                                  double* __restrict__ ana) {
  constexpr int LD = K + 1;
  __shared__ double Hs[K * LD];      // reflectors (zero padded), then rows of V
  __shared__ double st[K], sx[K], sy[K];
  const int64_t pair = blockIdx.x;
  if (!C.solve[pair] || status[pair] != 0) return;
  const int k = (int)S.k;
  const int lane = threadIdx.x;
  const double* Bp = B + pair * (int64_t)k * k;
  // reflector j occupies rows j+1..k-1 of column j (row j+1 is an implicit 1)
  for (int e = lane; e < K * K; e += K) {
    const int j = e / K, r = e - j * K;
    Hs[j * LD + r] = (r == j + 1) ? 1.0 : ((r > j + 1 && r < k && j <= k - 2) ? Bp[j * k + r] : 0.0);
  }
  st[lane] = tau[pair * K + lane];
  const double* Zp = Zin + pair * (int64_t)K * K;
  double zc[K];
#pragma unroll
  for (int r = 0; r < K; ++r) zc[r] = Zp[r * K + lane];
  __syncthreads();
  for (int j = k - 2; j >= 0; --j) {
    const double tj = st[j];
    if (tj == 0.0) continue;
    const double* hv = Hs + j * LD;
    double dot = 0.0;
#pragma unroll
    for (int r = 0; r < K; ++r) dot += hv[r] * zc[r];
    const double sc = tj * dot;
    // re-read the reflector from LDS rather than keeping K values live
    asm volatile("" ::: "memory");
#pragma unroll
    for (int r = 0; r < K; ++r) zc[r] -= sc * hv[r];
  }
  __syncthreads();
  const double lc = lane < k ? lam[pair * K + lane] : 1.0;
  const double fc = lane < k ? sqrt((double)(k - 1) / lc) : 0.0;
  const double ic = lane < k ? 1.0 / lc : 0.0;
#pragma unroll
  for (int r = 0; r < K; ++r) Hs[r * LD + lane] = zc[r];      // rows of V
  sx[lane] = lane < k ? bvec[pair * (int64_t)k + lane] : 0.0;
  __syncthreads();
  double u = 0.0;
#pragma unroll
  for (int r = 0; r < K; ++r) u += zc[r] * sx[r];
  sy[lane] = u * ic;
  __syncthreads();
  double wbar = 0.0;
#pragma unroll 8
  for (int c = 0; c < K; ++c) wbar += Hs[lane * LD + c] * sy[c];
  __syncthreads();

  const int64_t t_gp = pair / C.G;
  const int g = (int)(pair - t_gp * C.G);
  const int64_t ij = C.i0 + t_gp;
  const int mld = D.mld[ij];
  const float* xb0 = S.state + (int64_t)S.k * S.ns * ij;
  double* xa0 = ana + (int64_t)S.k * S.ns * ij;
  for (int s = 0; s < S.ns; ++s) {
    if (slab_group(L, V, s, mld) != g) continue;
    const double xr = lane < k ? (double)xb0[(int64_t)k * s + lane] : 0.0;
    sx[lane] = xr;
    const double wx = wsumK<K>(wbar * xr);
    __syncthreads();
    double y = 0.0;
#pragma unroll
    for (int r = 0; r < K; ++r) y += zc[r] * sx[r];
    sy[lane] = y * fc;
    __syncthreads();
    double o = 0.0;
#pragma unroll 8
    for (int c = 0; c < K; ++c) o += Hs[lane * LD + c] * sy[c];
    if (lane < k) xa0[(int64_t)k * s + lane] = o + wx;
    __syncthreads();
  }
}

int eigen_pipeline_K(int k) { return k <= 31 ? 32 : (k <= 63 ? 64 : 0); }

void eigen_caps(int k, int* cap_rot, int* cap_it) {
  const int K = eigen_pipeline_K(k);
  *cap_rot = 2 * K * K;
  *cap_it = 8 * K;
}

template <int KQ, int K>
static void tql_launch(hipStream_t st, unsigned nl, int k, int64_t np, const ChunkDev& C,
                       double* dd, double* ee, double2* rot, int cap_rot, int4* itr,
                       int cap_it, int* nit, int* status) {
  k_tql<KQ, K><<<nl, 64, 0, st>>>(k, np, C, dd, ee, rot, cap_rot, itr, cap_it, nit, status);
}

template <int K>
static void pipeline(hipStream_t st, const LocDev& L, const StateDev& S, const VarsDev& V,
                     const ChunkDev& C, const DiagDev& D, double* B, const double* bvec,
                     double* tau, double* dd, double* ee, double2* rot, int cap_rot,
                     int4* itr, int cap_it, int* nit, int* status, double* Z,
                     double* ana) {
  const int64_t np = C.ngp * C.G;
  const int k = (int)S.k;
  const unsigned nb = (unsigned)np;
  const unsigned nl = (unsigned)((np + 63) / 64);
  const size_t lds = sizeof(double) * (size_t)k * (k + 1);
  k_tridiag<K><<<nb, K, lds, st>>>(k, C, B, bvec, tau, dd, ee);
  tql_launch<K, K>(st, nl, k, np, C, dd, ee, rot, cap_rot, itr, cap_it, nit, status);
  k_replay<K><<<nb, K, 0, st>>>(C, rot, cap_rot, itr, cap_it, nit, status, Z);
  k_vec_apply<K><<<nb, K, 0, st>>>(L, S, V, C, D, B, bvec, tau, dd, Z, status, ana);
}

void launch_eigen_pipeline(hipStream_t st, const LocDev& L, const StateDev& S,
                           const VarsDev& V, const ChunkDev& C, const DiagDev& D,
                           double* B, const double* bvec, double* tau, double* dd,
                           double* ee, double2* rot, int cap_rot, int4* itr,
                           int cap_it, int* nit, int* status, double* Z, double* ana) {
  if (C.ngp * C.G <= 0) return;
  if (eigen_pipeline_K((int)S.k) == 32)
    pipeline<32>(st, L, S, V, C, D, B, bvec, tau, dd, ee, rot, cap_rot, itr, cap_it, nit,
                 status, Z, ana);
  else
    pipeline<64>(st, L, S, V, C, D, B, bvec, tau, dd, ee, rot, cap_rot, itr, cap_it, nit,
                 status, Z, ana);
}

}  // namespace lk
