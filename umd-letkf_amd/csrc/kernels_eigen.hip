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
// Kernels, each shaped for what it does:
//  k_tridiag  one wave per problem: Householder reduction A = Q Tri Q^T
//             (LAPACK dsytd2, lower), A in LDS, lane r owns row r;
//  k_tql      one LANE per problem: implicit QL with Wilkinson shift on the
//             tridiagonal (Numerical Recipes tqli), d and e held in registers
//             through a statically unrolled, predicated sweep; every rotation
//             (c, s) and every sweep header (l, m, count) is recorded;
// then, when a group has many slabs (full path):
//  k_replay   one wave per problem: replays the recorded rotations into Z
//             (lane r owns row r, registers; records prefetched two sweeps
//             ahead and broadcast with readlane) and stores Z;
//  k_vec_apply  one wave per problem: back-transforms V = Q Z (reflectors
//             bulk-loaded into LDS) and applies the transform to the group's
//             slabs with two matrix-vector products per slab (the k x k W is
//             never formed);
// or, when every group has few slabs (vector path), only the vectors that
// the transform touches are carried, V^T x = Z^T Q^T x:
//  k_tridiag  also applies each reflector to b and to the slabs' x_b;
//  k_rot_fwd  one lane per (problem, vector): Z^T u from the records;
//  k_rot_bwd  one lane per (problem, slab): y = f(lam) u, Z y, Q (Z y),
//             plus wbar.x_b = (V^T b)^T lam^-1 (V^T x_b).
// Problems whose record space overflows or whose QL does not converge are
// flagged and solved by the Jacobi kernel instead.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "eigen_common.h"
#include "letkf_internal.h"

namespace lk {

constexpr double kEps = 2.220446049250313e-16;



// ---------------------------------------------------------------------------
// Householder tridiagonalization, one K-lane block per problem (k < K <= 64).
// NV > 0 (vector path): the NV vectors b = Yb^T R^-1 d and x_b of each slab
// of the group (at most NV - 1 slabs) are carried along, u = Q^T x, by
// applying every reflector as it is formed; they are written to U with the
// slab count in nvec.
// ---------------------------------------------------------------------------
template <int K, int NV>
__global__ __launch_bounds__(64) void k_tridiag(int k, LocDev L, StateDev S, VarsDev V,
                                                ChunkDev C, DiagDev D, double* __restrict__ B,
                                                const double* __restrict__ bvec,
                                                double* __restrict__ tau,
                                                double* __restrict__ dd,
                                                double* __restrict__ ee,
                                                double* __restrict__ U,
                                                int* __restrict__ nvec,
                                                int* __restrict__ status) {
  extern __shared__ double As[];      // k x (k+1), column-major
  __shared__ double sv[K];
  __shared__ double sw[K];
  __shared__ int sl[NV > 0 ? NV : 1];
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
  double xv[NV > 0 ? NV : 1];
  int nsl = 0;
  if (NV > 0) {
    // slabs of this group in ascending order (ballot compaction)
    const int64_t t_gp = pair / C.G;
    const int g = (int)(pair - t_gp * C.G);
    const int64_t ij = C.i0 + t_gp;
    const int mld = D.mld[ij];
    for (int s0 = 0; s0 < S.ns; s0 += K) {
      const int s = s0 + r;
      const bool in = s < S.ns && slab_group(L, V, s, mld) == g;
      const unsigned long long bal = __ballot(in);
      const unsigned long long below = r == 0 ? 0ull : (bal & (~0ull >> (64 - r)));
      const int q = nsl + __popcll(below);
      if (in && q < NV - 1) sl[q] = s;
      nsl += __popcll(bal);
    }
    __syncthreads();
    xv[0] = r < k ? bvec[pair * (int64_t)k + r] : 0.0;
    const float* xb0 = S.state + (int64_t)S.k * S.ns * ij;
#pragma unroll
    for (int q = 1; q < NV; ++q)
      xv[q] = (q <= nsl && r < k) ? (double)xb0[(int64_t)k * sl[q - 1] + r] : 0.0;
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
    if (NV > 0) {
      // x <- H_j x for the carried vectors (H_j = I - t v v^T)
      double dq[NV > 0 ? NV : 1];
#pragma unroll
      for (int q = 0; q < NV; ++q) dq[q] = v * xv[q];
#pragma unroll
      for (int o = K / 2; o >= 1; o >>= 1)
#pragma unroll
        for (int q = 0; q < NV; ++q) dq[q] += __shfl_xor(dq[q], o, K);
#pragma unroll
      for (int q = 0; q < NV; ++q) xv[q] -= (t * dq[q]) * v;
    }
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
  if (NV > 0) {
    double* Up = U + pair * (int64_t)NV * K;
#pragma unroll
    for (int q = 0; q < NV; ++q) Up[q * K + r] = xv[q];
    if (r == 0) {
      nvec[pair] = nsl;
      status[pair] = nsl > NV - 1 ? 3 : 0;     // more slabs than carried: Jacobi
    }
  } else if (r == 0) {
    status[pair] = 0;
  }
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
    if (st) status[p] = st;
  }
}

// ---------------------------------------------------------------------------
// eigenvectors and application
// ---------------------------------------------------------------------------
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

// ---------------------------------------------------------------------------
// vector path: the QL rotations applied to the carried vectors only.
// Z = G_0 G_1 ... G_{T-1} (G_t the rotation of record t, acting on rows
// (i, i+1) of a row vector as z <- z G_t), so
//   Z^T u:  u^T <- u^T G_t for t = 0 .. T-1   (same update as the Z replay)
//   Z y:    y <- G_t y     for t = T-1 .. 0
// One lane per (problem, vector); u in registers, sweeps statically unrolled
// and predicated like the QL itself.
// ---------------------------------------------------------------------------
// record of the rotation at row i of a sweep (rows m-1 down to m-cnt, in
// record order); clamped inside the problem's record space
__device__ __forceinline__ double2 sweep_rot(const double2* rp, int64_t t0, int m, int cnt,
                                             int i, int cap_rot) {
  int q = m - 1 - i;
  q = q < 0 ? 0 : (q >= cnt ? cnt - 1 : q);
  int64_t a = t0 + (q < 0 ? 0 : q);
  a = a < cap_rot ? a : cap_rot - 1;
  return rp[a];
}

template <int K>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_rot_fwd(int k, int NV, int64_t np, ChunkDev C,
                                                const double2* __restrict__ rot, int cap_rot,
                                                const int4* __restrict__ itr, int cap_it,
                                                const int* __restrict__ nit_in,
                                                const int* __restrict__ nvec,
                                                const int* __restrict__ status,
                                                double* __restrict__ U) {
  const int64_t gid = (int64_t)blockIdx.x * 64 + threadIdx.x;
  const int64_t pair = gid / NV;
  const int q = (int)(gid - pair * NV);
  if (pair >= np || !C.solve[pair] || status[pair] != 0 || q > nvec[pair]) return;
  double* Up = U + (pair * NV + q) * K;
  double u[K];
#pragma unroll
  for (int i = 0; i < K; ++i) u[i] = Up[i];
  const double2* rp = rot + pair * (int64_t)cap_rot;
  const int4* hp = itr + pair * (int64_t)cap_it;
  const int nit = nit_in[pair];
  int64_t t0 = 0;
  for (int it = 0; it < nit; ++it) {
    const int4 h = hp[it];
    const int m = h.y, cnt = h.z, lo = m - cnt;
#pragma unroll
    for (int i = K - 2; i >= 0; --i) {
      const double2 cs = sweep_rot(rp, t0, m, cnt, i, cap_rot);
      const bool act = i < m && i >= lo;
      const double f = u[i + 1], e = u[i];
      const double n1 = cs.y * e + cs.x * f;
      const double n0 = cs.x * e - cs.y * f;
      u[i + 1] = act ? n1 : f;
      u[i] = act ? n0 : e;
    }
    t0 += cnt;
  }
#pragma unroll
  for (int i = 0; i < K; ++i) Up[i] = u[i];
  (void)k;
}

// y = sqrt((k-1)/lam) u_slab, wbar.x_b = u_b . (u_slab / lam); then Z y and
// Q (Z y) (reflectors from B, tau); x_a = W x_b + wbar.x_b
template <int K>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_rot_bwd(LocDev L, StateDev S, VarsDev V, ChunkDev C,
                                                DiagDev D, int NV, int64_t np,
                                                const double* __restrict__ B,
                                                const double* __restrict__ tau,
                                                const double* __restrict__ lam,
                                                const double2* __restrict__ rot, int cap_rot,
                                                const int4* __restrict__ itr, int cap_it,
                                                const int* __restrict__ nit_in,
                                                const int* __restrict__ nvec,
                                                const int* __restrict__ status,
                                                double* __restrict__ U,
                                                double* __restrict__ ana) {
  const int NS = NV - 1;
  const int64_t gid = (int64_t)blockIdx.x * 64 + threadIdx.x;
  const int64_t pair = gid / NS;
  const int q = 1 + (int)(gid - pair * NS);
  if (pair >= np || !C.solve[pair] || status[pair] != 0 || q > nvec[pair]) return;
  const int k = (int)S.k;
  double* u1 = U + (pair * NV + q) * K;
  const double* u2 = U + pair * NV * K;
  const double* lp = lam + pair * K;
  // scale in place (rolled loop: the FP64 divide and square root expand to
  // long sequences), then load y into registers
  double wx = 0.0;
#pragma unroll 2
  for (int i = 0; i < k; ++i) {
    const double a = u1[i], l = lp[i];
    wx += u2[i] * (a / l);
    u1[i] = sqrt((double)(k - 1) / l) * a;
  }
  double y[K];
#pragma unroll
  for (int i = 0; i < K; ++i) y[i] = i < k ? u1[i] : 0.0;
  const double2* rp = rot + pair * (int64_t)cap_rot;
  const int4* hp = itr + pair * (int64_t)cap_it;
  const int nit = nit_in[pair];
  int64_t t0 = 0;
  for (int it = 0; it < nit; ++it) t0 += hp[it].z;
  for (int it = nit - 1; it >= 0; --it) {
    const int4 h = hp[it];
    const int m = h.y, cnt = h.z, lo = m - cnt;
    t0 -= cnt;
#pragma unroll
    for (int i = 0; i <= K - 2; ++i) {
      const double2 cs = sweep_rot(rp, t0, m, cnt, i, cap_rot);
      const bool act = i < m && i >= lo;
      const double e = y[i], f = y[i + 1];
      const double n0 = cs.x * e + cs.y * f;
      const double n1 = cs.x * f - cs.y * e;
      y[i] = act ? n0 : e;
      y[i + 1] = act ? n1 : f;
    }
  }
  // Q y = H_0 H_1 ... H_{k-2} y; reflector j: 1 at row j+1, B column j below
  const double* Bp = B + pair * (int64_t)k * k;
  const double* tp = tau + pair * K;
  for (int j = k - 2; j >= 0; --j) {
    const double tj = tp[j];
    if (tj == 0.0) continue;
    const double* hv = Bp + (int64_t)j * k;
    double dot = 0.0;
#pragma unroll
    for (int r = 1; r < K; ++r)
      if (r > j && r < k) dot += hv[r] * y[r];
    const double sc = tj * dot;
    // re-read the reflector rather than keeping K values live
    asm volatile("" ::: "memory");
#pragma unroll
    for (int r = 1; r < K; ++r)
      if (r > j && r < k) y[r] -= sc * hv[r];
  }
  const int64_t t_gp = pair / C.G;
  const int g = (int)(pair - t_gp * C.G);
  const int64_t ij = C.i0 + t_gp;
  const int mld = D.mld[ij];
  int s = -1;
  for (int c = 0, n = 0; c < S.ns; ++c)
    if (slab_group(L, V, c, mld) == g && ++n == q) { s = c; break; }
  if (s < 0) return;
  double* xa = ana + (int64_t)S.k * S.ns * ij + (int64_t)k * s;
#pragma unroll
  for (int r = 0; r < K; ++r)
    if (r < k) xa[r] = y[r] + wx;
}

// ---------------------------------------------------------------------------
// batched (mixed-precision) vector path without the eigenvectors:
//   W x_b = sqrt(k-1) A^(-1/2) x_b = Q [sqrt(k-1) T^(-1/2)] Q^T x_b
//   wbar . x_b = (Q^T x_b) . T^-1 (Q^T b)
// with T^(-1/2) u ~= sum_j sqrt(m) w_j (T + m s_j I)^-1 u, m = k-1 (a lower
// bound of the spectrum, A = (k-1)I + PSD) and the node table chosen by the
// Gershgorin bound of the largest eigenvalue.  Each (T + sigma I) is an SPD
// tridiagonal system solved by LDL^T in O(k).
// ---------------------------------------------------------------------------
constexpr int kRatLanes = 32;        // lanes per problem: 31 nodes + the T^-1 b lane

// 1/x: the hardware estimate and one Newton step, which squares its
// relative error (the pivots are >= k - 1 > 0 and the batched path's
// tolerance is far above the result)
__device__ __forceinline__ double rcp_nr(double x) {
  const double y = __builtin_amdgcn_rcp(x);
  return y * (2.0 - x * y);
}

// one wave = two problems x 32 lanes; lane j < 31 of a problem solves with
// node j, lane 31 solves T x = Q^T b.  U holds Q^T b (vector 0) and Q^T x_b
// (vectors 1..nvec); on exit vector q holds y = sqrt(k-1) T^(-1/2) u_q and its
// last slot (index K-1 >= k) holds wbar . x_b.
template <int KQ>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
void k_ratsolve(int k, int K, int NV, int64_t np, ChunkDev C, const double* __restrict__ dd,
                const double* __restrict__ ee, const int* __restrict__ nvec,
                const int* __restrict__ status, const double* __restrict__ tab,
                const int* __restrict__ tabn, int nbucket, double* __restrict__ U) {
  const int lane = threadIdx.x;
  const int j = lane & (kRatLanes - 1);
  const int64_t pair_raw = (int64_t)blockIdx.x * 2 + (lane >> 5);
  const bool valid = pair_raw < np && C.solve[pair_raw] && status[pair_raw] == 0;
  const int64_t pair = valid ? pair_raw : 0;
  const double* dp = dd + pair * K;
  const double* ep = ee + pair * K;
  // Gershgorin bound of the largest eigenvalue of T (SPD)
  double gm = 0.0;
  for (int i = j; i < k; i += kRatLanes) {
    const double em = i > 0 ? fabs(ep[i - 1]) : 0.0;
    gm = fmax(gm, dp[i] + em + fabs(ep[i]));
  }
#pragma unroll
  for (int o = kRatLanes / 2; o >= 1; o >>= 1) gm = fmax(gm, __shfl_xor(gm, o, kRatLanes));
  const double m = (double)(k - 1);
  const double kappa = gm / m;
  int b = (int)ceil(log2(fmax(kappa, 2.0)));
  b = b < 1 ? 1 : (b > nbucket - 1 ? nbucket - 1 : b);
  const int nn = tabn[b];
  const bool node = j < kRatLanes - 1;
  const double sigma = node ? (j < nn ? m * tab[b * 64 + j] : m) : 0.0;
  const double coef = (node && j < nn) ? m * tab[b * 64 + 32 + j] : 0.0;
  const int nsl = valid ? nvec[pair] : 0;
  const int nsl_all = max(nsl, __shfl_xor(nsl, 32, 64));
  for (int q = 1; q <= nsl_all; ++q) {
    const bool act = valid && q <= nsl;
    const double* ux = U + (pair * NV + (act ? q : 0)) * K;
    const double* u = node ? ux : U + pair * NV * K;
    // LDL^T of T + sigma I with the forward substitution, then back
    double l[KQ], p[KQ];
    double rprev = 0.0, zprev = 0.0, eprev = 0.0;
#pragma unroll
    for (int i = 0; i < KQ; ++i) {
      if (i < k) {
        const double li = eprev * rprev;
        const double Di = dp[i] + sigma - li * eprev;
        const double ri = rcp_nr(Di);
        const double zi = u[i] - li * zprev;
        l[i] = li;
        p[i] = zi * ri;
        rprev = ri;
        zprev = zi;
        eprev = ep[i];
      } else {
        l[i] = 0.0;
        p[i] = 0.0;
      }
    }
#pragma unroll
    for (int i = KQ - 2; i >= 0; --i)
      if (i + 1 < k) p[i] -= l[i + 1] * p[i + 1];
    // lane 31: wbar . x_b = u_q . T^-1 u_b
    double wx = 0.0;
    if (!node) {
#pragma unroll
      for (int i = 0; i < KQ; ++i)
        if (i < k) wx += ux[i] * p[i];
    }
    // y_i = sum over node lanes of coef * x_i, gathered by DPP in the last
    // lane of each half (the T^-1 b lane, whose coef is 0); rows past k
    // stay unused
#pragma unroll
    for (int i = 0; i < KQ; ++i)
      if (i < k) p[i] = half_sum_dpp(coef * p[i]);
    if (act && !node) {
      double* uo = U + (pair * NV + q) * K;
#pragma unroll
      for (int i = 0; i < KQ; ++i)
        if (i < k) uo[i] = p[i];
      uo[K - 1] = wx;
    }
  }
}

// x_a = Q y + wbar . x_b for every slab of a problem; one wave per problem,
// lane r owns row r
template <int K>
__global__ __launch_bounds__(64) void k_qapply(LocDev L, StateDev S, VarsDev V, ChunkDev C,
                                               DiagDev D, int NV, const double* __restrict__ B,
                                               const double* __restrict__ tau,
                                               const int* __restrict__ nvec,
                                               const int* __restrict__ status,
                                               const double* __restrict__ U,
                                               double* __restrict__ ana) {
  const int64_t pair = blockIdx.x;
  if (!C.solve[pair] || status[pair] != 0) return;
  const int k = (int)S.k;
  const int r = threadIdx.x;
  const double* Bp = B + pair * (int64_t)k * k;
  const double* tp = tau + pair * K;
  const int64_t t_gp = pair / C.G;
  const int g = (int)(pair - t_gp * C.G);
  const int64_t ij = C.i0 + t_gp;
  const int mld = D.mld[ij];
  const int nsl = nvec[pair];
  // every reflector's row r and tau (lane jj holds tau_jj) are loaded once,
  // up front, so the chain of wave sums below never waits on memory; the
  // step loop is unrolled so that the reflectors stay in registers
  double hv[K - 1];
#pragma unroll
  for (int c = 0; c < K - 1; ++c)
    hv[c] = (c <= k - 2 && r > c && r < k) ? Bp[(int64_t)c * k + r] : 0.0;
  const double tv = tp[r];
  int s = -1;
  for (int q = 1; q <= nsl; ++q) {
    for (++s; s < S.ns && slab_group(L, V, s, mld) != g; ++s) {
    }
    if (s >= S.ns) return;
    const double* up = U + (pair * NV + q) * K;
    double y = r < k ? up[r] : 0.0;
    const double wx = up[K - 1];
#pragma unroll
    for (int jj = K - 2; jj >= 0; --jj) {
      if (jj > k - 2) continue;
      const double tj = bcast(tv, jj);
      if (tj == 0.0) continue;
      const double dot = K == 64 ? wave_sum_dpp(hv[jj] * y) : wsumK<K>(hv[jj] * y);
      y -= (tj * dot) * hv[jj];
    }
    if (r < k) ana[(int64_t)S.k * S.ns * ij + (int64_t)k * s + r] = y + wx;
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

// returns 2 if the rational solve ran in the same kernel (R given), else 1
template <int K, int NV>
static int tridiag_launch(hipStream_t st, unsigned nb, size_t lds, int k, const LocDev& L,
                          const StateDev& S, const VarsDev& V, const ChunkDev& C,
                          const DiagDev& D, double* B, const double* bvec, double* tau,
                          double* dd, double* ee, double* U, int* nvec, int* status,
                          int reg, const int32_t* perm, const RatTab& R) {
  // k >= 32 (K = 64), rows padded to a multiple of 8 (to 50 for k = 49 and
  // 50): reg 2 the two-problem register kernel (k <= 50), reg 1 (or 2 past
  // k = 50) the one-problem register kernel, reg 0 the LDS kernel
  const int kq = (k == 49 || k == 50) ? 50 : (k + 7) / 8 * 8;
  if (K == 64 && reg == 2) {
    const int rc = launch_tridiag_pair(st, (int64_t)nb, perm, kq, NV, k, K, L, S, V, C, D, B,
                                       bvec, tau, dd, ee, U, nvec, status, R.tab, R.n, R.nb);
    if (rc) return rc;
  }
  if (K == 64 && reg != 0 &&
      launch_tridiag_reg(st, nb, kq, NV, k, K, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec,
                         status))
    return 1;
  k_tridiag<K, NV><<<nb, K, lds, st>>>(k, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status);
  return 1;
}

template <int KQ>
static void ratsolve_launch(hipStream_t st, unsigned nb, int k, int K, int NV, int64_t np,
                            const ChunkDev& C, const double* dd, const double* ee,
                            const int* nvec, const int* status, const RatTab& R, double* U) {
  k_ratsolve<KQ><<<nb, 64, 0, st>>>(k, K, NV, np, C, dd, ee, nvec, status, R.tab, R.n, R.nb, U);
}

template <int K>
static void pipeline(hipStream_t st, const LocDev& L, const StateDev& S, const VarsDev& V,
                     const ChunkDev& C, const DiagDev& D, double* B, const double* bvec,
                     double* tau, double* dd, double* ee, double2* rot, int cap_rot,
                     int4* itr, int cap_it, int* nit, int* status, double* Z, int NV,
                     double* U, int* nvec, const RatTab& R, double* ana, int reg,
                     const int32_t* perm, bool fuse) {
  const int64_t np = C.ngp * C.G;
  const int k = (int)S.k;
  const unsigned nb = (unsigned)np;
  const unsigned nl = (unsigned)((np + 63) / 64);
  const size_t lds = sizeof(double) * (size_t)k * (k + 1);
  // the rational solve may run inside the two-problem tridiagonalization
  const RatTab RF = (NV > 0 && fuse) ? R : RatTab{};
  int tri = 1;
  switch (NV) {
    case 0: tri = tridiag_launch<K, 0>(st, nb, lds, k, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status, reg, perm, RF); break;
    case 2: tri = tridiag_launch<K, 2>(st, nb, lds, k, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status, reg, perm, RF); break;
    case 3: tri = tridiag_launch<K, 3>(st, nb, lds, k, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status, reg, perm, RF); break;
    case 5: tri = tridiag_launch<K, 5>(st, nb, lds, k, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status, reg, perm, RF); break;
    default: tri = tridiag_launch<K, 9>(st, nb, lds, k, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status, reg, perm, RF); break;
  }
  if (NV > 0 && R.nb > 0) {
    // rational T^(-1/2): no QL, no eigenvectors
    const unsigned nr = (unsigned)((np + 1) / 2);
    if (tri == 2) {   // solved inside the tridiagonalization
      k_qapply<K><<<nb, K, 0, st>>>(L, S, V, C, D, NV, B, tau, nvec, status, U, ana);
      return;
    }
    // rows of 50 for k = 49 and 50 (as the register tridiagonalization)
    switch (k == 49 || k == 50 ? 50 : (k + 7) / 8) {
      case 1: ratsolve_launch<8>(st, nr, k, K, NV, np, C, dd, ee, nvec, status, R, U); break;
      case 2: ratsolve_launch<16>(st, nr, k, K, NV, np, C, dd, ee, nvec, status, R, U); break;
      case 3: ratsolve_launch<24>(st, nr, k, K, NV, np, C, dd, ee, nvec, status, R, U); break;
      case 4: ratsolve_launch<32>(st, nr, k, K, NV, np, C, dd, ee, nvec, status, R, U); break;
      case 5: ratsolve_launch<40>(st, nr, k, K, NV, np, C, dd, ee, nvec, status, R, U); break;
      case 6: ratsolve_launch<48>(st, nr, k, K, NV, np, C, dd, ee, nvec, status, R, U); break;
      case 7: ratsolve_launch<56>(st, nr, k, K, NV, np, C, dd, ee, nvec, status, R, U); break;
      case 50: ratsolve_launch<50>(st, nr, k, K, NV, np, C, dd, ee, nvec, status, R, U); break;
      default: ratsolve_launch<64>(st, nr, k, K, NV, np, C, dd, ee, nvec, status, R, U); break;
    }
    k_qapply<K><<<nb, K, 0, st>>>(L, S, V, C, D, NV, B, tau, nvec, status, U, ana);
    return;
  }
  tql_launch<K, K>(st, nl, k, np, C, dd, ee, rot, cap_rot, itr, cap_it, nit, status);
  if (NV == 0) {
    k_replay<K><<<nb, K, 0, st>>>(C, rot, cap_rot, itr, cap_it, nit, status, Z);
    k_vec_apply<K><<<nb, K, 0, st>>>(L, S, V, C, D, B, bvec, tau, dd, Z, status, ana);
  } else {
    const unsigned nf = (unsigned)((np * NV + 63) / 64);
    const unsigned nbk = (unsigned)((np * (NV - 1) + 63) / 64);
    k_rot_fwd<K><<<nf, 64, 0, st>>>(k, NV, np, C, rot, cap_rot, itr, cap_it, nit, nvec, status, U);
    k_rot_bwd<K><<<nbk, 64, 0, st>>>(L, S, V, C, D, NV, np, B, tau, dd, rot, cap_rot, itr,
                                     cap_it, nit, nvec, status, U, ana);
  }
}

int eigen_vec_nv(int maxslab) {
  const int n = 1 + maxslab;
  return n <= 2 ? 2 : (n <= 3 ? 3 : (n <= 5 ? 5 : (n <= 9 ? 9 : 0)));
}

void launch_eigen_pipeline(hipStream_t st, const LocDev& L, const StateDev& S,
                           const VarsDev& V, const ChunkDev& C, const DiagDev& D,
                           double* B, const double* bvec, double* tau, double* dd,
                           double* ee, double2* rot, int cap_rot, int4* itr,
                           int cap_it, int* nit, int* status, double* Z, int NV, double* U,
                           int* nvec, const RatTab& R, double* ana, int trireg,
                           const int32_t* perm, bool fuse) {
  if (C.ngp * C.G <= 0) return;
  if (eigen_pipeline_K((int)S.k) == 32)
    pipeline<32>(st, L, S, V, C, D, B, bvec, tau, dd, ee, rot, cap_rot, itr, cap_it, nit,
                 status, Z, NV, U, nvec, R, ana, trireg, perm, fuse);
  else
    pipeline<64>(st, L, S, V, C, D, B, bvec, tau, dd, ee, rot, cap_rot, itr, cap_it, nit,
                 status, Z, NV, U, nvec, R, ana, trireg, perm, fuse);
}

}  // namespace lk
