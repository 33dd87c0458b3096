// tridiag_reg.h -- Householder tridiagonalization with the matrix in
// registers (the template; instantiated per row length KQ in
// tridiag_reg_kq*.hip so that the fully unrolled kernels compile in parallel).
#pragma once

#include <hip/hip_runtime.h>

#include "eigen_common.h"
#include "letkf_internal.h"

namespace lk {

// ---------------------------------------------------------------------------
// The same reduction with the matrix in registers: lane r keeps row r of A
// (a[c] = A(r, c), KQ >= k even).  The Householder vector v goes
// through LDS for the symmetric product A v (read by all lanes), and the
// rank-2 update takes v_c and w_c = (t A v - (t^2 v.Av / 2) v)_c from lane c
// by readlane, so the LDS and the VALU share the broadcast work.  The step
// loop is unrolled so that every register index is static; columns c <= j
// are no longer touched at step j.
// ---------------------------------------------------------------------------
template <int KQ, int NV>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
void k_tridiag_reg(int k, int K, LocDev L, StateDev S, VarsDev V, ChunkDev C, DiagDev D,
                   double* __restrict__ B, const double* __restrict__ bvec,
                   double* __restrict__ tau, double* __restrict__ dd,
                   double* __restrict__ ee, double* __restrict__ U,
                   int* __restrict__ nvec, int* __restrict__ status) {
  __shared__ double sv[KQ];
  __shared__ int sl[NV > 0 ? NV : 1];
  const int64_t pair = blockIdx.x;
  if (!C.solve[pair]) return;
  const int r = threadIdx.x;
  double* Bp = B + pair * (int64_t)k * k;
  double a[KQ];
#pragma unroll
  for (int c = 0; c < KQ; ++c) {
    double v = (r < k && c < k) ? Bp[r + (int64_t)k * c] : 0.0;
    a[c] = (c == r && r < k) ? v + (double)(k - 1) : v;
  }
  double xv[NV > 0 ? NV : 1];
  int nsl = 0;
  if (NV > 0) {
    const int64_t t_gp = pair / C.G;
    const int g = (int)(pair - t_gp * C.G);
    const int64_t ij = C.i0 + t_gp;
    const int mld = D.mld[ij];
    for (int s0 = 0; s0 < S.ns; s0 += 64) {
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
  double* taup = tau + pair * K;
  double* ep = ee + pair * K;
#pragma unroll
  for (int j = 0; j < KQ - 1; ++j) {
    // static j: every a[] index below is a compile-time constant
    if (j <= k - 2) {
      const double x = (r > j && r < k) ? a[j] : 0.0;
      const double alpha = bcast(a[j], j + 1);
      const double sig = wsumK<64>((r > j + 1 && r < k) ? x * x : 0.0);
      if (sig == 0.0) {
        if (r == 0) { taup[j] = 0.0; ep[j] = alpha; }
      } else {
        const double mu = sqrt(alpha * alpha + sig);
        const double beta = alpha >= 0.0 ? -mu : mu;
        const double t = (beta - alpha) / beta;
        const double scal = 1.0 / (alpha - beta);
        const double v = (r == j + 1) ? 1.0 : ((r > j + 1 && r < k) ? x * scal : 0.0);
        if (r < KQ) sv[r] = v;
        if (NV > 0) {
          double dq[NV > 0 ? NV : 1];
#pragma unroll
          for (int q = 0; q < NV; ++q) dq[q] = v * xv[q];
#pragma unroll
          for (int o = 32; o >= 1; o >>= 1)
#pragma unroll
            for (int q = 0; q < NV; ++q) dq[q] += __shfl_xor(dq[q], o, 64);
#pragma unroll
          for (int q = 0; q < NV; ++q) xv[q] -= (t * dq[q]) * v;
        }
        __syncthreads();
        double p0 = 0.0, p1 = 0.0;
#pragma unroll
        for (int c = j + 1; c < KQ; ++c) {
          if (c & 1) p1 = fma(a[c], sv[c], p1);
          else p0 = fma(a[c], sv[c], p0);
          if ((c & 7) == 7) asm volatile("" : "+v"(p0), "+v"(p1) :: "memory");
        }
        const double p = (r > j && r < k) ? t * (p0 + p1) : 0.0;
        const double kc = 0.5 * t * wsumK<64>(p * v);
        const double w = p - kc * v;
        // rank-2 update; v_c and w_c come from lane c by readlane (VALU),
        // which leaves the LDS to the symv above
#pragma unroll
        for (int c = j + 1; c < KQ; ++c) {
          const double vc = bcast(v, c), wc = bcast(w, c);
          a[c] = fma(-w, vc, fma(-v, wc, a[c]));
          if ((c & 3) == 3) asm volatile("" : "+v"(a[c]) :: "memory");
          else asm volatile("" : "+v"(a[c]));
        }
        if (r > j && r < k) Bp[(int64_t)j * k + r] = v;  // reflector j, rows j+1..k-1
        if (r == 0) { taup[j] = t; ep[j] = beta; }
        __syncthreads();
      }
    }
    // keep the scheduler from interleaving steps (register pressure)
    __builtin_amdgcn_sched_barrier(0);
  }
  double dg = 0.0;
#pragma unroll
  for (int c = 0; c < KQ; ++c) dg = (c == r) ? a[c] : dg;
  if (r < K) {
    dd[pair * K + r] = r < k ? dg : 0.0;
    if (r >= k - 1) { ep[r] = 0.0; taup[r] = 0.0; }
  }
  if (NV > 0) {
    double* Up = U + pair * (int64_t)NV * K;
#pragma unroll
    for (int q = 0; q < NV; ++q)
      if (r < K) Up[q * K + r] = xv[q];
    if (r == 0) {
      nvec[pair] = nsl;
      status[pair] = nsl > NV - 1 ? 3 : 0;
    }
  } else if (r == 0) {
    status[pair] = 0;
  }
}

template <int KQ>
bool launch_tridiag_reg_kq(hipStream_t st, unsigned nb, int NV, int k, int K, const LocDev& L,
                           const StateDev& S, const VarsDev& V, const ChunkDev& C,
                           const DiagDev& D, double* B, const double* bvec, double* tau,
                           double* dd, double* ee, double* U, int* nvec, int* status) {
  switch (NV) {
    case 0: k_tridiag_reg<KQ, 0><<<nb, 64, 0, st>>>(k, K, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status); return true;
    case 2: k_tridiag_reg<KQ, 2><<<nb, 64, 0, st>>>(k, K, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status); return true;
    case 3: k_tridiag_reg<KQ, 3><<<nb, 64, 0, st>>>(k, K, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status); return true;
    case 5: k_tridiag_reg<KQ, 5><<<nb, 64, 0, st>>>(k, K, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status); return true;
    case 9: k_tridiag_reg<KQ, 9><<<nb, 64, 0, st>>>(k, K, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status); return true;
    default: return false;
  }
}

}  // namespace lk
