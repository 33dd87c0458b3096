// transform_mat.h -- the LETKF transform as an explicit k x k matrix for
// large ensembles (64 <= k <= 128), where a group can hold hundreds of slabs
// (src/letkf_core.F90:100-141):
//
//   T = W + 1 wbar^T,  W = sqrt(k-1) A^(-1/2),  wbar = A^-1 Yb^T R^-1 d,
//   A = (k-1) I + Yb^T R^-1 Yb.
//
// Three kernels per chunk; B (k x k per problem, row major) carries the data
// between them:
//  k_mat_tridiag<NR>  one 256-thread block per problem; A, padded to
//                     N = 2 NR, in registers: thread (c, h) holds column c,
//                     rows [h NR, h NR + NR) (c = lane + 64 (wave & 1),
//                     h = wave >> 1).  Householder reduction A = Q Tri Q^T
//                     (LAPACK dsytd2, lower) by two-sided rank-2 updates
//                     A -= v w^T + w v^T; reflector j goes to row j of B
//                     (columns > j), tau, d and e to global arrays.
//  k_mat_rational     one 128-thread block per problem: the lower triangle of
//                     M = sqrt(k-1) Tri^(-1/2) = sum_n c_n (Tri + s_n I)^-1
//                     with the rational nodes of the node table (picked by
//                     the Gershgorin bound of Tri) into B (columns <= row).
//                     Column c of (Tri + s I)^-1 below the diagonal follows
//                     from the LDL^T pivots of Tri + s I: x_c = 1/g_c with
//                     g_c = D+_c + D-_c - (d_c + s), x_i = -(e_(i-1) / D-_i)
//                     x_(i-1) for i > c (D+ top-down, D- bottom-up pivots).
//                     Launched once per class of node counts (<= 16, 24,
//                     32), its LDS sized for the class.
//  k_mat_back<NR>     the register layout of k_mat_tridiag: W = Q M Q^T by
//                     the reflectors applied from both sides in reverse
//                     order (the same two-sided rank-2 update), then
//                     wbar = W (W b) / (k-1) and T = W + 1 wbar^T over B.
#pragma once

#include <hip/hip_runtime.h>

#include "eigen_common.h"
#include "letkf_internal.h"

namespace lk {

constexpr int kMatStride = 128;   // d, e, tau entries per problem
// rows per fenced group in the unrolled row loops: the LDS reads of a group
// are in flight together (more rows, more registers)
constexpr int kFenceRows = 8;

// keeps the scheduler from hoisting the LDS reads of an unrolled loop far
// ahead of their use (which would need more registers than a wave has)
#define LK_FENCE(x) asm volatile("" : "+v"(x) :: "memory")

template <int NR>
__global__ __launch_bounds__(256) void k_mat_tridiag(int k, ChunkDev C, double* __restrict__ B,
                                                     double* __restrict__ tau,
                                                     double* __restrict__ dd,
                                                     double* __restrict__ ee) {
  static_assert(2 * NR <= 128, "two row halves of at most 64");
  __shared__ double sx[128];
  __shared__ double2 svw[128];
  __shared__ double pp[2][128];
  __shared__ double se[128], stau[128];
  const int64_t pair = blockIdx.x;
  if (!C.solve[pair]) return;
  const int tid = threadIdx.x, l = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int h = wv >> 1, c = l + 64 * (wv & 1);
  const double m = (double)(k - 1);
  double* Bp = B + pair * (int64_t)k * k;
  // first row of this half, re-fenced in every step so that the per-row
  // indices and tests derived from it are not kept live across the loop
  int r0 = NR * h;
  double a[NR];
#pragma unroll
  for (int s = 0; s < NR; ++s) {
    const int r = r0 + s;
    double v = 0.0;
    if (r < k && c < k) {
      v = Bp[(int64_t)r * k + c];           // symmetric: row r read along c
      if (r == c) v += m;
    }
    a[s] = v;
  }
  if (tid < 128) { sx[tid] = 0.0; se[tid] = 0.0; stau[tid] = 0.0; }
  __syncthreads();
  for (int j = 0; j < k - 2; ++j) {
    asm volatile("" : "+s"(r0));
    const int j1 = j + 1;
    if (c == j) {
#pragma unroll
      for (int s = 0; s < NR; ++s) sx[r0 + s] = a[s];
    }
    __syncthreads();
    const double x0 = sx[l], x1 = sx[l + 64];
    double q = 0.0;
    if (l > j1 && l < k) q += x0 * x0;
    if (l + 64 > j1 && l + 64 < k) q += x1 * x1;
    const double sig = wave_sum_dpp(q);        // the same in every wave
    const double alpha = sx[j1];
    // column already reduced (sig = 0): tau = 0, beta = alpha, and the step
    // below changes nothing (kept in line: a separate path would double
    // the registers the matrix needs)
    const bool refl = sig != 0.0;
    const double mu = sqrt(alpha * alpha + sig);
    const double beta = refl ? (alpha >= 0.0 ? -mu : mu) : alpha;
    const double t = refl ? (beta - alpha) / beta : 0.0;
    const double sc = refl ? 1.0 / (alpha - beta) : 0.0;
    // p = t A v over this half's rows (v: 1 at j+1, x sc below, 0 above)
    // a wave whose rows or whose columns all lie at or above row j has
    // nothing left to do: v and w vanish there
    const bool live = r0 + NR - 1 > j && 64 * (wv & 1) + 63 > j;
    double pc = 0.0;
    if (live) {
#pragma unroll
      for (int s = 0; s < NR; ++s) {
        // branch-free: the LDS read is unconditional, the row tests only
        // pick scalar factors (a conditional read became a branch and a
        // full LDS wait per row)
        const int r = r0 + s;
        const double fr = r > j1 ? sc : 0.0;
        const double ar = r == j1 ? 1.0 : 0.0;
        const double vr = fma(sx[r], fr, ar);
        pc = fma(a[s], vr, pc);
        if ((s & (kFenceRows - 1)) == kFenceRows - 1) LK_FENCE(pc);
      }
    }
    pp[h][c] = pc;
    __syncthreads();
    const double v0 = l == j1 ? 1.0 : (l > j1 ? x0 * sc : 0.0);
    const double v1 = l + 64 == j1 ? 1.0 : (l + 64 > j1 ? x1 * sc : 0.0);
    const double p0 = t * (pp[0][l] + pp[1][l]);
    const double p1 = t * (pp[0][l + 64] + pp[1][l + 64]);
    const double kc = 0.5 * t * wave_sum_dpp(v0 * p0 + v1 * p1);
    const double vc = c == j1 ? 1.0 : (c > j1 ? sx[c] * sc : 0.0);
    const double wc = c > j ? t * (pp[0][c] + pp[1][c]) - kc * vc : 0.0;
    if (h == 0) {
      svw[c] = make_double2(vc, wc);
      if (c < k) Bp[(int64_t)j * k + c] = vc;        // reflector j (columns > j)
    }
    if (tid == 0) { stau[j] = t; se[j] = beta; }
    __syncthreads();
    if (live) {
#pragma unroll
      for (int s = 0; s < NR; ++s) {
        const double2 q2 = svw[r0 + s];
        a[s] = fma(-q2.x, wc, fma(-q2.y, vc, a[s]));
        if ((s & (kFenceRows - 1)) == kFenceRows - 1) LK_FENCE(a[s]);
      }
    }
  }
  // diagonal, and the last subdiagonal entry A[k-1][k-2]
  asm volatile("" : "+s"(r0));
  double* dp = dd + pair * kMatStride;
  double* ep = ee + pair * kMatStride;
#pragma unroll
  for (int s = 0; s < NR; ++s) {
    const int r = r0 + s;
    if (r == c && c < k) dp[c] = a[s];
    if (c == k - 2 && r == k - 1) ep[k - 2] = a[s];
  }
  if (tid < k - 2) {
    ep[tid] = se[tid];
    tau[pair * kMatStride + tid] = stau[tid];
  }
}

template <int NR>
__global__ __launch_bounds__(256) void k_mat_back(int k, ChunkDev C, double* __restrict__ B,
                                                  const double* __restrict__ bvec,
                                                  const double* __restrict__ tau) {
  __shared__ double sx[128];
  __shared__ double2 svw[2][128];
  __shared__ double pp[2][128];
  __shared__ double stau[128];
  const int64_t pair = blockIdx.x;
  if (!C.solve[pair]) return;
  const int tid = threadIdx.x, l = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int h = wv >> 1, c = l + 64 * (wv & 1);
  const double m = (double)(k - 1);
  double* Bp = B + pair * (int64_t)k * k;
  int r0 = NR * h;
  // M from the lower triangle of B (symmetric)
  double a[NR];
#pragma unroll
  for (int s = 0; s < NR; ++s) {
    const int r = r0 + s;
    double v = 0.0;
    if (r < k && c < k) v = r >= c ? Bp[(int64_t)r * k + c] : Bp[(int64_t)c * k + r];
    a[s] = v;
  }
  if (tid < 128) stau[tid] = tid < k - 2 ? tau[pair * kMatStride + tid] : 0.0;
  __syncthreads();
  for (int j = k - 3; j >= 0; --j) {
    asm volatile("" : "+s"(r0));
    const int buf = j & 1;
    const double t = stau[j];
    const double vc = (c > j && c < k) ? Bp[(int64_t)j * k + c] : 0.0;
    if (h == 0) svw[buf][c].x = vc;
    __syncthreads();
    double pc = 0.0;
    if (r0 + NR - 1 > j) {
#pragma unroll
      for (int s = 0; s < NR; ++s) {
        pc = fma(a[s], svw[buf][r0 + s].x, pc);
        if ((s & (kFenceRows - 1)) == kFenceRows - 1) LK_FENCE(pc);
      }
    }
    pp[h][c] = pc;
    __syncthreads();
    const double v0 = svw[buf][l].x, v1 = svw[buf][l + 64].x;
    const double p0 = t * (pp[0][l] + pp[1][l]);
    const double p1 = t * (pp[0][l + 64] + pp[1][l + 64]);
    const double kc = 0.5 * t * wave_sum_dpp(v0 * p0 + v1 * p1);
    const double wc = t * (pp[0][c] + pp[1][c]) - kc * vc;
    if (h == 0) svw[buf][c].y = wc;
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NR; ++s) {
      const double2 q2 = svw[buf][r0 + s];
      a[s] = fma(-q2.x, wc, fma(-q2.y, vc, a[s]));
      if ((s & (kFenceRows - 1)) == kFenceRows - 1) LK_FENCE(a[s]);
    }
  }
  // wbar = W (W b) / (k-1); T = W + 1 wbar^T, row major over B
  asm volatile("" : "+s"(r0));
  __syncthreads();
  if (h == 0) sx[c] = c < k ? bvec[pair * (int64_t)k + c] : 0.0;
  __syncthreads();
  double q = 0.0;
#pragma unroll
  for (int s = 0; s < NR; ++s) {
    q = fma(a[s], sx[r0 + s], q);
    if ((s & (kFenceRows - 1)) == kFenceRows - 1) LK_FENCE(q);
  }
  pp[h][c] = q;
  __syncthreads();
  const double t1 = pp[0][c] + pp[1][c];
  __syncthreads();
  if (h == 0) sx[c] = t1;
  __syncthreads();
  q = 0.0;
#pragma unroll
  for (int s = 0; s < NR; ++s) {
    q = fma(a[s], sx[r0 + s], q);
    if ((s & (kFenceRows - 1)) == kFenceRows - 1) LK_FENCE(q);
  }
  pp[h][c] = q;
  __syncthreads();
  const double wb = (pp[0][c] + pp[1][c]) / m;
  asm volatile("" : "+s"(r0));
#pragma unroll
  for (int s = 0; s < NR; ++s) {
    const int r = r0 + s;
    if (r < k && c < k) Bp[(int64_t)r * k + c] = a[s] + wb;
  }
}

// one instantiation per translation unit (transform_mat_nr*.hip)
template <int NR>
void launch_mat_tridiag_nr(hipStream_t st, unsigned nb, int k, const ChunkDev& C, double* B,
                           double* tau, double* dd, double* ee) {
  k_mat_tridiag<NR><<<nb, 256, 0, st>>>(k, C, B, tau, dd, ee);
}
template <int NR>
void launch_mat_back_nr(hipStream_t st, unsigned nb, int k, const ChunkDev& C, double* B,
                        const double* bvec, const double* tau) {
  k_mat_back<NR><<<nb, 256, 0, st>>>(k, C, B, bvec, tau);
}

}  // namespace lk
