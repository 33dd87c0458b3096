// transform_mat.h -- the LETKF transform as an explicit k x k matrix
// and its application to every slab of the group on MFMA: the batched path
// for large ensembles (64 <= k <= 128), where a group can hold hundreds of
// slabs (src/letkf_core.F90:100-141, src/letkf_solver.F90:284-290):
//
//   T = W + 1 wbar^T,  W = sqrt(k-1) A^(-1/2),  wbar = A^-1 Yb^T R^-1 d,
//   A = (k-1) I + Yb^T R^-1 Yb,      x_a = T x_b for each slab x_b.
//
// k_transform_mat  one 256-thread block per problem.  A, padded to
//                  N = 2 NR, lives in registers: thread (c, h) holds column
//                  c, rows [h NR, h NR + NR) (c = lane + 64 (wave & 1),
//                  h = wave >> 1).
//   1. Householder reduction A = Q Tri Q^T (LAPACK dsytd2, lower) by
//      two-sided rank-2 updates A -= v w^T + w v^T; reflector j is stored in
//      row j of B, tau and the tridiagonal in LDS.
//   2. M = sqrt(k-1) Tri^(-1/2) = sum_n c_n (Tri + s_n I)^-1 with the
//      rational nodes of the node table (the Gershgorin bound of Tri picks
//      the table).  Column c of (Tri + s I)^-1 follows from the top-down and
//      bottom-up LDL^T pivots D+, D- of Tri + s I: x_c = 1/g_c with
//      g_c = D+_c + D-_c - (d_c + s), x_i = -(e_i / D+_i) x_(i+1) above c and
//      x_i = -(e_(i-1) / D-_i) x_(i-1) below it.  Each thread runs these
//      two-term recurrences for its own column over its own half of the
//      rows; the values at the half boundary pass through LDS.
//   3. W = Q M Q^T: the reflectors applied from both sides in reverse order,
//      the same two-sided rank-2 update as step 1.
//   4. wbar = W (W b) / (k-1); T = W + 1 wbar^T written over B, row major.
// k_apply_mat      one 256-thread block per problem; wave w computes members
//                  32w..32w+31 of x_a = T x_b for 32 slabs at a time on
//                  MFMA (32x32x2 FP32, or 32x32x16 BF16 / bf16x3 with FP32
//                  accumulation) and writes the FP64 analysis.
#pragma once

#include <hip/hip_runtime.h>

#include "eigen_common.h"
#include "letkf_internal.h"

namespace lk {

constexpr int kNodeBatch = 8;     // rational nodes per factorization batch

// keeps the scheduler from hoisting the LDS reads of an unrolled loop far
// ahead of their use (which would need more registers than a wave has)
#define LK_FENCE(x) asm volatile("" : "+v"(x) :: "memory")

__device__ __forceinline__ double wmax64(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

template <int NR>
__global__ __launch_bounds__(256) void k_transform_mat(int k, ChunkDev C, double* __restrict__ B,
                                                       const double* __restrict__ bvec,
                                                       int* __restrict__ status,
                                                       const double* __restrict__ tab,
                                                       const int* __restrict__ tabn,
                                                       int nbucket) {
  static_assert(2 * NR <= 128, "two row halves of at most 64");
  __shared__ double sx[128];
  __shared__ double2 svw[2][128];
  __shared__ double pp[2][128];
  __shared__ double sd[128], se[128], stau[128];
  __shared__ double Ls[kNodeBatch][128], Us[kNodeBatch][128];
  __shared__ double Gs[kNodeBatch][128], Ds[kNodeBatch][128];
  __shared__ double xbu[kNodeBatch][128], xbd[kNodeBatch][128];
  const int64_t pair = blockIdx.x;
  if (!C.solve[pair]) return;
  const int tid = threadIdx.x, l = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);      // wave-uniform
  const int h = wv >> 1, c = l + 64 * (wv & 1);
  const double m = (double)(k - 1);
  double* Bp = B + pair * (int64_t)k * k;
  // first row of this half; re-fenced at the top of each section so that
  // the per-row indices and tests derived from it are not kept live (as
  // hundreds of registers) across the whole kernel
  int r0 = NR * h;
#define LK_ROWS() asm volatile("" : "+s"(r0))

  // A = (k-1) I + B (symmetric: row c of B read as column c, coalesced)
  double a[NR];
#pragma unroll
  for (int s = 0; s < NR; ++s) {
    const int r = r0 + s;
    double v = 0.0;
    if (r < k && c < k) {
      v = Bp[(int64_t)r * k + c];
      if (r == c) v += m;
    }
    a[s] = v;
  }
  if (tid < 128) {
    sx[tid] = 0.0;
    sd[tid] = 0.0;
    se[tid] = 0.0;
    stau[tid] = 0.0;
#pragma unroll
    for (int t = 0; t < kNodeBatch; ++t) { Ls[t][tid] = 0.0; Us[t][tid] = 0.0; }
  }
  __syncthreads();

  // ---- 1. tridiagonalization ------------------------------------------
  for (int j = 0; j < k - 2; ++j) {
    LK_ROWS();
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
    const double sig = wsumK<64>(q);     // the same in every wave
    const double alpha = sx[j1];
    if (sig == 0.0) {
      if (tid == 0) { stau[j] = 0.0; se[j] = alpha; }
      if (h == 0 && c < k) Bp[(int64_t)j * k + c] = 0.0;
      __syncthreads();
      continue;
    }
    const double mu = sqrt(alpha * alpha + sig);
    const double beta = alpha >= 0.0 ? -mu : mu;
    const double t = (beta - alpha) / beta;
    const double sc = 1.0 / (alpha - beta);
    // p = t A v over this half's rows (v: 1 at j+1, x sc below, 0 above)
    double pc = 0.0;
    if (r0 + NR - 1 > j) {
#pragma unroll
      for (int s = 0; s < NR; ++s) {
        const int r = r0 + s;
        const double vr = r == j1 ? 1.0 : (r > j1 ? sx[r] * sc : 0.0);
        pc = fma(a[s], vr, pc);
        if ((s & 3) == 3) LK_FENCE(pc);
      }
    }
    pp[h][c] = pc;
    __syncthreads();
    const double v0 = l == j1 ? 1.0 : (l > j1 ? x0 * sc : 0.0);
    const double v1 = l + 64 == j1 ? 1.0 : (l + 64 > j1 ? x1 * sc : 0.0);
    const double p0 = t * (pp[0][l] + pp[1][l]);
    const double p1 = t * (pp[0][l + 64] + pp[1][l + 64]);
    const double kc = 0.5 * t * wsumK<64>(v0 * p0 + v1 * p1);
    const double vc = c == j1 ? 1.0 : (c > j1 ? sx[c] * sc : 0.0);
    const double wc = c > j ? t * (pp[0][c] + pp[1][c]) - kc * vc : 0.0;
    if (h == 0) {
      svw[0][c] = make_double2(vc, wc);
      if (c < k) Bp[(int64_t)j * k + c] = vc;        // reflector j
    }
    if (tid == 0) { stau[j] = t; se[j] = beta; }
    __syncthreads();
    if (r0 + NR - 1 > j) {
#pragma unroll
      for (int s = 0; s < NR; ++s) {
        const double2 q2 = svw[0][r0 + s];
        a[s] = fma(-q2.x, wc, fma(-q2.y, vc, a[s]));
        if ((s & 3) == 3) LK_FENCE(a[s]);
      }
    }
  }
  // diagonal, and the last subdiagonal entry A[k-1][k-2]
  LK_ROWS();
#pragma unroll
  for (int s = 0; s < NR; ++s) {
    const int r = r0 + s;
    if (r == c && c < k) sd[c] = a[s];
    if (c == k - 2 && r == k - 1) se[k - 2] = a[s];
  }
  __syncthreads();

  // ---- 2. M = sqrt(k-1) Tri^(-1/2) ------------------------------------
  double gm = 0.0;
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int i = l + 64 * u;
    if (i < k) {
      const double em = i > 0 ? fabs(se[i - 1]) : 0.0;
      const double ep = i < k - 1 ? fabs(se[i]) : 0.0;
      gm = fmax(gm, sd[i] + em + ep);
    }
  }
  gm = wmax64(gm);
  const double kappa = gm / m;
  int bk = (int)ceil(log2(fmax(kappa, 2.0)));
  bk = bk < 1 ? 1 : (bk > nbucket - 1 ? nbucket - 1 : bk);
  bk = __builtin_amdgcn_readfirstlane(bk);     // the same in every lane
  const int nn = __builtin_amdgcn_readfirstlane(tabn[bk]);
  const double* sn = tab + bk * 64;
  const double* wn = sn + 32;
#pragma unroll
  for (int s = 0; s < NR; ++s) a[s] = 0.0;
  for (int n0 = 0; n0 < nn; n0 += kNodeBatch) {
    LK_ROWS();
    const int nb = min(kNodeBatch, nn - n0);
    // pivots: wave 0 top-down, wave 1 bottom-up, one lane per node
    if (wv == 0 && l < nb) {
      const double sg = m * sn[n0 + l];
      double Lp = 0.0;
      for (int i = 0; i < k; ++i) {
        const double Dp = sd[i] + sg - (i > 0 ? se[i - 1] * Lp : 0.0);
        const double Li = i < k - 1 ? se[i] / Dp : 0.0;
        Ls[l][i] = Li;
        Gs[l][i] = Dp;
        Lp = Li;
      }
    } else if (wv == 1 && l < nb) {
      const double sg = m * sn[n0 + l];
      double Up = 0.0;
      for (int i = k - 1; i >= 0; --i) {
        const double Dm = sd[i] + sg - (i < k - 1 ? se[i] * Up : 0.0);
        const double Ui = i > 0 ? se[i - 1] / Dm : 0.0;
        Us[l][i] = Ui;
        Ds[l][i] = Dm;
        Up = Ui;
      }
    }
    __syncthreads();
    if (tid < k) {
      for (int t = 0; t < nb; ++t) {
        const double sg = m * sn[n0 + t];
        Gs[t][tid] = 1.0 / (Gs[t][tid] + Ds[t][tid] - (sd[tid] + sg));
      }
    }
    __syncthreads();
    // an opaque copy of c: the row tests below are then recomputed in each
    // batch instead of being hoisted out of the batch loop as hundreds of
    // live masks
    int cq = c;
    asm volatile("" : "+v"(cq));
    double gc[kNodeBatch], cf[kNodeBatch], x[kNodeBatch];
#pragma unroll
    for (int t = 0; t < kNodeBatch; ++t) {
      gc[t] = (t < nb && c < k) ? Gs[t][c] : 0.0;
      cf[t] = t < nb ? m * wn[n0 + t] : 0.0;
      x[t] = 0.0;
    }
    // phase 1: upward recurrence through half 1, downward through half 0
    if (h == 1) {
#pragma unroll
      for (int s = NR - 1; s >= 0; --s) {
        const int r = NR + s;
        double acc = 0.0;
#pragma unroll
        for (int t = 0; t < kNodeBatch; ++t) {
          x[t] = r == cq ? gc[t] : (r < cq ? -Ls[t][r] * x[t] : 0.0);
          acc = fma(cf[t], x[t], acc);
        }
        a[s] += acc;
        asm volatile("" : "+v"(a[s]), "+v"(cq) :: "memory");
      }
#pragma unroll
      for (int t = 0; t < kNodeBatch; ++t) xbu[t][c] = x[t];
    } else {
#pragma unroll
      for (int s = 0; s < NR; ++s) {
        const int r = s;
        double acc = 0.0;
#pragma unroll
        for (int t = 0; t < kNodeBatch; ++t) {
          x[t] = r == cq ? gc[t] : (r > cq ? -Us[t][r] * x[t] : 0.0);
          acc = fma(cf[t], x[t], acc);
        }
        if (r > cq) a[s] += acc;
        asm volatile("" : "+v"(a[s]), "+v"(cq) :: "memory");        // the diagonal comes with the upward pass
      }
#pragma unroll
      for (int t = 0; t < kNodeBatch; ++t) xbd[t][c] = x[t];
    }
    __syncthreads();
    // phase 2: the other direction through each half, continued from the
    // boundary value of the other half
    if (h == 0) {
#pragma unroll
      for (int t = 0; t < kNodeBatch; ++t) x[t] = cq >= NR ? xbu[t][c] : 0.0;
#pragma unroll
      for (int s = NR - 1; s >= 0; --s) {
        const int r = s;
        double acc = 0.0;
#pragma unroll
        for (int t = 0; t < kNodeBatch; ++t) {
          x[t] = r == cq ? gc[t] : (r < cq ? -Ls[t][r] * x[t] : 0.0);
          acc = fma(cf[t], x[t], acc);
        }
        a[s] += acc;
        asm volatile("" : "+v"(a[s]), "+v"(cq) :: "memory");
      }
    } else {
#pragma unroll
      for (int t = 0; t < kNodeBatch; ++t) x[t] = cq < NR ? xbd[t][c] : 0.0;
#pragma unroll
      for (int s = 0; s < NR; ++s) {
        const int r = NR + s;
        double acc = 0.0;
#pragma unroll
        for (int t = 0; t < kNodeBatch; ++t) {
          x[t] = r == cq ? gc[t] : (r > cq ? -Us[t][r] * x[t] : 0.0);
          acc = fma(cf[t], x[t], acc);
        }
        if (r > cq) a[s] += acc;
        asm volatile("" : "+v"(a[s]), "+v"(cq) :: "memory");
      }
    }
    __syncthreads();
  }

  // ---- 3. W = Q M Q^T ---------------------------------------------------
  for (int j = k - 3; j >= 0; --j) {
    LK_ROWS();
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
        if ((s & 3) == 3) LK_FENCE(pc);
      }
    }
    pp[h][c] = pc;
    __syncthreads();
    const double v0 = svw[buf][l].x, v1 = svw[buf][l + 64].x;
    const double p0 = t * (pp[0][l] + pp[1][l]);
    const double p1 = t * (pp[0][l + 64] + pp[1][l + 64]);
    const double kc = 0.5 * t * wsumK<64>(v0 * p0 + v1 * p1);
    const double wc = t * (pp[0][c] + pp[1][c]) - kc * vc;
    if (h == 0) svw[buf][c].y = wc;
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NR; ++s) {
      const double2 q2 = svw[buf][r0 + s];
      a[s] = fma(-q2.x, wc, fma(-q2.y, vc, a[s]));
      if ((s & 3) == 3) LK_FENCE(a[s]);
    }
  }

  // ---- 4. wbar = W (W b) / (k-1); T = W + 1 wbar^T ----------------------
  LK_ROWS();
  __syncthreads();
  if (h == 0) sx[c] = c < k ? bvec[pair * (int64_t)k + c] : 0.0;
  __syncthreads();
  double q = 0.0;
#pragma unroll
  for (int s = 0; s < NR; ++s) {
    q = fma(a[s], sx[r0 + s], q);
    if ((s & 3) == 3) LK_FENCE(q);
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
    if ((s & 3) == 3) LK_FENCE(q);
  }
  pp[h][c] = q;
  __syncthreads();
  const double wb = (pp[0][c] + pp[1][c]) / m;
  LK_ROWS();
#pragma unroll
  for (int s = 0; s < NR; ++s) {
    const int r = r0 + s;
    if (r < k && c < k) Bp[(int64_t)r * k + c] = a[s] + wb;
  }
  if (tid == 0) status[pair] = 0;
#undef LK_ROWS
}

// one instantiation per translation unit (transform_mat_nr*.hip)
template <int NR>
void launch_transform_nr(hipStream_t st, unsigned nb, int k, const ChunkDev& C, double* B,
                         const double* bvec, int* status, const RatTab& R) {
  k_transform_mat<NR><<<nb, 256, 0, st>>>(k, C, B, bvec, status, R.tab, R.n, R.nb);
}

}  // namespace lk
