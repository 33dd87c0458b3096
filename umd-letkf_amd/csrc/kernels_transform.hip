// kernels_transform.hip -- the LETKF transform as an explicit k x k matrix
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
#include <hip/hip_runtime.h>

#include "letkf_internal.h"
#include "mx_common.h"

namespace lk {

// k_transform_mat<NR> launchers (transform_mat.h, one file per NR)
#define LK_NR(n)                                                                         \
  void launch_transform_nr##n(hipStream_t st, unsigned nb, int k, const ChunkDev& C,     \
                              double* B, const double* bvec, int* status, const RatTab& R);
LK_NR(40) LK_NR(48) LK_NR(56) LK_NR(64)
#undef LK_NR

// x_a = T x_b for every slab of the group: A operand T (rows = members),
// B operand x_b (columns = slabs), K = the ensemble index.  Lane (i, half)
// holds K = 16 step + 8 half + e (e = 0..7) of row / column i in both
// operands (the same K order on both sides, any order contracts the same).
template <int PREC>
__global__ __launch_bounds__(256) void k_apply_mat(LocDev L, StateDev S, VarsDev V, ChunkDev C,
                                                   DiagDev D, const double* __restrict__ B,
                                                   double* __restrict__ ana) {
  constexpr int KS = 8;              // K steps of 16, k <= 128
  __shared__ int sl[2048];
  __shared__ int snsl;
  const int64_t pair = blockIdx.x;
  if (!C.solve[pair]) return;
  const int k = (int)S.k;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int64_t t_gp = pair / C.G;
  const int g = (int)(pair - t_gp * C.G);
  const int64_t ij = C.i0 + t_gp;
  if (wv == 0) {
    // slabs of this group in ascending order (ballot compaction)
    const int mld = D.mld[ij];
    int nsl = 0;
    for (int s0 = 0; s0 < S.ns; s0 += 64) {
      const int s = s0 + lane;
      const bool in = s < S.ns && slab_group(L, V, s, mld) == g;
      const unsigned long long bal = __ballot(in);
      const int q = nsl + __popcll(bal & ((1ull << lane) - 1ull));
      if (in) sl[q] = s;
      nsl += __popcll(bal);
    }
    if (lane == 0) snsl = nsl;
  }
  __syncthreads();
  const int nsl = snsl;
  if (32 * wv >= k || nsl == 0) return;
  const int i = lane & 31, half = lane >> 5;
  const int ks = (k + 15) / 16;
  // T rows 32 wv + i, converted once
  const int row = 32 * wv + i;
  const double* Tp = B + pair * (int64_t)k * k + (int64_t)row * k;
  float af[KS][8];
#pragma unroll
  for (int st = 0; st < KS; ++st)
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int kk = 16 * st + 8 * half + e;
      af[st][e] = (st < ks && row < k && kk < k) ? (float)Tp[kk] : 0.0f;
    }
  bf16x8 ahi[KS], alo[KS];
  if (PREC != LETKF_PREC_FP32) {
#pragma unroll
    for (int st = 0; st < KS; ++st) {
      if (PREC == LETKF_PREC_BF16) ahi[st] = to_bf16(af[st]);
      else split_bf16(af[st], ahi[st], alo[st]);
    }
  }
  const float* xb0 = S.state + (int64_t)S.k * S.ns * ij;
  double* xa0 = ana + (int64_t)S.k * S.ns * ij;
  for (int q0 = 0; q0 < nsl; q0 += 32) {
    const int qi = q0 + i;
    const int slab = qi < nsl ? sl[qi] : sl[0];
    const float* xp = xb0 + (int64_t)k * slab;
    f32x16 acc = (f32x16){};
#pragma unroll
    for (int st = 0; st < KS; ++st) {
      if (st < ks) {
        float xf[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int kk = 16 * st + 8 * half + e;
          xf[e] = kk < k ? xp[kk] : 0.0f;
        }
        if (PREC == LETKF_PREC_FP32) {
#pragma unroll
          for (int e = 0; e < 8; ++e)
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(af[st][e], xf[e], acc, 0, 0, 0);
        } else if (PREC == LETKF_PREC_BF16) {
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[st], to_bf16(xf), acc, 0, 0, 0);
        } else {
          bf16x8 xh, xl;
          split_bf16(xf, xh, xl);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo[st], xh, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[st], xl, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[st], xh, acc, 0, 0, 0);
        }
      }
    }
    // register r of lane (i, half): member 32 wv + 8 (r/4) + 4 half + r%4 of slab q0 + i
    if (qi < nsl) {
      double* op = xa0 + (int64_t)k * slab;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int mem = 32 * wv + 8 * (r >> 2) + 4 * half + (r & 3);
        if (mem < k) op[mem] = (double)acc[r];
      }
    }
  }
}

// FP64 apply (the plain path): thread (i, half) accumulates member i of
// every other slab, x_a(i) = sum_j T[i][j] x_b(j); x_b(j) is the same
// address across the 128 threads of a slab (broadcast), T rows stay in L2
__global__ __launch_bounds__(256) void k_apply_mat_f64(LocDev L, StateDev S, VarsDev V,
                                                       ChunkDev C, DiagDev D,
                                                       const double* __restrict__ B,
                                                       double* __restrict__ ana) {
  __shared__ int sl[2048];
  __shared__ int snsl;
  const int64_t pair = blockIdx.x;
  if (!C.solve[pair]) return;
  const int k = (int)S.k;
  const int tid = threadIdx.x, lane = tid & 63;
  const int64_t t_gp = pair / C.G;
  const int g = (int)(pair - t_gp * C.G);
  const int64_t ij = C.i0 + t_gp;
  if (tid < 64) {
    const int mld = D.mld[ij];
    int nsl = 0;
    for (int s0 = 0; s0 < S.ns; s0 += 64) {
      const int s = s0 + lane;
      const bool in = s < S.ns && slab_group(L, V, s, mld) == g;
      const unsigned long long bal = __ballot(in);
      const int q = nsl + __popcll(bal & ((1ull << lane) - 1ull));
      if (in) sl[q] = s;
      nsl += __popcll(bal);
    }
    if (lane == 0) snsl = nsl;
  }
  __syncthreads();
  const int nsl = snsl;
  const int i = tid & 127, par = tid >> 7;
  if (i >= k) return;
  const double* Ti = B + pair * (int64_t)k * k + (int64_t)i * k;
  const float* xb0 = S.state + (int64_t)S.k * S.ns * ij;
  double* xa0 = ana + (int64_t)S.k * S.ns * ij;
  for (int q = par; q < nsl; q += 2) {
    const float* xp = xb0 + (int64_t)k * sl[q];
    double a0 = 0.0, a1 = 0.0;
    int j = 0;
    for (; j + 1 < k; j += 2) {
      a0 = fma(Ti[j], (double)xp[j], a0);
      a1 = fma(Ti[j + 1], (double)xp[j + 1], a1);
    }
    if (j < k) a0 = fma(Ti[j], (double)xp[j], a0);
    xa0[(int64_t)k * sl[q] + i] = a0 + a1;
  }
}

bool transform_mat_supported(int k) { return k >= 64 && k <= 128; }

void launch_transform_mat(hipStream_t st, int prec, const LocDev& L, const StateDev& S,
                          const VarsDev& V, const ChunkDev& C, const DiagDev& D, double* B,
                          const double* bvec, int* status, const RatTab& R, double* ana) {
  const int64_t np = C.ngp * C.G;
  if (np <= 0) return;
  const int k = (int)S.k;
  const unsigned nb = (unsigned)np;
  // rows per half: k <= 2 NR, NR even
  const int nr = ((k + 1) / 2 + 1) & ~1;
  if (nr <= 40) launch_transform_nr40(st, nb, k, C, B, bvec, status, R);
  else if (nr <= 48) launch_transform_nr48(st, nb, k, C, B, bvec, status, R);
  else if (nr <= 56) launch_transform_nr56(st, nb, k, C, B, bvec, status, R);
  else launch_transform_nr64(st, nb, k, C, B, bvec, status, R);
  if (prec == LETKF_PREC_FP64)
    k_apply_mat_f64<<<nb, 256, 0, st>>>(L, S, V, C, D, B, ana);
  else if (prec == LETKF_PREC_FP32)
    k_apply_mat<LETKF_PREC_FP32><<<nb, 256, 0, st>>>(L, S, V, C, D, B, ana);
  else if (prec == LETKF_PREC_BF16)
    k_apply_mat<LETKF_PREC_BF16><<<nb, 256, 0, st>>>(L, S, V, C, D, B, ana);
  else
    k_apply_mat<LETKF_PREC_BF16X3><<<nb, 256, 0, st>>>(L, S, V, C, D, B, ana);
}

}  // namespace lk
