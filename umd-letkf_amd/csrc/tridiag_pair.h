// tridiag_pair.h -- Householder tridiagonalization of two problems per wave
// (the template; instantiated per row length KQ in tridiag_pair_kq*.hip so
// that the fully unrolled kernels compile in parallel).
#pragma once

#include <hip/hip_runtime.h>

#include "eigen_common.h"
#include "letkf_internal.h"

namespace lk {

// ---------------------------------------------------------------------------
// Each 32-lane half of the wave reduces one problem: lane r of a half keeps
// rows r and r + 32 of A (a0[c] = A(r, c), a1[c] = A(r + 32, c); KQ >= k,
// k <= 64).  The reductions are the same as the one-problem kernel's, with
// every broadcast and sum taken per half:
//   - v and w of a step go through LDS, one array per half, and the
//     symmetric product (v pairs) and the rank-2 update ((v, w) pairs) read
//     them as 16-byte broadcasts within the half (the two halves read two
//     addresses of different banks in one instruction), so the update needs no
//     v_readlane: 4 FMAs per column serve two problems where the one-problem
//     kernel spends 2 FMAs and 4 readlanes on one;
//   - sums over a half are DPP (row shifts, row_bcast 15) with the half's
//     result read from its last lane.
// A step whose column is already reduced (sig = 0 in one half only) is made
// a no-op in that half by zero v, w and tau, so both halves stay on one
// instruction stream.  For steps j >= 31 rows 0..31 are final and their
// updates are skipped (j is a compile-time constant: the step loop is
// unrolled).  The problems are taken in the order perm (the batched path's
// bucket order, solved problems first, so both halves are busy) or in
// gridpoint order; a half past the end or on an unsolved problem computes
// on zeros and stores nothing.
// ---------------------------------------------------------------------------

// value of lane l of this wave's half (l = 0..31, a compile-time constant)
__device__ __forceinline__ double half_bcast(double v, int l, bool upper) {
  const int lo0 = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi0 = __builtin_amdgcn_readlane(__double2hiint(v), l);
  const int lo1 = __builtin_amdgcn_readlane(__double2loint(v), l + 32);
  const int hi1 = __builtin_amdgcn_readlane(__double2hiint(v), l + 32);
  return upper ? __hiloint2double(hi1, lo1) : __hiloint2double(hi0, lo0);
}

// DPP move of a double with no old value: lanes without a source read 0
// (bound_ctrl), lanes of rows the mask disables are left undefined
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_z(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, ROWMASK, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, ROWMASK, 0xf, true);
  return __hiloint2double(hi, lo);
}

// sum over this wave's half, in every lane of the half (whole wave active):
// row_shr 1, 2, 4, 8 leave each 16-lane row's sum in its lane 15, row_bcast
// 15 adds row 0 into row 1 and row 2 into row 3 (lanes 31 and 63 hold the
// halves' sums; rows 0 and 2 are left with values that are not read)
__device__ __forceinline__ double half_sum(double x, bool upper) {
  x += dpp_z<0x111, 0xf>(x);
  x += dpp_z<0x112, 0xf>(x);
  x += dpp_z<0x114, 0xf>(x);
  x += dpp_z<0x118, 0xf>(x);
  x += dpp_z<0x142, 0xa>(x);
  return half_bcast(x, 31, upper);
}

// largest value over this wave's half (x >= 0; whole wave active)
__device__ __forceinline__ double half_max(double x, bool upper) {
  x = fmax(x, dpp_z<0x111, 0xf>(x));
  x = fmax(x, dpp_z<0x112, 0xf>(x));
  x = fmax(x, dpp_z<0x114, 0xf>(x));
  x = fmax(x, dpp_z<0x118, 0xf>(x));
  x = fmax(x, dpp_z<0x142, 0xa>(x));
  return half_bcast(x, 31, upper);
}

// 1/x: the hardware estimate and one Newton step (the pivots of the
// rational solve are >= k - 1 > 0)
__device__ __forceinline__ double rcp_nr1(double x) {
  const double y = __builtin_amdgcn_rcp(x);
  return y * (2.0 - x * y);
}

// The rational solve of one problem per half (see k_tridiag_pair, RAT); all
// lanes of the wave take part (barrier, DPP sums), lanes of an inactive half
// or of a problem with more slabs than carried vectors store nothing.
template <int KQ, int NV>
__device__ __forceinline__ void rat_tail(int k, int K, bool act, int64_t pq, int nsl, int r,
                                         bool upper, double dg0, double dg1,
                                         const double (&xv0)[NV], const double (&xv1)[NV],
                                         double* s_d, double* s_e, double* s_u,
                                         double* __restrict__ U, const double* __restrict__ rtab,
                                         const int* __restrict__ rtabn, int rnb) {
  constexpr int LDH = 66;
  const int hh = upper ? 1 : 0, r1 = r + 32;
  double* sdh = s_d + hh * LDH;
  const double* seh = s_e + hh * LDH;
  sdh[r] = r < k ? dg0 : 0.0;
  sdh[r1] = r1 < k ? dg1 : 0.0;
#pragma unroll
  for (int qv = 0; qv < NV; ++qv) {
    s_u[(hh * NV + qv) * LDH + r] = xv0[qv];
    s_u[(hh * NV + qv) * LDH + r1] = xv1[qv];
  }
  __syncthreads();
  const bool valid = act && nsl <= NV - 1;
  // Gershgorin bound of the largest eigenvalue of T (SPD)
  double gm = 0.0;
  if (r < k) gm = sdh[r] + (r > 0 ? fabs(seh[r - 1]) : 0.0) + fabs(seh[r]);
  if (r1 < k) gm = fmax(gm, sdh[r1] + fabs(seh[r1 - 1]) + fabs(seh[r1]));
  gm = half_max(gm, upper);
  const double m = (double)(k - 1);
  const double kappa = gm / m;
  int b = (int)ceil(log2(fmax(kappa, 2.0)));
  b = b < 1 ? 1 : (b > rnb - 1 ? rnb - 1 : b);
  const int nn = rtabn[b];
  const bool node = r < 31;
  const double sigma = node ? (r < nn ? m * rtab[b * 64 + r] : m) : 0.0;
  const double coef = (node && r < nn) ? m * rtab[b * 64 + 32 + r] : 0.0;
  const int nv_h = valid ? nsl : 0;
  const int nsl_all = max(nv_h, __shfl_xor(nv_h, 32, 64));
  for (int q = 1; q <= nsl_all; ++q) {
    const bool actq = valid && q <= nsl;
    const double* ux = s_u + (hh * NV + (actq ? q : 0)) * LDH;
    const double* u = node ? ux : s_u + hh * NV * LDH;
    // LDL^T of T + sigma I with the forward substitution, then back
    double l[KQ], p[KQ];
    double rprev = 0.0, zprev = 0.0, eprev = 0.0;
#pragma unroll
    for (int i = 0; i < KQ; ++i) {
      if (i < k) {
        const double li = eprev * rprev;
        const double Di = sdh[i] + sigma - li * eprev;
        const double ri = rcp_nr1(Di);
        const double zi = u[i] - li * zprev;
        l[i] = li;
        p[i] = zi * ri;
        rprev = ri;
        zprev = zi;
        eprev = seh[i];
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
    // y_i = sum over the node lanes of coef * x_i, gathered by DPP in the
    // last lane of each half (the T^-1 b lane, whose coef is 0)
#pragma unroll
    for (int i = 0; i < KQ; ++i)
      if (i < k) p[i] = half_sum_dpp(coef * p[i]);
    if (actq && !node) {
      double* uo = U + (pq * NV + q) * K;
#pragma unroll
      for (int i = 0; i < KQ; ++i)
        if (i < k) uo[i] = p[i];
      uo[K - 1] = wx;
    }
  }
}

// RAT: the rational solve of the batched vector path runs in the same
// kernel after the reduction (T, Q^T b and Q^T x_b of each half are in
// registers and go through LDS), in the layout of k_ratsolve: lane j < 31
// of a half solves (T + sigma_j I) x = Q^T x_b by LDL^T, lane 31 solves
// T x = Q^T b, and DPP sums over the half form y = sqrt(k-1) T^(-1/2) Q^T x_b
// and wbar . x_b, written to U for k_qapply (T, Q^T b and the inputs stay
// on chip: no dd, ee or input U round trip, no separate launch).
template <int KQ, int NV, bool RAT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
void k_tridiag_pair(int k, int K, int64_t np, const int32_t* __restrict__ perm, LocDev L,
                    StateDev S, VarsDev V, ChunkDev C, DiagDev D, double* __restrict__ B,
                    const double* __restrict__ bvec, double* __restrict__ tau,
                    double* __restrict__ dd, double* __restrict__ ee, double* __restrict__ U,
                    int* __restrict__ nvec, int* __restrict__ status,
                    const double* __restrict__ rtab, const int* __restrict__ rtabn, int rnb) {
  // per half: v and w of rows 0..63; the halves' arrays 66 entries apart
  // (a 16-byte read of each half falls in different banks)
  constexpr int LDH = 66;
  // columns between compiler barriers in the symmetric product and the
  // rank-2 update (fewer for long rows: registers)
  constexpr int GS = KQ >= 48 ? 4 : 8;
  constexpr int GU = KQ >= 48 ? 2 : 4;
  __shared__ double sv[2 * LDH];        // v, for the symmetric product
  __shared__ double2 svw[2 * LDH];      // (v, w), for the rank-2 update
  __shared__ int sl[2][NV > 0 ? NV : 1];
  // RAT: T (diagonal, off-diagonal) and the carried vectors of each half
  __shared__ double s_d[RAT ? 2 * LDH : 1], s_e[RAT ? 2 * LDH : 1];
  __shared__ double s_u[RAT ? 2 * NV * LDH : 1];
  const int lane = threadIdx.x;
  const bool upper = lane >= 32;
  const int hh = upper ? 1 : 0;
  const int r = lane & 31, r1 = r + 32;
  const int64_t q = 2 * (int64_t)blockIdx.x + hh;
  int64_t pair = -1;
  if (q < np) pair = perm ? (int64_t)perm[q] : q;
  const bool act = pair >= 0 && C.solve[pair] != 0;
  if (__ballot(act) == 0ull) return;   // wave-uniform
  const int64_t pq = act ? pair : 0;
  const double* Bp = B + pq * (int64_t)k * k;
  double* Bw = B + pq * (int64_t)k * k;
  double a0[KQ], a1[KQ];
#pragma unroll
  for (int c = 0; c < KQ; ++c) {
    const double u0 = (act && r < k && c < k) ? Bp[r + (int64_t)k * c] : 0.0;
    const double u1 = (act && r1 < k && c < k) ? Bp[r1 + (int64_t)k * c] : 0.0;
    a0[c] = (act && c == r && r < k) ? u0 + (double)(k - 1) : u0;
    a1[c] = (act && c == r1 && r1 < k) ? u1 + (double)(k - 1) : u1;
  }
  double xv0[NV > 0 ? NV : 1], xv1[NV > 0 ? NV : 1];
  int nsl = 0;
  if (NV > 0) {
    // slabs of each half's group in ascending order (ballot compaction
    // within the half)
    const int64_t t_gp = pq / C.G;
    const int g = (int)(pq - t_gp * C.G);
    const int64_t ij = C.i0 + t_gp;
    const int mld = act ? D.mld[ij] : 0;
    for (int s0 = 0; s0 < S.ns; s0 += 32) {
      const int s = s0 + r;
      const bool in = act && s < S.ns && slab_group(L, V, s, mld) == g;
      const unsigned long long bal = __ballot(in);
      const uint32_t hb = (uint32_t)(bal >> (32 * hh));
      const uint32_t below = r == 0 ? 0u : (hb & (0xffffffffu >> (32 - r)));
      const int qq = nsl + __popc(below);
      if (in && qq < NV - 1) sl[hh][qq] = s;
      nsl += __popc(hb);
    }
    __syncthreads();
    xv0[0] = (act && r < k) ? bvec[pq * (int64_t)k + r] : 0.0;
    xv1[0] = (act && r1 < k) ? bvec[pq * (int64_t)k + r1] : 0.0;
    const float* xb0 = S.state + (int64_t)S.k * S.ns * ij;
#pragma unroll
    for (int qv = 1; qv < NV; ++qv) {
      const bool have = act && qv <= nsl;
      const int64_t so = have ? (int64_t)k * sl[hh][qv - 1] : 0;
      xv0[qv] = (have && r < k) ? (double)xb0[so + r] : 0.0;
      xv1[qv] = (have && r1 < k) ? (double)xb0[so + r1] : 0.0;
    }
  }
  double* taup = tau + pq * K;
  double* ep = ee + pq * K;
  if (RAT) {   // off-diagonal entries past the last step stay 0
    s_e[hh * LDH + r] = 0.0;
    s_e[hh * LDH + r1] = 0.0;
  }
  double* svh = sv + hh * LDH;
  double2* svwh = svw + hh * LDH;
#pragma unroll
  for (int j = 0; j < KQ - 1; ++j) {
    // static j: every a0[] / a1[] index below is a compile-time constant
    if (j <= k - 2) {
      const bool lo_live = j < 31;   // rows 0..31 still have rows below j
      const double x0 = (r > j && r < k) ? a0[j] : 0.0;
      const double x1 = (r1 > j && r1 < k) ? a1[j] : 0.0;
      const double alpha = half_bcast(j + 1 < 32 ? a0[j] : a1[j], (j + 1) & 31, upper);
      const double sig = half_sum(((r > j + 1 && r < k) ? x0 * x0 : 0.0) +
                                      ((r1 > j + 1 && r1 < k) ? x1 * x1 : 0.0), upper);
      const bool nz = sig != 0.0;
      const double mu = sqrt(alpha * alpha + sig);
      const double beta = alpha >= 0.0 ? -mu : mu;
      // sig = 0: H_j = I (tau 0, v 0); the divisions' results are discarded
      const double t = nz ? (beta - alpha) / beta : 0.0;
      const double scal = nz ? 1.0 / (alpha - beta) : 0.0;
      const double v0 = !nz ? 0.0 : (r == j + 1 ? 1.0 : ((r > j + 1 && r < k) ? x0 * scal : 0.0));
      const double v1 = !nz ? 0.0 : (r1 == j + 1 ? 1.0 : ((r1 > j + 1 && r1 < k) ? x1 * scal : 0.0));
      svh[r] = v0;
      svh[r1] = v1;
      svwh[r].x = v0;
      svwh[r1].x = v1;
      if (NV > 0) {
        // x <- H_j x for the carried vectors
#pragma unroll
        for (int qv = 0; qv < NV; ++qv) {
          const double dq = half_sum(v0 * xv0[qv] + v1 * xv1[qv], upper);
          xv0[qv] -= (t * dq) * v0;
          xv1[qv] -= (t * dq) * v1;
        }
      }
      __syncthreads();
      // p = t A v over the columns right of j
      double p0a = 0.0, p0b = 0.0, p1a = 0.0, p1b = 0.0;
      // each column's v is read one column ahead of its use
      double vnx = svh[j + 1];
#pragma unroll
      for (int c = j + 1; c < KQ; ++c) {
        const double vc = vnx;
        if (c + 1 < KQ) vnx = svh[c + 1];
        if (c & 1) {
          if (lo_live) p0b = fma(a0[c], vc, p0b);
          p1b = fma(a1[c], vc, p1b);
        } else {
          if (lo_live) p0a = fma(a0[c], vc, p0a);
          p1a = fma(a1[c], vc, p1a);
        }
        // bound how far the LDS reads are hoisted (their registers): the
        // sums are pinned at a compiler barrier every GS columns
        if ((c & (GS - 1)) == GS - 1)
          asm volatile("" : "+v"(p0a), "+v"(p0b), "+v"(p1a), "+v"(p1b) :: "memory");
      }
      const double p0 = (lo_live && r > j && r < k) ? t * (p0a + p0b) : 0.0;
      const double p1 = (r1 > j && r1 < k) ? t * (p1a + p1b) : 0.0;
      const double kc = 0.5 * t * half_sum(p0 * v0 + p1 * v1, upper);
      const double w0 = p0 - kc * v0;
      const double w1 = p1 - kc * v1;
      svwh[r].y = w0;
      svwh[r1].y = w1;
      __syncthreads();
      // rank-2 update A -= v w^T + w v^T on the columns right of j
      // each column's (v, w) is read one column ahead of its use
      double2 vwnx = svwh[j + 1];
#pragma unroll
      for (int c = j + 1; c < KQ; ++c) {
        const double2 vw = vwnx;
        if (c + 1 < KQ) vwnx = svwh[c + 1];
        const double vc = vw.x, wc = vw.y;
        if (lo_live) a0[c] = fma(-w0, vc, fma(-v0, wc, a0[c]));
        a1[c] = fma(-w1, vc, fma(-v1, wc, a1[c]));
        // each updated column is pinned where it is computed, and every GU
        // columns at a compiler barrier for memory, so that the LDS reads
        // are not all hoisted ahead of the FMAs
        if (lo_live) {
          if ((c & (GU - 1)) == GU - 1) asm volatile("" : "+v"(a0[c]), "+v"(a1[c]) :: "memory");
          else asm volatile("" : "+v"(a0[c]), "+v"(a1[c]));
        } else {
          if ((c & (GU - 1)) == GU - 1) asm volatile("" : "+v"(a1[c]) :: "memory");
          else asm volatile("" : "+v"(a1[c]));
        }
      }
      if (act && nz) {   // reflector j, rows j+1..k-1
        if (r > j && r < k) Bw[(int64_t)j * k + r] = v0;
        if (r1 > j && r1 < k) Bw[(int64_t)j * k + r1] = v1;
      }
      if (act && r == 0) {
        taup[j] = t;
        if (!RAT) ep[j] = nz ? beta : alpha;
      }
      if (RAT && r == 0) s_e[hh * LDH + j] = nz ? beta : alpha;
      __syncthreads();
    }
    // keep the scheduler from interleaving steps (register pressure)
    __builtin_amdgcn_sched_barrier(0);
  }
  double dg0 = 0.0, dg1 = 0.0;
#pragma unroll
  for (int c = 0; c < KQ; ++c) {
    dg0 = (c == r) ? a0[c] : dg0;
    dg1 = (c == r1) ? a1[c] : dg1;
  }
  if constexpr (RAT) {
    rat_tail<KQ, NV>(k, K, act, pq, nsl, r, upper, dg0, dg1, xv0, xv1, s_d, s_e, s_u, U, rtab,
                     rtabn, rnb);
  }
  if (!act) return;
  if (r < K) {
    if (!RAT) dd[pq * K + r] = r < k ? dg0 : 0.0;
    if (r >= k - 1) {
      if (!RAT) ep[r] = 0.0;
      taup[r] = 0.0;
    }
  }
  if (r1 < K) {
    if (!RAT) dd[pq * K + r1] = r1 < k ? dg1 : 0.0;
    if (r1 >= k - 1) {
      if (!RAT) ep[r1] = 0.0;
      taup[r1] = 0.0;
    }
  }
  if (NV > 0 && !RAT) {
    double* Up = U + pq * (int64_t)NV * K;
#pragma unroll
    for (int qv = 0; qv < NV; ++qv) {
      if (r < K) Up[qv * K + r] = xv0[qv];
      if (r1 < K) Up[qv * K + r1] = xv1[qv];
    }
  }
  if (NV > 0) {
    if (r == 0) {
      nvec[pq] = nsl;
      status[pq] = nsl > NV - 1 ? 3 : 0;   // more slabs than carried: Jacobi
    }
  } else if (r == 0) {
    status[pq] = 0;
  }
}

template <int KQ>
int launch_tridiag_pair_kq(hipStream_t st, int64_t np, const int32_t* perm, int NV, int k, int K,
                           const LocDev& L, const StateDev& S, const VarsDev& V,
                           const ChunkDev& C, const DiagDev& D, double* B, const double* bvec,
                           double* tau, double* dd, double* ee, double* U, int* nvec,
                           int* status, const double* rtab, const int* rtabn, int rnb) {
  const unsigned nb = (unsigned)((np + 1) / 2);
  const bool rat = rnb > 0;
  switch (NV) {
    case 0: k_tridiag_pair<KQ, 0, false><<<nb, 64, 0, st>>>(k, K, np, perm, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status, rtab, rtabn, rnb); return 1;
    case 2:
      if (rat) { k_tridiag_pair<KQ, 2, true><<<nb, 64, 0, st>>>(k, K, np, perm, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status, rtab, rtabn, rnb); return 2; }
      k_tridiag_pair<KQ, 2, false><<<nb, 64, 0, st>>>(k, K, np, perm, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status, rtab, rtabn, rnb); return 1;
    case 3:
      if (rat) { k_tridiag_pair<KQ, 3, true><<<nb, 64, 0, st>>>(k, K, np, perm, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status, rtab, rtabn, rnb); return 2; }
      k_tridiag_pair<KQ, 3, false><<<nb, 64, 0, st>>>(k, K, np, perm, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status, rtab, rtabn, rnb); return 1;
    default: return 0;   // more carried vectors: too few registers for two problems
  }
}

}  // namespace lk
