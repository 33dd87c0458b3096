// transform_blk.h -- the reduction and the back-transform of the explicit
// transform matrix (64 <= k <= 128, see transform_mat.h) with A distributed
// over a 256-thread block in 2D: the matrix, padded to N = 16 BS, is a 16 x 16
// grid of BS x BS blocks, and thread t holds block (I, J) = (t / 16, t % 16)
// in registers.  Per step a thread needs v and w only for its BS rows and
// its BS columns (4 BS values from LDS), where the column layout of
// transform_mat.h needs v and w for all of its NR rows, broadcast to every
// thread of the block: the LDS traffic per step falls from about 24 NR to
// about 4 BS doubles per thread.  The symmetric product's row sums are DPP
// sums over the 16 threads of a grid row (consecutive lanes, one DPP row),
// scattered so that each lane finishes one row.  The reduction takes two
// barriers per step: after the column is in LDS, and after p = t A v and the
// partial sums of v.p are; the back-transform, whose reflectors are known in
// advance, only the second (its LDS arrays double buffered by step).  In the
// reduction the column goes to LDS with its
// reduced rows zeroed and the wholly reduced waves keep their p entries at
// zero, so v = x sc and w = p - kc v need no row tests.
//
//  k_blk_tridiag<BS>  A = Q Tri Q^T (LAPACK dsytd2, lower), the outputs of
//                     k_mat_tridiag: reflector j in row j of B (columns > j),
//                     tau, d and e.  Waves whose rows are all reduced skip
//                     the arithmetic of a step.
//  k_blk_back<BS>     W = Q M Q^T from M in the lower triangle of B, then
//                     wbar = W (W b) / (k-1) and T = W + 1 wbar^T over B, the
//                     outputs of k_mat_back.
#pragma once

#include <hip/hip_runtime.h>

#include "eigen_common.h"
#include "letkf_internal.h"

namespace lk {

constexpr int kBlkStride = 128;   // d, e, tau entries per problem (as kMatStride)

// DPP move of a double with no old value: lanes without a source read 0
template <int CTRL>
__device__ __forceinline__ double blk_dpp(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}

// sum over the 16 lanes of this lane's DPP row, complete in the row's lane 15
__device__ __forceinline__ double row16_sum(double x) {
  x += blk_dpp<0x111>(x);
  x += blk_dpp<0x112>(x);
  x += blk_dpp<0x114>(x);
  x += blk_dpp<0x118>(x);
  return x;
}

// sums of BS <= 8 values over the 16 lanes of this lane's DPP row, scattered:
// lane J of the row (J = lane % 16) ends with the whole sum of value J / 2
// (lanes J and J ^ 1 both hold it; values past BS are zero).  Four exchange
// stages, each with the lane at distance 8 (row_ror 8), at mirrored position
// in the half row (row_half_mirror), at distance 2 and 1 (quad_perm): each
// lane keeps half of the values it carries and sends the other half, so the
// stages move 4, 2, 1 and 1 values where four shift-and-add stages per value
// move 4 BS.  The whole wave must be active.
template <int BS>
__device__ __forceinline__ double row16_scatter_sums(const double (&s)[BS], int J) {
  static_assert(BS <= 8, "at most 8 values");
  const bool h1 = (J & 8) != 0, h2 = (J & 4) != 0, h3 = (J & 2) != 0;
  double t[4];
#pragma unroll
  for (int x = 0; x < 4; ++x) {
    const double a = s[x], b = x + 4 < BS ? s[x + 4 < BS ? x + 4 : 0] : 0.0;
    const double keep = h1 ? b : a, send = h1 ? a : b;
    t[x] = keep + blk_dpp<0x128>(send);
  }
  double u[2];
#pragma unroll
  for (int x = 0; x < 2; ++x) {
    const double keep = h2 ? t[x + 2] : t[x], send = h2 ? t[x] : t[x + 2];
    u[x] = keep + blk_dpp<0x141>(send);
  }
  const double keep = h3 ? u[1] : u[0], send = h3 ? u[0] : u[1];
  const double w = keep + blk_dpp<0x4E>(send);
  return w + blk_dpp<0xB1>(w);
}

template <int BS>
__global__ __launch_bounds__(256) void k_blk_tridiag(int k, ChunkDev C, double* __restrict__ B,
                                                     double* __restrict__ tau,
                                                     double* __restrict__ dd,
                                                     double* __restrict__ ee) {
  static_assert(16 * BS <= 128, "at most 128 rows");
  __shared__ double sx[128];        // column j of A
  __shared__ double sp[128];        // p = t A v
  __shared__ double skc[4];         // per-wave sums of v . p
  __shared__ double se[128], stau[128];
  const int64_t pair = blockIdx.x;
  if (!C.solve[pair]) return;
  const int tid = threadIdx.x, l = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int I = tid >> 4, J = tid & 15;
  const int R0 = BS * I, C0 = BS * J;
  const double m = (double)(k - 1);
  double* Bp = B + pair * (int64_t)k * k;
  // every load is issued (clamped index) before any value is used, so the
  // loads are in flight together; padding rows and columns are zeroed after
  double a[BS][BS];
#pragma unroll
  for (int x = 0; x < BS; ++x)
#pragma unroll
    for (int y = 0; y < BS; ++y) {
      const int r = min(R0 + x, k - 1), c = min(C0 + y, k - 1);
      a[x][y] = Bp[(int64_t)r * k + c];
    }
#pragma unroll
  for (int x = 0; x < BS; ++x)
#pragma unroll
    for (int y = 0; y < BS; ++y) {
      const int r = R0 + x, c = C0 + y;
      const double v = (r < k && c < k) ? a[x][y] : 0.0;
      a[x][y] = (r == c && r < k) ? v + m : v;
    }
  if (tid < 128) { sx[tid] = 0.0; sp[tid] = 0.0; se[tid] = 0.0; stau[tid] = 0.0; }
  __syncthreads();
  const int jend = k - 2;
  // last row of this wave's grid rows (wave-uniform)
  const int wlast = 4 * BS * wv + 4 * BS - 1;
  for (int jb = 0; jb * BS < jend; ++jb) {
#pragma unroll
    for (int bj = 0; bj < BS; ++bj) {
      const int j = jb * BS + bj;
      if (j >= jend) break;
      const int j1 = j + 1;
      // column j (local column bj of grid column jb) into LDS, zero in the
      // rows already reduced (r <= j)
      if (J == jb) {
#pragma unroll
        for (int x = 0; x < BS; ++x) sx[R0 + x] = R0 + x > j ? a[x][bj] : 0.0;
      }
      __syncthreads();
      // the reflector, the same in every wave
      const double x0 = sx[l], x1 = sx[l + 64];
      double q = 0.0;
      if (l > j1 && l < k) q += x0 * x0;
      if (l + 64 > j1 && l + 64 < k) q += x1 * x1;
      const double sig = wave_sum_dpp(q);
      const double alpha = sx[j1];
      // column already reduced (sig = 0): tau = 0, and the step changes nothing
      const bool refl = sig != 0.0;
      const double mu = sqrt(alpha * alpha + sig);
      const double beta = refl ? (alpha >= 0.0 ? -mu : mu) : alpha;
      const double t = refl ? (beta - alpha) / beta : 0.0;
      const double sc = refl ? 1.0 / (alpha - beta) : 0.0;
      // v = x sc below row j+1 (the column is zero in rows <= j) and 1 in
      // row j+1, which is local row (bj + 1) % BS of grid row jb + (bj + 1) / BS:
      // a register index known at compile time, so one select per thread
      const int xj1 = (bj + 1) % BS;
      const int gj1 = jb + (bj + 1) / BS;
      double vr[BS], vc[BS];
#pragma unroll
      for (int x = 0; x < BS; ++x) {
        vr[x] = sx[R0 + x] * sc;
        vc[x] = sx[C0 + x] * sc;
        if (x == xj1) {   // folded once the step loop is unrolled
          vr[x] = I == gj1 ? 1.0 : vr[x];
          vc[x] = J == gj1 ? 1.0 : vc[x];
        }
      }
      const bool live = wlast > j;   // this wave still has rows below j
      if (live) {
        // row sums of this block's part of A v, completed over the grid row:
        // lane J holds row R0 + J / 2 (the even lane writes it)
        double P[BS];
#pragma unroll
        for (int x = 0; x < BS; ++x) {
          double s0 = 0.0;
#pragma unroll
          for (int y = 0; y < BS; ++y) s0 = fma(a[x][y], vc[y], s0);
          P[x] = s0;
        }
        const double pz = row16_scatter_sums<BS>(P, J);
        const int xr = J >> 1, r = R0 + xr;
        const bool own = (J & 1) == 0 && xr < BS;
        const double px = r > j ? t * pz : 0.0;
        if (own) sp[r] = px;
        const double vrow = r == j1 ? 1.0 : sx[min(r, 127)] * sc;
        const double ws = wave_sum_dpp(own ? px * vrow : 0.0);
        if (l == 0) skc[wv] = ws;
      } else {
        // every row of this wave is reduced: its p entries stay zero, so
        // that w = p - kc v needs no row test
        if (J == 15) {
#pragma unroll
          for (int x = 0; x < BS; ++x) sp[R0 + x] = 0.0;
        }
        if (l == 0) skc[wv] = 0.0;
      }
      // reflector j (columns > j), one entry per thread of the first two
      // waves (one store each instead of BS predicated stores in the first
      // grid row), and tau, e
      if (tid > j && tid < k) Bp[(int64_t)j * k + tid] = tid == j1 ? 1.0 : sx[tid] * sc;
      if (tid == 0) { stau[j] = t; se[j] = beta; }
      __syncthreads();
      if (live) {
        const double kc = 0.5 * t * ((skc[0] + skc[1]) + (skc[2] + skc[3]));
        // p and v are zero in the reduced rows, and so is w
        double wr[BS], wc[BS];
#pragma unroll
        for (int x = 0; x < BS; ++x) {
          wr[x] = fma(-kc, vr[x], sp[R0 + x]);
          wc[x] = fma(-kc, vc[x], sp[C0 + x]);
        }
#pragma unroll
        for (int x = 0; x < BS; ++x)
#pragma unroll
          for (int y = 0; y < BS; ++y) a[x][y] = fma(-vr[x], wc[y], fma(-wr[x], vc[y], a[x][y]));
      }
    }
  }
  __syncthreads();
  double* dp = dd + pair * kBlkStride;
  double* ep = ee + pair * kBlkStride;
  if (I == J) {
#pragma unroll
    for (int x = 0; x < BS; ++x)
      if (R0 + x < k) dp[R0 + x] = a[x][x];
  }
  // the last subdiagonal entry A[k-1][k-2]
#pragma unroll
  for (int x = 0; x < BS; ++x)
#pragma unroll
    for (int y = 0; y < BS; ++y)
      if (R0 + x == k - 1 && C0 + y == k - 2) ep[k - 2] = a[x][y];
  if (tid < jend) {
    ep[tid] = se[tid];
    tau[pair * kBlkStride + tid] = stau[tid];
  }
}

template <int BS>
__global__ __launch_bounds__(256) void k_blk_back(int k, ChunkDev C, double* __restrict__ B,
                                                  const double* __restrict__ bvec,
                                                  const double* __restrict__ tau) {
  static_assert(16 * BS <= 128, "at most 128 rows");
  // one barrier per step: the reflectors, p and the wave sums are double
  // buffered by step, and the next step's reflector goes to LDS during this
  // step's product
  __shared__ double sv[2][128];     // reflectors of this step and the next
  __shared__ double sp[2][128];     // p = t A v; then W b, W (W b)
  __shared__ double skc[2][4];
  __shared__ double stau[128];
  const int64_t pair = blockIdx.x;
  if (!C.solve[pair]) return;
  const int tid = threadIdx.x, l = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int I = tid >> 4, J = tid & 15;
  const int R0 = BS * I, C0 = BS * J;
  const double m = (double)(k - 1);
  double* Bp = B + pair * (int64_t)k * k;
  // M from the lower triangle of B (symmetric); all loads issued first
  double a[BS][BS];
#pragma unroll
  for (int x = 0; x < BS; ++x)
#pragma unroll
    for (int y = 0; y < BS; ++y) {
      const int r = min(R0 + x, k - 1), c = min(C0 + y, k - 1);
      a[x][y] = Bp[r >= c ? (int64_t)r * k + c : (int64_t)c * k + r];
    }
#pragma unroll
  for (int x = 0; x < BS; ++x)
#pragma unroll
    for (int y = 0; y < BS; ++y)
      if (R0 + x >= k || C0 + y >= k) a[x][y] = 0.0;
  if (tid < 128) {
    stau[tid] = tid < k - 2 ? tau[pair * kBlkStride + tid] : 0.0;
    sp[0][tid] = 0.0;
    sp[1][tid] = 0.0;
  }
  // reflector rows go through LDS: reflector k-3 now, and threads 0..127
  // hold entry tid of the one after next (loaded two steps ahead)
  double vnext = 0.0;
  if (tid < 128) {
    const int j = k - 3;
    sv[j & 1][tid] = (tid > j && tid < k) ? Bp[(int64_t)j * k + tid] : 0.0;
    sv[(j + 1) & 1][tid] = 0.0;
    const int jn = j - 1;
    if (jn >= 0) vnext = (tid > jn && tid < k) ? Bp[(int64_t)jn * k + tid] : 0.0;
  }
  __syncthreads();
  for (int j = k - 3; j >= 0; --j) {
    const int buf = j & 1;
    const double t = stau[j];
    double vr[BS], vc[BS];
#pragma unroll
    for (int x = 0; x < BS; ++x) {
      vr[x] = sv[buf][R0 + x];
      vc[x] = sv[buf][C0 + x];
    }
    double P[BS];
#pragma unroll
    for (int x = 0; x < BS; ++x) {
      double s0 = 0.0;
#pragma unroll
      for (int y = 0; y < BS; ++y) s0 = fma(a[x][y], vc[y], s0);
      P[x] = s0;
    }
    // lane J holds row R0 + J / 2 of A v (the even lane writes it)
    const double pz = row16_scatter_sums<BS>(P, J);
    const int xr = J >> 1, r = R0 + xr;
    const bool own = (J & 1) == 0 && xr < BS;
    const double px = t * pz;
    if (own) sp[buf][r] = px;
    const double ws = wave_sum_dpp(own ? px * sv[buf][min(r, 127)] : 0.0);
    if (l == 0) skc[buf][wv] = ws;
    // reflector j-1 into the other buffer (its last reader, step j+1, has
    // passed the barrier of step j+1); reflector j-2 loaded
    if (tid < 128 && j > 0) {
      sv[buf ^ 1][tid] = vnext;
      const int jn = j - 2;
      if (jn >= 0) vnext = (tid > jn && tid < k) ? Bp[(int64_t)jn * k + tid] : 0.0;
    }
    __syncthreads();
    const double kc = 0.5 * t * ((skc[buf][0] + skc[buf][1]) + (skc[buf][2] + skc[buf][3]));
    double wr[BS], wc[BS];
#pragma unroll
    for (int x = 0; x < BS; ++x) {
      wr[x] = sp[buf][R0 + x] - kc * vr[x];
      wc[x] = sp[buf][C0 + x] - kc * vc[x];
    }
#pragma unroll
    for (int x = 0; x < BS; ++x)
#pragma unroll
      for (int y = 0; y < BS; ++y) a[x][y] = fma(-vr[x], wc[y], fma(-wr[x], vc[y], a[x][y]));
  }
  // wbar = W (W b) / (k-1); T = W + 1 wbar^T, row major over B
  __syncthreads();
  if (tid < 128) sv[0][tid] = tid < k ? bvec[pair * (int64_t)k + tid] : 0.0;
  __syncthreads();
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    double vc[BS];
#pragma unroll
    for (int y = 0; y < BS; ++y) vc[y] = sv[pass][C0 + y];
#pragma unroll
    for (int x = 0; x < BS; ++x) {
      double s0 = 0.0;
#pragma unroll
      for (int y = 0; y < BS; ++y) s0 = fma(a[x][y], vc[y], s0);
      s0 = row16_sum(s0);
      if (J == 15) {
        if (pass == 0) sv[1][R0 + x] = s0;
        else sp[0][R0 + x] = s0 / m;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < BS; ++x)
#pragma unroll
    for (int y = 0; y < BS; ++y) {
      const int r = R0 + x, c = C0 + y;
      if (r < k && c < k) Bp[(int64_t)r * k + c] = a[x][y] + sp[0][c];
    }
}

template <int BS>
void launch_blk_tridiag_bs(hipStream_t st, unsigned nb, int k, const ChunkDev& C, double* B,
                           double* tau, double* dd, double* ee) {
  k_blk_tridiag<BS><<<nb, 256, 0, st>>>(k, C, B, tau, dd, ee);
}
template <int BS>
void launch_blk_back_bs(hipStream_t st, unsigned nb, int k, const ChunkDev& C, double* B,
                        const double* bvec, const double* tau) {
  k_blk_back<BS><<<nb, 256, 0, st>>>(k, C, B, bvec, tau);
}

}  // namespace lk
