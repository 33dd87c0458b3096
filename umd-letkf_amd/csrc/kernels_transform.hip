// kernels_transform.hip -- the large-ensemble transform (64 <= k <= 128):
// the rational step of the explicit transform matrix and the application of
// T to every slab of the group (src/letkf_solver.F90:284-290):
//
//  k_mat_rational  the lower triangle of M = sqrt(k-1) Tri^(-1/2) (see
//                  transform_mat.h for the three-kernel scheme);
//  k_apply_mat     one 256-thread block per problem; wave w computes members
//                  32w..32w+31 of x_a = T x_b for 32 slabs at a time on
//                  MFMA (32x32x2 FP32, or 32x32x16 BF16 / bf16x3 with FP32
//                  accumulation), post-processes them (FIN) and writes the
//                  FP64 analysis;
//  k_apply_mat_f64 the same product in FP64 on the vector units (the plain
//                  path).
#include <hip/hip_runtime.h>

#include "letkf_internal.h"
#include "mx_common.h"
#include "transform_mat.h"

namespace lk {

// k_mat_tridiag / k_mat_back launchers (one file per NR)
#define LK_NR(n)                                                                          \
  void launch_mat_tridiag_nr##n(hipStream_t st, unsigned nb, int k, const ChunkDev& C,    \
                                double* B, double* tau, double* dd, double* ee);          \
  void launch_mat_back_nr##n(hipStream_t st, unsigned nb, int k, const ChunkDev& C,       \
                             double* B, const double* bvec, const double* tau);
LK_NR(40) LK_NR(48) LK_NR(56) LK_NR(64)
#undef LK_NR
// k_blk_tridiag / k_blk_back launchers (one file per block size)
#define LK_BS(n)                                                                          \
  void launch_blk_tridiag_bs##n(hipStream_t st, unsigned nb, int k, const ChunkDev& C,    \
                                double* B, double* tau, double* dd, double* ee);          \
  void launch_blk_back_bs##n(hipStream_t st, unsigned nb, int k, const ChunkDev& C,       \
                             double* B, const double* bvec, const double* tau);
LK_BS(4) LK_BS(5) LK_BS(6) LK_BS(7) LK_BS(8)
#undef LK_BS

// M below the diagonal, one thread per column c; the rows are visited in the
// same order by every lane (x starts at g_c in row c, is 0 above it), so the
// pivot reads are broadcasts and the stores of a row are contiguous
constexpr int kMaxNodes = 32;
constexpr int kNodeLd = 129;        // LDS row length of the node arrays

// a wave-uniform double in scalar registers
__device__ __forceinline__ double uniform_f64(double v) {
  const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
  const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
  return __hiloint2double(hi, lo);
}

// rows rfirst..k-1 of column c of M for the first NN nodes (nn <= NN).  x
// starts at g_c (the inverse's diagonal) and follows x <- -U_r x from row
// c + 1 on; rows above c are not stored.  The node weights are uniform
// (scalar registers), and each row's U values are read 8 at a time before
// they are used, all lanes reading the same entries.
template <int NN>
__device__ __forceinline__ void mat_rational_rows(int k, int c, int rfirst, int nn, double m,
                                                  const double* __restrict__ wn,
                                                  const double (*Us)[kNodeLd],
                                                  const double (*Gs)[kNodeLd], double* Bp) {
  double cf[NN], x[NN];
#pragma unroll
  for (int n = 0; n < NN; ++n) {
    cf[n] = uniform_f64(n < nn ? m * wn[n] : 0.0);
    x[n] = Gs[n][c];
  }
  for (int r = rfirst; r < k; ++r) {
    // four partial sums: the node chains are independent, the sum need not
    // be one dependent chain
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int n0 = 0; n0 < NN; n0 += 8) {
      double u[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) u[e] = Us[n0 + e][r];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int n = n0 + e;
        double xn = -u[e] * x[n];
        asm volatile("" : "+v"(xn));   // computed on every lane, then selected
        x[n] = r > c ? xn : x[n];
        acc[n & 3] = fma(cf[n], x[n], acc[n & 3]);
      }
    }
    if (r >= c) Bp[(int64_t)r * k + c] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  }
}

// 1/x: the hardware estimate and one Newton step, which squares its
// relative error (as the vector path's rational solve)
__device__ __forceinline__ double mat_rcp(double x) {
  const double y = __builtin_amdgcn_rcp(x);
  return y * (2.0 - x * y);
}
// One launch per class of node counts NNLO < nn <= NNL: the node arrays are
// sized for the class, so the problems that need few nodes (the better
// conditioned ones) run four blocks per CU (16 nodes, 35 KB of LDS) or three
// (24 nodes) where 32 nodes allow two; a problem of another class leaves once
// its node count is known (uniform over the block).  Each problem takes the
// same steps in its class's launch as in a single launch sized for 32 nodes.
template <int NNL, int NNLO>
__global__ __launch_bounds__(128) void k_mat_rational(int k, ChunkDev C, double* __restrict__ B,
                                                      const double* __restrict__ dd,
                                                      const double* __restrict__ ee,
                                                      const double* __restrict__ tab,
                                                      const int* __restrict__ tabn,
                                                      int nbucket) {
  static_assert(NNL % 8 == 0 && NNL <= kMaxNodes, "node class");
  __shared__ double sd[128], se[128];
  // rows of 129: lane n of the pivot recurrences writes row n, and the
  // lanes' entries then fall in different banks (with 128, all in one)
  __shared__ double Us[NNL][kNodeLd], Gs[NNL][kNodeLd];
  __shared__ double sgm[2];
  const int64_t pair = blockIdx.x;
  if (!C.solve[pair]) return;
  const int tid = threadIdx.x, l = tid & 63;
  const double m = (double)(k - 1);
  sd[tid] = tid < k ? dd[pair * kMatStride + tid] : 0.0;
  se[tid] = tid < k - 1 ? ee[pair * kMatStride + tid] : 0.0;
  __syncthreads();
  double gm = 0.0;
  if (tid < k)
    gm = sd[tid] + (tid > 0 ? fabs(se[tid - 1]) : 0.0) + (tid < k - 1 ? fabs(se[tid]) : 0.0);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) gm = fmax(gm, __shfl_xor(gm, o, 64));
  if (l == 0) sgm[tid >> 6] = gm;
  __syncthreads();
  gm = fmax(sgm[0], sgm[1]);
  int bk = (int)ceil(log2(fmax(gm / m, 2.0)));
  bk = bk < 1 ? 1 : (bk > nbucket - 1 ? nbucket - 1 : bk);
  bk = __builtin_amdgcn_readfirstlane(bk);
  const int nn = __builtin_amdgcn_readfirstlane(tabn[bk]);
  if (nn > NNL || nn <= NNLO) return;
  const double* sn = tab + bk * 64;
  const double* wn = sn + 32;
  // unused node rows stay 0, so that their chains stay 0
  for (int n = nn; n < NNL; ++n) { Us[n][tid] = 0.0; Gs[n][tid] = 0.0; }
  // one lane per node runs both pivot recurrences in one loop, step s
  // taking the top-down pivot D+ of row s and the bottom-up pivot D- of row
  // k-1-s (two independent chains).  U_i = e_(i-1) / D-_i goes to Us; the
  // diagonal of the inverse, 1 / g_i with g_i = D+_i + D-_i - (d_i + s),
  // goes to Gs once both pivots of row i are known (the first of the two
  // waits in Gs[n][i], written and read by the same lane)
  if (tid < nn) {
    const int n = tid;
    const double sg = m * sn[n];
    double Lp = 0.0, Up = 0.0;
#pragma unroll 2
    for (int st = 0; st < k; ++st) {
      const int it = st, ib = k - 1 - st;
      const double Dp = sd[it] + sg - (it > 0 ? se[it - 1] * Lp : 0.0);
      const double Dm = sd[ib] + sg - (ib < k - 1 ? se[ib] * Up : 0.0);
      Lp = it < k - 1 ? se[it] * mat_rcp(Dp) : 0.0;
      const double Ui = ib > 0 ? se[ib - 1] * mat_rcp(Dm) : 0.0;
      Us[n][ib] = Ui;
      Up = Ui;
      if (it < ib) {            // first half: park the pivots
        Gs[n][it] = Dp;
        Gs[n][ib] = Dm;
      } else if (it == ib) {    // middle row (k odd)
        Gs[n][it] = mat_rcp(Dp + Dm - (sd[it] + sg));
      } else {                  // second half: the other pivot is parked
        const double dmt = Gs[n][it];   // D- of row it
        const double dpb = Gs[n][ib];   // D+ of row ib
        Gs[n][it] = mat_rcp(Dp + dmt - (sd[it] + sg));
        Gs[n][ib] = mat_rcp(dpb + Dm - (sd[ib] + sg));
      }
    }
  }
  __syncthreads();
  const int c = tid;
  if (c >= k) return;
  // rows above a wave's first column hold zeros in every lane of the wave
  const int rfirst = __builtin_amdgcn_readfirstlane(tid & ~63);
  double* Bp = B + pair * (int64_t)k * k;
  mat_rational_rows<NNL>(k, c, rfirst, nn, m, wn, Us, Gs, Bp);
}

// sum over a slab of the tile: the lane's registers (summed by the caller),
// the two halves of the wave, then the four waves through LDS (every wave
// reads the four partial sums in the same order)
__device__ __forceinline__ double tile_slab_sum(double p, double (*red)[32], int wv, int i,
                                                int half) {
  p += __shfl_xor(p, 32);
  if (half == 0) red[wv][i] = p;
  __syncthreads();
  return (red[0][i] + red[1][i]) + (red[2][i] + red[3][i]);
}

// post-processing of one slab per lane pair (i, i + 32) of every wave, from
// the product's registers (register r: member 32 wv + 8 (r/4) + 4 half + r%4,
// valid where bit r of mvalid is set); the same steps and arithmetic as
// k_finalize.  Every wave of the block calls it (it holds barriers).
__device__ __forceinline__ void apply_finalize(const f32x16& acc, unsigned mvalid, bool slab_ok,
                                               const int* sl, int qi, int k, int64_t ij,
                                               int64_t ns, float bmf, float bsf, int v,
                                               const VarsDev& V, const FinArgs& F,
                                               const float* xbrow, double (*red)[4][32], int wv,
                                               int i, int half, double* xa0) {
  const double dk = (double)k, dk1 = (double)(k - 1);
  const double bm = (double)bmf, bs = (double)bsf;
  double x[16];
  auto msum = [&]() {
    double p = 0.0;
#pragma unroll
    for (int r = 0; r < 16; ++r) p += ((mvalid >> r) & 1u) ? x[r] : 0.0;
    return p;
  };
  auto msq = [&]() {
    double p = 0.0;
#pragma unroll
    for (int r = 0; r < 16; ++r) p += ((mvalid >> r) & 1u) ? x[r] * x[r] : 0.0;
    return p;
  };
  // recenter (src/letkf_solver.F90:299-303)
#pragma unroll
  for (int r = 0; r < 16; ++r) x[r] = (double)acc[r] + bm;
  double mean = tile_slab_sum(msum(), red[0], wv, i, half) / dk;
#pragma unroll
  for (int r = 0; r < 16; ++r) x[r] -= mean;
  // increment clipping (:308-318)
  double blo = 0.0, bhi = 0.0;
  if (v >= 0) {
    const double inc = (double)V.inc_max[v];
    blo = (double)V.bounds[2 * v];
    bhi = (double)V.bounds[2 * v + 1];
    double r = mean - bm;
    r = fmax(r, -inc);
    r = fmin(r, inc);
    r = fmax(r, blo - bm);
    r = fmin(r, bhi - bm);
    mean = r + bm;
  }
  // RTPS (:322-335)
  if (F.rtps > 0.0f) {
    const double sa = sqrt(tile_slab_sum(msq(), red[1], wv, i, half) / dk1);
    if (sa > 0.0) {
      const double f = (double)F.rtps * (bs - sa) / sa + 1.0;
#pragma unroll
      for (int r = 0; r < 16; ++r) x[r] *= f;
    }
  }
  // RTPP (:339-343), the background members from the staged tile
  if (F.rtpp > 0.0f) {
    const double a = (double)F.rtpp;
#pragma unroll
    for (int r4 = 0; r4 < 4; ++r4) {
      const float4 b = *reinterpret_cast<const float4*>(xbrow + 8 * r4);
      x[4 * r4 + 0] = a * (double)b.x + (1.0 - a) * x[4 * r4 + 0];
      x[4 * r4 + 1] = a * (double)b.y + (1.0 - a) * x[4 * r4 + 1];
      x[4 * r4 + 2] = a * (double)b.z + (1.0 - a) * x[4 * r4 + 2];
      x[4 * r4 + 3] = a * (double)b.w + (1.0 - a) * x[4 * r4 + 3];
    }
  }
  // multiplicative (:347-349)
  if (F.mul != 1.0f) {
#pragma unroll
    for (int r = 0; r < 16; ++r) x[r] *= (double)F.mul;
  }
  // member bounds and re-mean (:357-371)
  if (V.nvars > 0) {
    if (v >= 0) {
#pragma unroll
      for (int r = 0; r < 16; ++r) x[r] = fmin(fmax(x[r], blo - mean), bhi - mean);
    }
    const double r = tile_slab_sum(v >= 0 ? msum() : 0.0, red[2], wv, i, half) / dk;
    if (v >= 0) {
#pragma unroll
      for (int q = 0; q < 16; ++q) x[q] -= r;
      mean += r;
    }
  }
  // final spread and members (:375-379)
  const double sp = sqrt(tile_slab_sum(msq(), red[3], wv, i, half) / dk1);
  if (!slab_ok) return;
  const int s = sl[qi];
  double* op = xa0 + (int64_t)k * s;
  const bool vec = (k & 3) == 0;
#pragma unroll
  for (int r4 = 0; r4 < 4; ++r4) {
    const int mem = 32 * wv + 8 * r4 + 4 * half;
    if (vec && mem + 4 <= k) {
      double2* o2 = reinterpret_cast<double2*>(op + mem);
      o2[0] = make_double2(x[4 * r4] + mean, x[4 * r4 + 1] + mean);
      o2[1] = make_double2(x[4 * r4 + 2] + mean, x[4 * r4 + 3] + mean);
    } else {
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (mem + u < k) op[mem + u] = x[4 * r4 + u] + mean;
    }
  }
  if (wv == 0 && half == 0) {
    F.amean[ns * ij + s] = mean;
    F.asprd[ns * ij + s] = sp;
  }
}

// x_a = T x_b for every slab of the group: A operand T (rows = members),
// B operand x_b (columns = slabs), K = the ensemble index.  Lane (i, half)
// holds K = 16 step + 8 half + e (e = 0..7) of row / column i in both
// operands (the same K order on both sides, any order contracts the same).
//
// FIN: the post-processing of src/letkf_solver.F90:299-379 follows the
// product for the tile's slabs, in FP64 from the FP32 products, in
// k_finalize's order (recenter, increment clip, RTPS, RTPP, multiplicative,
// member bounds with re-mean, spread); its sums over a slab's members add
// the 16 registers of a lane, the two halves of the wave and the four waves
// (through LDS), and the background members for RTPP are read from the
// staged tile.  The analysis is then written once, post-processed.
template <int PREC, bool FIN>
__global__ __launch_bounds__(256) void k_apply_mat(LocDev L, StateDev S, VarsDev V, ChunkDev C,
                                                   DiagDev D, const double* __restrict__ B,
                                                   double* __restrict__ ana, FinArgs F) {
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
  if (nsl == 0) return;
  // waves past the last member only help to stage the slabs
  const bool rows = 32 * wv < k;
  const int i = lane & 31, half = lane >> 5;
  const int ks = (k + 15) / 16;
  // T rows 32 wv + i, converted once.  The loads are unconditional (a
  // conditional load waits for each value before the next is issued), 16 at
  // a time, from one lane offset with constant offsets through a buffer
  // resource over this problem's T (reads past it return 0); entries past k
  // are masked after.
  const int row = 32 * wv + i;
  const __amdgpu_buffer_rsrc_t trs = buf_rsrc(B + pair * (int64_t)k * k, (int64_t)k * k * 8);
  const uint32_t tvo = (uint32_t)(min(row, k - 1) * k + 8 * half) * 8u;
  // FP32 operand (PREC_FP32) or its BF16 split, converted group by group
  float af[PREC == LETKF_PREC_FP32 ? KS : 1][8];
  bf16x8 ahi[KS], alo[PREC == LETKF_PREC_BF16X3 ? KS : 1];
#pragma unroll
  for (int s0 = 0; s0 < KS; s0 += 2) {
    double tv[2][8];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int e = 0; e < 8; ++e) tv[u][e] = buf_f64(trs, tvo, 8 * (16 * (s0 + u) + e));
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      float c[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int kk = 16 * (s0 + u) + 8 * half + e;
        c[e] = (rows && row < k && kk < k) ? (float)tv[u][e] : 0.0f;
      }
      if (PREC == LETKF_PREC_FP32) {
#pragma unroll
        for (int e = 0; e < 8; ++e) af[s0 + u][e] = c[e];
      } else if (PREC == LETKF_PREC_BF16) {
        ahi[s0 + u] = to_bf16(c);
      } else {
        split_bf16(c, ahi[s0 + u], alo[s0 + u]);
      }
    }
    // one group of loads at a time (16 doubles in registers, not 64)
    __builtin_amdgcn_sched_barrier(0);
  }
  const float* xb0 = S.state + (int64_t)S.k * S.ns * ij;
  double* xa0 = ana + (int64_t)S.k * S.ns * ij;
  const bool vec = (k & 3) == 0;
  // Tiles of 32 slabs are staged in LDS once for the four waves (member
  // rows padded to 132 floats, so that the 16-byte reads of a half-wave,
  // one slab per lane, spread over the banks), and the next tile is loaded
  // while this one is multiplied.  Thread t stages slab t/8 of the tile,
  // members 4 (t%8) + 32u .. +3 for u < 4, zeros past k.
  constexpr int KPAD = 132;
  __shared__ __align__(16) float xs[2][32][KPAD];
  float4 pre[4];
  auto fetch = [&](int q0) {
    const int ql = q0 + (tid >> 3);
    const bool sv = ql < nsl;
    const float* xp = xb0 + (int64_t)k * (sv ? sl[ql] : sl[0]);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int m0 = 4 * (tid & 7) + 32 * u;
      if (vec) {
        pre[u] = (sv && m0 < k) ? *reinterpret_cast<const float4*>(xp + m0)
                                : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      } else {
        pre[u].x = (sv && m0 < k) ? xp[m0] : 0.0f;
        pre[u].y = (sv && m0 + 1 < k) ? xp[m0 + 1] : 0.0f;
        pre[u].z = (sv && m0 + 2 < k) ? xp[m0 + 2] : 0.0f;
        pre[u].w = (sv && m0 + 3 < k) ? xp[m0 + 3] : 0.0f;
      }
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
      *reinterpret_cast<float4*>(&xs[buf][tid >> 3][4 * (tid & 7) + 32 * u]) = pre[u];
  };
  const int ntile = (nsl + 31) / 32;
  // FIN: per-wave partial sums of the tile's slabs, one array per sum
  __shared__ double red[FIN ? 4 : 1][4][32];
  unsigned mvalid = 0;              // register r holds a member < k
#pragma unroll
  for (int r = 0; r < 16; ++r)
    if (32 * wv + 8 * (r >> 2) + 4 * half + (r & 3) < k) mvalid |= 1u << r;
  fetch(0);
  stash(0);
  __syncthreads();
  for (int tl = 0; tl < ntile; ++tl) {
    const int cur = tl & 1;
    const int q0 = 32 * tl;
    if (tl + 1 < ntile) fetch(q0 + 32);
    const int qi = q0 + i;
    // FIN: the slab's background mean, spread and variable, loaded before
    // the products
    float bmf = 0.0f, bsf = 0.0f;
    int var = -1;
    if (FIN) {
      const int s = sl[qi < nsl ? qi : 0];
      bmf = S.mean[S.ns * ij + s];
      bsf = S.sprd[S.ns * ij + s];
      var = V.slab_var[s];
    }
    f32x16 acc = (f32x16){};
    if (rows) {
#pragma unroll
      for (int st = 0; st < KS; ++st) {
        if (st < ks) {
          const int k0 = 16 * st + 8 * half;
          const float4 u0 = *reinterpret_cast<const float4*>(&xs[cur][i][k0]);
          const float4 u1 = *reinterpret_cast<const float4*>(&xs[cur][i][k0 + 4]);
          float xf[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
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
    }
    if (FIN) {
      apply_finalize(acc, mvalid, qi < nsl, sl, qi, k, ij, S.ns, bmf, bsf, var, V, F,
                     &xs[cur][i][32 * wv + 4 * half], red, wv, i, half, xa0);
    } else if (rows) {
      // register r of lane (i, half): member 32 wv + 8 (r/4) + 4 half + r%4 of slab q0 + i
      if (qi < nsl) {
        double* op = xa0 + (int64_t)k * sl[qi];
#pragma unroll
        for (int r4 = 0; r4 < 4; ++r4) {
          const int mem = 32 * wv + 8 * r4 + 4 * half;       // 4 consecutive members
          if (vec && mem + 4 <= k) {
            double2* o2 = reinterpret_cast<double2*>(op + mem);
            o2[0] = make_double2((double)acc[4 * r4], (double)acc[4 * r4 + 1]);
            o2[1] = make_double2((double)acc[4 * r4 + 2], (double)acc[4 * r4 + 3]);
          } else {
#pragma unroll
            for (int u = 0; u < 4; ++u)
              if (mem + u < k) op[mem + u] = (double)acc[4 * r4 + u];
          }
        }
      }
    }
    if (tl + 1 < ntile) stash(cur ^ 1);
    __syncthreads();
  }
}

// FP64 apply (the plain path) on FP64 MFMA (16x16x4): 16-member x 16-slab
// output tiles; wave w owns the member tiles w and w + 4, keeps their T rows
// as operand registers (lane l: member 16 t + l%16, K = 4 step + l/16) and
// walks the slab tiles.  Output register r of lane l is member 4r + l/16 of
// the tile, slab l%16.
typedef double f64x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_apply_mat_f64(LocDev L, StateDev S, VarsDev V,
                                                       ChunkDev C, DiagDev D,
                                                       const double* __restrict__ B,
                                                       double* __restrict__ ana) {
  constexpr int KS = 32;             // K steps of 4, k <= 128
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
  const int ks = (k + 3) / 4;
  const int i16 = lane & 15, kq = lane >> 4;
  const float* xb0 = S.state + (int64_t)S.k * S.ns * ij;
  double* xa0 = ana + (int64_t)S.k * S.ns * ij;
  const double* Tb = B + pair * (int64_t)k * k;
  // every load issued, from one lane offset with constant offsets through
  // buffer resources over this problem's T and this gridpoint's background
  // (reads past them return 0); entries past k masked after
  const __amdgpu_buffer_rsrc_t trs = buf_rsrc(Tb, (int64_t)k * k * 8);
  const __amdgpu_buffer_rsrc_t xrs = buf_rsrc(xb0, (int64_t)k * S.ns * 4);
  for (int mt = wv; 16 * mt < k; mt += 4) {
    const int row = 16 * mt + i16;
    const uint32_t tvo = (uint32_t)(min(row, k - 1) * k + kq) * 8u;
    double af[KS];
#pragma unroll
    for (int st = 0; st < KS; ++st) af[st] = buf_f64(trs, tvo, 32 * st);
#pragma unroll
    for (int st = 0; st < KS; ++st) {
      const int kk = 4 * st + kq;
      af[st] = (row < k && kk < k) ? af[st] : 0.0;
    }
    for (int q0 = 0; q0 < nsl; q0 += 16) {
      const int qi = q0 + i16;
      const int slab = qi < nsl ? sl[qi] : sl[0];
      const uint32_t xvo = (uint32_t)(slab * k + kq) * 4u;
      float xv[KS];
#pragma unroll
      for (int st = 0; st < KS; ++st) xv[st] = buf_f32(xrs, xvo, 16 * st);
      f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int st = 0; st < KS; ++st) {
        if (st < ks) {
          const int kk = 4 * st + kq;
          const double xb = kk < k ? (double)xv[st] : 0.0;
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(af[st], xb, acc, 0, 0, 0);
        }
      }
      const int q = q0 + i16;
      if (q < nsl) {
        double* op = xa0 + (int64_t)k * sl[q];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int mem = 16 * mt + 4 * r + kq;
          if (mem < k) op[mem] = acc[r];
        }
      }
    }
  }
}

bool transform_mat_supported(int k) { return k >= 64 && k <= 128; }

bool transform_mat_fuses_finalize(int prec) { return prec != LETKF_PREC_FP64; }

void launch_transform_mat(hipStream_t st, int prec, const LocDev& L, const StateDev& S,
                          const VarsDev& V, const ChunkDev& C, const DiagDev& D, double* B,
                          const double* bvec, double* tau, double* dd, double* ee,
                          const RatTab& R, double* ana, bool blk, const FinArgs* fin) {
  const int64_t np = C.ngp * C.G;
  if (np <= 0) return;
  const int k = (int)S.k;
  const unsigned nb = (unsigned)np;
  // blk: the 2D-block kernels, BS x BS blocks with 16 BS >= k; else the
  // column-half kernels, rows per half k <= 2 NR
  const int bs = (k + 15) / 16;
  const int nr = (k + 1) / 2;
  if (blk) {
    switch (bs) {
      case 4: launch_blk_tridiag_bs4(st, nb, k, C, B, tau, dd, ee); break;
      case 5: launch_blk_tridiag_bs5(st, nb, k, C, B, tau, dd, ee); break;
      case 6: launch_blk_tridiag_bs6(st, nb, k, C, B, tau, dd, ee); break;
      case 7: launch_blk_tridiag_bs7(st, nb, k, C, B, tau, dd, ee); break;
      default: launch_blk_tridiag_bs8(st, nb, k, C, B, tau, dd, ee); break;
    }
  } else if (nr <= 40) launch_mat_tridiag_nr40(st, nb, k, C, B, tau, dd, ee);
  else if (nr <= 48) launch_mat_tridiag_nr48(st, nb, k, C, B, tau, dd, ee);
  else if (nr <= 56) launch_mat_tridiag_nr56(st, nb, k, C, B, tau, dd, ee);
  else launch_mat_tridiag_nr64(st, nb, k, C, B, tau, dd, ee);
  k_mat_rational<16, 0><<<nb, 128, 0, st>>>(k, C, B, dd, ee, R.tab, R.n, R.nb);
  k_mat_rational<24, 16><<<nb, 128, 0, st>>>(k, C, B, dd, ee, R.tab, R.n, R.nb);
  k_mat_rational<kMaxNodes, 24><<<nb, 128, 0, st>>>(k, C, B, dd, ee, R.tab, R.n, R.nb);
  if (blk) {
    switch (bs) {
      case 4: launch_blk_back_bs4(st, nb, k, C, B, bvec, tau); break;
      case 5: launch_blk_back_bs5(st, nb, k, C, B, bvec, tau); break;
      case 6: launch_blk_back_bs6(st, nb, k, C, B, bvec, tau); break;
      case 7: launch_blk_back_bs7(st, nb, k, C, B, bvec, tau); break;
      default: launch_blk_back_bs8(st, nb, k, C, B, bvec, tau); break;
    }
  } else if (nr <= 40) launch_mat_back_nr40(st, nb, k, C, B, bvec, tau);
  else if (nr <= 48) launch_mat_back_nr48(st, nb, k, C, B, bvec, tau);
  else if (nr <= 56) launch_mat_back_nr56(st, nb, k, C, B, bvec, tau);
  else launch_mat_back_nr64(st, nb, k, C, B, bvec, tau);
  const FinArgs F = fin ? *fin : FinArgs{0.0f, 0.0f, 1.0f, nullptr, nullptr};
  const bool f = fin != nullptr && transform_mat_fuses_finalize(prec);
  if (prec == LETKF_PREC_FP64)
    k_apply_mat_f64<<<nb, 256, 0, st>>>(L, S, V, C, D, B, ana);
  else if (prec == LETKF_PREC_FP32)
    (f ? k_apply_mat<LETKF_PREC_FP32, true> : k_apply_mat<LETKF_PREC_FP32, false>)
        <<<nb, 256, 0, st>>>(L, S, V, C, D, B, ana, F);
  else if (prec == LETKF_PREC_BF16)
    (f ? k_apply_mat<LETKF_PREC_BF16, true> : k_apply_mat<LETKF_PREC_BF16, false>)
        <<<nb, 256, 0, st>>>(L, S, V, C, D, B, ana, F);
  else
    (f ? k_apply_mat<LETKF_PREC_BF16X3, true> : k_apply_mat<LETKF_PREC_BF16X3, false>)
        <<<nb, 256, 0, st>>>(L, S, V, C, D, B, ana, F);
}

}  // namespace lk
