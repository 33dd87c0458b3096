// kernels_product_mx.hip -- Yb^T R^-1 Yb and Yb^T R^-1 d on low-precision
// MFMA with FP32 accumulation (the batched mixed-precision path).
//
// One wave per (gridpoint, group) problem, launched in the order of a
// device-sorted permutation (k_bucket_keys below: by default buckets of the
// kept count m, largest first, gridpoint order inside).  The kept list of a
// problem holds (column byte offset, sqrt(rloc_j/err_j)) pairs, read with the
// padded operand columns hxp(kp, n) through buffer descriptors whose ranges
// end at the list's zero-padded length and at the column array, so no
// row or range test is needed.  Lists are loaded two 16-observation steps
// ahead and columns one step ahead of their MFMAs.  Both MFMA operands are
// sqrt(rloc_j/err_j) y_j, so each value is rounded or split once; row k of
// the padded columns carries d_j, so column k of the result is Yb^T R^-1 d.
// Only the upper-triangle 32x32 blocks are computed (or, as a launch
// choice, 16x16 tiles: k_product_t16 below).
//
//   PREC_FP32    v_mfma_f32_32x32x2_f32   (8 MFMA per block per 16 obs)
//   PREC_BF16    v_mfma_f32_32x32x16_bf16 (1 per block per 16 obs)
//   PREC_BF16X3  split x = hi + lo in BF16, hi*hi + hi*lo + lo*hi (3 per block)
//   PREC_BF16X4  the same split with lo*lo as well (4 per block): the exact
//                product of the split operands, whose error is then the
//                split's 2^-17 relative residual with a random sign instead
//                of the dropped lo*lo term, which is positive on the diagonal
//
// Accumulation.  The MFMAs accumulate in FP32.  For PREC_FP32 and
// PREC_BF16X4 the accumulators are compensated: every F observations (the
// fold period, a power of two >= 32) the partial sums are added into FP32
// running sums S by the error-free TwoSum (S + a = s + e exactly), and the
// partial sums restart from the rounding error e (Kahan's compensation
// carried forward: e, about 2^-24 |S|, joins the next block's sum, whose own
// rounding is of the same order).  The result S + partial is formed in
// FP64.  The error of an m-observation product then grows like
// F sqrt(m) rounding units instead of m^1.5 / 2 (products of FP32 operands
// summed exactly would be ~1e-9 from the FP64 analysis; the FP32 partial
// sums are what is left).  PREC_BF16X3 and PREC_BF16 keep plain FP32
// accumulation.
//
// The FP32 result is widened to FP64 for the transform.
#include <hip/hip_runtime.h>

#include "letkf_internal.h"
#include "mx_common.h"

namespace lk {


template <int NB>
struct MxFrag {
  float y[NB][8];   // y_j[32 rb + lane%32] of the lane's 8 obs, as loaded
  float w[8];       // their sqrt(rloc_j/err_j), applied by mx_step
};

// observation e (0..7) of a lane in a 16-observation step: the BF16 MFMA
// takes 8 consecutive K per lane half, the FP32 MFMA interleaves halves
template <int PREC>
__device__ __forceinline__ int obs_slot(int half, int e) {
  return PREC == LETKF_PREC_FP32 ? 2 * e + half : 8 * half + e;
}

// Operand registers of one 16-observation step.  The kept list of a problem
// holds (column byte offset, weight) pairs from a multiple of 8 entries on,
// zero-padded to the next multiple of 8; both are read through buffer
// descriptors whose range ends there, so entries past the list read as
// (0, 0): column 0 with weight 0.  The padded columns carry the departure in
// row k and zeros above it, so no row or range test is needed.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// column byte offsets of the lane's 8 observations of a step
struct ListRegs {
  uint32_t sidx[8];
};

// byte offset in the list of the lane's observation e of the step at j0
template <int PREC>
__device__ __forceinline__ uint32_t list_off(int j0, int half, int e) {
  return (uint32_t)(j0 + obs_slot<PREC>(half, e)) * 4u;
}

template <int PREC>
__device__ __forceinline__ void load_list(ListRegs& Lr, __amdgpu_buffer_rsrc_t ir, int j0,
                                          int lane) {
  const int half = lane >> 5;
  if (PREC == LETKF_PREC_FP32) {
#pragma unroll
    for (int e = 0; e < 8; ++e)
      Lr.sidx[e] = __builtin_amdgcn_raw_buffer_load_b32(ir, list_off<PREC>(j0, half, e), 0, 0);
  } else {
    const uint32_t vo = list_off<PREC>(j0, half, 0);
    const u32x4 a0 = __builtin_amdgcn_raw_buffer_load_b128(ir, vo, 0, 0);
    const u32x4 a1 = __builtin_amdgcn_raw_buffer_load_b128(ir, vo, 16, 0);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      Lr.sidx[e] = a0[e];
      Lr.sidx[4 + e] = a1[e];
    }
  }
}

// the step's columns (at the list's offsets) and its weights, which are
// loaded here rather than with the offsets so that no copy of them has to
// outlive the offsets' registers
template <int NB, int PREC>
__device__ __forceinline__ void load_cols(MxFrag<NB>& F, __amdgpu_buffer_rsrc_t hxr,
                                          __amdgpu_buffer_rsrc_t wr, const ListRegs& Lr,
                                          int j0, int lane) {
  const int i = lane & 31, half = lane >> 5;
  if (PREC == LETKF_PREC_FP32) {
#pragma unroll
    for (int e = 0; e < 8; ++e)
      F.w[e] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(wr, list_off<PREC>(j0, half, e), 0, 0));
  } else {
    const uint32_t vo = list_off<PREC>(j0, half, 0);
    const u32x4 b0 = __builtin_amdgcn_raw_buffer_load_b128(wr, vo, 0, 0);
    const u32x4 b1 = __builtin_amdgcn_raw_buffer_load_b128(wr, vo, 16, 0);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      F.w[e] = __uint_as_float(b0[e]);
      F.w[4 + e] = __uint_as_float(b1[e]);
    }
  }
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const uint32_t vo = Lr.sidx[e] + 4u * (uint32_t)i;
#pragma unroll
    for (int rb = 0; rb < NB; ++rb)
      F.y[rb][e] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(hxr, vo, 128 * rb, 0));
  }
}

// acc += the products of one 16-observation step
template <int NB, int PREC>
__device__ __forceinline__ void mx_step(f32x16* acc, const MxFrag<NB>& in) {
  // both operands are sqrt(R^-1 rloc) y: A B^T = Yb^T R^-1 Yb, and row k
  // (sqrt(R^-1 rloc) d) gives Yb^T R^-1 d in column k.  The weights are
  // applied here, where the columns are needed, and not next to their
  // loads, so that no wait for the loads is placed ahead of this step.
  MxFrag<NB> cur;
#pragma unroll
  for (int rb = 0; rb < NB; ++rb)
#pragma unroll
    for (int e = 0; e < 8; ++e) cur.y[rb][e] = in.y[rb][e] * in.w[e];
  if (PREC == LETKF_PREC_FP32) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      int b = 0;
#pragma unroll
      for (int rb = 0; rb < NB; ++rb)
#pragma unroll
        for (int cb = rb; cb < NB; ++cb, ++b)
          acc[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(cur.y[rb][e], cur.y[cb][e], acc[b], 0, 0, 0);
    }
  } else if (PREC == LETKF_PREC_BF16) {
    bf16x8 v16[NB];
#pragma unroll
    for (int rb = 0; rb < NB; ++rb) v16[rb] = to_bf16(cur.y[rb]);
    int b = 0;
#pragma unroll
    for (int rb = 0; rb < NB; ++rb)
#pragma unroll
      for (int cb = rb; cb < NB; ++cb, ++b)
        acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(v16[rb], v16[cb], acc[b], 0, 0, 0);
  } else {
    bf16x8 hi[NB], lo[NB];
#pragma unroll
    for (int rb = 0; rb < NB; ++rb) split_bf16(cur.y[rb], hi[rb], lo[rb]);
    int b = 0;
#pragma unroll
    for (int rb = 0; rb < NB; ++rb)
#pragma unroll
      for (int cb = rb; cb < NB; ++cb, ++b) {
        if (PREC == LETKF_PREC_BF16X4)
          acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lo[rb], lo[cb], acc[b], 0, 0, 0);
        acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lo[rb], hi[cb], acc[b], 0, 0, 0);
        acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi[rb], lo[cb], acc[b], 0, 0, 0);
        acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi[rb], hi[cb], acc[b], 0, 0, 0);
      }
  }
}

// compensated accumulation: S += acc by TwoSum, acc restarts from the
// rounding error; the running sums of block b, four at a time, are the
// lane's own LDS slots Sl[64 (4 b + q)], out of the registers
template <int NBLK>
__device__ __forceinline__ void mx_fold(f32x16* acc, f32x4* Sl) {
#pragma clang fp contract(off)
#pragma unroll
  for (int b = 0; b < NBLK; ++b)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const f32x4 x4 = Sl[64 * (4 * b + q)];
      f32x4 t4;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int r = 4 * q + u;
        const float a = acc[b][r], x = x4[u];
        const float t = x + a;
        const float bb = t - x;
        const float e = (x - (t - bb)) + (a - bb);
        t4[u] = t;
        acc[b][r] = e;
      }
      Sl[64 * (4 * b + q)] = t4;
    }
}

template <int KP, int PREC>
__global__ __launch_bounds__(64) void k_product_mx(ObsDev O, int k, ChunkDev C,
                                                   const int32_t* __restrict__ perm,
                                                   double* __restrict__ B,
                                                   double* __restrict__ bvec, int fold_mask) {
  constexpr int NB = KP / 32;
  constexpr int NBLK = NB * (NB + 1) / 2;
  const int64_t pair = perm[xcd_order(C.xcd, blockIdx.x, gridDim.x)];
  if (!C.solve[pair]) return;
  const int m = (int)C.cnt[pair];
  const int64_t off = C.off[pair];
  const int lane = threadIdx.x;
  // padded operand columns (the host guarantees n * kp * 4 < 2^31) and the
  // problem's list, ranged to its zero-padded length
  const __amdgpu_buffer_rsrc_t hxr = __builtin_amdgcn_make_buffer_rsrc(
      (void*)O.hxp, (short)0, (int)(O.n * O.kp * 4), 0x00020000);
  const int lbytes = ((m + kListAlign - 1) & ~(kListAlign - 1)) * 4;
  const __amdgpu_buffer_rsrc_t ir =
      __builtin_amdgcn_make_buffer_rsrc((void*)(C.list_idx + off), (short)0, lbytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t wr =
      __builtin_amdgcn_make_buffer_rsrc((void*)(C.list_w + off), (short)0, lbytes, 0x00020000);

  constexpr bool COMP = PREC == LETKF_PREC_FP32 || PREC == LETKF_PREC_BF16X4;
  f32x16 acc[NBLK];
#pragma unroll
  for (int b = 0; b < NBLK; ++b) acc[b] = (f32x16){};
  // compensated accumulation: running sums (LDS, the lane's own slots)
  __shared__ f32x4 Ssh[COMP ? NBLK * 4 * 64 : 1];
  if (COMP) {
#pragma unroll
    for (int q = 0; q < NBLK * 4; ++q) Ssh[64 * q + lane] = (f32x4){};
  }

  // software pipeline: the list entries of a step are loaded two steps
  // ahead, its columns one step ahead of its MFMAs; steps past m load zeros
  // and add nothing
  MxFrag<NB> fa, fb;
  ListRegs la, lb;
  load_list<PREC>(la, ir, 0, lane);
  load_list<PREC>(lb, ir, 16, lane);
  __builtin_amdgcn_sched_barrier(0);
  load_cols<NB, PREC>(fa, hxr, wr, la, 0, lane);
  __builtin_amdgcn_sched_barrier(0);
  load_list<PREC>(la, ir, 32, lane);
  __builtin_amdgcn_sched_barrier(0);
  // scheduling barriers keep each step's loads ahead of the other buffer's
  // products, so the waits before a step cover only that step's loads
  for (int j0 = 0; j0 < m; j0 += 32) {
    load_cols<NB, PREC>(fb, hxr, wr, lb, j0 + 16, lane);
    load_list<PREC>(lb, ir, j0 + 48, lane);
    __builtin_amdgcn_sched_barrier(0);
    mx_step<NB, PREC>(acc, fa);
    __builtin_amdgcn_sched_barrier(0);
    load_cols<NB, PREC>(fa, hxr, wr, la, j0 + 32, lane);
    load_list<PREC>(la, ir, j0 + 64, lane);
    __builtin_amdgcn_sched_barrier(0);
    mx_step<NB, PREC>(acc, fb);
    __builtin_amdgcn_sched_barrier(0);
    // every fold_mask + 1 observations (wave-uniform)
    if (COMP && ((j0 + 32) & fold_mask) == 0) mx_fold<NBLK>(acc, Ssh + lane);
  }

  // 32x32 f32 accumulator: register r of lane l holds row 8(r/4) + 4(l/32) + r%4
  double* Bp = B + pair * (int64_t)k * k;
  double* bp = bvec + pair * (int64_t)k;
  const int half = lane >> 5, i = lane & 31;
  int b = 0;
#pragma unroll
  for (int rb = 0; rb < NB; ++rb)
#pragma unroll
    for (int cb = rb; cb < NB; ++cb, ++b) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = 32 * rb + 8 * (r >> 2) + 4 * half + (r & 3);
        const int col = 32 * cb + i;
        double v = (double)acc[b][r];
        if (COMP) v += (double)Ssh[64 * (4 * b + (r >> 2)) + lane][r & 3];
        if (row < k && col < k) {
          Bp[row + (int64_t)k * col] = v;
          Bp[col + (int64_t)k * row] = v;
        } else if (col == k && row < k) {
          bp[row] = v;
        }
      }
    }
}

// ---------------------------------------------------------------------------
// The same products on 16x16 tiles (v_mfma_f32_16x16x32_bf16): the operand
// rows are padded to 16 NT >= k + 1 rather than to a multiple of 32, and
// the upper triangle is NT (NT + 1) / 2 tiles of 16x16 (k = 40: 6 tiles,
// 1,536 products per observation, against 3 blocks of 32x32, 3,072; k = 50:
// 2,560 against 3,072).  One step is 32 observations: lane l holds
// observations 8 (l / 16) + e (e = 0..7) of row 16 t + l % 16 of every row
// tile t, so each lane reads 8 list entries per 32 observations (half the
// 32x32 kernel's).  The columns of a step are loaded into one register set
// right after the previous step's were converted to BF16, and are in flight
// during that step's MFMAs; lists are loaded two steps ahead.  Split-BF16
// precisions and compensated accumulation as above (PREC_FP32 keeps the
// 32x32 kernel).
// ---------------------------------------------------------------------------

template <int NT>
struct T16Frag {
  float y[NT][8];   // y_j[16 t + lane%16] of the lane's 8 obs, as loaded
  float w[8];       // their sqrt(rloc_j/err_j)
};

__device__ __forceinline__ void t16_list(ListRegs& Lr, __amdgpu_buffer_rsrc_t ir, int j0,
                                         int lane) {
  const uint32_t vo = (uint32_t)(j0 + 8 * (lane >> 4)) * 4u;
  const u32x4 a0 = __builtin_amdgcn_raw_buffer_load_b128(ir, vo, 0, 0);
  const u32x4 a1 = __builtin_amdgcn_raw_buffer_load_b128(ir, vo, 16, 0);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    Lr.sidx[e] = a0[e];
    Lr.sidx[4 + e] = a1[e];
  }
}

template <int NT>
__device__ __forceinline__ void t16_cols(T16Frag<NT>& F, __amdgpu_buffer_rsrc_t hxr,
                                         __amdgpu_buffer_rsrc_t wr, const ListRegs& Lr, int j0,
                                         int lane) {
  const uint32_t vo = (uint32_t)(j0 + 8 * (lane >> 4)) * 4u;
  const u32x4 b0 = __builtin_amdgcn_raw_buffer_load_b128(wr, vo, 0, 0);
  const u32x4 b1 = __builtin_amdgcn_raw_buffer_load_b128(wr, vo, 16, 0);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    F.w[e] = __uint_as_float(b0[e]);
    F.w[4 + e] = __uint_as_float(b1[e]);
  }
  const uint32_t i4 = 4u * (uint32_t)(lane & 15);
#pragma unroll
  for (int e = 0; e < 8; ++e)
#pragma unroll
    for (int t = 0; t < NT; ++t)
      F.y[t][e] = __uint_as_float(
          __builtin_amdgcn_raw_buffer_load_b32(hxr, Lr.sidx[e] + i4, 64 * t, 0));
}

// weighted operands of a step, split (or rounded) to BF16
template <int NT, int PREC>
__device__ __forceinline__ void t16_convert(const T16Frag<NT>& F, bf16x8* hi, bf16x8* lo) {
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    float c[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) c[e] = F.y[t][e] * F.w[e];
    if (PREC == LETKF_PREC_BF16) hi[t] = to_bf16(c);
    else split_bf16(c, hi[t], lo[t]);
  }
}

// acc[tile (rt, ct), ct >= rt] += products of one step; each term sweeps all
// tiles, so consecutive MFMAs write different accumulators
template <int NT, int PREC>
__device__ __forceinline__ void t16_mfma(f32x4* acc, const bf16x8* hi, const bf16x8* lo) {
  if (PREC == LETKF_PREC_BF16X4) {
    int b = 0;
#pragma unroll
    for (int rt = 0; rt < NT; ++rt)
#pragma unroll
      for (int ct = rt; ct < NT; ++ct, ++b)
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lo[rt], lo[ct], acc[b], 0, 0, 0);
  }
  if (PREC != LETKF_PREC_BF16) {
    int b = 0;
#pragma unroll
    for (int rt = 0; rt < NT; ++rt)
#pragma unroll
      for (int ct = rt; ct < NT; ++ct, ++b)
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lo[rt], hi[ct], acc[b], 0, 0, 0);
    b = 0;
#pragma unroll
    for (int rt = 0; rt < NT; ++rt)
#pragma unroll
      for (int ct = rt; ct < NT; ++ct, ++b)
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi[rt], lo[ct], acc[b], 0, 0, 0);
  }
  int b = 0;
#pragma unroll
  for (int rt = 0; rt < NT; ++rt)
#pragma unroll
    for (int ct = rt; ct < NT; ++ct, ++b)
      acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi[rt], hi[ct], acc[b], 0, 0, 0);
}

// compensated accumulation on 16x16 tiles (4 values per lane per tile); the
// running sums of tile b are the lane's own LDS slot Sl[64 b], which keeps
// them out of the registers (two waves per SIMD become three at NT = 4)
template <int NTILE>
__device__ __forceinline__ void t16_fold(f32x4* acc, f32x4* Sl) {
#pragma clang fp contract(off)
#pragma unroll
  for (int b = 0; b < NTILE; ++b) {
    const f32x4 x4 = Sl[64 * b];
    f32x4 t4;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float a = acc[b][r], x = x4[r];
      const float t = x + a;
      const float bb = t - x;
      const float e = (x - (t - bb)) + (a - bb);
      t4[r] = t;
      acc[b][r] = e;
    }
    Sl[64 * b] = t4;
  }
}

template <int NT, int PREC>
__global__ __launch_bounds__(64) void k_product_t16(ObsDev O, int k, ChunkDev C,
                                                    const int32_t* __restrict__ perm,
                                                    double* __restrict__ B,
                                                    double* __restrict__ bvec, int fold_mask) {
  constexpr int NTILE = NT * (NT + 1) / 2;
  const int64_t pair = perm[xcd_order(C.xcd, blockIdx.x, gridDim.x)];
  if (!C.solve[pair]) return;
  const int m = (int)C.cnt[pair];
  const int64_t off = C.off[pair];
  const int lane = threadIdx.x;
  const __amdgpu_buffer_rsrc_t hxr = __builtin_amdgcn_make_buffer_rsrc(
      (void*)O.hxp, (short)0, (int)(O.n * O.kp * 4), 0x00020000);
  const int lbytes = ((m + kListAlign - 1) & ~(kListAlign - 1)) * 4;
  const __amdgpu_buffer_rsrc_t ir =
      __builtin_amdgcn_make_buffer_rsrc((void*)(C.list_idx + off), (short)0, lbytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t wr =
      __builtin_amdgcn_make_buffer_rsrc((void*)(C.list_w + off), (short)0, lbytes, 0x00020000);

  constexpr bool COMP = PREC == LETKF_PREC_BF16X4;
  f32x4 acc[NTILE];
#pragma unroll
  for (int b = 0; b < NTILE; ++b) acc[b] = (f32x4){};
  __shared__ f32x4 Ssh[COMP ? NTILE * 64 : 1];
  if (COMP) {
#pragma unroll
    for (int b = 0; b < NTILE; ++b) Ssh[64 * b + lane] = (f32x4){};
  }

  // f: the columns of the step in flight; la, lb: lists two and one steps
  // ahead (steps past m load zeros and add nothing)
  T16Frag<NT> f;
  ListRegs la, lb;
  bf16x8 hi[NT], lo[PREC == LETKF_PREC_BF16 ? 1 : NT];
  t16_list(la, ir, 0, lane);
  t16_list(lb, ir, 32, lane);
  t16_cols<NT>(f, hxr, wr, la, 0, lane);
  t16_list(la, ir, 64, lane);
  for (int j0 = 0; j0 < m; j0 += 32) {
    __builtin_amdgcn_sched_barrier(0);
    t16_convert<NT, PREC>(f, hi, lo);
    __builtin_amdgcn_sched_barrier(0);
    t16_cols<NT>(f, hxr, wr, lb, j0 + 32, lane);
    lb = la;
    t16_list(la, ir, j0 + 96, lane);
    __builtin_amdgcn_sched_barrier(0);
    t16_mfma<NT, PREC>(acc, hi, lo);
    if (COMP && ((j0 + 32) & fold_mask) == 0) t16_fold<NTILE>(acc, Ssh + lane);
  }

  // 16x16 f32 accumulator: register r of lane l holds row 4 (l / 16) + r,
  // column l % 16
  double* Bp = B + pair * (int64_t)k * k;
  double* bp = bvec + pair * (int64_t)k;
  const int g = lane >> 4, i = lane & 15;
  int b = 0;
#pragma unroll
  for (int rt = 0; rt < NT; ++rt)
#pragma unroll
    for (int ct = rt; ct < NT; ++ct, ++b) {
      const f32x4 sb = COMP ? Ssh[64 * b + lane] : (f32x4){};
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * rt + 4 * g + r;
        const int col = 16 * ct + i;
        double v = (double)acc[b][r];
        if (COMP) v += (double)sb[r];
        if (row < k && col < k) {
          Bp[row + (int64_t)k * col] = v;
          Bp[col + (int64_t)k * row] = v;
        } else if (col == k && row < k) {
          bp[row] = v;
        }
      }
    }
}

template <int KP>
static void launch_kp(hipStream_t st, int prec, unsigned nb, const ObsDev& O, int k,
                      const ChunkDev& C, const int32_t* perm, double* B, double* bvec,
                      int fm) {
  if (prec == LETKF_PREC_FP32)
    k_product_mx<KP, LETKF_PREC_FP32><<<nb, 64, 0, st>>>(O, k, C, perm, B, bvec, fm);
  else if (prec == LETKF_PREC_BF16)
    k_product_mx<KP, LETKF_PREC_BF16><<<nb, 64, 0, st>>>(O, k, C, perm, B, bvec, fm);
  else if (prec == LETKF_PREC_BF16X4)
    k_product_mx<KP, LETKF_PREC_BF16X4><<<nb, 64, 0, st>>>(O, k, C, perm, B, bvec, fm);
  else
    k_product_mx<KP, LETKF_PREC_BF16X3><<<nb, 64, 0, st>>>(O, k, C, perm, B, bvec, fm);
}

template <int NT>
static void launch_nt(hipStream_t st, int prec, unsigned nb, const ObsDev& O, int k,
                      const ChunkDev& C, const int32_t* perm, double* B, double* bvec,
                      int fm) {
  if (prec == LETKF_PREC_BF16)
    k_product_t16<NT, LETKF_PREC_BF16><<<nb, 64, 0, st>>>(O, k, C, perm, B, bvec, fm);
  else if (prec == LETKF_PREC_BF16X4)
    k_product_t16<NT, LETKF_PREC_BF16X4><<<nb, 64, 0, st>>>(O, k, C, perm, B, bvec, fm);
  else
    k_product_t16<NT, LETKF_PREC_BF16X3><<<nb, 64, 0, st>>>(O, k, C, perm, B, bvec, fm);
}

bool product_t16_supported(int prec, int k) {
  return prec != LETKF_PREC_FP32 && (k + 1 + 15) / 16 <= 7;
}

void launch_product_mx(hipStream_t st, int prec, const ObsDev& O, const StateDev& S,
                       const ChunkDev& C, const int32_t* perm, double* B, double* bvec,
                       int fold_log2, int tile) {
  const int64_t np = C.ngp * C.G;
  if (np <= 0) return;
  const int k = (int)S.k;
  const int kp = ((k + 1 + 31) / 32) * 32;
  const unsigned nb = (unsigned)np;
  const int fm = (1 << (fold_log2 < 5 ? 5 : fold_log2)) - 1;   // the loops step by 32
  if (tile == 1 && product_t16_supported(prec, k)) {
    switch ((k + 1 + 15) / 16) {
      case 1: launch_nt<1>(st, prec, nb, O, k, C, perm, B, bvec, fm); break;
      case 2: launch_nt<2>(st, prec, nb, O, k, C, perm, B, bvec, fm); break;
      case 3: launch_nt<3>(st, prec, nb, O, k, C, perm, B, bvec, fm); break;
      case 4: launch_nt<4>(st, prec, nb, O, k, C, perm, B, bvec, fm); break;
      case 5: launch_nt<5>(st, prec, nb, O, k, C, perm, B, bvec, fm); break;
      case 6: launch_nt<6>(st, prec, nb, O, k, C, perm, B, bvec, fm); break;
      default: launch_nt<7>(st, prec, nb, O, k, C, perm, B, bvec, fm); break;
    }
    return;
  }
  switch (kp) {
    case 32: launch_kp<32>(st, prec, nb, O, k, C, perm, B, bvec, fm); break;
    case 64: launch_kp<64>(st, prec, nb, O, k, C, perm, B, bvec, fm); break;
    case 96: launch_kp<96>(st, prec, nb, O, k, C, perm, B, bvec, fm); break;
    default: launch_kp<128>(st, prec, nb, O, k, C, perm, B, bvec, fm); break;
  }
}

// order of the problems of a chunk: mode 0 gridpoint order, mode 1 longest
// first (descending m), mode 2 buckets of m, largest first, gridpoint order
// inside a bucket, mode 3 the same buckets with a Morton (Z-order) walk over
// a 0.25-degree lat/lon lattice inside a bucket, so that problems in flight
// together share observation columns in L2.  Bucket edges: the powers of two
// split into 2^bbits equal parts (bbits 0..2), so bucket
// floor(log2 m) . (the next bbits bits of m).
__device__ __forceinline__ uint32_t spread_bits(uint32_t x) {   // 16 -> 32 bits
  x &= 0xffffu;
  x = (x | (x << 8)) & 0x00ff00ffu;
  x = (x | (x << 4)) & 0x0f0f0f0fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}

__device__ __forceinline__ uint32_t bucket_desc(int64_t m, int bbits) {
  // 0 for the largest bucket; unsolved problems (m = 0) last
  const uint32_t nb = 32u << bbits;
  if (m <= 0) return nb - 1u;
  const uint32_t L = 63u - (uint32_t)__clzll((unsigned long long)m);   // floor(log2 m)
  const uint32_t Lc = L > 30u ? 30u : L;
  const uint32_t mask = (1u << bbits) - 1u;
  const uint32_t sub = L >= (uint32_t)bbits ? (uint32_t)(m >> (L - bbits)) & mask
                                            : (uint32_t)(m << (bbits - L)) & mask;
  return nb - 2u - ((Lc << bbits) | sub);
}

__global__ void k_bucket_keys(int mode, int bbits, int64_t np, int G, int64_t i0,
                              const int64_t* __restrict__ cnt,
                              const uint8_t* __restrict__ solve,
                              const float* __restrict__ lat, const float* __restrict__ lon,
                              uint32_t* keys, int32_t* vals) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int64_t m = solve[p] ? cnt[p] : 0;
  uint32_t key = (uint32_t)p;
  // bucket id in the top 5 + bbits bits, the position inside in the rest
  // (27 - bbits >= 25 bits; a chunk holds at most 2^20 problems)
  const int low = 27 - bbits;
  const uint32_t bucket = bucket_desc(m, bbits);
  if (mode == 1) {
    key = 0xffffffffu - (uint32_t)(m > 0xfffffffe ? 0xfffffffe : m);
  } else if (mode == 2) {
    key = (bucket << low) | (uint32_t)p;
  } else if (mode == 3) {
    const int64_t ij = i0 + p / G;
    float fl = (lat[ij] + 90.0f) * 4.0f;
    float fo = lon[ij] - 360.0f * floorf(lon[ij] / 360.0f);
    fo *= 4.0f;
    const uint32_t li = (uint32_t)fminf(fmaxf(fl, 0.0f), 1023.0f);
    const uint32_t lo = (uint32_t)fminf(fmaxf(fo, 0.0f), 2047.0f);
    const uint32_t z = spread_bits(li) | (spread_bits(lo) << 1);   // 22 bits
    key = (bucket << low) | ((z >> bbits) << 5) | (uint32_t)(p % G & 31);
  }
  keys[p] = key;
  vals[p] = (int32_t)p;
}

void launch_bucket_keys(hipStream_t st, int mode, int bbits, int64_t np, int G, int64_t i0,
                        const int64_t* cnt, const uint8_t* solve, const float* lat,
                        const float* lon, uint32_t* keys, int32_t* vals) {
  if (np <= 0) return;
  k_bucket_keys<<<(unsigned)((np + 255) / 256), 256, 0, st>>>(mode, bbits, np, G, i0, cnt, solve,
                                                             lat, lon, keys, vals);
}

}  // namespace lk
