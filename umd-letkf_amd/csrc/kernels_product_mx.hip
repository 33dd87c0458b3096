// kernels_product_mx.hip -- Yb^T R^-1 Yb and Yb^T R^-1 d on low-precision
// MFMA with FP32 accumulation (the batched mixed-precision path).
//
// One wave per (gridpoint, group) problem.  Problems are visited by
// power-of-two buckets of their kept-observation count m, largest first
// (a permutation sorted on the device), in gridpoint order inside a bucket,
// and blocks are remapped so that each XCD walks a contiguous range.  The kept columns y_j (k values) are loaded
// straight from the sorted hx array into MFMA operand registers, 16
// observations per step, the next step's loads issued before the current
// step's MFMAs.  Both MFMA operands are sqrt(rloc_j/err_j) y_j (the square
// root written by the search's fill pass), so each value is rounded/split
// once.  Row k of the padded operand carries the scaled d_j, so column k of
// the result is Yb^T R^-1 d.  Only the upper-triangle 32x32 blocks are
// computed.
//
//   PREC_FP32    v_mfma_f32_32x32x2_f32   (8 MFMA per block per 16 obs)
//   PREC_BF16    v_mfma_f32_32x32x16_bf16 (1 per block per 16 obs)
//   PREC_BF16X3  split x = hi + lo in BF16, hi*hi + hi*lo + lo*hi (3 per block)
//
// The FP32 result is widened to FP64 for the eigensolve.
#include <hip/hip_runtime.h>

#include "letkf_internal.h"

namespace lk {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <int NB>
struct MxFrag {
  float y[NB][8];   // sqrt(rloc_j/err_j) y_j[32 rb + lane%32], the lane's 8 obs
};

// observation e (0..7) of a lane in a 16-observation step: the BF16 MFMA
// takes 8 consecutive K per lane half, the FP32 MFMA interleaves halves
template <int PREC>
__device__ __forceinline__ int obs_slot(int half, int e) {
  return PREC == LETKF_PREC_FP32 ? 2 * e + half : 8 * half + e;
}

template <int NB, int PREC>
__device__ __forceinline__ void load_frag(MxFrag<NB>& F, const ObsDev& O, const ChunkDev& C,
                                          int64_t off, int64_t m, int64_t j0, int k, int lane) {
  const int half = lane >> 5, i = lane & 31;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int64_t j = j0 + obs_slot<PREC>(half, e);
    int32_t s = 0;
    float w = 0.0f;
    if (j < m) {
      s = C.list_idx[off + j];
      w = C.list_w[off + j];            // sqrt(rloc / err), written by the search
    }
#pragma unroll
    for (int rb = 0; rb < NB; ++rb) {
      const int a = 32 * rb + i;
      float v = 0.0f;
      if (j < m) {
        if (a < k) v = O.hx[(int64_t)s * k + a];
        else if (a == k) v = O.dep[s];
      }
      F.y[rb][e] = v * w;
    }
  }
}

__device__ __forceinline__ bf16x8 to_bf16(const float* v) {
  bf16x8 r;
#pragma unroll
  for (int e = 0; e < 8; ++e) r[e] = (__bf16)v[e];
  return r;
}
__device__ __forceinline__ void split_bf16(const float* v, bf16x8& hi, bf16x8& lo) {
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const __bf16 h = (__bf16)v[e];
    hi[e] = h;
    lo[e] = (__bf16)(v[e] - (float)h);
  }
}

template <int KP, int PREC>
__global__ __launch_bounds__(64) void k_product_mx(ObsDev O, int k, ChunkDev C,
                                                   const int32_t* __restrict__ perm,
                                                   double* __restrict__ B,
                                                   double* __restrict__ bvec) {
  constexpr int NB = KP / 32;
  constexpr int NBLK = NB * (NB + 1) / 2;
  const int64_t pair = perm[C.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x];
  if (!C.solve[pair]) return;
  const int64_t m = C.cnt[pair];
  const int64_t off = C.off[pair];
  const int lane = threadIdx.x;

  f32x16 acc[NBLK];
#pragma unroll
  for (int b = 0; b < NBLK; ++b) acc[b] = (f32x16){};

  MxFrag<NB> cur, nxt;
  load_frag<NB, PREC>(cur, O, C, off, m, 0, k, lane);
  for (int64_t j0 = 0; j0 < m; j0 += 16) {
    if (j0 + 16 < m) load_frag<NB, PREC>(nxt, O, C, off, m, j0 + 16, k, lane);
    // both operands are sqrt(R^-1 rloc) y: A B^T = Yb^T R^-1 Yb, and row k
    // (sqrt(R^-1 rloc) d) gives Yb^T R^-1 d in column k
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
          acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lo[rb], hi[cb], acc[b], 0, 0, 0);
          acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi[rb], lo[cb], acc[b], 0, 0, 0);
          acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi[rb], hi[cb], acc[b], 0, 0, 0);
        }
    }
    cur = nxt;
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
        const double v = (double)acc[b][r];
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
                      const ChunkDev& C, const int32_t* perm, double* B, double* bvec) {
  if (prec == LETKF_PREC_FP32)
    k_product_mx<KP, LETKF_PREC_FP32><<<nb, 64, 0, st>>>(O, k, C, perm, B, bvec);
  else if (prec == LETKF_PREC_BF16)
    k_product_mx<KP, LETKF_PREC_BF16><<<nb, 64, 0, st>>>(O, k, C, perm, B, bvec);
  else
    k_product_mx<KP, LETKF_PREC_BF16X3><<<nb, 64, 0, st>>>(O, k, C, perm, B, bvec);
}

void launch_product_mx(hipStream_t st, int prec, const ObsDev& O, const StateDev& S,
                       const ChunkDev& C, const int32_t* perm, double* B, double* bvec) {
  const int64_t np = C.ngp * C.G;
  if (np <= 0) return;
  const int k = (int)S.k;
  const int kp = ((k + 1 + 31) / 32) * 32;
  const unsigned nb = (unsigned)np;
  switch (kp) {
    case 32: launch_kp<32>(st, prec, nb, O, k, C, perm, B, bvec); break;
    case 64: launch_kp<64>(st, prec, nb, O, k, C, perm, B, bvec); break;
    case 96: launch_kp<96>(st, prec, nb, O, k, C, perm, B, bvec); break;
    default: launch_kp<128>(st, prec, nb, O, k, C, perm, B, bvec); break;
  }
}

// order of the problems of a chunk: mode 0 gridpoint order, mode 1 longest
// first (descending m), mode 2 power-of-two buckets of m, largest first,
// gridpoint order inside a bucket
__global__ void k_bucket_keys(int mode, int64_t np, const int64_t* __restrict__ cnt,
                              const uint8_t* __restrict__ solve, uint32_t* keys,
                              int32_t* vals) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int64_t m = solve[p] ? cnt[p] : 0;
  uint32_t key = (uint32_t)p;
  if (mode == 1) {
    key = 0xffffffffu - (uint32_t)(m > 0xfffffffe ? 0xfffffffe : m);
  } else if (mode == 2) {
    const uint32_t b = m > 0 ? 64u - (uint32_t)__clzll((unsigned long long)m) : 0u;
    key = ((31u - (b > 31u ? 31u : b)) << 27) | (uint32_t)p;
  }
  keys[p] = key;
  vals[p] = (int32_t)p;
}

void launch_bucket_keys(hipStream_t st, int mode, int64_t np, const int64_t* cnt,
                        const uint8_t* solve, uint32_t* keys, int32_t* vals) {
  if (np <= 0) return;
  k_bucket_keys<<<(unsigned)((np + 255) / 256), 256, 0, st>>>(mode, np, cnt, solve, keys, vals);
}

}  // namespace lk
