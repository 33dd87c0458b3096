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

// Operand registers of one 16-observation step, without row or range tests:
// the kept list of a problem starts at a multiple of 8 entries and is
// followed by at least 8 readable entries, and the padded columns hxp carry
// the departure in row k and zeros above it.  Entries past m read column 0
// with weight 0.
template <int NB, int PREC>
__device__ __forceinline__ void load_frag(MxFrag<NB>& F, const ObsDev& O,
                                          __amdgpu_buffer_rsrc_t hxr, const ChunkDev& C,
                                          int64_t off, int64_t m, int64_t j0, int lane) {
  const int half = lane >> 5, i = lane & 31;
  int32_t sidx[8];
  float w[8];
  if (PREC == LETKF_PREC_FP32) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int64_t j = off + j0 + obs_slot<PREC>(half, e);
      sidx[e] = C.list_idx[j];
      w[e] = C.list_w[j];                 // sqrt(rloc / err), written by the search
    }
  } else {
    const int4* pi = reinterpret_cast<const int4*>(C.list_idx + off + j0 + 8 * half);
    const float4* pw = reinterpret_cast<const float4*>(C.list_w + off + j0 + 8 * half);
    const int4 a0 = pi[0], a1 = pi[1];
    const float4 b0 = pw[0], b1 = pw[1];
    sidx[0] = a0.x; sidx[1] = a0.y; sidx[2] = a0.z; sidx[3] = a0.w;
    sidx[4] = a1.x; sidx[5] = a1.y; sidx[6] = a1.z; sidx[7] = a1.w;
    w[0] = b0.x; w[1] = b0.y; w[2] = b0.z; w[3] = b0.w;
    w[4] = b1.x; w[5] = b1.y; w[6] = b1.z; w[7] = b1.w;
  }
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const bool valid = j0 + obs_slot<PREC>(half, e) < m;
    const float ww = valid ? w[e] : 0.0f;
    // byte offset of the lane's row in the column; past the buffer's range
    // (returns 0) for entries beyond m
    const uint32_t vo = valid ? ((uint32_t)sidx[e] * (uint32_t)O.kp + (uint32_t)i) * 4u
                              : 0x80000000u;
#pragma unroll
    for (int rb = 0; rb < NB; ++rb)
      F.y[rb][e] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(hxr, vo, 128 * rb, 0)) * ww;
  }
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// two values per conversion (v_cvt_pk_bf16_f32, round to nearest even)
__device__ __forceinline__ bf16x8 to_bf16(const float* v) {
  bf16x8 r;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const bf16x2 h = __builtin_convertvector((f32x2){v[2 * p], v[2 * p + 1]}, bf16x2);
    r[2 * p] = h[0];
    r[2 * p + 1] = h[1];
  }
  return r;
}

// x = hi + lo + O(2^-16 |x|): hi = bf16(x), lo = bf16(x - hi); x - hi is exact
// in FP32, and hi widens by a shift of the packed bits
__device__ __forceinline__ void split_bf16(const float* v, bf16x8& hi, bf16x8& lo) {
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const float x0 = v[2 * p], x1 = v[2 * p + 1];
    const bf16x2 h = __builtin_convertvector((f32x2){x0, x1}, bf16x2);
    const uint32_t hu = __builtin_bit_cast(uint32_t, h);
    const float r0 = x0 - __uint_as_float(hu << 16);
    const float r1 = x1 - __uint_as_float(hu & 0xffff0000u);
    const bf16x2 l = __builtin_convertvector((f32x2){r0, r1}, bf16x2);
    hi[2 * p] = h[0];
    hi[2 * p + 1] = h[1];
    lo[2 * p] = l[0];
    lo[2 * p + 1] = l[1];
  }
}

// acc += the products of one 16-observation step
template <int NB, int PREC>
__device__ __forceinline__ void mx_step(f32x16* acc, const MxFrag<NB>& cur) {
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
  // padded operand columns; the host guarantees n * kp * 4 < 2^31
  const __amdgpu_buffer_rsrc_t hxr = __builtin_amdgcn_make_buffer_rsrc(
      (void*)O.hxp, (short)0, (int)(O.n * O.kp * 4), 0x00020000);

  f32x16 acc[NBLK];
#pragma unroll
  for (int b = 0; b < NBLK; ++b) acc[b] = (f32x16){};

  // two fragment buffers in turn: the next step's loads are in flight while
  // the current step's MFMAs run
  MxFrag<NB> fa, fb;
  load_frag<NB, PREC>(fa, O, hxr, C, off, m, 0, lane);
  for (int64_t j0 = 0; j0 < m; j0 += 32) {
    const bool second = j0 + 16 < m;
    if (second) load_frag<NB, PREC>(fb, O, hxr, C, off, m, j0 + 16, lane);
    mx_step<NB, PREC>(acc, fa);
    if (j0 + 32 < m) load_frag<NB, PREC>(fa, O, hxr, C, off, m, j0 + 32, lane);
    if (second) mx_step<NB, PREC>(acc, fb);
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
