// In-bounds probe of the MFMA operand/output lane layouts this library
// assumes (gfx950).  Each kernel runs one wave: A and B are filled from
// dense row-major inputs following the assumed operand layout, one MFMA is
// issued, and the accumulator registers are written out following the
// assumed output layout.  The test compares with numpy's A @ B.
#include <hip/hip_runtime.h>

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

// 32x32x2 f32: A[m][k] lane m%32 + 32k; B[k][n] lane n + 32k;
// D: reg r of lane l -> row 8*(r/4) + 4*(l/32) + r%4, col l%32
__global__ void p_f32_32x32x2(const float* A, const float* B, float* D) {
  const int l = threadIdx.x;
  const float a = A[(l % 32) * 2 + l / 32];
  const float b = B[(l / 32) * 32 + l % 32];
  f32x16 c = {};
  c = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
  for (int r = 0; r < 16; ++r) D[(8 * (r / 4) + 4 * (l / 32) + r % 4) * 32 + l % 32] = c[r];
}

// 32x32x16 bf16: A[m][k] lane holds k = 8*(l/32)+e, m = l%32; B likewise
__global__ void p_bf16_32x32x16(const float* A, const float* B, float* D) {
  const int l = threadIdx.x;
  bf16x8 a, b;
  for (int e = 0; e < 8; ++e) {
    a[e] = (__bf16)A[(l % 32) * 16 + 8 * (l / 32) + e];
    b[e] = (__bf16)B[(8 * (l / 32) + e) * 32 + l % 32];
  }
  f32x16 c = {};
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  for (int r = 0; r < 16; ++r) D[(8 * (r / 4) + 4 * (l / 32) + r % 4) * 32 + l % 32] = c[r];
}

// 16x16x32 bf16: A[m][k] lane holds k = 8*(l/16)+e, m = l%16; B[k][n] likewise
// (n = l%16); D: reg r of lane l -> row 4*(l/16) + r, col l%16
__global__ void p_bf16_16x16x32(const float* A, const float* B, float* D) {
  const int l = threadIdx.x;
  bf16x8 a, b;
  for (int e = 0; e < 8; ++e) {
    a[e] = (__bf16)A[(l % 16) * 32 + 8 * (l / 16) + e];
    b[e] = (__bf16)B[(8 * (l / 16) + e) * 16 + l % 16];
  }
  f32x4 c = {};
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  for (int r = 0; r < 4; ++r) D[(4 * (l / 16) + r) * 16 + l % 16] = c[r];
}

// 16x16x4 f64: A[m][k] lane m + 16k; D reg r of lane l -> row 4r + l/16
__global__ void p_f64_16x16x4(const double* A, const double* B, double* D) {
  const int l = threadIdx.x;
  const double a = A[(l % 16) * 4 + l / 16];
  const double b = B[(l / 16) * 16 + l % 16];
  f64x4 c = {};
  c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  for (int r = 0; r < 4; ++r) D[(4 * r + l / 16) * 16 + l % 16] = c[r];
}

extern "C" int probe_run(int which, const void* A, const void* B, void* D) {
  size_t na, nb, nd, es = 4;
  if (which == 0) { na = 32 * 2; nb = 2 * 32; nd = 32 * 32; }
  else if (which == 1) { na = 32 * 16; nb = 16 * 32; nd = 32 * 32; }
  else if (which == 3) { na = 16 * 32; nb = 32 * 16; nd = 16 * 16; }
  else { na = 16 * 4; nb = 4 * 16; nd = 16 * 16; es = 8; }
  void *dA, *dB, *dD;
  if (hipMalloc(&dA, na * es) || hipMalloc(&dB, nb * es) || hipMalloc(&dD, nd * es)) return 1;
  hipMemcpy(dA, A, na * es, hipMemcpyHostToDevice);
  hipMemcpy(dB, B, nb * es, hipMemcpyHostToDevice);
  hipMemset(dD, 0, nd * es);
  if (which == 0) p_f32_32x32x2<<<1, 64>>>((const float*)dA, (const float*)dB, (float*)dD);
  else if (which == 1) p_bf16_32x32x16<<<1, 64>>>((const float*)dA, (const float*)dB, (float*)dD);
  else if (which == 3) p_bf16_16x16x32<<<1, 64>>>((const float*)dA, (const float*)dB, (float*)dD);
  else p_f64_16x16x4<<<1, 64>>>((const double*)dA, (const double*)dB, (double*)dD);
  hipError_t e = hipDeviceSynchronize();
  hipMemcpy(D, dD, nd * es, hipMemcpyDeviceToHost);
  hipFree(dA); hipFree(dB); hipFree(dD);
  return e == hipSuccess ? 0 : 2;
}
