/*
 * ref32_core.c -- the reference's FP32 transform core, in its order of
 * operations, for the CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY: bench.py times it as the CPU path beside the
 * GPU library (and tests/ check it); nothing in the product links it.
 *
 * r32_core_solve restates letkf_core_solve (src/letkf_core.F90:69-146) in
 * REAL(4): hdxb_rinv = hdxb / rdiag * rloc (:92-94); work1 = hdxb_rinv
 * hdxb^T (sgemm, :97-98); + (k-1)/infl on the diagonal (:101-104); the
 * eigendecomposition that ssyev('V','U') gives (:107-108), here FP32
 * Householder tridiagonalization with accumulated transformations followed
 * by the implicit QL iteration with Wilkinson shifts (the textbook
 * tred2/tql2 pair, Golub & Van Loan 8.3), eigenvalues ascending; Pa =
 * V L^-1 V^T (:111-115); work2 = Pa hdxb_rinv (:118-119, the reference's
 * k x k x m product that the transform itself does not need); work3 =
 * work2 dep (:122-125); trans = V sqrt((k-1)/L) V^T + work3 1^T (:128-141).
 * The GEMMs are plain loops in FP32 with the inner loop over contiguous
 * rows (what a BLAS does in FP32, not its blocking).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "letkf_oracle.h"

#define M2(a, n, i, j) (a)[(size_t)(j) * (n) + (i)]

/* c(k x n) = a(k x p) b(n x p)^T, FP32 accumulation */
static void gemm_nt(int k, int n, int p, const float* a, const float* b, float* c) {
  memset(c, 0, sizeof(float) * (size_t)k * n);
  for (int l = 0; l < p; ++l) {
    const float* al = a + (size_t)k * l;
    for (int j = 0; j < n; ++j) {
      const float bj = b[(size_t)l * n + j];
      float* cj = c + (size_t)k * j;
      for (int i = 0; i < k; ++i) cj[i] += al[i] * bj;
    }
  }
}

/* c(k x n) = a(k x k) b(k x n), FP32 accumulation */
static void gemm_nn(int k, int n, const float* a, const float* b, float* c) {
  memset(c, 0, sizeof(float) * (size_t)k * n);
  for (int j = 0; j < n; ++j) {
    float* cj = c + (size_t)k * j;
    const float* bj = b + (size_t)k * j;
    for (int l = 0; l < k; ++l) {
      const float blj = bj[l];
      const float* al = a + (size_t)k * l;
      for (int i = 0; i < k; ++i) cj[i] += al[i] * blj;
    }
  }
}

static float hyp(float a, float b) {
  const float x = fabsf(a), y = fabsf(b);
  if (x > y) { const float r = y / x; return x * sqrtf(1.0f + r * r); }
  if (y == 0.0f) return 0.0f;
  const float r = x / y;
  return y * sqrtf(1.0f + r * r);
}

/* Symmetric eigendecomposition in FP32.  a (n x n, column-major, both
 * triangles) is overwritten by the eigenvectors (columns), w receives the
 * eigenvalues in ascending order; e is n of work.  Returns 0, or -1 when
 * the QL iteration does not converge in 30 sweeps per eigenvalue. */
int r32_syev(int n, float* a, float* w, float* e) {
#define A(i, j) M2(a, n, i, j)
  float* d = w;
  /* Householder reduction to tridiagonal form, transformations kept */
  for (int i = n - 1; i > 0; --i) {
    const int l = i - 1;
    float h = 0.0f, scale = 0.0f;
    if (l > 0) {
      for (int k = 0; k <= l; ++k) scale += fabsf(A(i, k));
      if (scale == 0.0f) {
        e[i] = A(i, l);
      } else {
        for (int k = 0; k <= l; ++k) { A(i, k) /= scale; h += A(i, k) * A(i, k); }
        float f = A(i, l);
        float g = f >= 0.0f ? -sqrtf(h) : sqrtf(h);
        e[i] = scale * g;
        h -= f * g;
        A(i, l) = f - g;
        f = 0.0f;
        for (int j = 0; j <= l; ++j) {
          A(j, i) = A(i, j) / h;
          g = 0.0f;
          for (int k = 0; k <= j; ++k) g += A(j, k) * A(i, k);
          for (int k = j + 1; k <= l; ++k) g += A(k, j) * A(i, k);
          e[j] = g / h;
          f += e[j] * A(i, j);
        }
        const float hh = f / (h + h);
        for (int j = 0; j <= l; ++j) {
          f = A(i, j);
          g = e[j] - hh * f;
          e[j] = g;
          for (int k = 0; k <= j; ++k) A(j, k) -= f * e[k] + g * A(i, k);
        }
      }
    } else {
      e[i] = A(i, l);
    }
    d[i] = h;
  }
  d[0] = 0.0f;
  e[0] = 0.0f;
  for (int i = 0; i < n; ++i) {
    const int l = i - 1;
    if (d[i] != 0.0f) {
      for (int j = 0; j <= l; ++j) {
        float g = 0.0f;
        for (int k = 0; k <= l; ++k) g += A(i, k) * A(k, j);
        for (int k = 0; k <= l; ++k) A(k, j) -= g * A(k, i);
      }
    }
    d[i] = A(i, i);
    A(i, i) = 1.0f;
    for (int j = 0; j <= l; ++j) { A(j, i) = 0.0f; A(i, j) = 0.0f; }
  }
  /* implicit QL with shifts on (d, e); rotations applied to the columns */
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  e[n - 1] = 0.0f;
  for (int l = 0; l < n; ++l) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; ++m) {
        const float dd = fabsf(d[m]) + fabsf(d[m + 1]);
        if (fabsf(e[m]) <= 1.2e-7f * dd) break;
      }
      if (m != l) {
        if (iter++ == 30) return -1;
        float g = (d[l + 1] - d[l]) / (2.0f * e[l]);
        float r = hyp(g, 1.0f);
        g = d[m] - d[l] + e[l] / (g + (g >= 0.0f ? fabsf(r) : -fabsf(r)));
        float s = 1.0f, c = 1.0f, p = 0.0f;
        int i;
        for (i = m - 1; i >= l; --i) {
          float f = s * e[i];
          const float b = c * e[i];
          r = hyp(f, g);
          e[i + 1] = r;
          if (r == 0.0f) { d[i + 1] -= p; e[m] = 0.0f; break; }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0f * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          for (int k = 0; k < n; ++k) {
            f = A(k, i + 1);
            A(k, i + 1) = s * A(k, i) + c * f;
            A(k, i) = c * A(k, i) - s * f;
          }
        }
        if (r == 0.0f && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0f;
      }
    } while (m != l);
  }
  /* ascending, like xSYEV */
  for (int i = 0; i < n - 1; ++i) {
    int k = i;
    for (int j = i + 1; j < n; ++j) if (d[j] < d[k]) k = j;
    if (k != i) {
      const float t = d[i]; d[i] = d[k]; d[k] = t;
      for (int r = 0; r < n; ++r) { const float u = A(r, i); A(r, i) = A(r, k); A(r, k) = u; }
    }
  }
  return 0;
#undef A
}

size_t r32_core_work(int k, int m) {
  return (size_t)k * (size_t)(m > 0 ? m : 1) * 2 + (size_t)k * k * 4 + 4 * (size_t)k;
}

int r32_core_solve(int k, int m, const float* hdxb, const float* rdiag, const float* rloc,
                   const float* dep, float infl, float* trans, float* work) {
  float* hr = work;                                   /* hdxb_rinv (k, m) */
  float* work2 = hr + (size_t)k * (m > 0 ? m : 1);   /* (k, m)           */
  float* work1 = work2 + (size_t)k * (m > 0 ? m : 1);
  float* eivec = work1 + (size_t)k * k;
  float* pa = eivec + (size_t)k * k;
  float* tmp = pa + (size_t)k * k;
  float* eival = tmp + (size_t)k * k;
  float* work3 = eival + k;
  float* ework = work3 + k;
  for (int j = 0; j < m; ++j)                                           /* :92-94 */
    for (int a = 0; a < k; ++a) hr[a + (size_t)k * j] = hdxb[a + (size_t)k * j] / rdiag[j] * rloc[j];
  gemm_nt(k, k, m, hr, hdxb, work1);                                    /* :97-98 */
  const float r = (float)(k - 1) * 1.0f / infl;                        /* :101-104 */
  for (int i = 0; i < k; ++i) M2(work1, k, i, i) += r;
  for (int j = 0; j < k; ++j)       /* 'U': the upper triangle defines the matrix */
    for (int i = j + 1; i < k; ++i) M2(work1, k, i, j) = M2(work1, k, j, i);
  if (r32_syev(k, work1, eival, ework)) return -1;                      /* :107-108 */
  memcpy(eivec, work1, sizeof(float) * (size_t)k * k);
  for (int i = 0; i < k; ++i)                                           /* :111-115 */
    for (int a = 0; a < k; ++a) M2(tmp, k, a, i) = M2(eivec, k, a, i) / eival[i];
  gemm_nt(k, k, k, tmp, eivec, pa);
  gemm_nn(k, m, pa, hr, work2);                                         /* :118-119 */
  for (int a = 0; a < k; ++a) work3[a] = 0.0f;                          /* :122-125 */
  for (int j = 0; j < m; ++j)
    for (int a = 0; a < k; ++a) work3[a] = work3[a] + work2[a + (size_t)k * j] * dep[j];
  for (int j = 0; j < k; ++j) {                                         /* :128-132 */
    const float s = sqrtf((float)(k - 1) / eival[j]);
    for (int a = 0; a < k; ++a) M2(tmp, k, a, j) = M2(eivec, k, a, j) * s;
  }
  gemm_nt(k, k, k, tmp, eivec, trans);
  for (int j = 0; j < k; ++j)                                           /* :138-140 */
    for (int a = 0; a < k; ++a) M2(trans, k, a, j) += work3[a];
  return 0;
}
