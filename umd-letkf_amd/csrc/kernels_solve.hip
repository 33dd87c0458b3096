// kernels_solve.hip -- the FP64 local transform and its application.
//
//   k_product_f64   Yb^T R^-1 Yb and Yb^T R^-1 d of one (gridpoint, group)
//                   on FP64 MFMA (v_mfma_f64_16x16x4f64)
//                   (src/letkf_core.F90:92-98,118-125)
//   k_eigen_apply   (k-1)I + Yb^T R^-1 Yb = V L V^T, T = V sqrt((k-1)/L) V^T
//                   + V L^-1 V^T Yb^T R^-1 d 1^T, x_a = x_b^T T for each slab of
//                   the group (src/letkf_core.F90:100-141,
//                   src/letkf_solver.F90:284-290)
//   k_finalize      recenter, increment clip, RTPS, RTPP, multiplicative
//                   inflation, member bounds, spread
//                   (src/letkf_solver.F90:299-379)
#include <hip/hip_runtime.h>

#include <mutex>

#include <algorithm>

#include "eigen_common.h"
#include "letkf_internal.h"

namespace lk {

typedef double double4_t __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// products: one workgroup (4 waves) per (gridpoint, group).  The kept obs
// are gathered 64 at a time into LDS as FP64 rows y_j (k values, then d_j in
// row k, zero padding to KP).  The upper-triangle 16x16 blocks of the
// (KP x KP) result are spread over the 4 waves; column k of the result is
// Yb^T R^-1 d.
// ---------------------------------------------------------------------------
template <int KP>
__global__ __launch_bounds__(256) void k_product_f64(ObsDev O, int k, ChunkDev C,
                                                     double* __restrict__ B,
                                                     double* __restrict__ bvec) {
  constexpr int NB = KP / 16;
  constexpr int NBLK = NB * (NB + 1) / 2;
  constexpr int PER = (NBLK + 3) / 4;
  constexpr int LD = KP + 16;
  __shared__ double Ys[64 * LD];
  __shared__ double sc[64];
  const int64_t pair = xcd_order(C.xcd, blockIdx.x, gridDim.x);
  if (!C.solve[pair]) return;
  const int64_t m = C.cnt[pair];
  const int64_t off = C.off[pair];
  const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;

  int bi[PER], bj[PER];
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    int idx = wid + 4 * p, r = 0;
    bi[p] = -1; bj[p] = -1;
    if (idx < NBLK) {
      while (idx >= NB - r) { idx -= NB - r; ++r; }
      bi[p] = r;
      bj[p] = r + idx;
    }
  }
  double4_t acc[PER];
#pragma unroll
  for (int p = 0; p < PER; ++p) acc[p] = (double4_t){0.0, 0.0, 0.0, 0.0};

  for (int64_t j0 = 0; j0 < m; j0 += 64) {
    __syncthreads();
    for (int e = tid; e < 64 * KP; e += 256) {
      const int j = e / KP, a = e - j * KP;
      const int64_t jj = j0 + j;
      double v = 0.0;
      if (jj < m) {
        const int32_t s = C.list_idx[off + jj];
        if (a < k) v = (double)O.hx[(int64_t)s * k + a];
        else if (a == k) v = (double)O.dep[s];
      }
      Ys[j * LD + a] = v;
    }
    if (tid < 64) {
      const int64_t jj = j0 + tid;
      double v = 0.0;
      if (jj < m) v = (double)C.list_w[off + jj] / (double)O.err[C.list_idx[off + jj]];
      sc[tid] = v;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < 16; ++kk) {
      const int jr = kk * 4 + (lane >> 4);
      const double s = sc[jr];
      const double* row = &Ys[jr * LD + (lane & 15)];
#pragma unroll
      for (int p = 0; p < PER; ++p) {
        if (bi[p] >= 0) {
          const double a = row[16 * bi[p]] * s;
          const double b = row[16 * bj[p]];
          acc[p] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[p], 0, 0, 0);
        }
      }
    }
  }

  double* Bp = B + pair * (int64_t)k * k;
  double* bp = bvec + pair * (int64_t)k;
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    if (bi[p] < 0) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      // v_mfma_f64_16x16x4f64: register r of lane l holds row 4r + l/16
      const int row = 16 * bi[p] + 4 * r + (lane >> 4);
      const int col = 16 * bj[p] + (lane & 15);
      const double v = acc[p][r];
      if (row < k && col < k) {
        Bp[row + (int64_t)k * col] = v;
        Bp[col + (int64_t)k * row] = v;
      } else if (col == k && row < k) {
        bp[row] = v;
      }
    }
  }
}

void launch_product_f64(hipStream_t st, const ObsDev& O, const StateDev& S,
                        const ChunkDev& C, double* B, double* bvec) {
  const int64_t np = C.ngp * C.G;
  if (np <= 0) return;
  const int k = (int)S.k;
  const int kp = ((k + 1 + 15) / 16) * 16;
  const unsigned nb = (unsigned)np;
  switch (kp) {
    case 16: k_product_f64<16><<<nb, 256, 0, st>>>(O, k, C, B, bvec); break;
    case 32: k_product_f64<32><<<nb, 256, 0, st>>>(O, k, C, B, bvec); break;
    case 48: k_product_f64<48><<<nb, 256, 0, st>>>(O, k, C, B, bvec); break;
    case 64: k_product_f64<64><<<nb, 256, 0, st>>>(O, k, C, B, bvec); break;
    case 80: k_product_f64<80><<<nb, 256, 0, st>>>(O, k, C, B, bvec); break;
    case 96: k_product_f64<96><<<nb, 256, 0, st>>>(O, k, C, B, bvec); break;
    case 112: k_product_f64<112><<<nb, 256, 0, st>>>(O, k, C, B, bvec); break;
    default: k_product_f64<128><<<nb, 256, 0, st>>>(O, k, C, B, bvec); break;
  }
}

// ---------------------------------------------------------------------------
// eigensolve (parallel cyclic Jacobi, round-robin pair ordering) and the
// transform, one workgroup per (gridpoint, group); A and V live in LDS.
// ---------------------------------------------------------------------------
struct Rot { double c, s; int p, q; };

__device__ __forceinline__ void rr_pair(int n, int r, int i, int* p, int* q) {
  // round r of a round-robin tournament over n (even) players
  int a, b;
  if (i == 0) { a = 0; b = (r % (n - 1)) + 1; }
  else { a = ((r + i) % (n - 1)) + 1; b = ((r - i + (n - 1)) % (n - 1)) + 1; }
  *p = a < b ? a : b;
  *q = a < b ? b : a;
}

__global__ __launch_bounds__(256) void k_eigen_apply(LocDev L, StateDev S,
                                                     VarsDev V, ChunkDev C,
                                                     DiagDev D,
                                                     const double* __restrict__ B,
                                                     const double* __restrict__ bvec,
                                                     const int* __restrict__ only_flagged,
                                                     double* __restrict__ ana) {
  extern __shared__ double lds[];
  const int k = (int)S.k;
  const int n = (k + 1) & ~1;               // even number of players
  const int lda = n + 1;
  double* A = lds;                           // n x lda
  double* Vm = A + n * lda;                  // n x lda
  double* lam = Vm + n * lda;                // kMaxK
  double* u = lam + kMaxK;                   // kMaxK
  double* bb = u + kMaxK;                    // kMaxK
  double* fq = bb + kMaxK;                   // kMaxK
  Rot* rot = (Rot*)(fq + kMaxK);             // kMaxK / 2
  __shared__ int s_rot;
  __shared__ int s_nsl;
  __shared__ int s_sl[2048];

  // one block per problem, or, when only flagged problems are solved, a
  // bounded grid that strides over 256-problem runs of the chunk: one load
  // per thread finds whether a run holds a flagged problem, and the block
  // then walks that run's problems together, so the barriers below stay
  // uniform
  const int64_t npair = C.ngp * C.G;
  const int64_t rstep = only_flagged ? (int64_t)gridDim.x * 256 : (int64_t)gridDim.x;
  for (int64_t rbase = only_flagged ? (int64_t)blockIdx.x * 256 : (int64_t)blockIdx.x;
       rbase < npair; rbase += rstep) {
  int64_t rend = rbase + 1;
  if (only_flagged) {
    const int64_t mine = rbase + threadIdx.x;
    const int fl = (mine < npair && C.solve[mine] && only_flagged[mine] != 0) ? 1 : 0;
    if (!__syncthreads_or(fl)) continue;
    rend = rbase + 256 < npair ? rbase + 256 : npair;
  }
  for (int64_t pair = rbase; pair < rend; ++pair) {
  if (!C.solve[pair]) continue;
  if (only_flagged && only_flagged[pair] == 0) continue;
  const int tid = threadIdx.x;
  const int64_t t = pair / C.G;
  const int g = (int)(pair - t * C.G);
  const int64_t ij = C.i0 + t;
  const double* Bp = B + pair * (int64_t)k * k;

  for (int e = tid; e < n * n; e += 256) {
    const int i = e / n, j = e - i * n;
    double v = 0.0;
    // upper triangle only: the tridiagonalization overwrites the lower one
    if (i < k && j < k) v = i <= j ? Bp[i + (int64_t)k * j] : Bp[j + (int64_t)k * i];
    if (i == j) v += (i < k) ? (double)(k - 1) : 1.0;   // dummy player: isolated 1
    A[i * lda + j] = v;
    Vm[i * lda + j] = (i == j) ? 1.0 : 0.0;
  }
  for (int i = tid; i < k; i += 256) bb[i] = bvec[pair * (int64_t)k + i];
  __syncthreads();

  const int half = n / 2;
  for (int sweep = 0; sweep < 40; ++sweep) {
    if (tid == 0) s_rot = 0;
    __syncthreads();
    for (int r = 0; r < n - 1; ++r) {
      if (tid < half) {
        int p, q;
        rr_pair(n, r, tid, &p, &q);
        double c = 1.0, s = 0.0;
        if (q < k) {
          const double apq = A[p * lda + q];
          const double app = A[p * lda + p], aqq = A[q * lda + q];
          if (apq != 0.0 && fabs(apq) > 1e-18 * sqrt(fabs(app) * fabs(aqq))) {
            const double tau = (aqq - app) / (2.0 * apq);
            double tt;
            if (fabs(tau) > 1e150) tt = 0.5 / tau;
            else tt = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + tt * tt);
            s = tt * c;
            s_rot = 1;
          }
        }
        rot[tid] = Rot{c, s, p, q};
      }
      __syncthreads();
      for (int e = tid; e < half * n; e += 256) {       // rows p, q of A
        const int i = e / n, col = e - i * n;
        const Rot R = rot[i];
        if (R.s != 0.0) {
          const double apk = A[R.p * lda + col], aqk = A[R.q * lda + col];
          A[R.p * lda + col] = R.c * apk - R.s * aqk;
          A[R.q * lda + col] = R.s * apk + R.c * aqk;
        }
      }
      __syncthreads();
      for (int e = tid; e < half * n; e += 256) {       // columns p, q of A, V
        const int i = e / n, row = e - i * n;
        const Rot R = rot[i];
        if (R.s != 0.0) {
          const double akp = A[row * lda + R.p], akq = A[row * lda + R.q];
          A[row * lda + R.p] = R.c * akp - R.s * akq;
          A[row * lda + R.q] = R.s * akp + R.c * akq;
          const double vkp = Vm[row * lda + R.p], vkq = Vm[row * lda + R.q];
          Vm[row * lda + R.p] = R.c * vkp - R.s * vkq;
          Vm[row * lda + R.q] = R.s * vkp + R.c * vkq;
        }
      }
      __syncthreads();
      if (tid < half) {
        const Rot R = rot[tid];
        if (R.s != 0.0) { A[R.p * lda + R.q] = 0.0; A[R.q * lda + R.p] = 0.0; }
      }
      __syncthreads();
    }
    if (s_rot == 0) break;
    __syncthreads();
  }

  // T = V diag(sqrt((k-1)/lam)) V^T + wbar 1^T,  wbar = V diag(1/lam) V^T b
  for (int i = tid; i < k; i += 256) {
    lam[i] = A[i * lda + i];
    fq[i] = sqrt((double)(k - 1) / lam[i]);
  }
  __syncthreads();
  for (int i = tid; i < k; i += 256) {
    double acc = 0.0;
    for (int a = 0; a < k; ++a) acc += Vm[a * lda + i] * bb[a];
    u[i] = acc / lam[i];
  }
  __syncthreads();
  for (int a = tid; a < k; a += 256) {
    double acc = 0.0;
    for (int i = 0; i < k; ++i) acc += Vm[a * lda + i] * u[i];
    bb[a] = acc;                               // wbar
  }
  __syncthreads();
  for (int e = tid; e < k * k; e += 256) {
    const int a = e / k, c = e - a * k;
    double acc = 0.0;
    for (int i = 0; i < k; ++i) acc += Vm[a * lda + i] * fq[i] * Vm[c * lda + i];
    A[a * lda + c] = acc + bb[a];              // T(a, c)
  }
  if (tid == 0) {
    int nsl = 0;
    const int mld = D.mld[ij];
    for (int s = 0; s < S.ns && nsl < 2048; ++s)
      if (slab_group(L, V, s, mld) == g) s_sl[nsl++] = s;
    s_nsl = nsl;
  }
  __syncthreads();
  const int nsl = s_nsl;
  const float* xb0 = S.state + (int64_t)S.k * S.ns * ij;
  double* xa0 = ana + (int64_t)S.k * S.ns * ij;
  for (int e = tid; e < nsl * k; e += 256) {   // x_a = x_b^T T
    const int sl = s_sl[e / k], j = e % k;
    const float* xb = xb0 + (int64_t)k * sl;
    double acc = 0.0;
    for (int i = 0; i < k; ++i) acc += (double)xb[i] * A[i * lda + j];
    xa0[(int64_t)k * sl + j] = acc;
  }
  __syncthreads();                            // LDS reused by the next problem
  }
  }
}

size_t eigen_lds_bytes(int k) {
  const int n = (k + 1) & ~1;
  return sizeof(double) * (2 * (size_t)n * (n + 1) + 4 * kMaxK) + sizeof(Rot) * (kMaxK / 2 + 1);
}

hipError_t launch_eigen_apply_f64(hipStream_t st, const LocDev& L, const StateDev& S,
                                  const VarsDev& V, const ChunkDev& C,
                                  const DiagDev& D, const double* B,
                                  const double* bvec, const int* only_flagged, double* ana) {
  const int64_t np = C.ngp * C.G;
  if (np <= 0) return hipSuccess;
  const size_t lds = eigen_lds_bytes((int)S.k);
  // the dynamic LDS limit is raised once per process, thread-safely; its
  // result is kept and returned by every later launch
  static std::once_flag once;
  static hipError_t attr_err = hipSuccess;
  std::call_once(once, [] {
    attr_err = hipFuncSetAttribute((const void*)k_eigen_apply,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024);
  });
  if (attr_err != hipSuccess) return attr_err;
  const unsigned nb = only_flagged ? (unsigned)std::min<int64_t>(np, 1024) : (unsigned)np;
  k_eigen_apply<<<nb, 256, lds, st>>>(L, S, V, C, D, B, bvec, only_flagged, ana);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// post-processing: one wave per gridpoint, lanes over members (2 per lane
// for k > 64).  Every step of src/letkf_solver.F90:299-379 acts slab by slab,
// so each slab is carried through all of them in turn.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wsum(double v) { return wave_sum_dpp(v); }

__global__ __launch_bounds__(256) void k_finalize(LocDev L, StateDev S, VarsDev V,
                                                  ChunkDev C, DiagDev D, float rtps,
                                                  float rtpp, float mul,
                                                  double* __restrict__ ana,
                                                  double* __restrict__ amean,
                                                  double* __restrict__ asprd,
                                                  bool skip_solved) {
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t t = (int64_t)blockIdx.x * 4 + wid;
  if (t >= C.ngp) return;
  const int64_t ij = C.i0 + t;
  const int k = (int)S.k;
  const int64_t ns = S.ns;
  const float* xb0 = S.state + (int64_t)k * ns * ij;
  double* xa0 = ana + (int64_t)k * ns * ij;
  const bool m0 = lane < k, m1 = lane + 64 < k;
  if (S.mask[ij]) {
    // the background unchanged: the gridpoint's k x ns block is contiguous,
    // copied 512 elements per round with 8 loads of a lane in flight
    const int64_t n = (int64_t)k * ns;
    for (int64_t e0 = 0; e0 < n; e0 += 512) {
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = xb0[min(e0 + 64 * u + lane, n - 1)];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int64_t e = e0 + 64 * u + lane;
        if (e < n) xa0[e] = (double)v[u];
      }
    }
    for (int64_t s = lane; s < ns; s += 64) {
      amean[ns * ij + s] = (double)S.mean[ns * ij + s];
      asprd[ns * ij + s] = (double)S.sprd[ns * ij + s];
    }
    return;
  }
  const int mld = D.mld[ij];
  if (skip_solved) {
    // nothing to do where every slab's group was solved (and post-processed
    // by the apply): the slabs' flags are checked 64 at a time
    bool any = false;
    for (int64_t s0 = 0; s0 < ns && !any; s0 += 64) {
      bool need = false;
      if (s0 + lane < ns) {
        const int g = slab_group(L, V, (int)(s0 + lane), mld);
        need = !(g >= 0 && C.solve[t * C.G + g]);
      }
      any = __ballot(need) != 0;
    }
    if (!any) return;
  }
  const double dk = (double)k, dk1 = (double)(k - 1);
  // a slab's inputs are loaded two slabs ahead of its post-processing, so
  // that the loads of later slabs are in flight during the wave sums
  struct SlabIn {
    double x0, x1;      // transformed perturbation, or background member
    float b0, b1;       // background member
    float bm, bs;       // background mean and spread
    bool skip;          // post-processed elsewhere
  };
  auto fetch = [&](int64_t s) {
    SlabIn r;
    const int g = slab_group(L, V, (int)s, mld);
    const bool solved = g >= 0 && C.solve[t * C.G + g];
    r.skip = skip_solved && solved;
    if (r.skip) return r;
    r.b0 = m0 ? xb0[k * s + lane] : 0.0f;
    r.b1 = m1 ? xb0[k * s + lane + 64] : 0.0f;
    r.x0 = solved ? (m0 ? xa0[k * s + lane] : 0.0) : (double)r.b0;
    r.x1 = solved ? (m1 ? xa0[k * s + lane + 64] : 0.0) : (double)r.b1;
    r.bm = S.mean[ns * ij + s];
    r.bs = S.sprd[ns * ij + s];
    return r;
  };
  SlabIn cur = fetch(0);
  SlabIn nxt = ns > 1 ? fetch(1) : cur;
  for (int64_t s = 0; s < ns; ++s) {
    const SlabIn in = cur;
    cur = nxt;
    if (s + 2 < ns) nxt = fetch(s + 2);
    if (in.skip) continue;
    const double xb_0 = (double)in.b0;
    const double xb_1 = (double)in.b1;
    double x0 = in.x0;
    double x1 = in.x1;
    const double bm = (double)in.bm;
    const double bs = (double)in.bs;
    // recenter (:299-303)
    x0 += bm; x1 += bm;
    double mean = wsum((m0 ? x0 : 0.0) + (m1 ? x1 : 0.0)) / dk;
    x0 -= mean; x1 -= mean;
    const int v = V.slab_var[s];
    // increment clipping (:308-318)
    if (v >= 0) {
      const double inc = (double)V.inc_max[v];
      const double blo = (double)V.bounds[2 * v], bhi = (double)V.bounds[2 * v + 1];
      double r = mean - bm;
      r = fmax(r, -inc);
      r = fmin(r, inc);
      r = fmax(r, blo - bm);
      r = fmin(r, bhi - bm);
      mean = r + bm;
    }
    // RTPS (:322-335)
    if (rtps > 0.0f) {
      const double sa = sqrt(wsum((m0 ? x0 * x0 : 0.0) + (m1 ? x1 * x1 : 0.0)) / dk1);
      if (sa > 0.0) {
        const double f = (double)rtps * (bs - sa) / sa + 1.0;
        x0 *= f; x1 *= f;
      }
    }
    // RTPP (:339-343)
    if (rtpp > 0.0f) {
      const double a = (double)rtpp;
      x0 = a * xb_0 + (1.0 - a) * x0;
      x1 = a * xb_1 + (1.0 - a) * x1;
    }
    // multiplicative (:347-349)
    if (mul != 1.0f) { x0 *= (double)mul; x1 *= (double)mul; }
    // member bounds and re-mean (:357-371)
    if (v >= 0) {
      const double blo = (double)V.bounds[2 * v], bhi = (double)V.bounds[2 * v + 1];
      x0 = fmin(fmax(x0, blo - mean), bhi - mean);
      x1 = fmin(fmax(x1, blo - mean), bhi - mean);
      const double r = wsum((m0 ? x0 : 0.0) + (m1 ? x1 : 0.0)) / dk;
      x0 -= r; x1 -= r;
      mean += r;
    }
    // final spread and members (:375-379)
    const double sp = sqrt(wsum((m0 ? x0 * x0 : 0.0) + (m1 ? x1 * x1 : 0.0)) / dk1);
    if (m0) xa0[k * s + lane] = x0 + mean;
    if (m1) xa0[k * s + lane + 64] = x1 + mean;
    if (lane == 0) { amean[ns * ij + s] = mean; asprd[ns * ij + s] = sp; }
  }
}

void launch_finalize(hipStream_t st, const LocDev& L, const StateDev& S,
                     const VarsDev& V, const ChunkDev& C, const DiagDev& D,
                     const FinArgs& F, double* ana, bool skip_solved) {
  if (C.ngp <= 0) return;
  k_finalize<<<(unsigned)((C.ngp + 3) / 4), 256, 0, st>>>(
      L, S, V, C, D, F.rtps, F.rtpp, F.mul, ana, F.amean, F.asprd, skip_solved);
}

}  // namespace lk
