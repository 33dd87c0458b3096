// tridiag_pair.hip -- choice of the two-problem tridiagonalization by row
// length (a multiple of 8 >= k, or 50 for k = 49 and 50; k <= 50: longer
// rows leave too few registers for two of them per lane).
#include "letkf_internal.h"

namespace lk {

#define LK_DECL(KQ)                                                                     \
  int launch_tridiag_pair_##KQ(hipStream_t, int64_t, const int32_t*, int, int, int,     \
                               const LocDev&, const StateDev&, const VarsDev&,          \
                               const ChunkDev&, const DiagDev&, double*, const double*,  \
                               double*, double*, double*, double*, int*, int*,           \
                               const double*, const int*, int);
LK_DECL(32)
LK_DECL(40)
LK_DECL(48)
LK_DECL(50)
#undef LK_DECL

int launch_tridiag_pair(hipStream_t st, int64_t np, const int32_t* perm, int KQ, int NV, int k,
                        int K, const LocDev& L, const StateDev& S, const VarsDev& V,
                        const ChunkDev& C, const DiagDev& D, double* B, const double* bvec,
                        double* tau, double* dd, double* ee, double* U, int* nvec,
                        int* status, const double* rtab, const int* rtabn, int rnb) {
  switch (KQ) {
    case 32: return launch_tridiag_pair_32(st, np, perm, NV, k, K, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status, rtab, rtabn, rnb);
    case 40: return launch_tridiag_pair_40(st, np, perm, NV, k, K, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status, rtab, rtabn, rnb);
    case 48: return launch_tridiag_pair_48(st, np, perm, NV, k, K, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status, rtab, rtabn, rnb);
    case 50: return launch_tridiag_pair_50(st, np, perm, NV, k, K, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status, rtab, rtabn, rnb);
    default: return 0;
  }
}

}  // namespace lk
