// tridiag_reg.hip -- choice of the register-resident tridiagonalization by
// row length (a multiple of 8 >= k, or 50 for k = 49 and 50).
#include "letkf_internal.h"

namespace lk {

#define LK_DECL(KQ)                                                                   \
  bool launch_tridiag_reg_##KQ(hipStream_t, unsigned, int, int, int, const LocDev&,   \
                               const StateDev&, const VarsDev&, const ChunkDev&,      \
                               const DiagDev&, double*, const double*, double*,       \
                               double*, double*, double*, int*, int*);
LK_DECL(32)
LK_DECL(40)
LK_DECL(48)
LK_DECL(50)
LK_DECL(56)
LK_DECL(64)
#undef LK_DECL

bool launch_tridiag_reg(hipStream_t st, unsigned nb, int KQ, int NV, int k, int K,
                        const LocDev& L, const StateDev& S, const VarsDev& V,
                        const ChunkDev& C, const DiagDev& D, double* B, const double* bvec,
                        double* tau, double* dd, double* ee, double* U, int* nvec,
                        int* status) {
  switch (KQ) {
    case 32: return launch_tridiag_reg_32(st, nb, NV, k, K, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status);
    case 40: return launch_tridiag_reg_40(st, nb, NV, k, K, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status);
    case 48: return launch_tridiag_reg_48(st, nb, NV, k, K, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status);
    case 50: return launch_tridiag_reg_50(st, nb, NV, k, K, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status);
    case 56: return launch_tridiag_reg_56(st, nb, NV, k, K, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status);
    case 64: return launch_tridiag_reg_64(st, nb, NV, k, K, L, S, V, C, D, B, bvec, tau, dd, ee, U, nvec, status);
    default: return false;
  }
}

}  // namespace lk
