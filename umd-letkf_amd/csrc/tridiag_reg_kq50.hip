// tridiag_reg_kq50.hip -- register-resident tridiagonalization, rows of 50
// (ensembles of 49 and 50, which would otherwise pad rows to 56).
#include "tridiag_reg.h"

namespace lk {

bool launch_tridiag_reg_50(hipStream_t st, unsigned nb, int NV, int k, int K, const LocDev& L,
                           const StateDev& S, const VarsDev& V, const ChunkDev& C,
                           const DiagDev& D, double* B, const double* bvec, double* tau,
                           double* dd, double* ee, double* U, int* nvec, int* status) {
  return launch_tridiag_reg_kq<50>(st, nb, NV, k, K, L, S, V, C, D, B, bvec, tau, dd, ee, U,
                                    nvec, status);
}

}  // namespace lk
