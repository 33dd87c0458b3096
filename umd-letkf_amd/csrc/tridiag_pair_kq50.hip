// tridiag_pair_kq50.hip -- two-problem tridiagonalization, rows of 50.
#include "tridiag_pair.h"

namespace lk {

int launch_tridiag_pair_50(hipStream_t st, int64_t np, const int32_t* perm, int NV, int k, int K,
                           const LocDev& L, const StateDev& S, const VarsDev& V,
                           const ChunkDev& C, const DiagDev& D, double* B, const double* bvec,
                           double* tau, double* dd, double* ee, double* U, int* nvec,
                           int* status, const double* rtab, const int* rtabn, int rnb) {
  return launch_tridiag_pair_kq<50>(st, np, perm, NV, k, K, L, S, V, C, D, B, bvec, tau, dd, ee,
                                     U, nvec, status, rtab, rtabn, rnb);
}

}  // namespace lk
