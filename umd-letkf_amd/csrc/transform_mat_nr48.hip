// transform_mat_nr48.hip -- k_mat_tridiag and k_mat_back for rows of 48
// per half (2 x 48 >= k); one instantiation per file keeps the build
// parallel.
#include "transform_mat.h"

namespace lk {

void launch_mat_tridiag_nr48(hipStream_t st, unsigned nb, int k, const ChunkDev& C, double* B,
                              double* tau, double* dd, double* ee) {
  launch_mat_tridiag_nr<48>(st, nb, k, C, B, tau, dd, ee);
}

void launch_mat_back_nr48(hipStream_t st, unsigned nb, int k, const ChunkDev& C, double* B,
                           const double* bvec, const double* tau) {
  launch_mat_back_nr<48>(st, nb, k, C, B, bvec, tau);
}

}  // namespace lk
