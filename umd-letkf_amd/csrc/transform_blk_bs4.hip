// transform_blk_bs4.hip -- k_blk_tridiag and k_blk_back for 4x4 blocks
// (k <= 64); one instantiation per file keeps the build parallel.
#include "transform_blk.h"

namespace lk {

void launch_blk_tridiag_bs4(hipStream_t st, unsigned nb, int k, const ChunkDev& C, double* B,
                             double* tau, double* dd, double* ee) {
  launch_blk_tridiag_bs<4>(st, nb, k, C, B, tau, dd, ee);
}

void launch_blk_back_bs4(hipStream_t st, unsigned nb, int k, const ChunkDev& C, double* B,
                          const double* bvec, const double* tau) {
  launch_blk_back_bs<4>(st, nb, k, C, B, bvec, tau);
}

}  // namespace lk
