// transform_blk_bs8.hip -- k_blk_tridiag and k_blk_back for 8x8 blocks
// (k <= 128); one instantiation per file keeps the build parallel.
#include "transform_blk.h"

namespace lk {

void launch_blk_tridiag_bs8(hipStream_t st, unsigned nb, int k, const ChunkDev& C, double* B,
                             double* tau, double* dd, double* ee) {
  launch_blk_tridiag_bs<8>(st, nb, k, C, B, tau, dd, ee);
}

void launch_blk_back_bs8(hipStream_t st, unsigned nb, int k, const ChunkDev& C, double* B,
                          const double* bvec, const double* tau) {
  launch_blk_back_bs<8>(st, nb, k, C, B, bvec, tau);
}

}  // namespace lk
