// transform_mat_nr40.hip -- k_transform_mat for rows of 40 per half
// (2 x 40 >= k); one instantiation per file keeps the build parallel.
#include "transform_mat.h"

namespace lk {

void launch_transform_nr40(hipStream_t st, unsigned nb, int k, const ChunkDev& C, double* B,
                            const double* bvec, int* status, const RatTab& R) {
  launch_transform_nr<40>(st, nb, k, C, B, bvec, status, R);
}

}  // namespace lk
