/* sizes and field offsets of the C ABI structs, as the C compiler lays them out */
#include <stddef.h>
#include <stdio.h>

#include "../../include/letkf_gpu.h"

#define F(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))
int main(void) {
  printf("letkf_gpu_dims %zu\n", sizeof(letkf_gpu_dims));
  printf("letkf_linearinterp_lat %zu\n", sizeof(letkf_linearinterp_lat));
  printf("letkf_loc_desc %zu\n", sizeof(letkf_loc_desc));
  printf("letkf_gpu_diag %zu\n", sizeof(letkf_gpu_diag));
  printf("letkf_gpu_timing %zu\n", sizeof(letkf_gpu_timing));
  printf("letkf_gpu_tuning %zu\n", sizeof(letkf_gpu_tuning));
  F(letkf_gpu_dims, precision);
  F(letkf_gpu_dims, path);
  F(letkf_loc_desc, hzloc_sat);
  F(letkf_loc_desc, tloc_prof);
  F(letkf_loc_desc, sat_plats);
  F(letkf_loc_desc, vt_scale);
  F(letkf_loc_desc, lev_depth);
  F(letkf_gpu_timing, n_launch);
  F(letkf_gpu_timing, order_ms);
  F(letkf_gpu_tuning, n_tried);
  F(letkf_gpu_tuning, default_ms);
  F(letkf_gpu_tuning, chunk_log2);
  F(letkf_gpu_tuning, xcd);
  F(letkf_gpu_tuning, tile);
  return 0;
}
