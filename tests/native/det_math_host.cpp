// Host build of the library's det_math.h, so that a CPU test can check that
// it produces the same bits as the oracle's copy of the same functions.
#include "../../umd-letkf_amd/csrc/det_math.h"

extern "C" {
void dmh_sincosd(double d, double* s, double* c) { dm::sincosd(d, s, c); }
double dmh_asin(double x) { return dm::asin01(x); }
double dmh_exp_neg(double x) { return dm::exp_neg(x); }
double dmh_thr2(float r) { return dm::thr2(r); }
float dmh_dist(double c2) { return dm::dist(c2); }
float dmh_gc(float z, float L) { return dm::gc(z, L); }
void dmh_xyz(float lat, float lon, double* out) { dm::xyz(lat, lon, out, out + 1, out + 2); }
double dmh_chord2(const double* a, const double* b) {
  return dm::chord2(a[0], a[1], a[2], b[0], b[1], b[2]);
}
uint64_t dmh_mix64(uint64_t x) { return dm::mix64(x); }
float dmh_interp(int n, const float* lat, const float* dist, float x) {
  return dm::interp(n, lat, dist, x);
}
}
