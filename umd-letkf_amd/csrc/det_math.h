// det_math.h -- elementary functions with a fixed sequence of IEEE
// operations (+, -, *, / and sqrt only, no contraction), so that the GPU
// and any host copy of them produce the same bits.  They define this
// library's observation-distance and localization arithmetic:
//
//  * lat/lon -> unit vector: sin/cos of degrees reduced exactly by
//    quadrants, then Taylor polynomials on [-pi/4, pi/4];
//  * inclusion in the search cap: squared chord <= (2 sin(R/2Re))^2;
//  * great-circle distance 2 Re asin(chord/2), rounded to FP32 like the
//    distance the reference hands to its localizer (src/letkf_obs.F90:565);
//  * Gaspari-Cohn with the reference's truncated coefficients
//    (src/letkf_loc.F90:348-378), FP64 inside, FP32 result;
//  * linearinterp_lat get_dist in FP32 (src/letkf_loc.F90:302-320);
//  * exp(x) for x <= 0 (vertical Gaussian of LOC_VGAUSS).
//
// The reference's radius search lives in the absent geoKdTree submodule;
// Earth radius 6,371,000 m and the chord test are this library's own
// definition of "within radius".
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define DM_FN __host__ __device__ inline

namespace dm {

constexpr double kRE = 6371000.0;
constexpr double k2RE = 12742000.0;
constexpr double kPIO2 = 0x1.921fb54442d18p+0;
constexpr double kPIO4 = 0.78539816339744828;
constexpr double kDEG2RAD = 0x1.1df46a2529d39p-6;
constexpr double kSQRT03 = 0x1.186f174f88472p-1;
constexpr double kINVLN2 = 0x1.71547652b82fep+0;
constexpr double kLN2HI = 0x1.62e4200000000p-1;
constexpr double kLN2LO = 0x1.fdf473de6af28p-22;
constexpr float kSTDEV2MAX = 3.6514837741851807f;   // SQRT(40.0/3.0) in REAL

DM_FN double sin_core(double x) {
#pragma clang fp contract(off)
  const double x2 = x * x;
  double p = 0x1.71b8ef6dcf572p-66;
  p = p * x2 + -0x1.2f49b46814157p-57;
  p = p * x2 + 0x1.952c77030ad4ap-49;
  p = p * x2 + -0x1.ae7f3e733b81fp-41;
  p = p * x2 + 0x1.6124613a86d09p-33;
  p = p * x2 + -0x1.ae64567f544e4p-26;
  p = p * x2 + 0x1.71de3a556c734p-19;
  p = p * x2 + -0x1.a01a01a01a01ap-13;
  p = p * x2 + 0x1.1111111111111p-7;
  p = p * x2 + -0x1.5555555555555p-3;
  return x + (x * x2) * p;
}

DM_FN double cos_core(double x) {
#pragma clang fp contract(off)
  const double x2 = x * x;
  double p = -0x1.0ce396db7f853p-70;
  p = p * x2 + 0x1.e542ba4020225p-62;
  p = p * x2 + -0x1.6827863b97d97p-53;
  p = p * x2 + 0x1.ae7f3e733b81fp-45;
  p = p * x2 + -0x1.93974a8c07c9dp-37;
  p = p * x2 + 0x1.1eed8eff8d898p-29;
  p = p * x2 + -0x1.27e4fb7789f5cp-22;
  p = p * x2 + 0x1.a01a01a01a01ap-16;
  p = p * x2 + -0x1.6c16c16c16c17p-10;
  p = p * x2 + 0x1.5555555555555p-5;
  p = p * x2 + -0x1.0000000000000p-1;
  return 1.0 + x2 * p;
}

DM_FN void sincosd(double deg, double* s, double* c) {
#pragma clang fp contract(off)
  const double q = floor(deg / 90.0 + 0.5);
  const double r = deg - q * 90.0;
  const double x = r * kDEG2RAD;
  const double sn = sin_core(x), cs = cos_core(x);
  long iq = ((long)q % 4 + 4) % 4;
  if (iq == 0) { *s = sn; *c = cs; }
  else if (iq == 1) { *s = cs; *c = -sn; }
  else if (iq == 2) { *s = -sn; *c = -cs; }
  else { *s = -cs; *c = sn; }
}

DM_FN double sin_rad(double x) {   // 0 <= x < pi/2
#pragma clang fp contract(off)
  if (x <= kPIO4) return sin_core(x);
  return cos_core(kPIO2 - x);
}

DM_FN double asin_core(double x) {  // 0 <= x <= 0.5
#pragma clang fp contract(off)
  const double x2 = x * x;
  double p;
  if (x <= 0.2) {
    p = 0x1.51ba308d3dcb1p-8;
    p = p * x2 + 0x1.782dda12f684cp-8;
    p = p * x2 + 0x1.a6863d70a3d71p-8;
    p = p * x2 + 0x1.df3bd37a6f4dfp-8;
    p = p * x2 + 0x1.12ef3cf3cf3cfp-7;
    p = p * x2 + 0x1.3fde50d79435ep-7;
    p = p * x2 + 0x1.7a87878787878p-7;
    p = p * x2 + 0x1.c99999999999ap-7;
    p = p * x2 + 0x1.1c4ec4ec4ec4fp-6;
    p = p * x2 + 0x1.6e8ba2e8ba2e9p-6;
    p = p * x2 + 0x1.f1c71c71c71c7p-6;
    p = p * x2 + 0x1.6db6db6db6db7p-5;
    p = p * x2 + 0x1.3333333333333p-4;
    p = p * x2 + 0x1.5555555555555p-3;
  } else {
    p = 0x1.e82be60d9127ep-10;
    p = p * x2 + 0x1.018f963c229bfp-9;
    p = p * x2 + 0x1.1052bc5fa960ap-9;
    p = p * x2 + 0x1.208d3570ae5a6p-9;
    p = p * x2 + 0x1.3275586c5f2f0p-9;
    p = p * x2 + 0x1.464c0950f7d47p-9;
    p = p * x2 + 0x1.5c5f56efaaaabp-9;
    p = p * x2 + 0x1.750de64d7d05fp-9;
    p = p * x2 + 0x1.90cb77f60c7cep-9;
    p = p * x2 + 0x1.b026f57b13b14p-9;
    p = p * x2 + 0x1.d3d2a8e0dd67dp-9;
    p = p * x2 + 0x1.fcaf8fb6db6dbp-9;
    p = p * x2 + 0x1.15ee9d45d1746p-8;
    p = p * x2 + 0x1.31683bdef7bdfp-8;
    p = p * x2 + 0x1.51ba308d3dcb1p-8;
    p = p * x2 + 0x1.782dda12f684cp-8;
    p = p * x2 + 0x1.a6863d70a3d71p-8;
    p = p * x2 + 0x1.df3bd37a6f4dfp-8;
    p = p * x2 + 0x1.12ef3cf3cf3cfp-7;
    p = p * x2 + 0x1.3fde50d79435ep-7;
    p = p * x2 + 0x1.7a87878787878p-7;
    p = p * x2 + 0x1.c99999999999ap-7;
    p = p * x2 + 0x1.1c4ec4ec4ec4fp-6;
    p = p * x2 + 0x1.6e8ba2e8ba2e9p-6;
    p = p * x2 + 0x1.f1c71c71c71c7p-6;
    p = p * x2 + 0x1.6db6db6db6db7p-5;
    p = p * x2 + 0x1.3333333333333p-4;
    p = p * x2 + 0x1.5555555555555p-3;
  }
  return x + (x * x2) * p;
}

DM_FN double asin01(double x) {    // 0 <= x <= 1
#pragma clang fp contract(off)
  if (x <= 0.5) return asin_core(x);
  const double t = sqrt((1.0 - x) * 0.5);
  return kPIO2 - 2.0 * asin_core(t);
}

DM_FN double exp_neg(double x) {   // x <= 0
#pragma clang fp contract(off)
  if (x < -745.0) return 0.0;
  const double n = floor(x * kINVLN2 + 0.5);
  const double r = (x - n * kLN2HI) - n * kLN2LO;
  double p = 0x1.93974a8c07c9dp-37;
  p = p * r + 0x1.6124613a86d09p-33;
  p = p * r + 0x1.1eed8eff8d898p-29;
  p = p * r + 0x1.ae64567f544e4p-26;
  p = p * r + 0x1.27e4fb7789f5cp-22;
  p = p * r + 0x1.71de3a556c734p-19;
  p = p * r + 0x1.a01a01a01a01ap-16;
  p = p * r + 0x1.a01a01a01a01ap-13;
  p = p * r + 0x1.6c16c16c16c17p-10;
  p = p * r + 0x1.1111111111111p-7;
  p = p * r + 0x1.5555555555555p-5;
  p = p * r + 0x1.5555555555555p-3;
  p = p * r + 0x1.0000000000000p-1;
  p = p * r + 0x1.0000000000000p+0;
  p = p * r + 0x1.0000000000000p+0;
  return ldexp(p, (int)n);
}

DM_FN void xyz(float lat, float lon, double* x, double* y, double* z) {
#pragma clang fp contract(off)
  double sl, cl, so, co;
  sincosd((double)lat, &sl, &cl);
  sincosd((double)lon, &so, &co);
  *x = cl * co;
  *y = cl * so;
  *z = sl;
}

DM_FN double chord2(double ax, double ay, double az, double bx, double by, double bz) {
#pragma clang fp contract(off)
  const double dx = ax - bx, dy = ay - by, dz = az - bz;
  return (dx * dx + dy * dy) + dz * dz;
}

DM_FN double thr2(float radius_m) {
#pragma clang fp contract(off)
  const double half = (double)radius_m / k2RE;
  if (!(half < kPIO2)) return 5.0;
  const double t = 2.0 * sin_rad(half);
  return t * t;
}

DM_FN float dist(double c2) {
#pragma clang fp contract(off)
  double h = sqrt(c2) * 0.5;
  if (h > 1.0) h = 1.0;
  return (float)(k2RE * asin01(h));
}

DM_FN float gc(float z, float L) {
#pragma clang fp contract(off)
  const double c = (double)L / kSQRT03;
  const double abs_z = fabs((double)z);
  const double z_c = abs_z / c;
  double res;
  if (abs_z >= 2.0 * c) {
    res = 0.0;
  } else {
    const double z2 = z_c * z_c, z3 = z2 * z_c, z4 = z2 * z2, z5 = z4 * z_c;
    if (abs_z > c) {
      res = 0.08333 * z5 - 0.50000 * z4 + 0.62500 * z3 + 1.66667 * z2
            - 5.00000 * z_c + 4.0 - 0.66667 * c / abs_z;
    } else {
      res = -0.25000 * z5 + 0.50000 * z4 + 0.62500 * z3 - 1.66667 * z2 + 1.0;
    }
  }
  return (float)res;
}

// linearinterp_lat get_dist on an already-completed table (n >= 1)
DM_FN float interp(int n, const float* lat, const float* dist, float x) {
#pragma clang fp contract(off)
  const float a = fabsf(x);
  int i = 0;
  while (i < n - 1 && a > lat[i]) ++i;
  if (i == 0) return dist[0];
  const float d = (a - lat[i - 1]) / (lat[i] - lat[i - 1]);
  return d * dist[i] + (1.0f - d) * dist[i - 1];
}

DM_FN uint64_t mix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace dm
