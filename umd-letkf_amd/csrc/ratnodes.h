// ratnodes.h -- rational approximation of x^(-1/2) on [1, kappa] (host code).
//
//   x^(-1/2) = (2/pi) int_0^inf dt / (t^2 + x)
//
// With t = sc(u | k), k'^2 = 1/kappa (Jacobi elliptic functions, modulus
// k = sqrt(1 - 1/kappa)), the poles t^2 = -x, x in [1, kappa], move to the
// line Im u = K' and the integrand becomes even about u = 0 and u = K, so
// the midpoint rule on [0, K] converges like exp(-2 pi K' N / K)
// (about exp(-2 pi^2 N / log(16 kappa))):
//
//   x^(-1/2) ~= sum_j w_j / (x + s_j),   u_j = (j - 1/2) K / N,
//   s_j = sc^2(u_j),   w_j = (2 K / (pi N)) dn(u_j) nc^2(u_j).
//
// For a symmetric positive definite A with spectrum in [m, m kappa],
// A^(-1/2) = m^(-1/2) (A/m)^(-1/2) ~= sum_j sqrt(m) w_j (A + m s_j I)^-1.
//
// sn, cn, dn come from the arithmetic-geometric mean (descending Landen)
// method; for u > K/2 the quarter-period reflection u -> K - u keeps cn and
// the square of sc accurate when k is close to 1.
#pragma once

#include <cmath>

namespace lk {
namespace rat {

// complete elliptic integral K(k) = pi / (2 AGM(1, k'))
inline double ellipk_c(double kc) {
  double a = 1.0, b = kc;
  for (int i = 0; i < 64 && std::fabs(a - b) > 1e-17 * a; ++i) {
    const double an = 0.5 * (a + b);
    b = std::sqrt(a * b);
    a = an;
  }
  return M_PI / (2.0 * a);
}

// sn, cn, dn of u for modulus k (k' = kc), AGM method
inline void ellipj(double u, double k, double kc, double* sn, double* cn, double* dn) {
  double a[40], c[40];
  a[0] = 1.0;
  double b = kc;
  c[0] = k;
  int n = 0;
  while (n < 38 && std::fabs(c[n]) > 1e-17) {
    a[n + 1] = 0.5 * (a[n] + b);
    c[n + 1] = 0.5 * (a[n] - b);
    b = std::sqrt(a[n] * b);
    ++n;
  }
  double phi = std::ldexp(a[n] * u, n);
  double phi_prev = phi;
  for (int i = n; i >= 1; --i) {
    phi_prev = phi;
    phi = 0.5 * (phi + std::asin(c[i] * std::sin(phi) / a[i]));
  }
  *sn = std::sin(phi);
  *cn = std::cos(phi);
  *dn = n >= 1 ? std::cos(phi) / std::cos(phi_prev - phi) : 1.0;
}

// nodes for [1, kappa] with n points: s[j] = sc^2(u_j), w[j]
inline void nodes(double kappa, int n, double* s, double* w) {
  const double kc = 1.0 / std::sqrt(kappa);
  const double k = std::sqrt((1.0 - kc) * (1.0 + kc));
  const double K = ellipk_c(kc);
  for (int j = 0; j < n; ++j) {
    const double u = (j + 0.5) * K / n;
    double sc2, dnc2;
    if (u <= 0.5 * K) {
      double sn, cn, dn;
      ellipj(u, k, kc, &sn, &cn, &dn);
      sc2 = (sn / cn) * (sn / cn);
      dnc2 = dn / (cn * cn);
    } else {
      // u = K - v: sn = cn(v)/dn(v), cn = k' sn(v)/dn(v), dn = k'/dn(v)
      double sn, cn, dn;
      ellipj(K - u, k, kc, &sn, &cn, &dn);
      const double r = cn / (kc * sn);          // sc(u)
      sc2 = r * r;
      dnc2 = (kc / dn) * (dn / (kc * sn)) * (dn / (kc * sn));   // dn(u) nc(u)^2
    }
    s[j] = sc2;
    w[j] = 2.0 * K / (M_PI * n) * dnc2;
  }
}

// largest relative error of the n-point approximation over [1, kappa]
// (log-spaced samples; the error is a smooth function of log x)
inline double max_rel_err(double kappa, int n, const double* s, const double* w,
                          int samples = 1500) {
  double e = 0.0;
  for (int i = 0; i <= samples; ++i) {
    const double x = std::exp(std::log(kappa) * i / samples);
    double acc = 0.0;
    for (int j = 0; j < n; ++j) acc += w[j] / (x + s[j]);
    e = std::fmax(e, std::fabs(acc * std::sqrt(x) - 1.0));
  }
  return e;
}

// smallest n <= nmax meeting tol on [1, kappa]; fills s, w; returns n or -1
inline int choose(double kappa, double tol, int nmax, double* s, double* w) {
  for (int n = 2; n <= nmax; ++n) {
    nodes(kappa, n, s, w);
    if (max_rel_err(kappa, n, s, w) <= tol) return n;
  }
  return -1;
}

}  // namespace rat
}  // namespace lk
