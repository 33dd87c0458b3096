// Host build of umd-letkf_amd/csrc/ratnodes.h for the CPU tests.
#include "../../umd-letkf_amd/csrc/ratnodes.h"

extern "C" {
double rn_ellipk(double kc) { return lk::rat::ellipk_c(kc); }
void rn_ellipj(double u, double k, double kc, double* sn, double* cn, double* dn) {
  lk::rat::ellipj(u, k, kc, sn, cn, dn);
}
void rn_nodes(double kappa, int n, double* s, double* w) { lk::rat::nodes(kappa, n, s, w); }
int rn_choose(double kappa, double tol, int nmax, double* s, double* w) {
  return lk::rat::choose(kappa, tol, nmax, s, w);
}
}
