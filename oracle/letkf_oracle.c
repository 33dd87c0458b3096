/*
 * letkf_oracle.c -- FP64 CPU restatement of UMD-LETKF's local analysis.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU library.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this file's
 * library; nothing in the product links or calls it.
 *
 * What it restates (file:line in the reference tree):
 *   solver gridpoint loop ........ src/letkf_solver.F90:150-389
 *   transform core ............... src/letkf_core.F90:69-146
 *   Gaspari-Cohn (truncated) ..... src/letkf_loc.F90:348-378
 *   linearinterp_lat ............. src/letkf_loc.F90:251-320
 *   LOC_NOVRT .................... subclass.letkf_localizer/letkf_loc_novrt.F90:133-183
 *   LOC_OCEAN .................... subclass.letkf_localizer/letkf_loc_ocean.F90:331-506
 *   search radius L*sqrt(40/3) ... src/letkf_solver.F90:47,210-211
 *   radius search ................ src/letkf_obs.F90:560-573 (geoKdTree absent:
 *                                  distance definition below is this project's,
 *                                  parity unpinned against the reference)
 *
 * Arithmetic: FP64 throughout, from the reference's FP32 inputs.  Values the
 * reference keeps in REAL (the localization weight, the distance handed to
 * the localizer, the linear interpolation of the radius tables, the
 * departure val - hx_mean) are rounded to FP32 at the same points.
 * Build with -ffp-contract=off: the elementary functions below must give the
 * same bits as the GPU library's copy of them.
 */
#include "letkf_oracle.h"

#include <math.h>
#include <pthread.h>
#include <time.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* constants                                                                */
/* ------------------------------------------------------------------------ */
#define ORC_RE       6371000.0               /* Earth radius [m] (own choice) */
#define ORC_2RE      12742000.0
#define ORC_PIO2     0x1.921fb54442d18p+0
#define ORC_DEG2RAD  0x1.1df46a2529d39p-6
#define ORC_SQRT03   0x1.186f174f88472p-1    /* SQRT(0.3d0), letkf_loc.F90:355 */
#define ORC_INVLN2   0x1.71547652b82fep+0
#define ORC_LN2HI    0x1.62e4200000000p-1
#define ORC_LN2LO    0x1.fdf473de6af28p-22
/* stdev2max = SQRT(40.0/3.0) in REAL, src/letkf_solver.F90:47 */
#define ORC_STDEV2MAX 3.6514837741851807f

static const double SIN_C[10] = {
  -0x1.5555555555555p-3, 0x1.1111111111111p-7, -0x1.a01a01a01a01ap-13,
  0x1.71de3a556c734p-19, -0x1.ae64567f544e4p-26, 0x1.6124613a86d09p-33,
  -0x1.ae7f3e733b81fp-41, 0x1.952c77030ad4ap-49, -0x1.2f49b46814157p-57,
  0x1.71b8ef6dcf572p-66};
static const double COS_C[11] = {
  -0x1.0000000000000p-1, 0x1.5555555555555p-5, -0x1.6c16c16c16c17p-10,
  0x1.a01a01a01a01ap-16, -0x1.27e4fb7789f5cp-22, 0x1.1eed8eff8d898p-29,
  -0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-45, -0x1.6827863b97d97p-53,
  0x1.e542ba4020225p-62, -0x1.0ce396db7f853p-70};
static const double ASIN_C[28] = {
  0x1.5555555555555p-3, 0x1.3333333333333p-4, 0x1.6db6db6db6db7p-5,
  0x1.f1c71c71c71c7p-6, 0x1.6e8ba2e8ba2e9p-6, 0x1.1c4ec4ec4ec4fp-6,
  0x1.c99999999999ap-7, 0x1.7a87878787878p-7, 0x1.3fde50d79435ep-7,
  0x1.12ef3cf3cf3cfp-7, 0x1.df3bd37a6f4dfp-8, 0x1.a6863d70a3d71p-8,
  0x1.782dda12f684cp-8, 0x1.51ba308d3dcb1p-8, 0x1.31683bdef7bdfp-8,
  0x1.15ee9d45d1746p-8, 0x1.fcaf8fb6db6dbp-9, 0x1.d3d2a8e0dd67dp-9,
  0x1.b026f57b13b14p-9, 0x1.90cb77f60c7cep-9, 0x1.750de64d7d05fp-9,
  0x1.5c5f56efaaaabp-9, 0x1.464c0950f7d47p-9, 0x1.3275586c5f2f0p-9,
  0x1.208d3570ae5a6p-9, 0x1.1052bc5fa960ap-9, 0x1.018f963c229bfp-9,
  0x1.e82be60d9127ep-10};
static const double EXP_C[15] = {
  0x1.0000000000000p+0, 0x1.0000000000000p+0, 0x1.0000000000000p-1,
  0x1.5555555555555p-3, 0x1.5555555555555p-5, 0x1.1111111111111p-7,
  0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-16,
  0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26,
  0x1.1eed8eff8d898p-29, 0x1.6124613a86d09p-33, 0x1.93974a8c07c9dp-37};

/* ------------------------------------------------------------------------ */
/* deterministic elementary functions: only +,-,*,/ and sqrt, fixed order   */
/* ------------------------------------------------------------------------ */
static double sin_core(double x) {            /* |x| <= pi/4 (+ small) */
  double x2 = x * x;
  double p = SIN_C[9];
  for (int i = 8; i >= 0; --i) p = p * x2 + SIN_C[i];
  return x + (x * x2) * p;
}
static double cos_core(double x) {
  double x2 = x * x;
  double p = COS_C[10];
  for (int i = 9; i >= 0; --i) p = p * x2 + COS_C[i];
  return 1.0 + x2 * p;
}

void orc_sincosd(double deg, double* s, double* c) {
  double q = floor(deg / 90.0 + 0.5);
  double r = deg - q * 90.0;                  /* exact for FP32 degrees */
  double x = r * ORC_DEG2RAD;
  double sn = sin_core(x), cs = cos_core(x);
  long iq = ((long)q % 4 + 4) % 4;
  switch (iq) {
    case 0: *s = sn; *c = cs; break;
    case 1: *s = cs; *c = -sn; break;
    case 2: *s = -sn; *c = -cs; break;
    default: *s = -cs; *c = sn; break;
  }
}

double orc_sin_rad(double x) {                /* 0 <= x < pi/2 */
  if (x <= 0.78539816339744828) return sin_core(x);
  return cos_core(ORC_PIO2 - x);
}

static double asin_core(double x) {           /* 0 <= x <= 0.5 */
  double x2 = x * x;
  double p;
  if (x <= 0.2) {
    p = ASIN_C[13];
    for (int i = 12; i >= 0; --i) p = p * x2 + ASIN_C[i];
  } else {
    p = ASIN_C[27];
    for (int i = 26; i >= 0; --i) p = p * x2 + ASIN_C[i];
  }
  return x + (x * x2) * p;
}

double orc_asin(double x) {                   /* 0 <= x <= 1 */
  if (x <= 0.5) return asin_core(x);
  double t = sqrt((1.0 - x) * 0.5);
  return ORC_PIO2 - 2.0 * asin_core(t);
}

double orc_exp_neg(double x) {                /* x <= 0 */
  if (x < -745.0) return 0.0;
  double n = floor(x * ORC_INVLN2 + 0.5);
  double r = (x - n * ORC_LN2HI) - n * ORC_LN2LO;
  double p = EXP_C[14];
  for (int i = 13; i >= 0; --i) p = p * r + EXP_C[i];
  return ldexp(p, (int)n);
}

void orc_xyz(float lat, float lon, double* xyz) {
  double sl, cl, so, co;
  orc_sincosd((double)lat, &sl, &cl);
  orc_sincosd((double)lon, &so, &co);
  xyz[0] = cl * co;
  xyz[1] = cl * so;
  xyz[2] = sl;
}

double orc_chord2(const double* a, const double* b) {
  double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return (dx * dx + dy * dy) + dz * dz;
}

/* squared chord of a search radius in metres; 5 (> any chord^2) when the
 * cap covers the sphere */
double orc_thr2(float radius_m) {
  double half = (double)radius_m / ORC_2RE;
  if (!(half < ORC_PIO2)) return 5.0;
  double t = 2.0 * orc_sin_rad(half);
  return t * t;
}

/* great-circle distance of a squared chord, rounded to REAL like the
 * kd-tree's rdist (src/letkf_obs.F90:565) */
float orc_dist(double chord2) {
  double h = sqrt(chord2) * 0.5;
  if (h > 1.0) h = 1.0;
  return (float)(ORC_2RE * orc_asin(h));
}

uint64_t orc_mix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* ------------------------------------------------------------------------ */
/* Gaspari-Cohn with the reference's truncated coefficients                  */
/* src/letkf_loc.F90:348-378                                                 */
/* ------------------------------------------------------------------------ */
float orc_gc(float z, float L) {
  double c = (double)L / ORC_SQRT03;
  double abs_z = fabs((double)z);
  double z_c = abs_z / c;
  double res;
  if (abs_z >= 2.0 * c) {
    res = 0.0;
  } else {
    double z2 = z_c * z_c, z3 = z2 * z_c, z4 = z2 * z2, z5 = z4 * z_c;
    if (abs_z > c) {
      res = 0.08333 * z5 - 0.50000 * z4 + 0.62500 * z3 + 1.66667 * z2
            - 5.00000 * z_c + 4.0 - 0.66667 * c / abs_z;
    } else {
      res = -0.25000 * z5 + 0.50000 * z4 + 0.62500 * z3 - 1.66667 * z2 + 1.0;
    }
  }
  return (float)res;
}

/* ------------------------------------------------------------------------ */
/* linearinterp_lat  (src/letkf_loc.F90:251-320)                            */
/* ------------------------------------------------------------------------ */
int orc_interp_build(const letkf_linearinterp_lat* cfg, float* lat, float* dist) {
  int n = 0;
  float l = 0.0f, r = 0.0f;
  for (int i = 0; i < cfg->npoints; ++i) {
    l = cfg->lat[i];
    r = cfg->radius[i];
    if (l > 0.0f && n == 0) { lat[n] = 0.0f; dist[n] = r; ++n; }   /* :277-281 */
    lat[n] = l; dist[n] = r; ++n;
  }
  if (l < 90.0f) { lat[n] = 90.0f; dist[n] = r; ++n; }              /* :289-293 */
  return n;
}

float orc_interp_get(int n, const float* lat, const float* dist, float x) {
  float a = fabsf(x);
  int i = 0;
  while (i < n - 1 && a > lat[i]) ++i;        /* :309-311 (bounded) */
  if (i == 0) return dist[0];
  float d = (a - lat[i - 1]) / (lat[i] - lat[i - 1]);
  return d * dist[i] + (1.0f - d) * dist[i - 1];
}

/* ------------------------------------------------------------------------ */
/* localizers                                                                */
/* ------------------------------------------------------------------------ */
typedef struct {
  int n; float lat[LETKF_MAX_INTERP + 2]; float dist[LETKF_MAX_INTERP + 2];
} interp_t;

typedef struct {
  const letkf_loc_desc* d;
  interp_t hz, prof, sat;
  int maxgroups;
} loc_t;

static void interp_init(interp_t* t, const letkf_linearinterp_lat* cfg) {
  t->n = cfg->npoints > 0 ? orc_interp_build(cfg, t->lat, t->dist) : 0;
}
static float interp_get(const interp_t* t, float x) {
  return orc_interp_get(t->n, t->lat, t->dist, x);
}

int orc_maxgroups(const letkf_loc_desc* loc) {
  switch (loc->cls) {
    case LETKF_LOC_NOVRT: return 1;
    case LETKF_LOC_OCEAN: return loc->vtloc_surf == 1 ? 2 : 1;
    case LETKF_LOC_VGAUSS: return loc->nlev;
    default: return -1;
  }
}

static void loc_init(loc_t* L, const letkf_loc_desc* d) {
  memset(L, 0, sizeof(*L));
  L->d = d;
  interp_init(&L->hz, &d->hzloc);
  interp_init(&L->prof, &d->hzloc_prof);
  interp_init(&L->sat, &d->hzloc_sat);
  L->maxgroups = orc_maxgroups(d);
}

/* maxhz: novrt :155-161, ocean :435-444 */
static float loc_maxhz(const loc_t* L, float lat) {
  if (L->d->cls == LETKF_LOC_OCEAN) {
    float a = interp_get(&L->prof, lat), b = interp_get(&L->sat, lat);
    return a > b ? a : b;                                /* MAX(prof, sat) */
  }
  return interp_get(&L->hz, lat);
}

/* groups(): returns the number of groups, writes group id and the group of
 * every slab (-1 = in no group).  novrt :133-147, ocean :331-427 */
static int loc_groups(const loc_t* L, const orc_problem* p, int64_t ij,
                      int* gid, int* slab_group) {
  const letkf_loc_desc* d = L->d;
  int64_t ns = p->ns;
  if (d->cls == LETKF_LOC_OCEAN && d->vtloc_surf == 1) {
    const float* mean = p->mean + ns * ij;
    float t0 = mean[d->bkg_t_var_idx];
    int mld = d->bkg_t_var_len;
    for (int i = 1; i <= d->bkg_t_var_len - 1; ++i) {   /* calc_mldlvl_bkgt */
      float t1 = mean[d->bkg_t_var_idx + i];
      if (fabsf(t0 - t1) > d->bkg_t_delta) { mld = i; break; }
    }
    for (int64_t s = 0; s < ns; ++s) slab_group[s] = -1;
    for (int v = 0; v < p->nvars; ++v)
      for (int j = 1; j <= p->var_lev[v]; ++j)
        slab_group[p->var_s[v] + j - 1] = (j <= mld) ? 0 : 1;
    gid[0] = 0;  /* LOC_GROUP_SURF  */
    gid[1] = 1;  /* LOC_GROUP_SUBML */
    return 2;
  }
  if (d->cls == LETKF_LOC_VGAUSS) {
    for (int64_t s = 0; s < ns; ++s) slab_group[s] = -1;
    for (int v = 0; v < p->nvars; ++v)
      for (int j = 0; j < p->var_lev[v]; ++j) slab_group[p->var_s[v] + j] = j;
    for (int l = 0; l < d->nlev; ++l) gid[l] = l;
    return d->nlev;
  }
  /* one group with every slab; the ocean class leaves its id unset
   * (:392-396), taken here as LOC_GROUP_SURF = 0 */
  for (int64_t s = 0; s < ns; ++s) slab_group[s] = 0;
  gid[0] = 0;
  return 1;
}

static int in_list(int32_t v, const int32_t* list, int n) {
  for (int i = 0; i < n; ++i) if (list[i] == v) return 1;
  return 0;
}

/* localize(): novrt :170-183, ocean :451-506, vgauss (new class) */
static float loc_localize(const loc_t* L, const orc_problem* p, float lat_ij,
                          int group_id, int64_t o, float dist) {
  const letkf_loc_desc* d = L->d;
  if (d->cls == LETKF_LOC_OCEAN) {
    int surf = 1;
    if (d->n_sat_obs > 0 && !in_list(p->obsid[o], d->sat_obs, d->n_sat_obs)) surf = 0;
    if (d->n_sat_plats > 0 && !in_list(p->platid[o], d->sat_plats, d->n_sat_plats)) surf = 0;
    float loc;
    if (surf) {
      if (group_id == 1) return 0.0f;
      loc = orc_gc(dist, interp_get(&L->sat, lat_ij));
      if (d->tloc_sat > 0.0f) loc = loc * orc_gc(p->otime[o], d->tloc_sat);
    } else {
      loc = orc_gc(dist, interp_get(&L->prof, lat_ij));
      if (d->tloc_prof > 0.0f) loc = loc * orc_gc(p->otime[o], d->tloc_prof);
    }
    return loc;
  }
  float loc = orc_gc(dist, interp_get(&L->hz, lat_ij));
  if (d->cls == LETKF_LOC_VGAUSS) {
    double dz = (double)p->ozdim[o] - (double)d->lev_depth[group_id];
    double sc = (double)d->vt_scale;
    double v = orc_exp_neg(-(dz * dz) / (2.0 * sc * sc));
    loc = loc * (float)v;
  }
  return loc;
}

/* ------------------------------------------------------------------------ */
/* symmetric eigensolver: cyclic Jacobi (Golub & Van Loan 4th ed. 8.5.2)    */
/* a: n x n column-major, destroyed; w ascending; v column eigenvectors     */
/* ------------------------------------------------------------------------ */
int orc_eigh(int n, double* a, double* w, double* v) {
#define A(i, j) a[(size_t)(j) * n + (i)]
#define V(i, j) v[(size_t)(j) * n + (i)]
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) V(i, j) = (i == j) ? 1.0 : 0.0;
  int sweep;
  for (sweep = 0; sweep < 60; ++sweep) {
    int rotated = 0;
    for (int p = 0; p < n - 1; ++p) {
      for (int q = p + 1; q < n; ++q) {
        double apq = A(p, q);
        if (apq == 0.0) continue;
        double app = A(p, p), aqq = A(q, q);
        if (fabs(apq) <= 1e-18 * sqrt(fabs(app) * fabs(aqq))) { A(p, q) = A(q, p) = 0.0; continue; }
        rotated = 1;
        double tau = (aqq - app) / (2.0 * apq);
        double t;
        if (fabs(tau) > 1e150) t = 0.5 / tau;
        else t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        for (int k = 0; k < n; ++k) {          /* A <- A J (columns p, q) */
          double akp = A(k, p), akq = A(k, q);
          A(k, p) = c * akp - s * akq;
          A(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {          /* A <- J^T A (rows p, q) */
          double apk = A(p, k), aqk = A(q, k);
          A(p, k) = c * apk - s * aqk;
          A(q, k) = s * apk + c * aqk;
        }
        A(p, q) = A(q, p) = 0.0;
        for (int k = 0; k < n; ++k) {          /* V <- V J */
          double vkp = V(k, p), vkq = V(k, q);
          V(k, p) = c * vkp - s * vkq;
          V(k, q) = s * vkp + c * vkq;
        }
      }
    }
    if (!rotated) break;
  }
  for (int i = 0; i < n; ++i) w[i] = A(i, i);
  /* ascending order, like LAPACK xSYEV */
  for (int i = 0; i < n - 1; ++i) {
    int m = i;
    for (int j = i + 1; j < n; ++j) if (w[j] < w[m]) m = j;
    if (m != i) {
      double t = w[i]; w[i] = w[m]; w[m] = t;
      for (int k = 0; k < n; ++k) { double u = V(k, i); V(k, i) = V(k, m); V(k, m) = u; }
    }
  }
  return sweep;
#undef A
#undef V
}

/* ------------------------------------------------------------------------ */
/* letkf_core_solve  (src/letkf_core.F90:69-146), FP64                      */
/* hdxb (k, m) column-major; trans (k, k) column-major                      */
/* ------------------------------------------------------------------------ */
int orc_core_solve(int k, int m, const double* hdxb, const double* rdiag,
                   const double* rloc, const double* dep, double infl,
                   double* trans, double* eival) {
  double* hr = (double*)malloc(sizeof(double) * (size_t)k * (m > 0 ? m : 1));
  double* w1 = (double*)malloc(sizeof(double) * (size_t)k * k * 3 + sizeof(double) * 3 * k);
  if (!hr || !w1) { free(hr); free(w1); return -1; }
  double* ev = w1 + (size_t)k * k;
  double* pa = ev + (size_t)k * k;
  double* lam = pa + (size_t)k * k;
  double* hd = lam + k;
  double* w3 = hd + k;
  /* hdxb_rinv(:,j) = hdxb(:,j) / rdiag(j) * rloc(j)          :92-94 */
  for (int j = 0; j < m; ++j)
    for (int a = 0; a < k; ++a)
      hr[a + (size_t)k * j] = hdxb[a + (size_t)k * j] / rdiag[j] * rloc[j];
  /* work1 = hdxb_rinv * hdxb^T                                :97-98 */
  memset(w1, 0, sizeof(double) * (size_t)k * k);
  for (int j = 0; j < m; ++j) {
    const double* hrj = hr + (size_t)k * j;
    const double* hj = hdxb + (size_t)k * j;
    for (int b = 0; b < k; ++b) {
      double hb = hj[b];
      double* col = w1 + (size_t)k * b;
      for (int a = 0; a < k; ++a) col[a] += hrj[a] * hb;
    }
  }
  /* + (k-1)/infl I                                            :101-104 */
  double r = (double)(k - 1) / infl;
  for (int i = 0; i < k; ++i) w1[i + (size_t)k * i] += r;
  /* eigen-decomposition                                       :107 */
  orc_eigh(k, w1, lam, ev);
  /* pa = eivec diag(1/eival) eivec^T                          :111-115 */
  for (int b = 0; b < k; ++b)
    for (int a = 0; a < k; ++a) {
      double s = 0.0;
      for (int i = 0; i < k; ++i) s += ev[a + (size_t)k * i] / lam[i] * ev[b + (size_t)k * i];
      pa[a + (size_t)k * b] = s;
    }
  /* work3 = pa * (hdxb_rinv * dep)                            :118-125 */
  for (int a = 0; a < k; ++a) hd[a] = 0.0;
  for (int j = 0; j < m; ++j)
    for (int a = 0; a < k; ++a) hd[a] += hr[a + (size_t)k * j] * dep[j];
  for (int a = 0; a < k; ++a) {
    double s = 0.0;
    for (int b = 0; b < k; ++b) s += pa[a + (size_t)k * b] * hd[b];
    w3[a] = s;
  }
  /* trans = eivec diag(sqrt((k-1)/eival)) eivec^T + work3 1^T  :128-141 */
  for (int b = 0; b < k; ++b)
    for (int a = 0; a < k; ++a) {
      double s = 0.0;
      for (int i = 0; i < k; ++i)
        s += ev[a + (size_t)k * i] * sqrt((double)(k - 1) / lam[i]) * ev[b + (size_t)k * i];
      trans[a + (size_t)k * b] = s + w3[a];
    }
  if (eival) for (int i = 0; i < k; ++i) eival[i] = lam[i];
  free(hr);
  free(w1);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* search structure: obs sorted by the z coordinate                         */
/* ------------------------------------------------------------------------ */
typedef struct {
  const orc_problem* p;
  loc_t loc;
  double* oxyz;      /* (3, nobs) */
  double* zs;        /* sorted z */
  int32_t* zperm;    /* sorted -> original index */
} orc_ctx;

static const double* g_sort_z;
static int cmp_z(const void* a, const void* b) {
  double za = g_sort_z[3 * (size_t)*(const int32_t*)a + 2];
  double zb = g_sort_z[3 * (size_t)*(const int32_t*)b + 2];
  if (za < zb) return -1;
  if (za > zb) return 1;
  int32_t ia = *(const int32_t*)a, ib = *(const int32_t*)b;
  return (ia > ib) - (ia < ib);
}
static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

static pthread_mutex_t g_sort_lock = PTHREAD_MUTEX_INITIALIZER;

static int ctx_init(orc_ctx* c, const orc_problem* p) {
  memset(c, 0, sizeof(*c));
  c->p = p;
  loc_init(&c->loc, p->loc);
  if (c->loc.maxgroups <= 0) return -1;
  int64_t n = p->nobs;
  c->oxyz = (double*)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
  c->zs = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  c->zperm = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  if (!c->oxyz || !c->zs || !c->zperm) return -1;
  for (int64_t i = 0; i < n; ++i) {
    orc_xyz(p->olat[i], p->olon[i], c->oxyz + 3 * i);
    c->zperm[i] = (int32_t)i;
  }
  pthread_mutex_lock(&g_sort_lock);
  g_sort_z = c->oxyz;
  qsort(c->zperm, (size_t)n, sizeof(int32_t), cmp_z);
  pthread_mutex_unlock(&g_sort_lock);
  for (int64_t i = 0; i < n; ++i) c->zs[i] = c->oxyz[3 * (size_t)c->zperm[i] + 2];
  return 0;
}
static void ctx_free(orc_ctx* c) {
  free(c->oxyz); free(c->zs); free(c->zperm);
}

static int64_t lower_bound(const double* a, int64_t n, double x) {
  int64_t lo = 0, hi = n;
  while (lo < hi) { int64_t mid = (lo + hi) / 2; if (a[mid] < x) lo = mid + 1; else hi = mid; }
  return lo;
}

/* radius search (letkf_obs_get, src/letkf_obs.F90:560-573): every obs with
 * chord^2 <= thr2, returned sorted by index with FP32 distances */
static int64_t search(const orc_ctx* c, const double* g, double thr2,
                      int32_t* idx, float* dist) {
  const orc_problem* p = c->p;
  if (p->nobs <= 0) return 0;
  double thr = sqrt(thr2);
  double lo = g[2] - thr * (1.0 + 1e-9) - 1e-12, hi = g[2] + thr * (1.0 + 1e-9) + 1e-12;
  int64_t b = lower_bound(c->zs, p->nobs, lo);
  int64_t n = 0;
  for (int64_t s = b; s < p->nobs && c->zs[s] <= hi; ++s) {
    int32_t o = c->zperm[s];
    double c2 = orc_chord2(c->oxyz + 3 * (size_t)o, g);
    if (c2 <= thr2) idx[n++] = o;
  }
  qsort(idx, (size_t)n, sizeof(int32_t), cmp_i32);
  for (int64_t i = 0; i < n; ++i)
    dist[i] = orc_dist(orc_chord2(c->oxyz + 3 * (size_t)idx[i], g));
  return n;
}

/* ------------------------------------------------------------------------ */
/* the gridpoint loop  (src/letkf_solver.F90:203-384)                       */
/* ------------------------------------------------------------------------ */
typedef struct {
  int32_t* cidx; float* cdist;
  int32_t* kidx; double* krloc;
  double* hdxb; double* rdiag; double* rloc; double* dep;
  int64_t cap_m;
  double* bkg; double* ana; double* trans;
  int* gid; int* slab_group;
  double t_search, t_core, t_trans;   /* phase thread-seconds */
} orc_scratch;

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int scratch_init(orc_scratch* s, const orc_problem* p, int maxgroups) {
  memset(s, 0, sizeof(*s));
  int64_t n = p->nobs > 0 ? p->nobs : 1;
  s->cidx = (int32_t*)malloc(sizeof(int32_t) * n);
  s->cdist = (float*)malloc(sizeof(float) * n);
  s->kidx = (int32_t*)malloc(sizeof(int32_t) * n);
  s->krloc = (double*)malloc(sizeof(double) * n);
  s->bkg = (double*)malloc(sizeof(double) * p->k * p->ns);
  s->ana = (double*)malloc(sizeof(double) * p->k * p->ns);
  s->trans = (double*)malloc(sizeof(double) * p->k * p->k);
  s->gid = (int*)malloc(sizeof(int) * (maxgroups > 0 ? maxgroups : 1));
  s->slab_group = (int*)malloc(sizeof(int) * p->ns);
  return (s->cidx && s->cdist && s->kidx && s->krloc && s->bkg && s->ana &&
          s->trans && s->gid && s->slab_group) ? 0 : -1;
}
static int scratch_reserve(orc_scratch* s, int64_t k, int64_t m) {
  if (m <= s->cap_m) return 0;
  int64_t cap = m + m / 2 + 64;
  free(s->hdxb); free(s->rdiag); free(s->rloc); free(s->dep);
  s->hdxb = (double*)malloc(sizeof(double) * k * cap);
  s->rdiag = (double*)malloc(sizeof(double) * cap);
  s->rloc = (double*)malloc(sizeof(double) * cap);
  s->dep = (double*)malloc(sizeof(double) * cap);
  s->cap_m = cap;
  return (s->hdxb && s->rdiag && s->rloc && s->dep) ? 0 : -1;
}
static void scratch_free(orc_scratch* s) {
  free(s->cidx); free(s->cdist); free(s->kidx); free(s->krloc);
  free(s->hdxb); free(s->rdiag); free(s->rloc); free(s->dep);
  free(s->bkg); free(s->ana); free(s->trans); free(s->gid); free(s->slab_group);
}

/* kept set of one group (:244-263): returns m */
static int64_t localize_group(const orc_ctx* c, float lat_ij, int gidv, int64_t n,
                              const int32_t* cidx, const float* cdist,
                              int32_t* kidx, double* krloc) {
  int64_t m = 0;
  for (int64_t i = 0; i < n; ++i) {
    float r = loc_localize(&c->loc, c->p, lat_ij, gidv, cidx[i], cdist[i]);
    if (r > 1.0e-4f) { kidx[m] = cidx[i]; krloc[m] = (double)r; ++m; }
  }
  return m;
}

static int analyse_point(const orc_ctx* c, orc_scratch* s, int64_t ij,
                         int64_t slot, const orc_output* out) {
  const orc_problem* p = c->p;
  const int64_t k = p->k, ns = p->ns;
  const int G = c->loc.maxgroups;
  const float* st = p->state + k * ns * ij;
  double* ost = out->state + k * ns * slot;
  double* omean = out->mean + ns * slot;
  double* osprd = out->sprd + ns * slot;
  if (out->maxhz) out->maxhz[slot] = 0.0f;
  if (out->obscount) out->obscount[slot] = 0;
  for (int g = 0; g < G; ++g) {
    if (out->lg_m) out->lg_m[g + (int64_t)G * slot] = 0;
    if (out->lg_obsloc) out->lg_obsloc[g + (int64_t)G * slot] = 0.0;
    if (out->lg_hash) out->lg_hash[g + (int64_t)G * slot] = 0;
  }
  if (p->mask[ij]) {                                          /* :207 */
    for (int64_t i = 0; i < k * ns; ++i) ost[i] = (double)st[i];
    for (int64_t i = 0; i < ns; ++i) {
      omean[i] = (double)p->mean[ns * ij + i];
      osprd[i] = (double)p->sprd[ns * ij + i];
    }
    return 0;
  }
  double tp = now_s();
  float lat_ij = p->lat[ij];
  float maxhz = loc_maxhz(&c->loc, lat_ij);                  /* :210-212 */
  float rs = maxhz * ORC_STDEV2MAX;
  if (out->maxhz) out->maxhz[slot] = maxhz;
  double g[3];
  orc_xyz(p->lat[ij], p->lon[ij], g);
  int64_t n = search(c, g, orc_thr2(rs), s->cidx, s->cdist);  /* :217-220 */
  if (out->obscount) out->obscount[slot] = n;
  {
    const double t = now_s();
    s->t_search += t - tp;
    tp = t;
  }

  int ngr = loc_groups(&c->loc, p, ij, s->gid, s->slab_group);  /* :232 */
  for (int64_t i = 0; i < k * ns; ++i) { s->bkg[i] = (double)st[i]; s->ana[i] = s->bkg[i]; }
  const float* bmean = p->mean + ns * ij;
  const float* bsprd = p->sprd + ns * ij;

  for (int lg = 0; lg < ngr; ++lg) {                          /* :241-292 */
    double tq = now_s();
    s->t_trans += tq - tp;
    int64_t m = localize_group(c, lat_ij, s->gid[lg], n, s->cidx, s->cdist,
                               s->kidx, s->krloc);
    tp = now_s();
    s->t_search += tp - tq;
    double sumr = 0.0;
    uint64_t h = 0;
    for (int64_t i = 0; i < m; ++i) { sumr += s->krloc[i]; h += orc_mix64((uint64_t)s->kidx[i]); }
    if (out->lg_m) out->lg_m[lg + (int64_t)G * slot] = m;
    if (out->lg_obsloc) out->lg_obsloc[lg + (int64_t)G * slot] = sumr;
    if (out->lg_hash) out->lg_hash[lg + (int64_t)G * slot] = h;
    if (m == 0) continue;
    int has_slab = 0;
    for (int64_t sl = 0; sl < ns; ++sl) if (s->slab_group[sl] == lg) { has_slab = 1; break; }
    if (!has_slab) continue;
    if (scratch_reserve(s, k, m)) return -1;
    for (int64_t i = 0; i < m; ++i) {                          /* :224-229 */
      int64_t o = s->kidx[i];
      for (int64_t a = 0; a < k; ++a) s->hdxb[a + k * i] = (double)p->hx[a + k * o];
      s->rdiag[i] = (double)p->oerr[o];
      s->dep[i] = (double)(p->oval[o] - p->hxmean[o]);       /* REAL subtraction */
      s->rloc[i] = s->krloc[i];
    }
    if (orc_core_solve((int)k, (int)m, s->hdxb, s->rdiag, s->rloc, s->dep, 1.0,
                       s->trans, NULL)) return -1;
    tq = now_s();
    s->t_core += tq - tp;
    tp = tq;
    for (int64_t sl = 0; sl < ns; ++sl) {                      /* :285-289 */
      if (s->slab_group[sl] != lg) continue;
      const double* xb = s->bkg + k * sl;
      double* xa = s->ana + k * sl;
      for (int64_t j = 0; j < k; ++j) {
        double acc = 0.0;
        for (int64_t i = 0; i < k; ++i) acc += xb[i] * s->trans[i + k * j];
        xa[j] = acc;
      }
    }
  }

  /* recenter (:299-303) */
  for (int64_t sl = 0; sl < ns; ++sl) {
    double* x = s->ana + k * sl;
    double sum = 0.0;
    for (int64_t j = 0; j < k; ++j) { x[j] += (double)bmean[sl]; sum += x[j]; }
    omean[sl] = sum / (double)k;
    for (int64_t j = 0; j < k; ++j) x[j] -= omean[sl];
  }
  /* increment clipping (:308-318) */
  for (int v = 0; v < p->nvars; ++v) {
    double inc = (double)p->inc_max[v];
    double blo = (double)p->bounds[2 * v], bhi = (double)p->bounds[2 * v + 1];
    for (int j = p->var_s[v]; j < p->var_s[v] + p->var_lev[v]; ++j) {
      double bm = (double)bmean[j];
      double r = omean[j] - bm;
      r = fmax(r, -inc);
      r = fmin(r, inc);
      r = fmax(r, blo - bm);
      r = fmin(r, bhi - bm);
      omean[j] = r + bm;
    }
  }
  /* RTPS (:322-335) */
  if (p->rtps > 0.0f) {
    double a = (double)p->rtps;
    for (int64_t sl = 0; sl < ns; ++sl) {
      double* x = s->ana + k * sl;
      double ss = 0.0;
      for (int64_t j = 0; j < k; ++j) ss += x[j] * x[j];
      double sa = sqrt(ss / (double)(k - 1));
      if (sa <= 0.0) continue;
      double f = a * ((double)bsprd[sl] - sa) / sa + 1.0;
      for (int64_t j = 0; j < k; ++j) x[j] *= f;
    }
  }
  /* RTPP (:339-343) */
  if (p->rtpp > 0.0f) {
    double a = (double)p->rtpp;
    for (int64_t i = 0; i < k * ns; ++i) s->ana[i] = a * s->bkg[i] + (1.0 - a) * s->ana[i];
  }
  /* multiplicative (:347-349) */
  if (p->mul != 1.0f) {
    double f = (double)p->mul;
    for (int64_t i = 0; i < k * ns; ++i) s->ana[i] *= f;
  }
  /* member bounds with re-mean (:357-371) */
  for (int v = 0; v < p->nvars; ++v) {
    double blo = (double)p->bounds[2 * v], bhi = (double)p->bounds[2 * v + 1];
    for (int j = p->var_s[v]; j < p->var_s[v] + p->var_lev[v]; ++j) {
      double* x = s->ana + k * j;
      double sum = 0.0;
      for (int64_t mbr = 0; mbr < k; ++mbr) {
        double r = x[mbr];
        r = fmax(r, blo - omean[j]);
        r = fmin(r, bhi - omean[j]);
        x[mbr] = r;
      }
      for (int64_t mbr = 0; mbr < k; ++mbr) sum += x[mbr];
      double r = sum / (double)k;
      for (int64_t mbr = 0; mbr < k; ++mbr) x[mbr] -= r;
      omean[j] += r;
    }
  }
  /* final spread and full members (:375-379) */
  for (int64_t sl = 0; sl < ns; ++sl) {
    double* x = s->ana + k * sl;
    double ss = 0.0;
    for (int64_t j = 0; j < k; ++j) ss += x[j] * x[j];
    osprd[sl] = sqrt(ss / (double)(k - 1));
    for (int64_t j = 0; j < k; ++j) ost[j + k * sl] = x[j] + omean[sl];
  }
  s->t_trans += now_s() - tp;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* The reference's loop in REAL(4), for the CPU baseline: the same search,  */
/* groups and localization as above (the reference computes those in REAL  */
/* with FP64 GC internals), then letkf_core_solve and the apply and        */
/* post-processing of src/letkf_solver.F90:224-379 in FP32.                 */
/* ------------------------------------------------------------------------ */
typedef struct {
  float *hdxb, *rdiag, *rloc, *dep, *work, *trans, *bkg, *ana, *mean;
  int64_t cap_m;
} r32_scratch;

static int r32_reserve(r32_scratch* s, int64_t k, int64_t m) {
  if (m <= s->cap_m) return 0;
  const int64_t cap = m + m / 2 + 64;
  free(s->hdxb); free(s->rdiag); free(s->rloc); free(s->dep); free(s->work);
  s->hdxb = (float*)malloc(sizeof(float) * k * cap);
  s->rdiag = (float*)malloc(sizeof(float) * cap);
  s->rloc = (float*)malloc(sizeof(float) * cap);
  s->dep = (float*)malloc(sizeof(float) * cap);
  s->work = (float*)malloc(sizeof(float) * r32_core_work((int)k, (int)cap));
  s->cap_m = cap;
  return (s->hdxb && s->rdiag && s->rloc && s->dep && s->work) ? 0 : -1;
}

static int analyse_point_f32(const orc_ctx* c, orc_scratch* s, r32_scratch* r, int64_t ij,
                             int64_t slot, const orc_output* out) {
  const orc_problem* p = c->p;
  const int64_t k = p->k, ns = p->ns;
  const int G = c->loc.maxgroups;
  const float* st = p->state + k * ns * ij;
  double* ost = out->state + k * ns * slot;
  double* omean = out->mean + ns * slot;
  double* osprd = out->sprd + ns * slot;
  for (int g = 0; g < G; ++g)
    if (out->lg_m) out->lg_m[g + (int64_t)G * slot] = 0;
  if (p->mask[ij]) {                                          /* :207 */
    for (int64_t i = 0; i < k * ns; ++i) ost[i] = (double)st[i];
    for (int64_t i = 0; i < ns; ++i) {
      omean[i] = (double)p->mean[ns * ij + i];
      osprd[i] = (double)p->sprd[ns * ij + i];
    }
    return 0;
  }
  double tp = now_s();
  const float lat_ij = p->lat[ij];
  const float rs = loc_maxhz(&c->loc, lat_ij) * ORC_STDEV2MAX;
  double g3[3];
  orc_xyz(p->lat[ij], p->lon[ij], g3);
  const int64_t n = search(c, g3, orc_thr2(rs), s->cidx, s->cdist);
  { const double t = now_s(); s->t_search += t - tp; tp = t; }
  const int ngr = loc_groups(&c->loc, p, ij, s->gid, s->slab_group);
  float* bkg = r->bkg;
  float* x = r->ana;
  float* mean = r->mean;
  for (int64_t i = 0; i < k * ns; ++i) { bkg[i] = st[i]; x[i] = st[i]; }
  const float* bmean = p->mean + ns * ij;
  const float* bsprd = p->sprd + ns * ij;
  for (int lg = 0; lg < ngr; ++lg) {                          /* :241-292 */
    double tq = now_s();
    s->t_trans += tq - tp;
    const int64_t m = localize_group(c, lat_ij, s->gid[lg], n, s->cidx, s->cdist, s->kidx,
                                     s->krloc);
    tp = now_s();
    s->t_search += tp - tq;
    if (out->lg_m) out->lg_m[lg + (int64_t)G * slot] = m;
    if (m == 0) continue;
    if (r32_reserve(r, k, m)) return -1;
    for (int64_t i = 0; i < m; ++i) {                          /* :224-229, 251-262 */
      const int64_t o = s->kidx[i];
      memcpy(r->hdxb + k * i, p->hx + k * o, sizeof(float) * k);
      r->rdiag[i] = p->oerr[o];
      r->dep[i] = p->oval[o] - p->hxmean[o];
      r->rloc[i] = (float)s->krloc[i];
    }
    if (r32_core_solve((int)k, (int)m, r->hdxb, r->rdiag, r->rloc, r->dep, 1.0f, r->trans,
                       r->work)) return -1;
    tq = now_s();
    s->t_core += tq - tp;
    tp = tq;
    for (int64_t sl = 0; sl < ns; ++sl) {                      /* :285-289, sgemv */
      if (s->slab_group[sl] != lg) continue;
      const float* xb = bkg + k * sl;
      float* xa = x + k * sl;
      for (int64_t j = 0; j < k; ++j) {
        float acc = 0.0f;
        for (int64_t i = 0; i < k; ++i) acc += xb[i] * r->trans[i + k * j];
        xa[j] = acc;
      }
    }
  }
  const float fk = (float)k;
  for (int64_t sl = 0; sl < ns; ++sl) {                        /* :299-303 */
    float* xs = x + k * sl;
    float sum = 0.0f;
    for (int64_t j = 0; j < k; ++j) xs[j] = xs[j] + bmean[sl];
    for (int64_t j = 0; j < k; ++j) sum += xs[j];
    mean[sl] = sum / fk;
    for (int64_t j = 0; j < k; ++j) xs[j] = xs[j] - mean[sl];
  }
  for (int v = 0; v < p->nvars; ++v) {                         /* :308-318 */
    const float inc = p->inc_max[v], blo = p->bounds[2 * v], bhi = p->bounds[2 * v + 1];
    for (int j = p->var_s[v]; j < p->var_s[v] + p->var_lev[v]; ++j) {
      float rr = mean[j] - bmean[j];
      rr = fmaxf(rr, -inc);
      rr = fminf(rr, inc);
      rr = fmaxf(rr, blo - bmean[j]);
      rr = fminf(rr, bhi - bmean[j]);
      mean[j] = rr + bmean[j];
    }
  }
  if (p->rtps > 0.0f) {                                        /* :322-335 */
    for (int64_t sl = 0; sl < ns; ++sl) {
      float* xs = x + k * sl;
      float ss = 0.0f;
      for (int64_t j = 0; j < k; ++j) ss += xs[j] * xs[j];
      const float sa = sqrtf(ss / (float)(k - 1));
      if (sa <= 0.0f) continue;
      const float f = p->rtps * (bsprd[sl] - sa) / sa + 1.0f;
      for (int64_t j = 0; j < k; ++j) xs[j] *= f;
    }
  }
  if (p->rtpp > 0.0f)                                          /* :339-343 */
    for (int64_t i = 0; i < k * ns; ++i) x[i] = p->rtpp * bkg[i] + (1.0f - p->rtpp) * x[i];
  if (p->mul != 1.0f)                                          /* :347-349 */
    for (int64_t i = 0; i < k * ns; ++i) x[i] *= p->mul;
  for (int v = 0; v < p->nvars; ++v) {                         /* :357-371 */
    const float blo = p->bounds[2 * v], bhi = p->bounds[2 * v + 1];
    for (int j = p->var_s[v]; j < p->var_s[v] + p->var_lev[v]; ++j) {
      float* xs = x + k * j;
      float sum = 0.0f;
      for (int64_t mb = 0; mb < k; ++mb) {
        float rr = xs[mb];
        rr = fmaxf(rr, blo - mean[j]);
        rr = fminf(rr, bhi - mean[j]);
        xs[mb] = rr;
      }
      for (int64_t mb = 0; mb < k; ++mb) sum += xs[mb];
      const float rr = sum / fk;
      for (int64_t mb = 0; mb < k; ++mb) xs[mb] -= rr;
      mean[j] += rr;
    }
  }
  for (int64_t sl = 0; sl < ns; ++sl) {                        /* :375-379 */
    const float* xs = x + k * sl;
    float ss = 0.0f;
    for (int64_t j = 0; j < k; ++j) ss += xs[j] * xs[j];
    osprd[sl] = (double)sqrtf(ss / (float)(k - 1));
    omean[sl] = (double)mean[sl];
    for (int64_t j = 0; j < k; ++j) ost[j + k * sl] = (double)(xs[j] + mean[sl]);
  }
  s->t_trans += now_s() - tp;
  return 0;
}

typedef struct {
  const orc_ctx* c;
  const int64_t* pts;
  int64_t npts;
  int tid, nthreads;
  const orc_output* out;
  int status;
  double t_search, t_core, t_trans;
  int f32;
} worker_arg;

static void* worker(void* argp) {
  worker_arg* a = (worker_arg*)argp;
  orc_scratch s;
  r32_scratch r;
  memset(&r, 0, sizeof(r));
  const orc_problem* p = a->c->p;
  if (scratch_init(&s, p, a->c->loc.maxgroups)) { a->status = -1; scratch_free(&s); return NULL; }
  if (a->f32) {
    r.trans = (float*)malloc(sizeof(float) * p->k * p->k);
    r.bkg = (float*)malloc(sizeof(float) * p->k * p->ns);
    r.ana = (float*)malloc(sizeof(float) * p->k * p->ns);
    r.mean = (float*)malloc(sizeof(float) * p->ns);
    if (!r.trans || !r.bkg || !r.ana || !r.mean) a->status = -1;
  }
  /* round-robin ownership like letkf_mpi_setgrid (src/letkf_mpi.F90:213-225) */
  for (int64_t i = a->tid; i < a->npts && !a->status; i += a->nthreads) {
    const int rc = a->f32 ? analyse_point_f32(a->c, &s, &r, a->pts[i], i, a->out)
                          : analyse_point(a->c, &s, a->pts[i], i, a->out);
    if (rc) { a->status = -1; break; }
  }
  a->t_search = s.t_search;
  a->t_core = s.t_core;
  a->t_trans = s.t_trans;
  scratch_free(&s);
  free(r.hdxb); free(r.rdiag); free(r.rloc); free(r.dep); free(r.work);
  free(r.trans); free(r.bkg); free(r.ana); free(r.mean);
  return NULL;
}


static int run_analysis(const orc_problem* p, const int64_t* pts, int64_t npts,
                        int nthreads, const orc_output* out, int f32) {
  orc_ctx c;
  const double t0 = now_s();
  if (ctx_init(&c, p)) { ctx_free(&c); return -1; }
  const double t1 = now_s();
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  worker_arg args[256];
  int status = 0;
  for (int t = 0; t < nthreads; ++t) {
    args[t] = (worker_arg){&c, pts, npts, t, nthreads, out, 0, 0.0, 0.0, 0.0, f32};
    if (nthreads == 1) worker(&args[t]);
    else pthread_create(&th[t], NULL, worker, &args[t]);
  }
  for (int t = 0; t < nthreads; ++t) {
    if (nthreads > 1) pthread_join(th[t], NULL);
    if (args[t].status) status = -1;
  }
  if (out->seconds) {
    out->seconds[0] = t1 - t0;
    out->seconds[1] = now_s() - t1;
    out->seconds[2] = out->seconds[3] = out->seconds[4] = 0.0;
    for (int t = 0; t < nthreads; ++t) {
      out->seconds[2] += args[t].t_search;
      out->seconds[3] += args[t].t_core;
      out->seconds[4] += args[t].t_trans;
    }
  }
  ctx_free(&c);
  return status;
}

int orc_analysis(const orc_problem* p, const int64_t* pts, int64_t npts,
                 int nthreads, const orc_output* out) {
  return run_analysis(p, pts, npts, nthreads, out, 0);
}

/* the reference's loop in FP32 (CPU baseline); outputs widened to FP64,
 * only state, mean, sprd, lg_m and seconds are written */
int orc_analysis_f32(const orc_problem* p, const int64_t* pts, int64_t npts,
                     int nthreads, const orc_output* out) {
  return run_analysis(p, pts, npts, nthreads, out, 1);
}

int64_t orc_selection(const orc_problem* p, int64_t ij, int32_t group,
                      int32_t* idx_out, float* rloc_out, float* dist_out,
                      int64_t cap) {
  orc_ctx c;
  orc_scratch s;
  int64_t m = -1;
  if (ctx_init(&c, p) || scratch_init(&s, p, c.loc.maxgroups)) goto done;
  if (p->mask[ij]) { m = 0; goto done_s; }
  {
    float lat_ij = p->lat[ij];
    float rs = loc_maxhz(&c.loc, lat_ij) * ORC_STDEV2MAX;
    double g[3];
    orc_xyz(p->lat[ij], p->lon[ij], g);
    int64_t n = search(&c, g, orc_thr2(rs), s.cidx, s.cdist);
    int ngr = loc_groups(&c.loc, p, ij, s.gid, s.slab_group);
    if (group < 0 || group >= ngr) goto done_s;
    m = 0;
    for (int64_t i = 0; i < n; ++i) {
      float r = loc_localize(&c.loc, p, lat_ij, s.gid[group], s.cidx[i], s.cdist[i]);
      if (r > 1.0e-4f) {
        if (m < cap) { idx_out[m] = s.cidx[i]; rloc_out[m] = r; if (dist_out) dist_out[m] = s.cdist[i]; }
        ++m;
      }
    }
  }
done_s:
  scratch_free(&s);
done:
  ctx_free(&c);
  return m;
}
