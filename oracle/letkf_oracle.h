/*
 * letkf_oracle.h -- CPU restatement of the reference local analysis.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the GPU
 * library: only the tests, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  The product path never links it.
 *
 * Every routine computes in FP64 from the reference's FP32 inputs and
 * follows the reference file:line cited beside it.  The pieces that the
 * reference delegates to the absent geoKdTree submodule (distance formula,
 * Earth radius, boundary inclusion) are this project's own definition and
 * are marked "parity unpinned" against the reference; the GPU library is
 * pinned bit-exactly against them.
 */
#ifndef LETKF_ORACLE_H
#define LETKF_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "../include/letkf_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_problem {
  int64_t k, ns, nij, nobs;
  /* background, reference layout (src/letkf_state.F90:182-192) */
  const float* state;   /* (k, ns, nij) perturbations */
  const float* mean;    /* (ns, nij) */
  const float* sprd;    /* (ns, nij) */
  const float* lat;     /* (nij) */
  const float* lon;
  const int32_t* mask;
  /* observations (src/letkf_obs.F90:57-68,165-171) */
  const float* olat; const float* olon; const float* ozdim; const float* otime;
  const float* oval; const float* oerr;
  const int32_t* obsid; const int32_t* platid;
  const float* hx;      /* (k, nobs) perturbations */
  const float* hxmean;  /* (nobs) */
  /* state variables */
  int32_t nvars;
  const int32_t* var_s; const int32_t* var_lev;
  const float* inc_max; const float* bounds;
  /* solver */
  float rtps, rtpp, mul;
  const letkf_loc_desc* loc;
} orc_problem;

/* outputs for a list of gridpoints, packed in list order */
typedef struct orc_output {
  double* state;        /* (k, ns, npts) */
  double* mean;         /* (ns, npts) */
  double* sprd;         /* (ns, npts) */
  float* maxhz;         /* (npts) */
  int64_t* obscount;    /* (npts) */
  int64_t* lg_m;        /* (maxgroups, npts) */
  double* lg_obsloc;    /* (maxgroups, npts) */
  uint64_t* lg_hash;    /* (maxgroups, npts) */
  double* seconds;      /* optional [5]: search-structure set-up, gridpoint loop
                           (wall), then thread-seconds summed over the
                           workers in the reference's timer phases:
                           obs_search (search, localization, filter),
                           letkf_core (core solve), letkf_trans (apply and
                           post-processing) */
} orc_output;

/* deterministic elementary functions (shared definition with the GPU) */
void   orc_sincosd(double deg, double* s, double* c);
double orc_sin_rad(double x);
double orc_asin(double x);
double orc_exp_neg(double x);
void   orc_xyz(float lat, float lon, double* xyz);
double orc_chord2(const double* a, const double* b);
double orc_thr2(float radius_m);
float  orc_dist(double chord2);
uint64_t orc_mix64(uint64_t x);

/* reference routines */
float  orc_gc(float z, float L);
int    orc_interp_build(const letkf_linearinterp_lat* cfg, float* lat, float* dist);
float  orc_interp_get(int n, const float* lat, const float* dist, float x);
int    orc_maxgroups(const letkf_loc_desc* loc);

int orc_core_solve(int k, int m, const double* hdxb, const double* rdiag,
                   const double* rloc, const double* dep, double infl,
                   double* trans, double* eival);
int orc_eigh(int n, double* a, double* w, double* v);

int orc_analysis(const orc_problem* p, const int64_t* pts, int64_t npts,
                 int nthreads, const orc_output* out);

int orc_analysis_f32(const orc_problem* p, const int64_t* pts, int64_t npts,
                     int nthreads, const orc_output* out);

/* FP32 core in the reference's order (ref32_core.c) */
int r32_syev(int n, float* a, float* w, float* e);
size_t r32_core_work(int k, int m);
int r32_core_solve(int k, int m, const float* hdxb, const float* rdiag, const float* rloc,
                   const float* dep, float infl, float* trans, float* work);

int64_t orc_selection(const orc_problem* p, int64_t ij, int32_t group,
                      int32_t* idx_out, float* rloc_out, float* dist_out,
                      int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
