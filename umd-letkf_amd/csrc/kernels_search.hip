// kernels_search.hip -- local-observation search and localization.
//
// Replaces, per gridpoint, the reference's
//   maxhz / search radius ........ src/letkf_solver.F90:210-212
//   kd_search_radius ............. src/letkf_obs.F90:560-573
//   groups() ..................... letkf_loc_novrt.F90:133-147,
//                                  letkf_loc_ocean.F90:331-427
//   localize() + keep r > 1e-4 ... src/letkf_solver.F90:244-268,
//                                  letkf_loc_novrt.F90:170-183,
//                                  letkf_loc_ocean.F90:451-506
//
// One 64-lane wavefront owns one gridpoint.  The gridpoint's search cap is
// covered by whole lat/lon cells (a conservative superset computed with
// margins); lanes stride over the contiguous obs ranges of those cells and
// apply the exact test (squared chord <= threshold) and the localizer in
// the deterministic arithmetic of det_math.h.  Kept observations are counted
// per localization group with ballots.  The same kernel runs twice: the
// count pass sizes the lists, the fill pass writes each (gridpoint,
// group)'s kept (obs index, weight) list in scan order.  With several
// groups the count pass localizes (exact counts, diagnostics); with one it
// only sums the cover's cell sizes, and the fill pass writes the
// diagnostics.
#include <hip/hip_runtime.h>

#include "det_math.h"
#include "letkf_internal.h"

namespace lk {

constexpr double kRad2Deg = 57.295779513082320876;

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// exclusive prefix sum over the wave's lanes, and the total
__device__ __forceinline__ int wave_excl_scan(int v, int lane, int* total) {
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  *total = __shfl(x, 63, 64);
  return x - v;
}
__device__ __forceinline__ int lane_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// the localizer's scalars that the weights use (copied out of the kernel
// arguments, so that no reference to them is held)
struct LocScalars {
  float tloc_prof, tloc_sat, vt_scale;
};

// localization weight of one kept-candidate obs for group g (FP32 result)
__device__ __forceinline__ float loc_weight(int cls, const LocScalars& L, int g, float hz,
                                            float d, float Lp, float Ls,
                                            uint8_t surf, float twp, float tws,
                                            float zdim, const float* lev_depth) {
#pragma clang fp contract(off)
  if (cls == LETKF_LOC_OCEAN) {
    if (surf) {
      if (g == 1) return 0.0f;
      float r = dm::gc(d, Ls);
      if (L.tloc_sat > 0.0f) r = r * tws;
      return r;
    }
    float r = dm::gc(d, Lp);
    if (L.tloc_prof > 0.0f) r = r * twp;
    return r;
  }
  if (cls == LETKF_LOC_VGAUSS) {
    const double dz = (double)zdim - (double)lev_depth[g];
    const double sc = (double)L.vt_scale;
    const double v = dm::exp_neg(-(dz * dz) / (2.0 * sc * sc));
    return hz * (float)v;
  }
  return hz;
}

__device__ __forceinline__ bool group_has_slabs(int cls, const LocDev& L, int g, int mld) {
  if (cls == LETKF_LOC_OCEAN && L.vtloc_surf == 1) return g == 0 || L.max_var_lev > mld;
  if (cls == LETKF_LOC_VGAUSS) return g < L.max_var_lev;
  return true;
}

// dm::interp on a table held in the kernel arguments, with constant indices
// only (a table read at a computed index would make the compiler copy the
// whole argument block to private memory); the same arithmetic
__device__ __forceinline__ float interp_arg(int n, const float (&lat)[kTab],
                                            const float (&dist)[kTab], float x) {
#pragma clang fp contract(off)
  const float a = fabsf(x);
  int i = n - 1;
#pragma unroll
  for (int j = kTab - 1; j >= 0; --j)
    if (j < n - 1 && !(a > lat[j])) i = j;
  if (i == 0) return dist[0];
  float l1 = 0.0f, d1 = 0.0f, l0 = 0.0f, d0 = 0.0f;
#pragma unroll
  for (int j = 0; j < kTab; ++j) {
    if (j == i) { l1 = lat[j]; d1 = dist[j]; }
    if (j == i - 1) { l0 = lat[j]; d0 = dist[j]; }
  }
  const float d = (a - l0) / (l1 - l0);
  return d * d1 + (1.0f - d) * d0;
}

// Search radius of a gridpoint (src/letkf_solver.F90:210-212) and the
// localization distances it comes from.
struct Radii {
  float Lh, Lp, Ls, maxhz, rs;
};

__device__ __forceinline__ Radii search_radii(int cls, const LocDev& L, float lat_ij) {
  Radii R{0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if (cls == LETKF_LOC_OCEAN) {
    R.Lp = interp_arg(L.n_prof, L.prof_lat, L.prof_dist, lat_ij);
    R.Ls = interp_arg(L.n_sat, L.sat_lat, L.sat_dist, lat_ij);
    R.maxhz = R.Lp > R.Ls ? R.Lp : R.Ls;
  } else {
    R.Lh = interp_arg(L.n_hz, L.hz_lat, L.hz_dist, lat_ij);
    R.maxhz = R.Lh;
  }
  {
#pragma clang fp contract(off)
    R.rs = R.maxhz * dm::kSTDEV2MAX;                  // src/letkf_solver.F90:211
  }
  return R;
}

// Conservative cell cover of a spherical cap: bands b0..b1 and, in each
// band, the longitude cells the cap can reach within that band.  The count
// pass and the fill pass use these same functions without contraction, so
// both compute the same cells and the list space the count pass sizes holds
// every candidate the fill pass scans.
struct CapCover {
  double lat0, lon0, thd, hav_cap, cos_lat0;
  int b0, b1;
  bool all_lon;   // the cap reaches a pole: whole bands
};

__device__ __forceinline__ CapCover cap_cover(float rs, float lat, float lon) {
#pragma clang fp contract(off)
  CapCover cv;
  const double theta = (double)rs / dm::kRE * (1.0 + 1e-6) + 1e-9;
  cv.lat0 = (double)lat;
  cv.lon0 = (double)lon;
  cv.lon0 = cv.lon0 - 360.0 * floor(cv.lon0 / 360.0);
  cv.b0 = 0;
  cv.b1 = kBands - 1;
  cv.all_lon = true;
  cv.thd = 180.0;
  if (theta < 3.14159) {
    cv.thd = theta * kRad2Deg;
    int b0 = (int)floor(cv.lat0 - cv.thd + 90.0);
    int b1 = (int)floor(cv.lat0 + cv.thd + 90.0);
    cv.b0 = b0 < 0 ? 0 : (b0 >= kBands ? kBands - 1 : b0);
    cv.b1 = b1 < 0 ? 0 : (b1 >= kBands ? kBands - 1 : b1);
    cv.all_lon = !(cv.lat0 + cv.thd < 90.0 && cv.lat0 - cv.thd > -90.0);
  }
  const double sh = sin(0.5 * theta);
  cv.hav_cap = sh * sh;
  cv.cos_lat0 = cos(cv.lat0 / kRad2Deg);
  return cv;
}

// longitude half-width (degrees) that bounds the cap within band b, from
//   hav(dlon) = (hav(cap) - hav(dlat)) / (cos lat0 cos lat),
// with dlat at its smallest over the band and |lat| at its largest over the
// band's part of the cap's latitude range; < 0: the band misses the cap.
// Bands past the 64th of the cover (caps wider than 32 degrees) take whole
// longitude circles.
__device__ __forceinline__ double cover_halfwidth(const CapCover& cv, int b) {
#pragma clang fp contract(off)
  if (cv.all_lon || b - cv.b0 >= 64) return 180.0;
  const double la = (double)b - 90.0, lb = la + 1.0;
  const double dmin = cv.lat0 < la ? la - cv.lat0 : (cv.lat0 > lb ? cv.lat0 - lb : 0.0);
  const double sd = sin(0.5 * dmin / kRad2Deg);
  const double num = cv.hav_cap - sd * sd;
  if (num <= 0.0) return -1.0;
  const double lo = fmax(la, cv.lat0 - cv.thd), hi = fmin(lb, cv.lat0 + cv.thd);
  const double amax = lo <= hi ? fmax(fabs(lo), fabs(hi)) : fmax(fabs(la), fabs(lb));
  const double q = num / (cv.cos_lat0 * cos(amax / kRad2Deg));
  if (!(q < 1.0)) return 180.0;
  return 2.0 * asin(sqrt(q)) * kRad2Deg * (1.0 + 1e-9) + 1e-6;
}

// the one or two longitude cell ranges [rc0, rc1] of half-width dl (< 180)
__device__ __forceinline__ int cover_ranges(const CapCover& cv, double dl, int rc0[2],
                                            int rc1[2]) {
#pragma clang fp contract(off)
  rc0[0] = 0; rc0[1] = 0;
  rc1[0] = kLonCells - 1; rc1[1] = kLonCells - 1;
  if (!(dl < 180.0)) return 1;
  const int c0 = (int)floor(cv.lon0 - dl), c1 = (int)floor(cv.lon0 + dl);
  if (c0 < 0) { rc0[0] = c0 + kLonCells; rc1[0] = kLonCells - 1; rc0[1] = 0; rc1[1] = c1; return 2; }
  if (c1 >= kLonCells) { rc0[0] = c0; rc1[0] = kLonCells - 1; rc0[1] = 0; rc1[1] = c1 - kLonCells; return 2; }
  rc0[0] = c0;
  rc1[0] = c1;
  return 1;
}

// Passes (one wave per gridpoint):
//   kCountExact  kept counts and diagnostics per group (weights evaluated)
//   kFill        lists written at the offsets of the exact counts
//   kFillDiag    one group: lists written at the offsets of the bounds that
//                k_count_cover (below) sizes, then the kept count, the solve
//                flag, the search-set size and the other diagnostics
constexpr int kCountExact = 0, kFill = 1, kFillDiag = 3;
constexpr int kQueue = 128;   // per-wave candidate ring (two batches)

// orders this wave's LDS stores and loads on either side (the queue is
// private to one wave; the block's waves retire independently, so no
// block barrier)
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

// the one-group fill pass (the common case) keeps one group in LDS, which
// leaves six waves per SIMD to hide the candidates' load latency
template <int PASS, int CLS>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(PASS == 3 ? 6 : 1, 8))) void k_search(LocDev L, ObsDev O, StateDev S,
                                                VarsDev V, ChunkDev C, DiagDev D) {
  constexpr bool FILL = PASS == kFill || PASS == kFillDiag;
  constexpr bool STATS = PASS == kCountExact || PASS == kFillDiag;   // sums, hash
  constexpr int NG = PASS == kFillDiag ? 1 : kMaxGroups;   // groups held in LDS
  __shared__ int64_t s_cnt[4][NG];
  __shared__ double s_sum[4][NG];
  __shared__ uint64_t s_hash[4][NG];
  __shared__ int64_t s_lim[4][NG];               // fill: end of each list's space
  __shared__ int32_t s_qs[4][kQueue];             // queued candidates: obs index,
  __shared__ double s_qc[4][kQueue];              //   squared chord,
  __shared__ float s_qe[4][kQueue];               //   error variance
  __shared__ int32_t s_qo[4][kQueue];             //   and original index
  __shared__ int32_t s_rsb[4][129];               // cover ranges: first obs
  __shared__ int32_t s_roff[4][130];              //   and offset in the candidate stream
  const int wid = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t t = (int64_t)blockIdx.x * 4 + wid;
  if (t >= C.ngp) return;
  const int64_t ij = C.i0 + t;
  const int G = C.G;

  if (lane == 0) {
    for (int g = 0; g < G; ++g) {
      s_cnt[wid][g] = FILL ? C.off[t * G + g] : 0;
      s_lim[wid][g] = FILL ? C.off[t * G + g] + C.cnt[t * G + g] : 0;
      s_sum[wid][g] = 0.0;
      s_hash[wid][g] = 0;
    }
  }
  if (S.mask[ij]) {                                   // src/letkf_solver.F90:207
    if (PASS == kCountExact) {
      if (lane != 0) return;
      for (int g = 0; g < G; ++g) {
        C.cnt[t * G + g] = 0;
        C.solve[t * G + g] = 0;
        D.lg_m[ij * G + g] = 0;
        D.lg_obsloc[ij * G + g] = 0.0;
        D.lg_hash[ij * G + g] = 0;
      }
      D.obscount[ij] = 0;
      D.maxhz[ij] = 0.0f;
      D.mld[ij] = 0;
    }
    return;
  }

  const float lat_ij = S.lat[ij];
  const Radii R = search_radii(CLS, L, lat_ij);
  const LocScalars LS{L.tloc_prof, L.tloc_sat, L.vt_scale};
  const float Lh = R.Lh, Lp = R.Lp, Ls = R.Ls, maxhz = R.maxhz, rs = R.rs;
  const double thr2 = dm::thr2(rs);

  int mld = 0;                                        // calc_mldlvl_bkgt
  if (CLS == LETKF_LOC_OCEAN && L.vtloc_surf == 1) {
    const float* mean = S.mean + S.ns * ij;
    const float t0 = mean[L.bkg_t_var_idx];
    mld = L.bkg_t_var_len;
    for (int i = 1; i <= L.bkg_t_var_len - 1; ++i) {
      const float t1 = mean[L.bkg_t_var_idx + i];
      if (fabsf(t0 - t1) > L.bkg_t_delta) { mld = i; break; }
    }
  }

  const double gx = S.gx[ij], gy = S.gy[ij], gz = S.gz[ij];

  const CapCover cv = cap_cover(rs, lat_ij, S.lon[ij]);
  const int b0 = cv.b0, b1 = cv.b1;

  int64_t nin = 0, ncand = 0;
  double lsum = 0.0;          // kFillDiag: this lane's weights and hashes
  uint64_t lhash = 0;

  // Candidates inside the cap go through a per-wave queue in LDS (a ring of
  // 128 entries, in scan order) and are localized 64 at a time, so that the
  // distance, weight, hash and list work runs on full waves; lists are
  // written in the same scan order as without the queue.
  int qh = 0, qn = 0;          // ring head and length (wave-uniform)
  // one group: the list cursor and end stay in (wave-uniform) registers
  int64_t fd_cur = 0, fd_lim = 0;
  if (PASS == kFillDiag) {
    fd_cur = C.off[t];
    fd_lim = fd_cur + C.cnt[t];
  }
  auto localize = [&](int cnt) __attribute__((always_inline)) {
    wave_lds_fence();
    const bool inside = lane < cnt;
    int64_t s = 0;
    double c2 = 9.0;
    float oerr = 1.0f;
    int32_t oid = 0;
    if (inside) {
      const int e = (qh + lane) & (kQueue - 1);
      s = s_qs[wid][e];
      c2 = s_qc[wid][e];
      oerr = s_qe[wid][e];
      oid = s_qo[wid][e];
    }
    wave_lds_fence();
    float d = 0.0f, hz = 0.0f, zd = 0.0f, twp = 1.0f, tws = 1.0f;
    uint8_t sf = 0;
    uint64_t mix = 0;
    if (inside) {
      d = dm::dist(c2);
      if (CLS != LETKF_LOC_OCEAN) hz = dm::gc(d, Lh);
      if (CLS == LETKF_LOC_OCEAN) { sf = O.surf[s]; twp = O.tw_prof[s]; tws = O.tw_sat[s]; }
      if (CLS == LETKF_LOC_VGAUSS) zd = O.zdim[s];
      if (STATS) mix = dm::mix64((uint64_t)oid);
    }
    for (int g = 0; g < G; ++g) {
      float w = 0.0f;
      if (inside) w = loc_weight(CLS, LS, g, hz, d, Lp, Ls, sf, twp, tws, zd, V.lev_depth);
      const bool keep = inside && (w > 1.0e-4f);   // src/letkf_solver.F90:251
      const uint64_t kmask = __ballot(keep);
      if (kmask == 0) continue;
      if (PASS == kFillDiag) {
        // one group: per-lane partial sums, reduced once per gridpoint
        lsum += keep ? (double)w : 0.0;
        lhash += keep ? mix : 0ull;
      } else if (STATS) {
        const double sr = wave_sum_f64(keep ? (double)w : 0.0);
        const uint64_t hs = wave_sum_u64(keep ? mix : 0ull);
        if (lane == 0) {
          if (!FILL) s_cnt[wid][g] += __popcll(kmask);
          s_sum[wid][g] += sr;
          s_hash[wid][g] += hs;
        }
      }
      if (FILL) {
        // writes never pass the end of this pair's list space
        int64_t cur = fd_cur, lim = fd_lim;
        if (PASS != kFillDiag) {
          if (lane == 0) { cur = s_cnt[wid][g]; lim = s_lim[wid][g]; }
          cur = __shfl(cur, 0, 64);
          lim = __shfl(lim, 0, 64);
        }
        if (keep) {
          const int64_t pos = cur + lane_rank(kmask);
          if (pos < lim) {
            // mixed products: byte offset of the padded operand column and
            // sqrt(w / err) by the hardware reciprocal and square root
            // (1 ulp each; the products round far more)
            C.list_idx[pos] = C.list_mode ? (int32_t)(s * O.kp * 4) : (int32_t)s;
            C.list_w[pos] = C.list_mode
                                ? __builtin_amdgcn_sqrtf(w * __builtin_amdgcn_rcpf(oerr))
                                : w;
          }
        }
        if (PASS == kFillDiag) fd_cur = cur + __popcll(kmask);
        else if (lane == 0) s_cnt[wid][g] = cur + __popcll(kmask);
      }
    }
    qh = (qh + cnt) & (kQueue - 1);
    qn -= cnt;
  };

  // The cover's cells: lane i finds the one or two longitude ranges of band
  // b0 + i and their obs ranges; bands past the 64th of the cover take whole
  // longitude circles, which follow one another in the cell order and make
  // one more range.  The ranges, in band order, make one stream of
  // candidates, scanned 64 at a time in the same order as band by band.  The
  // positions (and the fields the localization needs) of a batch are loaded
  // two batches ahead.
  const bool need_err = FILL && C.list_mode;
  struct Cand {
    double x, y, z;
    float e;
    int32_t o, s;       // s < 0: past the stream
  };
  {
    const int b = b0 + lane;
    int nr = 0, sb0 = 0, len0 = 0, sb1 = 0, len1 = 0;
    if (b <= b1) {
      const double dl = cover_halfwidth(cv, b);
      if (dl >= 0.0) {
        int rc0[2], rc1[2];
        nr = cover_ranges(cv, dl, rc0, rc1);
        const int32_t* cs = O.cell_start + b * kLonCells;
        sb0 = cs[rc0[0]];
        len0 = cs[rc1[0] + 1] - sb0;
        if (nr == 2) {
          sb1 = cs[rc0[1]];
          len1 = cs[rc1[1] + 1] - sb1;
        }
      }
    }
    int nrt = 0, T = 0;
    const int slot = wave_excl_scan(nr, lane, &nrt);
    const int off = wave_excl_scan(len0 + len1, lane, &T);
    if (nr >= 1) {
      s_rsb[wid][slot] = sb0;
      s_roff[wid][slot] = off;
    }
    if (nr == 2) {
      s_rsb[wid][slot + 1] = sb1;
      s_roff[wid][slot + 1] = off + len0;
    }
    if (b1 - b0 >= 64) {
      const int st0 = O.cell_start[(b0 + 64) * kLonCells];
      const int st1 = O.cell_start[(b1 + 1) * kLonCells];
      if (lane == 0) {
        s_rsb[wid][nrt] = st0;
        s_roff[wid][nrt] = T;
      }
      nrt += 1;
      T += st1 - st0;
    }
    if (lane == 0) s_roff[wid][nrt] = T;
    wave_lds_fence();
    if (PASS != kFill) ncand += T;
    int rl = 0;                       // range of this lane's next load
    // the loads are issued by every lane (past the stream: observation 0),
    // so that the wait for a batch counts the loads issued after it
    auto fetch = [&](Cand& c, int q) __attribute__((always_inline)) {
      const int p = q + lane;
      int so = 0;
      if (p < T) {
        while (p >= s_roff[wid][rl + 1]) ++rl;
        so = s_rsb[wid][rl] + (p - s_roff[wid][rl]);
      }
      c.s = p < T ? so : -1;
      c.x = O.x[so];
      c.y = O.y[so];
      c.z = O.z[so];
      if (FILL) c.e = O.err[so];
      if (STATS) c.o = O.oidx[so];
    };
    auto scan = [&](const Cand& c) __attribute__((always_inline)) {
      double c2 = 9.0;
      if (c.s >= 0) c2 = dm::chord2(c.x, c.y, c.z, gx, gy, gz);
      const bool inside = c2 <= thr2;
      const uint64_t imask = __ballot(inside);
      if (imask == 0) return;
      nin += __popcll(imask);
      // qn < 64 here, so the ring (128) has room for the whole batch
      if (inside) {
        const int e = (qh + qn + lane_rank(imask)) & (kQueue - 1);
        s_qs[wid][e] = c.s;
        s_qc[wid][e] = c2;
        s_qe[wid][e] = need_err ? c.e : 1.0f;
        s_qo[wid][e] = STATS ? c.o : 0;
      }
      qn += __popcll(imask);
      if (qn >= 64) localize(64);
    };
    // three batches in registers, each in the same ones throughout (the
    // loop is unrolled three times, so no batch is moved while its loads
    // are in flight): the one scanned and the two loading
    Cand ca, cb, cc;
    if (T > 0) {
      fetch(ca, 0);
      fetch(cb, 64);
    }
    for (int q = 0; q < T; q += 192) {
      fetch(cc, q + 128);
      scan(ca);
      if (q + 64 >= T) break;
      fetch(ca, q + 192);
      scan(cb);
      if (q + 128 >= T) break;
      fetch(cb, q + 256);
      scan(cc);
    }
    wave_lds_fence();
  }
  if (qn > 0) localize(qn);

  if (PASS == kFillDiag) {
    const double sr = wave_sum_f64(lsum);
    const uint64_t hs = wave_sum_u64(lhash);
    if (lane == 0) {
      s_sum[wid][0] = sr;
      s_hash[wid][0] = hs;
      // the kept count never exceeds the list space the count pass sized
      // (the cover bound); a pair whose space was withdrawn (list capacity
      // guard) ends with no list
      s_cnt[wid][0] = fd_cur < fd_lim ? fd_cur : fd_lim;
    }
  }
  if (FILL && C.list_mode) {
    // zero the alignment padding after each list: the mixed products read
    // whole 8-entry groups
    for (int g = 0; g < G; ++g) {
      int64_t end = 0;
      if (lane == 0) {
        end = s_cnt[wid][g];
        const int64_t lim = PASS == kFillDiag ? fd_lim : s_lim[wid][g];
        if (end > lim) end = lim;      // only past the space of a guarded pair
      }
      end = __shfl(end, 0, 64);
      const int64_t pad_end = C.off[t * G + g] + PadCount()(end - C.off[t * G + g]);
      if (end + lane < pad_end) {
        C.list_idx[end + lane] = 0;
        C.list_w[end + lane] = 0.0f;
      }
    }
  }
  if (lane != 0 || PASS == kFill) return;
  unsigned long long* c = D.ctr + (blockIdx.x & (kCtrSlots - 1)) * kCtrStride;
  if (PASS == kCountExact || PASS == kFillDiag) {
    unsigned long long nsolve = 0, npair = 0;
    for (int g = 0; g < G; ++g) {
      const int64_t m = PASS == kFillDiag ? s_cnt[wid][g] - C.off[t * G + g] : s_cnt[wid][g];
      const bool sv = m > 0 && group_has_slabs(CLS, L, g, mld);
      nsolve += sv ? 1 : 0;
      npair += sv ? (unsigned long long)m : 0ull;
      C.cnt[t * G + g] = m;
      C.solve[t * G + g] = sv ? 1 : 0;
      D.lg_m[ij * G + g] = (int32_t)m;
      D.lg_obsloc[ij * G + g] = s_sum[wid][g];
      D.lg_hash[ij * G + g] = s_hash[wid][g];
    }
    atomicAdd(c + kCtrSolves, nsolve);
    atomicAdd(c + kCtrPairs, npair);
  }
  if (PASS == kCountExact || PASS == kFillDiag) {
    D.obscount[ij] = (int32_t)nin;
    D.maxhz[ij] = maxhz;
    D.mld[ij] = mld;
    atomicMax(c + kCtrMax, (unsigned long long)nin);
    atomicAdd(c + kCtrCand, (unsigned long long)ncand);
  }
}

// One-group count pass, one thread per gridpoint: the number of
// observations in the cells of the cap's cover, an upper bound of the kept
// count that sizes the list (the fill pass scans the same cells).  Masked
// gridpoints get zero counts and diagnostics.
__global__ __launch_bounds__(256) void k_count_cover(LocDev L, ObsDev O, StateDev S,
                                                     ChunkDev C, DiagDev D) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= C.ngp) return;
  const int64_t ij = C.i0 + t;
  if (S.mask[ij]) {                                   // src/letkf_solver.F90:207
    C.cnt[t] = 0;
    C.solve[t] = 0;
    D.lg_m[ij] = 0;
    D.lg_obsloc[ij] = 0.0;
    D.lg_hash[ij] = 0;
    D.obscount[ij] = 0;
    D.maxhz[ij] = 0.0f;
    D.mld[ij] = 0;
    return;
  }
  const float lat_ij = S.lat[ij];
  const CapCover cv = cap_cover(search_radii(L.cls, L, lat_ij).rs, lat_ij, S.lon[ij]);
  int64_t ncand = 0;
  for (int b = cv.b0; b <= cv.b1; ++b) {
    const double dl = cover_halfwidth(cv, b);
    if (dl < 0.0) continue;
    int rc0[2], rc1[2];
    const int nr = cover_ranges(cv, dl, rc0, rc1);
    for (int r = 0; r < nr; ++r)
      ncand += O.cell_start[b * kLonCells + rc1[r] + 1] - O.cell_start[b * kLonCells + rc0[r]];
  }
  C.cnt[t] = ncand;
}

// The graph-captured analysis.  Its counters are cleared by a kernel (a
// captured graph holds no memset nodes), and the list space is checked on
// the device instead of by the host: a (gridpoint, group) pair whose padded
// list (plus the alignment slack of the vector loads) would end past the
// list capacity gets no list space and no solve, and the overflow is
// counted for the host, which reads the counters after the replay.
__global__ void k_clear_u64(int64_t n, unsigned long long* p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

__global__ void k_list_guard(int64_t np, int64_t* cnt, const int64_t* off, uint8_t* solve,
                             int64_t cap, unsigned long long* ctr) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  if (off[p] + PadCount()(cnt[p]) + kListAlign > cap) {
    cnt[p] = 0;
    solve[p] = 0;
    atomicAdd(ctr + kCtrOverflow, 1ull);
  }
}

void launch_clear_u64(hipStream_t st, int64_t n, unsigned long long* p) {
  if (n <= 0) return;
  k_clear_u64<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, p);
}

void launch_list_guard(hipStream_t st, int64_t np, int64_t* cnt, const int64_t* off,
                       uint8_t* solve, int64_t cap, unsigned long long* ctr) {
  if (np <= 0) return;
  k_list_guard<<<(unsigned)((np + 255) / 256), 256, 0, st>>>(np, cnt, off, solve, cap, ctr);
}

// the search kernel for localizer class CLS (its branches folded away)
template <int CLS>
static void launch_search_cls(hipStream_t st, int pass, unsigned nb, const LocDev& L,
                              const ObsDev& O, const StateDev& S, const VarsDev& V,
                              const ChunkDev& C, const DiagDev& D) {
  if (pass == kFillDiag) k_search<kFillDiag, CLS><<<nb, 256, 0, st>>>(L, O, S, V, C, D);
  else if (pass == kFill) k_search<kFill, CLS><<<nb, 256, 0, st>>>(L, O, S, V, C, D);
  else k_search<kCountExact, CLS><<<nb, 256, 0, st>>>(L, O, S, V, C, D);
}

void launch_search(hipStream_t st, bool fill, const LocDev& L, const ObsDev& O,
                   const StateDev& S, const VarsDev& V, const ChunkDev& C,
                   const DiagDev& D) {
  if (C.ngp <= 0) return;
  const unsigned nb = (unsigned)((C.ngp + 3) / 4);
  // one group: size the lists by the search set (chord test only) and count
  // the kept observations while writing them
  if (C.G == 1 && !fill) {
    k_count_cover<<<(unsigned)((C.ngp + 255) / 256), 256, 0, st>>>(L, O, S, C, D);
    return;
  }
  const int pass = C.G == 1 ? kFillDiag : (fill ? kFill : kCountExact);
  switch (L.cls) {
    case LETKF_LOC_OCEAN: launch_search_cls<LETKF_LOC_OCEAN>(st, pass, nb, L, O, S, V, C, D); break;
    case LETKF_LOC_VGAUSS: launch_search_cls<LETKF_LOC_VGAUSS>(st, pass, nb, L, O, S, V, C, D); break;
    default: launch_search_cls<LETKF_LOC_NOVRT>(st, pass, nb, L, O, S, V, C, D); break;
  }
}

}  // namespace lk
