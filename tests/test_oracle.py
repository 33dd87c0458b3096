"""Pins the FP64 CPU oracle (oracle/letkf_oracle.c) before it is trusted.

Every check below comes from outside the oracle's own formulas: published
values of the truncated Gaspari-Cohn polynomial, closed forms of the LETKF
transform, numpy/scipy library routines, brute force, and invariants that a
dropped term, a wrong sign or a transposed operand would break.
"""
import math

import numpy as np
import pytest

import oracle as O
from letkf_amd import synth
from letkf_amd.problem import LinearInterpLat

RE = 6371000.0


# ------------------------------------------------------------ Gaspari-Cohn
def test_gc_known_values():
    """Values of the truncated polynomial of src/letkf_loc.F90:362-376 at
    z/c = 0, 0.5, 1, 1.5, 1.9, 2 (evaluated by hand from the coefficients):
    1.0, 0.684895, 0.20833, 0.016473, -4.195e-5, 0."""
    L = 300e3
    c = L / math.sqrt(0.3)
    want = {0.0: 1.0, 0.5: 0.684895, 1.0: 0.20833, 1.5: 0.016473, 1.9: -4.195e-5}
    for zc, v in want.items():
        got = O.gc(np.float32(zc * c), L)
        assert abs(got - v) <= 2e-6 * max(1.0, abs(v)) + 5e-8, (zc, got, v)
    assert O.gc(np.float32(2.0 * c * 1.0000001), L) == 0.0
    assert O.gc(np.float32(5.0 * c), L) == 0.0


def test_gc_keep_edge_and_continuity():
    """r > 1e-4 switches off near z/c = 1.84728 (exact GC: 1.86485);
    the two branches meet at z/c = 1 within the truncation (1e-5)."""
    L = 500e3
    c = L / math.sqrt(0.3)
    assert O.gc(np.float32(1.8460 * c), L) > 1e-4
    assert O.gc(np.float32(1.8485 * c), L) <= 1e-4
    lo = O.gc(np.float32(c * (1 - 1e-7)), L)
    hi = O.gc(np.float32(c * (1 + 1e-7)), L)
    assert abs(lo - hi) < 2e-5


def test_gc_close_to_exact_gc():
    """The truncated coefficients stay within 1e-4 of the exact
    Gaspari-Cohn (1999, eq. 4.10) function: a wrong coefficient fails."""
    L = 400e3
    c = L / math.sqrt(0.3)
    for zc in np.linspace(0.0, 1.99, 200):
        z = zc
        if z <= 1:
            ex = -0.25 * z**5 + 0.5 * z**4 + 0.625 * z**3 - 5.0 / 3.0 * z**2 + 1
        else:
            ex = (z**5 / 12 - 0.5 * z**4 + 0.625 * z**3 + 5.0 / 3.0 * z**2
                  - 5 * z + 4 - 2.0 / (3.0 * z))
        got = O.gc(np.float32(zc * c), L)
        assert abs(got - ex) < 1.2e-4, (zc, got, ex)


# ------------------------------------------------------------ elementary
def test_sincos_asin_exp_against_libm():
    g = np.random.default_rng(1)
    for d in np.concatenate([g.uniform(-360, 720, 2000), np.arange(-360, 721, 15.0)]):
        d = float(np.float32(d))
        s, c = O.sincosd(d)
        rd = math.radians(d - 360.0 * round(d / 360.0))   # exact reduction first
        assert abs(s - math.sin(rd)) < 6e-16
        assert abs(c - math.cos(rd)) < 6e-16
    L = O.lib()
    for x in np.concatenate([g.uniform(0, 1, 2000), [0.0, 0.2, 0.5, 0.9999, 1.0]]):
        a = L.orc_asin(float(x))
        assert abs(a - math.asin(x)) <= 4e-16 * max(1.0, math.asin(x)), x
    for x in np.concatenate([-g.exponential(5.0, 2000), [0.0, -700.0]]):
        e = L.orc_exp_neg(float(x))
        assert abs(e - math.exp(x)) <= 4e-16 * math.exp(x) + 1e-300, x
    for x in g.uniform(0, math.pi / 2 - 1e-9, 500):
        assert abs(L.orc_sin_rad(float(x)) - math.sin(x)) < 3e-16


def test_distance_and_threshold_closed_forms():
    """chord threshold (2 sin(R/2Re))^2 and great-circle distance against
    the haversine formula."""
    g = np.random.default_rng(2)
    L = O.lib()
    for r in [1.0, 1e3, 3e5, 1.8e6, 1e7, 1.9e7]:
        want = (2 * math.sin(float(np.float32(r)) / (2 * RE))) ** 2
        assert abs(L.orc_thr2(r) - want) <= 1e-15 * want
    assert L.orc_thr2(2.1e7) == 5.0                      # whole sphere
    for _ in range(300):
        la1, la2 = np.float32(g.uniform(-90, 90, 2))
        lo1, lo2 = np.float32(g.uniform(0, 360, 2))
        a, b = O.xyz(la1, lo1), O.xyz(la2, lo2)
        c2 = float(((a - b) ** 2).sum())
        p1, p2 = math.radians(la1), math.radians(la2)
        h = math.sin((p2 - p1) / 2) ** 2 + math.cos(p1) * math.cos(p2) * math.sin(
            math.radians(lo2 - lo1) / 2) ** 2
        hv = 2 * RE * math.asin(min(1.0, math.sqrt(h)))
        d = L.orc_dist(c2)
        assert abs(d - hv) <= 2e-7 * hv + 1e-3


# ------------------------------------------------------------ linearinterp
def test_linearinterp_lat_endpoints():
    """src/letkf_loc.F90:275-293 adds 0 and 90 degree end points with the
    nearest radius; get_dist interpolates linearly in |lat|."""
    t = LinearInterpLat([10.0, 50.0], [1000e3, 500e3])
    lat, dist = O.interp_table(t)
    assert lat.tolist() == [0.0, 10.0, 50.0, 90.0]
    assert dist.tolist() == [1000e3, 1000e3, 500e3, 500e3]
    assert O.interp(t, 5.0) == 1000e3
    assert O.interp(t, -30.0) == pytest.approx(750e3, rel=1e-6)
    assert O.interp(t, 30.0) == pytest.approx(750e3, rel=1e-6)
    assert O.interp(t, 70.0) == 500e3
    assert O.interp(t, 0.0) == 1000e3
    # the ioda/default.yaml table: explicit 0, missing 90
    t2 = LinearInterpLat([0.0, 10.0, 50.0], [1000e3, 1000e3, 500e3])
    lat2, _ = O.interp_table(t2)
    assert lat2.tolist() == [0.0, 10.0, 50.0, 90.0]
    # simple3d: 500 km at 0, 50 km at 90
    t3 = LinearInterpLat([0.0, 90.0], [500e3, 50e3])
    assert O.interp(t3, 45.0) == pytest.approx(275e3, rel=1e-6)


# ------------------------------------------------------------ eigen / core
def test_eigh_against_numpy():
    g = np.random.default_rng(3)
    for n in (5, 20, 50):
        x = g.standard_normal((n, 3 * n))
        a = x @ x.T + (n - 1) * np.eye(n)
        w, v = O.eigh(a)
        w0 = np.linalg.eigvalsh(a)
        assert np.allclose(w, w0, rtol=1e-13, atol=0)
        assert np.allclose(v @ np.diag(w) @ v.T, a, rtol=0, atol=1e-12 * np.abs(a).max())
        assert np.allclose(v.T @ v, np.eye(n), atol=1e-13)


def _core_numpy(hdxb, rdiag, rloc, dep):
    """Independent restatement with numpy's LAPACK eigensolver."""
    k = hdxb.shape[0]
    hr = hdxb / rdiag * rloc
    a = hr @ hdxb.T + (k - 1) * np.eye(k)
    w, v = np.linalg.eigh(a)
    pa = (v / w) @ v.T
    wbar = pa @ (hr @ dep)
    return (v * np.sqrt((k - 1) / w)) @ v.T + wbar[:, None]


@pytest.mark.parametrize("k,m", [(5, 1), (20, 7), (20, 19), (20, 20), (50, 300), (40, 2000)])
def test_core_against_numpy(k, m):
    g = np.random.default_rng(k * 1000 + m)
    y = g.standard_normal((k, m))
    y -= y.mean(0, keepdims=True)
    rdiag = g.uniform(0.05, 1.0, m)
    rloc = g.uniform(1e-4, 1.0, m)
    dep = g.standard_normal(m)
    t, _ = O.core_solve(y, rdiag, rloc, dep)
    t0 = _core_numpy(y, rdiag, rloc, dep)
    assert np.abs(t - t0).max() <= 1e-12 * np.abs(t0).max()


def test_core_single_obs_closed_form():
    """m = 1 (Sherman-Morrison): w = (rloc/r) d y / ((k-1) + (rloc/r)|y|^2);
    W = I + (sqrt((k-1)/lam) - 1) y y^T/|y|^2 with lam = (k-1)+(rloc/r)|y|^2."""
    g = np.random.default_rng(5)
    k = 20
    y = g.standard_normal(k)
    y -= y.mean()
    r, rl, d = 0.3, 0.7, 1.3
    t, ev = O.core_solve(y[:, None], [r], [rl], [d])
    q = rl / r
    lam = (k - 1) + q * y @ y
    wbar = q * d * y / lam
    W = np.eye(k) + (math.sqrt((k - 1) / lam) - 1) * np.outer(y, y) / (y @ y)
    assert np.abs(t - (W + wbar[:, None])).max() < 1e-14
    assert ev.max() == pytest.approx(lam, rel=1e-14)


def test_core_invariants():
    """Identical members -> T = I; W symmetric and W 1 = 1 when Yb 1 = 0;
    the transform keeps the analysis perturbations centred."""
    k, m = 30, 400
    t, _ = O.core_solve(np.zeros((k, m)), np.ones(m), np.ones(m), np.ones(m))
    assert np.abs(t - np.eye(k)).max() < 1e-15
    g = np.random.default_rng(9)
    y = g.standard_normal((k, m))
    y -= y.mean(0, keepdims=True)
    dep = g.standard_normal(m)
    t, _ = O.core_solve(y, np.full(m, 0.2), g.uniform(0, 1, m), dep)
    wbar = t.mean(1) - 1.0 / k * 0   # T = W + wbar 1^T, so W = T - wbar 1^T
    # recover wbar: W 1 = 1 -> T 1 = 1 + k wbar
    wbar = (t.sum(1) - 1.0) / k
    W = t - wbar[:, None]
    assert np.abs(W - W.T).max() < 1e-13
    assert np.abs(W.sum(1) - 1.0).max() < 1e-12
    xb = g.standard_normal((5, k))
    xb -= xb.mean(1, keepdims=True)
    assert np.abs((xb @ W).mean(1)).max() < 1e-12


# ------------------------------------------------------------ selection
def _gc_np(z, L):
    c = np.float64(L) / math.sqrt(0.3)
    az = np.abs(z.astype(np.float64))
    zc = az / c
    out = np.where(zc <= 1, -0.25 * zc**5 + 0.5 * zc**4 + 0.625 * zc**3 - 1.66667 * zc**2 + 1,
                   0.08333 * zc**5 - 0.5 * zc**4 + 0.625 * zc**3 + 1.66667 * zc**2 - 5 * zc
                   + 4 - 0.66667 * c / np.maximum(az, 1e-300))
    return np.where(az >= 2 * c, 0.0, out).astype(np.float32)


def test_selection_brute_force():
    """Kept set of every gridpoint equals a numpy brute force (haversine
    distance, numpy GC); only obs within 1e-6 of the keep edge may differ."""
    p = synth.random_problem(seed=11, nx=24, ny=12, nobs=1500, radius=900e3)
    out = O.analysis(p)
    plat, plon = np.radians(p.lat.astype(np.float64)), np.radians(p.lon.astype(np.float64))
    olat, olon = np.radians(p.olat.astype(np.float64)), np.radians(p.olon.astype(np.float64))
    lat_t, dist_t = O.interp_table(p.loc.hzloc)
    mismatch = 0
    for ij in range(p.nij):
        if p.mask[ij]:
            assert out["lg_m"][0, ij] == 0
            continue
        h = (np.sin((olat - plat[ij]) / 2) ** 2 + np.cos(plat[ij]) * np.cos(olat)
             * np.sin((olon - plon[ij]) / 2) ** 2)
        d = (2 * RE * np.arcsin(np.sqrt(np.minimum(h, 1.0)))).astype(np.float32)
        L = np.float32(np.interp(abs(p.lat[ij]), lat_t, dist_t))
        r = _gc_np(d, L)
        keep = np.nonzero(r > 1e-4)[0]
        idx, rl, _ = O.selection(p, ij)
        if not np.array_equal(np.sort(keep), idx):
            edge = np.abs(r - 1e-4) < 1e-6
            diff = np.setxor1d(keep, idx)
            assert edge[diff].all(), ij
            mismatch += diff.size
        assert out["lg_m"][0, ij] == idx.size
        h64 = np.uint64(0)
        for i in idx:
            h64 = np.uint64((int(h64) + int(O.lib().orc_mix64(int(i)))) % (1 << 64))
        assert out["lg_hash"][0, ij] == h64
    assert mismatch <= 2


# ------------------------------------------------------------ full analysis
def _single_obs_problem(k=20, r=0.5, d=0.8, rtps=0.0, rtpp=0.0, mul=1.0):
    """One obs exactly on a gridpoint, hx = that gridpoint's members."""
    p = synth.random_problem(seed=21, nx=24, ny=12, k=k, nobs=10, ns=1, with_mask=False,
                             rtps=rtps, rtpp=rtpp, mul=mul)
    ij = 5 * 24 + 7
    # exactly centred ensemble at ij (members +v, -v) so that the closed
    # forms hold to FP64 rounding
    st = p.state.copy(order="F")
    half = st[: k // 2, :, ij].copy()
    st[: k // 2, :, ij] = half
    st[k // 2:, :, ij] = -half
    p.state = st
    p.sprd = p.sprd.copy(order="F")
    p.sprd[:, ij] = np.sqrt((st[:, :, ij].astype(np.float64) ** 2).sum(0) / (k - 1))
    q = p.subset_obs(np.array([0]))
    q.olat = p.lat[ij:ij + 1].copy()
    q.olon = p.lon[ij:ij + 1].copy()
    q.hx = np.asfortranarray(p.state[:, 0, ij:ij + 1])
    q.hxmean = p.mean[0, ij:ij + 1].copy()
    q.oval = (q.hxmean + np.float32(d)).astype(np.float32)
    q.oerr = np.array([r], np.float32)
    return q, ij


def test_single_obs_kalman_gain():
    """Mean increment at the observed point = d sb^2/(sb^2 + R); analysis
    spread = sb sqrt(R/(sb^2+R)) (R = err, src/letkf_solver.F90:227)."""
    q, ij = _single_obs_problem()
    out = O.analysis(q, points=[ij])
    sb2 = float((q.state[:, 0, ij].astype(np.float64) ** 2).sum() / (q.k - 1))
    d = float(np.float32(q.oval[0] - q.hxmean[0]))
    r = 0.5
    inc = out["mean"][0, 0] - q.mean[0, ij]
    assert inc == pytest.approx(d * sb2 / (sb2 + r), rel=1e-12)
    assert out["sprd"][0, 0] == pytest.approx(math.sqrt(sb2 * r / (sb2 + r)), rel=1e-12)
    assert out["lg_m"][0, 0] == 1


@pytest.mark.parametrize("alpha", [0.6, 1.0])
def test_single_obs_rtps(alpha):
    q, ij = _single_obs_problem(rtps=alpha)
    out = O.analysis(q, points=[ij])
    sb = float(q.sprd[0, ij])
    sb2 = float((q.state[:, 0, ij].astype(np.float64) ** 2).sum() / (q.k - 1))
    sa = math.sqrt(sb2 * 0.5 / (sb2 + 0.5))
    assert out["sprd"][0, 0] == pytest.approx(alpha * sb + (1 - alpha) * sa, rel=1e-6)


def test_rtpp_one_restores_background_and_mul():
    q, ij = _single_obs_problem(rtpp=1.0)
    out = O.analysis(q, points=[ij])
    xb = q.state[:, 0, ij].astype(np.float64)
    assert np.abs((out["state"][:, 0, 0] - out["mean"][0, 0]) - xb).max() < 1e-12
    q2, _ = _single_obs_problem(mul=1.05)
    q3, _ = _single_obs_problem()
    a = O.analysis(q2, points=[ij])
    b = O.analysis(q3, points=[ij])
    assert a["sprd"][0, 0] == pytest.approx(float(np.float32(1.05)) * b["sprd"][0, 0], rel=1e-12)


def test_masked_and_unobserved_points():
    """Masked points return the background perturbations, mean and spread
    untouched (src/letkf_solver.F90:207); unobserved points return the full
    background members."""
    p = synth.random_problem(seed=31, nx=24, ny=12, nobs=40)
    out = O.analysis(p)
    for ij in range(p.nij):
        if p.mask[ij]:
            assert np.array_equal(out["state"][:, :, ij], p.state[:, :, ij].astype(np.float64))
            assert np.array_equal(out["mean"][:, ij], p.mean[:, ij].astype(np.float64))
        elif out["lg_m"][:, ij].sum() == 0:
            full = p.state[:, :, ij].astype(np.float64) + p.mean[:, ij].astype(np.float64)
            assert np.abs(out["state"][:, :, ij] - full).max() < 1e-5
            assert np.abs(out["mean"][:, ij] - p.mean[:, ij]).max() < 1e-5


def test_bounds_and_increment_clip():
    """simple3d bounds_check: ana_bounds [-2, 2] on state1, ana_inc_max 2
    on state2, 0.2 and [-2, 2] on state3 (test/simple3d/bounds_check.yaml)."""
    p = synth.simple3d("bounds_check")
    out = O.analysis(p)
    nlev = p.var_lev[0]
    inc = out["mean"] - p.mean
    s2 = slice(nlev, 2 * nlev)
    s3 = slice(2 * nlev, 3 * nlev)
    assert np.abs(inc[s2]).max() <= 2.0 + 1e-12
    # state3: the increment clip (0.2) acts first, then the member bounds
    # re-mean the ensemble (src/letkf_solver.F90:357-371), which can move
    # the mean again, so only the bounds are asserted there
    for s in list(range(0, nlev)) + list(range(2 * nlev, 3 * nlev)):
        full = out["state"][:, s, :]
        assert full.min() >= -2.0 - 1e-9 and full.max() <= 2.0 + 1e-9
    assert np.abs(inc).max() > 0.1            # the obs did something


def test_ocean_groups_split():
    """LOC_OCEAN with vtloc_surf bkg_t: the satellite obs (platid in
    sat_plats) only updates the mixed-layer group (letkf_loc_ocean.F90:483-493)."""
    p = synth.ocean_vtloc()
    out = O.analysis(p)
    assert out["lg_m"].shape[0] == 2
    ij_sat = int(np.argmax(out["lg_m"][0] - out["lg_m"][1]))
    assert out["lg_m"][0, ij_sat] > out["lg_m"][1, ij_sat]


def test_state_from_members_closed_forms():
    """src/letkf_state.F90:326-334 restated in FP32: integer-valued members
    give exact sums, so mean, perturbations and spread have closed forms; the
    layout is file order in, state_ij(k, ns, ij) out."""
    g = np.random.default_rng(21)
    nij, ns, k = 7, 3, 5
    mem = g.integers(-20, 20, size=(nij, ns, k)).astype(np.float32)
    st, mu, sp = O.state_from_members(mem)
    assert st.shape == (k, ns, nij) and mu.shape == (ns, nij) and sp.shape == (ns, nij)
    ref_mu = (mem.sum(axis=2) / k).astype(np.float32)
    assert np.array_equal(mu, ref_mu.T)
    pert = mem - ref_mu[:, :, None]
    assert np.array_equal(st, pert.transpose(2, 1, 0))
    ref_sp = np.sqrt((pert.astype(np.float64) ** 2).sum(axis=2) / (k - 1))
    assert np.allclose(sp, ref_sp.T, rtol=2e-7)
    # FP32 order: member by member, not pairwise
    big = np.full((1, 1, 3), [1e8, 1.0, -1e8], np.float32).reshape(1, 1, 3)
    assert O.state_from_members(big)[1][0, 0] == np.float32(0.0)


def test_obs_preprocess_qc_and_order():
    """src/letkf_obs.F90:459-489: the mean over members, perturbations,
    qc > 0 dropped, qc <= 0 (also negative) kept, input order kept."""
    hx = np.array([[1, 2, 3, 4], [3, 2, 5, 8]], np.float32)     # (k=2, n=4)
    qc = np.array([0, 1, -3, 0])
    keep, pert, mu = O.obs_preprocess(hx, qc)
    assert keep.tolist() == [0, 2, 3]
    assert mu.tolist() == [2.0, 4.0, 6.0]
    assert pert.tolist() == [[-1.0, -1.0, -2.0], [1.0, 1.0, 2.0]]
    assert O.obs_preprocess(hx, np.ones(4))[0].size == 0
