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


# ------------------------------------------------------------ LOC_OCEAN, VGAUSS
def _ocean(sat_obs, sat_plats, tloc_prof=-1.0, tloc_sat=-1.0, nlev=1, delta=0.2):
    return synth.Localizer(
        cls="LOC_OCEAN",
        hzloc_prof=LinearInterpLat([0.0, 90.0], [500e3, 500e3]),
        hzloc_sat=LinearInterpLat([0.0, 90.0], [300e3, 300e3]),
        tloc_prof=tloc_prof, tloc_sat=tloc_sat, vtloc_surf=1, bkg_t_delta=delta,
        bkg_t_var_idx=0, bkg_t_var_len=nlev, sat_obs=list(sat_obs),
        sat_plats=list(sat_plats))


# every (obsid, platid) pair of {1, 2} x {1, 2}, all at the gridpoint itself
_PAIRS = [(1, 1), (1, 2), (2, 1), (2, 2)]


@pytest.mark.parametrize("sat_obs,sat_plats,surface", [
    ([], [], {(1, 1), (1, 2), (2, 1), (2, 2)}),      # both lists empty: all surface
    ([1], [], {(1, 1), (1, 2)}),                      # obsid list only
    ([], [2], {(1, 2), (2, 2)}),                      # platid list only
    ([1], [2], {(1, 2)}),                             # both: AND, not OR
])
def test_ocean_surface_classification(sat_obs, sat_plats, surface):
    """letkf_loc_ocean.F90:467-479: an obs is surface when it matches every
    NON-EMPTY list (obsid in sat_obs AND platid in sat_plats); surface obs
    get weight 0 in the SUBML group (:485-487), all others keep GC(0) = 1
    there.  The SURF group keeps every obs at distance 0."""
    loc = _ocean(sat_obs, sat_plats)
    obs = [dict(obsid=a, platid=b) for a, b in _PAIRS]
    p = synth.point_problem(loc, obs)
    idx0, r0, _ = O.selection(p, 0, group=0)
    idx1, r1, _ = O.selection(p, 0, group=1)
    assert idx0.tolist() == [0, 1, 2, 3] and r0.tolist() == [1.0] * 4
    want1 = [i for i, pr in enumerate(_PAIRS) if pr not in surface]
    assert idx1.tolist() == want1
    assert r1.tolist() == [1.0] * len(want1)


def test_ocean_surface_and_profile_radii():
    """Surface obs use the sat radius, profile obs the prof radius
    (:489-497).  At 400 km, GC(400 km, L=300 km) and GC(400 km, L=500 km)
    against the truncated polynomial's value at z/c = d sqrt(0.3)/L."""
    loc = _ocean([1], [], nlev=1)
    obs = [dict(obsid=1, dist=400e3), dict(obsid=2, dist=400e3)]
    p = synth.point_problem(loc, obs)
    idx, r, d = O.selection(p, 0, group=0)
    assert idx.tolist() == [0, 1]
    dd = float(d[0])
    assert dd == pytest.approx(400e3, rel=1e-6)
    assert r[0] == O.gc(dd, 300e3) and r[1] == O.gc(dd, 500e3)
    assert r[0] == pytest.approx(_gc_np(np.array([dd]), 300e3)[0], rel=1e-6)
    assert r[1] == pytest.approx(_gc_np(np.array([dd]), 500e3)[0], rel=1e-6)
    assert r[0] < r[1]


@pytest.mark.parametrize("lat,want", [(9.0, 460e3), (81.0, 300e3), (-81.0, 300e3)])
def test_ocean_maxhz_is_max_of_tables(lat, want):
    """loc_ocean_maxhz (letkf_loc_ocean.F90:435-444) is the MAX of the prof
    and sat radii: prof falls 500 -> 100 km from 0 to 90 degrees (460 km at
    9 degrees, 140 km at 81), sat is 300 km everywhere.  The search radius
    is that times sqrt(40/3), so a profile obs at 1.5 * 460 km = 690 km is
    in the search set at every latitude here (MIN would give 511 km at
    81 degrees)."""
    loc = synth.Localizer(
        cls="LOC_OCEAN", hzloc_prof=LinearInterpLat([0.0, 90.0], [500e3, 100e3]),
        hzloc_sat=LinearInterpLat([0.0, 90.0], [300e3, 300e3]), vtloc_surf=0,
        sat_obs=[1], sat_plats=[])
    p = synth.point_problem(loc, [dict(obsid=2, dist=1.5 * 460e3)], lat=lat)
    out = O.analysis(p)
    assert out["maxhz"][0] == pytest.approx(want, rel=1e-6)
    assert out["obscount"][0] == 1


@pytest.mark.parametrize("which", ["sat", "prof"])
def test_ocean_time_localization(which):
    """GC(obs%time, tloc) multiplies the horizontal weight when tloc > 0
    (:490-491 surface, :496-497 profile).  Obs at the gridpoint (GC = 1) with
    time = z_c * tloc / sqrt(0.3) for z_c = 0, 0.5, 1, 1.5 give the known
    values 1, 0.684895, 0.20833, 0.016473; tloc <= 0 leaves weight 1."""
    T = 6.0
    c = T / math.sqrt(0.3)
    zc = [0.0, 0.5, 1.0, 1.5]
    want = [1.0, 0.684895, 0.20833, 0.016473]
    oid = 1 if which == "sat" else 2
    kw = dict(tloc_sat=T) if which == "sat" else dict(tloc_prof=T)
    obs = [dict(obsid=oid, time=float(np.float32(z * c))) for z in zc]
    p = synth.point_problem(_ocean([1], [], **kw), obs)
    _, r, _ = O.selection(p, 0, group=0)
    for got, v in zip(r, want):
        assert abs(got - v) <= 2e-6 * v + 5e-8, (got, v)
    # the other class's tloc is not applied, and negative times count by |t|
    kw2 = dict(tloc_prof=T) if which == "sat" else dict(tloc_sat=T)
    p2 = synth.point_problem(_ocean([1], [], **kw2), obs)
    assert O.selection(p2, 0, group=0)[1].tolist() == [1.0] * 4
    obs_neg = [dict(obsid=oid, time=-float(np.float32(0.5 * c)))]
    p3 = synth.point_problem(_ocean([1], [], **kw), obs_neg)
    assert abs(O.selection(p3, 0, group=0)[1][0] - 0.684895) < 2e-6


@pytest.mark.parametrize("profile,mld", [
    ([20.0, 20.1, 20.15, 19.5, 15.0, 10.0], 3),     # first |t0 - t(1+i)| > delta at i = 3
    ([20.0, 20.5, 20.5, 20.5, 20.5, 20.5], 1),      # at i = 1
    ([20.0, 20.1, 19.9, 20.15, 19.85, 20.0], 6),    # never: mldlvl = bkg_t_var_len
    ([20.0, 19.9, 19.7, 19.7, 19.7, 19.7], 2),      # |.| with a falling profile
])
def test_ocean_mixed_layer_level(profile, mld):
    """calc_mldlvl_bkgt (letkf_loc_ocean.F90:408-427): the first i in
    1..len-1 with |T(1) - T(1+i)| > bkg_t_delta, else len.  Levels j <= mld
    of every variable are SURF, the rest SUBML, and single-level variables
    are SURF (:375-391).  With one surface obs, only SURF slabs change: the
    SUBML slabs keep the background exactly (an empty group, :272)."""
    nlev = 6
    prof = list(profile) + [35.0] * nlev + [1.0]          # T, S (6 levels), SSH (1)
    p = synth.point_problem(_ocean([], [1], nlev=nlev), [dict(platid=1, dep=0.5)],
                            var_lev=(nlev, nlev, 1), mean_profile=prof)
    out = O.analysis(p)
    assert out["lg_m"][:, 0].tolist() == [1, 0]
    changed = np.abs(out["mean"][:, 0] - p.mean[:, 0]) > 1e-4   # increments ~0.1
    want = np.array([j < mld for j in range(nlev)] * 2 + [True])
    assert changed.tolist() == want.tolist()
    # a profile obs (platid 2) updates every slab
    p2 = synth.point_problem(_ocean([], [1], nlev=nlev), [dict(platid=2, dep=0.5)],
                             var_lev=(nlev, nlev, 1), mean_profile=prof)
    out2 = O.analysis(p2)
    assert out2["lg_m"][:, 0].tolist() == [1, 1]
    assert (np.abs(out2["mean"][:, 0] - p2.mean[:, 0]) > 1e-4).all()


def test_vgauss_single_obs_weight():
    """LOC_VGAUSS: GC(d) * exp(-(zdim - depth_l)^2 / (2 s^2)) per level
    group.  An obs at the gridpoint (GC = 1) and zdim 0 with levels at
    0, s, 2s, 3s gives exp(-l^2/2) = 1, 0.606531, 0.135335, 0.011109; at
    400 km the horizontal GC multiplies in."""
    s = 40.0
    depth = np.array([0.0, s, 2 * s, 3 * s], np.float32)
    loc = synth.Localizer(cls="LOC_VGAUSS",
                          hzloc=LinearInterpLat([0.0, 90.0], [500e3, 500e3]),
                          vt_scale=s, lev_depth=depth)
    p = synth.point_problem(loc, [dict(dist=0.0), dict(dist=400e3)], var_lev=(4,))
    for l in range(4):
        idx, r, d = O.selection(p, 0, group=l)
        want = math.exp(-l * l / 2.0)
        assert idx.tolist() == [0, 1]
        assert r[0] == pytest.approx(want, rel=1e-6)
        assert r[1] == pytest.approx(want * O.gc(float(d[1]), 500e3), rel=1e-6)


# ------------------------------------------------------------ increment clip
def test_increment_clip_single_obs_closed_form():
    """ana_inc_max (src/letkf_solver.F90:308-318) clips the MEAN increment
    only: with one obs at the gridpoint, the unclipped increment is
    d sb^2/(sb^2 + R); clipped it is +-inc_max exactly, while the analysis
    perturbations (and so the spread sb sqrt(R/(sb^2+R))) are untouched."""
    for dep in (0.8, -0.8):
        q, ij = _single_obs_problem(d=dep)
        free = O.analysis(q, points=[ij])
        sb2 = float((q.state[:, 0, ij].astype(np.float64) ** 2).sum() / (q.k - 1))
        d = float(np.float32(q.oval[0] - q.hxmean[0]))
        inc_free = free["mean"][0, 0] - q.mean[0, ij]
        assert inc_free == pytest.approx(d * sb2 / (sb2 + 0.5), rel=1e-12)
        cap = np.float32(0.1)
        assert abs(inc_free) > cap
        q.inc_max = np.array([cap], np.float32)
        out = O.analysis(q, points=[ij])
        inc = out["mean"][0, 0] - q.mean[0, ij]
        assert inc == pytest.approx(math.copysign(float(cap), dep), rel=1e-12)
        pa = out["state"][:, 0, 0] - out["mean"][0, 0]
        pa_free = free["state"][:, 0, 0] - free["mean"][0, 0]
        assert np.abs(pa - pa_free).max() < 1e-12
        assert out["sprd"][0, 0] == pytest.approx(math.sqrt(sb2 * 0.5 / (sb2 + 0.5)), rel=1e-12)


def test_bounds_on_mean_then_members_closed_form():
    """ana_bounds clip the mean increment to bounds - x_b (:314-315), then
    each member to bounds - x_a (:357-362) and re-mean (:364-367).  With the
    upper bound at x_b + 0.05 below the unclipped analysis mean, the mean
    becomes the bound, every member perturbation is clamped to <= 0, and the
    re-mean gives members = bound + min(a y_k, 0), a = sqrt(R/(sb^2 + R))."""
    q, ij = _single_obs_problem(d=0.8)
    y = q.state[:, 0, ij].astype(np.float64)
    sb2 = float((y ** 2).sum() / (q.k - 1))
    a = math.sqrt(0.5 / (sb2 + 0.5))
    hi = np.float32(q.mean[0, ij] + np.float32(0.05))
    q.bounds = np.asfortranarray(np.array([[-9.99e36], [hi]], np.float32))
    out = O.analysis(q, points=[ij])
    want = float(hi) + np.minimum(a * y, 0.0)
    assert np.abs(out["state"][:, 0, 0] - want).max() < 1e-12
    assert out["mean"][0, 0] == pytest.approx(want.mean(), rel=1e-12)
    assert out["state"][:, 0, 0].max() <= float(hi) + 1e-12


def test_bounds_check_clip_binds():
    """The rebuilt simple3d bounds_check: obs2's -5.0 departure on state2
    asks for more than ana_inc_max = 2 (and state3 for more than 0.2);
    with the clip, those gridpoints get exactly -2 (and +-0.2 before the
    state3 member bounds re-mean), and every other gridpoint is unchanged."""
    p = synth.simple3d("bounds_check")
    nlev = int(p.var_lev[0])
    s2 = slice(nlev, 2 * nlev)
    out = O.analysis(p)
    import copy
    q = copy.copy(p)
    q.inc_max = np.full(3, 9.99e36, np.float32)
    free = O.analysis(q)
    inc_free = free["mean"] - p.mean
    inc = out["mean"] - p.mean
    assert inc_free[s2].min() < -2.5                 # the clip must bind
    assert np.abs(inc_free[2 * nlev:]).max() > 0.4
    bind = np.abs(inc_free[s2]) > 2.0
    assert bind.sum() >= 3
    assert np.allclose(inc[s2][bind], 2.0 * np.sign(inc_free[s2][bind]), atol=1e-6)
    assert np.array_equal(inc[s2][~bind], inc_free[s2][~bind])
    # the perturbations of state2 are untouched by the clip
    pa = out["state"][:, s2] - out["mean"][None, s2]
    pf = free["state"][:, s2] - free["mean"][None, s2]
    assert np.abs(pa - pf).max() < 1e-9


# ------------------------------------------------------------ FP32 CPU baseline
def test_syev32_against_numpy():
    """The FP32 eigensolver of the CPU baseline (Householder + implicit QL)
    against numpy's FP64 eigh: eigenvalues to FP32 precision relative to the
    largest, orthonormal vectors, A reconstructed."""
    g = np.random.default_rng(8)
    for n in (2, 5, 20, 50, 100):
        x = g.standard_normal((n, 3 * n))
        a = x @ x.T + (n - 1) * np.eye(n)
        w, v = O.syev32(a)
        w0 = np.linalg.eigvalsh(a)
        assert np.all(np.diff(w) >= 0)
        assert np.abs(w - w0).max() <= 2e-6 * n * w0.max()
        v = v.astype(np.float64)
        assert np.abs(v.T @ v - np.eye(n)).max() < 1e-5 * n
        assert np.abs(v @ np.diag(w.astype(np.float64)) @ v.T - a).max() <= 2e-6 * n * w0.max()


@pytest.mark.parametrize("k,m", [(5, 1), (20, 7), (20, 40), (50, 300)])
def test_core32_against_fp64_core(k, m):
    """The FP32 core in the reference's order against the FP64 oracle core
    (itself pinned against numpy and closed forms): FP32-level agreement."""
    g = np.random.default_rng(k * 7 + m)
    y = g.standard_normal((k, m))
    y -= y.mean(0, keepdims=True)
    rdiag = g.uniform(0.2, 1.0, m)
    rloc = g.uniform(0.1, 1.0, m)
    dep = g.standard_normal(m)
    y32 = y.astype(np.float32)
    t64, _ = O.core_solve(y32.astype(np.float64), np.float32(rdiag), np.float32(rloc),
                          np.float32(dep))
    t32 = O.core_solve32(y32, rdiag, rloc, dep)
    assert np.abs(t32 - t64).max() <= 2e-5 * np.abs(t64).max()


def test_analysis_f32_matches_fp64_oracle():
    """The FP32 gridpoint loop selects the same observations as the FP64
    oracle and agrees with it to FP32 precision (config 1)."""
    p = synth.config1()
    g = np.random.default_rng(3)
    pts = np.sort(g.choice(np.nonzero(p.mask == 0)[0], 300, replace=False))
    a = O.analysis(p, points=pts)
    b = O.analysis(p, points=pts, f32=True)
    assert np.array_equal(a["lg_m"], b["lg_m"])
    assert np.abs(b["mean"] - a["mean"]).max() <= 1e-5 * np.abs(a["mean"]).max()
    assert np.abs(b["sprd"] - a["sprd"]).max() <= 1e-4 * np.abs(a["sprd"]).max()
    pa, pb = a["state"] - a["mean"][None], b["state"] - b["mean"][None]
    assert np.abs(pb - pa).max() <= 1e-4 * np.abs(pa).max()
