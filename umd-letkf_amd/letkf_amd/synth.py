"""Seeded synthetic analyses in the reference's in-memory layout.

All random numbers come from numpy's counter-based Philox generator
(Salmon et al., SC'11), so a (config, seed) pair fixes every value.

The generator builds an ensemble of full fields, then reduces it exactly
like the reference's ingest:

* state: mean = SUM/k, perturbation = member - mean, spread with the (k-1)
  denominator, all in FP32 (src/letkf_state.F90:326-334);
* obs: hx = model equivalent of every member, obs_hx_mean = SUM/k,
  obs_hx = hx - mean (src/letkf_obs.F90:459-466).  Observation errors are
  written so that R = err holds under the reference's convention
  (src/letkf_solver.F90:227): err = sigma_o**2.

Configurations (sizes scale down for tests with ``scale``):

* ``config1``: 4-degree 90x45 SST, k=20, 5,000 random ocean obs, GC 500 km,
  RTPS 0.6, exact-Cholesky Gaussian-correlated perturbations (800 km).
* ``config2``: 1-degree 360x180 x 50 levels T/S, k=40, 4,000 profiles x 50
  depths, latitude-dependent radius 700 -> 200 km, vertical Gaussian 300 m
  (LOC_VGAUSS).
* ``config3``: 0.25-degree 1440x720 surface, k=50, 2M obs on 30 swath arcs
  with lognormal along-track density, 300 km cut-off radius (hzloc
  L = 300 km / sqrt(40/3) = 82.2 km).  With 2M obs the mean local count is
  N_obs x (keep-cap area) / (sphere area); a 300 km sigma-equivalent L would
  give ~12,600 on average, against the stated 0..~5,000 range, which the
  cut-off reading reproduces (mean ~1,000, 99th percentile ~4,300).
* ``simple3d``: the reference's two-observation test cases
  (test/simple3d/*.yaml) on a synthetic 3-variable grid.
"""
from __future__ import annotations

import math

import numpy as np

from .problem import LinearInterpLat, Localizer, Problem

RE = 6371000.0


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(int(seed)))


# ---------------------------------------------------------------- geometry
def grid(nx: int, ny: int):
    """Cell-centred regular lat-lon grid: lon [0,360), lat (-90,90)."""
    dlon, dlat = 360.0 / nx, 180.0 / ny
    lon = (np.arange(nx) + 0.5) * dlon
    lat = -90.0 + (np.arange(ny) + 0.5) * dlat
    return lon, lat


def to_xyz(lat_deg, lon_deg):
    la, lo = np.radians(lat_deg), np.radians(lon_deg)
    return np.stack([np.cos(la) * np.cos(lo), np.cos(la) * np.sin(lo), np.sin(la)], -1)


def land_mask(lat2d, lon2d):
    """Analytic land: six spherical caps plus the Antarctic cap (|lat|>70 S).
    True = land (masked)."""
    caps = [(45.0, 100.0, 3000e3), (10.0, 20.0, 2500e3), (40.0, 260.0, 2500e3),
            (-15.0, 300.0, 2000e3), (-25.0, 135.0, 1800e3), (70.0, 320.0, 1200e3)]
    p = to_xyz(lat2d, lon2d)
    land = lat2d < -70.0
    for la, lo, r in caps:
        c = to_xyz(np.array(la), np.array(lo))
        ang = np.arccos(np.clip(p @ c, -1, 1))
        land |= ang * RE < r
    return land


# ---------------------------------------------------------------- fields
def gaussian_cholesky_perts(g, lat1d_pts, lon1d_pts, nsamp, sigma, L):
    """Exact Gaussian-correlated draws: C = sigma^2 exp(-d^2/(2L^2))."""
    p = to_xyz(lat1d_pts, lon1d_pts)
    d = RE * np.arccos(np.clip(p @ p.T, -1.0, 1.0))
    cov = sigma ** 2 * np.exp(-d ** 2 / (2.0 * L ** 2))
    n = cov.shape[0]
    try:
        import scipy.linalg as sl
        chol = sl.cholesky(cov + 1e-6 * sigma ** 2 * np.eye(n), lower=True)
    except Exception:  # fall back to a clipped eigen square root
        w, v = np.linalg.eigh(cov)
        chol = v * np.sqrt(np.clip(w, 0.0, None))
    z = g.standard_normal((n, nsamp))
    return chol @ z        # (npts, nsamp)


def smooth_perts(g, nsamp, lat, lon, L, sigma):
    """Separable Gaussian smoothing of white noise on a lat-lon grid.

    Meridional: Gaussian kernel of scale L along latitude.  Zonal: per-row
    circular Gaussian of scale L / (RE cos(lat)).  Every gridpoint is then
    rescaled to standard deviation ``sigma`` over the draws' row.
    Returns (nsamp, ny, nx) float64.
    """
    ny, nx = lat.size, lon.size
    noise = g.standard_normal((nsamp, ny, nx))
    ylat = np.radians(lat) * RE
    ky = np.exp(-0.5 * ((ylat[:, None] - ylat[None, :]) / L) ** 2)
    ky /= np.sqrt((ky ** 2).sum(1, keepdims=True))
    f = np.matmul(ky[None], noise)
    dx = np.radians(360.0 / nx) * RE * np.maximum(np.cos(np.radians(lat)), 1e-3)
    j = np.arange(nx)
    jj = np.minimum(j, nx - j)[None, :] * dx[:, None]       # (ny, nx)
    kx = np.exp(-0.5 * (jj / L) ** 2)
    kx /= np.sqrt((kx ** 2).sum(1, keepdims=True))
    fk = np.fft.rfft(kx, axis=-1)
    f = np.fft.irfft(np.fft.rfft(f, axis=-1) * fk[None], n=nx, axis=-1)
    sd = f.std(axis=(0, 2), keepdims=True)
    return sigma * f / np.where(sd > 0, sd, 1.0)


def bilinear(field, lat, lon, olat, olon):
    """Bilinear interpolation of field(..., ny, nx) at points; lon periodic,
    lat clamped to the first/last row.  Returns (..., npts)."""
    ny, nx = lat.size, lon.size
    dlon = 360.0 / nx
    dlat = lat[1] - lat[0]
    fx = (np.mod(olon, 360.0) - lon[0]) / dlon
    i0 = np.floor(fx).astype(np.int64)
    wx = fx - i0
    i0 %= nx
    i1 = (i0 + 1) % nx
    fy = np.clip((olat - lat[0]) / dlat, 0.0, ny - 1.0)
    j0 = np.minimum(np.floor(fy).astype(np.int64), ny - 2)
    wy = fy - j0
    j1 = j0 + 1
    f00 = field[..., j0, i0]
    f01 = field[..., j0, i1]
    f10 = field[..., j1, i0]
    f11 = field[..., j1, i1]
    return ((1 - wy) * ((1 - wx) * f00 + wx * f01) + wy * ((1 - wx) * f10 + wx * f11))


# ---------------------------------------------------------------- assembly
def ensemble_to_state(members):
    """members (k, ns, nij) float32 full fields -> reference background."""
    m = members.astype(np.float32)
    k = m.shape[0]
    mean = (m.sum(0, dtype=np.float32) / np.float32(k)).astype(np.float32)
    pert = (m - mean[None]).astype(np.float32)
    sprd = np.sqrt((pert * pert).sum(0, dtype=np.float32) / np.float32(k - 1)).astype(np.float32)
    state = np.asfortranarray(np.transpose(pert, (0, 1, 2)))
    return state, np.asfortranarray(mean), np.asfortranarray(sprd)


def obs_from_hx(hx_full):
    """hx (k, nobs) full model equivalents -> (perturbations, mean)."""
    h = hx_full.astype(np.float32)
    k = h.shape[0]
    mean = (h.sum(0, dtype=np.float32) / np.float32(k)).astype(np.float32)
    pert = np.asfortranarray((h - mean[None]).astype(np.float32))
    return pert, mean


def make_problem(members, lat_ij, lon_ij, mask, olat, olon, ozdim, otime,
                 oval, oerr, obsid, platid, hx_full, var_s, var_lev, loc,
                 rtps=0.0, rtpp=0.0, mul=1.0, inc_max=None, bounds=None,
                 meta=None) -> Problem:
    state, mean, sprd = ensemble_to_state(members)
    hx, hxmean = obs_from_hx(hx_full)
    k, ns, nij = state.shape
    nv = len(var_s)
    if inc_max is None:
        inc_max = np.full(nv, 9.99e36, np.float32)
    if bounds is None:
        bounds = np.array([[-9.99e36] * nv, [9.99e36] * nv], np.float32)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    p = Problem(k=k, ns=ns, nij=nij, state=state, mean=mean, sprd=sprd,
                lat=f32(lat_ij), lon=f32(lon_ij),
                mask=np.ascontiguousarray(mask, dtype=np.int32),
                olat=f32(olat), olon=f32(olon), ozdim=f32(ozdim), otime=f32(otime),
                oval=f32(oval), oerr=f32(oerr),
                obsid=np.ascontiguousarray(obsid, dtype=np.int32),
                platid=np.ascontiguousarray(platid, dtype=np.int32),
                hx=hx, hxmean=hxmean,
                var_s=np.ascontiguousarray(var_s, dtype=np.int32),
                var_lev=np.ascontiguousarray(var_lev, dtype=np.int32),
                inc_max=f32(inc_max), bounds=np.asfortranarray(bounds, dtype=np.float32),
                loc=loc, rtps=float(rtps), rtpp=float(rtpp), mul=float(mul),
                meta=meta or {})
    p.validate()
    return p


def flat_grid(nx, ny):
    lon, lat = grid(nx, ny)
    lon2, lat2 = np.meshgrid(lon, lat)          # (ny, nx), lon fastest
    return lon, lat, lon2.ravel(), lat2.ravel()


def uniform_sphere(g, n):
    lat = np.degrees(np.arcsin(g.uniform(-1.0, 1.0, n)))
    lon = g.uniform(0.0, 360.0, n)
    return lat, lon


def novrt(radius):
    tab = LinearInterpLat([0.0, 90.0], [radius, radius])
    return Localizer(cls="LOC_NOVRT", hzloc=tab)


# ---------------------------------------------------------------- config 1
def config1(seed=0x1E7F0001, nx=90, ny=45, k=20, nobs=5000, with_mask=True):
    g = rng(seed)
    lon, lat, lon_ij, lat_ij = flat_grid(nx, ny)
    mean = 15.0 + 10.0 * np.cos(np.radians(lat_ij)) + 2.0 * np.sin(2.0 * np.radians(lon_ij))
    d = gaussian_cholesky_perts(g, lat_ij, lon_ij, k + 1, 0.5, 800e3)
    members = (mean[:, None] + d[:, :k]).T[:, None, :]          # (k, 1, nij)
    truth = mean + d[:, k]
    mask = land_mask(lat_ij, lon_ij) if with_mask else np.zeros(lat_ij.size, bool)
    olat, olon = np.zeros(0), np.zeros(0)
    while olat.size < nobs:                                     # rejection sample ocean
        a, b = uniform_sphere(g, 2 * nobs)
        keep = ~land_mask(a, b) if with_mask else np.ones(a.size, bool)
        olat = np.concatenate([olat, a[keep]])
        olon = np.concatenate([olon, b[keep]])
    olat, olon = olat[:nobs], olon[:nobs]
    f2 = members[:, 0, :].reshape(k, ny, nx)
    hx = bilinear(f2, lat, lon, olat, olon)
    ht = bilinear(truth.reshape(ny, nx), lat, lon, olat, olon)
    val = ht + g.normal(0.0, 0.3, nobs)
    z = np.zeros(nobs)
    return make_problem(members, lat_ij, lon_ij, mask, olat, olon, z, z, val,
                        np.full(nobs, 0.09), np.ones(nobs), np.ones(nobs), hx,
                        [0], [1], novrt(500e3), rtps=0.6,
                        meta=dict(config=1, nx=nx, ny=ny, seed=seed))


# ---------------------------------------------------------------- config 3
def swath_points(g, nobs, ntracks=30, width=100e3, arc=(15000e3, 30000e3),
                 sigma_ln=0.6, seg=50e3):
    """Observation positions on great-circle arcs with lognormal
    along-track density.  Returns (lat, lon) in degrees."""
    segs = []
    for _ in range(ntracks):
        p0 = g.standard_normal(3)
        p0 /= np.linalg.norm(p0)
        t0 = np.cross(p0, g.standard_normal(3))
        t0 /= np.linalg.norm(t0)
        L = g.uniform(*arc)
        ns = max(1, int(L / seg))
        w = np.exp(g.normal(0.0, sigma_ln, ns))
        segs.append((p0, t0, L / ns, w))
    weights = np.concatenate([s[3] for s in segs])
    owner = np.concatenate([np.full(s[3].size, i) for i, s in enumerate(segs)])
    sidx = np.concatenate([np.arange(s[3].size) for s in segs])
    pick = g.choice(weights.size, size=nobs, p=weights / weights.sum())
    out = np.zeros((nobs, 3))
    u = g.uniform(0.0, 1.0, nobs)
    cr = g.uniform(-0.5 * width, 0.5 * width, nobs)
    for i, (p0, t0, sl, _) in enumerate(segs):
        sel = np.nonzero(owner[pick] == i)[0]
        if sel.size == 0:
            continue
        s = (sidx[pick[sel]] + u[sel]) * sl / RE
        p = np.cos(s)[:, None] * p0 + np.sin(s)[:, None] * t0
        t = -np.sin(s)[:, None] * p0 + np.cos(s)[:, None] * t0
        n = np.cross(p, t)
        c = (cr[sel] / RE)[:, None]
        out[sel] = np.cos(c) * p + np.sin(c) * n
    lat = np.degrees(np.arcsin(np.clip(out[:, 2], -1, 1)))
    lon = np.mod(np.degrees(np.arctan2(out[:, 1], out[:, 0])), 360.0)
    return lat, lon


# 300 km read as the cut-off (search) radius: hzloc L = 300 km / sqrt(40/3)
C3_RADIUS = 300e3 / 3.6514837741851807


def config3(seed=0x1E7F0003, nx=1440, ny=720, k=50, nobs=2_000_000,
            radius=C3_RADIUS, corr=500e3, sigma=0.5, sigma_o=0.25, rtps=0.6,
            **swath):
    g = rng(seed)
    lon, lat, lon_ij, lat_ij = flat_grid(nx, ny)
    mean = 15.0 + 10.0 * np.cos(np.radians(lat_ij)) + 2.0 * np.sin(2.0 * np.radians(lon_ij))
    d = smooth_perts(g, k + 1, lat, lon, corr, sigma)            # (k+1, ny, nx)
    members = (mean.reshape(ny, nx)[None] + d[:k]).astype(np.float32)
    truth = mean.reshape(ny, nx) + d[k]
    del d
    olat, olon = swath_points(g, nobs, **swath)
    hx = bilinear(members, lat, lon, olat, olon)                 # (k, nobs)
    ht = bilinear(truth, lat, lon, olat, olon)
    val = ht + g.normal(0.0, sigma_o, nobs)
    z = np.zeros(nobs)
    members = members.reshape(k, 1, ny * nx)
    return make_problem(members, lat_ij, lon_ij, np.zeros(ny * nx, bool), olat, olon,
                        z, z, val, np.full(nobs, sigma_o ** 2), np.ones(nobs),
                        np.ones(nobs), hx, [0], [1], novrt(radius), rtps=rtps,
                        meta=dict(config=3, nx=nx, ny=ny, seed=seed,
                                  truth=truth.astype(np.float32)))


def config5_cycle(prev, members_fo, cycle, seed=0x1E7F0005, nobs=(1_000_000, 2_000_000),
                  sigma_noise=0.1, sigma_o=0.25, **swath):
    """Cycle ``cycle`` (>= 1) of the cycled 0.25-degree analysis (config 5).

    The background is the previous analysis ``members_fo`` (nij, 1, k) in
    file order, advected by one gridpoint eastward plus N(0, sigma_noise^2)
    per member and gridpoint; the truth moves one gridpoint with it.  A
    fresh swath set (its own seed, so a different track phase) of between
    nobs[0] and nobs[1] observations samples the new truth.  ``prev`` is the
    previous cycle's Problem (grid, localizer and settings are kept)."""
    g = rng(seed + 7919 * cycle)
    nx, ny = prev.meta["nx"], prev.meta["ny"]
    k = prev.k
    lon, lat, lon_ij, lat_ij = flat_grid(nx, ny)
    mem = np.asarray(members_fo, np.float32).reshape(ny, nx, k)
    mem = np.roll(mem, 1, axis=1) + g.normal(0.0, sigma_noise, mem.shape).astype(np.float32)
    members = np.ascontiguousarray(mem.transpose(2, 0, 1))           # (k, ny, nx)
    truth = np.roll(prev.meta["truth"], 1, axis=1)
    n = int(g.integers(nobs[0], nobs[1] + 1))
    olat, olon = swath_points(g, n, **swath)
    hx = bilinear(members, lat, lon, olat, olon)
    val = bilinear(truth, lat, lon, olat, olon) + g.normal(0.0, sigma_o, n)
    z = np.zeros(n)
    return make_problem(members.reshape(k, 1, ny * nx), lat_ij, lon_ij, prev.mask != 0,
                        olat, olon, z, z, val, np.full(n, sigma_o ** 2), np.ones(n),
                        np.ones(n), hx, [0], [1], prev.loc, rtps=prev.rtps,
                        meta=dict(config=5, cycle=cycle, nx=nx, ny=ny, seed=seed,
                                  truth=truth))


# ---------------------------------------------------------------- config 2
def config2(seed=0x1E7F0002, nx=360, ny=180, nlev=50, k=40, nprof=4000,
            hcorr=500e3, vcorr=200.0, vt_scale=300.0, rtps=0.0):
    g = rng(seed)
    lon, lat, lon_ij, lat_ij = flat_grid(nx, ny)
    depth = 5000.0 * (np.arange(nlev) / max(nlev - 1, 1)) ** 2
    mask = land_mask(lat_ij, lon_ij)
    kz = np.exp(-0.5 * ((depth[:, None] - depth[None, :]) / vcorr) ** 2)
    kz /= np.sqrt((kz ** 2).sum(1, keepdims=True))
    nij = nx * ny
    members = np.zeros((k, 2 * nlev, nij), np.float32)
    tmean = (2.0 + 18.0 * np.exp(-depth / 800.0))[:, None] * \
        (0.6 + 0.4 * np.cos(np.radians(lat_ij)))[None, :]
    smean = 34.5 + 0.5 * np.exp(-depth / 1000.0)[:, None] * np.cos(np.radians(lat_ij))[None, :]
    for v, (mfield, sg) in enumerate(((tmean, 0.5), (smean, 0.05))):
        d = smooth_perts(g, k * nlev, lat, lon, hcorr, 1.0).reshape(k, nlev, nij)
        d = np.einsum("ab,kbn->kan", kz, d)
        d *= sg / np.maximum(d.std(axis=(0, 2), keepdims=True), 1e-12)
        members[:, v * nlev:(v + 1) * nlev, :] = (mfield[None] + d).astype(np.float32)
        del d
    truth = members.mean(0) + (members[0] - members.mean(0))  # one member as truth
    olat, olon = np.zeros(0), np.zeros(0)
    while olat.size < nprof:
        a, b = uniform_sphere(g, 2 * nprof)
        keep = ~land_mask(a, b)
        olat, olon = np.concatenate([olat, a[keep]]), np.concatenate([olon, b[keep]])
    olat, olon = olat[:nprof], olon[:nprof]
    var = np.arange(nprof) % 2                       # even profiles T, odd S
    plat = np.repeat(olat, nlev)
    plon = np.repeat(olon, nlev)
    pvar = np.repeat(var, nlev)
    plev = np.tile(np.arange(nlev), nprof)
    slab = pvar * nlev + plev
    hx = np.zeros((k, plat.size), np.float64)
    ht = np.zeros(plat.size)
    for s in np.unique(slab):
        sel = slab == s
        hx[:, sel] = bilinear(members[:, s, :].reshape(k, ny, nx), lat, lon, plat[sel], plon[sel])
        ht[sel] = bilinear(truth[s].reshape(ny, nx), lat, lon, plat[sel], plon[sel])
    so = np.where(pvar == 0, 0.2, 0.02)
    val = ht + g.normal(0.0, 1.0, plat.size) * so
    tab = LinearInterpLat([0.0, 90.0], [700e3, 200e3])
    loc = Localizer(cls="LOC_VGAUSS", hzloc=tab, vt_scale=vt_scale,
                    lev_depth=depth.astype(np.float32))
    return make_problem(members, lat_ij, lon_ij, mask, plat, plon, depth[plev],
                        np.zeros(plat.size), val, so ** 2, pvar + 1, np.ones(plat.size),
                        hx, [0, nlev], [nlev, nlev], loc, rtps=rtps,
                        meta=dict(config=2, nx=nx, ny=ny, nlev=nlev, seed=seed))


# ---------------------------------------------------------------- config 4
C4_ERR = (0.2, 0.02, 0.05, 0.05)           # T (K), S (psu), u, v (m/s)
C4_SIGMA = (0.5, 0.05, 0.1, 0.1)           # background perturbation scale
# C2's radii read as cut-off radii, as for C3: L = radius / sqrt(40/3)
C4_L = (700e3 / 3.6514837741851807, 200e3 / 3.6514837741851807)


def config4(seed=0x1E7F0004, nx=720, ny=360, nlev=75, k=100, nprof=13000, nsurf=25000,
            modes=32, hcorr=500e3, rtps=0.6, chunk=4096):
    """Config 4: 0.5-degree 720x360 grid, 75 levels, T, S, u, v (ns = 300),
    k = 100, FP32 state (31.1 GB at full size), 13,000 profiles x 75 levels
    (T or S) plus 25,000 surface u or v points (1M obs), per-variable errors
    0.2 K, 0.02 psu, 0.05 m/s, LOC_NOVRT with 700 km (equator) to 200 km
    (poles) cut-off radii, RTPS 0.6.

    Perturbations are low rank and separable, so that the 31 GB background
    is written once, straight into the reference layout: member e of
    variable v at level l is sigma_v sum_r C_v[e, r] Z_v[r, l] F[r, :] with
    `modes` smooth horizontal fields F (Gaussian correlation hcorr), smooth
    random vertical modes Z_v and N(0, 1) coefficients C_v.  The background
    is mean = base + mean over members, perturbations = members - mean, in
    FP32.  hx is the bilinear interpolation of each member (base + the same
    modes interpolated, since interpolation is linear); the truth is one
    more coefficient draw; val = H(truth) + N(0, err_v^2)."""
    g = rng(seed)
    lon, lat, lon_ij, lat_ij = flat_grid(nx, ny)
    nij = nx * ny
    ns = 4 * nlev
    depth = 5000.0 * (np.arange(nlev) / max(nlev - 1, 1)) ** 2
    mask = land_mask(lat_ij, lon_ij)
    cl = np.cos(np.radians(lat_ij))
    F = smooth_perts(g, modes, lat, lon, hcorr, 1.0).reshape(modes, nij).astype(np.float32)
    # vertical modes: Gaussian bumps of random centre, width and sign,
    # normalised to unit variance at every level
    Z = np.zeros((4, modes, nlev))
    for v in range(4):
        z0 = g.uniform(0.0, 5000.0, modes)
        wz = g.uniform(150.0, 1500.0, modes)
        Z[v] = np.exp(-0.5 * ((depth[None, :] - z0[:, None]) / wz[:, None]) ** 2)
        Z[v] *= g.choice([-1.0, 1.0], modes)[:, None]
        Z[v] /= np.sqrt((Z[v] ** 2).sum(0, keepdims=True))
    Cv = g.standard_normal((4, k + 1, modes))          # member k is the truth
    sig = np.array(C4_SIGMA)

    def base(v, lev, ij):
        z = depth[lev][:, None]
        c = cl[ij][None, :] if np.ndim(ij) else cl[ij]
        if v == 0:
            return (2.0 + 18.0 * np.exp(-z / 800.0)) * (0.6 + 0.4 * c)
        if v == 1:
            return 34.5 + 0.5 * np.exp(-z / 1000.0) * c
        if v == 2:
            return 0.2 * c * np.exp(-z / 500.0)
        return 0.05 * np.sin(2.0 * np.radians(lon_ij[ij]))[None, :] * np.exp(-z / 500.0)

    state = np.empty((k, ns, nij), np.float32, order="F")
    mean = np.empty((ns, nij), np.float32, order="F")
    sprd = np.empty((ns, nij), np.float32, order="F")
    lev = np.arange(nlev)
    Ct = [np.ascontiguousarray(Cv[v, :k].T * sig[v], np.float32) for v in range(4)]  # (modes, k)
    Zt = [np.ascontiguousarray(Z[v].T, np.float32) for v in range(4)]                 # (nlev, modes)
    for i0 in range(0, nij, chunk):
        ij = np.arange(i0, min(nij, i0 + chunk))
        n = ij.size
        Fi = np.ascontiguousarray(F[:, ij].T)                                        # (n, modes)
        for v in range(4):
            # P[i, l, e]: member fastest, the order of the state's columns
            H = Fi[:, None, :] * Zt[v][None, :, :]                                   # (n, nlev, modes)
            P = (H.reshape(n * nlev, modes) @ Ct[v]).reshape(n, nlev, k)
            pm = P.sum(2, dtype=np.float32) / np.float32(k)                          # (n, nlev)
            P -= pm[:, :, None]
            sl = slice(v * nlev, (v + 1) * nlev)
            state[:, sl, i0:i0 + n] = P.transpose(2, 1, 0)
            mean[sl, i0:i0 + n] = (base(v, lev, ij) + pm.T).astype(np.float32)
            sprd[sl, i0:i0 + n] = np.sqrt(np.einsum("ile,ile->li", P, P) / np.float32(k - 1))
    # observations: profiles (T for even, S for odd) and surface u / v
    def ocean_points(n):
        a, b = np.zeros(0), np.zeros(0)
        while a.size < n:
            x, y = uniform_sphere(g, 2 * n)
            keep = ~land_mask(x, y)
            a, b = np.concatenate([a, x[keep]]), np.concatenate([b, y[keep]])
        return a[:n], b[:n]
    plat0, plon0 = ocean_points(nprof)
    slat, slon = ocean_points(nsurf)
    pvar = np.concatenate([np.repeat(np.arange(nprof) % 2, nlev), 2 + np.arange(nsurf) % 2])
    plev = np.concatenate([np.tile(lev, nprof), np.zeros(nsurf, np.int64)])
    olat = np.concatenate([np.repeat(plat0, nlev), slat])
    olon = np.concatenate([np.repeat(plon0, nlev), slon])
    nobs = olat.size
    HF = bilinear(F.reshape(modes, ny, nx).astype(np.float64), lat, lon, olat, olon)  # (modes, nobs)
    hx = np.zeros((k, nobs))
    ht = np.zeros(nobs)
    for v in range(4):
        sel = np.nonzero(pvar == v)[0]
        if sel.size == 0:
            continue
        Hs = Z[v][:, plev[sel]] * HF[:, sel]                       # (modes, nsel)
        dv = sig[v] * (Cv[v] @ Hs)                                 # (k + 1, nsel)
        hb = np.zeros(sel.size)
        for l in np.unique(plev[sel]):
            q = sel[plev[sel] == l]
            bf = base(v, np.array([l]), np.arange(nij)).reshape(ny, nx)
            hb[plev[sel] == l] = bilinear(bf, lat, lon, olat[q], olon[q])
        hx[:, sel] = hb[None] + dv[:k]
        ht[sel] = hb + dv[k]
    err = np.array(C4_ERR)[pvar]
    val = ht + g.normal(0.0, 1.0, nobs) * err
    hxp, hxm = obs_from_hx(hx)
    loc = Localizer(cls="LOC_NOVRT", hzloc=LinearInterpLat([0.0, 90.0], list(C4_L)))
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    p = Problem(k=k, ns=ns, nij=nij, state=state, mean=mean, sprd=sprd,
                lat=f32(lat_ij), lon=f32(lon_ij), mask=np.ascontiguousarray(mask, np.int32),
                olat=f32(olat), olon=f32(olon), ozdim=f32(depth[plev]), otime=f32(np.zeros(nobs)),
                oval=f32(val), oerr=f32(err ** 2),
                obsid=np.ascontiguousarray(pvar + 1, np.int32),
                platid=np.ascontiguousarray(np.where(pvar < 2, 1, 2), np.int32),
                hx=hxp, hxmean=hxm,
                var_s=np.ascontiguousarray([0, nlev, 2 * nlev, 3 * nlev], np.int32),
                var_lev=np.ascontiguousarray([nlev] * 4, np.int32),
                inc_max=f32(np.full(4, 9.99e36)),
                bounds=np.asfortranarray(np.array([[-9.99e36] * 4, [9.99e36] * 4]), np.float32),
                loc=loc, rtps=float(rtps), rtpp=0.0, mul=1.0,
                meta=dict(config=4, nx=nx, ny=ny, nlev=nlev, seed=seed))
    p.validate()
    return p


# ---------------------------------------------------------------- simple3d
SIMPLE3D_OBS = {
    # ob_type, plat_type, state_var, lat, lon, depth, time, val (O-F), err
    "default": [(1, 1, 0, 40.5, 220.5, 1, 0, 1.0, 1.0), (2, 1, 1, 1.5, 220.5, 1, 10, -1.0, 1.0)],
    "bounds_check": [(1, 1, 0, 40.5, 220.5, 1, 0, 5.0, 1.0), (2, 1, 1, 1.5, 220.5, 1, 10, -5.0, 1.0)],
    "inflation_rtps": [(1, 1, 0, 40.5, 220.5, 1, 0, 1.0, 1.0), (2, 1, 1, -1.5, 220.5, 1, 0, -1.0, 1.0)],
    "inflation_rtpp": [(1, 1, 0, 40.5, 220.5, 1, 0, 1.0, 1.0), (2, 1, 1, -1.5, 220.5, 1, 0, -1.0, 1.0)],
    "inflation_mul": [(1, 1, 0, 40.5, 220.5, 1, 0, 1.0, 1.0), (2, 1, 1, -1.5, 220.5, 1, 0, -1.0, 1.0)],
}


def simple3d(case="default", seed=0x1E7F0100, nx=72, ny=36, nlev=4, k=50):
    """test/simple3d/<case>.yaml on a synthetic 5-degree grid: 3 variables,
    OBSIO_TEST semantics (hx = nearest gridpoint, val = O-F + mean(hx);
    src/subclass.letkf_obsio/letkf_obsio_test.F90:146-220), LOC_NOVRT with
    500 km at the equator and 50 km at the pole."""
    g = rng(seed)
    lon, lat, lon_ij, lat_ij = flat_grid(nx, ny)
    nij = nx * ny
    depth = np.arange(nlev, dtype=float) * 10.0 + 1.0
    members = np.zeros((k, 3 * nlev, nij), np.float32)
    # bounds_check: state2 and state3 get a 2.0 spread, and state3 shares
    # 0.6 of state2's perturbation, so that obs2's -5.0 departure asks for
    # increments beyond ana_inc_max (2 on state2, 0.2 on state3) and the
    # clip of src/letkf_solver.F90:308-318 binds
    sig = [0.5, 2.0, 2.0] if case == "bounds_check" else [0.5, 0.5, 0.5]
    pv = []
    for v in range(3):
        d = smooth_perts(g, k * nlev, lat, lon, 1000e3, sig[v]).reshape(k, nlev, nij)
        if case == "bounds_check" and v == 2:
            d = 0.6 * pv[1] + 0.8 * d
        pv.append(d)
        members[:, v * nlev:(v + 1) * nlev] = (float(v) + d).astype(np.float32)
    state, mean, _ = ensemble_to_state(members)
    obs = SIMPLE3D_OBS[case]
    pts = to_xyz(lat_ij, lon_ij)
    cols = []
    for (oid, pid, var, la, lo, dep, tm, inc, err) in obs:
        j = int(np.argmax(pts @ to_xyz(np.array(la), np.array(lo))))
        lvl = int(np.argmin(np.abs(depth - dep)))
        s = var * nlev + lvl
        cols.append(members[:, s, j].astype(np.float32))
    hx = np.stack(cols, 1)
    hxm = hx.sum(0, dtype=np.float32) / np.float32(k)
    val = np.array([o[7] for o in obs], np.float32) + hxm
    rtps = rtpp = 0.0
    mul = 1.0
    inc_max = np.full(3, 9.99e36, np.float32)
    bounds = np.array([[-9.99e36] * 3, [9.99e36] * 3], np.float32)
    if case == "bounds_check":
        bounds[:, 0] = (-2, 2)
        inc_max[1] = 2.0
        inc_max[2] = 0.2
        bounds[:, 2] = (-2, 2)
    elif case == "inflation_rtps":
        rtps = 0.8
    elif case == "inflation_rtpp":
        rtpp = 0.8
    elif case == "inflation_mul":
        mul = 1.05
    loc = Localizer(cls="LOC_NOVRT", hzloc=LinearInterpLat([0.0, 90.0], [500e3, 50e3]))
    return make_problem(members, lat_ij, lon_ij, np.zeros(nij, bool),
                        [o[3] for o in obs], [o[4] for o in obs], [o[5] for o in obs],
                        [o[6] for o in obs], val, [o[8] for o in obs],
                        [o[0] for o in obs], [o[1] for o in obs], hx,
                        [0, nlev, 2 * nlev], [nlev] * 3, loc, rtps=rtps, rtpp=rtpp,
                        mul=mul, inc_max=inc_max, bounds=bounds,
                        meta=dict(case=case, seed=seed))


# ---------------------------------------------------------------- ocean
def ocean_vtloc(seed=0x1E7F0200, nx=72, ny=36, nlev=6, k=10):
    """test/ocean/vtloc.yaml pattern: LOC_OCEAN with a bkg_t mixed-layer
    split, one satellite obs and four profile obs (OBSIO_TEST semantics)."""
    g = rng(seed)
    lon, lat, lon_ij, lat_ij = flat_grid(nx, ny)
    nij = nx * ny
    depth = np.array([5.0, 15.0, 30.0, 60.0, 120.0, 250.0][:nlev])
    members = np.zeros((k, 4 * nlev, nij), np.float32)
    # temperature: mixed layer whose depth varies with longitude
    mld = 20.0 + 80.0 * (0.5 + 0.5 * np.sin(np.radians(lon_ij)))
    tm = 20.0 - 0.5 * np.maximum(depth[:, None] - mld[None, :], 0.0) / 10.0
    for v in range(4):
        d = smooth_perts(g, k * nlev, lat, lon, 800e3, 0.3).reshape(k, nlev, nij)
        base = tm if v == 0 else (35.0 if v == 1 else 0.0)
        members[:, v * nlev:(v + 1) * nlev] = (base + d).astype(np.float32)
    obs = [(1, 1, 0, 20.0, -140.0, 5.0, 0.0, -1.0, 0.1),
           (1, 2, 0, 0.0, -140.0, 5.0, 0.0, -1.0, 1.0),
           (1, 2, 0, 0.0, -150.0, 5.0, 0.0, -1.0, 1.0),
           (1, 2, 0, 0.0, -160.0, 5.0, 0.0, -1.0, 1.0),
           (1, 2, 1, 0.0, -160.0, 5.0, 0.0, 0.5, 1.0)]
    pts = to_xyz(lat_ij, lon_ij)
    cols = []
    for (oid, pid, var, la, lo, dep, tmv, inc, err) in obs:
        j = int(np.argmax(pts @ to_xyz(np.array(la), np.array(np.mod(lo, 360.0)))))
        lvl = int(np.argmin(np.abs(depth - dep)))
        cols.append(members[:, var * nlev + lvl, j])
    hx = np.stack(cols, 1)
    hxm = hx.sum(0, dtype=np.float32) / np.float32(k)
    val = np.array([o[7] for o in obs], np.float32) + hxm
    loc = Localizer(cls="LOC_OCEAN",
                    hzloc_prof=LinearInterpLat([0.0, 90.0], [720e3, 200e3]),
                    hzloc_sat=LinearInterpLat([0.0, 90.0], [500e3, 50e3]),
                    vtloc_surf=1, bkg_t_delta=0.2, bkg_t_var_idx=0, bkg_t_var_len=nlev,
                    sat_obs=[], sat_plats=[1])
    return make_problem(members, lat_ij, lon_ij, np.zeros(nij, bool),
                        [o[3] for o in obs], [np.mod(o[4], 360.0) for o in obs],
                        [o[5] for o in obs], [o[6] for o in obs], val,
                        [o[8] for o in obs], [o[0] for o in obs], [o[1] for o in obs],
                        hx, [0, nlev, 2 * nlev, 3 * nlev], [nlev] * 4, loc,
                        meta=dict(case="ocean_vtloc", seed=seed))


# ---------------------------------------------------------------- one point
def offset_point(lat, lon, dist, bearing=0.0):
    """(lat, lon) in degrees at great-circle distance ``dist`` (m) from
    (lat, lon) along ``bearing`` (degrees from north)."""
    p1, l1, b, a = (np.radians(lat), np.radians(lon), np.radians(bearing), dist / RE)
    p2 = np.arcsin(np.sin(p1) * np.cos(a) + np.cos(p1) * np.sin(a) * np.cos(b))
    l2 = l1 + np.arctan2(np.sin(b) * np.sin(a) * np.cos(p1), np.cos(a) - np.sin(p1) * np.sin(p2))
    return float(np.degrees(p2)), float(np.mod(np.degrees(l2), 360.0))


def point_problem(loc, obs, var_lev=(1,), mean_profile=None, k=10, seed=0x1E7F0300,
                  lat=10.0, lon=200.0, inc_max=None, bounds=None):
    """One unmasked gridpoint at (lat, lon) and a handful of observations,
    the single-obs OBSIO_TEST pattern (src/subclass.letkf_obsio/
    letkf_obsio_test.F90:146-220) reduced to one column.

    ``obs`` is a list of dicts with keys dist (m), bearing (deg), obsid,
    platid, zdim, time, dep (O-F), err and slab (the slab whose members are
    the obs' hx).  ``mean_profile`` (ns,) fixes the background mean of the
    column (the mixed-layer level reads it; default 15 everywhere, an SST-like
    value); every slab's perturbations are one common draw, so each
    observation correlates with every slab."""
    g = rng(seed)
    ns = int(sum(var_lev))
    base = np.full(ns, 15.0) if mean_profile is None else np.asarray(mean_profile, float)
    e = g.standard_normal(k)
    e = 0.5 * (e - e.mean()) / e.std()
    members = (base[None, :, None] + e[:, None, None]).astype(np.float32)
    var_s = np.concatenate([[0], np.cumsum(var_lev)[:-1]]).astype(int)
    olat, olon, cols = [], [], []
    for o in obs:
        la, lo = offset_point(lat, lon, o.get("dist", 0.0), o.get("bearing", 0.0))
        olat.append(la)
        olon.append(lo)
        cols.append(members[:, o.get("slab", 0), 0])
    hx = np.stack(cols, 1)
    hxm = hx.sum(0, dtype=np.float32) / np.float32(k)
    val = np.array([o.get("dep", 1.0) for o in obs], np.float32) + hxm
    return make_problem(members, [lat], [lon], np.zeros(1, bool), olat, olon,
                        [o.get("zdim", 0.0) for o in obs], [o.get("time", 0.0) for o in obs],
                        val, [o.get("err", 1.0) for o in obs],
                        [o.get("obsid", 1) for o in obs], [o.get("platid", 1) for o in obs],
                        hx, var_s, list(var_lev), loc, inc_max=inc_max, bounds=bounds,
                        meta=dict(case="point", seed=seed))


# ---------------------------------------------------------------- random
def random_problem(seed=7, nx=48, ny=24, k=20, nobs=3000, ns=2, radius=600e3,
                   loc_cls="LOC_NOVRT", rtps=0.0, rtpp=0.0, mul=1.0,
                   err=0.05, with_mask=True, qc_like_times=False):
    """Small dense random analysis for unit and parity tests."""
    g = rng(seed)
    lon, lat, lon_ij, lat_ij = flat_grid(nx, ny)
    nij = nx * ny
    d = smooth_perts(g, k * ns, lat, lon, 700e3, 0.5).reshape(k, ns, nij)
    base = (10.0 + 5.0 * np.cos(np.radians(lat_ij)))[None, None, :]
    members = (base + d).astype(np.float32)
    olat, olon = uniform_sphere(g, nobs)
    oslab = g.integers(0, ns, nobs)
    hx = np.zeros((k, nobs))
    for s in range(ns):
        sel = oslab == s
        hx[:, sel] = bilinear(members[:, s, :].reshape(k, ny, nx), lat, lon, olat[sel], olon[sel])
    val = hx.mean(0) + g.normal(0.0, math.sqrt(err), nobs) + 0.3 * g.standard_normal(nobs)
    otime = g.uniform(-6, 6, nobs) if qc_like_times else np.zeros(nobs)
    zd = np.where(oslab == 0, 0.0, 50.0)
    obsid = 1 + oslab
    platid = g.integers(1, 3, nobs)
    mask = land_mask(lat_ij, lon_ij) if with_mask else np.zeros(nij, bool)
    if loc_cls == "LOC_NOVRT":
        loc = Localizer(cls="LOC_NOVRT",
                        hzloc=LinearInterpLat([10.0, 60.0], [radius, 0.5 * radius]))
        var_s, var_lev = list(range(ns)), [1] * ns
    elif loc_cls == "LOC_OCEAN":
        loc = Localizer(cls="LOC_OCEAN",
                        hzloc_prof=LinearInterpLat([0.0, 90.0], [radius, 0.5 * radius]),
                        hzloc_sat=LinearInterpLat([0.0, 30.0], [0.6 * radius, 0.3 * radius]),
                        tloc_prof=-1.0, tloc_sat=10.0 if qc_like_times else -1.0,
                        vtloc_surf=1, bkg_t_delta=0.05, bkg_t_var_idx=0, bkg_t_var_len=ns,
                        sat_obs=[1], sat_plats=[1])
        var_s, var_lev = [0], [ns]
    else:
        depth = np.linspace(0.0, 100.0, ns).astype(np.float32)
        loc = Localizer(cls="LOC_VGAUSS",
                        hzloc=LinearInterpLat([0.0, 90.0], [radius, 0.5 * radius]),
                        vt_scale=40.0, lev_depth=depth)
        var_s, var_lev = [0], [ns]
    return make_problem(members, lat_ij, lon_ij, mask, olat, olon, zd, otime, val,
                        np.full(nobs, err), obsid, platid, hx, var_s, var_lev, loc,
                        rtps=rtps, rtpp=rtpp, mul=mul,
                        meta=dict(case="random", seed=seed))
