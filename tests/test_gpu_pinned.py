"""The behaviours the oracle is pinned on by closed forms (LOC_OCEAN
surface classification, mixed-layer level, time localization, maxhz,
LOC_VGAUSS weights, the increment clip and the bounds), run through the GPU
library on both paths and compared with the oracle.

Bar: bit-exact selection (counts, kept-set hashes, search-set sizes, radii)
on every path; plain FP64 within 1e-10 of the oracle on mean, spread and
members; batched FP32 / bf16x3 within 1e-5 on mean and spread.
"""
import numpy as np
import pytest

from helpers import compare, oracle_run, rel, run_gpu
from letkf_amd import synth
from letkf_amd.problem import LinearInterpLat

pytestmark = pytest.mark.gpu

PATHS = [("plain", "fp64"), ("batched", "fp32"), ("batched", "bf16x3")]


def _ocean(sat_obs, sat_plats, tloc_prof=-1.0, tloc_sat=-1.0, nlev=1, delta=0.2,
           prof=(500e3, 500e3), vtloc=1):
    return synth.Localizer(
        cls="LOC_OCEAN", hzloc_prof=LinearInterpLat([0.0, 90.0], list(prof)),
        hzloc_sat=LinearInterpLat([0.0, 90.0], [300e3, 300e3]),
        tloc_prof=tloc_prof, tloc_sat=tloc_sat, vtloc_surf=vtloc, bkg_t_delta=delta,
        bkg_t_var_idx=0, bkg_t_var_len=nlev, sat_obs=list(sat_obs), sat_plats=list(sat_plats))


def _cases():
    pairs = [(1, 1), (1, 2), (2, 1), (2, 2)]
    out = {}
    for so, sp in (([], []), ([1], []), ([], [2]), ([1], [2])):
        out[f"surface_{so}_{sp}"] = lambda so=so, sp=sp: synth.point_problem(
            _ocean(so, sp), [dict(obsid=a, platid=b, dist=d, dep=0.3 * (1 + i))
                             for i, ((a, b), d) in enumerate(zip(pairs, (0.0, 50e3, 120e3, 200e3)))],
            k=20)
    for which in ("sat", "prof"):
        kw = dict(tloc_sat=6.0) if which == "sat" else dict(tloc_prof=6.0)
        oid = 1 if which == "sat" else 2
        out[f"tloc_{which}"] = lambda kw=kw, oid=oid: synth.point_problem(
            _ocean([1], [], **kw), [dict(obsid=oid, time=t, dep=0.4) for t in (0.0, 3.0, -6.0, 9.0)],
            k=20)
    for lat in (9.0, 81.0):
        out[f"maxhz_{lat}"] = lambda lat=lat: synth.point_problem(
            _ocean([1], [], prof=(500e3, 100e3), vtloc=0),
            [dict(obsid=2, dist=690e3, dep=0.5), dict(obsid=1, dist=100e3, dep=-0.2)], k=20, lat=lat)
    profiles = {3: [20.0, 20.1, 20.15, 19.5, 15.0, 10.0], 1: [20.0, 20.5, 20.5, 20.5, 20.5, 20.5],
                6: [20.0, 20.1, 19.9, 20.15, 19.85, 20.0], 2: [20.0, 19.9, 19.7, 19.7, 19.7, 19.7]}
    for mld, prof in profiles.items():
        out[f"mld_{mld}"] = lambda prof=prof: synth.point_problem(
            _ocean([], [1], nlev=6), [dict(platid=1, dep=0.5), dict(platid=2, dist=80e3, dep=-0.3)],
            var_lev=(6, 6, 1), mean_profile=list(prof) + [35.0] * 6 + [1.0], k=20)
    depth = np.array([0.0, 40.0, 80.0, 120.0], np.float32)
    out["vgauss"] = lambda: synth.point_problem(
        synth.Localizer(cls="LOC_VGAUSS", hzloc=LinearInterpLat([0.0, 90.0], [500e3, 500e3]),
                        vt_scale=40.0, lev_depth=depth),
        [dict(dist=0.0, zdim=0.0, dep=0.5), dict(dist=400e3, zdim=60.0, dep=-0.4, slab=2)],
        var_lev=(4,), k=20)
    out["clip"] = lambda: synth.point_problem(
        synth.novrt(500e3), [dict(dep=2.0)], k=20, inc_max=np.array([0.1], np.float32))
    out["bounds"] = lambda: synth.point_problem(
        synth.novrt(500e3), [dict(dep=2.0)], k=20,
        bounds=np.array([[-9.99e36], [15.05]], np.float32))    # background mean 15
    return out


CASES = _cases()


@pytest.mark.parametrize("path,prec", PATHS)
@pytest.mark.parametrize("case", sorted(CASES))
def test_pinned_point_cases(case, path, prec):
    p = CASES[case]()
    print(f"case {case} {path} {prec} k={p.k} ns={p.ns} nobs={p.nobs}", flush=True)
    gpu = run_gpu(p, precision=prec, path=path)
    orc = oracle_run(p)
    if path == "plain":
        errs, _ = compare(p, gpu, orc)
    else:
        assert np.array_equal(gpu["lg_obscount"].T.astype(np.int64), orc["lg_m"])
        assert np.array_equal(gpu["lg_selhash"].T, orc["lg_hash"])
        assert np.array_equal(gpu["obscount"].astype(np.int64), orc["obscount"])
        assert np.array_equal(gpu["maxhz"], orc["maxhz"])
        errs = dict(mean=rel(gpu["mean"], orc["mean"]), sprd=rel(gpu["sprd"], orc["sprd"]))
        assert errs["mean"] <= 1e-5 and errs["sprd"] <= 1e-5, errs
    print("  ", errs, flush=True)


@pytest.mark.parametrize("path,prec", PATHS)
def test_bounds_check_clip_binds(path, prec):
    """test/simple3d/bounds_check rebuilt so that ana_inc_max binds on state2
    (-5.0 departure, increments beyond 2) and state3 (beyond 0.2): the GPU
    analysis equals the oracle's, whose clip is pinned by closed forms."""
    p = synth.simple3d("bounds_check")
    print(f"bounds_check {path} {prec}", flush=True)
    gpu = run_gpu(p, precision=prec, path=path)
    orc = oracle_run(p)
    nlev = int(p.var_lev[0])
    inc = gpu["mean"][nlev:2 * nlev] - p.mean[nlev:2 * nlev]
    assert np.abs(inc).max() == pytest.approx(2.0, abs=1e-6)      # the clip binds
    if path == "plain":
        compare(p, gpu, orc)
    else:
        assert np.array_equal(gpu["lg_selhash"].T, orc["lg_hash"])
        assert rel(gpu["mean"], orc["mean"]) <= 1e-5
        assert rel(gpu["sprd"], orc["sprd"]) <= 1e-5
