"""The configurations at their full sizes on one GPU.

The oracle checks a random sample of unmasked gridpoints (selection
bit-exact, FP64 within 1e-10, mixed mean and spread within 1e-5).  Over
every gridpoint, properties that hold at any size:
- the plain and the batched path select the same observations (same kept
  counts and order-free hashes everywhere: they share the search);
- masked gridpoints keep the background exactly;
- gridpoints without kept observations keep the background mean (to FP32
  rounding of the recentering);
- the batched mean agrees with the FP64 mean everywhere within the mixed
  tolerance.
"""
import numpy as np
import pytest

from helpers import oracle_run, rel
from letkf_amd import synth
from letkf_amd.solver import LetkfGpuSolver

pytestmark = pytest.mark.gpu


def _run(prob, prec, path, members):
    with LetkfGpuSolver(prob, precision=prec, path=path) as s:
        s.run()
        st, mn, sp = s.state_f64(members=members)
        d = s.diag()
        t = s.timing()
    print(f"  {path} {prec}: {t['total_ms']:.1f} ms, {t['n_problems']} solves", flush=True)
    return st, mn, sp, d


def _check(prob, nsample, seed, members=True,
           mixed=(("bf16x4", 1e-5, 1e-5, None, None), ("bf16x3", 1e-5, 1e-5, None, None))):
    g = np.random.default_rng(seed)
    pool = np.nonzero(prob.mask == 0)[0]
    pts = np.sort(g.choice(pool, min(nsample, pool.size), replace=False))
    orc = oracle_run(prob, pts)
    st, mn, sp, d = _run(prob, "fp64", "plain", members)
    assert np.array_equal(d["lg_selhash"][pts, :].T, orc["lg_hash"])
    assert np.array_equal(d["lg_obscount"][pts, :].T.astype(np.int64), orc["lg_m"])
    assert rel(mn[:, pts], orc["mean"]) <= 1e-10 and rel(sp[:, pts], orc["sprd"]) <= 1e-10
    if members:
        assert rel(st[:, :, pts], orc["state"]) <= 1e-10
    # masked gridpoints keep the background exactly
    msk = prob.mask != 0
    assert np.array_equal(mn[:, msk], prob.mean[:, msk].astype(np.float64))
    assert np.array_equal(sp[:, msk], prob.sprd[:, msk].astype(np.float64))
    # unobserved, unmasked gridpoints keep the background mean
    free = (d["obscount"] == 0) & ~msk
    if free.any():
        assert rel(mn[:, free], prob.mean[:, free]) <= 1e-6
    for prec, tol_mean, tol_sprd, tol_mem, tol_pert in mixed:
        stb, mb, sb, db = _run(prob, prec, "batched", tol_mem is not None)
        assert np.array_equal(db["lg_selhash"], d["lg_selhash"])
        assert np.array_equal(db["lg_obscount"], d["lg_obscount"])
        errs = (rel(mb[:, pts], orc["mean"]), rel(sb[:, pts], orc["sprd"]), rel(mb, mn),
                rel(sb, sp))
        print(f"  {prec}: sample mean {errs[0]:.2e} sprd {errs[1]:.2e}; whole grid against "
              f"FP64 mean {errs[2]:.2e} sprd {errs[3]:.2e}", flush=True)
        assert errs[0] <= tol_mean and errs[2] <= tol_mean
        assert errs[1] <= tol_sprd and errs[3] <= tol_sprd
        if tol_mem is not None:
            em = rel(stb[:, :, pts], orc["state"])
            ep = rel(stb[:, :, pts] - mb[None, :, pts], orc["state"] - orc["mean"][None])
            print(f"  {prec}: sample members {em:.2e} perturbations {ep:.2e}", flush=True)
            assert em <= tol_mem and ep <= tol_pert


def test_config2_full():
    """1-degree 360x180 x 50 levels T/S (ns = 100), k = 40, 200k profile obs,
    LOC_VGAUSS (one group per level)."""
    print("config2 full", flush=True)
    _check(synth.config2(), 300, 21)


def test_config3_full():
    """0.25-degree 1440x720 surface, k = 50, 2M swath obs: the headline
    configuration.  The members of bf16x4 (compensated, four-term split)
    meet 1e-5 and their perturbations 2e-5; bf16x3 is bounded at 1e-4 and
    1e-3 (its measured 2e-5 and 3e-4 come from the dropped lo*lo term and
    plain FP32 accumulation over thousands of observations)."""
    print("config3 full", flush=True)
    _check(synth.config3(), 3000, 31,
           mixed=(("bf16x4", 1e-5, 1e-5, 1e-5, 2e-5), ("bf16x3", 1e-5, 1e-5, 1e-4, 1e-3)))


def test_config4_full():
    """0.5-degree 720x360 x 75 levels T/S/u/v (ns = 300), k = 100, 1M obs,
    31 GB of state: the explicit-transform path (mean and spread fetched;
    the 62 GB FP64 members stay on the device).  FP32 and bf16x4 products
    (compensated) meet 1e-5 over the whole grid.  With bf16x3 products the
    mean does; the spread reaches 1.2e-5 at the worst gridpoint (the lo x lo
    term the split drops shows in the smallest eigenvalues at k = 100): its
    stated bound is 2e-5."""
    print("config4 full", flush=True)
    _check(synth.config4(), 150, 41, members=False,
           mixed=(("fp32", 1e-5, 1e-5, None, None), ("bf16x4", 1e-5, 1e-5, None, None),
                  ("bf16x3", 1e-5, 2e-5, None, None)))
