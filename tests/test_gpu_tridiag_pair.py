"""The two-problem tridiagonalization (launch choice tridiag_reg = 2: two
problems per wave, rows r and r + 32 in a lane, k <= 50 and at most 3
carried vectors; otherwise the one-problem kernel runs) against the
one-problem register kernel and the FP64 oracle.

The grid has an odd number of gridpoints (the last wave's second half has
no problem) and masked gridpoints (unsolved problems, skipped in the batched
path's order, interleaved on the plain path).  Both kernels compute the same
reflectors with sums in a different order, so the analyses agree to
rounding (1e-12 relative); each meets the oracle as the path does (FP64
plain path 1e-10, batched bf16x4 1e-5 on the mean and spread)."""
import numpy as np
import pytest

from helpers import oracle_run, rel
from letkf_amd import synth
from letkf_amd.solver import LetkfGpuSolver

pytestmark = pytest.mark.gpu


def _run(p, precision, path, tr):
    with LetkfGpuSolver(p, precision=precision, path=path) as s:
        s.set_tuning(2, 0, tr)
        s.run()
        st, mn, sp = s.state_f64()
        hs = s.diag()["lg_selhash"]
    return st, mn, sp, hs


@pytest.mark.parametrize("ns", [1, 2, 4])
@pytest.mark.parametrize("k", [32, 40, 49, 50])
@pytest.mark.parametrize("precision,path", [("fp64", "plain"), ("bf16x4", "batched")])
def test_pair_matches_one_problem_and_oracle(precision, path, k, ns):
    p = synth.random_problem(seed=900 + k + ns, nx=47, ny=23, k=k, ns=ns, nobs=2500, rtps=0.3)
    print("tridiag pair", precision, path, "k", k, "ns", ns, "nij", p.nij, flush=True)
    one = _run(p, precision, path, 1)
    two = _run(p, precision, path, 2)
    assert np.array_equal(one[3], two[3])
    d = {f: rel(a, b) for f, a, b in zip(("state", "mean", "sprd"), one[:3], two[:3])}
    print("pair vs one-problem", d, flush=True)
    assert max(d.values()) <= 1e-12, d
    orc = oracle_run(p)
    st, mn, sp, hs = two
    assert np.array_equal(hs.T, orc["lg_hash"])
    tol = 1e-10 if precision == "fp64" else 1e-5
    e = dict(mean=rel(mn, orc["mean"]), sprd=rel(sp, orc["sprd"]))
    if precision == "fp64":
        e["state"] = rel(st, orc["state"])
    print("pair vs oracle", e, flush=True)
    assert max(e.values()) <= tol, e


@pytest.mark.parametrize("ns", [1, 2])
@pytest.mark.parametrize("k", [32, 49, 50])
def test_rational_solve_fused(k, ns, monkeypatch):
    """On the batched path the rational solve runs inside the two-problem
    tridiagonalization (default) or as its own kernel (LETKF_GPU_FUSE=0):
    the same arithmetic on the same values, so the analyses agree to
    rounding (1e-12 relative), and the fused one meets the oracle (1e-5)."""
    p = synth.random_problem(seed=950 + k + ns, nx=47, ny=23, k=k, ns=ns, nobs=2500, rtps=0.3)
    print("fused rational solve k", k, "ns", ns, flush=True)
    fused = _run(p, "bf16x4", "batched", 2)
    monkeypatch.setenv("LETKF_GPU_FUSE", "0")
    apart = _run(p, "bf16x4", "batched", 2)
    d = {f: rel(a, b) for f, a, b in zip(("state", "mean", "sprd"), fused[:3], apart[:3])}
    print("fused vs separate", d, flush=True)
    assert max(d.values()) <= 1e-12, d
    orc = oracle_run(p)
    e = dict(mean=rel(fused[1], orc["mean"]), sprd=rel(fused[2], orc["sprd"]))
    assert max(e.values()) <= 1e-5, e
