"""Shared helpers for the parity tests: run the GPU library through its C
ABI and compare with the FP64 oracle."""
import numpy as np

import oracle as O
from letkf_amd.solver import LetkfGpuSolver


def rel(a, b):
    """max |a - b| / max |b| (0 when both are empty)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if b.size == 0:
        return 0.0
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def run_gpu(prob, precision="fp64", path="plain", max_obs=0):
    with LetkfGpuSolver(prob, precision=precision, path=path, max_obs=max_obs) as s:
        s.run()
        st, mn, sp = s.state_f64()
        d = s.diag()
        t = s.timing()
    return dict(state=st, mean=mn, sprd=sp, timing=t, **d)


def compare(prob, gpu, orc, points=None, tol=1e-10):
    """Bit-exact selection and FP64 agreement of mean, spread and members."""
    pts = np.arange(prob.nij) if points is None else np.asarray(points)
    G = prob.loc.maxgroups()
    lg_m = gpu["lg_obscount"][pts, :].T.astype(np.int64)
    assert np.array_equal(lg_m, orc["lg_m"]), "kept-observation counts differ"
    assert np.array_equal(gpu["lg_selhash"][pts, :].T, orc["lg_hash"]), "kept sets differ"
    assert np.array_equal(gpu["obscount"][pts].astype(np.int64), orc["obscount"])
    assert np.array_equal(gpu["maxhz"][pts], orc["maxhz"])
    lo = gpu["lg_obsloc"][pts, :].T
    assert np.allclose(lo, orc["lg_obsloc"], rtol=2e-6, atol=1e-6)
    errs = dict(mean=rel(gpu["mean"][:, pts], orc["mean"]),
                sprd=rel(gpu["sprd"][:, pts], orc["sprd"]),
                state=rel(gpu["state"][:, :, pts], orc["state"]),
                inc=rel(gpu["mean"][:, pts] - prob.mean[:, pts],
                        orc["mean"] - prob.mean[:, pts]))
    assert errs["mean"] <= tol and errs["sprd"] <= tol and errs["state"] <= tol, errs
    return errs, G


def oracle_run(prob, points=None):
    return O.analysis(prob, points=points)
