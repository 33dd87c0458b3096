"""Post-processing inside the large-ensemble apply (64 <= k <= 128, mixed
precisions): the apply kernel recenters, clips the mean increment, relaxes
the spread (RTPS) or the perturbations (RTPP), inflates, bounds the members
and takes the spread of the slabs it has just transformed, and the separate
post-processing kernel only handles masked gridpoints and the slabs of
groups without a solved problem.  LETKF_GPU_FUSEFIN=0 keeps the apply
writing the transformed perturbations and the separate kernel doing every
slab.

Both compute the same FP64 steps from the same FP32 products, with the sums
over members in another order, so the analyses agree to rounding (1e-12
relative); the fused one meets the oracle (1e-5 on mean and spread).  The
cases make the increment clip and the member bounds bind (checked on the
oracle: the limits change its analysis), cover a ragged last tile of slabs,
ensemble sizes that leave whole waves of the block without members (k = 64)
or with one (k = 97), and k not a multiple of 4 (element stores)."""
import numpy as np
import pytest

from helpers import oracle_run, rel, run_gpu
from letkf_amd import synth

pytestmark = pytest.mark.gpu

CASES = [
    # k, ns, rtps, rtpp, mul, limits
    (64, 5, 0.5, 0.0, 1.0, True),
    (71, 37, 0.0, 0.4, 1.1, False),
    (97, 4, 0.0, 0.0, 1.2, True),
    (100, 6, 0.3, 0.0, 1.0, True),
    (127, 35, 0.0, 0.3, 1.0, True),
]


def _problem(k, ns, rtps, rtpp, mul, limits):
    p = synth.random_problem(seed=1200 + k + ns, nx=47, ny=23, k=k, ns=ns, nobs=2500,
                             rtps=rtps, rtpp=rtpp, mul=mul)
    if limits:
        # every slab is its own variable: clip the mean increment of the
        # even ones, bound the members of the odd ones
        nv = len(p.var_s)
        inc = np.full(nv, 9.99e36, np.float32)
        inc[0::2] = 0.02
        bnd = np.array([[-9.99e36] * nv, [9.99e36] * nv], np.float32)
        bnd[0, 1::2] = 11.0
        bnd[1, 1::2] = 14.0
        p.inc_max = np.ascontiguousarray(inc)
        p.bounds = np.asfortranarray(bnd)
        p.validate()
    return p


@pytest.mark.parametrize("k,ns,rtps,rtpp,mul,limits", CASES)
def test_fused_post_processing(k, ns, rtps, rtpp, mul, limits, monkeypatch):
    p = _problem(k, ns, rtps, rtpp, mul, limits)
    print("fused post-processing k", k, "ns", ns, "rtps", rtps, "rtpp", rtpp, "mul", mul,
          "limits", limits, flush=True)
    fused = run_gpu(p, precision="bf16x3", path="batched")
    monkeypatch.setenv("LETKF_GPU_FUSEFIN", "0")
    apart = run_gpu(p, precision="bf16x3", path="batched")
    d = {f: rel(fused[f], apart[f]) for f in ("state", "mean", "sprd")}
    print("fused vs separate", d, flush=True)
    assert max(d.values()) <= 1e-12, d
    orc = oracle_run(p)
    assert np.array_equal(fused["lg_selhash"].T, orc["lg_hash"])
    e = dict(mean=rel(fused["mean"], orc["mean"]), sprd=rel(fused["sprd"], orc["sprd"]))
    print("fused vs oracle", e, flush=True)
    assert max(e.values()) <= 1e-5, e
    if limits:
        free = _problem(k, ns, rtps, rtpp, mul, False)
        assert rel(oracle_run(free)["mean"], orc["mean"]) > 1e-4
