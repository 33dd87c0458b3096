"""Batched mixed-precision path against the FP64 oracle.

Selection is the same search as the FP64 path, so it must stay bit-exact.
Analysis errors (max |gpu - oracle| / max |oracle|) are bounded per
precision of the products: FP32 and split-BF16 (bf16x3) within 1e-5 on the
mean and spread, plain BF16 within 1e-2 (reported, not a target).
"""
import numpy as np
import pytest

from helpers import compare, oracle_run, rel, run_gpu
from letkf_amd import synth

pytestmark = pytest.mark.gpu

TOL = {"fp32": 1e-5, "bf16x3": 1e-5, "bf16": 1e-2}


def _mixed(prob, prec, points=None):
    gpu = run_gpu(prob, precision=prec, path="batched")
    orc = oracle_run(prob, points)
    pts = np.arange(prob.nij) if points is None else np.asarray(points)
    assert np.array_equal(gpu["lg_selhash"][pts, :].T, orc["lg_hash"])
    errs = dict(mean=rel(gpu["mean"][:, pts], orc["mean"]),
                sprd=rel(gpu["sprd"][:, pts], orc["sprd"]),
                members=rel(gpu["state"][:, :, pts], orc["state"]),
                inc=rel(gpu["mean"][:, pts] - prob.mean[:, pts], orc["mean"] - prob.mean[:, pts]))
    print(prec, prob.meta, errs, flush=True)
    assert errs["mean"] <= TOL[prec] and errs["sprd"] <= TOL[prec], errs
    return errs


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "bf16"])
def test_config1(prec):
    _mixed(synth.config1(), prec)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "bf16"])
@pytest.mark.parametrize("k", [20, 50])
def test_random(prec, k):
    _mixed(synth.random_problem(seed=200 + k, k=k, ns=2, nobs=4000, rtps=0.5), prec)


@pytest.mark.parametrize("prec", ["bf16x3"])
@pytest.mark.parametrize("loc", ["LOC_OCEAN", "LOC_VGAUSS"])
def test_localizers(prec, loc):
    _mixed(synth.random_problem(seed=300, k=40, ns=3, nobs=3000, loc_cls=loc,
                                qc_like_times=(loc == "LOC_OCEAN")), prec)
