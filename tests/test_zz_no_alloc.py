"""letkf_gpu_get_state_members_noalloc (no device allocation; staging in
the library's file-scope device array) against letkf_gpu_get_state_members
(staging buffers allocated per call), through the C ABI: bit-identical
members, mean and spread on two calls in a row, the second on a smaller
analysis, and with the staging forced into ragged slices of a few
gridpoints."""
import numpy as np
import pytest

from letkf_amd import synth
from letkf_amd.solver import LetkfGpuSolver

pytestmark = pytest.mark.gpu


def _both(prob, precision="fp64", path="plain"):
    with LetkfGpuSolver(prob, precision=precision, path=path) as s:
        s.run()
        ref = s.state_members()
        got = s.state_members_noalloc()
    return ref, got


@pytest.mark.parametrize("slice_env", [None, "97"])
def test_noalloc_egress_matches(slice_env, monkeypatch):
    if slice_env:
        monkeypatch.setenv("LETKF_GPU_NOALLOC_SLICE", slice_env)
    big = synth.random_problem(seed=51, nx=48, ny=24, k=20, ns=3, nobs=2000, rtps=0.5)
    small = synth.random_problem(seed=52, nx=20, ny=10, k=24, ns=2, nobs=500)
    for prob, prec, path in ((big, "fp64", "plain"), (small, "bf16x3", "batched")):
        print("noalloc", prob.nij, prob.k, prob.ns, prec, slice_env, flush=True)
        ref, got = _both(prob, prec, path)
        for a, b in zip(ref, got):
            assert a.shape == b.shape and np.array_equal(a, b)
        assert np.abs(ref[0]).max() > 0
