"""Failure paths of the C ABI, forced from the host so that every kernel
still exits normally.

* Illegal inflation settings return LETKF_EINVAL with the messages of
  letkf_solver_init (src/letkf_solver.F90:114-125), where the reference
  aborts through letkf_mpi_abort.
* An allocation beyond the handle's device-memory cap (LETKF_GPU_MEM_LIMIT,
  read at create) returns LETKF_ENOMEM, as a full device would.
* New inputs invalidate the previous analysis (LETKF_ESTATE until the next
  run succeeds).
"""
import numpy as np
import pytest

from letkf_amd import abi, synth
from letkf_amd.solver import LetkfError, LetkfGpuSolver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rtps,rtpp,mul,msg", [
    (1.5, 0.0, 1.0, "illegal value for infl_rtps (should be <1.0 and >0.0)."),
    (-0.1, 0.0, 1.0, "illegal value for infl_rtps (should be <1.0 and >0.0)."),
    (0.0, 1.2, 1.0, "illegal value for infl_rtpp (should be <1.0 and >0.0)."),
    (0.0, -0.5, 1.0, "illegal value for infl_rtpp (should be <1.0 and >0.0)."),
    (0.5, 0.5, 1.0, "cannot have both RTPS and RTPP enabled at the same time."),
    (0.0, 0.0, 0.99, "infl_mul must be >=1.0."),
])
def test_illegal_inflation(rtps, rtpp, mul, msg):
    p = synth.random_problem(seed=1, nx=12, ny=6, nobs=50, rtps=rtps, rtpp=rtpp, mul=mul)
    print("inflation", rtps, rtpp, mul, flush=True)
    with pytest.raises(LetkfError) as e:
        LetkfGpuSolver(p)
    assert e.value.code == abi.LETKF_EINVAL
    assert msg in str(e.value)


@pytest.mark.parametrize("rtps,rtpp,mul", [(1.0, 0.0, 1.0), (0.0, 1.0, 1.0), (0.0, 0.0, 1.0),
                                           (0.0, 0.0, 1.5)])
def test_legal_inflation_edges(rtps, rtpp, mul):
    p = synth.random_problem(seed=2, nx=12, ny=6, nobs=50, rtps=rtps, rtpp=rtpp, mul=mul)
    with LetkfGpuSolver(p) as s:
        s.run()


def test_memory_cap_enomem(monkeypatch):
    """A cap below the background state fails set_state; a cap that lets the
    inputs in but not the run's work arrays fails run; both with
    LETKF_ENOMEM, and the failed run leaves no analysis to fetch.  Without
    the cap the same problem runs."""
    p = synth.random_problem(seed=3, nx=96, ny=48, k=20, nobs=3000, ns=2)
    state_bytes = p.state.nbytes
    print("state bytes", state_bytes, flush=True)
    monkeypatch.setenv("LETKF_GPU_MEM_LIMIT", str(state_bytes // 2))
    with pytest.raises(LetkfError) as e:
        LetkfGpuSolver(p)
    assert e.value.code == abi.LETKF_ENOMEM and "out of memory" in str(e.value)
    # the inputs (0.74 MB of state, about 2 MB of observation arrays and
    # upload temporaries) fit in 6 MB, the run's work arrays do not (the
    # FP64 products alone take 4608 x 20 x 20 x 8 B = 14.7 MB)
    monkeypatch.setenv("LETKF_GPU_MEM_LIMIT", str(6 << 20))
    s = LetkfGpuSolver(p)
    try:
        with pytest.raises(LetkfError) as e:
            s.run()
        assert e.value.code == abi.LETKF_ENOMEM
        with pytest.raises(LetkfError) as e2:
            s.state_f64()
        assert e2.value.code == abi.LETKF_ESTATE
    finally:
        s.final()
    monkeypatch.delenv("LETKF_GPU_MEM_LIMIT")
    with LetkfGpuSolver(p) as s:
        s.run()
        s.state_f64()


def test_new_inputs_invalidate_analysis():
    p = synth.random_problem(seed=4, nx=24, ny=12, nobs=400)
    with LetkfGpuSolver(p) as s:
        s.run()
        s.state_f64()
        s.cycle(p)                               # new background and obs
        with pytest.raises(LetkfError) as e:
            s.state_f64()
        assert e.value.code == abi.LETKF_ESTATE
        s.run()
        st, mn, sp = s.state_f64()
        assert np.isfinite(st).all()
