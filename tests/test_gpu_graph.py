"""letkf_gpu_run_graph: the analysis captured once as a HIP graph and
replayed, against letkf_gpu_run (eager) and the FP64 oracle.

* Every call gives the eager analysis bit for bit: the first call runs
  eagerly and captures, later calls replay.
* A cycled analysis (new background and a new observation set of another
  size each cycle) replays the same graph: the handle's buffers keep their
  addresses.
* A cycle that needs more list space than the capture falls back to an
  eager run with larger lists and is still exact; a changed setting is
  captured again.
* Cycled configuration on the 0.25-degree grid (k = 50, 1-2M fresh swath
  observations per cycle, background = previous analysis moved one
  gridpoint east + noise): three cycles replayed from one capture, with
  bit-exact selection and mean / spread within 1e-5 (batched bf16x3) and
  1e-10 (plain FP64) of the oracle on a sample of gridpoints.
"""
import dataclasses

import numpy as np
import pytest

import oracle as O
from helpers import compare, rel
from letkf_amd import synth
from letkf_amd.solver import LetkfGpuSolver

pytestmark = pytest.mark.gpu


def _eager(prob, precision, path, members=None, obs_members=None):
    with LetkfGpuSolver(prob, precision=precision, path=path, members=members,
                        obs_members=obs_members) as s:
        s.run()
        return s.state_f64(), s.diag()["lg_selhash"]


@pytest.mark.parametrize("precision,path", [("fp64", "plain"), ("bf16x3", "batched")])
def test_graph_replays_equal_eager(precision, path):
    p = synth.random_problem(seed=61, k=24, ns=2, nobs=3000, rtps=0.5)
    ref, href = _eager(p, precision, path)
    with LetkfGpuSolver(p, precision=precision, path=path) as s:
        for call in range(3):
            print("graph call", call, precision, path, flush=True)
            s.run_graph()
            t = s.timing()
            replayed = t["search_ms"] == 0.0 and t["n_launch"] == 1
            assert replayed == (call > 0), t
            got = s.state_f64()
            for a, b in zip(got, ref):
                assert np.array_equal(a, b)
            assert np.array_equal(s.diag()["lg_selhash"], href)


def _cycle_inputs(prob, members, g, nobs):
    """A new background (members moved one gridpoint east + noise) and a
    new random observation set of nobs observations."""
    nx, ny = prob.meta.get("nx", 48), prob.meta.get("ny", 24)
    k = prob.k
    mem = np.asarray(members, np.float32).reshape(ny, nx, prob.ns, k)
    mem = np.roll(mem, 1, axis=1) + g.normal(0.0, 0.1, mem.shape).astype(np.float32)
    mem = np.asfortranarray(mem.reshape(nx * ny, prob.ns, k))
    q = synth.random_problem(seed=int(g.integers(1, 1 << 30)), nx=nx, ny=ny, k=k, ns=prob.ns,
                             nobs=nobs, rtps=prob.rtps)
    qc = np.zeros(q.nobs, np.int32)
    hxm = np.asfortranarray((q.hx + q.hxmean[None]).astype(np.float32))
    return q, mem, qc, hxm


@pytest.mark.parametrize("precision,path", [("fp64", "plain"), ("bf16x3", "batched")])
def test_graph_cycles_replay_one_capture(precision, path):
    g = np.random.default_rng(7)
    p = synth.random_problem(seed=62, nx=48, ny=24, k=20, ns=1, nobs=4000, rtps=0.5)
    p.meta.update(nx=48, ny=24)
    mem = np.asfortranarray((p.state + p.mean[None]).transpose(2, 1, 0), np.float32)
    qc = np.zeros(p.nobs, np.int32)
    hxm = np.asfortranarray((p.hx + p.hxmean[None]).astype(np.float32))
    with LetkfGpuSolver(p, precision=precision, path=path, members=mem,
                        obs_members=(qc, hxm)) as s:
        s.run_graph()                                  # eager + capture
        for c, nobs in enumerate((3500, 3900, 2000)):  # within the obs buffers' room
            q, mem, qc, hxm = _cycle_inputs(p, mem, g, nobs)
            print("cycle", c, nobs, precision, path, flush=True)
            s.cycle(q, members=mem, obs_members=(qc, hxm))
            s.run_graph()
            t = s.timing()
            assert t["search_ms"] == 0.0 and t["n_launch"] == 1, t   # a replay
            got = s.state_f64()
            ref, href = _eager(q, precision, path, members=mem, obs_members=(qc, hxm))
            for a, b in zip(got, ref):
                assert np.array_equal(a, b)
            assert np.array_equal(s.diag()["lg_selhash"], href)
            mem = s.state_members()[0]


def test_graph_list_overflow_falls_back_and_settings_recapture():
    """New observations of the same count but all inside the polar caps,
    where each observation falls in the covers of many more gridpoints:
    the replay runs out of list space, is redone eagerly (exact), and the
    next call captures again.  More observations than the buffers hold,
    and a changed inflation setting, capture again too."""
    p = synth.random_problem(seed=63, k=20, ns=2, nobs=400, rtps=0.5)
    q = synth.random_problem(seed=64, k=20, ns=2, nobs=4000, rtps=0.5)
    polar = synth.random_problem(seed=66, k=20, ns=2, nobs=4000, rtps=0.5)
    g = np.random.default_rng(3)
    polar.olat = np.where(g.random(4000) < 0.5, 1.0, -1.0).astype(np.float32) * \
        g.uniform(84.0, 89.9, 4000).astype(np.float32)
    with LetkfGpuSolver(p) as s:
        s.run_graph()
        s.run_graph()
        assert s.timing()["n_launch"] == 1
        s.cycle(q)                        # more obs than the buffers hold: new capture
        s.run_graph()
        assert s.timing()["search_ms"] > 0.0
        ref, _ = _eager(q, "fp64", "plain")
        for a, b in zip(s.state_f64(), ref):
            assert np.array_equal(a, b)
        s.run_graph()                     # replay of the new capture
        assert s.timing()["n_launch"] == 1
        for a, b in zip(s.state_f64(), ref):
            assert np.array_equal(a, b)
        s.cycle(polar)                    # same count: replay, out of list space
        s.run_graph()
        assert s.timing()["search_ms"] > 0.0               # redone eagerly
        pref, _ = _eager(polar, "fp64", "plain")
        for a, b in zip(s.state_f64(), pref):
            assert np.array_equal(a, b)
        s.run_graph()                     # lists grew: captured again
        assert s.timing()["search_ms"] > 0.0
        s.run_graph()
        assert s.timing()["n_launch"] == 1
        for a, b in zip(s.state_f64(), pref):
            assert np.array_equal(a, b)
        s._chk(s.lib.letkf_gpu_set_solver(s.h, 0.0, 0.3, 1.0))   # RTPP instead of RTPS
        s.run_graph()
        assert s.timing()["search_ms"] > 0.0               # captured again


def test_cycled_config5_three_cycles():
    """Three cycles of the cycled 0.25-degree analysis (k = 50, 1-2M fresh
    swath observations per cycle) on both paths from one capture each:
    selection bit-exact and the analysis against the FP64 oracle on 300
    gridpoints per cycle (mean, spread: 1e-5 batched bf16x3, 1e-10 plain)."""
    g = np.random.default_rng(4321)
    prob = synth.config3()
    mem = np.asfortranarray((prob.state + prob.mean[None]).transpose(2, 1, 0), np.float32)
    pool = np.nonzero(prob.mask == 0)[0]
    solvers = {}
    try:
        for c in range(4):
            if c > 0:
                prob = synth.config5_cycle(prob, mem, c)
            qc = np.zeros(prob.nobs, np.int32)
            hxm = np.asfortranarray((prob.hx + prob.hxmean[None]).astype(np.float32))
            ost, omn, osp = O.state_from_members(mem)
            _, ohx, ohm = O.obs_preprocess(hxm, qc)
            oprob = dataclasses.replace(prob, state=ost, mean=omn, sprd=osp, hx=ohx, hxmean=ohm)
            pts = np.sort(g.choice(pool, 300, replace=False))
            orc = O.analysis(oprob, points=pts)
            nxt = None
            for name, prec, path, tol in (("batched", "bf16x3", "batched", 1e-5),
                                          ("plain", "fp64", "plain", 1e-10)):
                print("config5 cycle", c, name, prob.nobs, flush=True)
                s = solvers.get(name)
                if s is None:
                    s = LetkfGpuSolver(prob, precision=prec, path=path, members=mem,
                                       obs_members=(qc, hxm))
                    solvers[name] = s
                else:
                    s.cycle(prob, members=mem, obs_members=(qc, hxm))
                s.run_graph()
                if c > 0:
                    assert s.timing()["n_launch"] == 1, (name, c)   # replayed
                st, mn, sp = s.state_f64()
                d = s.diag()
                assert np.array_equal(d["lg_selhash"][pts, :].T, orc["lg_hash"])
                em, es = rel(mn[:, pts], orc["mean"]), rel(sp[:, pts], orc["sprd"])
                print("  mean", em, "sprd", es, flush=True)
                assert em <= tol and es <= tol
                if name == "batched":
                    nxt = s.state_members()[0]
            mem = nxt
    finally:
        for s in solvers.values():
            s.final()
