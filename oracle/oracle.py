"""ctypes front end of the FP64 CPU oracle (oracle/letkf_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")

_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)
_LP = C.POINTER(C.c_int64)
_UP = C.POINTER(C.c_uint64)


class OrcProblem(C.Structure):
    _fields_ = [("k", C.c_int64), ("ns", C.c_int64), ("nij", C.c_int64),
                ("nobs", C.c_int64),
                ("state", _FP), ("mean", _FP), ("sprd", _FP), ("lat", _FP),
                ("lon", _FP), ("mask", _IP),
                ("olat", _FP), ("olon", _FP), ("ozdim", _FP), ("otime", _FP),
                ("oval", _FP), ("oerr", _FP), ("obsid", _IP), ("platid", _IP),
                ("hx", _FP), ("hxmean", _FP),
                ("nvars", C.c_int32), ("var_s", _IP), ("var_lev", _IP),
                ("inc_max", _FP), ("bounds", _FP),
                ("rtps", C.c_float), ("rtpp", C.c_float), ("mul", C.c_float),
                ("loc", C.c_void_p)]


class OrcOutput(C.Structure):
    _fields_ = [("state", _DP), ("mean", _DP), ("sprd", _DP), ("maxhz", _FP),
                ("obscount", _LP), ("lg_m", _LP), ("lg_obsloc", _DP),
                ("lg_hash", _UP), ("seconds", _DP)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_gc.restype = C.c_float
        L.orc_gc.argtypes = [C.c_float, C.c_float]
        L.orc_sincosd.argtypes = [C.c_double, _DP, _DP]
        L.orc_sin_rad.restype = C.c_double
        L.orc_sin_rad.argtypes = [C.c_double]
        L.orc_asin.restype = C.c_double
        L.orc_asin.argtypes = [C.c_double]
        L.orc_exp_neg.restype = C.c_double
        L.orc_exp_neg.argtypes = [C.c_double]
        L.orc_xyz.argtypes = [C.c_float, C.c_float, _DP]
        L.orc_thr2.restype = C.c_double
        L.orc_thr2.argtypes = [C.c_float]
        L.orc_dist.restype = C.c_float
        L.orc_dist.argtypes = [C.c_double]
        L.orc_mix64.restype = C.c_uint64
        L.orc_mix64.argtypes = [C.c_uint64]
        L.orc_interp_build.restype = C.c_int
        L.orc_interp_build.argtypes = [C.c_void_p, _FP, _FP]
        L.orc_interp_get.restype = C.c_float
        L.orc_interp_get.argtypes = [C.c_int, _FP, _FP, C.c_float]
        L.orc_maxgroups.restype = C.c_int
        L.orc_maxgroups.argtypes = [C.c_void_p]
        L.orc_core_solve.restype = C.c_int
        L.orc_core_solve.argtypes = [C.c_int, C.c_int, _DP, _DP, _DP, _DP,
                                     C.c_double, _DP, _DP]
        L.orc_eigh.restype = C.c_int
        L.orc_eigh.argtypes = [C.c_int, _DP, _DP, _DP]
        L.orc_analysis.restype = C.c_int
        L.orc_analysis.argtypes = [C.POINTER(OrcProblem), _LP, C.c_int64,
                                   C.c_int, C.POINTER(OrcOutput)]
        L.orc_analysis_f32.restype = C.c_int
        L.orc_analysis_f32.argtypes = [C.POINTER(OrcProblem), _LP, C.c_int64,
                                       C.c_int, C.POINTER(OrcOutput)]
        L.r32_syev.restype = C.c_int
        L.r32_syev.argtypes = [C.c_int, _FP, _FP, _FP]
        L.r32_core_work.restype = C.c_size_t
        L.r32_core_work.argtypes = [C.c_int, C.c_int]
        L.r32_core_solve.restype = C.c_int
        L.r32_core_solve.argtypes = [C.c_int, C.c_int, _FP, _FP, _FP, _FP, C.c_float, _FP, _FP]
        L.orc_selection.restype = C.c_int64
        L.orc_selection.argtypes = [C.POINTER(OrcProblem), C.c_int64,
                                    C.c_int32, _IP, _FP, _FP, C.c_int64]
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


# ---------------------------------------------------------------- scalars
def gc(z, L):
    return lib().orc_gc(float(z), float(L))


def sincosd(deg):
    s, c = C.c_double(), C.c_double()
    lib().orc_sincosd(float(deg), C.byref(s), C.byref(c))
    return s.value, c.value


def xyz(lat, lon):
    out = np.zeros(3)
    lib().orc_xyz(float(lat), float(lon), _p(out, C.c_double))
    return out


def interp(table, x):
    """linearinterp_lat get_dist on a table given as a LinearInterpLat."""
    cfg = table.to_c()
    lat = np.zeros(40, np.float32)
    dist = np.zeros(40, np.float32)
    n = lib().orc_interp_build(C.byref(cfg), _p(lat, C.c_float), _p(dist, C.c_float))
    return lib().orc_interp_get(n, _p(lat, C.c_float), _p(dist, C.c_float), float(x))


def interp_table(table):
    cfg = table.to_c()
    lat = np.zeros(40, np.float32)
    dist = np.zeros(40, np.float32)
    n = lib().orc_interp_build(C.byref(cfg), _p(lat, C.c_float), _p(dist, C.c_float))
    return lat[:n].copy(), dist[:n].copy()


# ---------------------------------------------------------------- core
def core_solve(hdxb, rdiag, rloc, dep, infl=1.0):
    """letkf_core_solve in FP64: returns (trans (k,k), eigenvalues)."""
    hdxb = np.asfortranarray(hdxb, dtype=np.float64)
    k, m = hdxb.shape
    rdiag = np.ascontiguousarray(rdiag, np.float64)
    rloc = np.ascontiguousarray(rloc, np.float64)
    dep = np.ascontiguousarray(dep, np.float64)
    trans = np.zeros((k, k), order="F")
    ev = np.zeros(k)
    rc = lib().orc_core_solve(k, m, _p(hdxb, C.c_double), _p(rdiag, C.c_double),
                              _p(rloc, C.c_double), _p(dep, C.c_double),
                              float(infl), _p(trans, C.c_double), _p(ev, C.c_double))
    assert rc == 0
    return trans, ev


def syev32(a):
    """The FP32 eigensolver of the CPU baseline (ascending, vectors as
    columns)."""
    a = np.array(a, dtype=np.float32, order="F")
    n = a.shape[0]
    w = np.zeros(n, np.float32)
    e = np.zeros(n, np.float32)
    rc = lib().r32_syev(n, _p(a, C.c_float), _p(w, C.c_float), _p(e, C.c_float))
    assert rc == 0
    return w, a


def core_solve32(hdxb, rdiag, rloc, dep, infl=1.0):
    """letkf_core_solve in FP32, the reference's order (CPU baseline)."""
    hdxb = np.asfortranarray(hdxb, dtype=np.float32)
    k, m = hdxb.shape
    f = lambda x: np.ascontiguousarray(x, np.float32)
    rdiag, rloc, dep = f(rdiag), f(rloc), f(dep)
    trans = np.zeros((k, k), np.float32, order="F")
    work = np.zeros(lib().r32_core_work(k, m), np.float32)
    rc = lib().r32_core_solve(k, m, _p(hdxb, C.c_float), _p(rdiag, C.c_float),
                              _p(rloc, C.c_float), _p(dep, C.c_float), float(infl),
                              _p(trans, C.c_float), _p(work, C.c_float))
    assert rc == 0
    return trans


def eigh(a):
    a = np.array(a, dtype=np.float64, order="F")
    n = a.shape[0]
    w = np.zeros(n)
    v = np.zeros((n, n), order="F")
    lib().orc_eigh(n, _p(a, C.c_double), _p(w, C.c_double), _p(v, C.c_double))
    return w, v


# ---------------------------------------------------------------- analysis
class _Bound:
    """Keeps the numpy arrays alive while the C struct points at them."""

    def __init__(self, prob):
        self.prob = prob
        self.desc, self.keep = prob.loc.to_c()
        self.bounds = np.ascontiguousarray(prob.bounds.T, np.float32)  # (nvars, 2)
        s = OrcProblem()
        s.k, s.ns, s.nij, s.nobs = prob.k, prob.ns, prob.nij, prob.nobs
        s.state = _p(prob.state, C.c_float)
        s.mean = _p(prob.mean, C.c_float)
        s.sprd = _p(prob.sprd, C.c_float)
        s.lat = _p(prob.lat, C.c_float)
        s.lon = _p(prob.lon, C.c_float)
        s.mask = _p(prob.mask, C.c_int32)
        for name in ("olat", "olon", "ozdim", "otime", "oval", "oerr"):
            setattr(s, name, _p(getattr(prob, name), C.c_float))
        s.obsid = _p(prob.obsid, C.c_int32)
        s.platid = _p(prob.platid, C.c_int32)
        s.hx = _p(prob.hx, C.c_float)
        s.hxmean = _p(prob.hxmean, C.c_float)
        s.nvars = prob.nvars
        s.var_s = _p(prob.var_s, C.c_int32)
        s.var_lev = _p(prob.var_lev, C.c_int32)
        s.inc_max = _p(prob.inc_max, C.c_float)
        s.bounds = _p(self.bounds, C.c_float)
        s.rtps, s.rtpp, s.mul = prob.rtps, prob.rtpp, prob.mul
        s.loc = C.cast(C.pointer(self.desc), C.c_void_p)
        self.s = s


def analysis(prob, points=None, nthreads=None, f32=False):
    """Run the FP64 restated loop over ``points`` (default: all).

    Returns a dict with state (k, ns, npts), mean, sprd (ns, npts) in FP64
    and the diagnostics maxhz, obscount, lg_m, lg_obsloc, lg_hash
    (maxgroups, npts).  With f32=True the reference's loop runs in FP32
    instead (the CPU baseline; results widened to FP64, diagnostics other
    than lg_m not written).
    """
    prob.validate()
    b = _Bound(prob)
    pts = np.arange(prob.nij, dtype=np.int64) if points is None else \
        np.ascontiguousarray(points, dtype=np.int64)
    n = pts.size
    G = prob.loc.maxgroups()
    out = dict(
        state=np.zeros((prob.k, prob.ns, n), order="F"),
        mean=np.zeros((prob.ns, n), order="F"),
        sprd=np.zeros((prob.ns, n), order="F"),
        maxhz=np.zeros(n, np.float32),
        obscount=np.zeros(n, np.int64),
        lg_m=np.zeros((G, n), np.int64, order="F"),
        lg_obsloc=np.zeros((G, n), order="F"),
        lg_hash=np.zeros((G, n), np.uint64, order="F"),
    )
    o = OrcOutput()
    o.state = _p(out["state"], C.c_double)
    o.mean = _p(out["mean"], C.c_double)
    o.sprd = _p(out["sprd"], C.c_double)
    o.maxhz = _p(out["maxhz"], C.c_float)
    o.obscount = _p(out["obscount"], C.c_int64)
    o.lg_m = _p(out["lg_m"], C.c_int64)
    o.lg_obsloc = _p(out["lg_obsloc"], C.c_double)
    o.lg_hash = _p(out["lg_hash"], C.c_uint64)
    secs = np.zeros(5)
    o.seconds = _p(secs, C.c_double)
    nt = nthreads or min(os.cpu_count() or 1, 16)
    fn = lib().orc_analysis_f32 if f32 else lib().orc_analysis
    rc = fn(C.byref(b.s), _p(pts, C.c_int64), n, nt, C.byref(o))
    if rc != 0:
        raise RuntimeError("oracle analysis failed")
    out["points"] = pts
    out["setup_s"], out["loop_s"] = float(secs[0]), float(secs[1])
    # thread-seconds in the reference's timer phases
    out["obs_search_s"], out["letkf_core_s"], out["letkf_trans_s"] = (float(x) for x in secs[2:])
    return out


def selection(prob, ij, group=0):
    """Kept (index, rloc, dist) of one gridpoint and group, sorted by index."""
    b = _Bound(prob)
    cap = max(prob.nobs, 1)
    idx = np.zeros(cap, np.int32)
    rloc = np.zeros(cap, np.float32)
    dist = np.zeros(cap, np.float32)
    m = lib().orc_selection(C.byref(b.s), int(ij), int(group), _p(idx, C.c_int32),
                            _p(rloc, C.c_float), _p(dist, C.c_float), cap)
    if m < 0:
        raise RuntimeError("oracle selection failed")
    return idx[:m].copy(), rloc[:m].copy(), dist[:m].copy()


# --------------------------------------------------------------------------
# the steps either side of the path, FP32 in the reference's order
# --------------------------------------------------------------------------
def state_from_members(members):
    """letkf_state_init's mean, perturbations and spread
    (src/letkf_state.F90:326-334) in FP32: members (nij, ns, k) in file order
    (gridpoint fastest), summed member by member.  Returns state (k, ns, nij)
    perturbations, mean and sprd (ns, nij)."""
    members = np.asarray(members, np.float32)
    k = members.shape[2]
    acc = np.zeros(members.shape[:2], np.float32)
    for m in range(k):
        acc = acc + members[:, :, m]
    mu = acc / np.float32(k)
    pert = members - mu[:, :, None]
    dot = np.zeros_like(acc)
    for m in range(k):
        dot = dot + pert[:, :, m] * pert[:, :, m]
    sprd = np.sqrt(dot / np.float32(k - 1))
    # column-major like the reference's arrays (member fastest)
    return (np.asfortranarray(pert.transpose(2, 1, 0), np.float32),
            np.asfortranarray(mu.T), np.asfortranarray(sprd.T))


def obs_preprocess(hx_members, qc):
    """letkf_obs_read's obs_hx_mean, perturbations (src/letkf_obs.F90:459-466)
    and QC compaction (:479-489) in FP32: hx_members (k, n).  Returns the
    indices of the kept obs (qc <= 0, input order), hx (k, nkept) and
    hx_mean (nkept)."""
    hx_members = np.asarray(hx_members, np.float32)
    k = hx_members.shape[0]
    acc = np.zeros(hx_members.shape[1], np.float32)
    for m in range(k):
        acc = acc + hx_members[m]
    mu = acc / np.float32(k)
    pert = hx_members - mu[None, :]
    keep = np.nonzero(np.asarray(qc) <= 0)[0]
    return keep, np.asfortranarray(pert[:, keep]), np.ascontiguousarray(mu[keep])
