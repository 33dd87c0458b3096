"""Host-side solver object over the C ABI of libletkf_gpu.so.

``LetkfGpuSolver`` mirrors the reference's solver module
(src/letkf_solver.F90:36-38): ``init`` validates the inflation settings
like ``letkf_solver_init`` (:114-125) and uploads the analysis inputs,
``run`` performs ``letkf_solver_run`` (:150-389) on the GPU, ``final``
releases the device state.  Errors raise ``LetkfError`` carrying the
library's message (the reference aborts through ``letkf_mpi_abort``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .problem import Problem


class LetkfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"letkf_gpu error {code}: {msg}")
        self.code = code


class LetkfGpuSolver:
    def __init__(self, prob: Problem, precision: str = "fp64", path: str = "plain",
                 max_obs: int = 0, members=None, obs_members=None):
        """members: optional (nij, ns, k) full member values in file order,
        uploaded with letkf_gpu_set_state_members instead of prob's
        perturbations, mean and spread.  obs_members: optional (qc,
        hx_members (k, nobs)) uploaded with letkf_gpu_set_obs_members
        instead of prob's hx and hx mean."""
        self._members = members
        self._obs_members = obs_members
        self.lib = abi.load_library()
        self.prob = prob
        self.h = C.c_void_p()
        d = abi.Dims(prob.k, prob.ns, prob.nij, max_obs, abi.PRECISIONS[precision],
                     abi.PATH_PLAIN if path == "plain" else abi.PATH_BATCHED)
        rc = self.lib.letkf_gpu_create(C.byref(d), C.byref(self.h))
        if rc != 0:
            raise LetkfError(rc, self.lib.letkf_gpu_last_error(None).decode())
        self.precision = precision
        self.path = path
        try:
            self._upload()
        except Exception:
            self.final()
            raise

    # ------------------------------------------------------------ helpers
    def _chk(self, rc):
        if rc != 0:
            raise LetkfError(rc, self.lib.letkf_gpu_last_error(self.h).decode())

    def _upload(self):
        p = self.prob
        p.validate()
        self._upload_state()
        self._upload_obs()
        F, I = C.c_float, C.c_int32
        bounds = np.ascontiguousarray(p.bounds.T, np.float32)
        self._chk(self.lib.letkf_gpu_set_statevars(
            self.h, p.nvars, abi.ptr(p.var_s, I), abi.ptr(p.var_lev, I),
            abi.ptr(p.inc_max, F), abi.ptr(bounds, F)))
        desc, keep = p.loc.to_c()
        self._chk(self.lib.letkf_gpu_set_localizer(self.h, C.byref(desc)))
        del keep
        self._chk(self.lib.letkf_gpu_set_solver(self.h, p.rtps, p.rtpp, p.mul))

    def _upload_state(self):
        p = self.prob
        F, I = C.c_float, C.c_int32
        if self._members is None:
            self._chk(self.lib.letkf_gpu_set_state(
                self.h, abi.ptr(p.state, F), abi.ptr(p.mean, F), abi.ptr(p.sprd, F),
                abi.ptr(p.lat, F), abi.ptr(p.lon, F), abi.ptr(p.mask, I)))
        else:
            mem = np.asfortranarray(self._members, np.float32)
            assert mem.shape == (p.nij, p.ns, p.k)
            self._chk(self.lib.letkf_gpu_set_state_members(
                self.h, abi.ptr(mem, F), abi.ptr(p.lat, F), abi.ptr(p.lon, F),
                abi.ptr(p.mask, I)))

    def _upload_obs(self):
        p = self.prob
        F, I = C.c_float, C.c_int32
        if self._obs_members is None:
            self._chk(self.lib.letkf_gpu_set_obs(
                self.h, p.nobs, abi.ptr(p.olat, F), abi.ptr(p.olon, F), abi.ptr(p.ozdim, F),
                abi.ptr(p.otime, F), abi.ptr(p.oval, F), abi.ptr(p.oerr, F),
                abi.ptr(p.obsid, I), abi.ptr(p.platid, I), abi.ptr(p.hx, F),
                abi.ptr(p.hxmean, F)))
        else:
            qc, hxm = self._obs_members
            qc = np.ascontiguousarray(qc, np.int32)
            hxm = np.asfortranarray(hxm, np.float32)
            assert qc.shape == (p.nobs,) and hxm.shape == (p.k, p.nobs)
            self._chk(self.lib.letkf_gpu_set_obs_members(
                self.h, p.nobs, abi.ptr(p.olat, F), abi.ptr(p.olon, F), abi.ptr(p.ozdim, F),
                abi.ptr(p.otime, F), abi.ptr(p.oval, F), abi.ptr(p.oerr, F),
                abi.ptr(p.obsid, I), abi.ptr(p.platid, I), abi.ptr(qc, I), abi.ptr(hxm, F)))

    # ------------------------------------------------------------ API
    def cycle(self, prob: Problem, members=None, obs_members=None):
        """The next cycle's inputs on the same handle: a new background and
        new observations (prob's fields, or members / obs_members as in the
        constructor).  Grid, state variables, localizer and inflation stay
        as they were set; prob must have the same k, ns and nij."""
        assert (prob.k, prob.ns, prob.nij) == (self.prob.k, self.prob.ns, self.prob.nij)
        self.prob = prob
        self._members = members
        self._obs_members = obs_members
        self._upload_state()
        self._upload_obs()

    def run(self):
        self._chk(self.lib.letkf_gpu_run(self.h))

    def run_graph(self):
        """The analysis captured once as a HIP graph and replayed by later
        calls (letkf_gpu_run_graph); same results as run()."""
        self._chk(self.lib.letkf_gpu_run_graph(self.h))

    def maxgroups(self) -> int:
        g = C.c_int32()
        self._chk(self.lib.letkf_gpu_maxgroups(self.h, C.byref(g)))
        return g.value

    def state_f64(self, members: bool = True):
        """FP64 analysis (letkf_gpu_get_state_f64); members=False fetches
        only the mean and spread (state returned as None)."""
        p = self.prob
        st = np.zeros((p.k, p.ns, p.nij), order="F") if members else None
        mn = np.zeros((p.ns, p.nij), order="F")
        sp = np.zeros((p.ns, p.nij), order="F")
        D = C.c_double
        self._chk(self.lib.letkf_gpu_get_state_f64(
            self.h, abi.ptr(st, D) if members else None, abi.ptr(mn, D), abi.ptr(sp, D)))
        return st, mn, sp

    def state_members(self):
        """Analysis in file order: members (nij, ns, k), mean and spread
        (nij, ns), FP32 (letkf_gpu_get_state_members)."""
        p = self.prob
        mem = np.zeros((p.nij, p.ns, p.k), np.float32, order="F")
        mu = np.zeros((p.nij, p.ns), np.float32, order="F")
        sp = np.zeros((p.nij, p.ns), np.float32, order="F")
        F = C.c_float
        self._chk(self.lib.letkf_gpu_get_state_members(self.h, abi.ptr(mem, F),
                                                       abi.ptr(mu, F), abi.ptr(sp, F)))
        return mem, mu, sp

    def state_members_noalloc(self):
        """state_members through letkf_gpu_get_state_members_noalloc (no
        device allocation; the library's staging array)."""
        p = self.prob
        mem = np.zeros((p.nij, p.ns, p.k), np.float32, order="F")
        mu = np.zeros((p.nij, p.ns), np.float32, order="F")
        sp = np.zeros((p.nij, p.ns), np.float32, order="F")
        F = C.c_float
        self._chk(self.lib.letkf_gpu_get_state_members_noalloc(
            self.h, abi.ptr(mem, F), abi.ptr(mu, F), abi.ptr(sp, F)))
        return mem, mu, sp

    def state_f32(self):
        p = self.prob
        st = np.zeros((p.k, p.ns, p.nij), np.float32, order="F")
        mn = np.zeros((p.ns, p.nij), np.float32, order="F")
        sp = np.zeros((p.ns, p.nij), np.float32, order="F")
        F = C.c_float
        self._chk(self.lib.letkf_gpu_get_state(self.h, abi.ptr(st, F), abi.ptr(mn, F),
                                               abi.ptr(sp, F)))
        return st, mn, sp

    def diag(self):
        p = self.prob
        G = self.maxgroups()
        out = dict(maxhz=np.zeros(p.nij, np.float32), obscount=np.zeros(p.nij, np.float32),
                   lg_obscount=np.zeros((p.nij, G), np.float32, order="F"),
                   lg_obsloc=np.zeros((p.nij, G), np.float32, order="F"),
                   lg_selhash=np.zeros((p.nij, G), np.uint64, order="F"))
        d = abi.Diag(abi.ptr(out["maxhz"], C.c_float), abi.ptr(out["obscount"], C.c_float),
                     abi.ptr(out["lg_obscount"], C.c_float),
                     abi.ptr(out["lg_obsloc"], C.c_float),
                     abi.ptr(out["lg_selhash"], C.c_uint64))
        self._chk(self.lib.letkf_gpu_get_diag(self.h, C.byref(d)))
        return out

    def selection(self, ij: int, group: int = 0):
        cap = max(self.prob.nobs, 1)
        idx = np.zeros(cap, np.int32)
        w = np.zeros(cap, np.float32)
        m = C.c_int64()
        self._chk(self.lib.letkf_gpu_get_selection(self.h, int(ij), int(group),
                                                   abi.ptr(idx, C.c_int32),
                                                   abi.ptr(w, C.c_float), cap, C.byref(m)))
        n = min(m.value, cap)
        return idx[:n].copy(), w[:n].copy()

    def timing(self) -> dict:
        t = abi.Timing()
        self._chk(self.lib.letkf_gpu_get_timing(self.h, C.byref(t)))
        return {f: getattr(t, f) for f, _ in abi.Timing._fields_}

    def autotune(self, repeats: int = 2) -> dict:
        """Time the launch choices on the current inputs and keep the
        fastest (letkf_gpu_autotune)."""
        t = abi.Tuning()
        self._chk(self.lib.letkf_gpu_autotune(self.h, int(repeats), C.byref(t)))
        return {f: getattr(t, f) for f, _ in abi.Tuning._fields_}

    def tuning(self) -> dict:
        t = abi.Tuning()
        self._chk(self.lib.letkf_gpu_get_tuning(self.h, C.byref(t)))
        return {f: getattr(t, f) for f, _ in abi.Tuning._fields_}

    def set_tuning(self, order: int, bucket_bits: int = 0, tridiag_reg: int = 2,
                   overlap: int = 1, chunk_log2: int = 20, xcd: int = 4, tile: int = 1, **_):
        t = abi.Tuning(order, bucket_bits, tridiag_reg, 0, 0.0, 0.0, overlap, chunk_log2, xcd,
                       tile)
        self._chk(self.lib.letkf_gpu_set_tuning(self.h, C.byref(t)))

    def final(self):
        if self.h:
            self.lib.letkf_gpu_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.final()

    def __del__(self):
        try:
            self.final()
        except Exception:
            pass
