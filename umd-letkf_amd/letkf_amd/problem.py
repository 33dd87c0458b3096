"""In-memory description of one local analysis, in the reference's layout.

``Problem`` holds exactly what the reference's solver loop reads from its
module globals (src/letkf_solver.F90:150-389):

* state:   ``state_ij(k, ns, ij)`` background perturbations, ``mean_ij`` and
  ``sprd_ij`` (ns, ij), ``lat_ij``/``lon_ij``/``mask_ij``
  (src/letkf_state.F90:182-192);
* obs:     the QC-compacted ``obs_def`` fields, ``obs_hx`` perturbations
  (k, nobs) and ``obs_hx_mean`` (src/letkf_obs.F90:165-171);
* statedef: per-variable slab start, levels, ``ana_inc_max`` and
  ``ana_bounds`` (src/letkf_state.F90:83-98);
* the localizer and the inflation parameters.

Arrays are numpy, Fortran order, FP32 / int32 like the reference.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi


@dataclass
class LinearInterpLat:
    """``type: linearinterp_lat`` table as written in the YAML ``value:``."""
    lat: list
    radius: list

    def to_c(self) -> abi.LinearInterpLat:
        t = abi.LinearInterpLat()
        n = len(self.lat)
        if n != len(self.radius) or n > abi.MAX_INTERP:
            raise ValueError("bad linearinterp_lat table")
        t.npoints = n
        for i in range(n):
            t.lat[i] = float(self.lat[i])
            t.radius[i] = float(self.radius[i])
        return t


@dataclass
class Localizer:
    """A localizer instance (localization.class + its settings)."""
    cls: str = "LOC_NOVRT"
    hzloc: LinearInterpLat | None = None
    hzloc_prof: LinearInterpLat | None = None
    hzloc_sat: LinearInterpLat | None = None
    tloc_prof: float = -1.0
    tloc_sat: float = -1.0
    vtloc_surf: int = 0            # 0 none, 1 bkg_t
    bkg_t_delta: float = 0.0
    bkg_t_var_idx: int = 0         # 0-based slab start
    bkg_t_var_len: int = 1
    sat_obs: list = field(default_factory=list)
    sat_plats: list = field(default_factory=list)
    vt_scale: float = 0.0
    lev_depth: np.ndarray | None = None

    def maxgroups(self) -> int:
        c = self.cls.upper()
        if c == "LOC_NOVRT":
            return 1
        if c == "LOC_OCEAN":
            return 2 if self.vtloc_surf == 1 else 1
        if c == "LOC_VGAUSS":
            return int(len(self.lev_depth))
        raise ValueError(f"localizer class {self.cls} not found.")

    def to_c(self):
        """Return (LocDesc, keepalive) for the C ABI."""
        d = abi.LocDesc()
        c = self.cls.upper()
        d.cls = {"LOC_NOVRT": abi.LOC_NOVRT, "LOC_OCEAN": abi.LOC_OCEAN,
                 "LOC_VGAUSS": abi.LOC_VGAUSS}.get(c, -1)
        if self.hzloc is not None:
            d.hzloc = self.hzloc.to_c()
        if self.hzloc_prof is not None:
            d.hzloc_prof = self.hzloc_prof.to_c()
        if self.hzloc_sat is not None:
            d.hzloc_sat = self.hzloc_sat.to_c()
        d.tloc_prof = self.tloc_prof
        d.tloc_sat = self.tloc_sat
        d.vtloc_surf = self.vtloc_surf
        d.bkg_t_delta = self.bkg_t_delta
        d.bkg_t_var_idx = self.bkg_t_var_idx
        d.bkg_t_var_len = self.bkg_t_var_len
        if len(self.sat_obs) > abi.MAX_IDLIST or len(self.sat_plats) > abi.MAX_IDLIST:
            raise ValueError("too many sat_obs / sat_plats ids")
        d.n_sat_obs = len(self.sat_obs)
        for i, v in enumerate(self.sat_obs):
            d.sat_obs[i] = int(v)
        d.n_sat_plats = len(self.sat_plats)
        for i, v in enumerate(self.sat_plats):
            d.sat_plats[i] = int(v)
        d.vt_scale = self.vt_scale
        keep = None
        if self.lev_depth is not None:
            keep = np.ascontiguousarray(self.lev_depth, dtype=np.float32)
            d.nlev = int(keep.size)
            d.lev_depth = keep.ctypes.data_as(C.POINTER(C.c_float))
        return d, keep


@dataclass
class Problem:
    k: int
    ns: int
    nij: int
    state: np.ndarray        # (k, ns, nij) float32, F order
    mean: np.ndarray         # (ns, nij)
    sprd: np.ndarray         # (ns, nij)
    lat: np.ndarray          # (nij,)
    lon: np.ndarray
    mask: np.ndarray         # (nij,) int32
    olat: np.ndarray
    olon: np.ndarray
    ozdim: np.ndarray
    otime: np.ndarray
    oval: np.ndarray
    oerr: np.ndarray
    obsid: np.ndarray
    platid: np.ndarray
    hx: np.ndarray           # (k, nobs) perturbations
    hxmean: np.ndarray
    var_s: np.ndarray        # (nvars,) int32 0-based slab start
    var_lev: np.ndarray
    inc_max: np.ndarray      # (nvars,) float32
    bounds: np.ndarray       # (2, nvars) float32
    loc: Localizer
    rtps: float = 0.0
    rtpp: float = 0.0
    mul: float = 1.0
    meta: dict = field(default_factory=dict)

    @property
    def nobs(self) -> int:
        return int(self.olat.size)

    @property
    def nvars(self) -> int:
        return int(self.var_s.size)

    def validate(self):
        """Check every array against the layout the C ABI reads through raw
        pointers: shape, dtype, and Fortran (column-major) contiguity.  A
        C-ordered or wrongly typed array would otherwise be read as
        something else; raises ValueError naming the field."""
        k, ns, n, no, nv = self.k, self.ns, self.nij, self.nobs, self.nvars
        f32, i32 = np.float32, np.int32
        spec = [("state", (k, ns, n), f32), ("mean", (ns, n), f32), ("sprd", (ns, n), f32),
                ("lat", (n,), f32), ("lon", (n,), f32), ("mask", (n,), i32),
                ("olat", (no,), f32), ("olon", (no,), f32), ("ozdim", (no,), f32),
                ("otime", (no,), f32), ("oval", (no,), f32), ("oerr", (no,), f32),
                ("obsid", (no,), i32), ("platid", (no,), i32), ("hx", (k, no), f32),
                ("hxmean", (no,), f32), ("var_s", (nv,), i32), ("var_lev", (nv,), i32),
                ("inc_max", (nv,), f32), ("bounds", (2, nv), f32)]
        for name, shape, dt in spec:
            a = getattr(self, name)
            if not isinstance(a, np.ndarray):
                raise ValueError(f"Problem.{name}: not a numpy array")
            if a.shape != shape:
                raise ValueError(f"Problem.{name}: shape {a.shape}, expected {shape}")
            if a.dtype != dt:
                raise ValueError(f"Problem.{name}: dtype {a.dtype}, expected {np.dtype(dt)}")
            if not a.flags.f_contiguous:
                raise ValueError(f"Problem.{name}: not Fortran-contiguous (column-major, "
                                 "as the reference's arrays)")

    def subset_obs(self, sel):
        """New problem with only the observations ``sel`` (index array)."""
        import copy
        q = copy.copy(self)
        for name in ("olat", "olon", "ozdim", "otime", "oval", "oerr", "obsid",
                     "platid", "hxmean"):
            setattr(q, name, np.ascontiguousarray(getattr(self, name)[sel]))
        q.hx = np.asfortranarray(self.hx[:, sel])
        return q
