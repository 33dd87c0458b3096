"""Problem / result files of the Fortran host (umd-letkf_amd/fortran/
letkf_gpu_driver.F90); the layout is documented in the driver's header."""
import ctypes as C
import os
import subprocess

import numpy as np

from letkf_amd import abi, build
from letkf_amd.abi import PRECISIONS


def driver() -> str:
    if not os.path.exists(build.DRIVER):
        build.build_fortran(verbose=False)
    return build.DRIVER


def write_problem(path, p, precision="fp64", route="plain"):
    desc, keep = p.loc.to_c()
    lev = np.zeros(0, np.float32) if p.loc.lev_depth is None else \
        np.asarray(p.loc.lev_depth, np.float32)
    f32 = lambda a: np.asarray(a, np.float32).ravel(order="F").tobytes()
    i32 = lambda a: np.asarray(a, np.int32).ravel(order="F").tobytes()
    with open(path, "wb") as f:
        f.write(np.array([p.k, p.ns, p.nij, p.nobs], np.int64).tobytes())
        f.write(np.array([PRECISIONS[precision], 0 if route == "plain" else 1, p.nvars,
                          lev.size], np.int32).tobytes())
        f.write(np.array([p.rtps, p.rtpp, p.mul], np.float32).tobytes())
        f.write(bytes(desc))
        f.write(lev.tobytes())
        for a in (p.state, p.mean, p.sprd, p.lat, p.lon):
            f.write(f32(a))
        f.write(i32(p.mask))
        f.write(i32(p.var_s) + i32(p.var_lev) + f32(p.inc_max) + f32(p.bounds))
        for a in (p.olat, p.olon, p.ozdim, p.otime, p.oval, p.oerr):
            f.write(f32(a))
        f.write(i32(p.obsid) + i32(p.platid) + f32(p.hx) + f32(p.hxmean))
    del keep


def read_result(path, p):
    k, ns, n = p.k, p.ns, p.nij
    with open(path, "rb") as f:
        g = int(np.frombuffer(f.read(4), np.int32)[0])
        rd = lambda dt, shape: np.frombuffer(
            f.read(int(np.prod(shape)) * np.dtype(dt).itemsize), dt).reshape(shape, order="F")
        out = dict(maxgroups=g)
        out["state32"] = rd(np.float32, (k, ns, n))
        out["mean32"] = rd(np.float32, (ns, n))
        out["sprd32"] = rd(np.float32, (ns, n))
        out["state"] = rd(np.float64, (k, ns, n))
        out["mean"] = rd(np.float64, (ns, n))
        out["sprd"] = rd(np.float64, (ns, n))
        out["lg_obscount"] = rd(np.float32, (n, g))
        out["lg_selhash"] = rd(np.int64, (n, g)).view(np.uint64)
    return out


def run(p, workdir, precision="fp64", route="plain", timeout=300):
    fin = os.path.join(workdir, "problem.bin")
    fout = os.path.join(workdir, "result.bin")
    write_problem(fin, p, precision, route)
    r = subprocess.run([driver(), fin, fout], capture_output=True, text=True, timeout=timeout)
    return r, (read_result(fout, p) if r.returncode == 0 else None)


def struct_sizes():
    r = subprocess.run([driver(), "--sizes"], capture_output=True, text=True, check=True)
    line = [ln for ln in r.stdout.splitlines() if ln.startswith("sizes")][0]
    return [int(x) for x in line.split()[1:]]


def ctypes_sizes():
    return [C.sizeof(x) for x in (abi.Dims, abi.LinearInterpLat, abi.LocDesc, abi.Diag,
                                  abi.Timing, abi.Tuning)]
