"""Build libletkf_gpu.so in-tree with hipcc for gfx950, and the Fortran host.

Every .hip source under ``umd-letkf_amd/csrc`` is compiled to an object in
``umd-letkf_amd/_build`` (in parallel) and linked into
``umd-letkf_amd/letkf_amd/libletkf_gpu.so``.  ``build_fortran`` compiles the
iso_c_binding module, the solver host module and the stand-alone driver in
``umd-letkf_amd/fortran`` with amdflang into ``umd-letkf_amd/fortran/_build``.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(ROOT, "csrc")
BUILD = os.path.join(ROOT, "_build")
OUT = os.path.join(PKG, "libletkf_gpu.so")

ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 VALU ops stay off (v_pk_add/mul/fma_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-munsafe-fp-atomics",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


# per-source extra flags: the register-resident tridiagonalizations need their
# step loops fully unrolled (static register indices), beyond LLVM's default
# size limit for unroll pragmas
EXTRA = {f"tridiag_reg_kq{kq}.hip": ["-mllvm", "-pragma-unroll-threshold=2000000"]
         for kq in (32, 40, 48, 50, 56, 64)}
EXTRA.update({f"tridiag_pair_kq{kq}.hip": ["-mllvm", "-pragma-unroll-threshold=2000000"]
              for kq in (32, 40, 48, 50)})


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src).replace(".hip", ".o"))
    cmd = [hipcc(), *FLAGS, *EXTRA.get(os.path.basename(src), []), "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def _stale(src: str, newest_header: float) -> bool:
    obj = os.path.join(BUILD, os.path.basename(src).replace(".hip", ".o"))
    return (not os.path.exists(obj) or
            os.path.getmtime(obj) < max(os.path.getmtime(src), newest_header))


def build(verbose: bool = True, force: bool = True) -> str:
    """Compile and link the library; force=False recompiles only sources
    older than their object or than the newest header."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    headers = glob.glob(os.path.join(CSRC, "*.h")) + glob.glob(
        os.path.join(ROOT, "..", "include", "*.h"))
    newest = max((os.path.getmtime(h) for h in headers), default=0.0)
    todo = [s for s in srcs if force or _stale(s, newest)]
    if todo:
        with ThreadPoolExecutor(max_workers=min(12, len(todo))) as ex:
            list(ex.map(_compile, todo))
    objs = [os.path.join(BUILD, os.path.basename(s).replace(".hip", ".o")) for s in srcs]
    tmp = OUT + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,--no-undefined",
           "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


FORTRAN = os.path.join(ROOT, "fortran")
FBUILD = os.path.join(FORTRAN, "_build")
DRIVER = os.path.join(FBUILD, "letkf_gpu_driver")


def amdflang() -> str:
    for c in (os.environ.get("FC"), "/opt/rocm/bin/amdflang", shutil.which("amdflang")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("amdflang not found")


def build_fortran(verbose: bool = True) -> str:
    """Compile letkf_gpu_mod, letkf_gpu_loc, letkf_solver_gpu and
    letkf_gpu_driver; the driver links libletkf_gpu.so through an
    $ORIGIN-relative rpath.  (fortran/plugins/ holds a localizer plugin for
    the reference's own build; it uses the reference's modules and is not
    compiled here.)"""
    os.makedirs(FBUILD, exist_ok=True)
    objs = []
    for src in ("letkf_gpu_mod.F90", "letkf_gpu_loc.F90", "letkf_solver_gpu.F90",
                "letkf_gpu_driver.F90"):
        obj = os.path.join(FBUILD, src.replace(".F90", ".o"))
        cmd = [amdflang(), "-O2", "-c", os.path.join(FORTRAN, src), "-module-dir", FBUILD,
               "-I", FBUILD, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed for {src}:\n{r.stderr[-4000:]}")
        objs.append(obj)
    cmd = [amdflang(), *objs, "-L", PKG, "-lletkf_gpu",
           "-Wl,-rpath,$ORIGIN/../../letkf_amd", "-o", DRIVER]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran link failed:\n{r.stderr[-4000:]}")
    if verbose:
        print(f"built {DRIVER}")
    return DRIVER


if __name__ == "__main__":
    build()
    build_fortran()
