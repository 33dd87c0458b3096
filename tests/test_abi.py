"""CPU checks of the boundary: the library loads and exports every symbol
declared in include/letkf_gpu.h, and the ctypes mirror matches the header.
No compute call is made (no GPU here)."""
import ctypes as C
import os
import subprocess
import re

from letkf_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "letkf_gpu.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(letkf_gpu_[a-z0-9_]+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    lib = abi.load_library()
    names = declared()
    assert len(names) >= 16
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(abi.EXPORTED)
    assert lib.letkf_gpu_version().decode().startswith("letkf_gpu")


def test_struct_layouts(tmp_path):
    """The ctypes mirrors have the C compiler's sizes and field offsets for
    every struct of include/letkf_gpu.h."""
    exe = str(tmp_path / "abi_sizes")
    subprocess.run(["gcc", "-O0", os.path.join(ROOT, "tests", "native", "abi_sizes.c"),
                    "-o", exe], check=True)
    out = dict(ln.rsplit(" ", 1) for ln in
               subprocess.run([exe], capture_output=True, text=True, check=True)
               .stdout.splitlines())
    mirror = {"letkf_gpu_dims": abi.Dims, "letkf_linearinterp_lat": abi.LinearInterpLat,
              "letkf_loc_desc": abi.LocDesc, "letkf_gpu_diag": abi.Diag,
              "letkf_gpu_timing": abi.Timing, "letkf_gpu_tuning": abi.Tuning}
    for name, cls in mirror.items():
        assert C.sizeof(cls) == int(out[name]), name
    for key, val in out.items():
        if "." in key:
            struct, field = key.split(".")
            assert getattr(mirror[struct], field).offset == int(val), key


def test_no_cpu_fallback_in_package():
    """The product path must not route through the oracle or a CPU loop."""
    pkg = os.path.join(ROOT, "umd-letkf_amd", "letkf_amd")
    for f in os.listdir(pkg):
        if f.endswith(".py"):
            src = open(os.path.join(pkg, f)).read()
            assert "import oracle" not in src and "from oracle" not in src, f


def test_problem_validate_rejects_wrong_layout_and_dtype():
    """Arrays reach the C ABI as raw pointers, so a C-ordered multi-
    dimensional array or a wrong dtype must be refused before upload."""
    import numpy as np
    import pytest

    from letkf_amd import synth
    p = synth.random_problem(seed=1, nx=12, ny=6, nobs=40, ns=2)
    p.validate()
    for name, bad in [("state", np.ascontiguousarray(p.state)),
                      ("hx", np.ascontiguousarray(p.hx)),
                      ("mean", np.asfortranarray(p.mean, np.float64)),
                      ("sprd", np.ascontiguousarray(p.sprd)),
                      ("olat", p.olat.astype(np.float64)),
                      ("mask", p.mask.astype(np.int64)),
                      ("inc_max", p.inc_max.astype(np.float64)),
                      ("bounds", np.ascontiguousarray(p.bounds)),
                      ("oerr", p.oerr[:-1])]:
        q = synth.random_problem(seed=1, nx=12, ny=6, nobs=40, ns=2)
        setattr(q, name, bad)
        with pytest.raises(ValueError, match=name):
            q.validate()
    with pytest.raises(TypeError):
        abi.ptr(np.zeros(4), C.c_float)
    with pytest.raises(ValueError):
        abi.ptr(np.zeros((4, 4), np.float32)[:, ::2], C.c_float)
