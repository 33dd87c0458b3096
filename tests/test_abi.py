"""CPU checks of the boundary: the library loads and exports every symbol
declared in include/letkf_gpu.h, and the ctypes mirror matches the header.
No compute call is made (no GPU here)."""
import ctypes as C
import os
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


def test_struct_layouts():
    # sizes follow the header (natural C alignment)
    assert C.sizeof(abi.Dims) == 40
    assert C.sizeof(abi.LinearInterpLat) == 4 + 8 * abi.MAX_INTERP
    assert C.sizeof(abi.Diag) == 5 * 8
    assert C.sizeof(abi.Timing) == 5 * 8 + 4 * 8


def test_no_cpu_fallback_in_package():
    """The product path must not route through the oracle or a CPU loop."""
    pkg = os.path.join(ROOT, "umd-letkf_amd", "letkf_amd")
    for f in os.listdir(pkg):
        if f.endswith(".py"):
            src = open(os.path.join(pkg, f)).read()
            assert "import oracle" not in src and "from oracle" not in src, f
