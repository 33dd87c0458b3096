"""In-bounds probe of the MFMA lane layouts the kernels assume (gfx950):
32x32x2 f32 and 32x32x16 bf16 (operands and 16-register accumulator),
16x16x32 bf16 (lane l holds k = 8(l/16)..+7 of row l%16; accumulator
register r of lane l holds row 4(l/16) + r, column l%16) and 16x16x4 f64
(accumulator register r of lane l holds row 4r + l/16)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "mfma_probe.hip")
OUT = os.path.join(HERE, "native", "_build", "libmfma_probe.so")


@pytest.fixture(scope="module")
def probe():
    if not os.path.exists(OUT):
        os.makedirs(os.path.dirname(OUT), exist_ok=True)
        subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-fPIC",
                        "-shared", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops", SRC, "-o", OUT], check=True)
    L = C.CDLL(OUT)
    L.probe_run.restype = C.c_int
    L.probe_run.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    return L


def _run(L, which, A, B, D):
    rc = L.probe_run(which, A.ctypes.data, B.ctypes.data, D.ctypes.data)
    assert rc == 0


def test_f32_32x32x2(probe):
    g = np.random.default_rng(0)
    A = g.integers(-4, 5, (32, 2)).astype(np.float32)
    B = g.integers(-4, 5, (2, 32)).astype(np.float32)
    D = np.zeros((32, 32), np.float32)
    _run(probe, 0, A, B, D)
    assert np.array_equal(D, A @ B)


def test_bf16_32x32x16(probe):
    g = np.random.default_rng(1)
    A = g.integers(-4, 5, (32, 16)).astype(np.float32)
    B = g.integers(-4, 5, (16, 32)).astype(np.float32)
    D = np.zeros((32, 32), np.float32)
    _run(probe, 1, A, B, D)
    assert np.array_equal(D, A @ B)


def test_bf16_16x16x32(probe):
    g = np.random.default_rng(3)
    A = g.integers(-4, 5, (16, 32)).astype(np.float32)
    B = g.integers(-4, 5, (32, 16)).astype(np.float32)
    D = np.zeros((16, 16), np.float32)
    _run(probe, 3, A, B, D)
    assert np.array_equal(D, A @ B)


def test_f64_16x16x4(probe):
    g = np.random.default_rng(2)
    A = g.standard_normal((16, 4))
    B = g.standard_normal((4, 16))
    D = np.zeros((16, 16))
    _run(probe, 2, A, B, D)
    assert np.allclose(D, A @ B, rtol=1e-14, atol=1e-14)
