"""The library's deterministic arithmetic (umd-letkf_amd/csrc/det_math.h)
built for the host gives the same bits as the oracle's copy: the basis of
bit-exact observation selection between the GPU and the oracle."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "det_math_host.cpp")
OUT = os.path.join(HERE, "native", "_build", "libdmh.so")


@pytest.fixture(scope="module")
def dmh():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    if not os.path.exists(OUT) or os.path.getmtime(OUT) < os.path.getmtime(SRC):
        subprocess.run(["/opt/rocm/bin/hipcc", "-O3", "-fPIC", "-shared", "-std=c++17",
                        "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", SRC, "-o", OUT],
                       check=True)
    L = C.CDLL(OUT)
    D, F, P = C.c_double, C.c_float, C.POINTER(C.c_double)
    for n, r, a in [("dmh_asin", D, [D]), ("dmh_exp_neg", D, [D]), ("dmh_thr2", D, [F]),
                    ("dmh_dist", F, [D]), ("dmh_gc", F, [F, F]), ("dmh_chord2", D, [P, P]),
                    ("dmh_mix64", C.c_uint64, [C.c_uint64])]:
        getattr(L, n).restype = r
        getattr(L, n).argtypes = a
    L.dmh_sincosd.argtypes = [D, P, P]
    L.dmh_xyz.argtypes = [F, F, P]
    return L


def test_bits_equal_oracle(dmh):
    g = np.random.default_rng(0)
    L = O.lib()
    L.orc_chord2.restype = C.c_double
    L.orc_chord2.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double)]
    for d in np.concatenate([g.uniform(-400, 400, 20000), np.arange(-360, 361, 0.5)]):
        d = float(np.float32(d))
        s1, c1 = O.sincosd(d)
        s2, c2 = C.c_double(), C.c_double()
        dmh.dmh_sincosd(d, C.byref(s2), C.byref(c2))
        assert s1 == s2.value and c1 == c2.value, d
    for x in g.uniform(0, 1, 20000):
        assert L.orc_asin(float(x)) == dmh.dmh_asin(float(x))
    for x in -g.exponential(10, 20000):
        assert L.orc_exp_neg(float(x)) == dmh.dmh_exp_neg(float(x))
    for r in g.uniform(1, 2.5e7, 5000).astype(np.float32):
        assert L.orc_thr2(float(r)) == dmh.dmh_thr2(float(r))
    for c2 in g.uniform(0, 4.0, 20000):
        assert L.orc_dist(float(c2)) == dmh.dmh_dist(float(c2))
    for z, l in zip(g.uniform(0, 3e6, 20000).astype(np.float32),
                    g.uniform(1e4, 1e6, 20000).astype(np.float32)):
        assert O.gc(z, l) == dmh.dmh_gc(float(z), float(l))
    a = np.zeros(3)
    b = np.zeros(3)
    pa = a.ctypes.data_as(C.POINTER(C.c_double))
    pb = b.ctypes.data_as(C.POINTER(C.c_double))
    for la, lo, la2, lo2 in g.uniform([-90, -180, -90, -180], [90, 540, 90, 540], (5000, 4)):
        x1 = O.xyz(np.float32(la), np.float32(lo))
        dmh.dmh_xyz(float(np.float32(la)), float(np.float32(lo)), pa)
        assert np.array_equal(x1, a)
        dmh.dmh_xyz(float(np.float32(la2)), float(np.float32(lo2)), pb)
        assert dmh.dmh_chord2(pa, pb) == L.orc_chord2(pa, pb)
    for x in g.integers(0, 2**62, 1000):
        assert L.orc_mix64(int(x)) == dmh.dmh_mix64(int(x))
