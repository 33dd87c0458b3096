"""Rational approximation of x^(-1/2) used by the batched path to apply
A^(-1/2) (umd-letkf_amd/csrc/ratnodes.h): Jacobi elliptic functions against
scipy, the approximation error on dense samples against the closed form, and
the matrix version against an eigendecomposition."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest
import scipy.special as sp

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "ratnodes_host.cpp")
HDR = os.path.join(HERE, "..", "umd-letkf_amd", "csrc", "ratnodes.h")
OUT = os.path.join(HERE, "native", "_build", "libratnodes.so")


@pytest.fixture(scope="module")
def rn():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    if (not os.path.exists(OUT) or os.path.getmtime(OUT) < os.path.getmtime(SRC)
            or os.path.getmtime(OUT) < os.path.getmtime(HDR)):
        subprocess.run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", SRC, "-o", OUT],
                       check=True)
    L = C.CDLL(OUT)
    D, P = C.c_double, C.POINTER(C.c_double)
    L.rn_ellipk.restype = D
    L.rn_ellipk.argtypes = [D]
    L.rn_ellipj.argtypes = [D, D, D, P, P, P]
    L.rn_nodes.argtypes = [D, C.c_int, P, P]
    L.rn_choose.restype = C.c_int
    L.rn_choose.argtypes = [D, D, C.c_int, P, P]
    return L


def _nodes(rn, kappa, n):
    s = np.zeros(n)
    w = np.zeros(n)
    rn.rn_nodes(kappa, n, s.ctypes.data_as(C.POINTER(C.c_double)),
                w.ctypes.data_as(C.POINTER(C.c_double)))
    return s, w


@pytest.mark.parametrize("kappa", [1.5, 10.0, 1e3, 1e6, 1e8])  # scipy switches to an approximation for 1 - m < 1e-9
def test_ellipj_and_ellipk_against_scipy(rn, kappa):
    kc = 1.0 / np.sqrt(kappa)
    k = np.sqrt((1 - kc) * (1 + kc))
    K = rn.rn_ellipk(kc)
    assert abs(K - sp.ellipkm1(kc * kc)) <= 1e-13 * K
    sn, cn, dn = C.c_double(), C.c_double(), C.c_double()
    for u in np.linspace(0.0, K, 17)[:-1]:
        rn.rn_ellipj(u, k, kc, C.byref(sn), C.byref(cn), C.byref(dn))
        ssn, scn, sdn, _ = sp.ellipj(u, k * k)
        assert abs(sn.value - ssn) <= 1e-12
        assert abs(cn.value - scn) <= 1e-12
        assert abs(dn.value - sdn) <= 1e-12


@pytest.mark.parametrize("kappa", [2.0, 50.0, 1e3, 1e5, 1e8])
def test_error_decays_and_meets_tolerance(rn, kappa):
    x = np.concatenate([[1.0, kappa], np.exp(np.random.default_rng(1).uniform(0, np.log(kappa), 20000))])
    errs = []
    for n in (4, 8, 12, 16, 20):
        s, w = _nodes(rn, kappa, n)
        approx = (w[None, :] / (x[:, None] + s[None, :])).sum(1)
        errs.append(np.abs(approx * np.sqrt(x) - 1).max())
    # geometric decay at about the predicted rate exp(-2 pi^2 n / log(16 kappa))
    rate = 2 * np.pi ** 2 / np.log(16 * kappa)
    for n, e in zip((4, 8, 12, 16, 20), errs):
        assert e <= max(20 * np.exp(-rate * n), 1e-14), (n, e)
    s = np.zeros(40)
    w = np.zeros(40)
    n = rn.rn_choose(kappa, 1e-12, 40, s.ctypes.data_as(C.POINTER(C.c_double)),
                     w.ctypes.data_as(C.POINTER(C.c_double)))
    assert 2 <= n <= 40
    approx = (w[None, :n] / (x[:, None] + s[None, :n])).sum(1)
    assert np.abs(approx * np.sqrt(x) - 1).max() <= 1.5e-12


def test_matrix_inverse_sqrt(rn):
    g = np.random.default_rng(3)
    k = 50
    m = k - 1.0
    Y = g.normal(size=(k, 400)) * np.linspace(0.01, 3, 400)
    A = m * np.eye(k) + Y @ Y.T                 # spectrum in [k-1, ...]
    lam, V = np.linalg.eigh(A)
    kappa = 2.0 ** np.ceil(np.log2(lam.max() / m))
    s, w = np.zeros(40), np.zeros(40)
    n = rn.rn_choose(kappa, 1e-12, 40, s.ctypes.data_as(C.POINTER(C.c_double)),
                     w.ctypes.data_as(C.POINTER(C.c_double)))
    v = g.normal(size=k)
    ref = V @ ((V.T @ v) / np.sqrt(lam))
    got = sum(np.sqrt(m) * w[j] * np.linalg.solve(A + m * s[j] * np.eye(k), v) for j in range(n))
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()
