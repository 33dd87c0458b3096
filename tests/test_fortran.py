"""Fortran host (iso_c_binding module + letkf_solver_gpu + driver).

CPU: the BIND(C) derived types have the C structs' sizes, and the driver
reports the library's error through letkf_gpu_errmsg when no GPU is present.
GPU: the Fortran host runs the analysis and matches the FP64 oracle like the
Python host does (bit-exact selection, 1e-10 relative)."""
import numpy as np
import pytest
import torch

import fortran_io as FIO
from helpers import oracle_run, rel
from letkf_amd import synth


def test_struct_layouts_match_c():
    assert FIO.struct_sizes() == FIO.ctypes_sizes()


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU error path")
def test_driver_reports_library_error_without_gpu(tmp_path):
    p = synth.random_problem(seed=1, nx=12, ny=6, nobs=50)
    r, out = FIO.run(p, str(tmp_path))
    assert r.returncode != 0 and out is None
    assert "letkf_solver_gpu_init: status" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["config1", "ocean", "vgauss"])
def test_fortran_host_matches_oracle(case, tmp_path):
    if case == "config1":
        p = synth.config1()
    elif case == "ocean":
        p = synth.ocean_vtloc()
    else:
        p = synth.random_problem(seed=71, k=24, ns=4, nobs=2000, loc_cls="LOC_VGAUSS", rtps=0.5)
    print("fortran host case", case, flush=True)
    r, out = FIO.run(p, str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    orc = oracle_run(p)
    assert np.array_equal(out["lg_selhash"].T, orc["lg_hash"])
    assert np.array_equal(out["lg_obscount"].T.astype(np.int64), orc["lg_m"])
    for f in ("mean", "sprd", "state"):
        assert rel(out[f], orc[f]) <= 1e-10, (f, rel(out[f], orc[f]))
    # the in-place FP32 arrays are the FP64 analysis rounded
    assert np.array_equal(out["state32"], out["state"].astype(np.float32))
    assert np.array_equal(out["mean32"], out["mean"].astype(np.float32))


def test_vgauss_vweight_matches_closed_form():
    """letkf_gpu_loc's vertical LOC_VGAUSS factor (used by the reference-side
    plugin, fortran/plugins/letkf_loc_vgauss.F90) is exp(-dz^2 / (2 s^2))
    rounded to REAL: 1, e^-1/2, e^-2 at dz = 0, s, 2s, and the oracle's
    weight for arbitrary arguments."""
    import math
    import subprocess
    for z, d, s, want in [(0.0, 0.0, 40.0, 1.0), (0.0, 40.0, 40.0, math.exp(-0.5)),
                          (10.0, 90.0, 40.0, math.exp(-2.0)), (5.0, 35.0, 300.0, None),
                          (-20.0, 1500.0, 300.0, None)]:
        r = subprocess.run([FIO.driver(), "--vweight", str(z), str(d), str(s)],
                           capture_output=True, text=True, check=True)
        got = float([ln for ln in r.stdout.splitlines() if ln.startswith("vweight")][0].split()[1])
        if want is None:
            want = math.exp(-((z - d) ** 2) / (2.0 * s * s))
        assert got == float(np.float32(want)), (z, d, s, got, want)
