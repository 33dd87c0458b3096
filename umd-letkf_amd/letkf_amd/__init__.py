"""MI355X-native local analysis of UMD-LETKF (host side).

The analysis itself runs in the HIP library ``libletkf_gpu.so`` behind the C
ABI of ``include/letkf_gpu.h``; this package is its Python host: the ctypes
binding (``abi``), the problem description in the reference's layout
(``problem``), the solver object that mirrors
``letkf_solver_init/run/final`` (``solver``) and the seeded synthetic
analyses used by the tests and the benchmark (``synth``).  The YAML
configuration stays with the reference's Fortran host.
"""
from .problem import LinearInterpLat, Localizer, Problem  # noqa: F401
