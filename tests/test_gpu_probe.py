"""The FP64 issue-rate probe behind bench.py's roofline.valu: on the GPU it
returns a rate that a 16-lane SIMD can reach (at least 4 cycles per wave64
FP64 instruction, with some slack for the timing window) and a plausible
shader clock, for every instruction kind and wave count."""
import os
import sys

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("waves", [1, 2, 4])
@pytest.mark.parametrize("op", [0, 1, 2, 3])
def test_fp64_issue_probe_is_physical(op, waves):
    sys.path.insert(0, ROOT)
    import bench
    ns, cyc, ghz = bench.fp64_issue(op, waves, launches=20)
    assert 0.5 < ghz < 3.0
    assert 3.8 < cyc < 12.0, (ns, cyc, ghz)
    assert abs(ns * ghz - cyc) < 1e-9 * cyc


@pytest.mark.parametrize("fn", ["sqrt", "rcp"])
def test_device_sqrt_and_reciprocal_round_correctly(fn):
    """The solver's device square root (hardware rsq estimate, a Goldschmidt
    step and one Newton correction) and reciprocal (estimate and two Newton
    steps) against IEEE sqrt and 1/x on 2^22 log-uniform arguments in
    [1e-8, 1e8], plus the edges sqrtp gives its own values for (0 -> 0,
    negative -> NaN). Bar: within 1 ulp everywhere and correctly rounded in
    all but a millionth of the arguments (none, measured)."""
    import ctypes
    import numpy as np
    import torch  # noqa: F401  (same HIP runtime as the engine)
    path = os.path.join(ROOT, "multiblockmomentsfv_amd", "libmbmprobe.so")
    L = ctypes.CDLL(path)
    L.mbm_probe_math.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong]
    rng = np.random.default_rng(11)
    x = np.exp(rng.uniform(np.log(1e-8), np.log(1e8), 1 << 22))
    y = np.empty_like(x)
    assert L.mbm_probe_math(0 if fn == "sqrt" else 1, x.ctypes.data, y.ctypes.data, x.size) == 0
    ref = np.sqrt(x) if fn == "sqrt" else 1.0 / x
    ulp = np.abs(y - ref) / np.spacing(ref)
    print(f"\n{fn}: max {ulp.max():.3g} ulp, {(ulp > 0).sum()} of {x.size} not correctly rounded")
    assert ulp.max() <= 1.0 and (ulp > 0).mean() <= 1e-6
    if fn == "sqrt":
        e = np.array([0.0, -1.0, -1e-300])
        ye = np.empty_like(e)
        assert L.mbm_probe_math(0, e.ctypes.data, ye.ctypes.data, e.size) == 0
        assert ye[0] == 0.0 and np.isnan(ye[1:]).all()
