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
