"""mbm_output_values_noalloc (static staging buffer, no device allocation)
against mbm_output_values, through the C ABI: bitwise equal values on two
calls in a row (the second on a smaller block), for Float32 and Float64, and
a clear rejection of a block whose output exceeds the staging buffer."""
import ctypes

import numpy as np
import pytest

from conftest import deck_path

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 * 1024 * 1024


def convert(L, eng, k, float32, fn):
    b = eng.blocks[k]
    v = np.full(((b["nx"] + 2) * (b["ny"] + 2), 10), np.nan,
                dtype=np.float32 if float32 else np.float64)
    rc = getattr(L, fn)(eng.h, k, int(float32), v.ctypes.data_as(ctypes.c_void_p))
    return rc, v


def test_noalloc_output_matches_allocating_output(mbm):
    L = mbm.lib()
    L.mbm_output_values_noalloc.restype = ctypes.c_int
    L.mbm_output_values_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_void_p]
    deck = mbm.read_deck(deck_path("squareSmall"))
    eng = mbm.Engine.from_deck(deck)
    eng.init()
    eng.steps(0.0, deck.tout[0], deck.cfl, maxsteps=7)
    sizes = [(b["nx"] + 2) * (b["ny"] + 2) for b in deck.blocks]
    big = int(np.argmax(sizes))
    small = int(np.argmin(sizes))
    assert sizes[small] < sizes[big]
    for float32 in (True, False):
        for k in (big, small):          # the second call converts a smaller block
            rc0, ref = convert(L, eng, k, float32, "mbm_output_values")
            rc1, got = convert(L, eng, k, float32, "mbm_output_values_noalloc")
            assert rc0 == 0 and rc1 == 0, mbm.lib().mbm_last_error()
            assert got.tobytes() == ref.tobytes(), (float32, k)
    eng.close()


def test_noalloc_output_rejects_blocks_beyond_the_buffer(mbm):
    L = mbm.lib()
    L.mbm_output_values_noalloc.restype = ctypes.c_int
    L.mbm_output_values_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_void_p]
    # 1024 x 1024: 1026^2 points x 10 doubles = 84 MB > 64 MiB; as float32 42 MB fits
    blocks = [dict(nx=1024, ny=1024, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, bc=[0, 0, 0, 0])]
    eng = mbm.Engine(blocks, kn=0.01)
    eng.init()
    assert 1026 * 1026 * 10 * 8 > SCRATCH_BYTES >= 1026 * 1026 * 10 * 4
    rc, _ = convert(L, eng, 0, False, "mbm_output_values_noalloc")
    assert rc == -1 and b"static staging buffer" in L.mbm_last_error()
    rc0, ref = convert(L, eng, 0, True, "mbm_output_values")
    rc1, got = convert(L, eng, 0, True, "mbm_output_values_noalloc")
    assert rc0 == 0 and rc1 == 0 and got.tobytes() == ref.tobytes()
    eng.close()
