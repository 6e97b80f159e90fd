import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
DECKS = os.path.join(GOLDEN, "decks")
FIXTURES = os.path.join(GOLDEN, "fixtures")

FIELD_NAMES = ["rho", "rho_vx", "rho_vy", "E", "Fxx", "Fxy", "Fyy", "F1x", "F1y"]

# FP64 parity bound of the engine against the reference path, per field:
# max|a-b| / max|b| over all cells (Main.c state, ghosts included).
REL_LINF_TOL = 1e-12


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def deck_path(name):
    return os.path.join(DECKS, name + ".txt")


def load_fixture(name):
    z = np.load(os.path.join(FIXTURES, name + ".npz"), allow_pickle=False)
    return z["state"], z["log"]


def split_state(flat, blocks):
    """Flat reference dump -> list of (ny+6, nx+6, 9) arrays."""
    out, off = [], 0
    for b in blocks:
        n = (b["nx"] + 6) * (b["ny"] + 6) * 9
        out.append(flat[off:off + n].reshape(b["ny"] + 6, b["nx"] + 6, 9))
        off += n
    assert off == flat.size
    return out


def interior(a):
    return a[3:-3, 3:-3, :]


def field_scale(b_list):
    """Per-field max |value| over the finite values of a list of blocks."""
    return np.maximum.reduce([np.where(np.isfinite(b), np.abs(b), 0.0).reshape(-1, 9).max(axis=0)
                              for b in b_list])


def rel_linf(a_list, b_list, cells=None, scale=None):
    """Per-field relative L-inf difference over all blocks. Non-finite values
    must sit at the same places in both (NaN where the reference has NaN,
    +-inf with the same sign), else every field reports inf; the difference
    is then taken over the finite values, relative to max |b| per field or
    to the given per-field scale (for states that are mostly NaN, whose
    finite remainder can be cancellation residue far below the field's
    size)."""
    num = np.zeros(9)
    den = np.zeros(9)
    for a, b in zip(a_list, b_list):
        if cells is not None:
            a, b = cells(a), cells(b)
        a = a.reshape(-1, 9)
        b = b.reshape(-1, 9)
        fa, fb = np.isfinite(a), np.isfinite(b)
        if not (np.array_equal(fa, fb) and np.array_equal(np.isnan(a), np.isnan(b))
                and np.array_equal(a[~fa & ~np.isnan(a)], b[~fb & ~np.isnan(b)])):
            return np.full(9, np.inf)
        num = np.maximum(num, np.where(fb, np.abs(a - b), 0.0).max(axis=0))
        den = np.maximum(den, np.where(fb, np.abs(b), 0.0).max(axis=0))
    if scale is not None:
        den = np.asarray(scale, dtype=np.float64)
    return num / np.where(den > 0, den, 1.0)


def fixture_names():
    return sorted(f[:-4] for f in os.listdir(FIXTURES) if f.endswith(".npz"))


def scenario_of(name):
    return "ffs" if name.startswith("ffs") else "riemann2d"


@pytest.fixture(scope="session")
def mbm():
    import multiblockmomentsfv_amd as m
    m.lib()
    return m
