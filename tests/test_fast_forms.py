"""The engine's fast face flux against the reference's, on the CPU (no GPU).

mbm_phys.h evaluates the HLL flux (Flux.c:308-406) from the primitives in a
regrouped form: the shifted fluxes g = F - lam U of the two sides, combined
as (lamR g_L - lamL g_R) / (lamR - lamL). tests/native/libphyscheck.so (a
host build of the same device functions) is checked here against the
oracle's FnumHLLE restatement (oracle/mbm_oracle.c), which is bitwise equal
to the reference build on the fixture decks: subsonic faces, faces with one
side supersonic (lamL = 0 or lamR = 0), colliding supersonic faces
(lamL = lamR = 0, the mean of the two fluxes), both directions.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle

LIBP = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "libphyscheck.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIBP), reason="tests/native/libphyscheck.so not built")

_d, _i, _p = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
FACE_TOL = 1e-12     # the parity bar; a wrong term of the regrouped flux is O(1e-2)


@pytest.fixture(scope="module")
def pc():
    L = ctypes.CDLL(LIBP)
    L.pc_face.restype = None
    L.pc_face.argtypes = [_i, _p, _p, _d, _p]
    return L


def ptr(a):
    return a.ctypes.data_as(_p)


def prims(rng, un=None):
    """primitives {rho, v1, v2, theta, T11, T12, T22, q1, q2} with T11 and T22
    near theta (a relaxed state), optionally a given normal velocity"""
    th = rng.uniform(0.3, 2.0)
    P = np.array([rng.uniform(0.2, 5.0), rng.uniform(-0.8, 0.8), rng.uniform(-0.8, 0.8), th,
                  th * rng.uniform(0.7, 1.3), rng.uniform(-0.3, 0.3) * th, th * rng.uniform(0.7, 1.3),
                  rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5)])
    return P


def sound(P):
    return 2.0 * np.sqrt(5.0 * P[3] / 3.0)


@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("regime", ["subsonic", "left_supersonic", "right_supersonic", "colliding"])
def test_fast_face_matches_the_reference_flux(pc, direction, regime):
    O = oracle.lib()
    rng = np.random.default_rng(7 + direction)
    n = 1 + direction          # normal velocity component
    worst = 0.0
    for _ in range(400):
        L, R = prims(rng), prims(rng)
        if regime == "left_supersonic":        # un_L - c_L >= 0: lamL = 0
            L[n] = sound(L) * rng.uniform(1.0, 2.0)
            R[n] = sound(R) * rng.uniform(0.2, 2.0)
        elif regime == "right_supersonic":     # un_R + c_R <= 0: lamR = 0
            R[n] = -sound(R) * rng.uniform(1.0, 2.0)
            L[n] = -sound(L) * rng.uniform(0.2, 2.0)
        elif regime == "colliding":            # lamL = lamR = 0
            L[n] = sound(L) * rng.uniform(1.0, 2.0)
            R[n] = -sound(R) * rng.uniform(1.0, 2.0)
        dtdx = rng.uniform(0.01, 1.0)
        F, Fr = np.empty(9), np.empty(9)
        pc.pc_face(direction, ptr(L), ptr(R), dtdx, ptr(F))
        O.orc_hll(ptr(L), ptr(R), 1.0, direction + 1, ptr(Fr))
        Fr *= dtdx
        # per component, relative to the face's flux scale (a component can
        # cancel to far below the others)
        scale = np.maximum(np.abs(Fr), 1e-3 * np.abs(Fr).max())
        worst = max(worst, float((np.abs(F - Fr) / scale).max()))
    assert worst <= FACE_TOL, worst
