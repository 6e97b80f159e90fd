"""The engine's fast face flux and projection against the reference's, on the
CPU (no GPU).

mbm_phys.h evaluates the HLL flux (Flux.c:308-406) from the primitives in a
regrouped form: the shifted fluxes g = F - lam U of the two sides, combined
as (lamR g_L - lamL g_R) / (lamR - lamL). tests/native/libphyscheck.so (a
host build of the same device functions) is checked here against the
oracle's FnumHLLE restatement (oracle/mbm_oracle.c), which is bitwise equal
to the reference build on the fixture decks: subsonic faces, faces with one
side supersonic (lamL = 0 or lamR = 0), colliding supersonic faces
(lamL = lamR = 0, the mean of the two fluxes), both directions. The
projection's initial value of the new state (stage 1 takes the relaxed parts
off the conservative state directly) is checked against the oracle's
ProductionsDirect restatement.
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


INIT_TOL = 1e-12     # the parity bar (measured: <= 6.1e-14); a dropped or mis-signed
                     # relaxation term is O(1e-3) or more


@pytest.mark.parametrize("force", [0.0, 0.7])
@pytest.mark.parametrize("stage", [1, 2])
def test_projection_initial_value(pc, stage, force):
    """The projection's initial value of the new state against
    ProductionsDirect (Flux.c:230-267): flag 0 for stage 1, flag 2 of the
    stage input plus flag -1 of the step's input for stage 2, with and without
    a body force, collision numbers z = nu dt from 1e-5 to about 100. Stage 1
    without force takes the relaxed parts off U itself (no round trip through
    the primitives); the other cases go through the round trip."""
    O = oracle.lib()
    pc.pc_project.restype = _d
    pc.pc_project.argtypes = [_i, _p, _p, _i, _d, _d, _d, _p, _p]
    kn = 0.001
    orc = oracle.Oracle([dict(nx=4, ny=4, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, bc=[0, 0, 0, 0])],
                        kn=kn, force=force)
    rng = np.random.default_rng(11 + stage)

    def cons(P):
        U = np.empty(9)
        O.orc_conservative(ptr(P), ptr(U))
        return U

    def relax(U, dt, flag):
        o = np.empty(9)
        cp = O.orc_relax(orc.h, ptr(U), ptr(o), dt, flag)
        return o, cp

    worst = 0.0
    for _ in range(400):
        U, U0 = cons(prims(rng)), cons(prims(rng))
        dt = kn * 10.0 ** rng.uniform(-5.0, 2.0)
        P, I = np.empty(9), np.empty(9)
        cp = pc.pc_project(stage, ptr(U), ptr(U0), 0, dt, kn, force, ptr(P), ptr(I))
        if stage == 1:
            Ir, cpr = relax(U, dt, 0)
            assert abs(cp - cpr) <= 1e-14 + 1e-13 * abs(cpr), (cp, cpr)
        else:
            Ir = relax(U, dt, 2)[0] + relax(U0, dt, -1)[0]
            assert cp == 0.0
        scale = np.maximum(np.abs(Ir), 1e-3 * np.abs(Ir).max())
        worst = max(worst, float((np.abs(I - Ir) / scale).max()))
    assert worst <= INIT_TOL, worst
