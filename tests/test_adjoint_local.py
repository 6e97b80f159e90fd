"""Hand-written VJPs of the engine's local functions vs central differences.

tests/native/libphyscheck.so is a host build of the engine's own device
functions (multiblockmomentsfv_amd/csrc/mbm_phys.h, mbm_adj.h). For random
states each VJP g^T J is compared with central finite differences of the
forward function along random directions v: g^T (J v) ~ (f(x+hv)-f(x-hv))/2h.
"""
import ctypes
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIBP = os.path.join(HERE, "native", "libphyscheck.so")


@pytest.fixture(scope="module")
def pc():
    if not os.path.exists(LIBP):
        pytest.skip("tests/native/libphyscheck.so not built")
    L = ctypes.CDLL(LIBP)
    L.pc_limit.restype = ctypes.c_double
    L.pc_limit.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
    L.pc_limit_vjp.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                               ctypes.c_void_p]
    L.pc_face.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                          ctypes.c_void_p]
    L.pc_face_vjp.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.pc_relax.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
    L.pc_relax_vjp.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                               ctypes.c_void_p, ctypes.c_void_p]
    L.pc_damp.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
    L.pc_damp_dz.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
    return L


def p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def rand_prim(rng):
    th = rng.uniform(0.5, 2.0)
    return np.array([rng.uniform(0.5, 5), rng.normal(), rng.normal(), th,
                     th + 0.2 * rng.normal(), 0.2 * rng.normal(), th + 0.2 * rng.normal(),
                     0.3 * rng.normal(), 0.3 * rng.normal()])


def fd_check(f, vjp, x, nout, rng, h=1e-6, tol=2e-7):
    for _ in range(4):
        g = rng.normal(size=nout)
        v = rng.normal(size=x.size)
        jv = (f(x + h * v) - f(x - h * v)) / (2 * h)
        lhs = g @ jv
        rhs = vjp(x, g) @ v
        assert abs(lhs - rhs) <= tol * max(1.0, abs(lhs)), (lhs, rhs)


def test_primitive_vjp(pc):
    rng = np.random.default_rng(0)
    for _ in range(10):
        P = rand_prim(rng)
        U = np.empty(9)
        pc.pc_conservative(p(P), p(U))

        def f(u):
            out = np.empty(9)
            pc.pc_primitive(p(np.ascontiguousarray(u)), p(out))
            return out

        def vjp(u, g):
            out = np.zeros(9)
            pc.pc_primitive_vjp(p(np.ascontiguousarray(u)), p(g), p(out))
            return out

        fd_check(f, vjp, U, 9, rng)


def test_conservative_vjp(pc):
    rng = np.random.default_rng(1)
    for _ in range(10):
        P = rand_prim(rng)

        def f(x):
            out = np.empty(9)
            pc.pc_conservative(p(np.ascontiguousarray(x)), p(out))
            return out

        def vjp(x, g):
            out = np.zeros(9)
            pc.pc_conservative_vjp(p(np.ascontiguousarray(x)), p(g), p(out))
            return out

        fd_check(f, vjp, P, 9, rng)


@pytest.mark.parametrize("direction", [0, 1])
def test_face_vjp(pc, direction):
    rng = np.random.default_rng(2 + direction)
    for trial in range(20):
        L = rand_prim(rng)
        R = rand_prim(rng)
        if trial % 4 == 1:        # supersonic to the right: S- clipped to 0
            L[1 + direction] += 8.0
            R[1 + direction] += 8.0
        if trial % 4 == 2:        # supersonic to the left
            L[1 + direction] -= 8.0
            R[1 + direction] -= 8.0
        x = np.concatenate([L, R])

        def f(z):
            out = np.empty(9)
            pc.pc_face(direction, p(np.ascontiguousarray(z[:9])), p(np.ascontiguousarray(z[9:])),
                       0.37, p(out))
            return out

        def vjp(z, g):
            gl, gr = np.zeros(9), np.zeros(9)
            pc.pc_face_vjp(direction, p(np.ascontiguousarray(z[:9])), p(np.ascontiguousarray(z[9:])),
                           0.37, p(g), p(gl), p(gr))
            return np.concatenate([gl, gr])

        fd_check(f, vjp, x, 9, rng)


@pytest.mark.parametrize("lim", [0, 1, 2, 3, 4])
def test_limiter_vjp(pc, lim):
    rng = np.random.default_rng(10 + lim)
    for _ in range(40):
        d = rng.normal(size=2) * rng.choice([1e-3, 1.0, 10.0])
        if rng.uniform() < 0.3:
            d[1] = d[0] * rng.choice([0.3, 0.7, 1.5, 3.0])   # all superbee branches
        h = 1e-7 * (abs(d).max() + 1e-9)

        def f(x):
            return np.array([pc.pc_limit(lim, x[0], x[1])])

        def vjp(x, g):
            out = np.zeros(2)
            pc.pc_limit_vjp(lim, x[0], x[1], g[0], p(out))
            return out

        fd_check(f, vjp, d, 1, rng, h=h, tol=1e-6)


def test_relax_and_damping_vjp(pc):
    rng = np.random.default_rng(5)
    for _ in range(10):
        P = rand_prim(rng)
        d = rng.uniform(-0.5, 1.0, size=2)
        x = np.concatenate([P, d])

        def f(z):
            out = np.empty(9)
            pc.pc_relax(p(np.ascontiguousarray(z[:9])), z[9], z[10], p(out))
            return out

        def vjp(z, g):
            gp, gd = np.zeros(9), np.zeros(2)
            pc.pc_relax_vjp(p(np.ascontiguousarray(z[:9])), z[9], z[10], p(g), p(gp), p(gd))
            return np.concatenate([gp, gd])

        fd_check(f, vjp, x, 9, rng)
    for flag in (0, 1, 3):
        for z in (0.0, 0.3, 2.0, 25.0):
            h = 1e-6 * max(1.0, z)
            a, b, dz = np.empty(2), np.empty(2), np.empty(2)
            pc.pc_damp(flag, z + h, p(a))
            pc.pc_damp(flag, z - h, p(b))
            pc.pc_damp_dz(flag, z, p(dz))
            assert np.allclose((a - b) / (2 * h), dz, rtol=1e-6, atol=1e-9), (flag, z)
