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

        # supersonic faces carry fluxes up to ~1e4 (rho un^4): at h = 1e-6 the
        # central difference's rounding error (eps |F| / h) reaches the 2e-7
        # bar; h = 1e-5 balances it against the O(h^2) truncation
        fd_check(f, vjp, x, 9, rng, h=1e-5)


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


def stage_fixture(seed, nx=7, ny=6):
    """One block with a generic smooth state (every primitive varies)."""
    import oracle
    rng = np.random.default_rng(seed)
    blocks = [dict(nx=nx, ny=ny, xmin=-0.5, xmax=0.6, ymin=-0.4, ymax=0.5, bc=[0, 0, 0, 0])]
    lx, ly = nx + 6, ny + 6
    a, b = np.meshgrid(np.arange(lx), np.arange(ly))
    def field(s):
        return s * np.sin(rng.uniform(0.3, 0.9) * a + rng.uniform(0.3, 0.9) * b + rng.uniform(0, 6))
    th = 1.0 + field(0.3)
    P = [1.0 + field(0.2), field(0.3), field(0.3), th, th + field(0.1), field(0.1), th + field(0.1),
         field(0.05), field(0.05)]
    U = np.empty((ly, lx, 9))
    Pc = np.stack(P, axis=-1)
    for j in range(ly):
        for i in range(lx):
            oracle.lib().orc_conservative(Pc[j, i].ctypes.data, U[j, i].ctypes.data)
    return blocks, U


@pytest.mark.parametrize("stage", [1, 2])
@pytest.mark.parametrize("lim", [0, 1, 2, 3, 4])
def test_stage_vjp_against_oracle_update(pc, stage, lim):
    """S^T of one stage (the GPU kernels' per-cell code, run on the CPU)
    against central differences of the oracle's Update (Main.c:297-396)."""
    import oracle
    pc.pc_stage_vjp.restype = ctypes.c_double
    pc.pc_stage_vjp.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_double, ctypes.c_double] + [ctypes.c_void_p] * 5 + \
                               [ctypes.c_double] * 3
    blocks, U = stage_fixture(stage * 10 + lim)
    _, U0 = stage_fixture(99)
    nx, ny = blocks[0]["nx"], blocks[0]["ny"]
    kn, dt, force = 0.02, 1e-3, 0.3
    orc = oracle.Oracle(blocks, kn=kn, force=force, limiter=lim)
    bx, by, dx, dy = (blocks[0]["xmax"] - blocks[0]["xmin"]) / nx, 0, 0, 0
    dx = (blocks[0]["xmax"] - blocks[0]["xmin"]) / nx
    dy = (blocks[0]["ymax"] - blocks[0]["ymin"]) / ny
    rng = np.random.default_rng(7)
    g = np.zeros_like(U)
    g[3:-3, 3:-3] = rng.normal(size=(ny, nx, 9))

    def fwd(Uin, Uin0):
        if stage == 1:
            orc.set([Uin], layer=0)
            oracle.lib().orc_update_block(orc.h, 0, dt, 0.0, 0)
            return orc.get(1)[0]
        orc.set([Uin], layer=1)
        orc.set([Uin0], layer=0)
        oracle.lib().orc_update_block(orc.h, 0, dt, 0.0, 1)
        return orc.get(2)[0]

    soa = lambda A: np.ascontiguousarray(A.transpose(2, 0, 1))
    gin = np.zeros((9, ny + 6, nx + 6))
    gin0 = np.zeros((9, ny + 6, nx + 6))
    gkn = pc.pc_stage_vjp(stage, lim, nx, ny, dx, dy, p(soa(U)), p(soa(U0)), p(soa(g)), p(gin), p(gin0),
                          dt, kn, force)
    gin = gin.transpose(1, 2, 0)
    gin0 = gin0.transpose(1, 2, 0)
    for trial in range(3):
        v = rng.normal(size=U.shape) * (np.abs(U) + 1e-2)
        h = 1e-6
        fd = (np.sum(g * fwd(U + h * v, U0)) - np.sum(g * fwd(U - h * v, U0))) / (2 * h)
        assert abs(np.sum(gin * v) - fd) <= 1e-6 * abs(fd) + 1e-9, ("in", trial, np.sum(gin * v), fd)
        if stage == 2:
            fd0 = (np.sum(g * fwd(U, U0 + h * v)) - np.sum(g * fwd(U, U0 - h * v))) / (2 * h)
            assert abs(np.sum(gin0 * v) - fd0) <= 1e-6 * abs(fd0) + 1e-9, ("in0", trial, np.sum(gin0 * v), fd0)
    orc2 = oracle.Oracle(blocks, kn=kn * (1 + 1e-5), force=force, limiter=lim)
    orc3 = oracle.Oracle(blocks, kn=kn * (1 - 1e-5), force=force, limiter=lim)
    vals = []
    for o in (orc2, orc3):
        orc = o
        vals.append(np.sum(g * fwd(U, U0)))
    fdk = (vals[0] - vals[1]) / (2e-5 * kn)
    assert abs(gkn - fdk) <= 1e-6 * abs(fdk) + 1e-9, (gkn, fdk)
