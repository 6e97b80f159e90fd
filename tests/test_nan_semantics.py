"""Non-finite data through the explicit update, on the CPU (no GPU).

The reference propagates NaN through its own expression structure: its
limiters compare against NaN (sgn(NaN) = 1, min(NaN, b) = b: Limiter.c:26-38,
tools.h:3-17), its flux divides by the face density after Conservative() and
rotates components into the face frame (Flux.c:79-172: a NaN times 0 is NaN,
and a zero face density is 0/0 in every component), and its production-CFL
fold max(CFL_prod, new) restarts after a NaN (Main.c:325-326). The engine's
fast forms reproduce that on finite data only, so the stage kernel switches a
row to the reference forms of mbm_phys.h (limit_ref, reconstruct_ref,
face_ref) when a primitive in its stencil is not finite, and decides exact
zero face densities in the reference's arithmetic.

Checked here with tests/native/libphyscheck.so, a host build of the same
device functions: each reference form against the oracle (bitwise, NaN
masks included), the fast projection's NaN pattern, and a host emulation of
one stage with the kernel's switching rule against the oracle's TimeUpdate
on small_average (the reference's own NaN fixture) and on states with
injected NaNs and negative temperatures. The GPU kernel itself is checked
against the oracle in test_gpu_parity.py.
"""
import ctypes
import itertools
import os

import numpy as np
import pytest

from conftest import REL_LINF_TOL, deck_path, field_scale, load_fixture, rel_linf, split_state

import oracle

LIBP = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "libphyscheck.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIBP), reason="tests/native/libphyscheck.so not built")

_d, _i, _p = ctypes.c_double, ctypes.c_int, ctypes.c_void_p


@pytest.fixture(scope="module")
def pc():
    L = ctypes.CDLL(LIBP)
    L.pc_limit_ref.restype = _d
    L.pc_limit_ref.argtypes = [_i, _d, _d]
    L.pc_face_ref.restype = _d
    L.pc_face_ref.argtypes = [_i, _p, _p, _d, _p]
    L.pc_project.restype = _d
    L.pc_project.argtypes = [_i, _p, _p, _i, _d, _d, _d, _p, _p]
    L.pc_emu_stage.restype = _d
    L.pc_emu_stage.argtypes = [_i, _i, _i, _i, _d, _d, _p, _p, _p, _d, _d, _d]
    return L


def ptr(a):
    return a.ctypes.data_as(_p)


def same(x, y):
    """bitwise equal, NaN equal to NaN"""
    return (np.isnan(x) and np.isnan(y)) or x == y


def test_limit_ref_is_the_reference_limiter(pc):
    O = oracle.lib()
    vals = [1.5, -0.7, 0.0, -0.0, np.nan, 2.0, -3.0, np.inf, -np.inf, 1e-300]
    for lim in range(5):
        for a, b in itertools.product(vals, vals):
            assert same(pc.pc_limit_ref(lim, a, b), O.orc_limit(lim, a, b)), (lim, a, b)


def generic_prims(rng):
    return np.array([rng.uniform(0.5, 2), rng.uniform(-.5, .5), rng.uniform(-.5, .5), rng.uniform(.5, 2),
                     rng.uniform(.5, 2), rng.uniform(-.3, .3), rng.uniform(.5, 2), rng.uniform(-.3, .3),
                     rng.uniform(-.3, .3)])


@pytest.mark.parametrize("direction", [0, 1])
def test_face_ref_has_the_reference_nan_pattern(pc, direction):
    """Every subset of NaN components on one side (with a few on the other),
    a negative temperature and a zero density: same NaN components as
    FnumHLLE, finite values to rounding, same face speed."""
    O = oracle.lib()
    rng = np.random.default_rng(3)
    worst = 0.0
    for case in range(3):
        L0, R0 = generic_prims(rng), generic_prims(rng)
        if case == 1:
            L0[3] = -0.2
        if case == 2:
            R0[0] = 0.0
        for ml in range(512):
            for mr in ((0, 1, 40, 511) if ml % 7 else range(0, 512, 5)):
                L, R = L0.copy(), R0.copy()
                L[[k for k in range(9) if ml >> k & 1]] = np.nan
                R[[k for k in range(9) if mr >> k & 1]] = np.nan
                F, Fr = np.empty(9), np.empty(9)
                s = pc.pc_face_ref(direction, ptr(L), ptr(R), 1.0, ptr(F))
                sr = O.orc_hll(ptr(L), ptr(R), 1.0, direction + 1, ptr(Fr))
                assert np.array_equal(np.isnan(F), np.isnan(Fr)), (case, ml, mr, F, Fr)
                f = np.isfinite(Fr)
                if f.any():
                    worst = max(worst, float((np.abs(F - Fr)[f] / np.maximum(np.abs(Fr[f]), 1e-300)).max()))
                assert same(s, sr) or abs(s - sr) <= 1e-15 * abs(sr)
    assert worst <= 1e-13, worst


@pytest.mark.parametrize("stage", [1, 2])
def test_projection_nan_pattern(pc, stage):
    """The fast projection (primitive, relaxation, initial value) gives the
    reference's NaN components for every subset of NaN inputs and for a
    negative temperature, in the stage input and in the step's input (stage
    2's identity relaxation is the round trip of Flux.c:230-267 when the data
    are not finite; so is the relaxation of DATA)."""
    O = oracle.lib()
    orc = oracle.Oracle([dict(nx=4, ny=4, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, bc=[0, 0, 0, 0])],
                        kn=0.001)
    dt = 1e-4
    rng = np.random.default_rng(1)

    def relax(U, flag):
        o = np.empty(9)
        O.orc_relax(orc.h, ptr(np.ascontiguousarray(U)), ptr(o), dt, flag)
        return o

    def prim(U):
        q = np.empty(9)
        O.orc_primitive(ptr(np.ascontiguousarray(U)), ptr(q))
        return q

    def cons(P):
        U = np.empty(9)
        O.orc_conservative(ptr(P), ptr(U))
        return U

    for trial in range(4):
        P0, Q0 = generic_prims(rng), generic_prims(rng)
        if trial == 2:
            P0[3] = -0.3
        if trial == 3:
            Q0[3] = -0.3
        base, base0 = cons(P0), cons(Q0)
        for mask in range(512):
            # stage 2 takes the DATA part off DATA itself unless that is not
            # finite: every single NaN component of DATA, and some sets
            for mask0 in ((0,) if stage == 1 else (0, 1, 2, 4, 8, 16, 32, 64, 128, 256, 272, 511)):
                U, U0 = base.copy(), base0.copy()
                U[[k for k in range(9) if mask >> k & 1]] = np.nan
                U0[[k for k in range(9) if mask0 >> k & 1]] = np.nan
                P, I = np.empty(9), np.empty(9)
                pc.pc_project(stage, ptr(U), ptr(U0), 0, dt, 0.001, 0.0, ptr(P), ptr(I))
                if stage == 1:
                    Pr, Ir = prim(relax(U, 1)), relax(U, 0)
                else:
                    Pr, Ir = prim(relax(U, 3)), relax(U, 2) + relax(U0, -1)
                assert np.array_equal(np.isnan(P), np.isnan(Pr)), (trial, mask)
                assert np.array_equal(np.isnan(I), np.isnan(Ir)), (trial, mask, mask0)


def emulate_time_update(pc, mbm, deck, limiter, data, dt):
    """TimeUpdate from the host build of the kernel's functions and its
    switching rule, with the engine's ghost map on the CPU."""
    lim = mbm.LIMITERS[limiter]
    geo = [deck.geom(k) for k in range(deck.nblocks)]
    tmp, new = [a.copy() for a in data], [a.copy() for a in data]
    c1 = c2 = 0.0
    for k, b in enumerate(deck.blocks):
        c = pc.pc_emu_stage(1, lim, b["nx"], b["ny"], geo[k][2], geo[k][3], ptr(data[k]), ptr(data[k]),
                            ptr(tmp[k]), dt, deck.kn, deck.force)
        c1 = c if c > c1 else c1
    mbm.host_set_boundary(deck.blocks, tmp, kn=deck.kn, limiter=limiter)
    for k, b in enumerate(deck.blocks):
        c = pc.pc_emu_stage(2, lim, b["nx"], b["ny"], geo[k][2], geo[k][3], ptr(tmp[k]), ptr(data[k]),
                            ptr(new[k]), dt, deck.kn, deck.force)
        c2 = c if c > c2 else c2
    mbm.host_set_boundary(deck.blocks, new, kn=deck.kn, limiter=limiter)
    return tmp, new, max(c1, c2)


def test_small_average_emulated_matches_the_reference(pc, mbm):
    """Four steps of small_average (all 9216 values NaN at the end) against
    the oracle, and the first against the compiled reference's fixture."""
    deck = mbm.read_deck(deck_path("small_average"))
    orc = oracle.Oracle(deck.blocks, kn=deck.kn, force=deck.force, limiter=deck.limiter)
    orc.init()
    data = [a.copy() for a in orc.get(0)]
    state, log = load_fixture("small_average")
    scale = None
    for step in range(4):
        tmp, new, cfl = emulate_time_update(pc, mbm, deck, deck.limiter, data, 1e-4)
        c_ref = orc.time_update(1e-4, 0.0)
        if scale is None:
            # the field sizes of the first step; later the finite remainder
            # of a mostly-NaN field is cancellation residue (~1e-9)
            scale = field_scale(orc.get(2))
        for layer, mine in ((1, tmp), (2, new)):
            err = rel_linf(mine, orc.get(layer), scale=scale)
            assert err.max() <= REL_LINF_TOL, (step, layer, err)
        assert abs(cfl - c_ref) <= REL_LINF_TOL * c_ref
        if step == 0:
            err = rel_linf(new, split_state(state, deck.blocks))
            assert err.max() <= REL_LINF_TOL
            assert sum(int(np.isnan(a).sum()) for a in new) == 2655
            assert abs(cfl - log[0, 0]) <= REL_LINF_TOL * log[0, 0]
        orc.commit()
        data = [a.copy() for a in orc.get(0)]


@pytest.mark.parametrize("limiter", ["none", "average", "minmod", "superbee", "harmonic"])
def test_injected_nonfinite_states_emulated(pc, mbm, limiter):
    deck = mbm.read_deck(deck_path("small_kn01"))
    for seed in range(6):
        rng = np.random.default_rng(seed)
        orc = oracle.Oracle(deck.blocks, kn=deck.kn, force=deck.force, limiter=limiter)
        orc.init()
        orc.steps(deck.tout[0], deck.cfl, maxsteps=3)
        data = [a.copy() for a in orc.get(0)]
        for a in data:
            ny, nx, _ = a.shape
            for _ in range(3):
                a[rng.integers(0, ny), rng.integers(0, nx), rng.integers(0, 9)] = np.nan
            if seed % 2:
                b, c = rng.integers(3, ny - 3), rng.integers(3, nx - 3)
                a[b, c, 3] = 0.1 * a[b, c, 3] - 1.0
        orc.set(data, 0)
        tmp, new, cfl = emulate_time_update(pc, mbm, deck, limiter, data, 1e-4)
        c_ref = orc.time_update(1e-4, 0.0)
        for layer, mine in ((1, tmp), (2, new)):
            err = rel_linf(mine, orc.get(layer))
            assert err.max() <= REL_LINF_TOL, (seed, layer, err)
        assert abs(cfl - c_ref) <= REL_LINF_TOL * c_ref, (seed, cfl, c_ref)
