"""Discrete adjoint of the Heun step on the GPU (no reference counterpart).

Checked against central finite differences of the GPU forward map with the dt
schedule held fixed, and against central differences of the CPU oracle's
step (the reference algorithm). Base states are generic (smooth fields in
every primitive), see multiblockmomentsfv_amd/adjoint.py.
"""
import math

import numpy as np
import pytest

from conftest import deck_path

import oracle

pytestmark = pytest.mark.gpu

try:
    import torch
    HAVE_TORCH = torch.cuda.is_available()
except Exception:   # pragma: no cover
    HAVE_TORCH = False


def generic_state(mbm, eng, deck, seed=0, amp=1.0):
    """Layer-layout state: tanh temperature jump + smooth fields in every primitive."""
    from multiblockmomentsfv_amd.adjoint import smooth_fields
    rng = np.random.default_rng(seed)
    U = eng.layer_get(0)
    cells = eng.layer_to_blocks(U)
    for b, k in zip(deck.blocks, range(deck.nblocks)):
        bx, by, dx, dy = deck.geom(k)
        X, Y = np.meshgrid(bx + dx * np.arange(3, b["nx"] + 3), by + dy * np.arange(3, b["ny"] + 3))
        dr, v1, v2, d11, d22, t12, q1, q2 = smooth_fields(X, Y, rng, amp)
        # off-centre: a profile symmetric about the block interfaces x=0, y=0 makes
        # the theta jump across them vanish to round-off, a limiter kink
        th = 1.0 + 0.5 * (1 - np.tanh((np.hypot(X - 0.137, Y + 0.071) - 0.4) / 0.1))
        r = 1.0 + dr
        t11, t22 = th + d11, th + d22
        E = r * (0.5 * (v1 ** 2 + v2 ** 2) + 1.5 * th)
        comps = [r, r * v1, r * v2, E, r * (t11 + v1 * v1), r * (t12 + v1 * v2), r * (t22 + v2 * v2),
                 q1 + E * v1 + r * (t11 * v1 + t12 * v2), q2 + E * v2 + r * (t12 * v1 + t22 * v2)]
        for m, c in enumerate(comps):
            cells[k][3:-3, 3:-3, m] = c
    U = eng.blocks_to_layer(cells, U)
    eng.layer_set(0, U)
    eng.set_boundary(0, 0.0)
    return eng.layer_get(0)


def interior_mask(eng):
    cells = []
    for b in eng.blocks:
        a = np.zeros((b["ny"] + 6, b["nx"] + 6, 9))
        a[3:-3, 3:-3] = 1.0
        cells.append(a)
    return eng.blocks_to_layer(cells)


def run(eng, U, dts, kn=None, stride=None):
    if kn is not None:
        eng.set_kn(kn)
    eng.layer_set(0, U)
    eng.set_boundary(0, 0.0)
    eng.adj_forward(tend=math.inf, cflreg=1.0, dts=dts, stride=stride)
    return eng.layer_get(0)


def smooth_direction(eng, deck, U, rng):
    """A smooth perturbation of every component (low Fourier modes on the
    cell centres, relative amplitude ~10 % of the state), layer layout."""
    from multiblockmomentsfv_amd.adjoint import smooth_fields
    cells = eng.layer_to_blocks(U)
    out = []
    for k, b in enumerate(deck.blocks):
        bx, by, dx, dy = deck.geom(k)
        X, Y = np.meshgrid(bx + dx * np.arange(b["nx"] + 6), by + dy * np.arange(b["ny"] + 6))
        f = smooth_fields(X, Y, rng, 1.0) + smooth_fields(X, Y, rng, 1.0)[:1]
        out.append(np.stack(f, axis=-1) * (np.abs(cells[k]) + 1e-2))
    return eng.blocks_to_layer(out)


def gpu_to_blocks(eng, U):
    return eng.layer_to_blocks(U)


@pytest.mark.parametrize("name", ["box_small", "lshape_small", "periodic", "mixed_bc"])
@pytest.mark.parametrize("limiter", ["harmonic", "minmod", "superbee", "average", "none"])
def test_one_step_adjoint_matches_cpu_reference(mbm, name, limiter):
    """GPU adjoint of one step == the CPU evaluation of the same adjoint code
    (tests/adjoint_ref.py, itself checked against central differences of the
    oracle), cell by cell."""
    one_step_vs_cpu_reference(mbm, name, limiter, 0.0)


@pytest.mark.parametrize("name", ["box_small", "mixed_bc"])
@pytest.mark.parametrize("limiter", ["harmonic", "minmod"])
def test_one_step_adjoint_with_body_force(mbm, name, limiter):
    """With a body force the two stages' initial-value VJPs are taken
    separately on the GPU (without one they are merged); both against the CPU
    reference, which always takes them separately."""
    one_step_vs_cpu_reference(mbm, name, limiter, 0.3)


def one_step_vs_cpu_reference(mbm, name, limiter, force):
    import adjoint_ref as ar
    deck = mbm.read_deck(deck_path(name))
    lim = mbm.LIMITERS[limiter]
    eng = mbm.Engine(deck.blocks, kn=0.01, force=force, limiter=limiter, scenario="tempjump")
    eng.init()
    U0 = generic_state(mbm, eng, deck)
    mask = interior_mask(eng)
    rng = np.random.default_rng(11)
    w = rng.normal(size=U0.size) * mask
    dt = 1e-3
    run(eng, U0, [dt])
    gI, gkn = eng.adj_backward(w)
    D = gpu_to_blocks(eng, U0)
    W = gpu_to_blocks(eng, w)
    orc = oracle.Oracle(deck.blocks, kn=0.01, force=force, limiter=limiter, scenario="tempjump")
    fmap = ar.fill_transpose_map(deck.blocks)
    geoms = [deck.geom(k) for k in range(deck.nblocks)]
    gD, gk = ar.step_adjoint(orc, deck.blocks, geoms, lim, D, W, dt, 0.01, force, fmap)
    gD = ar.fill_T(gD, fmap)
    G = gpu_to_blocks(eng, gI)
    scale = max(np.abs(a).max() for a in gD)
    worst = (0.0, None)
    for k, (a, b) in enumerate(zip(G, gD)):
        d = np.abs(a - b)
        j = np.unravel_index(np.argmax(d), d.shape)
        if d[j] > worst[0]:
            worst = (float(d[j]), (k, j))
    # The two forward states differ in the last bits (the GPU's reciprocals and
    # fused products); the harmonic limiter's derivative varies on the scale
    # of its 1e-10 regularisation, so near-zero slopes amplify those bits.
    tol = 1e-9 if limiter in ("none", "average") else 1e-7
    assert worst[0] <= tol * scale, (worst, scale)
    assert abs(gkn - gk) <= tol * abs(gk), (gkn, gk)


def test_backward_repeats_and_release(mbm):
    """A backward pass leaves the recorded run intact: a second one with the
    same cotangent gives the same gradient bit for bit, also after a second
    forward run of another length has reused the pooled layers; after
    mbm_adj_release a backward pass is refused."""
    deck = mbm.read_deck(deck_path("mixed_bc"))
    eng = mbm.Engine(deck.blocks, kn=0.01, scenario="tempjump")
    eng.init()
    U0 = generic_state(mbm, eng, deck)
    w = np.random.default_rng(3).normal(size=U0.size) * interior_mask(eng)
    dts = [1e-3] * 9
    run(eng, U0, dts, stride=3)
    g1, k1 = eng.adj_backward(w)
    g2, k2 = eng.adj_backward(w)
    assert np.array_equal(g1, g2) and k1 == k2
    run(eng, U0, [1e-3] * 16, stride=2)        # more checkpoints and segment layers
    eng.adj_backward(w)
    run(eng, U0, dts, stride=3)
    g3, k3 = eng.adj_backward(w)
    assert np.array_equal(g1, g3) and k1 == k3
    eng.adj_release()
    with pytest.raises(mbm.MbmError):
        eng.adj_backward(w)


def test_zero_step_adjoint_is_fill_transpose(mbm):
    """<fill^T g, u> == <g, fill(u)> with the GPU fill and its transpose."""
    for name in ("box_small", "lshape_small", "periodic", "mixed_bc"):
        deck = mbm.read_deck(deck_path(name))
        eng = mbm.Engine(deck.blocks, kn=0.01)
        eng.init()
        rng = np.random.default_rng(5)
        mask = interior_mask(eng)
        u = rng.normal(size=mask.size) * mask
        zero = np.zeros_like(u)
        f_u = run(eng, u, [])
        f_0 = run(eng, zero, [])            # Dirichlet ghosts are constants: use differences
        # a cotangent on every cell (ghosts included), zero between the cells
        g = eng.blocks_to_layer(eng.layer_to_blocks(rng.normal(size=mask.size)))
        run(eng, u, [])
        gT, _ = eng.adj_backward(g)
        lhs = gT @ u
        rhs = g @ (f_u - f_0)
        assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(rhs)), (name, lhs, rhs)


@pytest.mark.parametrize("limiter", ["average", "none"])
def test_one_step_adjoint_vs_fd(mbm, limiter):
    deck = mbm.read_deck(deck_path("box_small"))
    eng = mbm.Engine(deck.blocks, kn=0.01, limiter=limiter, scenario="tempjump")
    eng.init()
    U0 = generic_state(mbm, eng, deck)
    mask = interior_mask(eng)
    rng = np.random.default_rng(1)
    w = rng.normal(size=U0.size) * mask
    dts = [2e-3]
    run(eng, U0, dts)
    gI, gkn = eng.adj_backward(w)
    assert np.all(gI[mask == 0] == 0)
    for trial in range(3):
        v = rng.normal(size=U0.size) * mask * (np.abs(U0) + 1e-2)
        h = 1e-6
        fd = (w @ run(eng, U0 + h * v, dts) - w @ run(eng, U0 - h * v, dts)) / (2 * h)
        assert abs(gI @ v - fd) <= 1e-6 * abs(fd), (trial, gI @ v, fd)
    kn, h = 0.01, 1e-5
    fdk = (w @ run(eng, U0, dts, kn=kn * (1 + h)) - w @ run(eng, U0, dts, kn=kn * (1 - h))) / (2 * h * kn)
    assert abs(gkn - fdk) <= 1e-6 * abs(fdk), (gkn, fdk)


@pytest.mark.parametrize("limiter", ["average", "none"])
def test_one_step_adjoint_vs_oracle_step(mbm, limiter):
    """Dot product with central differences of the CPU oracle's step (the
    reference algorithm), same state, same dt."""
    deck = mbm.read_deck(deck_path("lshape_small"))
    eng = mbm.Engine(deck.blocks, kn=deck.kn, limiter=limiter, scenario="tempjump")
    eng.init()
    U0 = generic_state(mbm, eng, deck, seed=3)
    mask = interior_mask(eng)
    rng = np.random.default_rng(2)
    w = rng.normal(size=U0.size) * mask
    dt = 1e-3
    run(eng, U0, [dt])
    gI, _ = eng.adj_backward(w)
    orc = oracle.Oracle(deck.blocks, kn=deck.kn, limiter=limiter, scenario="tempjump")
    def to_blocks(U):
        return eng.layer_to_blocks(U)

    def oracle_step(U):
        orc.set(to_blocks(U), layer=0)
        orc.set_boundary(0, 0.0)
        orc.time_update(dt, 0.0)
        return to_blocks_inv(orc.get(2))

    def to_blocks_inv(blocks):
        return eng.blocks_to_layer(blocks)

    for trial in range(3):
        v = rng.normal(size=U0.size) * mask * (np.abs(U0) + 1e-2)
        h = 1e-6
        fd = (w @ oracle_step(U0 + h * v) - w @ oracle_step(U0 - h * v)) / (2 * h)
        assert abs(gI @ v - fd) <= 1e-6 * abs(fd), (trial, gI @ v, fd)


@pytest.mark.parametrize("limiter", ["harmonic", "average"])
def test_multi_step_adjoint_with_checkpoints_vs_fd(mbm, limiter):
    """Checkpoint stride does not change the gradient (bitwise); for the
    smooth limiter the gradient matches central differences of the run."""
    deck = mbm.read_deck(deck_path("small_kn01"))
    eng = mbm.Engine(deck.blocks, kn=deck.kn, limiter=limiter, scenario="tempjump")
    eng.init()
    U0 = generic_state(mbm, eng, deck, seed=5)
    eng.layer_set(0, U0)
    n, t = eng.adj_forward(tend=deck.tout[0], cflreg=deck.cfl, maxsteps=23, stride=5)
    assert n == 23
    dts = eng.adj_schedule()
    mask = interior_mask(eng)
    rng = np.random.default_rng(4)
    w = rng.normal(size=U0.size) * mask
    run(eng, U0, dts, stride=5)
    g5, gkn5 = eng.adj_backward(w)
    run(eng, U0, dts, stride=23)
    g23, gkn23 = eng.adj_backward(w)
    assert np.array_equal(g5, g23) and gkn5 == gkn23      # checkpointing changes nothing
    if limiter == "harmonic":
        # the harmonic limiter has kinks where a slope crosses zero: white-noise
        # directions at h = 1e-6 measure secants there. Smooth directions (low
        # Fourier modes, scaled by the state) at h = 1e-7 move no slope across
        # a kink on this generic state, as in the 4x512^2 benchmark check.
        for trial in range(3):
            v = smooth_direction(eng, deck, U0, rng) * mask
            # the two runs are differenced cell by cell before the weighting:
            # w @ U(+) - w @ U(-) would cancel two sums of size ~300 at h ~ 1e-7
            h = 1e-7
            fd = w @ (run(eng, U0 + h * v, dts) - run(eng, U0 - h * v, dts)) / (2 * h)
            assert abs(g5 @ v - fd) <= 1e-6 * abs(fd), (trial, g5 @ v, fd)
        return
    for trial in range(2):
        v = rng.normal(size=U0.size) * mask * (np.abs(U0) + 1e-2)
        h = 1e-6
        fd = (w @ run(eng, U0 + h * v, dts) - w @ run(eng, U0 - h * v, dts)) / (2 * h)
        assert abs(g5 @ v - fd) <= 1e-5 * abs(fd), (trial, g5 @ v, fd)
    kn, h = deck.kn, 1e-5
    fdk = (w @ run(eng, U0, dts, kn=kn * (1 + h)) - w @ run(eng, U0, dts, kn=kn * (1 - h))) / (2 * h * kn)
    assert abs(gkn5 - fdk) <= 1e-5 * abs(fdk), (gkn5, fdk)


@pytest.mark.skipif(not HAVE_TORCH, reason="torch with a GPU is required")
def test_torch_autograd_temperature_inversion(mbm):
    from multiblockmomentsfv_amd.adjoint import TemperatureInversion
    prob = TemperatureInversion(deck_path("small_kn01"), kn=0.01, generic=1.0, seed=1,
                                limiter="average")
    prob.record_schedule(tend=0.05, maxsteps=40)
    th0 = [t.clone().requires_grad_(True) for t in prob.theta0]
    kn = torch.tensor(0.01, dtype=torch.float64, requires_grad=True)
    J = prob.objective(th0, kn, stride=7)
    J.backward()
    rng = np.random.default_rng(0)
    for trial in range(3):
        dirs = [torch.tensor(rng.normal(size=t.shape), dtype=torch.float64, device=t.device)
                for t in prob.theta0]
        h = 1e-6
        with torch.no_grad():
            jp = prob.objective([t + h * d for t, d in zip(prob.theta0, dirs)], kn.detach())
            jm = prob.objective([t - h * d for t, d in zip(prob.theta0, dirs)], kn.detach())
        fd = float((jp - jm) / (2 * h))
        ad = float(sum((g.grad * d).sum() for g, d in zip(th0, dirs)))
        assert abs(ad - fd) <= 1e-5 * abs(fd), (trial, ad, fd)
    h = 1e-5
    with torch.no_grad():
        jp = prob.objective(prob.theta0, torch.tensor(0.01 * (1 + h), dtype=torch.float64))
        jm = prob.objective(prob.theta0, torch.tensor(0.01 * (1 - h), dtype=torch.float64))
    fdk = float((jp - jm) / (2 * h * 0.01))
    assert abs(float(kn.grad) - fdk) <= 1e-5 * abs(fdk), (float(kn.grad), fdk)


@pytest.mark.skipif(not HAVE_TORCH, reason="torch with a GPU is required")
def test_two_objectives_on_one_engine(mbm):
    """The engine keeps one recorded run: a second objective (and a no-grad
    evaluation) recorded between the first one's forward and backward must
    not change the first gradient."""
    from multiblockmomentsfv_amd.adjoint import TemperatureInversion
    prob = TemperatureInversion(deck_path("small_kn01"), kn=0.01, generic=1.0, seed=2,
                                limiter="average")
    prob.record_schedule(tend=0.05, maxsteps=20)
    kn = torch.tensor(0.01, dtype=torch.float64)

    def grads(th):
        return [t.grad.clone() for t in th]

    ref = [t.clone().requires_grad_(True) for t in prob.theta0]
    prob.objective(ref, kn).backward()
    g_ref = grads(ref)

    a = [t.clone().requires_grad_(True) for t in prob.theta0]
    b = [(t * 1.05).requires_grad_(True) for t in prob.theta0]
    Ja = prob.objective(a, kn)
    Jb = prob.objective(b, kn, stride=3)
    prob.final_theta([t * 0.97 for t in prob.theta0], kn)   # another recording, no grad
    Ja.backward()
    Jb.backward()
    for x, y in zip(grads(a), g_ref):
        assert torch.equal(x, y)
    bb = [(t * 1.05).requires_grad_(True) for t in prob.theta0]
    prob.objective(bb, kn).backward()
    for x, y in zip(grads(b), grads(bb)):
        assert torch.equal(x, y)


def test_layer_buffers_are_checked(mbm):
    """Whole-layer copies and the adjoint take only float64 buffers of
    exactly layer_elems() elements (the count follows the tile height)."""
    deck = mbm.read_deck(deck_path("small_kn01"))
    eng = mbm.Engine(deck.blocks, kn=deck.kn, scenario="tempjump")
    eng.init()
    n = eng.layer_elems()
    with pytest.raises(mbm.MbmError):
        eng.layer_get(0, np.empty(n - 1))
    with pytest.raises(mbm.MbmError):
        eng.layer_set(0, np.zeros(n + 8))
    if HAVE_TORCH:
        with pytest.raises(mbm.MbmError):
            eng.layer_set(0, torch.zeros(n, dtype=torch.float32, device="cuda"))
        with pytest.raises(mbm.MbmError):
            eng.layer_get(0, torch.zeros(2 * n, dtype=torch.float64, device="cuda")[::2])
    U = eng.layer_get(0)
    eng.set_tile_rows(7)
    assert eng.layer_elems() != n
    with pytest.raises(mbm.MbmError):
        eng.layer_set(0, U)
    # the C entry point checks the count itself
    import ctypes
    assert mbm.lib().mbm_layer_get(eng.h, 0, U.ctypes.data_as(ctypes.c_void_p), n) != 0
    V = eng.layer_get(0)
    assert eng.layer_to_blocks(V)[0].shape == (deck.blocks[0]["ny"] + 6, deck.blocks[0]["nx"] + 6, 9)
    eng.close()


@pytest.mark.skipif(not HAVE_TORCH, reason="torch with a GPU is required")
def test_adjoint_at_the_configured_shape(mbm):
    """The adjoint workload itself: 4x512^2 square4 topology, Reg13, Kn 0.01,
    harmonic, t_end 0.3 (about 2200 steps, resident recording), gradient of
    sum (theta(t_end) - target)^2 w.r.t. theta0 and Kn, against central
    differences along 4 smooth directions (h = 1e-8) and in Kn (relative
    h = 1e-5): relative agreement <= 1e-6."""
    from multiblockmomentsfv_amd.adjoint import TemperatureInversion, smooth_fields
    import os
    path = os.path.join(os.path.dirname(mbm.HERE), "multiblockmomentsfv_amd", "decks", "square4_512.txt")
    prob = TemperatureInversion(path, kn=0.01, seed=0, generic=0.1)
    nsteps, t = prob.record_schedule(tend=0.3)
    assert nsteps > 2000 and abs(t - 0.3) < 1e-12
    th0 = [t.clone().requires_grad_(True) for t in prob.theta0]
    kn = torch.tensor(prob.kn, dtype=torch.float64, requires_grad=True)
    prob.objective(th0, kn).backward()
    rng = np.random.default_rng(0)
    h = 1e-8
    for trial in range(4):
        dirs = []
        for k, th in enumerate(prob.theta0):
            bx, by, dx, dy = prob.deck.geom(k)
            b = prob.deck.blocks[k]
            X, Y = np.meshgrid(bx + dx * np.arange(3, b["nx"] + 3), by + dy * np.arange(3, b["ny"] + 3))
            dirs.append(torch.tensor(smooth_fields(X, Y, rng, 10.0)[0], dtype=torch.float64,
                                     device=th.device))
        tp = prob.final_theta([t + h * d for t, d in zip(prob.theta0, dirs)], kn.detach())
        tm = prob.final_theta([t - h * d for t, d in zip(prob.theta0, dirs)], kn.detach())
        fd = prob.objective_difference(tp, tm) / (2 * h)
        ad = float(sum((g.grad * d).sum() for g, d in zip(th0, dirs)))
        assert abs(ad - fd) <= 1e-6 * abs(fd), (trial, ad, fd)
    hk = 1e-5
    with torch.no_grad():
        jp = float(prob.objective(prob.theta0, torch.tensor(prob.kn * (1 + hk), dtype=torch.float64)))
        jm = float(prob.objective(prob.theta0, torch.tensor(prob.kn * (1 - hk), dtype=torch.float64)))
    fdk = (jp - jm) / (2 * hk * prob.kn)
    assert abs(float(kn.grad) - fdk) <= 1e-6 * abs(fdk), (float(kn.grad), fdk)
    prob.engine.close()


def test_resident_recording_matches_checkpointing(mbm):
    """A replay that keeps its states resident (every DATA state, the stage
    states as far as the memory budget goes) gives the checkpointed run's
    final state and gradient bit for bit, whichever split the budget forces;
    a budget below N + 2 layers falls back to sqrt(N) checkpointing. The
    layers held show which mode ran."""
    deck = mbm.read_deck(deck_path("small_kn01"))
    eng = mbm.Engine(deck.blocks, kn=deck.kn, scenario="tempjump")
    eng.init()
    U0 = generic_state(mbm, eng, deck, seed=9)
    eng.layer_set(0, U0)
    n, _ = eng.adj_forward(tend=deck.tout[0], cflreg=deck.cfl, maxsteps=23, stride=5)
    assert n == 23
    dts = eng.adj_schedule()
    w = np.random.default_rng(6).normal(size=U0.size) * interior_mask(eng)
    lb = eng.layer_elems() * 8
    eng.adj_release()
    UN = run(eng, U0, dts, stride=5)
    g, gk = eng.adj_backward(w)
    assert eng.adj_memory() == (5 + 2 * 9) * lb        # 5 checkpoints, 2 x (2*5 - 1) segment layers
    # budget in layers -> layers held after forward + backward
    cases = [(0, 23 + 23),                 # automatic: everything kept
             (23 + 2 + 10, 23 + 10 + 2),   # 10 stage states kept, 13 recomputed in 2 layers
             (23 + 2, 23 + 2),             # no stage state kept
             (23 + 1, 5 + 2 * 9)]          # too small: checkpoints every 5 steps
    for budget, held in cases:
        eng.adj_release()
        eng.adj_set_budget(budget * lb)
        assert np.array_equal(run(eng, U0, dts), UN), budget
        g2, gk2 = eng.adj_backward(w)
        assert np.array_equal(g2, g) and gk2 == gk, budget
        assert eng.adj_memory() == held * lb, (budget, eng.adj_memory() // lb)
        g3, gk3 = eng.adj_backward(w)                   # the recording stays intact
        assert np.array_equal(g3, g) and gk3 == gk, budget
    # changes of mode without a release: the held layers are reused
    eng.adj_set_budget(0)
    for stride in (5, None, 2, None):
        assert np.array_equal(run(eng, U0, dts, stride=stride), UN), stride
        g2, gk2 = eng.adj_backward(w)
        assert np.array_equal(g2, g) and gk2 == gk, stride
    assert eng.adj_memory() == (23 + 23) * lb           # the largest set of layers used
    eng.close()


def test_failed_forward_leaves_no_recording(mbm):
    """A forward run that fails (here an invalid stride) discards the
    recording, so a backward pass is refused instead of reversing a stale or
    partial run."""
    deck = mbm.read_deck(deck_path("box_small"))
    eng = mbm.Engine(deck.blocks, kn=0.01, scenario="tempjump")
    eng.init()
    U0 = generic_state(mbm, eng, deck)
    w = np.random.default_rng(2).normal(size=U0.size) * interior_mask(eng)
    run(eng, U0, [1e-3] * 4)
    eng.adj_backward(w)
    with pytest.raises(mbm.MbmError):
        eng.adj_forward(tend=math.inf, cflreg=1.0, dts=[1e-3] * 4, stride=-1)
    with pytest.raises(mbm.MbmError):
        eng.adj_backward(w)
    eng.close()
