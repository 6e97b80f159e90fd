"""Full-size checks on the GPU: the benchmark deck (4 x 1024^2) against the
oracle for a few steps, and size-independent properties at 10^6-cell sizes
(decomposition invariance, conservation and mirror symmetry of a wall box,
a uniform state staying uniform)."""
import os

import numpy as np
import pytest

from conftest import REL_LINF_TOL, ROOT, rel_linf

import oracle

pytestmark = pytest.mark.gpu

DECK = os.path.join(ROOT, "multiblockmomentsfv_amd", "decks", "square4_1024.txt")


def test_square4_1024_against_oracle(mbm):
    deck = mbm.read_deck(DECK)
    eng = mbm.Engine.from_deck(deck)
    eng.init()
    orc = oracle.Oracle(deck.blocks, kn=deck.kn, limiter=deck.limiter, threads=min(8, os.cpu_count() or 1))
    orc.init()
    t, n, _, log = eng.steps(0.0, deck.tout[0], deck.cfl, maxsteps=2)
    n2, log2 = orc.steps(deck.tout[0], deck.cfl, maxsteps=2)
    assert n == n2 == 2
    assert np.allclose(log[:, :3], log2[:, :3], rtol=REL_LINF_TOL, atol=0)
    err = rel_linf(eng.download(0), orc.get(0))
    assert err.max() <= REL_LINF_TOL, err


def test_lshape_full_size_against_oracle(mbm):
    """The L-shaped 3-block deck at full size (2048x1024, 1024x1024, 1024x2048
    cells, walls, Kn 0.05, minmod): two steps against the oracle."""
    deck = mbm.read_deck(os.path.join(ROOT, "multiblockmomentsfv_amd", "decks", "lshape_2048.txt"))
    assert deck.interior_cells() == 5242880
    eng = mbm.Engine.from_deck(deck)
    eng.init()
    orc = oracle.Oracle(deck.blocks, kn=deck.kn, limiter=deck.limiter, threads=min(8, os.cpu_count() or 1))
    orc.init()
    t, n, _, log = eng.steps(0.0, deck.tout[0], deck.cfl, maxsteps=2)
    n2, log2 = orc.steps(deck.tout[0], deck.cfl, maxsteps=2)
    assert n == n2 == 2
    assert np.allclose(log[:, :3], log2[:, :3], rtol=REL_LINF_TOL, atol=0)
    err = rel_linf(eng.download(0), orc.get(0))
    assert err.max() <= REL_LINF_TOL, err


def test_split_block_is_bitwise_identical(mbm):
    one = [dict(nx=2048, ny=1024, xmin=-1.0, xmax=1.0, ymin=-0.5, ymax=0.5, bc=[0, 0, -1, 0])]
    two = [dict(nx=1024, ny=1024, xmin=-1.0, xmax=0.0, ymin=-0.5, ymax=0.5, bc=[0, 2, -1, 0]),
           dict(nx=1024, ny=1024, xmin=0.0, xmax=1.0, ymin=-0.5, ymax=0.5, bc=[1, 0, -1, 0])]
    e1 = mbm.Engine(one, kn=0.01)
    e2 = mbm.Engine(two, kn=0.01)
    e1.init()
    e2.init()
    for k in range(5):
        e1.time_update(2e-4, 2e-4 * k)
        e1.commit()
        e2.time_update(2e-4, 2e-4 * k)
        e2.commit()
    a = e1.download(0)[0][3:-3, 3:-3]
    b2 = e2.download(0)
    b = np.concatenate([b2[0][3:-3, 3:-3], b2[1][3:-3, 3:-3]], axis=1)
    assert a.tobytes() == b.tobytes()


def box_blocks(n):
    return [dict(nx=n, ny=n, xmin=-1.0, xmax=0.0, ymin=-1.0, ymax=0.0, bc=[-1, 2, -1, 4]),
            dict(nx=n, ny=n, xmin=0.0, xmax=1.0, ymin=-1.0, ymax=0.0, bc=[1, -1, -1, 3]),
            dict(nx=n, ny=n, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, bc=[4, -1, 2, -1]),
            dict(nx=n, ny=n, xmin=-1.0, xmax=0.0, ymin=0.0, ymax=1.0, bc=[-1, 3, 1, -1])]


def test_wall_box_conservation_and_symmetry(mbm):
    """square4Box semantics (cases/square4Box.txt:1-6) at 4 x 512^2, 200 steps."""
    eng = mbm.Engine(box_blocks(512), kn=0.001)
    eng.init()
    s0 = eng.download(0)
    m0 = sum(math_sum(a) for a in s0)
    eng.steps(0.0, 1.0, 0.95, maxsteps=200)
    s1 = eng.download(0)
    m1 = sum(math_sum(a) for a in s1)
    for c in (0, 3):     # mass, energy
        assert abs(m1[c] - m0[c]) <= 1e-12 * abs(m0[c]), (c, m0[c], m1[c])
    odd_x = np.array([1, -1, 1, 1, 1, -1, 1, -1, 1])
    for a, b in ((0, 1), (3, 2)):
        A = s1[a][3:-3, 3:-3]
        B = s1[b][3:-3, 3:-3][:, ::-1] * odd_x
        scale = np.abs(A).reshape(-1, 9).max(axis=0)
        assert (np.abs(A - B).reshape(-1, 9).max(axis=0) <= 1e-12 * scale).all()


def math_sum(a):
    import math
    return np.array([math.fsum(a[3:-3, 3:-3, c].ravel()) for c in range(9)])


def test_uniform_state_stays_uniform(mbm):
    blk = [dict(nx=1024, ny=1024, xmin=-0.2, xmax=0.2, ymin=-0.2, ymax=0.2, bc=[0, 0, 0, 0])]
    eng = mbm.Engine(blk, kn=0.01)
    eng.init()
    u0 = eng.download(0)[0][3:-3, 3:-3].reshape(-1, 9)[0].copy()
    eng.steps(0.0, 1.0, 0.9, maxsteps=20)
    s = eng.download(0)[0][3:-3, 3:-3].reshape(-1, 9)
    assert np.abs(s - u0).max() <= 1e-13 * np.abs(u0).max()
