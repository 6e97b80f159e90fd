"""The device row order on the GPU: the whole-layer view agrees with the
per-block copies, and a change of tile height moves the state to the new
order without changing it (the result of a run does not depend on when the
tile height changed)."""
import numpy as np
import pytest

from conftest import deck_path

pytestmark = pytest.mark.gpu


def test_layer_view_matches_block_copies(mbm):
    deck = mbm.read_deck(deck_path("lshape_small"))
    eng = mbm.Engine(deck.blocks, kn=deck.kn, limiter=deck.limiter)
    eng.init()
    eng.steps(0.0, deck.tout[0], deck.cfl, maxsteps=3)
    blocks = eng.download(0)
    via_layer = eng.layer_to_blocks(eng.layer_get(0))
    for k, (a, b) in enumerate(zip(blocks, via_layer)):
        assert a.shape == b.shape and a.tobytes() == b.tobytes(), k
    # and back: a layer assembled from the blocks reads back through the C copies
    eng.layer_set(1, eng.blocks_to_layer(blocks))
    for k, (a, b) in enumerate(zip(blocks, eng.download(1))):
        assert a.tobytes() == b.tobytes(), k


def test_tile_height_change_moves_the_state(mbm):
    deck = mbm.read_deck(deck_path("lshape_small"))
    a = mbm.Engine(deck.blocks, kn=deck.kn, limiter=deck.limiter)
    b = mbm.Engine(deck.blocks, kn=deck.kn, limiter=deck.limiter)
    out = []
    for e in (a, b):
        e.init()
        out.append(e.steps(0.0, deck.tout[0], deck.cfl, maxsteps=6))
    assert out[0][0] == out[1][0]
    before = a.download(0)
    rows0 = [ix[3].copy() for ix in a.layer_index()]
    old_rows = a.tile_rows()
    new_rows = 5 if old_rows != 5 else 7
    a.set_tile_rows(new_rows)
    assert a.tile_rows() == new_rows
    assert any(not np.array_equal(r, ix[3]) for r, ix in zip(rows0, a.layer_index()))
    for k, (x, y) in enumerate(zip(before, a.download(0))):
        assert x.tobytes() == y.tobytes(), k
    t = out[0][0]
    ra = a.steps(t, deck.tout[0], deck.cfl, maxsteps=6, dt=a.next_dt)
    rb = b.steps(t, deck.tout[0], deck.cfl, maxsteps=6, dt=b.next_dt)
    assert ra[:3] == rb[:3]
    for k, (x, y) in enumerate(zip(a.download(0), b.download(0))):
        assert x.tobytes() == y.tobytes(), k


def test_tile_height_change_drops_a_recorded_adjoint_run(mbm):
    deck = mbm.read_deck(deck_path("mixed_bc"))
    eng = mbm.Engine(deck.blocks, kn=0.01, scenario="tempjump")
    eng.init()
    eng.adj_forward(tend=float("inf"), cflreg=1.0, dts=[1e-3] * 4, stride=2)
    g = np.zeros(eng.layer_elems())
    eng.adj_backward(g)
    eng.set_tile_rows(eng.tile_rows() + 3)
    with pytest.raises(mbm.MbmError):
        eng.adj_backward(g)
