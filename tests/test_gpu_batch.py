"""Batched decks (no reference counterpart): every deck of a batch must
advance exactly as it does alone through the plain engine (same kernels,
same arithmetic: bit-identical), and as the oracle (1e-12)."""
import numpy as np
import pytest

from conftest import REL_LINF_TOL, rel_linf

import oracle

pytestmark = pytest.mark.gpu


def small_sweep(mbm, n=6, seed=3):
    return mbm.sweep_decks(n=n, seed=seed, sides=(8, 12, 16, 24, 32))


def test_batch_decks_equal_single_runs(mbm):
    decks = small_sweep(mbm)
    bat = mbm.Batch(decks)
    bat.init()
    ns, nr = bat.run(maxsteps=25)
    assert (ns == 25).all()
    for i, d in enumerate(decks):
        eng = mbm.Engine(d["blocks"], kn=d["kn"])
        eng.init()
        t, n, rej, _ = eng.steps(0.0, 1e30, d["cfl"], maxsteps=25)
        assert n == 25 and rej == nr[i]
        assert t == bat.t[i]
        for g, r in zip(bat.download(i), eng.download(0)):
            assert g.tobytes() == r.tobytes(), f"deck {i}"


def test_batch_against_oracle_and_second_run(mbm):
    decks = small_sweep(mbm, n=4, seed=9)
    bat = mbm.Batch(decks)
    bat.init()
    bat.run(maxsteps=10)
    bat.run(maxsteps=7)          # second call: dt restarts from DT_INITIAL, DATA carries over
    for i, d in enumerate(decks):
        orc = oracle.Oracle(d["blocks"], kn=d["kn"])
        orc.init()
        orc.steps(1e30, d["cfl"], maxsteps=10)
        orc.steps(1e30, d["cfl"], maxsteps=7)
        assert rel_linf(bat.download(i), orc.get(0)).max() <= REL_LINF_TOL


def test_batch_tend_and_autotune(mbm):
    decks = small_sweep(mbm, n=5, seed=1)
    for i, d in enumerate(decks):
        d["tend"] = 0.01 * (i + 1)
    bat = mbm.Batch(decks)
    bat.init()
    best, ms = bat.autotune(rows=(4, 8, 16), reps=2)
    assert best in (4, 8, 16) and ms > 0
    ns, nr = bat.run(maxsteps=0)
    for i, d in enumerate(decks):
        assert bat.t[i] == pytest.approx(d["tend"], abs=1e-15)
        eng = mbm.Engine(d["blocks"], kn=d["kn"])
        eng.init()
        t, n, rej, _ = eng.steps(0.0, d["tend"], d["cfl"])
        assert n == ns[i]
        assert rel_linf(bat.download(i), eng.download(0)).max() == 0.0


def test_batch_autotune_shape_keeps_results(mbm):
    """The shape autotuner picks a (tile height, wave group) pair from the
    candidates, and the batch then advances exactly as the decks alone."""
    decks = small_sweep(mbm, n=5, seed=2)
    bat = mbm.Batch(decks)
    bat.init()
    rows, waves, ms = bat.autotune_shape(rows=(4, 16), waves=(1, 2, 8), reps=1)
    assert rows in (4, 16) and waves in (1, 2, 8) and ms > 0
    assert bat.wave_group() == waves
    bat.init()
    ns, nr = bat.run(maxsteps=12)
    for i, d in enumerate(decks):
        eng = mbm.Engine(d["blocks"], kn=d["kn"])
        eng.init()
        t, n, rej, _ = eng.steps(0.0, 1e30, d["cfl"], maxsteps=12)
        assert n == ns[i] and rej == nr[i] and t == bat.t[i]
        for g, r in zip(bat.download(i), eng.download(0)):
            assert g.tobytes() == r.tobytes(), f"deck {i}"
        eng.close()
    bat.close()


@pytest.mark.parametrize("waves", [1, 2, 8])
def test_wave_group_does_not_change_results(mbm, waves):
    """Waves per workgroup of the stage kernels is a launch shape only: the
    step log and the state are bit-identical to the default group (4)."""
    d = small_sweep(mbm, n=1, seed=7)[0]
    out = []
    for w in (4, waves):
        eng = mbm.Engine(d["blocks"], kn=d["kn"])
        eng.set_wave_group(w)
        assert eng.wave_group() == w
        eng.init()
        t, n, rej, log = eng.steps(0.0, 1e30, d["cfl"], maxsteps=20)
        out.append((t, n, rej, [a.tobytes() for a in eng.download(0)]))
        eng.close()
    assert out[0] == out[1]


def test_wave_group_rejects_other_sizes(mbm):
    d = small_sweep(mbm, n=1, seed=7)[0]
    eng = mbm.Engine(d["blocks"], kn=d["kn"])
    with pytest.raises(mbm.MbmError):
        eng.set_wave_group(3)
    assert eng.wave_group() == 4
    eng.close()


def test_full_size_sweep_subset_equals_single_runs(mbm):
    """The 256-deck sweep at its full sizes (block sides 64..512, 88 M cells)
    with an autotuned tile height: a few steps, then four decks against their
    single-deck runs, bit for bit, and every deck's state finite."""
    decks = mbm.sweep_decks(256, seed=0)
    bat = mbm.Batch(decks)
    bat.init()
    bat.autotune(rows=(8, 16, 32), reps=1)
    bat.init()
    ns, nr = bat.run(maxsteps=6)
    assert (ns == 6).all()
    for i in (0, 77, 151, 255):
        d = decks[i]
        eng = mbm.Engine(d["blocks"], kn=d["kn"])
        eng.init()
        t, n, rej, _ = eng.steps(0.0, 1e30, d["cfl"], maxsteps=6)
        assert n == 6 and rej == nr[i] and t == bat.t[i]
        for g, r in zip(bat.download(i), eng.download(0)):
            assert g.tobytes() == r.tobytes(), f"deck {i}"
        eng.close()
    for i in range(0, 256, 51):
        assert all(np.isfinite(b).all() for b in bat.download(i))
    bat.close()


def test_batch_of_boundary_decks_equals_single_runs(mbm):
    """Decks with walls, Dirichlet edges, periodic sync and an L-shape in one
    batch (two groups on two streams): the stage kernels' ghost copies and
    the per-group fill of the scenario (Dirichlet) ghosts leave every deck
    bit-identical to its run alone."""
    from conftest import deck_path
    decks = []
    for name in ("mixed_bc", "box_small", "periodic", "lshape_small", "mixed_bc"):
        deck = mbm.read_deck(deck_path(name))
        decks.append(dict(blocks=deck.blocks, kn=deck.kn, cfl=deck.cfl))
    bat = mbm.Batch(decks)
    bat.init()
    ns, nr = bat.run(maxsteps=12)
    assert (ns == 12).all()
    for i, d in enumerate(decks):
        eng = mbm.Engine(d["blocks"], kn=d["kn"])
        eng.init()
        t, n, rej, _ = eng.steps(0.0, 1e30, d["cfl"], maxsteps=12)
        assert n == 12 and rej == nr[i] and t == bat.t[i]
        for g, r in zip(bat.download(i), eng.download(0)):
            assert g.tobytes() == r.tobytes(), f"deck {i}"
        eng.close()
    bat.close()
