"""Parity of the gfx950 engine with the reference path (through the C ABI).

Bar: FP64 relative L-inf per field <= 1e-12 (REL_LINF_TOL) against the FP64
fixtures written by the compiled reference and against the CPU oracle, with
non-finite values at exactly the reference's places (small_average turns
2655 of 9216 values NaN in one step: a zero face density under the
unlimited average limiter); the step controller must take the same
accept/reject decisions; integer/index work (ghost map, initial state) must
be bit-exact.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import (GOLDEN, REL_LINF_TOL, deck_path, fixture_names, interior, load_fixture,
                      rel_linf, scenario_of, split_state)

import oracle

pytestmark = pytest.mark.gpu


def engine_for(mbm, name):
    deck = mbm.read_deck(deck_path(name))
    eng = mbm.Engine.from_deck(deck, scenario=scenario_of(name))
    return deck, eng


def oracle_for(deck, name):
    return oracle.Oracle(deck.blocks, kn=deck.kn, force=deck.force, limiter=deck.limiter,
                         scenario=scenario_of(name))


@pytest.mark.parametrize("name", fixture_names())
def test_initial_state_bitwise(mbm, name):
    deck, eng = engine_for(mbm, name)
    eng.init()
    orc = oracle_for(deck, name)
    orc.init()
    for k, (g, r) in enumerate(zip(eng.download(0), orc.get(0))):
        assert g.tobytes() == r.tobytes(), f"block {k + 1}"


@pytest.mark.parametrize("name", fixture_names())
def test_ghost_fill_matches_host_map_and_oracle(mbm, name):
    deck, eng = engine_for(mbm, name)
    rng = np.random.default_rng(7)
    arrays = [rng.uniform(0.5, 2.0, eng.shape(k)) for k in range(deck.nblocks)]
    eng.upload(arrays, layer=1)
    eng.set_boundary(1, 0.0)
    got = eng.download(1)
    host = [a.copy() for a in arrays]
    mbm.host_set_boundary(deck.blocks, host, kn=deck.kn, limiter=deck.limiter,
                          scenario=scenario_of(name))
    orc = oracle_for(deck, name)
    orc.set([a.copy() for a in arrays], layer=1)
    orc.set_boundary(1, 0.0)
    for k in range(deck.nblocks):
        assert got[k].tobytes() == host[k].tobytes(), f"GPU vs host map, block {k + 1}"
        assert got[k].tobytes() == orc.get(1)[k].tobytes(), f"GPU vs oracle, block {k + 1}"


@pytest.mark.parametrize("name", fixture_names())
def test_one_time_update_matches_oracle(mbm, name):
    deck, eng = engine_for(mbm, name)
    eng.init()
    orc = oracle_for(deck, name)
    orc.init()
    dt = 1e-4
    c_gpu = eng.time_update(dt, 0.0)
    c_ref = orc.time_update(dt, 0.0)
    for layer in (1, 2):
        g, r = eng.download(layer), orc.get(layer)
        err = rel_linf(g, r)
        assert err.max() <= REL_LINF_TOL, (layer, err)
    assert abs(c_gpu - c_ref) <= REL_LINF_TOL * abs(c_ref), (c_gpu, c_ref)


@pytest.mark.parametrize("name", fixture_names())
def test_steps_match_reference_fixture(mbm, name):
    state, ref_log = load_fixture(name)
    nsteps = int(ref_log[:, 3].sum())
    deck, eng = engine_for(mbm, name)
    eng.init()
    t, n, nrej, log = eng.steps(0.0, deck.tout[0], deck.cfl, maxsteps=nsteps)
    assert n == nsteps
    assert log.shape == ref_log.shape
    assert np.array_equal(log[:, 3], ref_log[:, 3]), "accept/reject sequence differs"
    assert np.allclose(log[:, :3], ref_log[:, :3], rtol=REL_LINF_TOL, atol=0.0, equal_nan=True)
    ref = split_state(state, deck.blocks)
    got = eng.download(0)
    err = rel_linf(got, ref)
    assert err.max() <= REL_LINF_TOL, err


def test_output_conversion_bitwise(mbm):
    """Output.c:217-243 on the device == the oracle's conversion of the same state."""
    deck, eng = engine_for(mbm, "small_harmonic")
    eng.init()
    eng.steps(0.0, deck.tout[0], deck.cfl, maxsteps=5)
    orc = oracle_for(deck, "small_harmonic")
    orc.set(eng.download(0), layer=0)
    for k in range(deck.nblocks):
        assert eng.output_values(k).tobytes() == orc.output_values(k).tobytes()
        f = eng.output_values(k, float32=True)
        assert f.tobytes() == orc.output_values(k).astype(np.float32).tobytes()


def test_output_fields_at_output_times(mbm):
    """FP64 output fields (rho vx vy T p Pxx Pxy Pyy qx qy) of a full squareSmall
    run at every output time, against the oracle: rel-L-inf per field."""
    deck, eng = engine_for(mbm, "squareSmall")
    eng.init()
    orc = oracle_for(deck, "squareSmall")
    orc.init()
    t = 0.0
    for tend in deck.tout:
        t, n, _, _ = eng.steps(t, tend, deck.cfl)
        n2, _ = orc.steps(tend, deck.cfl)
        assert n == n2
        g = np.concatenate([eng.output_values(k) for k in range(deck.nblocks)])
        r = np.concatenate([orc.output_values(k) for k in range(deck.nblocks)])
        err = np.abs(g - r).max(axis=0) / np.abs(r).max(axis=0)
        assert err.max() <= REL_LINF_TOL, (tend, err)


@pytest.mark.parametrize("limiter", ["average", "none", "minmod", "superbee", "harmonic"])
def test_every_limiter_on_a_smooth_deck(mbm, limiter):
    """A smooth temperature jump (no negative temperatures) for every limiter,
    30 steps of the Steps controller against the oracle."""
    deck = mbm.read_deck(deck_path("small_kn01"))
    params = (1.5, 1.0, 0.4, 0.2, 1.0)
    eng = mbm.Engine(deck.blocks, kn=deck.kn, limiter=limiter, scenario="tempjump",
                     scenario_params=params)
    eng.init()
    orc = oracle.Oracle(deck.blocks, kn=deck.kn, limiter=limiter, scenario="tempjump")
    orc.set_tempjump(*params)
    orc.init()
    _, n, _, log = eng.steps(0.0, deck.tout[0], deck.cfl, maxsteps=20)
    n2, log2 = orc.steps(deck.tout[0], deck.cfl, maxsteps=20)
    assert n == n2 == 20
    assert np.allclose(log[:, :3], log2[:, :3], rtol=REL_LINF_TOL, atol=0)
    err = rel_linf(eng.download(0), orc.get(0))
    assert err.max() <= REL_LINF_TOL, err


def run_moments(args, cwd):
    exe = os.path.join(os.path.dirname(GOLDEN), "..", "build", "moments")
    return subprocess.run([os.path.abspath(exe)] + args, cwd=cwd, capture_output=True,
                          text=True, timeout=600)


@pytest.mark.parametrize("name", ["squareSmall", "square4"])
def test_moments_cli_against_golden(mbm, name):
    """Drop-in run: same deck, same output files, compared like `make check`."""
    from multiblockmomentsfv_amd.compare_solution import compare
    import lzma
    with tempfile.TemporaryDirectory() as tmp:
        r = run_moments(["-bfile", deck_path(name), "-out", tmp], tmp)
        assert r.returncode == 0, r.stdout + r.stderr
        nst = 400 if name == "squareSmall" else 22500
        base = os.path.join(tmp, f"Grad13Riemann2D_nst{nst}_Kn0p001_tAll")
        ref = os.path.join(tmp, "ref")
        with open(ref + ".hdr", "w") as f:
            f.write(open(os.path.join(GOLDEN, name + ".ref.hdr")).read())
        p = os.path.join(GOLDEN, name + ".ref.dat")
        data = open(p, "rb").read() if os.path.exists(p) else lzma.decompress(open(p + ".xz", "rb").read())
        with open(ref + ".dat", "wb") as f:
            f.write(data)
        msgs = []
        assert compare(base, ref, 1e-6, 1e-5, out=msgs.append) == 0, msgs
        mine = [l for l in open(base + ".hdr") if l.startswith("ZONE")]
        gold = [l for l in open(ref + ".hdr") if l.startswith("ZONE")]
        assert mine == gold
        assert f"Done after {352 if name == 'square4' else 112} steps." in r.stdout
        if name == "square4":
            # the step log (every 20th step: CFL, time, next dt to 6 decimals)
            # against the reference's own log of this run
            import re
            pat = re.compile(r"step:\s*(\d+)\.\.\. done\.  CFL: ([0-9.]+), time: ([0-9.]+), next dt: ([0-9.]+)")
            mine_log = pat.findall(r.stdout)
            gold_log = pat.findall(open(os.path.join(GOLDEN, "square4.ref.txt")).read())
            assert len(gold_log) == 16 and mine_log == gold_log, (mine_log, gold_log)


LIMITER_NAMES = ["none", "average", "minmod", "superbee", "harmonic"]


@pytest.mark.parametrize("limiter", LIMITER_NAMES)
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_nonfinite_states_propagate_as_in_the_reference(mbm, limiter, seed):
    """States holding NaN in scattered components (ghosts included) and,
    for odd seeds, a cell with negative temperature: one TimeUpdate on the
    GPU against the oracle -- the same NaN cells and components in TMP and
    NEW, finite values within 1e-12, and the same CFL (the reference's
    folds ignore a NaN face speed but restart the production maximum after
    a NaN, Main.c:325-326,364,381)."""
    deck = mbm.read_deck(deck_path("small_kn01"))
    orc = oracle.Oracle(deck.blocks, kn=deck.kn, force=deck.force, limiter=limiter)
    orc.init()
    orc.steps(deck.tout[0], deck.cfl, maxsteps=3)
    rng = np.random.default_rng(seed)
    data = [a.copy() for a in orc.get(0)]
    for a in data:
        ny, nx, _ = a.shape
        for _ in range(3):
            a[rng.integers(0, ny), rng.integers(0, nx), rng.integers(0, 9)] = np.nan
        if seed % 2:
            b, c = rng.integers(3, ny - 3), rng.integers(3, nx - 3)
            a[b, c, 3] = 0.1 * a[b, c, 3] - 1.0
    orc.set(data, 0)
    eng = mbm.Engine(deck.blocks, kn=deck.kn, force=deck.force, limiter=limiter)
    eng.upload(data, layer=0)
    c_gpu = eng.time_update(1e-4, 0.0)
    c_ref = orc.time_update(1e-4, 0.0)
    for layer in (1, 2):
        ref = orc.get(layer)
        assert sum(int(np.isnan(r).sum()) for r in ref) > 0
        err = rel_linf(eng.download(layer), ref)
        assert err.max() <= REL_LINF_TOL, (layer, err)
    assert abs(c_gpu - c_ref) <= REL_LINF_TOL * abs(c_ref), (c_gpu, c_ref)
    eng.close()


def test_nan_production_cfl_restarts_the_maximum(mbm):
    """A NaN production CFL at the last cell of the reference's scan makes
    the block's CFL NaN, and a NaN block CFL is skipped in the stage maximum
    (Main.c:268-271): the GPU controller takes the same CFL and dt."""
    deck = mbm.read_deck(deck_path("small_kn01"))
    orc = oracle.Oracle(deck.blocks, kn=deck.kn, limiter="harmonic")
    orc.init()
    data = [a.copy() for a in orc.get(0)]
    for k, a in enumerate(data):
        ly, lx, _ = a.shape
        # the last cell of the i-outer, j-inner scan over rings >= 1 of block k
        # (column lx-2, row ly-2, a corner ghost of ring 1) gets a negative
        # temperature in every other block
        if k % 2 == 0:
            a[ly - 2, lx - 2, 3] = 0.1 * a[ly - 2, lx - 2, 3] - 1.0
    orc.set(data, 0)
    eng = mbm.Engine(deck.blocks, kn=deck.kn, limiter="harmonic")
    eng.upload(data, layer=0)
    c_gpu = eng.time_update(1e-4, 0.0)
    c_ref = orc.time_update(1e-4, 0.0)
    assert abs(c_gpu - c_ref) <= REL_LINF_TOL * abs(c_ref), (c_gpu, c_ref)
    for layer in (1, 2):
        err = rel_linf(eng.download(layer), orc.get(layer))
        assert err.max() <= REL_LINF_TOL, (layer, err)
    eng.close()


@pytest.mark.parametrize("name", ["lshape_small", "periodic", "mixed_bc", "ffs_small"])
def test_tile_ghost_copies_equal_set_boundary(mbm, name):
    """The stage kernels copy the interior-sourced ghosts at the end of the
    tiles that store their sources, in the graph-replayed steps and in
    TimeUpdate; only scenario (Dirichlet) ghosts keep a fill launch. The
    ghosts they leave equal a stand-alone SetBoundary (Boundary.c:173-193,
    the fill kernel over the whole ghost map) bit for bit: DATA after the
    steps, TMP and NEW after a TimeUpdate (both filled at t + dt,
    Main.c:253-284)."""
    deck, eng = engine_for(mbm, name)
    eng.init()
    t, n, r, _ = eng.steps(0.0, deck.tout[0], deck.cfl, maxsteps=4, reccap=0)
    assert n == 4
    U = eng.layer_get(0)
    eng.set_boundary(0, t)
    assert U.tobytes() == eng.layer_get(0).tobytes()
    dt = eng.next_dt
    eng.time_update(dt, t)
    for layer in (1, 2):
        V = eng.layer_get(layer)
        eng.set_boundary(layer, t + dt)
        assert V.tobytes() == eng.layer_get(layer).tobytes(), layer
    eng.close()
