"""Pins the CPU oracle bit-for-bit against FP64 fixtures of the reference.

The fixtures (tests/golden/fixtures, written by tests/golden/make_fixtures.py
from the compiled reference) cover every limiter, a body force, Kn=0.01 and
0.05, walls on all sides, Dirichlet edges, a periodic (self-sync) block, the
step-rejection branch and the ForwardFacingStep scenario. Where oracle/_ref is
built, the oracle is also compared with a fresh reference run.
"""
import os
import tempfile

import numpy as np
import pytest

from conftest import deck_path, fixture_names, load_fixture, scenario_of, split_state

import oracle


def run_oracle(mbm, name, nsteps):
    deck = mbm.read_deck(deck_path(name))
    orc = oracle.Oracle(deck.blocks, kn=deck.kn, force=deck.force, limiter=deck.limiter,
                        scenario=scenario_of(name))
    orc.init()
    n, log = orc.steps(deck.tout[0], deck.cfl, maxsteps=nsteps)
    return deck, orc, n, log


@pytest.mark.parametrize("name", fixture_names())
def test_oracle_matches_reference_fixture(mbm, name):
    state, ref_log = load_fixture(name)
    nsteps = int(ref_log[:, 3].sum())
    deck, orc, n, log = run_oracle(mbm, name, nsteps)
    assert n == nsteps
    assert log.shape == ref_log.shape
    assert np.array_equal(log, ref_log, equal_nan=True), "step log differs"
    mine = np.concatenate([a.reshape(-1) for a in orc.get(0)])
    assert mine.tobytes() == state.tobytes(), "state differs from the reference"


def test_rejection_branch_is_exercised():
    _, log = load_fixture("reject")
    assert (log[:, 3] == 0).sum() >= 1


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("name", ["small_harmonic", "lshape_small", "mixed_bc"])
def test_oracle_matches_fresh_reference_run(mbm, name):
    nsteps = 7
    with tempfile.TemporaryDirectory() as tmp:
        ref_log, state = oracle.ref_steplog(deck_path(name), nsteps, os.path.join(tmp, "r"),
                                            ffs=name.startswith("ffs"))
    deck, orc, n, log = run_oracle(mbm, name, nsteps)
    assert np.array_equal(log, ref_log)
    mine = np.concatenate([a.reshape(-1) for a in orc.get(0)])
    assert mine.tobytes() == state.tobytes()
