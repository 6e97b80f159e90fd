"""Pins the CPU oracle against the reference's own goldens.

reference/squareSmall.ref.* and reference/square4.ref.* (copied as data into
tests/golden) are the Float32 outputs of the reference's Riemann2D runs; the
oracle must reproduce every byte of their payload, every ZONE's STRANDID and
SOLUTIONTIME, and the dt/CFL trajectory printed in square4.ref.txt.
"""
import lzma
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, deck_path

import oracle


def golden_zones(hdr):
    zones = []
    for line in open(hdr):
        if line.startswith("ZONE"):
            kv = dict(re.findall(r"(\w+)=(\S+)", line))
            zones.append(kv)
    return zones


def golden_payload(name):
    p = os.path.join(GOLDEN, name + ".ref.dat")
    if os.path.exists(p):
        return open(p, "rb").read()
    return lzma.decompress(open(p + ".xz", "rb").read())


def run_oracle_outputs(mbm, deck_name, threads=1):
    deck = mbm.read_deck(deck_path(deck_name))
    orc = oracle.Oracle(deck.blocks, kn=deck.kn, force=deck.force, limiter=deck.limiter,
                        threads=threads)
    orc.init()
    records = [(0.0, 1, [orc.output_values(k) for k in range(deck.nblocks)])]
    total = 0
    logs = []
    for tend in deck.tout:
        n, log = orc.steps(tend, deck.cfl)
        total += n
        logs.append(log)
        records.append((orc.t, total + 1, [orc.output_values(k) for k in range(deck.nblocks)]))
    return deck, records, logs


@pytest.mark.parametrize("name,threads", [("squareSmall", 1), ("square4", 4)])
def test_oracle_reproduces_golden_bytes(mbm, name, threads):
    deck, records, _ = run_oracle_outputs(mbm, name, threads)
    zones = golden_zones(os.path.join(GOLDEN, name + ".ref.hdr"))
    payload = golden_payload(name)
    assert len(zones) == deck.nblocks * len(records)
    zi = 0
    for t, strand, vals in records:
        for k in range(deck.nblocks):
            z = zones[zi]
            zi += 1
            assert int(z["block"]) == k + 1
            assert int(z["STRANDID"]) == strand
            assert float(z["SOLUTIONTIME"]) == float("%.16g" % t)
            off, nb = int(z["OFFSET"]), int(z["NBYTES"])
            mine = vals[k].astype(np.float32).tobytes()
            assert len(mine) == nb
            assert mine == payload[off:off + nb], f"zone {zi} (block {k + 1}, t={t}) differs"


def test_oracle_square4_step_log(mbm):
    """square4.ref.txt prints CFL, t and next dt (6 decimals) every 20 steps."""
    deck, _, logs = run_oracle_outputs(mbm, "square4", threads=4)
    printed = []
    for line in open(os.path.join(GOLDEN, "square4.ref.txt")):
        m = re.search(r"step:\s+(\d+)\.\.\. done\.\s+CFL: ([\d.]+), time: ([\d.]+), next dt: ([\d.]+)", line)
        if m:
            printed.append(tuple(m.groups()))
    ours = []
    for log in logs:
        acc = log[log[:, 3] == 1]
        for i in range(19, len(acc), 20):
            nxt = acc[i + 1, 1] if i + 1 < len(acc) else 0.0
            ours.append(("%3d" % (i + 1), "%f" % acc[i, 0], "%f" % acc[i, 2], "%f" % nxt))
    assert len(ours) == len(printed)
    for a, b in zip(ours, printed):
        assert a[0].strip() == b[0]
        assert a[1:] == b[1:], (a, b)
    assert sum(int((log[:, 3] == 1).sum()) for log in logs) == 352
