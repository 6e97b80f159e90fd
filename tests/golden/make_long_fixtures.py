"""Compact fixtures of the full-size runs, from the reference build.

    python tests/golden/make_long_fixtures.py square4_1024 [lshape_500] [lshape_2000] [--run]

A full-size run of the reference (oracle/_ref/ref_harness steplog, OpenMP
over blocks) is too large to keep (305 MB of FP64 state for 4 x 1024^2), so
what is kept of it is:
    log     every TimeUpdate call: cfl, dt_used, t_after, accepted (exact)
    stats   per block and conservative field over the interior cells:
            max, min, sum |x|, sqrt(sum x^2)          shape (nblocks, 9, 4)
    lines   full FP64 interior rows/columns through the initial jump: for each
            block containing y = 0 its row nearest to y = 0, for each block
            containing x = 0 its column nearest to x = 0, as
            (block, 'row'|'col', index, values (n, 9))
written to tests/golden/long/<name>.npz with the deck and step limit used.

--run runs the reference first (hours for square4 to t = 0.3 and for 2000
steps of the L-shape; its run to t = 0.6 would take more than 12 hours on 3
cores); without it
the .log/.bin of a previous steplog run are read from --dir (default
_work/long, not part of the repository).
"""
import argparse
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
RUNS = {
    "square4_1024": ("multiblockmomentsfv_amd/decks/square4_1024.txt", 1000000, 4),
    "lshape_500": ("multiblockmomentsfv_amd/decks/lshape_2048.txt", 500, 3),
    "lshape_2000": ("multiblockmomentsfv_amd/decks/lshape_2048.txt", 2000, 3),
}


def summarise(name, deck_rel, nsteps, prefix):
    sys.path.insert(0, ROOT)
    import multiblockmomentsfv_amd as m
    deck = m.read_deck(os.path.join(ROOT, deck_rel))
    log = []
    with open(prefix + ".log") as f:
        for line in f:
            c, d, t, a = line.split()
            log.append((float.fromhex(c), float.fromhex(d), float.fromhex(t), float(a)))
    log = np.array(log)
    raw = np.memmap(prefix + ".bin", dtype="<f8", mode="r")
    stats, lines, names = [], [], []
    off = 0
    for k, b in enumerate(deck.blocks):
        lx, ly = b["nx"] + 6, b["ny"] + 6
        blk = np.asarray(raw[off:off + lx * ly * 9]).reshape(ly, lx, 9)
        off += lx * ly * 9
        inner = blk[3:-3, 3:-3, :].reshape(-1, 9)
        stats.append(np.stack([inner.max(axis=0), inner.min(axis=0), np.abs(inner).sum(axis=0),
                               np.sqrt((inner * inner).sum(axis=0))], axis=1))
        bx, by, dx, dy = deck.geom(k)
        if b["ymin"] <= 0.0 < b["ymax"]:
            j = int(np.argmin(np.abs(by + dy * np.arange(3, b["ny"] + 3))))
            lines.append(blk[3 + j, 3:-3, :].copy())
            names.append((k, 0, j))
        if b["xmin"] <= 0.0 < b["xmax"]:
            i = int(np.argmin(np.abs(bx + dx * np.arange(3, b["nx"] + 3))))
            lines.append(blk[3:-3, 3 + i, :].copy())
            names.append((k, 1, i))
    assert off == raw.size
    out = os.path.join(HERE, "long", name + ".npz")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    kw = {f"line{n}": v for n, v in enumerate(lines)}
    np.savez_compressed(out, log=log, stats=np.array(stats), line_index=np.array(names, dtype=np.int64),
                        deck=np.array(deck_rel), maxsteps=np.array(nsteps), **kw)
    acc = int(log[:, 3].sum())
    print(f"{out}: {acc} steps, {len(log) - acc} rejections, t = {log[-1, 2]!r}, {len(lines)} lines")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("names", nargs="+", choices=sorted(RUNS))
    ap.add_argument("--run", action="store_true")
    ap.add_argument("--dir", default=os.path.join(ROOT, "_work", "long"))
    a = ap.parse_args()
    for name in a.names:
        deck_rel, nsteps, threads = RUNS[name]
        prefix = os.path.join(a.dir, name)
        if a.run:
            sys.path.insert(0, ROOT)
            import oracle
            os.makedirs(a.dir, exist_ok=True)
            env = dict(os.environ, OMP_NUM_THREADS=str(threads))
            subprocess.run([oracle.REF_HARNESS, "steplog", "-bfile", os.path.join(ROOT, deck_rel),
                            "-nsteps", str(nsteps), "-dump", prefix], check=True, env=env)
        summarise(name, deck_rel, nsteps, prefix)


if __name__ == "__main__":
    main()
