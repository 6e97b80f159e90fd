"""Writes the FP64 fixtures of tests/golden/fixtures from the reference build.

Run where the reference sources exist (they are compiled in place by
``make -C oracle ref`` into oracle/_ref):

    python tests/golden/make_fixtures.py [case ...]     (default: every case)

For every case below it writes the block file to tests/golden/decks/<name>.txt
and runs ``oracle/_ref/ref_harness steplog`` (the reference's own SetBlocks,
SetInitial, SetBoundary and TimeUpdate, driven by the Steps controller) for the
given number of accepted steps.  The result is stored as
tests/golden/fixtures/<name>.npz with
    state : DATA layer of all blocks after the last step, float64, cell-major,
            9 components per cell, ghosts included, block order
    log   : one row per TimeUpdate call: cfl, dt_used, t_after, accepted
The decks use the reference's block-file format; parameters are set in the
file (not on the command line, which float-casts them).
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

SQUARE_SMALL = [(-1.0, 0.5, -1.0, 0.2, 15, 12, (0, 2, -1, 4)),
                (0.5, 1.0, -1.0, 0.2, 5, 12, (1, 0, -1, 3)),
                (0.5, 1.0, 0.2, 1.0, 5, 8, (4, 0, 2, -1)),
                (-1.0, 0.5, 0.2, 1.0, 15, 8, (0, 3, 1, -1))]
BOX_SMALL = [(-1.0, 0.0, -1.0, 0.0, 12, 12, (-1, 2, -1, 4)),
             (0.0, 1.0, -1.0, 0.0, 12, 12, (1, -1, -1, 3)),
             (0.0, 1.0, 0.0, 1.0, 12, 12, (4, -1, 2, -1)),
             (-1.0, 0.0, 0.0, 1.0, 12, 12, (-1, 3, 1, -1))]
LSHAPE_SMALL = [(-1.0, 0.5, -1.0, 0.2, 30, 24, (-2, -1, 0, 2)),
                (-1.0, 0.5, 0.2, 1.0, 30, 16, (-2, 3, 1, -1)),
                (0.5, 1.0, 0.2, 1.0, 10, 16, (2, 0, -1, -1))]
FFS_SMALL = [(0.0, 1.0, 0.0, 0.33333, 25, 10, (-2, -1, -1, 2)),
             (0.0, 1.0, 0.33333, 1.0, 25, 20, (-2, 3, 1, -1)),
             (1.0, 3.0, 0.33333, 1.0, 50, 20, (2, 0, -1, -1))]
MIXED = [(-1.0, 0.0, -1.0, 1.0, 10, 20, (-2, 2, 0, -1)),
         (0.0, 1.0, -1.0, 1.0, 10, 20, (1, -1, -2, 0))]
PERIODIC = [(-0.5, 0.5, -0.3, 0.3, 16, 12, (1, 1, -1, -1))]
REJECT = [(-0.1, 0.1, -0.1, 0.1, 20, 20, (0, 0, 0, 0))]


def grid_blocks(xs, ys, present, nx, ny, side_code):
    """Blocks on the cells (i, j) of a tensor grid with present(i, j), numbered
    row by row from the bottom; shared edges get sync codes, the others
    side_code(i, j, side) with side 0..3 = left, right, bottom, top."""
    ids = {}
    for j in range(len(ys) - 1):
        for i in range(len(xs) - 1):
            if present(i, j):
                ids[(i, j)] = len(ids) + 1
    out = []
    for (i, j), _ in sorted(ids.items(), key=lambda kv: kv[1]):
        nb = [(i - 1, j), (i + 1, j), (i, j - 1), (i, j + 1)]
        bc = tuple(ids.get(q, side_code(i, j, s)) if q in ids else side_code(i, j, s)
                   for s, q in enumerate(nb))
        out.append((xs[i], xs[i + 1], ys[j], ys[j + 1], nx, ny, bc))
    return out


# forward-facing step on 28 blocks of 0.5 x 1/6 (the step occupies x >= 1,
# y < 1/3): inflow (Dirichlet) on the left, outflow on the right, walls
# elsewhere
FFS28 = grid_blocks([0.0, 0.5, 1.0, 1.5, 2.0, 2.5, 3.0],
                    [0.0, 0.166667, 0.333333, 0.5, 0.666667, 0.833333, 1.0],
                    lambda i, j: j >= 2 or i < 2, 30, 10,
                    lambda i, j, s: -2 if s == 0 else (0 if s == 1 and i == 5 else -1))
# 4 x 4 uniform blocks on [-1, 1]^2, walls outside
SQUARE16 = grid_blocks([-1.0, -0.5, 0.0, 0.5, 1.0], [-1.0, -0.5, 0.0, 0.5, 1.0],
                       lambda i, j: True, 12, 12, lambda i, j, s: -1)

# name: (blocks, header parameters, accepted steps, ffs scenario)
CASES = {
    "small_harmonic": (SQUARE_SMALL, dict(limiter="harmonic"), 30, False),
    "small_minmod": (SQUARE_SMALL, dict(limiter="minmod"), 20, False),
    "small_superbee": (SQUARE_SMALL, dict(limiter="superbee"), 20, False),
    "small_average": (SQUARE_SMALL, dict(limiter="average"), 1, False),   # NaN already after step 1 in the reference
    "small_none": (SQUARE_SMALL, dict(limiter="none"), 20, False),
    "small_force": (SQUARE_SMALL, dict(limiter="harmonic", F=0.5), 20, False),
    "small_kn01": (SQUARE_SMALL, dict(limiter="harmonic", Kn=0.01, R13=1, CFL=0.9), 20, False),
    "box_small": (BOX_SMALL, dict(limiter="harmonic"), 20, False),
    "lshape_small": (LSHAPE_SMALL, dict(limiter="minmod", Kn=0.05, R13=1), 20, False),
    "mixed_bc": (MIXED, dict(limiter="harmonic"), 15, False),
    "periodic": (PERIODIC, dict(limiter="harmonic", Kn=0.01), 10, False),
    "reject": (REJECT, dict(limiter="harmonic", Kn=1e-6), 5, False),
    "ffs_small": (FFS_SMALL, dict(limiter="harmonic"), 10, True),
    "ffs28": (FFS28, dict(limiter="average"), 20, True),
    "square16": (SQUARE16, dict(limiter="harmonic", Kn=0.01, R13=1, CFL=0.9), 20, False),
}


def deck_text(blocks, params):
    p = dict(Kn=0.001, R13=0, CFL=0.95, limiter="harmonic", dt_out=0.0, t_list=0.3)
    p.update(params)
    lines = ["# fixture deck (reference block-file format)"]
    for k, v in p.items():
        lines.append(f"{k} = {v}")
    lines += ["", "Block data (Maximal NB blocks!):",
              "number/xmin,xmax/ymin,ymax/nstx,nsty/leftbound,rightbound,bottombound,topbound/"]
    for i, (x0, x1, y0, y1, nx, ny, bc) in enumerate(blocks, 1):
        lines.append(f"{i}/{x0},{x1}/{y0},{y1}/{nx},{ny}/{bc[0]},{bc[1]},{bc[2]},{bc[3]}/")
    return "\n".join(lines) + "\n"


def main():
    if not oracle.ref_available() or not oracle.ref_available(ffs=True):
        sys.exit("oracle/_ref is not built (make -C oracle ref)")
    os.makedirs(os.path.join(HERE, "fixtures"), exist_ok=True)
    names = sys.argv[1:] or list(CASES)
    for name in names:
        blocks, params, nsteps, ffs = CASES[name]
        path = os.path.join(HERE, "decks", name + ".txt")
        with open(path, "w") as f:
            f.write(deck_text(blocks, params))
        with tempfile.TemporaryDirectory() as tmp:
            log, state = oracle.ref_steplog(path, nsteps, os.path.join(tmp, "run"), ffs=ffs)
        np.savez_compressed(os.path.join(HERE, "fixtures", name + ".npz"), state=state, log=log)
        print(f"{name}: {len(log)} attempts, {state.size} values")


if __name__ == "__main__":
    main()
