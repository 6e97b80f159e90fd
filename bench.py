"""Benchmark of the explicit update on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--deck PATH] [--no-cpu]

Workload: square4 topology with every block 1024x1024 (4 194 304 cells),
Reg13, Kn=0.01, CFL=0.9, harmonic limiter, Riemann2D initial state, FP64,
adaptive dt (multiblockmomentsfv_amd/decks/square4_1024.txt).  A step is one
accepted Heun step (TimeUpdate + commit) driven by the Steps controller; the
state stays in HBM, the host reads one 8-byte CFL per step.

Prints one JSON line: value = cell-updates/s (interior cells x accepted
steps / wall time of the K timed steps).  roofline: the stage kernel's
algorithmic HBM bytes (stage 1 reads U^n and writes U*, stage 2 reads U*,
U^n and writes U^n+1: 360 B per cell-update, 180 B per cell per launch on
average) over its mean launch time measured with HIP events on the engine's
stream.  cpu_baseline: the reference itself (oracle/_ref) on the host cores,
a bounded sample of the same deck.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

DECK = os.path.join(ROOT, "multiblockmomentsfv_amd", "decks", "square4_1024.txt")
METRIC = "FP64 cell-updates/s + % HBM roofline; batched & adjoint vs plain step"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_UPDATE = 360.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "stage_kernel_square4_1024_traffic.json")


def cpu_baseline(deck_path, cells, steps=12):
    """Times the reference build (or the oracle port) on a few steps."""
    import oracle
    threads = min(4, os.cpu_count() or 1)
    if oracle.ref_available():
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
                   OMP_PLACES="cores")
        r = subprocess.run([oracle.REF_HARNESS, "steplog", "-bfile", deck_path, "-nsteps",
                            str(steps), "-time"], capture_output=True, text=True, env=env,
                           timeout=600)
        line = [l for l in r.stdout.splitlines() if l.startswith("{")][-1]
        d = json.loads(line)
        rate = d["cells"] * (d["steps"] + d["rejected"]) / d["seconds"]
        return {"value": rate, "unit": "cell-updates/s", "cores": d["threads"],
                "kind": "reference",
                "sample": f"{d['steps']} Heun steps of the same deck from t=0 "
                          f"(reference Main.c/Flux.c build, OpenMP over blocks, "
                          f"{d['seconds']:.1f} s)"}
    import multiblockmomentsfv_amd as m
    deck = m.read_deck(deck_path)
    orc = oracle.Oracle(deck.blocks, kn=deck.kn, force=deck.force, limiter=deck.limiter,
                        threads=threads)
    orc.init()
    t0 = time.perf_counter()
    n, log = orc.steps(deck.tout[0], deck.cfl, maxsteps=steps)
    sec = time.perf_counter() - t0
    return {"value": cells * len(log) / sec, "unit": "cell-updates/s", "cores": threads,
            "kind": "port", "sample": f"{n} Heun steps of the same deck (C oracle)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--deck", default=DECK)
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    import multiblockmomentsfv_amd as m
    deck = m.read_deck(args.deck)
    eng = m.Engine.from_deck(deck)
    if args.tile_rows:
        eng.set_tile_rows(args.tile_rows)
    eng.init()
    cells = eng.interior_cells()
    tend = deck.tout[0]

    t, n_w, rej_w, _ = eng.steps(0.0, tend, deck.cfl, maxsteps=args.warmup, reccap=0)
    dt = eng.next_dt
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t, n, rej, _ = eng.steps(t, tend, deck.cfl, maxsteps=args.steps, reccap=0, dt=dt)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if n < args.steps:
        print(f"bench.py: run reached t_end after {n} timed steps", file=sys.stderr)

    # Stage-kernel duration: the same kernels launched one TimeUpdate at a time
    # on the engine's stream, each launch bracketed by HIP events (the timed
    # loop above replays a captured graph, which events cannot bracket).
    eng.enable_timing(True)
    for _ in range(args.kernel_reps):
        eng.time_update(eng.next_dt, t)
    stage_ms, nstage, fill_ms, nfill = eng.kernel_time()
    eng.enable_timing(False)

    attempts = n + rej
    value = cells * attempts / wall
    mean_launch_s = stage_ms / nstage / 1e3
    bytes_per_launch = cells * BYTES_PER_CELL_UPDATE / 2
    achieved = bytes_per_launch / mean_launch_s / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("bytes_per_launch")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": n,
        "warmup": n_w,
        "ms_per_step": 1e3 * wall / max(n, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "square4 topology, 4 blocks x 1024x1024 (4194304 cells), Reg13, "
                        "Kn=0.01, CFL=0.9, harmonic, Riemann2D initial state, adaptive dt",
            "cells": cells, "blocks": deck.nblocks, "kn": deck.kn, "r13": deck.r13,
            "cfl": deck.cfl, "limiter": deck.limiter, "t_start": round(t - 0, 6),
            "rejected_steps": rej, "tile_rows": eng.tile_rows(),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "stage_kernel (stage 1 + stage 2 launches)",
            "mean_launch_ms": mean_launch_s * 1e3,
            "timed_launches": nstage,
            "fill_mean_launch_ms": fill_ms / max(nfill, 1),
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.deck, cells)
    eng.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
