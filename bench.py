"""Benchmark of the explicit update on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--deck PATH] [--no-cpu]
                    [--no-batched] [--no-adjoint]
    python bench.py --calibration-copy | --stream-rates | --valu-rate

Workload: square4 topology with every block 1024x1024 (4 194 304 cells),
Reg13, Kn=0.01, CFL=0.9, harmonic limiter, Riemann2D initial state, FP64,
adaptive dt (multiblockmomentsfv_amd/decks/square4_1024.txt).  A step is one
Heun step attempt (TimeUpdate, then accept + commit or reject) driven by the
Steps controller on the device; the state stays in HBM.

Prints one JSON line: value = cell-updates/s (interior cells x attempts /
wall time of the K timed steps).  roofline: the stage kernel's algorithmic
HBM bytes (stage 1 reads U^n and writes U*, stage 2 reads U*, U^n and writes
U^n+1: 360 B per cell-update, 180 B per cell per launch on average) over its
mean launch time measured with HIP events on the engine's stream.
batched: 256 seeded decks of uneven sizes and Kn advanced together, against
the same decks run one after another.  adjoint: cost of the gradient of a
temperature-matching objective over ~2000 steps on 4x512^2 relative to the
plain forward run, and its agreement with central differences.
cpu_baseline: the reference itself (oracle/_ref) on the host cores, a bounded
sample of the same deck.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

DECK = os.path.join(ROOT, "multiblockmomentsfv_amd", "decks", "square4_1024.txt")
ADJ_DECK = os.path.join(ROOT, "multiblockmomentsfv_amd", "decks", "square4_512.txt")
METRIC = "FP64 cell-updates/s + % HBM roofline; batched & adjoint vs plain step"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_UPDATE = 360.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "stage_kernel_square4_1024_traffic.json")
VALU_FILE = os.path.join(ROOT, "profiles", "stage_kernel_square4_1024_valu.json")
COMPULSORY_BYTES_PER_CELL_UPDATE = 144.0     # U^n in, U^n+1 out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(deck_path, cells, nblocks, serial_steps=5, omp_steps=20):
    """Times the reference build (or the oracle port) on the host cores: a
    serial leg and an OpenMP leg with min(blocks, nproc) threads (the
    reference parallelises over blocks only), each over the first steps of
    the same deck from t=0, step loop only (no I/O)."""
    import oracle
    nproc = os.cpu_count() or 1
    legs = []
    for threads, steps in ((1, serial_steps), (min(nblocks, nproc), omp_steps)):
        if oracle.ref_available():
            env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
                       OMP_PLACES="cores")
            r = subprocess.run([oracle.REF_HARNESS, "steplog", "-bfile", deck_path, "-nsteps",
                                str(steps), "-time"], capture_output=True, text=True, env=env,
                               timeout=900)
            line = [l for l in r.stdout.splitlines() if l.startswith("{")][-1]
            d = json.loads(line)
            legs.append({"threads": d["threads"], "steps": d["steps"], "rejected": d["rejected"],
                         "seconds": d["seconds"], "kind": "reference",
                         "value": d["cells"] * (d["steps"] + d["rejected"]) / d["seconds"]})
        else:
            import multiblockmomentsfv_amd as m
            deck = m.read_deck(deck_path)
            orc = oracle.Oracle(deck.blocks, kn=deck.kn, force=deck.force, limiter=deck.limiter,
                                threads=threads)
            orc.init()
            t0 = time.perf_counter()
            n, log = orc.steps(deck.tout[0], deck.cfl, maxsteps=steps)
            sec = time.perf_counter() - t0
            legs.append({"threads": threads, "steps": n, "rejected": len(log) - n, "seconds": sec,
                         "kind": "port", "value": cells * len(log) / sec})
    best = max(legs, key=lambda l: l["value"])
    return {"value": best["value"], "unit": "cell-updates/s", "cores": best["threads"],
            "kind": best["kind"], "nproc": nproc, "cpu_model": cpu_model(), "legs": legs,
            "sample": (f"first Heun steps of the same deck from t=0, step loop only: serial "
                       f"{legs[0]['steps']} steps ({legs[0]['seconds']:.1f} s), OpenMP over blocks "
                       f"with {legs[1]['threads']} threads {legs[1]['steps']} steps "
                       f"({legs[1]['seconds']:.1f} s); value = the faster leg"
                       + (" (reference Main.c/Flux.c build)" if best["kind"] == "reference"
                          else " (C oracle)"))}


def batched_leg(ndecks, steps, plain_steps, seed=0):
    """256 seeded square4-topology decks (block sides 64..512, Kn log-uniform
    in [1e-3, 1e-1]) advanced together for `steps` accepted steps each, against
    the same decks run one after another through the plain single-deck path."""
    import torch
    import multiblockmomentsfv_amd as m
    decks = m.sweep_decks(ndecks, seed=seed)
    bat = m.Batch(decks)
    bat.init()
    best, best_w, best_ms = bat.autotune_shape(rows=(16, 24, 32, 40, 48, 64, 80, 96), waves=(1, 2, 4, 8),
                                               reps=4)
    bat.init()
    cells = np.array([sum(b["nx"] * b["ny"] for b in d["blocks"]) for d in decks], dtype=np.float64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ns, nr = bat.run(maxsteps=steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    bat.close()
    batched = float((cells * (ns + nr)).sum() / wall)

    plain_cu, plain_s = 0.0, 0.0
    for d, c in zip(decks, cells):
        eng = m.Engine(d["blocks"], kn=d["kn"])
        eng.init()
        t, _, _, _ = eng.steps(0.0, 1e30, d["cfl"], maxsteps=3, reccap=0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, n, rej, _ = eng.steps(t, 1e30, d["cfl"], maxsteps=plain_steps, reccap=0, dt=eng.next_dt)
        torch.cuda.synchronize()
        plain_s += time.perf_counter() - t0
        plain_cu += c * (n + rej)
        eng.close()
    plain = plain_cu / plain_s
    return {"decks": ndecks, "cells": int(cells.sum()), "steps_per_deck": steps,
            "attempts": int((ns + nr).sum()), "value": batched, "unit": "cell-updates/s",
            "seconds": wall, "tile_rows": best, "wave_group": best_w,
            "plain_per_deck": plain, "plain_steps_per_deck": plain_steps,
            "ratio_vs_plain": batched / plain,
            "config": f"seed {seed}: square4 topology, block sides from {{64,...,512}}, "
                      "Kn log-uniform [1e-3,1e-1], CFL 0.95, harmonic, Riemann2D"}


def adjoint_leg(deck_path, fd_dirs, tend=0.3):
    """Gradient of J = sum (theta(t_end) - theta_target)^2 with respect to the
    initial temperature field and Kn on 4x512^2, against central differences."""
    import torch
    import multiblockmomentsfv_amd as m
    from multiblockmomentsfv_amd.adjoint import TemperatureInversion
    # generic=0.1 adds smooth fields of 1-5 % to every primitive: the plain jump
    # starts at rest, where every slope is exactly zero and the limiters sit on
    # their kinks (finite differences there measure the limiter's nonlinear
    # response, not a derivative)
    prob = TemperatureInversion(deck_path, kn=0.01, seed=0, generic=0.1)
    nsteps, _ = prob.record_schedule(tend=tend)
    eng = prob.engine
    U0 = prob.initial_state(prob.theta0)

    def plain_forward():
        # the same steps with the graph-replayed adaptive controller
        eng.layer_set(m.L_DATA, U0)
        eng.set_boundary(m.L_DATA, 0.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.steps(0.0, 1e30, prob.deck.cfl, maxsteps=nsteps, reccap=0)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def gradient():
        th0 = [t.clone().requires_grad_(True) for t in prob.theta0]
        kn = torch.tensor(prob.kn, dtype=torch.float64, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J = prob.objective(th0, kn)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        J.backward()
        torch.cuda.synchronize()
        return th0, kn, t1 - t0, time.perf_counter() - t1

    # one untimed pass of each first: graph capture of the controller and the
    # first-use allocations (checkpoint and segment layers) are set-up costs;
    # then the faster of two timed passes of each
    plain_forward()
    gradient()
    t_plain = min(plain_forward(), plain_forward())
    th0, kn, t_fwd, t_bwd = min((gradient() for _ in range(2)), key=lambda g: g[2] + g[3])
    rng = np.random.default_rng(0)
    errs = []
    # Central differences with h = 1e-8 on theta ~ 1..5: the limiters are only
    # piecewise smooth, and cells whose slopes cross a kink within +-h add an
    # error proportional to h, which falls to the rounding floor near 1e-8. The
    # two objectives are differenced cell by cell.
    h = 1e-8
    from multiblockmomentsfv_amd.adjoint import smooth_fields
    knd = kn.detach()
    for _ in range(fd_dirs):
        # smooth random directions (low Fourier modes): white noise would move
        # every limited slope by O(h) relative to slopes that are themselves
        # small where theta is nearly flat, probing the limiter's curvature
        dirs = []
        for k, t in enumerate(prob.theta0):
            bx, by, dx, dy = prob.deck.geom(k)
            b = prob.deck.blocks[k]
            X, Y = np.meshgrid(bx + dx * np.arange(3, b["nx"] + 3), by + dy * np.arange(3, b["ny"] + 3))
            dirs.append(torch.tensor(smooth_fields(X, Y, rng, 10.0)[0], dtype=torch.float64,
                                     device=t.device))
        tp = prob.final_theta([t + h * d for t, d in zip(prob.theta0, dirs)], knd)
        tm = prob.final_theta([t - h * d for t, d in zip(prob.theta0, dirs)], knd)
        fd = prob.objective_difference(tp, tm) / (2 * h)
        ad = float(sum((g.grad * d).sum() for g, d in zip(th0, dirs)))
        errs.append(abs(ad - fd) / max(abs(fd), 1e-300))
    with torch.no_grad():
        hk = 1e-5
        jp = float(prob.objective(prob.theta0, torch.tensor(prob.kn * (1 + hk), dtype=torch.float64)))
        jm = float(prob.objective(prob.theta0, torch.tensor(prob.kn * (1 - hk), dtype=torch.float64)))
    fdk = (jp - jm) / (2 * hk * prob.kn)
    return {"cells": prob.deck.interior_cells(), "steps": nsteps,
            "plain_forward_s": t_plain, "forward_with_checkpoints_s": t_fwd, "backward_s": t_bwd,
            "gradient_over_forward": (t_fwd + t_bwd) / t_plain, "backward_over_forward": t_bwd / t_plain,
            "fd_directions": fd_dirs, "fd_rel_err_max": max(errs) if errs else None,
            "fd_rel_err_median": float(np.median(errs)) if errs else None,
            "dJ_dKn": float(kn.grad), "dJ_dKn_fd": fdk,
            "dJ_dKn_rel_err": abs(float(kn.grad) - fdk) / abs(fdk),
            "config": "square4 topology 4x512^2, Reg13, Kn 0.01, CFL 0.9, harmonic, "
                      "tanh temperature jump (5 inside r<0.4, 1 outside, width 0.05) plus smooth "
                      "fields of 1-5 % in every primitive, t_end 0.3; FD: smooth random "
                      "directions on theta0 (low Fourier modes), h=1e-8, objectives differenced "
                      "cell by cell; timings: the faster of two passes after one untimed pass"}


def calibration_copy():
    """Five plain device-to-device copies of 256 MiB: the reference dispatches
    against which rocprofv3's FETCH_SIZE / WRITE_SIZE are scaled to bytes."""
    import torch
    x = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibration_copy_bytes": x.numel() * 8, "copies": 5}))


def fp64_issue(op=3, waves=2, launches=200):
    """FP64 VALU issue of this GPU (libmbmprobe.so): ns and shader cycles per
    wave-instruction per SIMD of independent FP64 work over the last half of
    `launches` back-to-back launches (about 70 us each), and the shader clock
    under that load. op 0 fma, 1 mul, 2 add, 3 their mix; waves resident per
    SIMD 1, 2 or 4 (the stage kernels run two)."""
    import ctypes
    import torch  # noqa: F401  (same HIP runtime as the engine)
    path = os.path.join(ROOT, "multiblockmomentsfv_amd", "libmbmprobe.so")
    if not os.path.exists(path):
        raise SystemExit(f"{path} is not built: run `make` at the repository root")
    L = ctypes.CDLL(path)
    d = ctypes.c_double
    L.mbm_probe_fp64_rate.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(d),
                                      ctypes.POINTER(d), ctypes.POINTER(d)]
    ns, cyc, ghz = d(), d(), d()
    rc = L.mbm_probe_fp64_rate(op, waves, launches, ctypes.byref(ns), ctypes.byref(cyc), ctypes.byref(ghz))
    if rc:
        raise RuntimeError(f"mbm_probe_fp64_rate failed ({rc})")
    return ns.value, cyc.value, ghz.value


def valu_rate():
    """The probe over fma, mul, add and their mix at 1, 2 and 4 waves per SIMD."""
    out = {}
    for waves in (1, 2, 4):
        for op, name in enumerate(("fma", "mul", "add", "mix")):
            ns, cyc, ghz = fp64_issue(op, waves)
            out[f"{name}_{waves}w"] = {"ns_per_instr": round(ns, 4), "cycles_per_instr": round(cyc, 3),
                                       "clock_GHz": round(ghz, 3)}
    print(json.dumps({"fp64_valu_issue_per_simd": out}))


def stream_rates():
    """Streaming HBM rates of plain torch kernels: a copy (one read, one write)
    and an add (two reads, one write, the byte mix of stage 2), at the size of
    one 4x1024^2 layer (305 MB per array) and at 2 GiB per array (no reuse
    through the 256 MB memory-side cache)."""
    import torch
    out = {}
    for label, n in (("305MB", 4243600 * 9), ("2GiB", 256 * 1024 * 1024)):
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.ones_like(x)
        z = torch.empty_like(x)
        for name, fn, arrays in (("copy", lambda: z.copy_(x), 2), ("add", lambda: torch.add(x, y, out=z), 3)):
            for _ in range(3):
                fn()
            torch.cuda.synchronize()
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            ev[0].record()
            for _ in range(20):
                fn()
            ev[1].record()
            torch.cuda.synchronize()
            sec = ev[0].elapsed_time(ev[1]) / 1e3 / 20
            out[f"{name}_{label}_TBps"] = arrays * n * 8 / sec / 1e12
        del x, y, z
        torch.cuda.empty_cache()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--clock-warmup", type=float, default=0.3,
                    help="seconds of untimed TimeUpdate calls (state unchanged) before the timed steps")
    ap.add_argument("--deck", default=DECK)
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-decks", type=int, default=256)
    ap.add_argument("--batched-steps", type=int, default=1000)
    ap.add_argument("--plain-steps", type=int, default=50)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--fd-dirs", type=int, default=32)
    ap.add_argument("--calibration-copy", action="store_true",
                    help="only run the 256 MiB reference copies for the PMC calibration")
    ap.add_argument("--stream-rates", action="store_true",
                    help="only measure the streaming copy/add rates")
    ap.add_argument("--valu-rate", action="store_true",
                    help="only measure the FP64 VALU issue rate per SIMD")
    args = ap.parse_args()
    if args.valu_rate:
        valu_rate()
        return
    if args.calibration_copy:
        calibration_copy()
        return
    if args.stream_rates:
        stream_rates()
        return
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
    t_first = t
    import torch
    torch.cuda.synchronize()
    # The GPU's clock rises over the first ~20 ms of sustained load: after only
    # a few warm-up steps the first timed steps run 5-7 % slower than the
    # steady rate, and after a pause of 0.5 s 20 % slower. TimeUpdate calls
    # (TMP and NEW rewritten, DATA and t unchanged, so the timed steps start
    # from the state after the warm-up steps) keep the GPU busy for
    # --clock-warmup seconds before the timed window.
    cw0 = time.perf_counter()
    n_cw = 0
    while time.perf_counter() - cw0 < args.clock_warmup:
        eng.time_update(dt, t)
        n_cw += 1
    torch.cuda.synchronize()
    cw_s = time.perf_counter() - cw0
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
    traffic = None    # measured (rocprofv3 FETCH_SIZE + WRITE_SIZE) for the default deck only
    valu = None       # FP64 issue: rocprofv3 SQ_INSTS_VALU per launch (default deck) and a live probe
    default_deck = os.path.abspath(args.deck) == os.path.abspath(DECK)
    if os.path.exists(TRAFFIC_FILE) and default_deck:
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("bytes_per_launch")
    if os.path.exists(VALU_FILE) and default_deck:
        with open(VALU_FILE) as f:
            vf = json.load(f)
        ns, cyc, ghz = fp64_issue()
        floor_s = vf["insts_per_launch"] / vf["simds"] * ns * 1e-9
        valu = {"insts_per_launch": vf["insts_per_launch"], "probe_ns_per_instr": ns,
                "probe_cycles_per_instr": cyc, "probe_clock_GHz": ghz,
                "floor_ms": floor_s * 1e3, "frac": floor_s / mean_launch_s,
                "note": "SQ_INSTS_VALU per stage launch (rocprofv3, committed profile) at the FP64 "
                        "issue rate the probe measures on this GPU under sustained load, two waves "
                        "per SIMD (bench.py --valu-rate) = floor_ms; frac = floor_ms / mean_launch_ms"}
    if os.path.abspath(args.deck) == os.path.abspath(DECK):
        workload = ("square4 topology, 4 blocks x 1024x1024 (4194304 cells), Reg13, "
                    "Kn=0.01, CFL=0.9, harmonic, Riemann2D initial state, adaptive dt")
    else:
        sizes = ", ".join(f"{b['nx']}x{b['ny']}" for b in deck.blocks)
        workload = (f"{os.path.basename(args.deck)}: {deck.nblocks} blocks ({sizes}), {cells} cells, "
                    f"Kn={deck.kn:g}, CFL={deck.cfl:g}, {deck.limiter}, Riemann2D initial state, "
                    "adaptive dt")
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
            "workload": workload,
            "cells": cells, "blocks": deck.nblocks, "kn": deck.kn, "r13": deck.r13,
            "cfl": deck.cfl, "limiter": deck.limiter, "t_window": [round(t_first, 6), round(t, 6)],
            "rejected_steps": rej, "tile_rows": eng.tile_rows(),
            "clock_warmup": {"seconds": round(cw_s, 3), "time_updates": n_cw,
                             "note": "untimed TimeUpdate calls after the warm-up steps; DATA and t "
                                     "unchanged"},
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
            "compulsory_frac": cells * COMPULSORY_BYTES_PER_CELL_UPDATE / 2 / mean_launch_s / 1e9 / HBM_PEAK_GBS,
            "valu": valu,
        },
    }
    eng.close()
    if not args.no_batched:
        out["batched"] = batched_leg(args.batched_decks, args.batched_steps, args.plain_steps)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg(ADJ_DECK, args.fd_dirs)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.deck, cells, deck.nblocks)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
