"""The drop-in CLI (build/moments) against the reference's own main on the
flag paths and the restart reader (Main.c:61-164, tools.c:38-46,
Mesh.c:156-157,198-283, Limiter.c:66-74, Boundary.c:101-103).

Fixtures: tests/golden/cli/<case>, written by tests/golden/make_cli_fixtures.py
from the unchanged reference main (output files, stdout, stderr). Bar: the
same output file names and zone records, the comparator's pass at the
payload's precision defaults, FP64 payloads within 1e-12 relative L-inf per
output variable (Float32 payloads within one float rounding), the same step
count and 20-step log lines, and the same warning text.
"""
import lzma
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

CLI = os.path.join(GOLDEN, "cli")
CASES = sorted(os.listdir(CLI)) if os.path.isdir(CLI) else []
LOG = re.compile(r"step:\s*(\d+)\.\.\. done\.  CFL: ([0-9.]+), time: ([0-9.]+), next dt: ([0-9.]+)")


def run_moments(args, cwd):
    exe = os.path.join(ROOT, "build", "moments")
    return subprocess.run([exe] + args, cwd=cwd, capture_output=True, text=True, timeout=600)


def payload(hdr_path, dat_bytes):
    """(zone records, values as float64 per zone) of an output pair"""
    text = open(hdr_path).read()
    prec = re.search(r"PRECISION\s*=\s*(\S+)", text).group(1)
    dt = np.dtype("<f4") if prec == "Float32" else np.dtype("<f8")
    zones = [l for l in text.splitlines() if l.startswith("ZONE")]
    vals = []
    for z in zones:
        g = {k: int(re.search(rf"\b{k}=(\d+)", z).group(1)) for k in ("I", "J", "NVARS", "OFFSET")}
        n = g["I"] * g["J"] * g["NVARS"]
        vals.append(np.frombuffer(dat_bytes, dtype=dt, count=n, offset=g["OFFSET"]).astype(np.float64)
                    .reshape(-1, g["NVARS"]))
    return zones, vals, prec


@pytest.mark.parametrize("case", CASES)
def test_cli_case_matches_the_reference_main(mbm, case):
    from multiblockmomentsfv_amd.compare_solution import compare
    ref_dir = os.path.join(CLI, case)
    args = [a if not a.startswith("tests/") else os.path.join(ROOT, a)
            for a in open(os.path.join(ref_dir, "args.txt")).read().split()]
    with tempfile.TemporaryDirectory() as tmp:
        r = run_moments(args + ["-out", tmp], tmp)
        assert r.returncode == 0, r.stdout + r.stderr
        ref_hdrs = sorted(f for f in os.listdir(ref_dir) if f.endswith(".hdr"))
        assert sorted(f for f in os.listdir(tmp) if f.endswith(".hdr")) == ref_hdrs
        for hdr in ref_hdrs:
            base = hdr[:-4]
            ref_dat = lzma.decompress(open(os.path.join(ref_dir, base + ".dat.xz"), "rb").read())
            zr, vr, pr = payload(os.path.join(ref_dir, hdr), ref_dat)
            zm, vm, pm = payload(os.path.join(tmp, hdr), open(os.path.join(tmp, base + ".dat"), "rb").read())
            assert pm == pr and zm == zr      # precision and ZONE records (STRANDID, SOLUTIONTIME)
            with open(os.path.join(tmp, "ref.hdr"), "w") as f:
                f.write(open(os.path.join(ref_dir, hdr)).read())
            with open(os.path.join(tmp, "ref.dat"), "wb") as f:
                f.write(ref_dat)
            msgs = []
            assert compare(os.path.join(tmp, base), os.path.join(tmp, "ref"), out=msgs.append) == 0, msgs
            num = np.max([np.abs(a - b).max(axis=0) for a, b in zip(vm, vr)], axis=0)
            den = np.max([np.abs(b).max(axis=0) for b in vr], axis=0)
            err = num / np.where(den > 0, den, 1.0)
            assert err.max() <= (1e-12 if pr == "Float64" else 1.2e-7), err
    ref_out = open(os.path.join(ref_dir, "stdout.txt")).read()
    done = re.search(r"Done after \d+ steps\.", ref_out).group(0)
    assert done in r.stdout
    assert LOG.findall(r.stdout) == LOG.findall(ref_out)
    for line in open(os.path.join(ref_dir, "stderr.txt")).read().splitlines():
        assert line in r.stderr, (line, r.stderr)
    assert re.search(r"Running: .*limiter = \w+", r.stdout).group(0) == \
        re.search(r"Running: .*limiter = \w+", ref_out).group(0)


def test_nonconforming_deck_aborts_like_the_reference(mbm):
    """A sync interface between edges of different length: the reference
    fails its assert in FillEdge (Boundary.c:101-103) and aborts."""
    deck = ("Kn = 0.01\nt_list = 0.1\ndt_out = 0\n\nBlock data (Maximal NB blocks!):\n"
            "number/xmin,xmax/ymin,ymax/nstx,nsty/leftbound,rightbound,bottombound,topbound/\n"
            "1/-1.0,0.0/-1.0,1.0/10,10/-1,2,-1,-1/\n2/0.0,1.0/-1.0,1.0/10,12/1,-1,-1,-1/\n")
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "nonconf.txt")
        open(path, "w").write(deck)
        r = run_moments(["-bfile", path, "-out", os.path.join(tmp, "o")], tmp)
    assert r.returncode == -6          # SIGABRT, exit status 134 in a shell
    assert r.stderr.strip() == ("moments: src/Boundary.c:102: FillEdge: Assertion `(side<2) ? "
                                "(Rd.rbt.j-Rd.lfb.j)==(Rs.rbt.j-Rs.lfb.j) : "
                                "(Rd.rbt.i-Rd.lfb.i)==(Rs.rbt.i-Rs.lfb.i)' failed.")
