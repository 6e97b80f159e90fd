"""Numeric comparison of two .hdr/.dat solution files.

Same rules and exit codes as the reference's regression gate
(tools/compare_solution.py:30-122):
  * each argument may be the .hdr, the .dat or the shared basename;
  * payload precision from the [BINARY] PRECISION field (Float64 if absent);
  * zones are matched by their (block, I, J, NVARS) fingerprint, STRANDID,
    SOLUTIONTIME and the provenance lines are ignored;
  * a value passes if |new - ref| <= abstol + reltol*|ref|; the defaults are
    1e-6/1e-5 when either side is Float32, else 1e-10/1e-8;
  * exit 0 = PASS, 1 = FAIL, 2 = usage.

    python -m multiblockmomentsfv_amd.compare_solution <new> <ref> [abstol] [reltol]
"""
import re
import sys

import numpy as np


def companion(path):
    base = path[:-4] if path.endswith((".hdr", ".dat")) else path
    return base + ".hdr", base + ".dat"


def precision(hdr):
    with open(hdr) as f:
        for line in f:
            m = re.match(r"\s*PRECISION\s*=\s*(\S+)", line)
            if m:
                return m.group(1)
    return "Float64"


def zones(hdr):
    out = []
    with open(hdr) as f:
        for line in f:
            s = line.strip()
            if not s.startswith("ZONE"):
                continue

            def g(k):
                return int(re.search(rf"\b{k}\s*=\s*(-?\d+)", s).group(1))

            out.append(dict(block=g("block"), I=g("I"), J=g("J"), nvars=g("NVARS"),
                            offset=g("OFFSET"), nbytes=g("NBYTES")))
    return out


def load(path):
    """(fingerprint list, float64 values over all zones, precision)."""
    hdr, dat = companion(path)
    prec = precision(hdr)
    dtype = np.dtype("<f4") if prec == "Float32" else np.dtype("<f8")
    zs = zones(hdr)
    keys = [(z["block"], z["I"], z["J"], z["nvars"]) for z in zs]
    parts = []
    with open(dat, "rb") as f:
        for z in zs:
            f.seek(z["offset"])
            n = z["nvars"] * z["I"] * z["J"]
            parts.append(np.frombuffer(f.read(n * dtype.itemsize), dtype=dtype, count=n))
    vals = np.concatenate(parts).astype(np.float64) if parts else np.zeros(0)
    return keys, vals, prec


def compare(new, ref, abstol=None, reltol=None, out=print):
    kn, dn, pn = load(new)
    kr, dr, pr = load(ref)
    single = "Float32" in (pn, pr)
    abstol = (1e-6 if single else 1e-10) if abstol is None else abstol
    reltol = (1e-5 if single else 1e-8) if reltol is None else reltol
    if kn != kr:
        out(f"FAIL: zone structure differs ({len(kn)} vs {len(kr)} zones)")
        for a, b in zip(kn, kr):
            if a != b:
                out(f"   new: {a}\n   ref: {b}")
                break
        return 1
    if dn.size != dr.size:
        out(f"FAIL: value count differs ({dn.size} vs {dr.size})")
        return 1
    d = np.abs(dn - dr)
    bad = np.nonzero(~(d <= abstol + reltol * np.abs(dr)))[0]
    if bad.size:
        i = bad[0]
        out(f"FAIL: value mismatch at index {i}: {dn[i]:.12g} vs {dr[i]:.12g} (|diff| = {d[i]:.3e})")
        return 1
    i = int(np.argmax(d)) if d.size else 0
    mx = float(d[i]) if d.size else 0.0
    out(f"PASS: {dn.size} values match within tol (abstol={abstol:g}, reltol={reltol:g}, "
        f"precision new={pn}/ref={pr}); max |diff| = {mx:.3e}")
    if mx > 0:
        out(f"      largest diff at index {i}: {dn[i]:.12g} vs {dr[i]:.12g}")
    return 0


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) < 2:
        print("usage: compare_solution.py <new> <reference> [abstol] [reltol]")
        return 2
    a = float(argv[2]) if len(argv) > 2 else None
    r = float(argv[3]) if len(argv) > 3 else None
    return compare(argv[0], argv[1], a, r)


if __name__ == "__main__":
    sys.exit(main())
