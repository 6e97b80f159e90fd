"""Regression check of one solution file pair against another.

Usage (the argument order and exit codes of the reference's `make check`
gate, tools/compare_solution.py:80-122):

    python -m multiblockmomentsfv_amd.compare_solution NEW REF [ABSTOL [RELTOL]]

NEW and REF each name a solution written by the output writer: the `.hdr`
manifest, the `.dat` payload, or their common stem.  The contract kept from
the reference gate:

* the payload precision is the manifest's `PRECISION = ...` entry, Float64
  when the entry is missing;
* the zones must have the same structure, zone by zone in file order:
  (block, I, J, NVARS).  Everything else in the manifest (STRANDID,
  SOLUTIONTIME, provenance) is free to differ;
* a value fails when |new - ref| > abstol + reltol * |ref| (so a NaN on
  either side does not fail, as in the reference gate);
* default tolerances: 1e-6 / 1e-5 if either payload is Float32, otherwise
  1e-10 / 1e-8;
* exit status 0 = every value within tolerance, 1 = a mismatch, 2 = fewer
  than two paths (arguments after RELTOL are ignored).

Implementation: one pass over each manifest builds a `Solution` (precision
and zone table); each zone's values are mapped from the payload with
np.memmap and compared as float64 arrays, zone by zone.
"""
import os
import sys
from dataclasses import dataclass, field

import numpy as np

STATUS_PASS, STATUS_FAIL, STATUS_USAGE = 0, 1, 2

# tolerance defaults keyed by "is either payload single precision"
DEFAULT_TOL = {True: (1e-6, 1e-5), False: (1e-10, 1e-8)}


@dataclass
class Zone:
    block: int
    ni: int
    nj: int
    nvars: int
    offset: int

    @property
    def shape_key(self):
        return (self.block, self.ni, self.nj, self.nvars)

    @property
    def count(self):
        return self.ni * self.nj * self.nvars


@dataclass
class Solution:
    stem: str
    precision: str = "Float64"
    zones: list = field(default_factory=list)

    @property
    def dtype(self):
        return np.dtype("<f4") if self.precision == "Float32" else np.dtype("<f8")

    def values(self, zone):
        """The zone's payload as a float64 array (memory-mapped read)."""
        if zone.count == 0:
            return np.zeros(0)
        m = np.memmap(self.stem + ".dat", dtype=self.dtype, mode="r", offset=zone.offset,
                      shape=(zone.count,))
        return np.asarray(m, dtype=np.float64)


def _stem(path):
    root, ext = os.path.splitext(path)
    return root if ext in (".hdr", ".dat") else path


def _fields(line):
    """KEY=VALUE tokens of a manifest line (spaces around '=' allowed)."""
    out = {}
    toks = line.replace("=", " = ").split()
    for i in range(1, len(toks) - 1):
        if toks[i] == "=":
            out[toks[i - 1]] = toks[i + 1]
    return out


def read_solution(path):
    sol = Solution(_stem(path))
    with open(sol.stem + ".hdr") as f:
        for raw in f:
            line = raw.strip()
            if line.startswith("ZONE"):
                kv = _fields(line)
                sol.zones.append(Zone(int(kv["block"]), int(kv["I"]), int(kv["J"]),
                                      int(kv["NVARS"]), int(kv["OFFSET"])))
            elif line.startswith("PRECISION"):
                kv = _fields(line)
                if "PRECISION" in kv:
                    sol.precision = kv["PRECISION"]
    return sol


def compare(new, ref, abstol=None, reltol=None, out=print):
    """Compares two solutions; returns the exit status (0 pass, 1 fail) and
    reports through `out`."""
    a, b = read_solution(new), read_solution(ref)
    single = "Float32" in (a.precision, b.precision)
    d_abs, d_rel = DEFAULT_TOL[single]
    abstol = d_abs if abstol is None else abstol
    reltol = d_rel if reltol is None else reltol

    ka = [z.shape_key for z in a.zones]
    kb = [z.shape_key for z in b.zones]
    if ka != kb:
        first = next((i for i, (x, y) in enumerate(zip(ka, kb)) if x != y), min(len(ka), len(kb)))
        out(f"FAIL structure: {len(ka)} zones against {len(kb)}; first difference at zone {first}: "
            f"{ka[first] if first < len(ka) else None} against {kb[first] if first < len(kb) else None} "
            "(block, I, J, NVARS)")
        return STATUS_FAIL

    seen, worst = 0, (0.0, -1, 0.0, 0.0)
    for za, zb in zip(a.zones, b.zones):
        x, y = a.values(za), b.values(zb)
        diff = np.abs(x - y)
        over = np.flatnonzero(diff > abstol + reltol * np.abs(y))
        if over.size:
            j = int(over[0])
            out(f"FAIL value: block {za.block} zone value {j} (flat index {seen + j}): new "
                f"{x[j]:.12g}, ref {y[j]:.12g}, |diff| {diff[j]:.3e} > {abstol:g} + {reltol:g}*|ref|")
            return STATUS_FAIL
        if diff.size:
            j = int(np.nanargmax(diff)) if np.isfinite(diff).any() else 0
            if diff[j] > worst[0]:
                worst = (float(diff[j]), seen + j, float(x[j]), float(y[j]))
        seen += x.size

    out(f"PASS {seen} values of {len(ka)} zones (new {a.precision}, ref {b.precision}; "
        f"abstol {abstol:g}, reltol {reltol:g}); largest |diff| {worst[0]:.3e}")
    if worst[0] > 0:
        out(f"     at flat index {worst[1]}: new {worst[2]:.12g}, ref {worst[3]:.12g}")
    return STATUS_PASS


def main(argv=None):
    args = sys.argv[1:] if argv is None else list(argv)
    if len(args) < 2:
        print("usage: python -m multiblockmomentsfv_amd.compare_solution NEW REF [ABSTOL [RELTOL]]")
        return STATUS_USAGE
    try:
        tol = [float(v) for v in args[2:4]]
    except ValueError:
        print(f"compare_solution: tolerances must be numbers, got {args[2:4]}")
        return STATUS_USAGE
    tol += [None] * (2 - len(tol))
    return compare(args[0], args[1], tol[0], tol[1])


if __name__ == "__main__":
    sys.exit(main())
