"""CPU oracle for the explicit update of MultiBlockMomentsFV (test infrastructure).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
use this package.  It wraps ``oracle/liboracle.so`` (a C restatement of the
reference path, see ``mbm_oracle.h``) and, where it was built, the reference
itself under ``oracle/_ref`` (``ref_harness``).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")
REF_HARNESS_FFS = os.path.join(HERE, "_ref", "ref_harness_ffs")

LIMITERS = {"none": 0, "average": 1, "minmod": 2, "superbee": 3, "harmonic": 4}
SCENARIOS = {"riemann2d": 0, "ffs": 1, "tempjump": 2}

_lib = None


def build():
    """Compile liboracle.so, and the reference build when its sources exist."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        d, i, p = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        L.orc_create.restype = p
        L.orc_create.argtypes = [i, p, p, p, p, p, p, p, d, d, i, i]
        L.orc_destroy.argtypes = [p]
        L.orc_set_threads.argtypes = [i]
        L.orc_set_tempjump.argtypes = [p, d, d, d, d, d]
        L.orc_init.argtypes = [p]
        L.orc_set_boundary.argtypes = [p, i, d]
        L.orc_time_update.restype = d
        L.orc_time_update.argtypes = [p, d, d]
        L.orc_commit.argtypes = [p]
        L.orc_update_block.restype = d
        L.orc_update_block.argtypes = [p, i, d, d, i]
        L.orc_steps.restype = i
        L.orc_steps.argtypes = [p, p, d, d, d, i, p, i, p, p]
        L.orc_block_cells.restype = i
        L.orc_block_cells.argtypes = [p, i]
        L.orc_get_layer.argtypes = [p, i, i, p]
        L.orc_set_layer.argtypes = [p, i, i, p]
        L.orc_output_values.argtypes = [p, i, p]
        L.orc_primitive.argtypes = [p, p]
        L.orc_conservative.argtypes = [p, p]
        L.orc_flux_dir.argtypes = [p, p, d, d]
        L.orc_relax.restype = d
        L.orc_relax.argtypes = [p, p, p, d, i]
        L.orc_limit.restype = d
        L.orc_limit.argtypes = [i, d, d]
        L.orc_hll.restype = d
        L.orc_hll.argtypes = [p, p, d, i, p]
        L.orc_initial_state.argtypes = [p, d, d, d, p]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """One multi-block deck held on the CPU, advanced by the oracle."""

    def __init__(self, blocks, kn=0.001, force=0.0, limiter="harmonic",
                 scenario="riemann2d", threads=None):
        L = lib()
        self.blocks = blocks
        nb = len(blocks)
        self.nx = np.array([b["nx"] for b in blocks], dtype=np.int32)
        self.ny = np.array([b["ny"] for b in blocks], dtype=np.int32)
        arrs = [np.array([b[k] for b in blocks], dtype=np.float64)
                for k in ("xmin", "xmax", "ymin", "ymax")]
        bc = np.array([b["bc"] for b in blocks], dtype=np.int32).reshape(-1)
        lim = LIMITERS[limiter] if isinstance(limiter, str) else int(limiter)
        scn = SCENARIOS[scenario] if isinstance(scenario, str) else int(scenario)
        if threads:
            L.orc_set_threads(int(threads))
        self._keep = (self.nx, self.ny, arrs, bc)
        self.h = L.orc_create(nb, _ptr(self.nx), _ptr(self.ny), *[_ptr(a) for a in arrs],
                              _ptr(bc), float(kn), float(force), lim, scn)
        if not self.h:
            raise ValueError("orc_create failed")
        self.t = 0.0

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_tempjump(self, th_in, th_out, radius, width, rho):
        lib().orc_set_tempjump(self.h, th_in, th_out, radius, width, rho)

    def init(self):
        lib().orc_init(self.h)
        self.t = 0.0

    def shape(self, k):
        return (int(self.ny[k]) + 6, int(self.nx[k]) + 6, 9)

    def get(self, layer=0):
        out = []
        for k in range(len(self.blocks)):
            a = np.empty(self.shape(k), dtype=np.float64)
            lib().orc_get_layer(self.h, layer, k, _ptr(a))
            out.append(a)
        return out

    def set(self, arrays, layer=0):
        for k, a in enumerate(arrays):
            a = np.ascontiguousarray(a, dtype=np.float64)
            assert a.shape == self.shape(k)
            lib().orc_set_layer(self.h, layer, k, _ptr(a))

    def set_boundary(self, layer, t):
        lib().orc_set_boundary(self.h, layer, t)

    def time_update(self, dt, t):
        return lib().orc_time_update(self.h, dt, t)

    def commit(self):
        lib().orc_commit(self.h)

    def steps(self, tend, cflreg, cflmax=1.0, maxsteps=0, logcap=100000):
        """Runs the Steps controller; returns (accepted steps, log array)."""
        t = ctypes.c_double(self.t)
        log = np.zeros((logcap, 4))
        nlog = ctypes.c_int(0)
        nrej = ctypes.c_int(0)
        n = lib().orc_steps(self.h, ctypes.byref(t), tend, cflmax, cflreg, maxsteps,
                            _ptr(log), logcap, ctypes.byref(nlog), ctypes.byref(nrej))
        self.t = t.value
        return n, log[:min(nlog.value, logcap)]

    def output_values(self, k):
        n = (int(self.nx[k]) + 2) * (int(self.ny[k]) + 2)
        v = np.empty((n, 10), dtype=np.float64)
        lib().orc_output_values(self.h, k, _ptr(v))
        return v


def ref_available(ffs=False):
    return os.access(REF_HARNESS_FFS if ffs else REF_HARNESS, os.X_OK)


def ref_steplog(deck_path, nsteps, out_prefix, ffs=False, threads=1):
    """Runs the reference build for nsteps; returns (log, state bytes)."""
    exe = REF_HARNESS_FFS if ffs else REF_HARNESS
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    subprocess.run([exe, "steplog", "-bfile", deck_path, "-nsteps", str(nsteps),
                    "-dump", out_prefix], check=True, env=env,
                   stdout=subprocess.DEVNULL)
    log = []
    with open(out_prefix + ".log") as f:
        for line in f:
            a = line.split()
            log.append((float.fromhex(a[0]), float.fromhex(a[1]),
                        float.fromhex(a[2]), int(a[3])))
    state = np.fromfile(out_prefix + ".bin", dtype=np.float64)
    return np.array(log), state
