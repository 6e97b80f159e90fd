"""MI355X (gfx950) build of the explicit update of MultiBlockMomentsFV.

Python mirror of the C ABI in ``include/mbmfv.h`` and ``include/mbmfv_host.h``
(loaded from the in-tree ``libmbmfv.so`` with ctypes).  The engine keeps the
three FP64 layers DATA/TMP/NEW on the GPU; the host only sees the CFL scalar
per step and the fields at output times.

    deck = read_deck("cases/square4.txt")
    eng = Engine.from_deck(deck)
    eng.init()
    t, nsteps, nrej, recs = eng.steps(0.0, deck.tout[0], cflreg=deck.cfl)
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmbmfv.so")
MOMENTS_BIN = os.path.join(os.path.dirname(HERE), "build", "moments")

NV = 9
MBC = 3
MAX_BLOCKS = 32
MAX_OUT = 50
L_DATA, L_TMP, L_NEW = 0, 1, 2

LIMITERS = {"none": 0, "average": 1, "minmod": 2, "superbee": 3, "harmonic": 4}
LIMITER_NAMES = {v: k for k, v in LIMITERS.items()}
SCENARIOS = {"riemann2d": 0, "ffs": 1, "tempjump": 2}
TEMPJUMP_DEFAULT = (5.0, 1.0, 0.4, 0.05, 1.0)


class MbmError(RuntimeError):
    pass


class BlockDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("xmin", ctypes.c_double), ("xmax", ctypes.c_double),
                ("ymin", ctypes.c_double), ("ymax", ctypes.c_double),
                ("bc", ctypes.c_int * 4)]


class Params(ctypes.Structure):
    _fields_ = [("kn", ctypes.c_double), ("force", ctypes.c_double),
                ("limiter", ctypes.c_int), ("scenario", ctypes.c_int),
                ("scenario_params", ctypes.c_double * 5)]


class StepRecord(ctypes.Structure):
    _fields_ = [("cfl", ctypes.c_double), ("dt", ctypes.c_double), ("t", ctypes.c_double),
                ("accepted", ctypes.c_int), ("step", ctypes.c_int)]


class DeckStruct(ctypes.Structure):
    _fields_ = [("nblocks", ctypes.c_int),
                ("nx", ctypes.c_int * MAX_BLOCKS), ("ny", ctypes.c_int * MAX_BLOCKS),
                ("xmin", ctypes.c_double * MAX_BLOCKS), ("xmax", ctypes.c_double * MAX_BLOCKS),
                ("ymin", ctypes.c_double * MAX_BLOCKS), ("ymax", ctypes.c_double * MAX_BLOCKS),
                ("bc", (ctypes.c_int * 4) * MAX_BLOCKS),
                ("kn", ctypes.c_double), ("r13", ctypes.c_double),
                ("force", ctypes.c_double), ("cfl", ctypes.c_double),
                ("limiter", ctypes.c_int), ("refine", ctypes.c_int),
                ("nout", ctypes.c_int), ("tout", ctypes.c_double * (MAX_OUT + 1))]


class Geom(ctypes.Structure):
    _fields_ = [("bx", ctypes.c_double), ("by", ctypes.c_double),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double)]


_lib = None


def lib():
    """The loaded libmbmfv.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MbmError(f"{LIB_PATH} is not built: run `make` at the repository root")
    # PyTorch ships its own HIP runtime whose soname matches the one the
    # engine links against. Loading torch first makes the engine bind to that
    # same runtime, so both share one device context and torch tensors can be
    # handed to the engine; loaded the other way round, two runtimes end up in
    # the process and torch cannot open the device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    p, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    sigs = {
        "mbm_last_error": (ctypes.c_char_p, []),
        "mbm_create": (i, [ctypes.POINTER(BlockDesc), i, ctypes.POINTER(Params), p,
                           ctypes.POINTER(p)]),
        "mbm_destroy": (i, [p]),
        "mbm_set_stream": (i, [p, p]),
        "mbm_set_params": (i, [p, ctypes.POINTER(Params)]),
        "mbm_set_tile_rows": (i, [p, i]),
        "mbm_tile_rows": (i, [p]),
        "mbm_set_wave_group": (i, [p, i]),
        "mbm_wave_group": (i, [p]),
        "mbm_init": (i, [p]),
        "mbm_upload": (i, [p, i, i, p]),
        "mbm_download": (i, [p, i, i, p]),
        "mbm_layer_rows": (i, [p, i, p, i]),
        "mbm_row_order": (i, [i, i, p]),
        "mbm_layer_ptr": (i, [p, i, i, ctypes.POINTER(p), ctypes.POINTER(ctypes.c_longlong),
                              ctypes.POINTER(i)]),
        "mbm_set_boundary": (i, [p, i, d]),
        "mbm_time_update": (i, [p, d, d, ctypes.POINTER(d)]),
        "mbm_commit": (i, [p]),
        "mbm_steps": (i, [p, ctypes.POINTER(d), d, d, d, i, ctypes.POINTER(d),
                          ctypes.POINTER(StepRecord), i,
                          ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i)]),
        "mbm_output_values": (i, [p, i, i, p]),
        "mbm_interior_cells": (ctypes.c_longlong, [p]),
        "mbm_block_dims": (i, [p, i, ctypes.POINTER(i), ctypes.POINTER(i)]),
        "mbm_device": (i, [p]),
        "mbm_enable_timing": (i, [p, i]),
        "mbm_kernel_time": (i, [p, ctypes.POINTER(d), ctypes.POINTER(i), ctypes.POINTER(d),
                                ctypes.POINTER(i)]),
        "mbm_layer_elems": (ctypes.c_longlong, [p]),
        "mbm_layer_get": (i, [p, i, p, ctypes.c_longlong]),
        "mbm_layer_set": (i, [p, i, p, ctypes.c_longlong]),
        "mbm_adj_forward": (i, [p, d, d, d, d, i, p, i, i, ctypes.POINTER(i), ctypes.POINTER(d)]),
        "mbm_adj_schedule": (i, [p, p, i, ctypes.POINTER(i)]),
        "mbm_adj_backward": (i, [p, p, p, ctypes.c_longlong, ctypes.POINTER(d)]),
        "mbm_adj_release": (i, [p]),
        "mbm_adj_set_budget": (i, [p, ctypes.c_longlong]),
        "mbm_adj_memory": (i, [p, ctypes.POINTER(ctypes.c_longlong)]),
        "mbm_batch_create": (i, [ctypes.POINTER(BlockDesc), p, i, ctypes.POINTER(Params), p,
                                 ctypes.POINTER(p)]),
        "mbm_batch_ndecks": (i, [p]),
        "mbm_batch_run": (i, [p, p, p, d, i, p, p, p]),
        "mbm_batch_download": (i, [p, i, i, p]),
        "mbm_batch_autotune": (i, [p, p, i, i, ctypes.POINTER(i), ctypes.POINTER(d)]),
        "mbm_batch_autotune_shape": (i, [p, p, i, p, i, i, ctypes.POINTER(i), ctypes.POINTER(i),
                                         ctypes.POINTER(d)]),
        "mbm_host_set_boundary": (i, [ctypes.POINTER(BlockDesc), i, ctypes.POINTER(Params),
                                      ctypes.POINTER(p), d]),
        "mbm_host_last_error": (ctypes.c_char_p, []),
        "mbm_deck_defaults": (None, [ctypes.POINTER(DeckStruct)]),
        "mbm_deck_read": (i, [ctypes.c_char_p, ctypes.POINTER(DeckStruct), i, i]),
        "mbm_deck_check": (i, [ctypes.POINTER(DeckStruct)]),
        "mbm_deck_geom": (Geom, [ctypes.POINTER(DeckStruct), i]),
        "mbm_full_grid_size": (i, [ctypes.POINTER(DeckStruct)]),
        "mbm_limiter_code": (i, [ctypes.c_char_p]),
        "mbm_limiter_name": (ctypes.c_char_p, [i]),
        "mbm_convert": (ctypes.c_char_p, [d, ctypes.c_char_p]),
        "mbm_initial_state": (None, [i, p, d, d, d, p]),
        "mbm_output_write": (i, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(DeckStruct),
                                 i, d, i, ctypes.POINTER(p)]),
        "mbm_restart_read": (i, [ctypes.c_char_p, ctypes.POINTER(DeckStruct), ctypes.POINTER(p),
                                 ctypes.c_char_p]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols():
    """Names the C headers declare (checked against the library by the tests)."""
    return [n for n in ("mbm_last_error", "mbm_create", "mbm_destroy", "mbm_set_stream",
                        "mbm_set_params", "mbm_set_tile_rows", "mbm_tile_rows", "mbm_set_wave_group",
                        "mbm_wave_group", "mbm_init", "mbm_upload",
                        "mbm_download", "mbm_layer_ptr", "mbm_layer_rows", "mbm_row_order",
                        "mbm_set_boundary",
                        "mbm_time_update", "mbm_commit", "mbm_steps", "mbm_output_values",
                        "mbm_interior_cells", "mbm_block_dims", "mbm_device", "mbm_enable_timing",
                        "mbm_kernel_time", "mbm_layer_elems", "mbm_layer_get", "mbm_layer_set",
                        "mbm_adj_forward", "mbm_adj_schedule", "mbm_adj_backward",
                        "mbm_adj_release", "mbm_adj_set_budget", "mbm_adj_memory",
                        "mbm_batch_create",
                        "mbm_batch_ndecks",
                        "mbm_batch_run", "mbm_batch_download", "mbm_batch_autotune",
                        "mbm_batch_autotune_shape")]


def _check(rc, what):
    if rc != 0:
        msg = lib().mbm_last_error()
        raise MbmError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def layer_to_blocks(U, index, blocks):
    """Per block (ny+6, nx+6, 9) arrays of a whole-layer array, given the
    layout index (offset, plane, pitch, rows) of every block."""
    out = []
    for (off, plane, pitch, rows), b in zip(index, blocks):
        npos = int(rows.max()) + 1
        blk = U[off:off + npos * pitch].reshape(npos, NV, plane)
        out.append(np.ascontiguousarray(blk[rows, :, :b["nx"] + 6].transpose(0, 2, 1)))
    return out


def blocks_to_layer(arrays, index, blocks, out):
    """Writes per block (ny+6, nx+6, 9) arrays into the whole-layer array out."""
    for (off, plane, pitch, rows), b, a in zip(index, blocks, arrays):
        npos = int(rows.max()) + 1
        blk = out[off:off + npos * pitch].reshape(npos, NV, plane)
        blk[rows, :, :b["nx"] + 6] = np.asarray(a).transpose(0, 2, 1)
    return out


def row_order(ny, band):
    """Row positions of a block of ny interior rows in the device layout for
    tiles of `band` rows (no device needed): (pos per row, number of positions)."""
    pos = np.empty(ny + 6, dtype=np.int32)
    n = lib().mbm_row_order(int(ny), int(band), _ptr(pos))
    if n < 0:
        _check(n, "mbm_row_order")
    return pos.astype(np.int64), n


# ---------------------------------------------------------------- decks
class Deck:
    """A parsed block file (Mesh.c:60-184)."""

    def __init__(self, st):
        self._st = st
        n = st.nblocks
        self.nblocks = n
        self.blocks = [dict(nx=st.nx[k], ny=st.ny[k], xmin=st.xmin[k], xmax=st.xmax[k],
                            ymin=st.ymin[k], ymax=st.ymax[k], bc=list(st.bc[k]))
                       for k in range(n)]
        self.kn, self.r13, self.force, self.cfl = st.kn, st.r13, st.force, st.cfl
        self.limiter = LIMITER_NAMES[st.limiter]
        self.refine = st.refine
        self.tout = [st.tout[i] for i in range(st.nout)]

    def struct(self):
        st = DeckStruct()
        lib().mbm_deck_defaults(ctypes.byref(st))
        st.nblocks = self.nblocks
        for k, b in enumerate(self.blocks):
            st.nx[k], st.ny[k] = b["nx"], b["ny"]
            st.xmin[k], st.xmax[k], st.ymin[k], st.ymax[k] = b["xmin"], b["xmax"], b["ymin"], b["ymax"]
            for s in range(4):
                st.bc[k][s] = b["bc"][s]
        st.kn, st.r13, st.force, st.cfl = self.kn, self.r13, self.force, self.cfl
        st.limiter = LIMITERS[self.limiter]
        st.refine = self.refine
        st.nout = len(self.tout)
        for i, t in enumerate(self.tout):
            st.tout[i] = t
        st.tout[len(self.tout)] = 0.0
        return st

    def geom(self, k):
        g = lib().mbm_deck_geom(ctypes.byref(self.struct()), k)
        return g.bx, g.by, g.dx, g.dy

    def interior_cells(self):
        return sum(b["nx"] * b["ny"] for b in self.blocks)


def read_deck(path, refine=0, verbose=False):
    st = DeckStruct()
    rc = lib().mbm_deck_read(path.encode(), ctypes.byref(st), int(refine), int(verbose))
    if rc != 0:
        raise MbmError(f"reading {path}: {lib().mbm_host_last_error().decode()}")
    return Deck(st)


def make_deck(blocks, kn=0.001, r13=0.0, force=0.0, cfl=0.95, limiter="harmonic", tout=(0.3,)):
    st = DeckStruct()
    lib().mbm_deck_defaults(ctypes.byref(st))
    d = Deck(st)
    d.nblocks = len(blocks)
    d.blocks = [dict(b) for b in blocks]
    d.kn, d.r13, d.force, d.cfl, d.limiter = kn, r13, force, cfl, limiter
    d.tout = list(tout)
    return d


# ---------------------------------------------------------------- engine
class Engine:
    """Device state of one deck (mbm_create .. mbm_destroy)."""

    def __init__(self, blocks, kn=0.001, force=0.0, limiter="harmonic", scenario="riemann2d",
                 scenario_params=TEMPJUMP_DEFAULT, stream=None):
        L = lib()
        self.blocks = [dict(b) for b in blocks]
        n = len(blocks)
        arr = (BlockDesc * n)()
        for k, b in enumerate(blocks):
            arr[k].nx, arr[k].ny = b["nx"], b["ny"]
            arr[k].xmin, arr[k].xmax, arr[k].ymin, arr[k].ymax = b["xmin"], b["xmax"], b["ymin"], b["ymax"]
            for s in range(4):
                arr[k].bc[s] = b["bc"][s]
        self._desc = arr
        self.params = Params()
        self.params.kn = kn
        self.params.force = force
        self.params.limiter = LIMITERS[limiter] if isinstance(limiter, str) else int(limiter)
        self.params.scenario = SCENARIOS[scenario] if isinstance(scenario, str) else int(scenario)
        for i, v in enumerate(scenario_params):
            self.params.scenario_params[i] = v
        h = ctypes.c_void_p()
        _check(L.mbm_create(arr, n, ctypes.byref(self.params),
                            ctypes.c_void_p(stream) if stream else None, ctypes.byref(h)),
               "mbm_create")
        self.h = h
        self.adj_runs = 0          # recordings made by adj_forward (one is kept)

    @classmethod
    def from_deck(cls, deck, scenario="riemann2d", **kw):
        args = dict(kn=deck.kn, force=deck.force, limiter=deck.limiter, scenario=scenario)
        args.update(kw)
        return cls(deck.blocks, **args)

    def close(self):
        if getattr(self, "h", None):
            lib().mbm_destroy(self.h)
            self.h = None

    def __del__(self):
        # at interpreter exit the module globals may already be gone; the
        # process is ending and the GPU driver releases the device memory
        try:
            self.close()
        except Exception:
            pass

    def shape(self, k):
        return (self.blocks[k]["ny"] + 6, self.blocks[k]["nx"] + 6, NV)

    def set_tile_rows(self, rows):
        """Rows per wavefront of the stage kernel; 0 = automatic."""
        _check(lib().mbm_set_tile_rows(self.h, int(rows)), "mbm_set_tile_rows")
        self.adj_runs += 1         # a tile-height change drops the recorded run

    def tile_rows(self):
        return lib().mbm_tile_rows(self.h)

    def set_wave_group(self, waves):
        """Waves (tiles) per workgroup of the stage kernels: 1, 2, 4 or 8."""
        _check(lib().mbm_set_wave_group(self.h, int(waves)), "mbm_set_wave_group")

    def wave_group(self):
        return lib().mbm_wave_group(self.h)

    def init(self):
        _check(lib().mbm_init(self.h), "mbm_init")

    def upload(self, arrays, layer=L_DATA):
        for k, a in enumerate(arrays):
            a = np.ascontiguousarray(a, dtype=np.float64)
            assert a.shape == self.shape(k), (a.shape, self.shape(k))
            _check(lib().mbm_upload(self.h, layer, k, _ptr(a)), "mbm_upload")

    def download(self, layer=L_DATA):
        out = []
        for k in range(len(self.blocks)):
            a = np.empty(self.shape(k), dtype=np.float64)
            _check(lib().mbm_download(self.h, layer, k, _ptr(a)), "mbm_download")
            out.append(a)
        return out

    def set_boundary(self, layer, t):
        _check(lib().mbm_set_boundary(self.h, layer, t), "mbm_set_boundary")

    def time_update(self, dt, t):
        c = ctypes.c_double()
        _check(lib().mbm_time_update(self.h, dt, t, ctypes.byref(c)), "mbm_time_update")
        return c.value

    def commit(self):
        _check(lib().mbm_commit(self.h), "mbm_commit")

    def steps(self, t, tend, cflreg, cflmax=1.0, maxsteps=0, reccap=100000, dt=None):
        """Steps controller (Main.c:189-244). dt=None starts from DT_INITIAL;
        otherwise continues with dt. Returns (t, accepted, rejected, log) and
        keeps the next dt in self.next_dt."""
        tt = ctypes.c_double(t)
        dtio = ctypes.c_double(dt if dt else 0.0)
        recs = (StepRecord * reccap)() if reccap else None
        nrec, ns, nr = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().mbm_steps(self.h, ctypes.byref(tt), tend, cflmax, cflreg, maxsteps,
                               ctypes.byref(dtio), recs, reccap, ctypes.byref(nrec),
                               ctypes.byref(ns), ctypes.byref(nr)), "mbm_steps")
        self.next_dt = dtio.value
        n = min(nrec.value, reccap)
        log = np.array([(r.cfl, r.dt, r.t, r.accepted) for r in recs[:n]]) if n else np.zeros((0, 4))
        return tt.value, ns.value, nr.value, log

    def output_values(self, k, float32=False):
        n = (self.blocks[k]["nx"] + 2) * (self.blocks[k]["ny"] + 2)
        v = np.empty((n, 10), dtype=np.float32 if float32 else np.float64)
        _check(lib().mbm_output_values(self.h, k, int(bool(float32)), _ptr(v)), "mbm_output_values")
        return v

    def interior_cells(self):
        return lib().mbm_interior_cells(self.h)

    def enable_timing(self, on=True):
        _check(lib().mbm_enable_timing(self.h, int(on)), "mbm_enable_timing")

    # ---- whole-layer access (device layout) and the discrete adjoint ----
    def layer_elems(self):
        return lib().mbm_layer_elems(self.h)

    def _layer_buffer(self, a, what, writable):
        """(pointer, keep-alive object) of a whole-layer buffer: a float64
        numpy array or torch tensor of exactly layer_elems() contiguous
        elements (a torch tensor on this engine's device)."""
        n = self.layer_elems()
        if hasattr(a, "data_ptr"):
            import torch
            if a.dtype != torch.float64 or not a.is_contiguous() or a.numel() != n:
                raise MbmError(f"{what}: need a contiguous float64 tensor of {n} elements, got "
                               f"{a.dtype} {tuple(a.shape)} contiguous={a.is_contiguous()}")
            if a.is_cuda and a.device.index != self.device():
                raise MbmError(f"{what}: tensor on cuda:{a.device.index}, engine on cuda:{self.device()}")
            return a.data_ptr(), a
        if writable:
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
                raise MbmError(f"{what}: need a C-contiguous float64 numpy array")
            arr = a
        else:
            arr = np.ascontiguousarray(a, dtype=np.float64)
        if arr.size != n:
            raise MbmError(f"{what}: buffer holds {arr.size} doubles, the layout needs {n}")
        return arr.ctypes.data, arr

    def device(self):
        return lib().mbm_device(self.h)

    def layer_get(self, layer, out=None):
        """Copies a whole layer (device layout) to `out` (a float64 numpy array
        or torch tensor of layer_elems() elements)."""
        if out is None:
            out = np.empty(self.layer_elems())
        ptr, keep = self._layer_buffer(out, "layer_get", True)
        _check(lib().mbm_layer_get(self.h, layer, ctypes.c_void_p(ptr), self.layer_elems()),
               "mbm_layer_get")
        del keep
        return out

    def layer_set(self, layer, src):
        ptr, keep = self._layer_buffer(src, "layer_set", False)
        _check(lib().mbm_layer_set(self.h, layer, ctypes.c_void_p(ptr), self.layer_elems()),
               "mbm_layer_set")
        del keep

    def set_kn(self, kn):
        self.params.kn = float(kn)
        _check(lib().mbm_set_params(self.h, ctypes.byref(self.params)), "mbm_set_params")

    def set_stream(self, stream_handle):
        _check(lib().mbm_set_stream(self.h, ctypes.c_void_p(stream_handle) if stream_handle else None),
               "mbm_set_stream")

    def adj_forward(self, tend, cflreg, cflmax=1.0, maxsteps=0, dts=None, stride=None, t0=0.0):
        """Advances DATA recording (dts=None) or replaying a dt schedule,
        with checkpoints every `stride` steps. A replay with stride None (or
        0) keeps the states of every step in device memory when the budget
        (adj_set_budget) holds them, else checkpoints every ceil(sqrt(N))
        steps. Returns (nsteps, t_end)."""
        n, te = ctypes.c_int(), ctypes.c_double()
        self.adj_runs += 1
        if dts is not None:
            dts = np.ascontiguousarray(dts, dtype=np.float64)
            _check(lib().mbm_adj_forward(self.h, t0, tend, cflmax, cflreg, 0, _ptr(dts), len(dts),
                                         int(stride or 0), ctypes.byref(n), ctypes.byref(te)),
                   "mbm_adj_forward")
        else:
            stride = stride or max(1, int(np.ceil(np.sqrt(maxsteps if maxsteps > 0 else 1000))))
            _check(lib().mbm_adj_forward(self.h, t0, tend, cflmax, cflreg, maxsteps, None, 0, stride,
                                         ctypes.byref(n), ctypes.byref(te)), "mbm_adj_forward")
        return n.value, te.value

    def adj_schedule(self):
        n = ctypes.c_int()
        _check(lib().mbm_adj_schedule(self.h, None, 0, ctypes.byref(n)), "mbm_adj_schedule")
        out = np.empty(n.value)
        _check(lib().mbm_adj_schedule(self.h, _ptr(out), n.value, ctypes.byref(n)), "mbm_adj_schedule")
        return out

    def adj_backward(self, g_final, g_initial=None):
        """dJ/d(initial DATA) and dJ/dKn from dJ/d(final DATA) (layer layout;
        float64 numpy arrays or torch tensors of layer_elems() elements)."""
        if g_initial is None:
            g_initial = np.empty(self.layer_elems())
        gp, keep_g = self._layer_buffer(g_final, "adj_backward g_final", False)
        ip, keep_i = self._layer_buffer(g_initial, "adj_backward g_initial", True)
        gk = ctypes.c_double()
        _check(lib().mbm_adj_backward(self.h, ctypes.c_void_p(gp), ctypes.c_void_p(ip),
                                      self.layer_elems(), ctypes.byref(gk)), "mbm_adj_backward")
        del keep_g, keep_i
        return g_initial, gk.value

    def adj_set_budget(self, nbytes):
        """Device memory a resident recording may hold (0: the free memory
        less a reserve)."""
        _check(lib().mbm_adj_set_budget(self.h, int(nbytes)), "mbm_adj_set_budget")

    def adj_memory(self):
        """Device bytes held by the recorded states."""
        b = ctypes.c_longlong()
        _check(lib().mbm_adj_memory(self.h, ctypes.byref(b)), "mbm_adj_memory")
        return b.value

    def adj_release(self):
        _check(lib().mbm_adj_release(self.h), "mbm_adj_release")
        self.adj_runs += 1

    def layer_index(self):
        """Per block (offset, plane, pitch, rows) of the device layout: cell
        (a, b) of component m of block k is element
        offset + m*plane + rows[b]*pitch + a of a whole layer. The row order
        follows the tile height (set_tile_rows moves the layers)."""
        out = []
        base = None
        for k, b in enumerate(self.blocks):
            dp, pl, pi = ctypes.c_void_p(), ctypes.c_longlong(), ctypes.c_int()
            _check(lib().mbm_layer_ptr(self.h, 0, k, ctypes.byref(dp), ctypes.byref(pl), ctypes.byref(pi)),
                   "mbm_layer_ptr")
            if base is None:
                base = dp.value
            rows = np.empty(b["ny"] + 6, dtype=np.int32)
            _check(lib().mbm_layer_rows(self.h, k, _ptr(rows), len(rows)), "mbm_layer_rows")
            out.append(((dp.value - base) // 8, pl.value, pi.value, rows.astype(np.int64)))
        return out

    def layer_to_blocks(self, U):
        """Whole-layer array (device layout, numpy) -> per block (ny+6, nx+6, 9) arrays."""
        return layer_to_blocks(U, self.layer_index(), self.blocks)

    def blocks_to_layer(self, blocks, out=None):
        """Inverse of layer_to_blocks: writes (ny+6, nx+6, 9) arrays into a
        whole-layer array (a new one, zero between the cells, if out is None)."""
        if out is None:
            out = np.zeros(self.layer_elems())
        return blocks_to_layer(blocks, self.layer_index(), self.blocks, out)

    def kernel_time(self):
        sm, fm = ctypes.c_double(), ctypes.c_double()
        ns, nf = ctypes.c_int(), ctypes.c_int()
        _check(lib().mbm_kernel_time(self.h, ctypes.byref(sm), ctypes.byref(ns), ctypes.byref(fm),
                                     ctypes.byref(nf)), "mbm_kernel_time")
        return sm.value, ns.value, fm.value, nf.value


def _desc_array(blocks):
    arr = (BlockDesc * len(blocks))()
    for k, b in enumerate(blocks):
        arr[k].nx, arr[k].ny = b["nx"], b["ny"]
        arr[k].xmin, arr[k].xmax, arr[k].ymin, arr[k].ymax = b["xmin"], b["xmax"], b["ymin"], b["ymax"]
        for s in range(4):
            arr[k].bc[s] = b["bc"][s]
    return arr


class Batch:
    """Independent decks advanced together (mbm_batch_*): per-deck Kn and
    Steps controller, shared limiter and scenario."""

    def __init__(self, decks, limiter="harmonic", scenario="riemann2d",
                 scenario_params=TEMPJUMP_DEFAULT, stream=None):
        """decks: list of dicts with keys blocks (deck-local sync codes), kn,
        and optionally force, cfl, tend."""
        self.decks = decks
        blocks = [b for d in decks for b in d["blocks"]]
        self._desc = _desc_array(blocks)
        self._nb = np.array([len(d["blocks"]) for d in decks], dtype=np.int32)
        self._par = (Params * len(decks))()
        for i, d in enumerate(decks):
            self._par[i].kn = d["kn"]
            self._par[i].force = d.get("force", 0.0)
            self._par[i].limiter = LIMITERS[limiter]
            self._par[i].scenario = SCENARIOS[scenario]
            for j, v in enumerate(scenario_params):
                self._par[i].scenario_params[j] = v
        h = ctypes.c_void_p()
        _check(lib().mbm_batch_create(self._desc, _ptr(self._nb), len(decks), self._par,
                                      ctypes.c_void_p(stream) if stream else None, ctypes.byref(h)),
               "mbm_batch_create")
        self.h = h
        self.t = np.zeros(len(decks))

    def close(self):
        if getattr(self, "h", None):
            lib().mbm_destroy(self.h)
            self.h = None

    def __del__(self):
        # at interpreter exit the module globals may already be gone; the
        # process is ending and the GPU driver releases the device memory
        try:
            self.close()
        except Exception:
            pass

    def init(self):
        _check(lib().mbm_init(self.h), "mbm_init")
        self.t[:] = 0.0

    def cells(self):
        return sum(b["nx"] * b["ny"] for d in self.decks for b in d["blocks"])

    def autotune(self, rows=(8, 16, 24, 32, 48, 64), reps=3):
        r = np.array(rows, dtype=np.int32)
        best, ms = ctypes.c_int(), ctypes.c_double()
        _check(lib().mbm_batch_autotune(self.h, _ptr(r), len(r), reps, ctypes.byref(best),
                                        ctypes.byref(ms)), "mbm_batch_autotune")
        return best.value, ms.value

    def autotune_shape(self, rows=(16, 24, 32, 40, 48, 64, 80, 96), waves=(1, 2, 4, 8), reps=3):
        """Tile height and wave group of the fastest stage pair; returns
        (rows, waves, ms per stage pair)."""
        r = np.array(rows, dtype=np.int32)
        w = np.array(waves, dtype=np.int32)
        br, bw, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(lib().mbm_batch_autotune_shape(self.h, _ptr(r), len(r), _ptr(w), len(w), reps,
                                              ctypes.byref(br), ctypes.byref(bw), ctypes.byref(ms)),
               "mbm_batch_autotune_shape")
        return br.value, bw.value, ms.value

    def set_wave_group(self, waves):
        _check(lib().mbm_set_wave_group(self.h, int(waves)), "mbm_set_wave_group")

    def wave_group(self):
        return lib().mbm_wave_group(self.h)

    def run(self, maxsteps, tend=None, cflmax=1.0):
        n = len(self.decks)
        te = np.array([d.get("tend", 1e30) if tend is None else tend for d in self.decks], dtype=np.float64)
        cr = np.array([d.get("cfl", 0.95) for d in self.decks], dtype=np.float64)
        ns = np.zeros(n, dtype=np.int32)
        nr = np.zeros(n, dtype=np.int32)
        _check(lib().mbm_batch_run(self.h, _ptr(te), _ptr(cr), cflmax, maxsteps, _ptr(self.t),
                                   _ptr(ns), _ptr(nr)), "mbm_batch_run")
        return ns, nr

    def download(self, deck):
        out = []
        for k, b in enumerate(self.decks[deck]["blocks"]):
            a = np.empty((b["ny"] + 6, b["nx"] + 6, NV))
            _check(lib().mbm_batch_download(self.h, deck, k, _ptr(a)), "mbm_batch_download")
            out.append(a)
        return out


def square4_topology(a, b, c, d, x0=-1.0, xm=0.5, x1=1.0, y0=-1.0, ym=0.2, y1=1.0):
    """cases/square4.txt geometry with nx(B1)=nx(B4)=a, nx(B2)=nx(B3)=b,
    ny(B1)=ny(B2)=c, ny(B3)=ny(B4)=d (the sizes sync conformity allows)."""
    return [dict(nx=a, ny=c, xmin=x0, xmax=xm, ymin=y0, ymax=ym, bc=[0, 2, -1, 4]),
            dict(nx=b, ny=c, xmin=xm, xmax=x1, ymin=y0, ymax=ym, bc=[1, 0, -1, 3]),
            dict(nx=b, ny=d, xmin=xm, xmax=x1, ymin=ym, ymax=y1, bc=[4, 0, 2, -1]),
            dict(nx=a, ny=d, xmin=x0, xmax=xm, ymin=ym, ymax=y1, bc=[0, 3, 1, -1])]


def sweep_decks(n=256, seed=0, sides=tuple(range(64, 513, 32)), kn_range=(1e-3, 1e-1), cfl=0.95):
    """Seeded sweep of square4-topology decks: block sides drawn from
    {64, 96, ..., 512}, Kn log-uniform, R13 in {0, 1} (output naming only)."""
    rng = np.random.default_rng(seed)
    decks = []
    for _ in range(n):
        a, b, c, d = (int(rng.choice(sides)) for _ in range(4))
        kn = float(10 ** rng.uniform(np.log10(kn_range[0]), np.log10(kn_range[1])))
        decks.append(dict(blocks=square4_topology(a, b, c, d), kn=kn, r13=int(rng.integers(0, 2)),
                          cfl=cfl))
    return decks


def host_set_boundary(blocks, arrays, kn=0.001, limiter="harmonic", scenario="riemann2d",
                      scenario_params=TEMPJUMP_DEFAULT, t=0.0):
    """CPU application of the engine's ghost map (no GPU needed)."""
    L = lib()
    n = len(blocks)
    arr = (BlockDesc * n)()
    for k, b in enumerate(blocks):
        arr[k].nx, arr[k].ny = b["nx"], b["ny"]
        arr[k].xmin, arr[k].xmax, arr[k].ymin, arr[k].ymax = b["xmin"], b["xmax"], b["ymin"], b["ymax"]
        for s in range(4):
            arr[k].bc[s] = b["bc"][s]
    par = Params()
    par.kn = kn
    par.limiter = LIMITERS[limiter]
    par.scenario = SCENARIOS[scenario]
    for i, v in enumerate(scenario_params):
        par.scenario_params[i] = v
    ptrs = (ctypes.c_void_p * n)(*[a.ctypes.data for a in arrays])
    for a in arrays:
        assert a.dtype == np.float64 and a.flags.c_contiguous
    _check(L.mbm_host_set_boundary(arr, n, ctypes.byref(par), ptrs, t), "mbm_host_set_boundary")


def initial_state(scenario, x, y, t=0.0, params=TEMPJUMP_DEFAULT):
    U = np.empty(NV)
    p = np.asarray(params, dtype=np.float64)
    lib().mbm_initial_state(SCENARIOS[scenario], _ptr(p), x, y, t, _ptr(U))
    return U


def convert(a):
    buf = ctypes.create_string_buffer(32)
    return lib().mbm_convert(a, buf).decode()
