"""CPU reference of the adjoint of one Heun step, built from the engine's own
per-cell adjoint code (tests/native/libphyscheck.so, the same functions the
GPU kernels run) and the transpose of the ghost fill, which is read off the
engine's host ghost map by probing it with tagged values. Test helper."""
import ctypes
import os

import numpy as np

import multiblockmomentsfv_amd as mbm

LIBP = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "libphyscheck.so")
_L = None


def lib():
    global _L
    if _L is None:
        L = ctypes.CDLL(LIBP)
        L.pc_stage_vjp.restype = ctypes.c_double
        L.pc_stage_vjp.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 2 + \
                                  [ctypes.c_void_p] * 5 + [ctypes.c_double] * 3
        _L = L
    return _L


def fill_transpose_map(blocks):
    """For every ghost cell: (block, j, i) -> (source block, j, i, sign vector),
    sources None for Dirichlet ghosts."""
    tags = []
    for k, b in enumerate(blocks):
        a = np.zeros((b["ny"] + 6, b["nx"] + 6, 9))
        jj, ii = np.meshgrid(np.arange(b["ny"] + 6), np.arange(b["nx"] + 6), indexing="ij")
        code = (k + 1) * 1e6 + jj * 1e3 + ii + 0.25
        a[:] = code[..., None]
        a[:3] = a[-3:] = 0.0
        a[:, :3] = a[:, -3:] = 0.0
        tags.append(a)
    filled = [t.copy() for t in tags]
    mbm.host_set_boundary(blocks, filled)
    out = []
    for k, (b, f) in enumerate(zip(blocks, filled)):
        ly, lx = b["ny"] + 6, b["nx"] + 6
        for j in range(ly):
            for i in range(lx):
                if 3 <= j < ly - 3 and 3 <= i < lx - 3:
                    continue
                v = f[j, i]
                mag = np.abs(v)
                if not np.allclose(mag, mag[0]) or abs(mag[0] - round(mag[0] - 0.25) - 0.25) > 1e-6:
                    out.append((k, j, i, None, None, None, None))
                    continue
                c = int(round(mag[0] - 0.25))
                sk, sj, si = c // 1000000 - 1, (c // 1000) % 1000, c % 1000
                out.append((k, j, i, sk, sj, si, np.sign(v)))
    return out


def fill_T(g, fmap):
    """In place: fold ghost adjoints into their sources, clear ghosts."""
    add = [np.zeros_like(a) for a in g]
    for k, j, i, sk, sj, si, sg in fmap:
        if sk is not None:
            add[sk][sj, si] += sg * g[k][j, i]
    for k, j, i, *_ in fmap:
        g[k][j, i] = 0.0
    for a, b in zip(g, add):
        a += b
    return g


def stage_T(stage, lim, blocks, geoms, ins, ins0, gouts, dt, kn, force):
    """S^T per block (stages couple blocks only through the ghost fill)."""
    soa = lambda A: np.ascontiguousarray(A.transpose(2, 0, 1))
    gin, gin0, gkn = [], [], 0.0
    for b, (bx, by, dx, dy), U, U0, g in zip(blocks, geoms, ins, ins0, gouts):
        a = np.zeros((9, b["ny"] + 6, b["nx"] + 6))
        a0 = np.zeros_like(a)
        su, su0, sg = soa(U), soa(U0), soa(g)
        gkn += lib().pc_stage_vjp(stage, lim, b["nx"], b["ny"], dx, dy, su.ctypes.data,
                                  su0.ctypes.data, sg.ctypes.data, a.ctypes.data,
                                  a0.ctypes.data, dt, kn, force)
        gin.append(np.ascontiguousarray(a.transpose(1, 2, 0)))
        gin0.append(np.ascontiguousarray(a0.transpose(1, 2, 0)))
    return gin, gin0, gkn


def step_adjoint(orc, deck_blocks, geoms, lim, D, gN, dt, kn, force, fmap):
    """Adjoint of one TimeUpdate at DATA = D (ghosts filled): gN -> gD."""
    orc.set([a.copy() for a in D], 0)
    orc.time_update(dt, 0.0)
    T = orc.get(1)
    gN = fill_T([a.copy() for a in gN], fmap)
    gT, gD, k2 = stage_T(2, lim, deck_blocks, geoms, T, D, gN, dt, kn, force)
    gT = fill_T(gT, fmap)
    gD1, _, k1 = stage_T(1, lim, deck_blocks, geoms, D, D, gT, dt, kn, force)
    return [a + b for a, b in zip(gD, gD1)], k1 + k2
