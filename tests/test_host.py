"""Host-side pieces of the drop-in (no GPU): the C ABI exports, the block
file reader, the ghost map that the fill kernel applies, the output writer,
the restart reader and the comparator."""
import ctypes
import lzma
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, deck_path, fixture_names, scenario_of

import oracle


def declared_functions():
    names = []
    for h in ("mbmfv.h", "mbmfv_host.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(mbm_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(mbm):
    L = mbm.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert len(declared_functions()) >= 30


def test_fp64_rate_probe_is_built():
    """bench.py --valu-rate loads the FP64 issue-rate probe; it is built with
    the library and exports its one entry point (no GPU call here)."""
    path = os.path.join(ROOT, "multiblockmomentsfv_amd", "libmbmprobe.so")
    assert os.path.exists(path), "run make"
    import torch  # noqa: F401  (the probe links the same HIP runtime as torch)
    L = ctypes.CDLL(path)
    assert hasattr(L, "mbm_probe_fp64_rate")


@pytest.mark.parametrize("ny,band", [(1024, 37), (64, 20), (5, 2), (100, 1), (7, 4096), (2048, 45)])
def test_row_order_keeps_marching_rows_adjacent(mbm, ny, band):
    """The device row order is a bijection of the ny+6 rows onto positions,
    row 0 last, and the rows that tiles of `band` rows (starting at row 1)
    visit at the same step are consecutive positions, tile by tile."""
    pos, npos = mbm.row_order(ny, band)
    ly = ny + 6
    b = max(band, 2)
    nb = -(-(ly - 1) // b)
    assert len(pos) == ly and len(set(pos.tolist())) == ly
    assert pos.min() >= 0 and pos.max() < npos == b * nb + 1 and pos[0] == b * nb
    for t in range(b):
        rows = [1 + k * b + t for k in range(nb) if 1 + k * b + t < ly]
        assert [int(pos[r]) for r in rows] == list(range(t * nb, t * nb + len(rows)))


def test_layer_conversion_round_trip(mbm):
    """Block arrays -> whole layer -> block arrays is the identity, every cell
    lands at off + m*plane + rows[b]*pitch + a, and nothing else is written."""
    blocks = [dict(nx=7, ny=5), dict(nx=12, ny=9), dict(nx=3, ny=4)]
    index, off = [], 0
    for b in blocks:
        plane = (b["nx"] + 6 + 7) // 8 * 8
        rows, npos = mbm.row_order(b["ny"], 4)
        index.append((off, plane, 9 * plane, rows))
        off += npos * 9 * plane + 32
    rng = np.random.default_rng(0)
    arrays = [rng.normal(size=(b["ny"] + 6, b["nx"] + 6, 9)) for b in blocks]
    U = mbm.blocks_to_layer(arrays, index, blocks, np.zeros(off))
    for (o, plane, pitch, rows), b, a in zip(index, blocks, arrays):
        for bb, aa, m in ((0, 0, 0), (b["ny"] + 5, b["nx"] + 5, 8), (3, 2, 4)):
            assert U[o + m * plane + rows[bb] * pitch + aa] == a[bb, aa, m]
    assert np.count_nonzero(U) == sum(a.size for a in arrays)
    for a, c in zip(arrays, mbm.layer_to_blocks(U, index, blocks)):
        assert a.tobytes() == c.tobytes()


@pytest.mark.parametrize("name", ["squareSmall", "square4"])
def test_deck_geometry_matches_golden_zones(mbm, name):
    deck = mbm.read_deck(deck_path(name))
    zones = [dict(re.findall(r"(\w+)=(\S+)", l)) for l in open(os.path.join(GOLDEN, name + ".ref.hdr"))
             if l.startswith("ZONE")]
    for k in range(deck.nblocks):
        bx, by, dx, dy = deck.geom(k)
        z = zones[k]
        assert ("%.16g" % bx, "%.16g" % by, "%.16g" % dx, "%.16g" % dy) == (z["BX"], z["BY"], z["DX"], z["DY"])
        assert int(z["I"]) == deck.blocks[k]["ny"] + 2 and int(z["J"]) == deck.blocks[k]["nx"] + 2
    tout = [l for l in open(os.path.join(GOLDEN, name + ".ref.hdr")) if l.startswith("t_output")][0]
    assert tout.split()[2:] == ["0.0"] + ["%.10g" % t for t in deck.tout]


def test_convert_matches_golden_names(mbm):
    assert mbm.convert(0.001) == "0p001"
    assert mbm.convert(0.01) == "0p01"
    assert mbm.convert(0.05) == "0p05"
    assert mbm.convert(0.3) == "0p3"


def test_deck_reader_edge_cases(mbm, tmp_path):
    p = tmp_path / "d.txt"
    p.write_text("kn = 0.02  # inline comment\n// comment\nLIMITER = minmod\nrefine = 2\nfoo = 1\n"
                 "dt_out = 0.0\nt_list = 0.1 0.2 0.3\n\nBlock data\nheader line\n"
                 "1/0,1/0,1/4,5/0,0,-1,-1/\n")
    d = mbm.read_deck(str(p))
    assert d.kn == 0.02 and d.limiter == "minmod" and d.refine == 2
    assert d.tout == [0.1, 0.2, 0.3]
    assert d.blocks[0]["nx"] == 8 and d.blocks[0]["ny"] == 10
    d2 = mbm.read_deck(str(p), refine=3)
    assert d2.blocks[0]["nx"] == 12
    q = tmp_path / "bad.txt"
    q.write_text("Block data\nhdr\n40/0,1/0,1/4,4/0,0,0,0/\n")
    with pytest.raises(mbm.MbmError):
        mbm.read_deck(str(q))


@pytest.mark.parametrize("name", fixture_names())
def test_ghost_map_matches_oracle_set_boundary(mbm, name):
    deck = mbm.read_deck(deck_path(name))
    rng = np.random.default_rng(3)
    arrays = [rng.uniform(0.5, 2.0, (b["ny"] + 6, b["nx"] + 6, 9)) for b in deck.blocks]
    host = [a.copy() for a in arrays]
    mbm.host_set_boundary(deck.blocks, host, kn=deck.kn, limiter=deck.limiter,
                          scenario=scenario_of(name))
    orc = oracle.Oracle(deck.blocks, kn=deck.kn, limiter=deck.limiter, scenario=scenario_of(name))
    orc.set([a.copy() for a in arrays], layer=0)
    orc.set_boundary(0, 0.0)
    for k, (h, r) in enumerate(zip(host, orc.get(0))):
        assert h.tobytes() == r.tobytes(), f"block {k + 1}"


def test_nonconforming_sync_is_rejected(mbm):
    blocks = [dict(nx=4, ny=4, xmin=0, xmax=1, ymin=0, ymax=1, bc=[0, 2, 0, 0]),
              dict(nx=4, ny=5, xmin=1, xmax=2, ymin=0, ymax=1, bc=[1, 0, 0, 0])]
    arrays = [np.zeros((b["ny"] + 6, b["nx"] + 6, 9)) for b in blocks]
    with pytest.raises(mbm.MbmError, match="non-conforming"):
        mbm.host_set_boundary(blocks, arrays)


def write_with_product_writer(mbm, deck, vals_per_block, outdir, t, add, float32=True):
    st = deck.struct()
    ptrs = (ctypes.c_void_p * deck.nblocks)(*[v.ctypes.data for v in vals_per_block])
    rc = mbm.lib().mbm_output_write(outdir.encode(), b"Riemann2D", ctypes.byref(st), int(float32),
                                    t, add, ptrs)
    assert rc == 0


def test_output_writer_reproduces_golden_files(mbm, tmp_path):
    """Oracle values through the product writer == squareSmall.ref.{hdr,dat}."""
    deck = mbm.read_deck(deck_path("squareSmall"))
    orc = oracle.Oracle(deck.blocks, kn=deck.kn, limiter=deck.limiter)
    orc.init()
    out = str(tmp_path)
    vals = [np.ascontiguousarray(orc.output_values(k)) for k in range(deck.nblocks)]
    write_with_product_writer(mbm, deck, vals, out, 0.0, -1)
    total = 0
    for tend in deck.tout:
        n, _ = orc.steps(tend, deck.cfl)
        total += n
        vals = [np.ascontiguousarray(orc.output_values(k)) for k in range(deck.nblocks)]
        write_with_product_writer(mbm, deck, vals, out, orc.t, total)
    base = os.path.join(out, "Grad13Riemann2D_nst400_Kn0p001_tAll")
    assert open(base + ".dat", "rb").read() == open(os.path.join(GOLDEN, "squareSmall.ref.dat"), "rb").read()
    mine = [l for l in open(base + ".hdr") if not l.startswith(("# CREATED_UTC", "# GIT_COMMIT"))]
    gold = [l for l in open(os.path.join(GOLDEN, "squareSmall.ref.hdr"))
            if not l.startswith(("# CREATED_UTC", "# GIT_COMMIT"))]
    assert mine == gold


def test_restart_reader_is_bug_compatible(mbm):
    """LoadData (Mesh.c:263-273) puts p and P_ij into the theta and T_ij slots."""
    deck = mbm.read_deck(deck_path("squareSmall"))
    st = deck.struct()
    bufs = [np.zeros((b["ny"] + 6, b["nx"] + 6, 9)) for b in deck.blocks]
    ptrs = (ctypes.c_void_p * deck.nblocks)(*[b.ctypes.data for b in bufs])
    typ = ctypes.create_string_buffer(256)
    rc = mbm.lib().mbm_restart_read(os.path.join(GOLDEN, "squareSmall.ref").encode(), ctypes.byref(st), ptrs, typ)
    assert rc == 0 and typ.value == b"Riemann2D"
    raw = np.frombuffer(open(os.path.join(GOLDEN, "squareSmall.ref.dat"), "rb").read(), dtype="<f4")
    off = 0
    for k, b in enumerate(deck.blocks):
        n = (b["nx"] + 2) * (b["ny"] + 2)
        v = raw[off:off + 10 * n].astype(np.float64).reshape(b["nx"] + 2, b["ny"] + 2, 10)
        off += 10 * n
        for a in range(b["nx"] + 2):
            for j in range(b["ny"] + 2):
                x = v[a, j]
                P = [x[0], x[1], x[2], x[4], x[5], x[6], x[7], x[8], x[9]]
                U = np.empty(9)
                oracle.lib().orc_conservative(np.array(P).ctypes.data, U.ctypes.data)
                assert bufs[k][j + 2, a + 2].tobytes() == U.tobytes()


def test_comparator_semantics(tmp_path):
    from multiblockmomentsfv_amd.compare_solution import compare, main
    g = os.path.join(GOLDEN, "squareSmall.ref")
    assert compare(g, g, out=lambda s: None) == 0
    base = str(tmp_path / "p")
    open(base + ".hdr", "w").write(open(g + ".hdr").read())
    d = bytearray(open(g + ".dat", "rb").read())
    v = np.frombuffer(bytes(d), dtype="<f4").copy()
    v[1234] *= 1.0 + 3e-5
    open(base + ".dat", "wb").write(v.tobytes())
    assert compare(base, g, out=lambda s: None) == 1
    assert compare(base, g, 1e-3, 1e-3, out=lambda s: None) == 0
    assert main([g]) == 2
    assert main([g + ".hdr", g + ".dat"]) == 0          # any of hdr / dat / stem
    # the gate's rule is "fail when |diff| > tol": a NaN on either side passes
    v2 = np.frombuffer(bytes(d), dtype="<f4").copy()
    v2[77] = np.nan
    open(base + ".dat", "wb").write(v2.tobytes())
    assert compare(base, g, out=lambda s: None) == 0
    # the tolerance rule is relative to the reference value: 9e-6 relative
    # passes at the Float32 defaults (1e-6 + 1e-5 |ref|), 2e-5 does not
    v3 = np.frombuffer(bytes(d), dtype="<f4").copy().astype(np.float64)
    i = int(np.argmax(np.abs(v3)))
    for rel, status in ((9e-6, 0), (2e-5, 1)):
        v4 = v3.copy()
        v4[i] *= 1 + rel
        open(base + ".dat", "wb").write(v4.astype("<f4").tobytes())
        assert compare(base, g, out=lambda s: None) == status, rel
    # zone structure: a changed NVARS fails regardless of the values
    hdr = open(g + ".hdr").read().replace("NVARS=10", "NVARS=9", 1)
    open(base + ".hdr", "w").write(hdr)
    open(base + ".dat", "wb").write(bytes(d))
    assert compare(base, g, out=lambda s: None) == 1
    # no PRECISION entry means Float64: 8-byte values and the tighter defaults
    h64 = "\n".join(l for l in open(g + ".hdr").read().splitlines() if "PRECISION" not in l)
    open(base + ".hdr", "w").write(h64.replace("NBYTES=", "X="))
    v5 = np.frombuffer(bytes(d), dtype="<f4").astype("<f8")
    with open(base + ".dat", "wb") as f:
        f.write(v5.tobytes())
    hd = []
    off = 0
    for line in open(base + ".hdr").read().splitlines():
        if line.startswith("ZONE"):
            import re as _re
            n = [int(_re.search(rf"\b{k}=(\d+)", line).group(1)) for k in ("I", "J", "NVARS")]
            line = _re.sub(r"OFFSET=\d+", f"OFFSET={off}", line)
            off += 8 * n[0] * n[1] * n[2]
        hd.append(line)
    open(base + ".hdr", "w").write("\n".join(hd) + "\n")
    assert compare(base, base, out=lambda s: None) == 0
    msgs = []
    assert compare(base, g, out=msgs.append) == 0 and "new Float64, ref Float32" in msgs[0]
