"""The reference's on-disk formats (csrc/dpd_io.cpp) against byte-level
readers and writers written here from the reference's Fortran (CPU only):

* pyscf2cc/mod_interface.f90 node file: stream file [metaPos, nNode] (int64),
  data, then nNode records of (name char(20), pos int64, size int64); positions
  1-based; names lower case; a rewrite that fits keeps its place, a larger one
  moves to the end (get_node).
* moio lists (clmoioso0.f / updmoiso.f / getallso_d.f / getlstso_d.f): the data
  file holds irrep blocks (right-pair irrep ascending) of numl*numr doubles,
  each starting on a record of iprcwd = 4096 doubles at record rectmp(irrep)+1;
  the ".c" control file is Fortran sequential unformatted (4-byte markers) with
  IRREPX, popp..pops, ioppq, ioprs, irrepr, irrepi (8 int64 each), then
  moionuml, moionumr, rectmp (10 x 20 int64, Fortran order).
* IO_stream (IO.f90:87-154): element idx at byte 8*(idx-1); the Davidson
  eigenvector file puts root r of state istate at element (istate+r-2)*(1+dim)+1.
"""
import ctypes
import struct
import types

import numpy as np
import pytest

import oracle as O
from dpdhip import _lib
from dpdhip import io as dio


@pytest.fixture(scope="module")
def hctx():
    """a host-only context (no GPU needed), C2v"""
    h = ctypes.c_void_p()
    d = np.ascontiguousarray(O.xor_dirprd(4).reshape(-1))
    _lib.check(_lib.lib().dpd_ctx_create(ctypes.byref(h), 4, d.ctypes.data_as(_lib.c_i64_p), -1))
    yield types.SimpleNamespace(h=h)
    _lib.lib().dpd_ctx_destroy(h)


# ---- node file -------------------------------------------------------------------
def _parse_nodes(path):
    b = open(path, "rb").read()
    meta, n = struct.unpack_from("<qq", b, 0)
    nodes = {}
    for k in range(n):
        off = meta - 1 + 36 * k
        name = b[off: off + 20].decode()
        pos, size = struct.unpack_from("<qq", b, off + 20)
        nodes[name.rstrip()] = (pos, size, b[pos - 1: pos - 1 + size])
    return meta, nodes, b


def _write_nodes(path, items):
    """a node file as mod_interface.f90 lays it out (data in insertion order)"""
    data, meta = b"", []
    pos = 17
    for name, raw in items:
        meta.append((name.lower().ljust(20).encode(), pos, len(raw)))
        data += raw
        pos += len(raw)
    out = struct.pack("<qq", pos, len(items)) + data
    for nm, p, s in meta:
        out += nm + struct.pack("<qq", p, s)
    open(path, "wb").write(out)


def test_node_file_written_as_the_reference_lays_it_out(tmp_path):
    p = str(tmp_path / "interface")
    dirprd = (O.xor_dirprd(4) + 0).astype(np.int64)
    pop = np.array([[3, 0, 1, 2, 0, 0, 0, 0], [2, 0, 1, 1, 0, 0, 0, 0]], dtype=np.int64).T  # (8, 2)
    e = np.linspace(-1.2, 2.5, 24)
    dio.node_init(p)
    dio.node_write(p, "DIRPRD", dirprd)
    dio.node_write(p, "pop", pop)
    dio.node_write(p, "orb_energy_alpha_sym", e)
    dio.node_write(p, "nirrep", np.array([4]))
    dio.node_write(p, "group_n", "C2v")
    meta, nodes, raw = _parse_nodes(p)
    assert list(nodes) == ["dirprd", "pop", "orb_energy_alpha_sym", "nirrep", "group_n"]
    assert nodes["dirprd"][0] == 17 and nodes["dirprd"][1] == 64 * 8
    assert np.array_equal(np.frombuffer(nodes["pop"][2], "<i8").reshape(8, 2, order="F"), pop)
    assert np.array_equal(np.frombuffer(nodes["orb_energy_alpha_sym"][2], "<f8"), e)
    assert nodes["group_n"][2] == b"C2v"
    assert meta == 17 + 64 * 8 + 16 * 8 + 24 * 8 + 8 + 3
    assert len(raw) == meta - 1 + 36 * 5
    # a rewrite that fits stays in place, a larger one moves to the end
    dio.node_write(p, "pop", pop[:4])
    m2, n2, _ = _parse_nodes(p)
    assert n2["pop"][:2] == (nodes["pop"][0], 4 * 2 * 8) and m2 == meta
    dio.node_write(p, "nirrep", np.array([4, 5, 6]))
    m3, n3, _ = _parse_nodes(p)
    assert n3["nirrep"][0] == meta and m3 == meta + 24
    assert np.array_equal(dio.node_read(p, "nirrep", "i"), [4, 5, 6])
    assert dio.node_size(p, "missing") == -1


def test_node_file_written_elsewhere_reads_back(tmp_path):
    p = str(tmp_path / "interface")
    vrt = np.arange(16, dtype=np.int64).reshape(8, 2, order="F")
    _write_nodes(p, [("vrt", vrt.ravel(order="F").astype("<i8").tobytes()),
                     ("escf", struct.pack("<d", -75.25)),
                     ("irrep_name0", b"A1")])
    assert np.array_equal(dio.node_read(p, "VRT", "i", shape=(8, 2)), vrt)
    assert dio.node_read(p, "escf", "d")[0] == -75.25
    assert dio.node_read(p, "irrep_name0", "c") == "A1"


# ---- moio lists ------------------------------------------------------------------------
def _record(b, off):
    n = struct.unpack_from("<i", b, off)[0]
    data = b[off + 4: off + 4 + n]
    assert struct.unpack_from("<i", b, off + 4 + n)[0] == n
    return data, off + 8 + n


def _getallso(path):
    """getallso_d.f / getlstso_d.f on the raw files"""
    c = open(path + ".c", "rb").read()
    recs, off = [], 0
    for _ in range(12):
        r, off = _record(c, off)
        recs.append(r)
    irrepx = struct.unpack("<q", recs[0])[0]
    ioppq = np.frombuffer(recs[5], "<i8")
    numl = np.frombuffer(recs[9], "<i8").reshape(10, 20, order="F")
    numr = np.frombuffer(recs[10], "<i8").reshape(10, 20, order="F")
    rect = np.frombuffer(recs[11], "<i8").reshape(10, 20, order="F")
    d = np.fromfile(path, dtype="<f8")
    out = []
    for r in range(10):
        n = numl[r, 0] * numr[r, 0]
        if n:
            s = rect[r, 0] * 4096
            out.append(d[s: s + n])
    return irrepx, ioppq, numl[:, 0], rect[:, 0], (np.concatenate(out) if out else np.zeros(0))


@pytest.mark.parametrize("which,iops,sym", [("abcd", (0, 0), 1), ("iabj", (1, 1), 2), ("ijka", (0, 1), 3)])
def test_moio_list_layout(tmp_path, hctx, which, iops, sym):
    s = O.Sym(4)
    s.activate()
    V = np.array([30, 8, 14, 18, 0, 0, 0, 0])
    Oc = np.array([5, 2, 3, 3, 0, 0, 0, 0])
    pops = {"abcd": [V, V, V, V], "iabj": [V, Oc, V, Oc], "ijka": [Oc, Oc, Oc, V]}[which]
    d = O.Desc(pops, iops, sym)
    x = np.random.default_rng(0).standard_normal(d.length())
    p = str(tmp_path / which[:8])
    dio.moio_write(hctx, p, sym, pops[0], pops[1], iops[0], pops[2], pops[3], iops[1], x)
    irrepx, ioppq, numl, rect, got = _getallso(p)
    assert irrepx == sym and ioppq[0] == iops[0]
    np.testing.assert_array_equal(got, x)
    # blocks start on whole records, in irrep order: numl x numr of each irrep
    # (inipckso.f:36-46, irpdso of the left and right pairs)
    nrec = 0
    for r in range(1, 5):
        l = s.dp(r, sym)
        nl = O.irpdso(l, pops[0], pops[1], iops[0])
        nr = O.irpdso(r, pops[2], pops[3], iops[1])
        assert numl[r - 1] == nl and rect[r - 1] == nrec
        nrec += -(-(nl * nr) // 4096)
    info = dio.moio_info(p)
    assert info["length"] == len(x) and info["irrepx"] == sym and info["iop"] == iops
    np.testing.assert_array_equal(dio.moio_read(p), x)


def test_moio_records_are_4096_doubles(tmp_path, hctx):
    s = O.Sym(4)
    s.activate()
    V = np.array([40, 9, 17, 21, 0, 0, 0, 0])
    d = O.Desc([V, V, O.ZORB, O.ZORB], (1, 1), 1)
    x = np.arange(d.length(), dtype=np.float64) + 1.0
    p = str(tmp_path / "fab")
    dio.moio_write(hctx, p, 1, V, V, 1, O.ZORB, O.ZORB, 1, x)
    _, _, numl, rect, got = _getallso(p)
    # one right-pair irrep (the dummy pair) -> one block of all of x, padded
    assert rect[0] == 0 and numl[0] == len(x)
    assert np.fromfile(p, "<f8").size == 4096 * -(-len(x) // 4096)
    np.testing.assert_array_equal(got, x)


def test_moio_written_elsewhere_reads_back(tmp_path):
    """a list laid out here from clmoioso0.f + putallso (two irrep blocks of 5000
    and 3 doubles) is read by the library"""
    p = str(tmp_path / "abij")
    blocks = [np.arange(5000.0), np.array([7.0, 8.0, 9.0])]
    nrec = [2, 1]
    data = np.zeros(4096 * 3)
    data[:5000] = blocks[0]
    data[4096 * 2: 4096 * 2 + 3] = blocks[1]
    data.astype("<f8").tofile(p)
    numl = np.zeros((10, 20), np.int64)
    numr = np.zeros((10, 20), np.int64)
    rect = np.zeros((10, 20), np.int64)
    numl[0, 0], numr[0, 0], numl[1, 0], numr[1, 0] = 1000, 5, 3, 1
    rect[1, 0] = 2
    rect[0, 1] = 3

    def rec(a):
        b = np.asarray(a, dtype="<i8").ravel(order="F").tobytes()
        return struct.pack("<i", len(b)) + b + struct.pack("<i", len(b))
    c = rec([1]) + b"".join(rec(np.zeros(8)) for _ in range(8)) + rec(numl) + rec(numr) + rec(rect)
    open(p + ".c", "wb").write(c)
    np.testing.assert_array_equal(dio.moio_read(p), np.concatenate(blocks))


# ---- IO_stream / eigenvector file ---------------------------------------------------------
def test_stream_positions_and_eigvec_layout(tmp_path):
    p = str(tmp_path / "CIS_sym1")
    dim, nroot = 7, 3
    R = np.random.default_rng(1).standard_normal(nroot * dim)
    w = np.array([0.1, 0.2, 0.3])
    dio.eigvec_write(p, w, R, dim, istate=1)
    raw = np.fromfile(p, "<f8")
    for r in range(nroot):
        k = r * (1 + dim)  # (istate + iroot - 2)*(1 + dimA) with istate = 1
        assert raw[k] == w[r]
        np.testing.assert_array_equal(raw[k + 1: k + 1 + dim], R[r * dim:(r + 1) * dim])
    w2, R2 = dio.eigvec_read(p, nroot, dim)
    np.testing.assert_array_equal(w2, w)
    np.testing.assert_array_equal(R2, R)
    # a later state block (istate 4) starts at element (4-1)*(1+dim)+1
    dio.eigvec_write(p, [9.5], R[:dim], dim, istate=4, replace=False)
    raw = np.fromfile(p, "<f8")
    assert raw[3 * (1 + dim)] == 9.5
    # IO_stream '22' (value only) and '13' (vector only)
    dio.stream_write(p, 5, 42.0, mode=2)
    assert np.fromfile(p, "<f8")[4] == 42.0
    v, x = dio.stream_read(p, 2, n=3, mode=3)
    np.testing.assert_array_equal(x, raw[1:4])
    with pytest.raises(_lib.DpdError):
        dio.stream_read(p, 10 ** 6, n=2)
