"""The Fortran entry points (include/dpd_fortran.h) called the way the
reference's Fortran calls them -- every argument by reference, INTEGER*8,
hidden CHARACTER lengths at the end -- with host arrays (the igpu = 0 case,
staged through the device) and with device arrays (igpu = 1), against the
oracle on the same seeded inputs.  FP64: relative 1e-11 for the engines, exact
for the sorts."""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-11, 1e-12
I64 = ctypes.c_int64


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


@pytest.fixture(params=[4, 8], ids=["c2v", "d2h"])
def fsym(request, torch):
    from dpdhip import _lib
    L = _lib.lib()
    n = request.param
    dir_f = np.ascontiguousarray(O.xor_dirprd(n).reshape(8, 8).T.reshape(-1))
    L.dpd_fortran_init_(ctypes.byref(I64(n)), dir_f.ctypes.data_as(ctypes.POINTER(I64)), ctypes.byref(I64(0)))
    s = O.Sym(n)
    s.activate()
    yield s
    L.dpd_fortran_finalize_()


class Args:
    """keeps the by-reference scalars and population arrays alive for one call"""

    def __init__(self):
        self.keep = []

    def i(self, v):
        x = I64(int(v))
        self.keep.append(x)
        return ctypes.byref(x)

    def pop(self, p):
        a = np.ascontiguousarray(np.asarray(p, dtype=np.int64))
        self.keep.append(a)
        return a.ctypes.data_as(ctypes.POINTER(I64))

    def d(self, v):
        x = ctypes.c_double(v)
        self.keep.append(x)
        return ctypes.byref(x)

    def s(self, t):
        b = ctypes.create_string_buffer(t.encode(), len(t))
        self.keep.append(b)
        return b


def _spaces(sym):
    if sym.nirrep == 4:
        return np.array([3, 1, 2, 2, 0, 0, 0, 0]), np.array([2, 1, 0, 1, 0, 0, 0, 0])
    return np.array([2, 1, 1, 2, 1, 0, 1, 1]), np.array([1, 1, 0, 1, 1, 0, 0, 1])


def _ptr(torch, x, where):
    if where == "host":
        return x, x.ctypes.data_as(ctypes.c_void_p)
    t = torch.tensor(x, device="cuda:0")
    return t, ctypes.c_void_p(t.data_ptr())


def _back(x):
    return x if isinstance(x, np.ndarray) else x.cpu().numpy()


def _t4(a, d, x):
    return [x, a.pop(d.pops[0]), a.pop(d.pops[1]), a.i(d.iop[0]), a.pop(d.pops[2]), a.pop(d.pops[3]), a.i(d.iop[1]),
            a.i(d.sym)]


def _t2(a, p, q, iop, sym, x):
    return [x, a.pop(p), a.pop(q), a.i(iop), a.i(sym)]


@pytest.mark.parametrize("where", ["host", "device"])
def test_fortran_engines(torch, fsym, where):
    from dpdhip import _lib
    L = _lib.lib()
    V, Oc = _spaces(fsym)
    rng = np.random.default_rng(4)
    print(f"[fortran] engines nirrep={fsym.nirrep} {where}", flush=True)
    # 444 fast path (pp-ladder shape) and fallback (ring shape)
    for pops, iops, syms, t, f2 in [
        ([[V, V, V, V], [V, V, Oc, Oc], [V, V, Oc, Oc]], [(0, 0)] * 3, (1, 2, 2), ("abef", "efij", "abij"), 0.0),
        ([[V, V, Oc, Oc]] * 3, [(1, 1)] * 3, (1, 2, 2), ("aeim", "ebmj", "abij"), 0.5),
    ]:
        d = [O.Desc(pops[i], iops[i], syms[i]) for i in range(3)]
        x = [O.random_dpd(fsym, di, rng) for di in d]
        ref = x[2].copy()
        O.vv2v_dpd_444(x[0], d[0], t[0], x[1], d[1], t[1], ref, d[2], t[2], 1.25, f2)
        a = Args()
        h = [_ptr(torch, xi.copy(), where) for xi in x]
        L.vv2v_dpd_444_(*_t4(a, d[0], h[0][1]), a.s(t[0]), *_t4(a, d[1], h[1][1]), a.s(t[1]),
                        *_t4(a, d[2], h[2][1]), a.s(t[2]), a.d(1.25), a.d(f2),
                        ctypes.c_size_t(4), ctypes.c_size_t(4), ctypes.c_size_t(4))
        np.testing.assert_allclose(_back(h[2][0]), ref, rtol=RTOL, atol=ATOL, err_msg=str(t))
    # 424: abci(a,b,e,j) r1(e,i) -> c2(a,b,i,j)
    d1 = O.Desc([V, V, V, Oc], (1, 1), 1)
    d2 = O.desc2(V, Oc, 1, 2)
    d3 = O.Desc([V, V, Oc, Oc], (1, 1), 2)
    x1, x2, x3 = (O.random_dpd(fsym, di, rng) for di in (d1, d2, d3))
    ref = x3.copy()
    O.vv2v_dpd_424(x1, d1, "abej", x2, d2, "ei", ref, d3, "abij", 0.75, 1.0)
    a = Args()
    h1, h2, h3 = (_ptr(torch, xi.copy(), where) for xi in (x1, x2, x3))
    L.vv2v_dpd_424_(*_t4(a, d1, h1[1]), a.s("abej"), *_t2(a, V, Oc, 1, 2, h2[1]), a.s("ei"),
                    *_t4(a, d3, h3[1]), a.s("abij"), a.d(0.75), a.d(1.0),
                    ctypes.c_size_t(4), ctypes.c_size_t(2), ctypes.c_size_t(4))
    np.testing.assert_allclose(_back(h3[0]), ref, rtol=RTOL, atol=ATOL)
    # 442: abci(f,e,a,m) r2(f,e,i,m) -> c1(a,i)
    d1 = O.Desc([V, V, V, Oc], (1, 1), 1)
    d2 = O.Desc([V, V, Oc, Oc], (1, 1), 3)
    d3 = O.desc2(V, Oc, 1, 3)
    x1, x2, x3 = (O.random_dpd(fsym, di, rng) for di in (d1, d2, d3))
    ref = x3.copy()
    O.vv2v_dpd_442(x1, d1, "feam", x2, d2, "feim", ref, d3, "ai", -1.0, 0.0)
    a = Args()
    h1, h2, h3 = (_ptr(torch, xi.copy(), where) for xi in (x1, x2, x3))
    L.vv2v_dpd_442_(*_t4(a, d1, h1[1]), a.s("feam"), *_t4(a, d2, h2[1]), a.s("feim"),
                    *_t2(a, V, Oc, 1, 3, h3[1]), a.s("ai"), a.d(-1.0), a.d(0.0),
                    ctypes.c_size_t(4), ctypes.c_size_t(4), ctypes.c_size_t(2))
    np.testing.assert_allclose(_back(h3[0]), ref, rtol=RTOL, atol=ATOL)
    # 323: v(q,p,m) w(r,m) -> x(p,q,r) (the harness's scramble case)
    d1 = O.Desc([Oc, V, V, O.ZORB], (1, 1), 2)
    d2 = O.desc2(Oc, V, 1, 3)
    d3 = O.Desc([V, Oc, Oc, O.ZORB], (1, 1), O.xor_dirprd(fsym.nirrep)[1][2])
    x1, x2, x3 = (O.random_dpd(fsym, di, rng) for di in (d1, d2, d3))
    ref = x3.copy()
    O.vv2v_dpd_323(x1, d1, "qpm", x2, d2, "rm", ref, d3, "pqr", 0.5, 0.25)
    a = Args()
    h1, h2, h3 = (_ptr(torch, xi.copy(), where) for xi in (x1, x2, x3))
    L.vv2v_dpd_323_(h1[1], a.pop(Oc), a.pop(V), a.i(1), a.pop(V), a.i(2), a.s("qpm"),
                    *_t2(a, Oc, V, 1, 3, h2[1]), a.s("rm"),
                    h3[1], a.pop(V), a.pop(Oc), a.i(1), a.pop(Oc), a.i(d3.sym), a.s("pqr"), a.d(0.5), a.d(0.25),
                    ctypes.c_size_t(3), ctypes.c_size_t(2), ctypes.c_size_t(3))
    np.testing.assert_allclose(_back(h3[0]), ref, rtol=RTOL, atol=ATOL)


@pytest.mark.parametrize("where", ["host", "device"])
def test_fortran_sorts(torch, fsym, where):
    from dpdhip import _lib
    L = _lib.lib()
    V, Oc = _spaces(fsym)
    rng = np.random.default_rng(8)
    print(f"[fortran] sorts nirrep={fsym.nirrep} {where}", flush=True)

    def run(fn, x, nout, *args):
        hin = _ptr(torch, x.copy(), where)
        hout = _ptr(torch, np.zeros(max(nout, 1)), where)
        getattr(L, fn)(hin[1], hout[1], *args)
        return _back(hout[0])[:nout]

    for irrepx in range(1, fsym.nirrep + 1):
        a = Args()
        d0 = O.Desc([V, V, Oc, Oc], (0, 0), irrepx)
        x = O.random_dpd(fsym, d0, rng)
        ref = O.iexpso(x, V, V, 0, 1, Oc, Oc, 0, 1, irrepx)
        got = run("iexpso_", x, len(ref), a.pop(V), a.pop(V), a.i(0), a.i(1), a.pop(Oc), a.pop(Oc), a.i(0), a.i(1),
                  a.i(irrepx))
        assert np.array_equal(got, ref)
        full = O.iexpso(x, V, V, 0, 1, Oc, Oc, 0, 0, irrepx)
        ref = O.icompso(full, V, V, 1, 0, Oc, Oc, 0, 1, irrepx)
        got = run("icompso_", full, len(ref), a.pop(V), a.pop(V), a.i(1), a.i(0), a.pop(Oc), a.pop(Oc), a.i(0),
                  a.i(1), a.i(irrepx))
        assert np.array_equal(got, ref)
        d1 = O.Desc([V, Oc, V, Oc], (1, 1), irrepx)
        y = rng.standard_normal(d1.length())
        ref = O.isymtrso(y, V, Oc, 1, 1, V, Oc, 1, 1, irrepx)
        got = run("isymtrso_", y, len(ref), a.pop(V), a.pop(Oc), a.i(1), a.i(1), a.pop(V), a.pop(Oc), a.i(1), a.i(1),
                  a.i(irrepx))
        assert np.array_equal(got, ref)
        ref = O.itranspso(y, V, Oc, 1, V, Oc, 1, irrepx)
        got = run("itranspso_", y, len(ref), a.pop(V), a.pop(Oc), a.i(1), a.pop(V), a.pop(Oc), a.i(1), a.i(irrepx))
        assert np.array_equal(got, ref)
        ref = O.sstgenso(y, V, Oc, V, Oc, irrepx, "3142")
        got = run("sstgenso1_", y, len(ref), a.pop(V), a.pop(Oc), a.pop(V), a.pop(Oc), a.i(irrepx), a.s("3142"),
                  ctypes.c_size_t(4))
        assert np.array_equal(got, ref)
        io = {0: (0, 1), 1: (1, 1), 2: (1, 1), 3: (1, 1), 4: (1, 0), 5: (1, 1)}
        for iop in range(6):
            di = O.Desc([V, V, Oc, Oc], io[iop], irrepx)
            z = O.random_dpd(fsym, di, rng)
            ref = O.iassymso(z, irrepx, V, V, Oc, Oc, iop)
            got = run("iassymso_", z, len(ref), a.i(irrepx), a.pop(V), a.pop(V), a.pop(Oc), a.pop(Oc), a.i(iop))
            assert np.array_equal(got, ref), iop


@pytest.mark.parametrize("where", ["host", "device"])
def test_fortran_xgemmso(torch, fsym, where):
    from dpdhip import _lib
    L = _lib.lib()
    rng = np.random.default_rng(2)
    for ta, tb, m, n, k in [("N", "N", 37, 19, 23), ("T", "N", 64, 65, 33), ("N", "T", 5, 130, 7),
                            ("T", "T", 70, 3, 129)]:
        A = rng.standard_normal((k, m) if ta == "T" else (m, k))
        B = rng.standard_normal((n, k) if tb == "T" else (k, n))
        C = rng.standard_normal((m, n))
        ref = 0.5 * (A.T if ta == "T" else A) @ (B.T if tb == "T" else B) + 2.0 * C
        # column-major storage: the Fortran arrays are the transposes in C order
        af, bf, cf = (np.ascontiguousarray(X.T).reshape(-1) for X in (A, B, C))
        a = Args()
        ha, hb, hc = (_ptr(torch, X, where) for X in (af, bf, cf))
        L.xgemmso_(a.s(ta), a.s(tb), a.i(m), a.i(n), a.i(k), a.d(0.5), ha[1], a.i(A.shape[0]), hb[1],
                   a.i(B.shape[0]), a.d(2.0), hc[1], a.i(m), ctypes.c_size_t(1), ctypes.c_size_t(1))
        got = _back(hc[0]).reshape(n, m).T
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("where", ["host", "device"])
def test_fortran_xgemmso_empty_blocks(torch, fsym, where):
    """Callers loop over every irrep with leading dimensions equal to block sizes
    that can be 0 (e.g. VpqmnVrsmn_to_Vpqrs.f:93 passes LDA=INUMPQ(IRREPPQ));
    the reference wraps each one as MAX(1,LD) (xgemmso.f:11-22).  m = 0 or
    n = 0 leaves C alone; k = 0 gives C = beta*C."""
    from dpdhip import _lib
    L = _lib.lib()
    rng = np.random.default_rng(3)
    one = ctypes.c_size_t(1)
    # m = 0, ldc = lda = 0: nothing is read or written
    C = rng.standard_normal(6)
    a = Args()
    hc = _ptr(torch, C.copy(), where)
    hx = _ptr(torch, rng.standard_normal(4), where)
    L.xgemmso_(a.s("N"), a.s("N"), a.i(0), a.i(3), a.i(2), a.d(1.0), hx[1], a.i(0), hx[1], a.i(2), a.d(0.5),
               hc[1], a.i(0), one, one)
    np.testing.assert_array_equal(_back(hc[0]), C)
    # n = 0, ldb = 0
    a = Args()
    L.xgemmso_(a.s("N"), a.s("T"), a.i(3), a.i(0), a.i(2), a.d(1.0), hx[1], a.i(3), hx[1], a.i(0), a.d(0.5),
               hc[1], a.i(3), one, one)
    np.testing.assert_array_equal(_back(hc[0]), C)
    # k = 0 with lda = 3, ldb = 0: C = beta * C (3 x 2)
    a = Args()
    L.xgemmso_(a.s("N"), a.s("N"), a.i(3), a.i(2), a.i(0), a.d(1.0), hx[1], a.i(3), hx[1], a.i(0), a.d(0.5),
               hc[1], a.i(3), one, one)
    np.testing.assert_allclose(_back(hc[0]), 0.5 * C, rtol=0, atol=1e-15)
