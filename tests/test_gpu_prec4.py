"""FP32 tensor storage at the boundary: the reference's prec = 4 route
(libsrint/readinp.f90:120-127; vv2v_dpd_444.f90:206-229 -> vv2v_dpd_444_sg,
sstgenso.f90:35-42, xgemmso.f:21-22) through dpd_ctx_set_data_precision and the
Fortran dpd_fortran_set_prec_.

The oracle runs in FP64 on the same inputs rounded to FP32.  Engines and xgemm
accumulate in FP32 MFMA and round the update fact2*v3 + fact1*C once: relative
2e-5 of the result's scale.  Sorts move FP32 values exactly and iassymso rounds
one sum or difference once, so they equal the oracle's FP64 result rounded to
FP32 bit for bit."""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

PV = np.array([3, 1, 2, 2, 0, 0, 0, 0])
PO = np.array([2, 1, 0, 1, 0, 0, 0, 0])
PVL = np.array([21, 9, 14, 17, 0, 0, 0, 0])
POL = np.array([4, 2, 3, 3, 0, 0, 0, 0])
PV8 = np.array([2, 1, 1, 2, 1, 0, 1, 1])
PO8 = np.array([1, 1, 0, 1, 1, 0, 0, 1])


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


@pytest.fixture(params=[4, 8], ids=["c2v", "d2h"])
def sym(request, torch):
    import dpdhip
    s = O.Sym(request.param)
    s.activate()
    ctx = dpdhip.Context(request.param, 0)
    ctx.set_data_precision(4)
    s.ctx = ctx
    yield s
    ctx.close()


def f32(x):
    return np.asarray(x, dtype=np.float32)


def dev32(torch, x):
    return torch.from_numpy(np.ascontiguousarray(f32(x))).to("cuda:0")


def gopt(torch, sym, desc, x):
    import dpdhip
    return dpdhip.Opt(sym.ctx, *desc.pops, ioppq=desc.iop[0], ioprs=desc.iop[1], sym=desc.sym,
                      arr=dev32(torch, np.concatenate([x, [0.0]])))


def close32(got, exp, what=""):
    scale = max(1.0, float(np.abs(exp).max()))
    np.testing.assert_allclose(got, exp, rtol=2e-5, atol=2e-5 * scale, err_msg=what)


def _engine(torch, sym, kind, pops, iops, syms, t1, t2, t3, f1=1.0, f2=0.0, seed=0):
    import dpdhip
    rng = np.random.default_rng(seed)
    d = [O.Desc(pops[i], iops[i], syms[i]) for i in range(3)]
    x = [f32(O.random_dpd(sym, di, rng)).astype(np.float64) for di in d]
    ref = x[2].copy()
    {444: O.vv2v_dpd_444, 424: O.vv2v_dpd_424, 442: O.vv2v_dpd_442}[kind](x[0], d[0], t1, x[1], d[1], t2, ref, d[2],
                                                                          t3, f1, f2)
    g = [gopt(torch, sym, di, xi) for di, xi in zip(d, x)]
    {444: dpdhip.vv2v_dpd_444, 424: dpdhip.vv2v_dpd_424, 442: dpdhip.vv2v_dpd_442}[kind](g[0], t1, g[1], t2, g[2], t3,
                                                                                        f1, f2)
    assert g[2].arr.dtype == torch.float32
    close32(g[2].numpy(), ref, f"{kind} {t1}*{t2}->{t3}")


def test_fp32_engines(torch, sym):
    V, Oc = (PV, PO) if sym.nirrep == 4 else (PV8, PO8)
    print(f"[prec4] engines nirrep={sym.nirrep}", flush=True)
    # 444 fast paths (packed pp-ladder, hh-ladder accumulate), fallback and scrambles
    _engine(torch, sym, 444, [[V, V, V, V], [V, V, Oc, Oc], [V, V, Oc, Oc]], [(0, 0)] * 3, (1, 2, 2),
            "abef", "efij", "abij")
    _engine(torch, sym, 444, [[V, V, Oc, Oc], [V, V, Oc, Oc], [Oc, Oc, Oc, Oc]], [(0, 0)] * 3, (1, 1, 1),
            "efij", "efkl", "ijkl", f2=0.5)
    _engine(torch, sym, 444, [[V, V, Oc, Oc]] * 3, [(1, 1)] * 3, (1, 2, 2), "aeim", "ebmj", "abij", f1=1.25, f2=0.5)
    _engine(torch, sym, 444, [[Oc, V, V, Oc], [V, V, Oc, Oc], [Oc, V, Oc, V]], [(1, 1)] * 3, (1, 3, 3), "faeb",
            "iejf", "baji", f2=-1.0)
    # 424 and 442 (the rank-2 operand pairs with a 1-sized dummy pair)
    _engine(torch, sym, 424, [[V, V, V, Oc], [V, Oc, O.ZORB, O.ZORB], [V, V, Oc, Oc]], [(1, 1)] * 3, (1, 2, 2),
            "abej", "ei", "abij", f1=0.75, f2=1.0)
    _engine(torch, sym, 442, [[V, V, V, Oc], [V, V, Oc, Oc], [V, Oc, O.ZORB, O.ZORB]], [(1, 1)] * 3, (1, 3, 3),
            "feam", "feim", "ai", f1=-1.0)


def test_fp32_engine_large_blocks(torch):
    """tiles, several K segments and split-K with FP32 results accumulated in place"""
    import dpdhip
    s = O.Sym(4)
    s.activate()
    s.ctx = dpdhip.Context(4, 0)
    s.ctx.set_data_precision(4)
    try:
        _engine(torch, s, 444, [[PVL, PVL, PVL, PVL], [PVL, PVL, POL, POL], [PVL, PVL, POL, POL]], [(0, 0)] * 3,
                (1, 2, 2), "abef", "efij", "abij", f2=1.0, seed=3)
        _engine(torch, s, 444, [[PVL, PVL, POL, POL]] * 3, [(1, 1)] * 3, (2, 1, 2), "eaim", "ebmj", "abij", f2=0.5,
                seed=4)
    finally:
        s.ctx.close()


def test_fp32_engine_323(torch, sym):
    import dpdhip
    from test_oracle_engine import _cases323, _desc3
    V, Oc = (PV, PO) if sym.nirrep == 4 else (PV8, PO8)
    for case, (t1, p1, t2, p2, t3, p3, iop, f2) in enumerate(_cases323(V, Oc)):
        s1, s2 = 2, 3
        s3 = O.xor_dirprd(sym.nirrep)[s1 - 1][s2 - 1]
        rng = np.random.default_rng(case)
        d1 = _desc3(p1, iop if t1 == "pqm" else 1, s1)
        d2 = O.desc2(p2[0], p2[1], 1, s2)
        d3 = _desc3(p3, iop if t3 == "pqr" else 1, s3)
        x1, x2, x3 = (f32(O.random_dpd(sym, d, rng)).astype(np.float64) for d in (d1, d2, d3))
        ref = x3.copy()
        O.vv2v_dpd_323(x1, d1, t1, x2, d2, t2, ref, d3, t3, 1.5, f2)
        g1, g2, g3 = gopt(torch, sym, d1, x1), gopt(torch, sym, d2, x2), gopt(torch, sym, d3, x3)
        dpdhip.vv2v_dpd_323(g1, t1, g2, t2, g3, t3, 1.5, f2)
        close32(g3.numpy(), ref, f"case {case}")


def test_fp32_sorts_exact(torch, sym):
    import dpdhip
    c = sym.ctx
    V, Oc = (PV, PO) if sym.nirrep == 4 else (PV8, PO8)
    rng = np.random.default_rng(7)

    def run(fn, x, n, *args):
        out = torch.zeros(n + 1, dtype=torch.float32, device="cuda:0")
        fn(c, dev32(torch, x), out, *args)
        return out[:n].cpu().numpy()

    for irrepx in range(1, sym.nirrep + 1):
        x = f32(O.random_dpd(sym, O.Desc([V, V, Oc, Oc], (0, 0), irrepx), rng)).astype(np.float64)
        for iexp in [(1, 0), (0, 1), (1, 1)]:
            ref = O.iexpso(x, V, V, 0, iexp[0], Oc, Oc, 0, iexp[1], irrepx)
            got = run(dpdhip.iexpso, x, len(ref), V, V, 0, iexp[0], Oc, Oc, 0, iexp[1], irrepx)
            assert np.array_equal(got, f32(ref))
        full = O.iexpso(x, V, V, 0, 1, Oc, Oc, 0, 0, irrepx)
        ref = O.icompso(full, V, V, 1, 0, Oc, Oc, 0, 1, irrepx)
        assert np.array_equal(run(dpdhip.icompso, full, len(ref), V, V, 1, 0, Oc, Oc, 0, 1, irrepx), f32(ref))
        y = f32(rng.standard_normal(O.Desc([V, Oc, V, Oc], (1, 1), irrepx).length())).astype(np.float64)
        for icpq, icrs in [(1, 0), (0, 1), (1, 1)]:
            ref = O.isymtrso(y, V, Oc, 1, icpq, V, Oc, 1, icrs, irrepx)
            assert np.array_equal(run(dpdhip.isymtrso, y, len(y), V, Oc, 1, icpq, V, Oc, 1, icrs, irrepx), f32(ref))
        ref = O.itranspso(y, V, Oc, 1, V, Oc, 1, irrepx)
        assert np.array_equal(run(dpdhip.itranspso, y, len(y), V, Oc, 1, V, Oc, 1, irrepx), f32(ref))
        io = {0: (0, 1), 1: (1, 1), 2: (1, 1), 3: (1, 1), 4: (1, 0), 5: (1, 1)}
        for iop in range(6):
            z = f32(O.random_dpd(sym, O.Desc([V, V, Oc, Oc], io[iop], irrepx), rng)).astype(np.float64)
            ref = O.iassymso(z, irrepx, V, V, Oc, Oc, iop)
            assert np.array_equal(run(dpdhip.iassymso, z, len(ref), irrepx, V, V, Oc, Oc, iop), f32(ref)), iop
        for t in ("3142", "2413", "4321"):
            ref = O.sstgenso(y, V, Oc, V, Oc, irrepx, t)
            assert np.array_equal(run(dpdhip.sstgenso, y, len(y), V, Oc, V, Oc, irrepx, t), f32(ref))


@pytest.mark.parametrize("shape", [(70, 45, 37, "N", "N"), (65, 130, 100, "T", "N"), (128, 20, 300, "T", "T")])
def test_fp32_xgemm(torch, sym, shape):
    import dpdhip
    m, n, k, ta, tb = shape
    rng = np.random.default_rng(m + n)
    A = f32(rng.standard_normal((m, k) if ta == "N" else (k, m)))
    B = f32(rng.standard_normal((k, n) if tb == "N" else (n, k)))
    C = f32(rng.standard_normal((m, n)))
    opA = (A if ta == "N" else A.T).astype(np.float64)
    opB = (B if tb == "N" else B.T).astype(np.float64)
    exp = 1.5 * (opA @ opB) - 0.5 * C.astype(np.float64)
    gC = dev32(torch, C.flatten(order="F"))
    dpdhip.xgemm(sym.ctx, ta, tb, m, n, k, 1.5, dev32(torch, A.flatten(order="F")), A.shape[0],
                 dev32(torch, B.flatten(order="F")), B.shape[0], -0.5, gC, m)
    close32(gC.cpu().numpy().reshape(m, n, order="F"), exp)


def test_fp32_refusals(torch, sym):
    """FP64 tensors are refused under prec 4 before a pointer is passed; the
    contract surface (lib_cc runs with prec = 8) refuses FP32 data"""
    import dpdhip
    V, Oc = (PV, PO) if sym.nirrep == 4 else (PV8, PO8)
    d = O.Desc([V, V, Oc, Oc], (1, 1), 1)
    g = gopt(torch, sym, d, np.zeros(d.length()))
    with pytest.raises(ValueError):
        dpdhip.Opt(sym.ctx, *d.pops, ioppq=1, ioprs=1, sym=1,
                   arr=torch.zeros(d.length(), dtype=torch.float64, device="cuda:0"))._struct()
    with pytest.raises(dpdhip.DpdError):
        dpdhip.vv2v_444(g, g, g, 1.0, "abij", "ijcd", "abcd", 0)
    with pytest.raises(dpdhip.DpdError):
        sym.ctx.set_data_precision(2)
    assert sym.ctx.data_prec == 4


def test_fortran_prec4(torch):
    """the Fortran shims under prec = 4: REAL*4 arrays through the same dummy
    arguments, from host memory (staged) and from device memory"""
    from dpdhip import _lib
    L = _lib.lib()
    I64 = ctypes.c_int64
    n = 4
    dir_f = np.ascontiguousarray(O.xor_dirprd(n).reshape(8, 8).T.reshape(-1))
    L.dpd_fortran_init_(ctypes.byref(I64(n)), dir_f.ctypes.data_as(ctypes.POINTER(I64)), ctypes.byref(I64(0)))
    try:
        L.dpd_fortran_set_prec_(ctypes.byref(I64(4)))
        s = O.Sym(n)
        s.activate()
        keep = []

        def i(v):
            x = I64(int(v))
            keep.append(x)
            return ctypes.byref(x)

        def pop(p):
            a = np.ascontiguousarray(np.asarray(p, dtype=np.int64))
            keep.append(a)
            return a.ctypes.data_as(ctypes.POINTER(I64))

        def dbl(v):
            x = ctypes.c_double(v)
            keep.append(x)
            return ctypes.byref(x)

        def lab(t):
            b = ctypes.create_string_buffer(t.encode(), len(t))
            keep.append(b)
            return b

        def t4(dd, p):
            return [p, pop(dd.pops[0]), pop(dd.pops[1]), i(dd.iop[0]), pop(dd.pops[2]), pop(dd.pops[3]), i(dd.iop[1]),
                    i(dd.sym)]

        rng = np.random.default_rng(11)
        for where in ("host", "device"):
            d = [O.Desc([PV, PV, PO, PO], (1, 1), s_) for s_ in (1, 2, 2)]
            x = [f32(O.random_dpd(s, di, rng)) for di in d]
            ref = x[2].astype(np.float64)
            O.vv2v_dpd_444(x[0].astype(np.float64), d[0], "aeim", x[1].astype(np.float64), d[1], "ebmj", ref, d[2],
                           "abij", 1.25, 0.5)
            if where == "host":
                h = [xi.copy() for xi in x]
                ptr = [hi.ctypes.data_as(ctypes.c_void_p) for hi in h]
            else:
                h = [torch.from_numpy(xi.copy()).to("cuda:0") for xi in x]
                ptr = [ctypes.c_void_p(hi.data_ptr()) for hi in h]
            L.vv2v_dpd_444_(*t4(d[0], ptr[0]), lab("aeim"), *t4(d[1], ptr[1]), lab("ebmj"), *t4(d[2], ptr[2]),
                            lab("abij"), dbl(1.25), dbl(0.5), ctypes.c_size_t(4), ctypes.c_size_t(4),
                            ctypes.c_size_t(4))
            got = h[2] if where == "host" else h[2].cpu().numpy()
            assert got.dtype == np.float32
            close32(got, ref, where)
            y = f32(rng.standard_normal(O.Desc([PV, PO, PV, PO], (1, 1), 3).length()))
            ref = f32(O.sstgenso(y.astype(np.float64), PV, PO, PV, PO, 3, "3142"))
            if where == "host":
                out = np.zeros_like(y)
                L.sstgenso1_(y.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p), pop(PV),
                             pop(PO), pop(PV), pop(PO), i(3), lab("3142"), ctypes.c_size_t(4))
            else:
                gy = torch.from_numpy(y).to("cuda:0")
                go = torch.zeros_like(gy)
                L.sstgenso1_(ctypes.c_void_p(gy.data_ptr()), ctypes.c_void_p(go.data_ptr()), pop(PV), pop(PO),
                             pop(PV), pop(PO), i(3), lab("3142"), ctypes.c_size_t(4))
                out = go.cpu().numpy()
            assert np.array_equal(out, ref)
    finally:
        L.dpd_fortran_finalize_()
