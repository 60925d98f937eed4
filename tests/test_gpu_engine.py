"""GPU parity of the HIP engine against the CPU oracle (same seeded inputs).

Covers the MFMA lane layout, the engine level (vv2v_dpd_444/424/442 fast paths,
fallback and scrambles, accumulation), the mod_contract surface with its
expand/antisymmetrise/compress semantics, every sort kernel and the Davidson
vector primitives.  FP64 tolerance: relative 1e-11 (only the summation order
differs from the oracle's BLAS)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-11, 1e-12

PV = np.array([3, 1, 2, 2, 0, 0, 0, 0])
PO = np.array([2, 1, 0, 1, 0, 0, 0, 0])
PV8 = np.array([2, 1, 1, 2, 1, 0, 1, 1])
PO8 = np.array([1, 1, 0, 1, 1, 0, 0, 1])
# larger spaces so that tiles, split-K and several K segments are exercised
PVL = np.array([21, 9, 14, 17, 0, 0, 0, 0])
POL = np.array([4, 2, 3, 3, 0, 0, 0, 0])


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    return t


_ctxs = {}


def ctx_for(n):
    import dpdhip
    if n not in _ctxs:
        _ctxs[n] = dpdhip.Context(n, 0)
    return _ctxs[n]


# GEMM kernels (tile configuration, packing): automatic (timed) and automatic
# (cost model); first-generation
# kernel; gathering kernel with 64x64 / 128x64 / 64x128 / 128x128 tiles; the
# packed-panel kernel with automatic and with each fixed tile shape
GEMM_CFGS = [(-1, 1), (-2, 1), (0, 0), (1, 0), (2, 0), (3, 0), (4, 0), (5, 0), (-1, 2), (1, 2), (2, 2), (3, 2),
             (4, 2), (5, 2), (6, 2), (7, 2)]
GEMM_IDS = ["auto", "model", "v1", "g64x64", "g128x64", "g64x128", "g128x128", "g32x32", "pk_auto", "pk64x64",
            "pk128x64", "pk64x128", "pk128x128", "pk32x32", "pk64x192", "pk128x64d"]


@pytest.fixture(params=GEMM_CFGS, ids=GEMM_IDS)
def gemm_cfg(request, torch):
    cfg, packing = request.param
    for n in (4, 8):
        ctx_for(n).set_gemm_config(cfg, packing)
    yield request.param
    for n in (4, 8):
        ctx_for(n).set_gemm_config(-1, 1)


@pytest.fixture(params=[4, 8], ids=["c2v", "d2h"])
def sym(request, torch):
    s = O.Sym(request.param)
    s.activate()
    return s


def spaces(sym, large=False):
    if sym.nirrep == 4:
        return (PVL, POL) if large else (PV, PO)
    return PV8, PO8


def dev(torch, x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")


def gopt(torch, sym, desc, x):
    import dpdhip
    arr = dev(torch, np.concatenate([x, [0.0]]))
    return dpdhip.Opt(ctx_for(sym.nirrep), *desc.pops, ioppq=desc.iop[0], ioprs=desc.iop[1], sym=desc.sym, arr=arr)


def test_mfma_lane_layout_single_tile(torch, sym):
    """A 16x16x4 block: the smallest GEMM, checks the MFMA operand/result layout."""
    s = sym
    p = np.zeros(8, np.int64)
    p[0] = 16
    k = np.zeros(8, np.int64)
    k[0] = 4
    d1 = O.Desc([p, k, O.ZORB, O.ZORB], (1, 1), 1)
    d2 = O.Desc([k, p, O.ZORB, O.ZORB], (1, 1), 1)
    d3 = O.Desc([p, p, O.ZORB, O.ZORB], (1, 1), 1)
    rng = np.random.default_rng(0)
    a = rng.standard_normal(64)
    b = rng.standard_normal(64)
    g1, g2, g3 = gopt(torch, s, d1, a), gopt(torch, s, d2, b), gopt(torch, s, d3, np.zeros(256))
    import dpdhip
    dpdhip.vv2v_222(g1, g2, g3, 1.0, "ik", "kj", "ij")
    got = g3.numpy().reshape(16, 16, order="F")
    exp = a.reshape(16, 4, order="F") @ b.reshape(4, 16, order="F")
    np.testing.assert_allclose(got, exp, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", [(16, 16, 4, "N", "N"), (70, 45, 37, "N", "N"), (65, 130, 100, "T", "N"),
                                   (33, 64, 17, "N", "T"), (128, 20, 300, "T", "T"),
                                   (257, 190, 129, "T", "N")])
def test_xgemm_layout_and_transposes(torch, prec, shape, gemm_cfg):
    """dense GEMM through the same MFMA kernel: pins the FP64 and FP32 MFMA operand
    and result layouts directly (one 16x16x4 tile first, then ragged tiles)."""
    import dpdhip
    m, n, k, ta, tb = shape
    rng = np.random.default_rng(m * 7 + n)
    A = rng.standard_normal((m, k) if ta == "N" else (k, m))
    B = rng.standard_normal((k, n) if tb == "N" else (n, k))
    C = rng.standard_normal((m, n))
    opA = A if ta == "N" else A.T
    opB = B if tb == "N" else B.T
    if prec == "fp32":
        exp = 1.5 * (opA.astype(np.float32).astype(np.float64) @ opB.astype(np.float32).astype(np.float64)) - 0.5 * C
        tol = dict(rtol=2e-5, atol=2e-5 * np.sqrt(k))
    else:
        exp = 1.5 * (opA @ opB) - 0.5 * C
        tol = dict(rtol=1e-13, atol=1e-13 * k)
    gA = dev(torch, A.flatten(order="F"))
    gB = dev(torch, B.flatten(order="F"))
    gC = dev(torch, C.flatten(order="F"))
    dpdhip.xgemm(ctx_for(4), ta, tb, m, n, k, 1.5, gA, A.shape[0], gB, B.shape[0], -0.5, gC, m, prec)
    got = gC.cpu().numpy().reshape(m, n, order="F")
    np.testing.assert_allclose(got, exp, **tol)


@pytest.mark.parametrize("shape", [(16, 16, 32, "N", "N"), (70, 45, 37, "N", "N"), (65, 130, 100, "T", "N"),
                                   (257, 190, 129, "T", "N"), (128, 20, 300, "T", "T")])
def test_xgemm_bf16x3(torch, shape, gemm_cfg):
    """BF16x3 MFMA path: matches hi*hi + hi*lo + lo*hi of the emulated split to
    FP32-accumulation accuracy, and the exact product to the split's accuracy."""
    import dpdhip
    if gemm_cfg[1] != 2 and gemm_cfg != (-1, 1):
        pytest.skip("BF16x3 always runs the packed kernel")
    m, n, k, ta, tb = shape
    rng = np.random.default_rng(m * 11 + k)
    A = rng.standard_normal((m, k) if ta == "N" else (k, m))
    B = rng.standard_normal((k, n) if tb == "N" else (n, k))
    C = rng.standard_normal((m, n))
    opA = A if ta == "N" else A.T
    opB = B if tb == "N" else B.T
    from bf16_emulation import bf16x3_split
    ah, al = bf16x3_split(opA)
    bh, bl = bf16x3_split(opB)
    emu = 1.5 * (ah @ bh + ah @ bl + al @ bh) - 0.5 * C
    gA = dev(torch, A.flatten(order="F"))
    gB = dev(torch, B.flatten(order="F"))
    gC = dev(torch, C.flatten(order="F"))
    dpdhip.xgemm(ctx_for(4), ta, tb, m, n, k, 1.5, gA, A.shape[0], gB, B.shape[0], -0.5, gC, m, "bf16x3")
    got = gC.cpu().numpy().reshape(m, n, order="F")
    mag = 1.5 * (np.abs(opA) @ np.abs(opB))
    # FP32 accumulation of the three products: at most k roundings of 2^-24
    assert np.all(np.abs(got - emu) <= max(k, 32) * 2.0 ** -24 * mag + 1e-12), np.abs(got - emu).max()
    exact = 1.5 * (opA @ opB) - 0.5 * C
    assert np.all(np.abs(got - exact) <= 3 * 2.0 ** -16 * mag + 1e-12)


def _engine_case(torch, sym, kind, pops, iops, syms, t1, t2, t3, f1=1.0, f2=0.0, seed=0):
    import dpdhip
    rng = np.random.default_rng(seed)
    d1 = O.Desc(pops[0], iops[0], syms[0])
    d2 = O.Desc(pops[1], iops[1], syms[1])
    d3 = O.Desc(pops[2], iops[2], syms[2])
    x1, x2, x3 = O.random_dpd(sym, d1, rng), O.random_dpd(sym, d2, rng), O.random_dpd(sym, d3, rng)
    ref = x3.copy()
    {444: O.vv2v_dpd_444, 424: O.vv2v_dpd_424, 442: O.vv2v_dpd_442}[kind](x1, d1, t1, x2, d2, t2, ref, d3, t3, f1, f2)
    g1, g2, g3 = gopt(torch, sym, d1, x1), gopt(torch, sym, d2, x2), gopt(torch, sym, d3, x3)
    {444: dpdhip.vv2v_dpd_444, 424: dpdhip.vv2v_dpd_424, 442: dpdhip.vv2v_dpd_442}[kind](g1, t1, g2, t2, g3, t3, f1,
                                                                                        f2)
    np.testing.assert_allclose(g3.numpy(), ref, rtol=RTOL, atol=ATOL)


def test_engine_fast_paths(torch, sym, gemm_cfg):
    for large in (False, True):
        if large and sym.nirrep == 8:
            continue
        V, Oc = spaces(sym, large)
        _engine_case(torch, sym, 444, [[V, V, V, V], [V, V, Oc, Oc], [V, V, Oc, Oc]], [(0, 0)] * 3, (1, 2, 2),
                     "abef", "efij", "abij")
        _engine_case(torch, sym, 444, [[V, V, Oc, Oc], [V, V, Oc, Oc], [Oc, Oc, Oc, Oc]], [(0, 0)] * 3, (1, 1, 1),
                     "efij", "efkl", "ijkl", f2=0.5)
        _engine_case(torch, sym, 444, [[V, V, Oc, Oc], [V, V, Oc, Oc], [V, V, V, V]], [(0, 0)] * 3, (3, 3, 1),
                     "abmn", "cdmn", "abcd")
        _engine_case(torch, sym, 444, [[Oc, Oc, V, V], [V, V, Oc, Oc], [V, V, V, V]], [(1, 1)] * 3, (2, 2, 1),
                     "mnab", "cdmn", "abcd", f2=0.5)


def test_engine_fallback_and_scrambles(torch, sym, gemm_cfg):
    V, Oc = spaces(sym)
    full = [(1, 1)] * 3
    _engine_case(torch, sym, 444, [[V, V, Oc, Oc]] * 3, full, (1, 2, 2), "aeim", "ebmj", "abij")
    _engine_case(torch, sym, 444, [[V, V, Oc, Oc]] * 3, full, (2, 1, 2), "eaim", "ebmj", "abij", f2=0.5)
    _engine_case(torch, sym, 444, [[Oc, V, V, Oc], [V, V, Oc, Oc], [Oc, V, Oc, V]], full, (1, 3, 3), "faeb", "iejf",
                 "baji", f2=0.5)
    _engine_case(torch, sym, 444, [[V, Oc, V, Oc], [V, Oc, V, Oc], [V, V, V, V]], full, (2, 3, 4), "pmqn", "rmsn",
                 "prqs")


def test_engine_errors(torch, sym):
    import dpdhip
    V, Oc = spaces(sym)
    d1 = O.Desc([V, V, Oc, Oc], (1, 1), 1)
    d2 = O.Desc([V, V, Oc, Oc], (0, 0), 1)
    g1 = gopt(torch, sym, d1, np.zeros(d1.length()))
    g2 = gopt(torch, sym, d2, np.zeros(d2.length()))
    with pytest.raises(dpdhip.DpdError) as e:
        dpdhip.vv2v_dpd_444(g1, "aeim", g2, "ebmj", g2, "abij", 1.0, 0.0)
    assert e.value.code == -4  # IOP barrier of the fallback
    d3 = O.Desc([V, Oc, V, Oc], (1, 1), 1)
    g3 = gopt(torch, sym, d3, np.zeros(d3.length()))
    with pytest.raises(dpdhip.DpdError) as e:
        dpdhip.vv2v_dpd_444(g3, "aeim", g1, "ebmj", g1, "abij", 1.0, 0.0)
    assert e.value.code == -3  # contracted populations differ
    with pytest.raises(dpdhip.DpdError):
        dpdhip.vv2v_444(g2, g2, g2, 1.0, "abef", "efij", "abij", 4)


def _surface_case(torch, sym, kind, ops, t, fact, anti, seed=1):
    import dpdhip
    rng = np.random.default_rng(seed)
    d1, d2, d3 = ops
    x1, x2, x3 = O.random_dpd(sym, d1, rng), O.random_dpd(sym, d2, rng), O.random_dpd(sym, d3, rng)
    v3 = O.Opt(d3, x3.copy())
    t1, t2, t3 = t
    fo = {"444": O.vv2v_444, "424": O.vv2v_424, "442": O.vv2v_442, "222": O.vv2v_222, "422": O.vv2v_422,
          "224": O.vv2v_224}[kind]
    fg = {"444": dpdhip.vv2v_444, "424": dpdhip.vv2v_424, "442": dpdhip.vv2v_442, "222": dpdhip.vv2v_222,
          "422": dpdhip.vv2v_422, "224": dpdhip.vv2v_224}[kind]
    extra = (anti,) if kind in ("444", "424", "224") else ()
    fo(O.Opt(d1, x1), O.Opt(d2, x2), v3, fact, t1, t2, t3, *extra)
    g3 = gopt(torch, sym, d3, x3)
    fg(gopt(torch, sym, d1, x1), gopt(torch, sym, d2, x2), g3, fact, t1, t2, t3, *extra)
    np.testing.assert_allclose(g3.numpy(), v3.arr, rtol=RTOL, atol=ATOL, err_msg=f"{kind} {t}")


@pytest.mark.parametrize("large", [False, True], ids=["small", "large"])
def test_contract_surface_sigma_calls(torch, sym, large, gemm_cfg):
    if large and sym.nirrep == 8:
        pytest.skip("large spaces are C2v only")
    V, Oc = spaces(sym, large)
    Z = O.ZORB
    for S in (1, 2):
        r1 = O.Desc([V, Oc, Z, Z], (1, 1), S)
        r2 = O.Desc([V, V, Oc, Oc], (0, 0), S)
        fab = O.Desc([V, V, Z, Z], (1, 1), 1)
        fij = O.Desc([Oc, Oc, Z, Z], (1, 1), 1)
        iabj = O.Desc([V, Oc, V, Oc], (1, 1), 1)
        abci = O.Desc([V, V, V, Oc], (0, 1), 1)
        ijka = O.Desc([Oc, Oc, Oc, V], (0, 1), 1)
        ijkl = O.Desc([Oc, Oc, Oc, Oc], (0, 0), 1)
        abcd = O.Desc([V, V, V, V], (0, 0), 1)
        _surface_case(torch, sym, "222", (fab, r1, r1), ("ae", "ei", "ai"), 1.0, 0)
        _surface_case(torch, sym, "222", (r1, fij, r1), ("am", "mi", "ai"), -1.0, 0)
        _surface_case(torch, sym, "422", (iabj, r1, r1), ("fnai", "fn", "ai"), 1.0, 0)
        _surface_case(torch, sym, "442", (abci, r2, r1), ("feam", "feim", "ai"), 1.0, 0)
        _surface_case(torch, sym, "442", (r2, ijka, r1), ("aemn", "mnie", "ai"), -1.0, 0)
        _surface_case(torch, sym, "424", (abci, r1, r2), ("abej", "ei", "abij"), 1.0, 2)
        _surface_case(torch, sym, "424", (ijka, r1, r2), ("ijmb", "am", "abij"), -1.0, 1)
        _surface_case(torch, sym, "444", (abcd, r2, r2), ("abef", "efij", "abij"), 1.0, 0)
        _surface_case(torch, sym, "444", (r2, ijkl, r2), ("abmn", "mnij", "abij"), 1.0, 0)
        _surface_case(torch, sym, "444", (r2, iabj, r2), ("aeim", "embj", "abij"), 1.0, 3)
    t2 = O.Desc([V, V, Oc, Oc], (0, 0), 1)
    fab = O.Desc([V, V, Z, Z], (1, 1), 1)
    fij = O.Desc([Oc, Oc, Z, Z], (1, 1), 1)
    iabj = O.Desc([V, Oc, V, Oc], (1, 1), 1)
    _surface_case(torch, sym, "442", (t2, t2, fab), ("afmn", "efmn", "ae"), -1.0, 0)
    _surface_case(torch, sym, "442", (t2, t2, fij), ("efmn", "efin", "mi"), 1.0, 0)
    _surface_case(torch, sym, "444", (t2, t2, iabj), ("fenm", "bfjn", "embj"), 1.0, 0)


def test_contract_surface_224(torch, sym):
    """mod_contract.f90:203-219 (vv2v_224): two 2-index tensors into a 4-index
    one, r1(a,i) r1(b,j) -> c2(ab,ij) with every antisymmetriser, and
    f(a,b) t(i,j) -> x(ab,ij) into a packed and an unpacked output"""
    V, Oc = spaces(sym, False)
    Z = O.ZORB
    for S1, S2 in ((1, 1), (2, 3)):
        S3 = O.xor_dirprd(sym.nirrep)[S1 - 1][S2 - 1]
        r1a = O.Desc([V, Oc, Z, Z], (1, 1), S1)
        r1b = O.Desc([V, Oc, Z, Z], (1, 1), S2)
        for anti, iop in ((0, 1), (1, 1), (2, 1), (3, 0), (3, 1)):
            c2 = O.Desc([V, V, Oc, Oc], (iop, iop), S3)
            _surface_case(torch, sym, "224", (r1a, r1b, c2), ("ai", "bj", "abij"), 0.5, anti)
        fab = O.Desc([V, V, Z, Z], (1, 1), S1)
        tij = O.Desc([Oc, Oc, Z, Z], (1, 1), S2)
        x = O.Desc([V, V, Oc, Oc], (1, 1), S3)
        _surface_case(torch, sym, "224", (fab, tij, x), ("ab", "ij", "abij"), -1.25, 0)


def test_engine_323(torch, sym):
    """vv2v_dpd_323 (libsrint/vv2v_dpd_323.f90) on the seven cases of its harness
    (vv2v_dpd_323_check.f90:98-280) against the oracle, and its error codes"""
    import dpdhip
    from test_oracle_engine import _cases323, _desc3
    V, Oc = spaces(sym, False)
    for case, (t1, p1, t2, p2, t3, p3, iop, f2) in enumerate(_cases323(V, Oc)):
        for s1, s2 in ((1, 1), (2, 3)):
            s3 = O.xor_dirprd(sym.nirrep)[s1 - 1][s2 - 1]
            rng = np.random.default_rng(case * 10 + s1)
            d1 = _desc3(p1, iop if t1 == "pqm" else 1, s1)
            d2 = O.desc2(p2[0], p2[1], 1, s2)
            d3 = _desc3(p3, iop if t3 == "pqr" else 1, s3)
            x1, x2, x3 = (O.random_dpd(sym, d, rng) for d in (d1, d2, d3))
            ref = x3.copy()
            O.vv2v_dpd_323(x1, d1, t1, x2, d2, t2, ref, d3, t3, 1.5, f2)
            g1, g2, g3 = gopt(torch, sym, d1, x1), gopt(torch, sym, d2, x2), gopt(torch, sym, d3, x3)
            dpdhip.vv2v_dpd_323(g1, t1, g2, t2, g3, t3, 1.5, f2)
            np.testing.assert_allclose(g3.numpy(), ref, rtol=RTOL, atol=ATOL, err_msg=f"case {case}")
    d1 = _desc3((V, Oc, Oc), 1, 1)
    d2 = O.desc2(Oc, V, 1, 1)
    d3 = _desc3((V, Oc, V), 1, 1)
    g1, g2, g3 = (gopt(torch, sym, d, np.zeros(d.length())) for d in (d1, d2, d3))
    with pytest.raises(dpdhip.DpdError):  # the contracted index kept in v3
        dpdhip.vv2v_dpd_323(g1, "pqm", g2, "mr", g3, "pqm", 1.0, 0.0)
    gp = gopt(torch, sym, O.desc2(Oc, V, 0, 1), np.zeros(O.desc2(Oc, V, 0, 1).length()))
    with pytest.raises(dpdhip.DpdError):  # v2 packed
        dpdhip.vv2v_dpd_323(g1, "pqm", gp, "mr", g3, "pqr", 1.0, 0.0)


def test_sort_kernels(torch, sym):
    import dpdhip
    c = ctx_for(sym.nirrep)
    V, Oc = spaces(sym)
    rng = np.random.default_rng(3)
    for irrepx in range(1, sym.nirrep + 1):
        d0 = O.Desc([V, V, Oc, Oc], (0, 0), irrepx)
        x = O.random_dpd(sym, d0, rng)
        for iexp in [(1, 0), (0, 1), (1, 1)]:
            ref = O.iexpso(x, V, V, 0, iexp[0], Oc, Oc, 0, iexp[1], irrepx)
            out = torch.zeros(len(ref) + 1, dtype=torch.float64, device="cuda:0")
            dpdhip.iexpso(c, dev(torch, x), out, V, V, 0, iexp[0], Oc, Oc, 0, iexp[1], irrepx)
            assert np.array_equal(out[: len(ref)].cpu().numpy(), ref)
        full = O.iexpso(x, V, V, 0, 1, Oc, Oc, 0, 0, irrepx)
        ref = O.icompso(full, V, V, 1, 0, Oc, Oc, 0, 1, irrepx)
        out = torch.zeros(len(ref) + 1, dtype=torch.float64, device="cuda:0")
        dpdhip.icompso(c, dev(torch, full), out, V, V, 1, 0, Oc, Oc, 0, 1, irrepx)
        assert np.array_equal(out[: len(ref)].cpu().numpy(), ref)
        d1 = O.Desc([V, Oc, V, Oc], (1, 1), irrepx)
        y = rng.standard_normal(d1.length())
        for icpq, icrs in [(1, 0), (0, 1), (1, 1)]:
            ref = O.isymtrso(y, V, Oc, 1, icpq, V, Oc, 1, icrs, irrepx)
            out = torch.zeros(len(y) + 1, dtype=torch.float64, device="cuda:0")
            dpdhip.isymtrso(c, dev(torch, y), out, V, Oc, 1, icpq, V, Oc, 1, icrs, irrepx)
            assert np.array_equal(out[: len(y)].cpu().numpy(), ref)
        ref = O.itranspso(y, V, Oc, 1, V, Oc, 1, irrepx)
        out = torch.zeros(len(y) + 1, dtype=torch.float64, device="cuda:0")
        dpdhip.itranspso(c, dev(torch, y), out, V, Oc, 1, V, Oc, 1, irrepx)
        assert np.array_equal(out[: len(y)].cpu().numpy(), ref)
        io = {0: (0, 1), 1: (1, 1), 2: (1, 1), 3: (1, 1), 4: (1, 0), 5: (1, 1)}
        for iop in range(6):
            di = O.Desc([V, V, Oc, Oc], io[iop], irrepx)
            z = O.random_dpd(sym, di, rng)
            ref = O.iassymso(z, irrepx, V, V, Oc, Oc, iop)
            out = torch.zeros(len(ref) + 1, dtype=torch.float64, device="cuda:0")
            dpdhip.iassymso(c, dev(torch, z), out, irrepx, V, V, Oc, Oc, iop)
            assert np.array_equal(out[: len(ref)].cpu().numpy(), ref), iop
        for t in ("3142", "2413", "4321", "1243"):
            ref = O.sstgenso(y, V, Oc, V, Oc, irrepx, t)
            out = torch.zeros(len(y) + 1, dtype=torch.float64, device="cuda:0")
            dpdhip.sstgenso(c, dev(torch, y), out, V, Oc, V, Oc, irrepx, t)
            assert np.array_equal(out[: len(y)].cpu().numpy(), ref)


def test_vector_primitives(torch):
    import dpdhip
    c = ctx_for(4)
    rng = np.random.default_rng(9)
    for n in (1, 7, 1000, 300001):
        x = rng.standard_normal(n)
        y = rng.standard_normal(n)
        gx, gy = dev(torch, x), dev(torch, y)
        assert dpdhip.dot(c, gx, gy) == pytest.approx(float(x @ y), rel=1e-12, abs=1e-12)
        assert dpdhip.nrm2(c, gx) == pytest.approx(float(np.linalg.norm(x)), rel=1e-12)
        # idamax: first index of the max |x|, 1-based
        x2 = x.copy()
        k = n // 2
        x2[k] = 10.0
        if n > 3:
            x2[n - 1] = -10.0
        i, v = dpdhip.amax(c, dev(torch, x2))
        assert i == k + 1 and v == 10.0
        dpdhip.axpy(c, -0.5, gx, gy)
        np.testing.assert_allclose(gy.cpu().numpy(), y - 0.5 * x, rtol=1e-15)
        dpdhip.scal(c, 3.0, gx)
        np.testing.assert_allclose(gx.cpu().numpy(), 3.0 * x, rtol=1e-15)
    # preconditioner with the 1e-4 clamp (mod_davidson.f90:417-428)
    d = np.array([1.0, 2.0, 0.50005, 0.49999, 3.0])
    e = np.array([1.0, 1.0, 1.0, 1.0, 2.0])
    p = torch.zeros(5, dtype=torch.float64, device="cuda:0")
    dpdhip.precondition(c, 0.5, dev(torch, d), dev(torch, e), p)
    exp = e / np.array([0.5 - 1.0, 0.5 - 2.0, -1e-4, 1e-4, 0.5 - 3.0])
    np.testing.assert_allclose(p.cpu().numpy(), exp, rtol=1e-15)
