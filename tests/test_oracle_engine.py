"""Pins the oracle's contraction engines and contract surface against dense numpy
einsum on unpacked tensors (brute force on small populations).

The case list follows the reference's engine harnesses
(libsrint/vv2v_dpd_444_check.f90:87-251: fast paths 3412*3456, 1234*3456,
1234*5634 with IOP=0 packing, the fallback 3142*3546, the scrambles
4132*5364->2165 and pmqn*rmsn->prqs, accumulation with fact2=0.5;
vv2v_dpd_424_check_d.f90 and vv2v_dpd_442_check.f90 patterns), plus every
contraction the sigma builders make (lib_cc/c1_*.f90, c2_*.f90).
A contracted packed pair (iop=0) is summed over p<q only, which is half the
dense sum of two antisymmetric factors -- that factor is written into the
expected value below.
"""
import numpy as np
import pytest

import oracle as O

PV = np.array([3, 1, 2, 2, 0, 0, 0, 0])
PO = np.array([2, 1, 0, 1, 0, 0, 0, 0])
PV8 = np.array([2, 1, 1, 2, 1, 0, 1, 1])
PO8 = np.array([1, 1, 0, 1, 1, 0, 0, 1])


@pytest.fixture(params=[4, 8], ids=["c2v", "d2h"])
def sym(request):
    s = O.Sym(request.param)
    s.activate()
    return s


def _spaces(sym):
    return (PV, PO) if sym.nirrep == 4 else (PV8, PO8)


def _expected(sym, d1, x1, t1, d2, x2, t2, d3, x3, t3, f1, f2, half):
    D1 = O.unpack(sym, d1, x1)
    D2 = O.unpack(sym, d2, x2)
    if len(t1) == 2:
        D1 = D1[:, :, 0, 0]
    if len(t2) == 2:
        D2 = D2[:, :, 0, 0]
    D3 = O.unpack(sym, d3, x3)
    E = np.einsum(f"{t1},{t2}->{t3}", D1, D2) * half
    if len(t3) == 2:
        E = E[:, :, None, None]
    return O.pack(sym, d3, f1 * E + f2 * D3)


def _run444(sym, pops, iops, syms, t1, t2, t3, f1=1.0, f2=0.0, half=1.0, seed=0):
    rng = np.random.default_rng(seed)
    d1 = O.Desc(pops[0], iops[0], syms[0])
    d2 = O.Desc(pops[1], iops[1], syms[1])
    d3 = O.Desc(pops[2], iops[2], syms[2])
    x1 = O.random_dpd(sym, d1, rng)
    x2 = O.random_dpd(sym, d2, rng)
    x3 = O.random_dpd(sym, d3, rng)
    exp = _expected(sym, d1, x1, t1, d2, x2, t2, d3, x3, t3, f1, f2, half)
    out = x3.copy()
    O.vv2v_dpd_444(x1, d1, t1, x2, d2, t2, out, d3, t3, f1, f2)
    np.testing.assert_allclose(out, exp, rtol=1e-12, atol=1e-12)
    # the general engine (fallback restatement) must agree with the fast path too
    out2 = x3.copy()
    O.contract_general(x1, d1, t1, x2, d2, t2, out2, d3, t3, f1, f2)
    np.testing.assert_allclose(out2, exp, rtol=1e-12, atol=1e-12)


def test_fast_paths_iop0(sym):
    V, Oc = _spaces(sym)
    # 1234*3456->1256: pp-ladder c(ab,ij) += V(ab,ef) r(ef,ij), ef packed
    _run444(sym, [[V, V, V, V], [V, V, Oc, Oc], [V, V, Oc, Oc]], [(0, 0), (0, 0), (0, 0)], (1, 2, 2),
            "abef", "efij", "abij", half=0.5)
    # 3412*3456->1256 (transposed first operand)
    _run444(sym, [[V, V, Oc, Oc], [V, V, Oc, Oc], [Oc, Oc, Oc, Oc]], [(0, 0), (0, 0), (0, 0)], (1, 1, 1),
            "efij", "efkl", "ijkl", half=0.5, f2=0.5)
    # 1234*5634->1256
    _run444(sym, [[V, V, Oc, Oc], [V, V, Oc, Oc], [V, V, V, V]], [(0, 0), (0, 0), (0, 0)], (3, 3, 1),
            "abmn", "cdmn", "abcd", half=0.5)
    # 3412*5634->1256
    _run444(sym, [[Oc, Oc, V, V], [V, V, Oc, Oc], [V, V, V, V]], [(1, 1), (1, 1), (1, 1)], (2, 2, 1),
            "mnab", "cdmn", "abcd", f2=0.5)


def test_fallback_and_scrambles(sym):
    V, Oc = _spaces(sym)
    full = [(1, 1)] * 3
    _run444(sym, [[V, V, Oc, Oc], [V, V, Oc, Oc], [V, V, Oc, Oc]], full, (1, 2, 2), "aeim", "ebmj", "abij")
    # '3142'*'3546'
    _run444(sym, [[V, V, Oc, Oc], [V, V, Oc, Oc], [V, V, Oc, Oc]], full, (2, 1, 2), "eaim", "ebmj", "abij", f2=0.5)
    # ultimate scramble 4132*5364->2165
    _run444(sym, [[Oc, V, V, Oc], [V, V, Oc, Oc], [Oc, V, Oc, V]], full, (1, 3, 3), "faeb", "iejf", "baji", f2=0.5)
    # pmqn*rmsn->prqs
    _run444(sym, [[V, Oc, V, Oc], [V, Oc, V, Oc], [V, V, V, V]], full, (2, 3, 4), "pmqn", "rmsn", "prqs")


def test_fallback_rejects_packed(sym):
    V, Oc = _spaces(sym)
    d1 = O.Desc([V, Oc, V, Oc], (1, 1), 1)
    d2 = O.Desc([V, V, Oc, Oc], (0, 0), 1)
    d3 = O.Desc([V, V, Oc, Oc], (0, 0), 1)
    with pytest.raises(O.OracleError):
        O.vv2v_dpd_444(np.zeros(d1.length() + 1), d1, "aeim", np.zeros(d2.length() + 1), d2, "ebmj",
                       np.zeros(d3.length() + 1), d3, "abij", 1.0, 0.0)


def test_contracted_population_mismatch(sym):
    V, Oc = _spaces(sym)
    d1 = O.Desc([V, V, V, V], (1, 1), 1)
    d2 = O.Desc([Oc, Oc, Oc, Oc], (1, 1), 1)
    d3 = O.Desc([V, V, Oc, Oc], (1, 1), 1)
    with pytest.raises(O.OracleError):
        O.vv2v_dpd_444(np.zeros(d1.length()), d1, "abef", np.zeros(d2.length()), d2, "efij",
                       np.zeros(d3.length()), d3, "abij", 1.0, 0.0)


def _run_surface(sym, kind, ops, t, fact, anti, half, seed=1):
    """Check a mod_contract call (expand/antisym/compress) against dense einsum."""
    rng = np.random.default_rng(seed)
    (d1, d2, d3) = ops
    x1 = O.random_dpd(sym, d1, rng)
    x2 = O.random_dpd(sym, d2, rng)
    x3 = O.random_dpd(sym, d3, rng)
    D1 = O.unpack(sym, d1, x1)
    D2 = O.unpack(sym, d2, x2)
    t1, t2, t3 = t
    if len(t1) == 2:
        D1 = D1[:, :, 0, 0]
    if len(t2) == 2:
        D2 = D2[:, :, 0, 0]
    E = np.einsum(f"{t1},{t2}->{t3}", D1, D2) * half * fact
    if anti in (1, 3):
        E = E - E.swapaxes(0, 1)
    if anti in (2, 3):
        E = E - E.swapaxes(2, 3)
    if len(t3) == 2:
        E = E[:, :, None, None]
    exp = O.pack(sym, d3, O.unpack(sym, d3, x3) + E)
    v1, v2, v3 = O.Opt(d1, x1), O.Opt(d2, x2), O.Opt(d3, x3.copy())
    if kind == "444":
        O.vv2v_444(v1, v2, v3, fact, t1, t2, t3, anti)
    elif kind == "424":
        O.vv2v_424(v1, v2, v3, fact, t1, t2, t3, anti)
    elif kind == "442":
        O.vv2v_442(v1, v2, v3, fact, t1, t2, t3)
    elif kind == "222":
        O.vv2v_222(v1, v2, v3, fact, t1, t2, t3)
    elif kind == "422":
        O.vv2v_422(v1, v2, v3, fact, t1, t2, t3)
    elif kind == "224":
        O.vv2v_224(v1, v2, v3, fact, t1, t2, t3, anti)
    np.testing.assert_allclose(v3.arr, exp, rtol=1e-12, atol=1e-12)


def test_contract_surface_224(sym):
    """mod_contract.f90:203-219: outer products of two 2-index tensors into a
    4-index one with P(pq)/P(rs) antisymmetrisers, packed and unpacked outputs"""
    V, Oc = _spaces(sym)
    Z = O.ZORB
    for S1, S2 in ((1, 1), (2, 3)):
        S3 = O.xor_dirprd(sym.nirrep)[S1 - 1][S2 - 1]
        r1a = O.Desc([V, Oc, Z, Z], (1, 1), S1)
        r1b = O.Desc([V, Oc, Z, Z], (1, 1), S2)
        for anti, iop in ((0, 1), (1, 1), (2, 1), (3, 0), (3, 1)):
            c2 = O.Desc([V, V, Oc, Oc], (iop, iop), S3)
            _run_surface(sym, "224", (r1a, r1b, c2), ("ai", "bj", "abij"), 0.5, anti, 1.0)


def test_sigma_contractions(sym):
    """Every engine call the sigma builders make, with the reference's factors."""
    V, Oc = _spaces(sym)
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
        abij = O.Desc([V, V, Oc, Oc], (0, 0), 1)
        # c1_r1Fae.f90:15  c(a,i) += F(a,e) r(e,i)
        _run_surface(sym, "222", (fab, r1, r1), ("ae", "ei", "ai"), 1.0, 0, 1.0)
        # c1_r1Fmi.f90:15  c(a,i) -= r(a,m) F(m,i)
        _run_surface(sym, "222", (r1, fij, r1), ("am", "mi", "ai"), -1.0, 0, 1.0)
        # c1_r1iabj.f90:15 c(a,i) += W(fn,ai) r(fn)
        _run_surface(sym, "422", (iabj, r1, r1), ("fnai", "fn", "ai"), 1.0, 0, 1.0)
        # c1_r2abci.f90:17 c(a,i) += <fe||am> r(fe,im), fe packed
        _run_surface(sym, "442", (abci, r2, r1), ("feam", "feim", "ai"), 1.0, 0, 0.5)
        # c1_r2ijka.f90:16 c(a,i) -= r(ae,mn) <mn||ie>, mn packed
        _run_surface(sym, "442", (r2, ijka, r1), ("aemn", "mnie", "ai"), -1.0, 0, 0.5)
        # c2_r1abci.f90:16 c(ab,ij) += P(ij) <ab||ej> r(e,i)
        _run_surface(sym, "424", (abci, r1, r2), ("abej", "ei", "abij"), 1.0, 2, 1.0)
        # c2_r1ijka.f90:15 c(ab,ij) -= P(ab) <ij||mb> r(a,m)
        _run_surface(sym, "424", (ijka, r1, r2), ("ijmb", "am", "abij"), -1.0, 1, 1.0)
        # c2_r2abcd.f90:19 pp-ladder
        _run_surface(sym, "444", (abcd, r2, r2), ("abef", "efij", "abij"), 1.0, 0, 0.5)
        # c2_r2Wmnij.f90:16 hh-ladder
        _run_surface(sym, "444", (r2, ijkl, r2), ("abmn", "mnij", "abij"), 1.0, 0, 0.5)
        # c2_r2Wmbej.f90:15 ring with P(ab)P(ij)
        _run_surface(sym, "444", (r2, iabj, r2), ("aeim", "embj", "abij"), 1.0, 3, 1.0)
    # intermediates (sym 1): Fae_r2abij, Fmi_r2abij, Wmbej_r2abij
    t2 = O.Desc([V, V, Oc, Oc], (0, 0), 1)
    fab = O.Desc([V, V, Z, Z], (1, 1), 1)
    fij = O.Desc([Oc, Oc, Z, Z], (1, 1), 1)
    iabj = O.Desc([V, Oc, V, Oc], (1, 1), 1)
    _run_surface(sym, "442", (t2, t2, fab), ("afmn", "efmn", "ae"), -1.0, 0, 0.5)
    _run_surface(sym, "442", (t2, t2, fij), ("efmn", "efin", "mi"), 1.0, 0, 0.5)
    _run_surface(sym, "444", (t2, t2, iabj), ("fenm", "bfjn", "embj"), 1.0, 0, 1.0)


def test_424_442_error_paths(sym):
    V, Oc = _spaces(sym)
    abci = O.Desc([V, V, V, Oc], (0, 1), 1)
    r1 = O.Desc([V, Oc, O.ZORB, O.ZORB], (1, 1), 2)
    r2 = O.Desc([V, V, Oc, Oc], (0, 0), 1)  # wrong symmetry for v3
    with pytest.raises(O.OracleError):
        O.vv2v_424(O.Opt(abci), O.Opt(r1), O.Opt(r2), 1.0, "abej", "ei", "abij", 2)
    with pytest.raises(O.OracleError):  # anti out of range
        O.vv2v_444(O.Opt(r2), O.Opt(r2), O.Opt(r2), 1.0, "abef", "efij", "abij", 4)


# libsrint/vv2v_dpd_323_check.f90:98-280, the seven cases of the harness:
# (v1 labels, v1 pops, v2 labels, v2 pops, v3 labels, v3 pops, pair iop, fact2)
def _cases323(V, Oc):
    P, Q, M, R = V, Oc, Oc, V
    return [
        ("pqm", (P, Q, M), "mr", (M, R), "pqr", (P, Q, R), 1, 0.0),   # standard
        ("pqm", (P, Q, M), "rm", (R, M), "pqr", (P, Q, R), 1, 0.0),   # swapped 2-index
        ("qpm", (Q, P, M), "rm", (R, M), "pqr", (P, Q, R), 1, 0.0),   # extreme scramble
        ("pqm", (P, Q, M), "mr", (M, R), "pqr", (P, Q, R), 1, 0.5),   # accumulation
        ("pqm", (P, P, M), "mr", (M, R), "pqr", (P, P, R), 0, 0.0),   # intact pair, iop 0
        ("pqm", (P, P, M), "mr", (M, R), "pqr", (P, P, R), 2, 0.0),   # intact pair, iop 2
        ("pqm", (P, Q, M), "mr", (M, R), "rpq", (R, P, Q), 1, 0.0),   # output scramble
    ]


def _desc3(pops, iop, sym):
    return O.Desc([pops[0], pops[1], pops[2], O.ZORB], (iop, 1), sym)


@pytest.mark.parametrize("case", range(7))
def test_323_harness_cases(sym, case):
    """vv2v_dpd_323 against dense einsum 'pqm,mr->pqr' for every harness case and
    two symmetry combinations"""
    V, Oc = _spaces(sym)
    t1, p1, t2, p2, t3, p3, iop, f2 = _cases323(V, Oc)[case]
    for s1, s2 in ((1, 1), (2, 3)):
        s3 = O.xor_dirprd(sym.nirrep)[s1 - 1][s2 - 1]
        rng = np.random.default_rng(case * 10 + s1)
        d1 = _desc3(p1, iop if t1 == "pqm" else 1, s1)
        d2 = O.desc2(p2[0], p2[1], 1, s2)
        d3 = _desc3(p3, iop if t3 == "pqr" else 1, s3)
        x1, x2, x3 = (O.random_dpd(sym, d, rng, antisym=True) for d in (d1, d2, d3))
        D1 = O.unpack(sym, d1, x1)[:, :, :, 0]
        D2 = O.unpack(sym, d2, x2)[:, :, 0, 0]
        D3 = O.unpack(sym, d3, x3)[:, :, :, 0]
        E = 1.5 * np.einsum(f"{t1},{t2}->{t3}", D1, D2) + f2 * D3
        exp = O.pack(sym, d3, E[:, :, :, None])
        out = x3.copy()
        O.vv2v_dpd_323(x1, d1, t1, x2, d2, t2, out, d3, t3, 1.5, f2)
        np.testing.assert_allclose(out, exp, rtol=1e-12, atol=1e-12)


def test_323_error_paths(sym):
    V, Oc = _spaces(sym)
    d1 = _desc3((V, Oc, Oc), 1, 1)
    d2 = O.desc2(Oc, V, 1, 1)
    d3 = _desc3((V, Oc, V), 1, 1)
    z = [np.zeros(d.length() + 1) for d in (d1, d2, d3)]
    with pytest.raises(O.OracleError):  # contracted index kept in v3
        O.vv2v_dpd_323(z[0], d1, "pqm", z[1], d2, "mr", z[2], d3, "pqm", 1.0, 0.0)
    with pytest.raises(O.OracleError):  # v2 packed
        O.vv2v_dpd_323(z[0], d1, "pqm", z[1], O.desc2(Oc, V, 0, 1), "mr", z[2], d3, "pqr", 1.0, 0.0)
    dp = _desc3((V, V, Oc), 0, 1)
    dq = _desc3((V, V, V), 0, 1)
    with pytest.raises(O.OracleError):  # reordered pair with iop 0
        O.vv2v_dpd_323(np.zeros(dp.length() + 1), dp, "pqm", z[1], O.desc2(Oc, V, 1, 1), "mr",
                       np.zeros(dq.length() + 1), dq, "qpr", 1.0, 0.0)
