"""Pins the oracle's DPD metadata and sort kernels.

The reference ships no golden vectors for these (its check harnesses use random
data from helpers missing from the tree), so each function is pinned by what the
mathematics fixes: closed forms for pair counts, brute-force enumeration of every
(p,q,r,s) tuple, an independent numpy restatement of the storage order, round
trips and dense-array identities.  A plausible mistake (wrong triangular offset,
swapped sub-block order, missing mirror sign) fails at least one of them.
"""
import itertools

import numpy as np
import pytest

import oracle as O

POPS_C2V = [np.array([3, 1, 2, 2, 0, 0, 0, 0]), np.array([2, 0, 1, 3, 0, 0, 0, 0])]
POPS_D2H = [np.array([2, 1, 1, 2, 1, 0, 1, 1]), np.array([1, 1, 0, 2, 1, 1, 0, 1])]


@pytest.fixture(params=[4, 8], ids=["c2v", "d2h"])
def sym(request):
    s = O.Sym(request.param)
    s.activate()
    return s


def _pops(sym):
    return POPS_C2V if sym.nirrep == 4 else POPS_D2H


def test_dirprd_is_xor_table():
    s = O.Sym(8)
    # D2h product table from the reference example output (C-spin-flip out.pre):
    # row 2 = 2 1 4 3 6 5 8 7
    assert list(s.dirprd[1]) == [2, 1, 4, 3, 6, 5, 8, 7]
    for a in range(1, 9):
        for b in range(1, 9):
            assert s.dp(a, b) == s.dp(b, a)
            assert s.dp(s.dp(a, b), b) == a


def test_irpdso_closed_forms(sym):
    pa, pb = _pops(sym)
    n = int(pa.sum())
    tot1 = sum(O.irpdso(x, pa, pa, 1) for x in range(1, sym.nirrep + 1))
    tot0 = sum(O.irpdso(x, pa, pa, 0) for x in range(1, sym.nirrep + 1))
    tot2 = sum(O.irpdso(x, pa, pa, 2) for x in range(1, sym.nirrep + 1))
    assert tot1 == n * n
    assert tot0 == n * (n - 1) // 2
    assert tot2 == n * (n + 1) // 2
    # totally symmetric irrep, iop=0: sum_h n_h (n_h - 1) / 2
    assert O.irpdso(1, pa, pa, 0) == sum(int(k) * (int(k) - 1) // 2 for k in pa)
    assert O.irpdso(3, pa, pb, -1) == pa[2]
    # rectangular pair
    tot = sum(O.irpdso(x, pa, pb, 1) for x in range(1, sym.nirrep + 1))
    assert tot == int(pa.sum()) * int(pb.sum())


def _brute_count(sym, pops, iop, irrepx):
    h = [O.irrep_of(p) for p in pops]
    cnt = 0
    for p, q, r, s in itertools.product(*[range(len(x)) for x in h]):
        if sym.dp(sym.dp(h[0][p], h[1][q]), sym.dp(h[2][r], h[3][s])) != irrepx:
            continue
        if iop[0] == 0 and not p < q:
            continue
        if iop[0] == 2 and not p <= q:
            continue
        if iop[1] == 0 and not r < s:
            continue
        if iop[1] == 2 and not r <= s:
            continue
        cnt += 1
    return cnt


@pytest.mark.parametrize("iop", [(0, 0), (1, 1), (0, 1), (1, 0), (2, 2), (2, 1)])
def test_idsymsoc_brute_force(sym, iop):
    pa, pb = _pops(sym)
    for irrepx in range(1, sym.nirrep + 1):
        pops = [pa, pa, pb, pb]
        assert O.idsymsoc(irrepx, pa, pa, iop[0], pb, pb, iop[1]) == _brute_count(sym, pops, iop, irrepx)


@pytest.mark.parametrize("iop", [(0, 0), (1, 1), (0, 1), (2, 2)])
def test_storage_order_matches_numpy_restatement(sym, iop):
    """orc_putidx (the C address map) agrees with the numpy storage order at every
    element, and getidx inverts it."""
    pa, pb = _pops(sym)
    for irrepx in (1, 2, sym.nirrep):
        d = O.Desc([pa, pa, pb, pb], iop, irrepx)
        P, Q, R, S = O.dpd_index(sym, d)
        assert len(P) == d.length()
        assert len(set(zip(P, Q, R, S))) == len(P)
        offs = [np.concatenate([[0], np.cumsum(p)]) for p in d.pops]
        irr = [O.irrep_of(p) for p in d.pops]
        for k in range(0, len(P), max(1, len(P) // 97)):
            g = (P[k], Q[k], R[k], S[k])
            info = [irr[j][g[j]] for j in range(4)] + [g[j] - offs[j][irr[j][g[j]] - 1] + 1 for j in range(4)]
            assert O.putidx(info, d) == k + 1
            assert list(O.getidx(k + 1, d)) == info


def test_isymoffso_triangular_offsets():
    s = O.Sym(4)
    s.activate()
    pa = np.array([4, 3, 2, 5, 0, 0, 0, 0])
    # sym=1: sub-blocks are the diagonal triangles in irrep order
    assert O.isymoffso(1, 1, pa, pa, 0) == 1
    assert O.isymoffso(2, 1, pa, pa, 0) == 1 + 6
    assert O.isymoffso(3, 1, pa, pa, 0) == 1 + 6 + 3
    assert O.isymoffso(4, 1, pa, pa, 2) == 1 + 10 + 6 + 3
    # sym=2 (irreps 1x2, 3x4): only hp<hq sub-blocks stored for iop 0
    assert O.isymoffso(2, 2, pa, pa, 0) == 1
    assert O.isymoffso(4, 2, pa, pa, 0) == 1 + 4 * 3
    assert O.isymoffso(4, 2, pa, pa, 1) == 1 + 3 * 4 + 4 * 3 + 5 * 2


def _rand(sym, d, seed=0):
    return O.random_dpd(sym, d, np.random.default_rng(seed))


def test_iexpso_matches_dense(sym):
    pa, pb = _pops(sym)
    for iexp in [(1, 0), (0, 1), (1, 1)]:
        d = O.Desc([pa, pa, pb, pb], (0, 0), 2)
        x = _rand(sym, d)
        y = O.iexpso(x, pa, pa, 0, iexp[0], pb, pb, 0, iexp[1], 2)
        dy = O.Desc([pa, pa, pb, pb], (1 if iexp[0] else 0, 1 if iexp[1] else 0), 2)
        assert np.array_equal(O.unpack(sym, d, x), O.unpack(sym, dy, y))
    # p<=q symmetric expansion
    d = O.Desc([pa, pa, pb, pb], (2, 1), 1)
    x = _rand(sym, d, 3)
    y = O.iexpso(x, pa, pa, 2, 1, pb, pb, 1, 0, 1)
    assert np.array_equal(O.unpack(sym, d, x), O.unpack(sym, O.Desc([pa, pa, pb, pb], (1, 1), 1), y))


def test_iexpso_error_cases(sym):
    pa, pb = _pops(sym)
    x = np.zeros(10000)
    with pytest.raises(O.OracleError):
        O.iexpso(x, pa, pb, 0, 1, pb, pb, 1, 0, 1)  # popp /= popq
    with pytest.raises(O.OracleError):
        O.iexpso(x, pa, pa, 0, 1, pb, pb, 1, 1, 1)  # expand a full side


def test_compress_expand_roundtrip(sym):
    pa, pb = _pops(sym)
    for irrepx in range(1, sym.nirrep + 1):
        d0 = O.Desc([pa, pa, pb, pb], (0, 1), irrepx)
        x = _rand(sym, d0, irrepx)
        full = O.iexpso(x, pa, pa, 0, 1, pb, pb, 1, 0, irrepx)
        back = O.icompso(full, pa, pa, 1, 0, pb, pb, 1, 1, irrepx)
        assert np.array_equal(back, x)
        d1 = O.Desc([pa, pa, pb, pb], (1, 0), irrepx)
        y = _rand(sym, d1, 7 + irrepx)
        fy = O.iexpso(y, pa, pa, 1, 0, pb, pb, 0, 1, irrepx)
        assert np.array_equal(O.icompso(fy, pa, pa, 1, 1, pb, pb, 1, 0, irrepx), y)


def test_symtr_transpose_sstgenso_dense(sym):
    pa, pb = _pops(sym)
    rng = np.random.default_rng(5)
    d = O.Desc([pa, pb, pa, pb], (1, 1), 3)
    x = rng.standard_normal(d.length())
    D = O.unpack(sym, d, x)
    # swap inside pq
    y = O.isymtrso(x, pa, pb, 1, 1, pa, pb, 1, 0, 3)
    assert np.array_equal(O.unpack(sym, O.Desc([pb, pa, pa, pb], (1, 1), 3), y), D.transpose(1, 0, 2, 3))
    # swap inside rs
    y = O.isymtrso(x, pa, pb, 1, 0, pa, pb, 1, 1, 3)
    assert np.array_equal(O.unpack(sym, O.Desc([pa, pb, pb, pa], (1, 1), 3), y), D.transpose(0, 1, 3, 2))
    # (pq,rs) -> (rs,pq)
    y = O.itranspso(x, pa, pb, 1, pa, pb, 1, 3)
    assert np.array_equal(O.unpack(sym, O.Desc([pa, pb, pa, pb], (1, 1), 3), y), D.transpose(2, 3, 0, 1))
    # involutions
    assert np.array_equal(O.isymtrso(O.isymtrso(x, pa, pb, 1, 1, pa, pb, 1, 1, 3), pb, pa, 1, 1, pb, pa, 1, 1, 3), x)
    # sstgenso: new index i takes old index type[i]
    pops = [pa, pb, pa, pb]
    for t in ["3142", "2143", "4321", "1243", "2413"]:
        perm = [int(c) - 1 for c in t]
        y = O.sstgenso(x, *pops, 3, t)
        dn = O.Desc([pops[k] for k in perm], (1, 1), 3)
        assert np.array_equal(O.unpack(sym, dn, y), D.transpose(perm))


def test_packed_iop0_has_antisymmetric_dense(sym):
    pa, pb = _pops(sym)
    d = O.Desc([pa, pa, pb, pb], (0, 0), 1)
    D = O.unpack(sym, d, _rand(sym, d))
    assert np.allclose(D, -D.transpose(1, 0, 2, 3))
    assert np.allclose(D, -D.transpose(0, 1, 3, 2))
    n = D.shape[0]
    assert np.all(D[np.arange(n), np.arange(n)] == 0)


@pytest.mark.parametrize("iop", range(6))
def test_iassymso_dense(sym, iop):
    """libsrint/iassymso.f:1-142 against dense arrays: out = in -/+ in with the
    rs pair (iop 0, 1), both pairs (2, 3) or the pq pair (4, 5) swapped."""
    pa, pb = _pops(sym)
    io = {0: (0, 1, 0, 0), 1: (1, 1, 1, 0), 2: (1, 1, 1, 1), 3: (1, 1, 1, 1), 4: (1, 0, 0, 0), 5: (1, 1, 0, 1)}[iop]
    for irrepx in range(1, sym.nirrep + 1):
        din = O.Desc([pa, pa, pb, pb], io[:2], irrepx)
        dout = O.Desc([pa, pa, pb, pb], io[2:], irrepx)
        x = _rand(sym, din, 3 + irrepx)
        D = O.unpack(sym, din, x)
        if iop in (0, 1):
            E = D - D.transpose(0, 1, 3, 2)
        elif iop == 2:
            E = D - D.transpose(1, 0, 3, 2)
        elif iop == 3:
            E = D + D.transpose(1, 0, 3, 2)
        else:
            E = D - D.transpose(1, 0, 2, 3)
        y = O.iassymso(x, irrepx, pa, pa, pb, pb, iop)
        assert len(y) == dout.length()
        assert np.allclose(O.unpack(sym, dout, y), E, rtol=0, atol=1e-14)
