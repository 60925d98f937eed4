"""CPU-side checks of the product library: it loads, exports every symbol that
include/dpd_hip.h declares, its host-side metadata agrees with the oracle, and
the seeded synthetic Hamiltonian generator is pinned against a dense numpy
brute force (no GPU needed: these entry points are host code)."""
import ctypes
import itertools
import os
import re

import numpy as np
import pytest

import oracle as O
from dpdhip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "dpd_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dpd_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    L = _lib.lib()
    syms = _header_symbols()
    assert len(syms) >= 40
    for s in syms:
        assert hasattr(L, s), s
    # the Python binding declares a signature for each of them
    assert set(syms) <= set(_lib.public_symbols()), set(syms) - set(_lib.public_symbols())


def _host_ctx(nirrep):
    h = ctypes.c_void_p()
    d = np.ascontiguousarray(O.xor_dirprd(nirrep).reshape(-1))
    _lib.check(_lib.lib().dpd_ctx_create(ctypes.byref(h), nirrep, d.ctypes.data_as(_lib.c_i64_p), -1))
    return h


def test_host_context_metadata_matches_oracle():
    for nirrep in (4, 8):
        h = _host_ctx(nirrep)
        s = O.Sym(nirrep)
        s.activate()
        rng = np.random.default_rng(nirrep)
        for _ in range(20):
            pa = np.zeros(8, np.int64)
            pb = np.zeros(8, np.int64)
            pa[:nirrep] = rng.integers(0, 5, nirrep)
            pb[:nirrep] = rng.integers(0, 5, nirrep)
            pap, pbp = pa.ctypes.data_as(_lib.c_i64_p), pb.ctypes.data_as(_lib.c_i64_p)
            for sym in range(1, nirrep + 1):
                for iop in (0, 1, 2):
                    assert _lib.lib().dpd_irpdso(h, sym, pap, pap, iop) == O.irpdso(sym, pa, pa, iop)
                    assert _lib.lib().dpd_idsymsoc(h, sym, pap, pap, iop, pbp, pbp, 1) == \
                        O.idsymsoc(sym, pa, pa, iop, pb, pb, 1)
                    for hq in range(1, nirrep + 1):
                        assert _lib.lib().dpd_isymoffso(h, hq, sym, pap, pap, iop) == O.isymoffso(hq, sym, pa, pa, iop)
        # getidx/putidx agree with the oracle on a packed tensor
        t = _lib.Tensor_()
        pa = np.array([3, 1, 2, 2, 1, 0, 1, 1][:8], np.int64)
        pb = np.array([2, 1, 0, 1, 1, 1, 0, 1][:8], np.int64)
        pa[nirrep:] = 0
        pb[nirrep:] = 0
        d = O.Desc([pa, pa, pb, pb], (0, 0), 2)
        for k in range(4):
            for hh in range(8):
                t.pop[k][hh] = int(d.pops[k][hh])
        t.iop[0], t.iop[1], t.sym = 0, 0, 2
        n = _lib.lib().dpd_tensor_len(h, ctypes.byref(t))
        assert n == d.length()
        info = np.zeros(8, np.int64)
        for idx in range(1, n + 1, max(1, n // 50)):
            _lib.check(_lib.lib().dpd_getidx(h, idx, info.ctypes.data_as(_lib.c_i64_p), ctypes.byref(t)))
            assert list(info) == list(O.getidx(idx, d))
            assert _lib.lib().dpd_putidx(h, info.ctypes.data_as(_lib.c_i64_p), ctypes.byref(t)) == idx
        # device work on a host-only context is refused with ENODEV
        x = ctypes.c_double()
        assert _lib.lib().dpd_dot(h, 1, None, None, ctypes.byref(x)) == -7
        _lib.lib().dpd_ctx_destroy(h)


def _fortran_symbols():
    txt = open(os.path.join(ROOT, "include", "dpd_fortran.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b([a-z0-9_]+_)\s*\(", txt)))


def test_library_exports_fortran_entry_points():
    """the reference's external names (lower case, trailing underscore) that
    include/dpd_fortran.h declares"""
    L = _lib.lib()
    syms = _fortran_symbols()
    assert {"vv2v_dpd_444_", "vv2v_dpd_424_", "vv2v_dpd_442_", "iexpso_", "icompso_", "isymtrso_", "itranspso_",
            "iassymso_", "sstgenso1_", "xgemmso_", "irpdso_", "idsymsoc_", "isymoffso_"} <= set(syms)
    for s in syms:
        assert hasattr(L, s), s


def test_fortran_metadata_by_reference_host_only():
    """irpdso_/idsymsoc_/isymoffso_ called the Fortran way (every argument by
    reference, int64) on a host-only set-up agree with the oracle"""
    L = _lib.lib()
    i64 = ctypes.c_int64
    ref = lambda v: ctypes.byref(i64(v))
    arr = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.int64)).ctypes.data_as(_lib.c_i64_p)
    for fn in ("irpdso_", "idsymsoc_", "isymoffso_"):
        getattr(L, fn).restype = i64
    dir_f = np.ascontiguousarray(O.xor_dirprd(4).reshape(8, 8).T.reshape(-1))  # column major
    L.dpd_fortran_init_(ref(4), dir_f.ctypes.data_as(_lib.c_i64_p), ref(-1))
    try:
        s = O.Sym(4)
        s.activate()
        pa = np.array([5, 1, 3, 2, 0, 0, 0, 0])
        pb = np.array([4, 2, 0, 3, 0, 0, 0, 0])
        for sym in range(1, 5):
            for iop in (0, 1, 2):
                pq = pb if iop == 1 else pa
                assert L.irpdso_(ref(sym), arr(pa), arr(pq), ref(iop)) == O.irpdso(sym, pa, pq, iop)
                for iop2 in (0, 1, 2):
                    rs = pb if iop2 == 1 else pa
                    assert L.idsymsoc_(ref(sym), arr(pa), arr(pq), ref(iop), arr(pa), arr(rs), ref(iop2)) == \
                        O.idsymsoc(sym, pa, pq, iop, pa, rs, iop2)
                for irr in range(1, 5):
                    assert L.isymoffso_(ref(irr), ref(sym), arr(pa), arr(pq), ref(iop)) == \
                        O.isymoffso(irr, sym, pa, pq, iop)
    finally:
        L.dpd_fortran_finalize_()


def test_error_paths_without_gpu():
    h = _host_ctx(4)
    t = _lib.Tensor_()
    # anti out of range is rejected before any device work
    assert _lib.lib().dpd_vv2v_444(h, ctypes.byref(t), ctypes.byref(t), ctypes.byref(t), 1.0, b"abef", b"efij",
                                   b"abij", 7) != 0
    _lib.lib().dpd_ctx_destroy(h)


# ---------------------------------------------------------------------------
# synthetic generator
# ---------------------------------------------------------------------------
from dpdhip import Spec  # noqa: E402


def test_occupations_largest_remainder():
    # 24 orbitals [11,2,4,7], n_alpha=6, n_beta=4 -> alpha [3,0,1,2], beta [2,0,1,1]
    lay = Spec((11, 2, 4, 7), 6, 4, 1).layout()
    assert list(lay["popa"][:4]) == [3, 0, 1, 2]
    assert list(lay["popb"][:4]) == [2, 0, 1, 1]
    assert list(lay["popo"][:4]) == [5, 0, 2, 3]
    assert list(lay["popv"][:4]) == [8 + 9, 4, 6, 5 + 6]
    # D2h 120 orbitals: 22 occupied / 218 virtual spin orbitals
    lay = Spec((30, 8, 14, 18, 26, 6, 10, 8), 12, 10, 2).layout()
    assert lay["popo"].sum() == 22 and lay["popv"].sum() == 218
    # C2v 180 orbitals: 30 / 330
    lay = Spec((70, 18, 36, 56), 16, 14, 3).layout()
    assert lay["popo"].sum() == 30 and lay["popv"].sum() == 330


def test_eri_symmetry_and_selection_rule():
    sp = Spec((3, 1, 2, 2), 3, 2, 7)
    n = 8
    irr = np.repeat(np.arange(4), (3, 1, 2, 2))
    idx = np.array(list(itertools.product(range(n), repeat=4)))
    v = sp.eri(idx).reshape(n, n, n, n)
    # 8-fold symmetry
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.array_equal(v, v.transpose(perm))
    # selection rule: zero unless the product is totally symmetric (diagonal
    # Coulomb (pp|qq) is always allowed)
    x = irr[:, None, None, None] ^ irr[None, :, None, None] ^ irr[None, None, :, None] ^ irr[None, None, None, :]
    assert np.all(v[x != 0] == 0.0)
    # Coulomb diagonal in [0.3, 0.8] plus a small random part
    d = np.array([v[p, p, q, q] for p in range(n) for q in range(n)])
    assert d.min() > 0.2 and d.max() < 0.9
    off = v[x == 0]
    assert 0.005 < np.std(off[np.abs(off) < 0.2]) < 0.05


def _spin_orbitals(lay, nso):
    """(space, irrep, spin, spatial global index) for every spin orbital, in the
    DPD order: per irrep alpha then beta."""
    soff = np.concatenate([[0], np.cumsum(nso)])
    occ, vir = [], []
    for h in range(len(nso)):
        occ += [(h, 0, soff[h] + k) for k in range(lay["popa"][h])]
        occ += [(h, 1, soff[h] + k) for k in range(lay["popb"][h])]
        vir += [(h, 0, soff[h] + lay["popa"][h] + k) for k in range(lay["vrta"][h])]
        vir += [(h, 1, soff[h] + lay["popb"][h] + k) for k in range(lay["vrtb"][h])]
    return occ, vir


@pytest.mark.parametrize("nso,na,nb", [((3, 1, 2, 2), 3, 2), ((2, 1, 1, 2, 1, 1, 1, 1), 3, 1)])
def test_antisymmetrized_integrals_brute_force(nso, na, nb):
    """<pq||rs> = (pr|qs) d d - (ps|qr) d d built densely in numpy from the
    spatial ERIs agrees with every stored DPD element of every tensor."""
    sp = Spec(nso, na, nb, 11)
    lay = sp.layout()
    nirrep = len(nso)
    s = O.Sym(nirrep)
    s.activate()
    occ, vir = _spin_orbitals(lay, nso)
    N = int(sum(nso))
    idx = np.array(list(itertools.product(range(N), repeat=4)))
    E = sp.eri(idx).reshape(N, N, N, N)

    def phys(p, q, r, t):
        if p[1] != r[1] or q[1] != t[1]:
            return 0.0
        return E[p[2], r[2], q[2], t[2]]

    def anti(p, q, r, t):
        return phys(p, q, r, t) - phys(p, q, t, r)

    o, v = lay["popo"], lay["popv"]
    space = {"o": occ, "v": vir}
    cases = {"ijkl": ("oooo", (0, 0)), "ijka": ("ooov", (0, 1)), "abij": ("vvoo", (0, 0)),
             "abci": ("vvvo", (0, 1)), "abcd": ("vvvv", (0, 0)), "iabj": ("vovo", (1, 1))}
    for name, (sps, iop) in cases.items():
        pops = [v if c == "v" else o for c in sps]
        d = O.Desc(pops, iop, 1)
        arr = sp.fill(name, 0, d.length())
        P, Q, R, S_ = O.dpd_index(s, d)
        sl = [space[c] for c in sps]
        for k in range(len(arr)):
            a, b, c, e = sl[0][P[k]], sl[1][Q[k]], sl[2][R[k]], sl[3][S_[k]]
            ref = anti(b, c, a, e) if name == "iabj" else anti(a, b, c, e)
            assert arr[k] == pytest.approx(ref, abs=1e-15), (name, k)
    # canonical Fock and orbital energies
    eorb = sp.fill("eorb", 0, int(o.sum() + v.sum()))
    no = int(o.sum())
    assert np.all((eorb[:no] >= -1.5) & (eorb[:no] <= -0.3))
    assert np.all((eorb[no:] >= 0.2) & (eorb[no:] <= 3.0))
    dfab = O.Desc([v, v, O.ZORB, O.ZORB], (1, 1), 1)
    fab = sp.fill("fab", 0, dfab.length())
    D = O.unpack(s, dfab, fab)[:, :, 0, 0]
    assert np.allclose(np.diag(D), eorb[no:]) and np.count_nonzero(D - np.diag(np.diag(D))) == 0


def test_subspace_eigensolver_matches_lapack():
    """dpd_small_eig (Hessenberg + Francis QR + inverse iteration) against
    numpy.linalg.eig (LAPACK dgeev, what mod_davidson.f90:186 calls)."""
    from dpdhip import small_eig
    rng = np.random.default_rng(0)
    for n in (1, 2, 3, 5, 8, 13, 20, 40):
        for trial in range(5):
            # Davidson-like: diagonally dominant, mildly nonsymmetric
            A = np.diag(np.sort(rng.uniform(-1, 3, n))) + 0.05 * rng.standard_normal((n, n))
            wr, wi, V = small_eig(A)
            ref = np.linalg.eigvals(A)
            assert np.allclose(np.sort_complex(wr + 1j * wi), np.sort_complex(ref), atol=1e-12)
            for k in range(n):
                if wi[k] == 0.0:
                    v = V[:, k]
                    assert abs(np.linalg.norm(v) - 1.0) < 1e-12
                    assert np.linalg.norm(A @ v - wr[k] * v) < 1e-10
    # general random matrices with complex pairs: eigenvalues still agree
    for n in (3, 4, 7, 12, 25, 40):
        A = rng.standard_normal((n, n))
        wr, wi, _ = small_eig(A)
        assert np.allclose(np.sort_complex(wr + 1j * wi), np.sort_complex(np.linalg.eigvals(A)), atol=1e-10)
    # the cyclic shift (eigenvalues: the n-th roots of unity) stalls the standard
    # double shift on its first steps; the exceptional shifts must get it moving
    for n in (3, 4, 6, 9):
        A = np.roll(np.eye(n), 1, axis=0)
        wr, wi, _ = small_eig(A)
        ref = np.exp(2j * np.pi * np.arange(n) / n)
        assert np.allclose(np.sort_complex(wr + 1j * wi), np.sort_complex(ref), atol=1e-10)
    # symmetric (real spectrum, eigenvectors orthogonal) and a repeated eigenvalue
    B = rng.standard_normal((10, 10))
    A = B + B.T
    wr, wi, V = small_eig(A)
    assert np.all(wi == 0.0) and np.allclose(np.sort(wr), np.linalg.eigvalsh(A), atol=1e-11)
    A = np.diag([1.0, 1.0, 2.0, 3.0]) + np.triu(0.1 * np.ones((4, 4)), 2)
    wr, wi, _ = small_eig(A)
    assert np.allclose(np.sort(wr), [1.0, 1.0, 2.0, 3.0], atol=1e-12) and np.all(wi == 0.0)
