"""Pins the oracle's sigma builders, intermediates, orbital-energy helpers and
spin-flip guess against an independent dense restatement (CPU only).

The dense side never calls the oracle's contraction code: it builds every
antisymmetrised spin-orbital integral <pq||rs> from the spatial ERIs of the
synthetic generator (pinned element by element in test_abi.py), works on
unpacked tensors with numpy einsum, and writes each term from the equations
the reference states in its source comments:

  SF-CIS (EE-CIS.SF/calAR.f90:6-34), the textbook CIS matrix
      s(a,i) = (e_a - e_i) r(a,i) + <aj||ib> r(b,j)
  CIS(Doo) (CISdoo.SD.SF/run.f90:31-49, calAR.f90:9-79)
      t(ab,ij)  = -<ab||ij> / (e_a + e_b - e_i - e_j)            (denomT.f90)
      Fac(a,e)  = d_ae e_a - 1/2 t(af,mn) <ef||mn>                (Fae_r2abij.f90)
      Fki(m,i)  = d_mi e_m + 1/2 t(ef,in) <ef||mn>                (Fmi_r2abij.f90)
      W(em,bj)  = <mb||ej> + t(bf,jn) <fe||nm>                    (Wmbej_r2abij.f90, x2.0)
      s(a,i)    = Fac(a,e) r(e,i) - r(a,m) Fki(m,i) + W(fn,ai) r(f,n)
                  + 1/2 <fe||am> r(fe,im) - 1/2 r(ae,mn) <mn||ie>
      s(ab,ij)  = e(ab,ij) r(ab,ij) + P(ij){<ab||ej> r(e,i)} - P(ab){r(a,m) <ij||mb>}
  SF-CPSd adds J_dd (EE-CPSd.SF/calC2_VR2.f90:3)
      P(ab)P(ij){r(ae,im) <mb||ej>} + 1/2 r(ab,mn) <mn||ij> + 1/2 <ab||ef> r(ef,ij)

Every column of the oracle's J (sigma of every unit vector) must equal the
dense J to 1e-12.  The element formulas of Eorb_diff1/2, denomT2 and
putDiagFock are checked through getidx, the spin-flip guess against a brute
force enumeration, and a test plants mistakes in the oracle (a flipped MP2
sign, swapped occupied/virtual energy offsets, a wrong factor on a packed
contracted pair, the wrong spin block in the guess) to show each is caught.
"""
import itertools
import os

import numpy as np
import pytest

import oracle as O
import oracle_sigma as OS
from dpdhip import Spec

@pytest.fixture(autouse=True, scope="module")
def _one_thread():
    # thousands of sigma calls on tiny tensors: thread start-up would dominate
    O.set_threads(1)
    yield
    O.set_threads(os.cpu_count() or 1)


SYSTEMS = {
    "c2v": dict(nso=(3, 1, 2, 2), nalpha=3, nbeta=2, seed=11),
    "c2v-b": dict(nso=(4, 2, 2, 3), nalpha=4, nbeta=2, seed=3),
    "d2h": dict(nso=(2, 1, 1, 2, 1, 1, 1, 1), nalpha=3, nbeta=1, seed=11),
}


def _spin_orbitals(lay, nso):
    """(irrep, spin, spatial index) of every occupied and virtual spin orbital
    in the DPD order: irreps ascending, alpha then beta inside an irrep
    (lib_cc/readmodule.f90:40-43)."""
    soff = np.concatenate([[0], np.cumsum(nso)])
    occ, vir = [], []
    for h in range(len(nso)):
        occ += [(h, 0, soff[h] + k) for k in range(lay["popa"][h])]
        occ += [(h, 1, soff[h] + k) for k in range(lay["popb"][h])]
        vir += [(h, 0, soff[h] + lay["popa"][h] + k) for k in range(lay["vrta"][h])]
        vir += [(h, 1, soff[h] + lay["popb"][h] + k) for k in range(lay["vrtb"][h])]
    return occ, vir


class Dense:
    """The system in unpacked spin-orbital form, built without the oracle."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.sp = Spec(cfg["nso"], cfg["nalpha"], cfg["nbeta"], cfg["seed"])
        lay = self.lay = self.sp.layout()
        occ, vir = _spin_orbitals(lay, cfg["nso"])
        self.occ, self.vir = occ, vir
        self.no, self.nv = len(occ), len(vir)
        orbs = occ + vir
        N = int(sum(cfg["nso"]))
        E = self.sp.eri(np.array(list(itertools.product(range(N), repeat=4)))).reshape(N, N, N, N)
        sp_idx = np.array([o[2] for o in orbs])
        spin = np.array([o[1] for o in orbs])
        # <pq|rs> = (pr|qs) d(s_p,s_r) d(s_q,s_s);  <pq||rs> = <pq|rs> - <pq|sr>
        phys = E[np.ix_(sp_idx, sp_idx, sp_idx, sp_idx)].transpose(0, 2, 1, 3)
        phys = phys * (spin[:, None, None, None] == spin[None, None, :, None]) \
                    * (spin[None, :, None, None] == spin[None, None, None, :])
        self.G = phys - phys.transpose(0, 1, 3, 2)
        eorb = self.sp.fill("eorb", 0, self.no + self.nv)
        self.eo, self.ev = eorb[: self.no], eorb[self.no:]
        self.irr_o = np.array([o[0] for o in occ])
        self.irr_v = np.array([v[0] for v in vir])
        self.spin_o = np.array([o[1] for o in occ])
        self.spin_v = np.array([v[1] for v in vir])

    def g(self, spaces):
        o, v = np.arange(self.no), self.no + np.arange(self.nv)
        return self.G[np.ix_(*[o if c == "o" else v for c in spaces])]

    # ---- intermediates of CIS(Doo) -------------------------------------
    def d2(self):
        ev, eo = self.ev, self.eo
        return ev[:, None, None, None] + ev[None, :, None, None] - eo[None, None, :, None] - eo[None, None, None, :]

    def mp2(self):
        return -self.g("vvoo") / self.d2()

    def intermediates(self):
        t, vvoo = self.mp2(), self.g("vvoo")
        fac = np.diag(self.ev) - 0.5 * np.einsum("afmn,efmn->ae", t, vvoo)
        fki = np.diag(self.eo) + 0.5 * np.einsum("efin,efmn->mi", t, vvoo)
        w = np.einsum("mbej->embj", self.g("ovvo")) + np.einsum("bfjn,fenm->embj", t, vvoo)
        return fac, fki, w

    # ---- sigma on a batch of unpacked vectors --------------------------
    def sigma(self, method, R1, R2):
        eo, ev = self.eo, self.ev
        if method == "sf-cis":
            s1 = (ev[:, None] - eo[None, :]) * R1 + np.einsum("ajib,xbj->xai", self.g("voov"), R1)
            return s1, None
        fac, fki, w = self.intermediates()
        vvvo, ooov = self.g("vvvo"), self.g("ooov")
        s1 = (np.einsum("ae,xei->xai", fac, R1) - np.einsum("xam,mi->xai", R1, fki)
              + np.einsum("fnai,xfn->xai", w, R1)
              + 0.5 * np.einsum("feam,xfeim->xai", vvvo, R2)
              - 0.5 * np.einsum("xaemn,mnie->xai", R2, ooov))
        x = np.einsum("abej,xei->xabij", vvvo, R1)
        y = np.einsum("xam,ijmb->xabij", R1, ooov)
        s2 = (self.d2() * R2 + (x - x.transpose(0, 1, 2, 4, 3)) - (y - y.transpose(0, 2, 1, 3, 4)))
        if method == "sf-cpsd":
            ring = np.einsum("xaeim,mbej->xabij", R2, self.g("ovvo"), optimize=True)
            s2 = s2 + ring - ring.transpose(0, 2, 1, 3, 4) - ring.transpose(0, 1, 2, 4, 3) \
                + ring.transpose(0, 2, 1, 4, 3)
            s2 = s2 + 0.5 * np.einsum("xabmn,mnij->xabij", R2, self.g("oooo"), optimize=True)
            s2 = s2 + 0.5 * np.einsum("abef,xefij->xabij", self.g("vvvv"), R2, optimize=True)
        return s1, s2


def _oracle(cfg, method, symR, cls=OS.OracleProblem, eorb=None):
    sp = Spec(cfg["nso"], cfg["nalpha"], cfg["nbeta"], cfg["seed"])
    lay = sp.layout()
    s = O.Sym(len(cfg["nso"]))
    s.activate()
    v, o = lay["popv"], lay["popo"]
    descs = {"iabj": O.Desc([v, o, v, o], (1, 1), 1), "ijka": O.Desc([o, o, o, v], (0, 1), 1),
             "ijkl": O.Desc([o, o, o, o], (0, 0), 1), "abij": O.Desc([v, v, o, o], (0, 0), 1),
             "abci": O.Desc([v, v, v, o], (0, 1), 1), "abcd": O.Desc([v, v, v, v], (0, 0), 1)}
    ints = {k: sp.fill(k, 0, d.length()) for k, d in descs.items()}
    e = sp.fill("eorb", 0, int(o.sum() + v.sum())) if eorb is None else eorb
    return s, cls(s, o, v, lay["popa"], lay["popb"], lay["vrta"], lay["vrtb"], e, ints, method, symR)


def _dense_J(D, s, op, method):
    """J column by column from the dense sigma of unpacked unit vectors"""
    d1 = op.dR1
    n1, dim = op.l1, op.dim
    I = np.eye(dim)
    R1 = np.stack([O.unpack(s, d1, I[k, :n1])[:, :, 0, 0] for k in range(dim)])
    R2 = None
    if method != "sf-cis":
        R2 = np.stack([O.unpack(s, op.dR2, I[k, n1:]) for k in range(dim)])
    s1, s2 = D.sigma(method, R1, R2)
    J = np.zeros((dim, dim))
    p, q, _, _ = O.dpd_index(s, d1)
    J[:n1, :] = s1[:, p, q].T
    # the operator keeps the symmetry of R: nothing lands outside the stored block
    mask1 = np.ones(s1.shape[1:], bool)
    mask1[p, q] = False
    assert np.abs(s1[:, mask1]).max(initial=0.0) < 1e-13
    if method != "sf-cis":
        P, Q, Rr, S = O.dpd_index(s, op.dR2)
        J[n1:, :] = s2[:, P, Q, Rr, S].T
        # antisymmetry of the doubles output
        np.testing.assert_allclose(s2, -s2.transpose(0, 2, 1, 3, 4), atol=1e-13)
        np.testing.assert_allclose(s2, -s2.transpose(0, 1, 2, 4, 3), atol=1e-13)
    return J


def _oracle_J(op):
    I = np.eye(op.dim)
    return np.stack([op.sigma(I[k]) for k in range(op.dim)], axis=1)


_dense_cache = {}


def dense(name):
    if name not in _dense_cache:
        _dense_cache[name] = Dense(SYSTEMS[name])
    return _dense_cache[name]


@pytest.mark.parametrize("symR", [1, 2])
@pytest.mark.parametrize("method", ["sf-cis", "cisdoo", "sf-cpsd"])
@pytest.mark.parametrize("name", ["c2v", "d2h"])
def test_oracle_sigma_equals_dense_restatement(name, method, symR):
    D = dense(name)
    s, op = _oracle(SYSTEMS[name], method, symR)
    Jo = _oracle_J(op)
    Jd = _dense_J(D, s, op, method)
    scale = np.abs(Jd).max()
    assert np.abs(Jo - Jd).max() <= 1e-12 * scale, np.abs(Jo - Jd).max()


def test_reductions_and_structure():
    """J_dd = 0 is CIS(Doo); CIS(Doo) with the MP2 amplitudes removed, restricted
    to the singles, is SF-CIS; the SF-CIS block is symmetric."""
    D = dense("c2v-b")
    s, cis = _oracle(SYSTEMS["c2v-b"], "sf-cis", 1)
    _, doo = _oracle(SYSTEMS["c2v-b"], "cisdoo", 1)
    _, cps = _oracle(SYSTEMS["c2v-b"], "sf-cpsd", 1)
    Jc, Jd, Jp = _oracle_J(cis), _oracle_J(doo), _oracle_J(cps)
    n1 = cis.l1
    np.testing.assert_allclose(Jc, Jc.T, atol=1e-14)
    # singles<-doubles and doubles<-singles blocks are the same in both methods
    np.testing.assert_array_equal(Jd[:, :n1], Jp[:, :n1])
    np.testing.assert_array_equal(Jd[:n1, :], Jp[:n1, :])
    # the difference is the doubles-doubles first-order coupling alone
    diff = Jp[n1:, n1:] - Jd[n1:, n1:]
    Jdd = _dense_J(D, s, cps, "sf-cpsd")[n1:, n1:] - _dense_J(D, s, doo, "cisdoo")[n1:, n1:]
    np.testing.assert_allclose(diff, Jdd, atol=1e-12 * np.abs(Jdd).max())
    # with t = 0 the CIS(Doo) singles block is the CIS matrix
    class NoMP2(OS.OracleProblem):
        def _mp2_amplitudes(self):
            return 0.0 * super()._mp2_amplitudes()
    _, z = _oracle(SYSTEMS["c2v-b"], "cisdoo", 1, cls=NoMP2)
    np.testing.assert_allclose(_oracle_J(z)[:n1, :n1], Jc, atol=1e-14)


@pytest.mark.parametrize("name", ["c2v", "d2h", "c2v-b"])
def test_mp2_amplitudes_and_energy(name):
    """t(ab,ij) = -<ab||ij>/e(ab,ij) element by element, and the MP2 energy
    1/4 <ij||ab> t(ab,ij) is negative (every denominator is), which fixes the sign."""
    D = dense(name)
    s, op = _oracle(SYSTEMS[name], "cisdoo", 1)
    t = O.unpack(s, op.d["abij"], op.mp2t.arr)
    tref = D.mp2()
    np.testing.assert_allclose(t, tref, rtol=1e-14, atol=1e-16)
    e2 = 0.25 * np.sum(np.einsum("ijab->abij", D.g("oovv")) * t)
    assert e2 < 0.0
    assert e2 == pytest.approx(0.25 * np.sum(D.g("vvoo") ** 2 / -D.d2()), rel=1e-12)
    # Fac, Fki, Wkaci unpacked equal the dense intermediates
    fac, fki, w = D.intermediates()
    np.testing.assert_allclose(O.unpack(s, op.d["fab"], op.Fac.arr)[:, :, 0, 0], fac, atol=1e-14)
    np.testing.assert_allclose(O.unpack(s, op.d["fij"], op.Fki.arr)[:, :, 0, 0], fki, atol=1e-14)
    np.testing.assert_allclose(O.unpack(s, op.d["iabj"], op.Wkaci.arr), w, atol=1e-14)


@pytest.mark.parametrize("name", ["c2v", "d2h"])
@pytest.mark.parametrize("symR", [1, 3])
def test_orbital_energy_elements_through_getidx(name, symR):
    """Eorb_diff1 (Eorb_diff.f90:6-63), Eorb_diff2 (:68-), denomT2
    (denomT.f90:94-249) and putDiagFock (putDiagFock.f90:6-55), element by
    element: getidx gives the irreps and positions of each stored element, the
    energies are read with offsets computed here (occupied irreps first, then
    virtual, Eorb_diff.f90:26-35)."""
    cfg = SYSTEMS[name]
    s, op = _oracle(cfg, "cisdoo", symR)
    lay = Spec(cfg["nso"], cfg["nalpha"], cfg["nbeta"], cfg["seed"]).layout()
    n = len(cfg["nso"])
    po, pv = lay["popo"][:n], lay["popv"][:n]
    offo = np.concatenate([[0], np.cumsum(po)])
    offv = po.sum() + np.concatenate([[0], np.cumsum(pv)])
    e = op.eorb
    ev = lambda h, x: e[offv[h - 1] + x - 1]
    eo = lambda h, x: e[offo[h - 1] + x - 1]
    for k in range(op.l1):
        h = O.getidx(k + 1, op.dR1)
        assert op.oed1[k] == ev(h[0], h[4]) - eo(h[1], h[5])
    abij = op.ints["abij"]
    for k in range(op.l2):
        h = O.getidx(k + 1, op.dR2)
        assert op.oed2[k] == pytest.approx(ev(h[0], h[4]) + ev(h[1], h[5]) - eo(h[2], h[6]) - eo(h[3], h[7]),
                                           rel=1e-15, abs=1e-15)
    dT = op.d["abij"]
    for k in range(dT.length()):
        h = O.getidx(k + 1, dT)
        den = eo(h[2], h[6]) + eo(h[3], h[7]) - ev(h[0], h[4]) - ev(h[1], h[5])
        assert op.mp2t.arr[k] == pytest.approx(abij[k] / den, rel=1e-14, abs=1e-300)
    # putDiagFock: f(p,p) = e(p), nothing else
    for name_, off, pop in (("fab", offv, pv), ("fij", offo, po)):
        f = np.zeros(op.d[name_].length())
        O.lib().orc_putDiagFock(O._dp(f), O._dp(e), O._lp(lay["popv"])[1], O._lp(lay["popo"])[1],
                                (b"V" if name_ == "fab" else b"O")[0], O.ctypes.c_double(1.0))
        F = O.unpack(s, op.d[name_], f)[:, :, 0, 0]
        ref = np.diag(np.concatenate([e[off[h]: off[h] + pop[h]] for h in range(n)]))
        np.testing.assert_array_equal(F, ref)


def _brute_guess(D, s, op, nroot):
    """all alpha-occupied -> beta-virtual excitations of symmetry symR, lowest
    e_a - e_i first, ties in storage order (preDavidson.f90:120-223)"""
    p, q, _, _ = O.dpd_index(s, op.dR1)
    cand = []
    for k in range(op.l1):
        a, i = p[k], q[k]
        if D.spin_v[a] == 1 and D.spin_o[i] == 0:
            cand.append((D.ev[a] - D.eo[i], k))
    cand.sort()
    return cand[:nroot]


@pytest.mark.parametrize("name", ["c2v", "d2h", "c2v-b"])
@pytest.mark.parametrize("symR", [1, 2])
def test_spin_flip_guess_brute_force(name, symR):
    D = dense(name)
    s, op = _oracle(SYSTEMS[name], "sf-cpsd", symR)
    ref = _brute_guess(D, s, op, 3)
    if len(ref) < 3:
        pytest.skip("fewer than 3 spin-flip excitations of this symmetry")
    R, E = op.guess(3)
    for r, (e, k) in enumerate(ref):
        assert np.flatnonzero(R[r]).tolist() == [k] and R[r, k] == 1.0
        assert E[r] == e


def test_planted_mistakes_are_caught():
    """Each plausible oracle mistake, planted here on the CPU, fails one of the
    checks above (the real code passes them)."""
    cfg = SYSTEMS["c2v"]
    D = dense("c2v")

    def mismatch(op, method):
        return np.abs(_oracle_J(op) - _dense_J(D, O.Sym(4), op, method)).max()

    s, good = _oracle(cfg, "sf-cpsd", 1)
    assert mismatch(good, "sf-cpsd") < 1e-12

    # 1. MP2 amplitudes with the wrong sign
    class FlippedMP2(OS.OracleProblem):
        def _mp2_amplitudes(self):
            return -super()._mp2_amplitudes()
    _, bad = _oracle(cfg, "sf-cpsd", 1, cls=FlippedMP2)
    assert mismatch(bad, "sf-cpsd") > 1e-6

    # 2. occupied and virtual energy offsets exchanged: the energies are read
    #    from the other block
    e = good.eorb
    no = int(good.popo.sum())
    _, bad = _oracle(cfg, "sf-cpsd", 1, eorb=np.concatenate([e[no:], e[:no]]))
    assert mismatch(bad, "sf-cpsd") > 1e-3

    # 3. the 1/2 of a packed contracted pair not dropped in the pp-ladder
    real = O.vv2v_444

    def half_ladder(v1, v2, v3, fact, t1, t2, t3, anti=0):
        return real(v1, v2, v3, 0.5 * fact if t1 == "abef" else fact, t1, t2, t3, anti)
    O.vv2v_444 = half_ladder
    try:
        assert mismatch(good, "sf-cpsd") > 1e-6
    finally:
        O.vv2v_444 = real

    # 4. the guess from the wrong spin block (alpha occupied -> alpha virtual:
    #    the beta virtuals' offset inside an irrep left out)
    s, op = _oracle(SYSTEMS["c2v-b"], "sf-cpsd", 1)
    ref = _brute_guess(dense("c2v-b"), s, op, 3)
    op.vrta = np.zeros_like(op.vrta)
    R, _ = op.guess(3)
    assert [int(np.flatnonzero(R[r])[0]) for r in range(3)] != [k for _, k in ref]
