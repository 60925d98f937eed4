"""Pins the oracle of the perturbation drivers (oracle/oracle_perturb.py)
against an independent dense restatement (CPU only).

The dense side never calls the oracle's contraction code: it works on the
unpacked antisymmetrised spin-orbital integrals of test_oracle_sigma_dense.Dense
with numpy einsum and dense linear solves, and writes every term from the
equations the reference states in its source comments:

  CPS(D) ground state (CPSd.SF/run1.f90, run2.f90, run3.f90)
      T2^(1) = <ab||ij> / (e_i + e_j - e_a - e_b),  E^(2) = 1/4 <ij||ab> t(ab,ij)
      C1^(n) = <a,i|[V,T2^(n-1)]|0> = 1/2 <ja||bc> t(bc,ji) - 1/2 <jk||bi> t(ba,jk)
      C2^(n) = <ab,ij|[V,T^(n-1)]|0>: ring, hh- and pp-ladder (+ P(ij)<ab||ej>t(e,i)
               - P(ab)t(a,m)<ij||mb> at n = 3, + the [[V,T2^(1)],T2^(1)] terms of
               calC2_VTT.f90 at n = 3)
      -e(ab,ij) T2^(n) = C2^(n),  -e(a,i) T1^(n) = <a,i|[V,T1^(n)]|0> + C1^(n)
      E^(n+1) = 1/4 <ij||ab> t(ab,ij)^(n)
  EE-CPS(D) for a CIS root (w0, L) (EE-CPSd.SF/run_order.f90, calE_C.f90)
      [J^(0) - w0] R^(n) = sum_{k=1}^n w^(k) R^(n-k) - C^(n),   R^(n) _|_ L
      C^(n) = sum_k J^(k) R^(n-k),  w^(n) = <L|C1^(n)>
  with J^(2) = <u|[[V,T^(1)],tau]|0> and J^(3) = <u|[[V,T^(2)],tau]|0> restricted to
  the intermediates calJ.f90 keeps (Foo, Fvv, Wovoo, Wovvo from T^(1); Foo, Fvv,
  Wovvo from T^(2)) and the <u|[[V,T^(1)],R]|0> terms calC_VTR.f90 keeps.
The DIIS solve is pinned by its closed form (coefficients B^-1 1 / 1^T B^-1 1)
and by converging on linear problems to the dense solution; planted mistakes
in the oracle must fail the comparisons.
"""
import numpy as np
import pytest

import oracle as O
import oracle_perturb as OP
import oracle_sigma as OS
from test_oracle_sigma_dense import SYSTEMS, _oracle, _oracle_J, dense

TIGHT = dict(diis_dim=8, diis_cyc=200, diis_max_err=1e-13)


def P_ab(X):
    return X - X.transpose(1, 0, 2, 3)


def P_ij(X):
    return X - X.transpose(0, 1, 3, 2)


class DenseCPS:
    """CPS(D) / EE-CPS(D) on unpacked spin-orbital tensors"""

    def __init__(self, D):
        self.D = D
        g = D.g
        self.vvoo, self.ooov, self.vvvo = g("vvoo"), g("ooov"), g("vvvo")
        self.ovvo, self.oooo, self.vvvv, self.voov = g("ovvo"), g("oooo"), g("vvvv"), g("voov")
        self.d1 = D.ev[:, None] - D.eo[None, :]
        self.d2 = D.d2()
        nv, no = D.nv, D.no
        # CIS matrix on every (a,i): (e_a - e_i) d + <aj||ib>
        self.jcis = np.diag(self.d1.reshape(-1)) + np.einsum("ajib->aibj", self.voov).reshape(nv * no, nv * no)

    def c1_vr2(self, t):
        return 0.5 * np.einsum("feam,feim->ai", self.vvvo, t) - 0.5 * np.einsum("aemn,mnie->ai", t, self.ooov)

    def jdd(self, t):
        ring = np.einsum("aeim,mbej->abij", t, self.ovvo, optimize=True)
        return (P_ij(P_ab(ring)) + 0.5 * np.einsum("abmn,mnij->abij", t, self.oooo, optimize=True)
                + 0.5 * np.einsum("abef,efij->abij", self.vvvv, t, optimize=True))

    def vr1(self, r):
        x = np.einsum("abej,ei->abij", self.vvvo, r)
        y = np.einsum("am,ijmb->abij", r, self.ooov)
        return P_ij(x) - P_ab(y)

    def solve_t1(self, C1):
        """(e_a - e_i + V) T1 = -C1"""
        return np.linalg.solve(self.jcis, -C1.reshape(-1)).reshape(C1.shape)

    def vtt(self, t):
        vvoo = self.vvoo
        fmi = 0.5 * np.einsum("efin,efmn->mi", t, vvoo)
        fae = -0.5 * np.einsum("afmn,efmn->ae", t, vvoo)
        w = 0.5 * np.einsum("bfjn,fenm->embj", t, vvoo)
        wkl = 0.25 * np.einsum("efmn,efij->mnij", vvoo, t)
        c = -P_ij(np.einsum("abim,mj->abij", t, fmi)) + P_ab(np.einsum("aeij,be->abij", t, fae))
        c += P_ij(P_ab(np.einsum("aeim,embj->abij", t, w, optimize=True)))
        c += np.einsum("abmn,mnij->abij", t, wkl, optimize=True)
        return c

    def ground_state(self):
        D, vvoo = self.D, self.vvoo
        T1, T2, E = {1: np.zeros_like(self.d1)}, {1: D.mp2()}, {}
        E[2] = 0.25 * np.sum(vvoo * T2[1])
        for n in (2, 3):
            C1 = self.c1_vr2(T2[n - 1])
            C2 = self.jdd(T2[n - 1])
            if n == 3:
                C2 = C2 + self.vr1(T1[2]) + self.vtt(T2[1])
            T2[n] = -C2 / self.d2
            T1[n] = self.solve_t1(C1)
            E[n + 1] = 0.25 * np.sum(vvoo * T2[n])
        return dict(T1=T1, T2=T2, E=E)

    def excited_state(self, gs, w0, L):
        vvoo, ooov, vvvo = self.vvoo, self.ooov, self.vvvo
        t1 = gs["T2"][1]
        # J^(2) intermediates (calJ2) from T^(1)
        j2oo = 0.5 * np.einsum("efin,efmn->mi", t1, vvoo)
        j2vv = -0.5 * np.einsum("afmn,efmn->ae", t1, vvoo)
        j2w = np.einsum("bfjn,fenm->embj", t1, vvoo)
        X = np.einsum("klic,bcjl->ijkb", ooov, t1)
        j2wovoo = X - X.transpose(1, 0, 2, 3) - 0.5 * np.einsum("cdij,cdbk->ijkb", t1, vvvo)
        # J^(3) intermediates (calJ3) from T^(2)
        s1, s2 = gs["T1"][2], gs["T2"][2]
        j3oo = np.einsum("mnie,en->mi", ooov, s1) + 0.5 * np.einsum("efin,efmn->mi", s2, vvoo)
        j3vv = np.einsum("efam,fm->ae", vvvo, s1) - 0.5 * np.einsum("afmn,efmn->ae", s2, vvoo)
        j3w = (-np.einsum("nmje,bn->embj", ooov, s1) + np.einsum("febm,fj->embj", vvvo, s1)
               + np.einsum("bfjn,fenm->embj", s2, vvoo))

        def singles(w, foo, fvv, r):
            return (np.einsum("fnai,fn->ai", w, r) - np.einsum("am,mi->ai", r, foo)
                    + np.einsum("ae,ei->ai", fvv, r))

        r0 = L
        w = {1: 0.0}
        # order 1: C^(1) = (0, [V,R1^(0)]);  R^(1) = (0, -C2^(1) / (e(ab,ij) - w0))
        C2_1 = self.vr1(r0)
        R2_1 = -C2_1 / (self.d2 - w0)
        # order 2
        C1_2 = self.c1_vr2(R2_1) + singles(j2w, j2oo, j2vv, r0)
        w[2] = float(np.sum(L * C1_2))
        C2_2 = self.jdd(R2_1) - P_ab(np.einsum("ijmb,am->abij", j2wovoo, r0))
        w1 = P_ij(np.einsum("mnie,ej->mnij", ooov, r0))
        C2_2 = C2_2 + 0.5 * np.einsum("abmn,mnij->abij", t1, w1)
        w2 = np.einsum("febm,fj->embj", vvvo, r0)
        C2_2 = C2_2 + P_ij(P_ab(np.einsum("aeim,embj->abij", t1, w2, optimize=True)))
        g = np.einsum("mnie,en->mi", ooov, r0)
        C2_2 = C2_2 - P_ij(np.einsum("abim,mj->abij", t1, g))
        gad = np.einsum("efam,fm->ae", vvvo, r0)
        C2_2 = C2_2 + P_ab(np.einsum("aeij,be->abij", t1, gad))
        # order 3: Ctilde^(2) = w^(2) R^(0) - C^(2)
        Ct1 = -C1_2 + w[2] * r0
        R2_2 = -C2_2 / (self.d2 - w0)
        Lf = L.reshape(-1)
        A = self.jcis - w0 * np.eye(Lf.size) + np.outer(Lf, Lf)
        R1_2 = np.linalg.solve(A, Ct1.reshape(-1)).reshape(Ct1.shape)
        C1_3 = self.c1_vr2(R2_2) + singles(j3w, j3oo, j3vv, r0)
        w[3] = float(np.sum(L * C1_3))
        return dict(w=w, R1_2=R1_2, R2_1=R2_1, R2_2=R2_2)


_gs_cache = {}


def _setup(name):
    if name not in _gs_cache:
        D = dense(name)
        s, op = _oracle(SYSTEMS[name], "sf-cpsd", 1)
        pt = OP.Perturb(op, **TIGHT)
        _gs_cache[name] = (D, s, op, pt, pt.ground_state(), DenseCPS(D).ground_state())
    return _gs_cache[name]


def test_dodiis_closed_form():
    rng = np.random.default_rng(0)
    for n in (1, 2, 5):
        X = rng.standard_normal((n, 7))
        B = X @ X.T + 0.1 * np.eye(n)
        c = OP.dodiis(B, n)
        u = np.linalg.solve(B, np.ones(n))
        np.testing.assert_allclose(c, u / u.sum(), rtol=1e-12)
        assert c.sum() == pytest.approx(1.0, abs=1e-13)


def test_diis_solves_linear_fixed_point():
    """x = M x + b (|M| < 1) iterated with dosodiis2: converges to the solution"""
    rng = np.random.default_rng(1)
    n = 30
    M = 0.4 * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    ref = np.linalg.solve(np.eye(n) - M, b)
    d = OP.Diis(6)
    x = b.copy()
    for it in range(1, 60):
        x, err = d.step(it, M @ x + b if it > 1 else x)
        if it > 1 and abs(err) < 1e-13:
            break
    np.testing.assert_allclose(x, ref, atol=1e-11)


@pytest.mark.parametrize("name", ["c2v", "d2h", "c2v-b"])
def test_ground_state_equals_dense(name):
    D, s, op, pt, gs, ref = _setup(name)
    for n in (1, 2, 3):
        np.testing.assert_allclose(O.unpack(s, pt.dT2, gs["T2"][n].arr), ref["T2"][n], atol=1e-12)
        np.testing.assert_allclose(O.unpack(s, pt.dT1, gs["T1"][n].arr)[:, :, 0, 0], ref["T1"][n], atol=1e-11)
    for k in (2, 3, 4):
        assert gs["E"][k] == pytest.approx(ref["E"][k], rel=1e-10, abs=1e-14), k
    assert gs["E"][2] < 0.0  # MP2 energy (every denominator is negative)


def _cis_root(name, symR):
    s, cis = _oracle(SYSTEMS[name], "sf-cis", symR)
    J = _oracle_J(cis)
    w, V = np.linalg.eigh(0.5 * (J + J.T))
    return s, cis, float(w[0]), V[:, 0]


@pytest.mark.parametrize("name,symR", [("c2v", 1), ("c2v", 2), ("d2h", 1), ("c2v-b", 3)])
def test_excited_state_equals_dense(name, symR):
    D, s, op, pt, gs, gref = _setup(name)
    s, cis, w0, L = _cis_root(name, symR)
    J = pt.intermediates(gs)
    got = pt.excited_state(gs, J, w0, L, symR)
    Ld = O.unpack(s, cis.dR1, L)[:, :, 0, 0]
    ref = DenseCPS(D).excited_state(gref, w0, Ld)
    assert got["w"][1] == 0.0
    for n in (2, 3):
        assert got["w"][n] == pytest.approx(ref["w"][n], rel=1e-9, abs=1e-13), n
    dR2 = O.Desc([op.popv, op.popv, op.popo, op.popo], (0, 0), symR)
    np.testing.assert_allclose(O.unpack(s, dR2, got["R"][1][1]), ref["R2_1"], atol=1e-12)
    np.testing.assert_allclose(O.unpack(s, dR2, got["R"][2][1]), ref["R2_2"], atol=1e-12)
    np.testing.assert_allclose(O.unpack(s, cis.dR1, got["R"][2][0])[:, :, 0, 0], ref["R1_2"], atol=1e-10)
    # R^(2) is orthogonal to L (intermediate normalisation)
    assert abs(float(np.dot(got["R"][2][0], L))) < 1e-12


def test_planted_mistakes_are_caught():
    """mistakes planted in the oracle on the CPU fail the dense comparisons"""
    name = "c2v"
    D = dense(name)
    s, op = _oracle(SYSTEMS[name], "sf-cpsd", 1)
    ref = DenseCPS(D).ground_state()
    s, cis, w0, L = _cis_root(name, 2)
    Ld = O.unpack(s, cis.dR1, L)[:, :, 0, 0]
    eref = DenseCPS(D).excited_state(ref, w0, Ld)

    def run():
        pt = OP.Perturb(op, **TIGHT)
        gs = pt.ground_state()
        ex = pt.excited_state(gs, pt.intermediates(gs), w0, L, 2)
        return gs, ex

    gs, ex = run()
    assert abs(gs["E"][4] - ref["E"][4]) < 1e-10 and abs(ex["w"][3] - eref["w"][3]) < 1e-10

    def caught(mod, name_, bad, what):
        real = getattr(mod, name_)
        setattr(mod, name_, bad(real))
        try:
            g2, e2 = run()
        except np.linalg.LinAlgError:
            return  # the iteration broke down: caught
        finally:
            setattr(mod, name_, real)
        diff = max(abs(g2["E"][k] - ref["E"][k]) for k in (2, 3, 4))
        diff = max(diff, abs(e2["w"][2] - eref["w"][2]), abs(e2["w"][3] - eref["w"][3]))
        assert not diff <= 1e-8, what  # (nan: caught as well)

    # the 2.0 on the hh-ladder of calC2_VTT.f90 taken as 1.0
    caught(OP, "c2_r2Wmnij", lambda f: (lambda c2, r2, w, fact: f(c2, r2, w, 1.0 if fact == 2.0 else fact)),
           "VTT hh factor")
    # the sign of the abci part of Wovoo flipped
    caught(OP, "Wkbij_r2abci", lambda f: (lambda w, r2, abci, fact: f(w, r2, abci, -fact)), "Wkbij abci sign")
    # the DIIS Lagrange row with the wrong sign (coefficients no longer sum to 1)
    caught(OP, "dodiis", lambda f: (lambda B, n: -f(B, n)), "DIIS constraint")
    # the sign of c2_r1ijka (-P(ab) W(ij,mb) r(a,m)) flipped
    caught(OP, "c2_r1ijka", lambda f: (lambda c2, r1, w, fact: f(c2, r1, w, -fact)), "sign of c2_r1ijka")
    # the orbital-energy denominator of T1 with the wrong sign (e_a - e_i)
    caught(OP.Perturb, "denomT1", lambda f: (lambda self, r1, sym: -f(self, r1, sym)), "denomT1 sign")
