"""Pins the oracle's grCPS(D) ground state (oracle_perturb.Perturb.gr_ground_state,
grCPSd.SF/run.f90) against an independent dense restatement (CPU only).

grCPS(D) takes the CIS(Doo) Jacobian J0 as zeroth order (grCPSd.SF/README,
solver.f90:1-4) and solves, order by order,

    -J0 T^(n) = C^(n)              (solver.f90 separates the orbital-energy diagonal)
    C^(1) = (0, <ab||ij>)                                           (run1.f90:18-19)
    C^(2) = (0, Jdd T2^(1))                                          (calA1R2.f90)
    C^(3) = (-D1 T1^(1) + L1[T^(1)],  Jdd T2^(2) + L2[T^(1)])         (calC1.f90, calC2.f90)
    E^(n+1) = 1/4 <ij||ab> (T2^(n) + tau^(n))                        (calEcorr.f90)
    tau^(2) = P(ij) t1(a,i) t1(b,j),  tau^(3) = P(ab)P(ij) t1^(2)(a,i) t1^(1)(b,j)  (effectT2.f90)

with Jdd the ring + hh + pp ladder, D1 the MP2 dressing of J0's singles block
(calA2R1 in calA1R1.f90) and L1/L2 the [[V,T^(1)],T^(1)] terms of calL1_VTT.f90
and calL2_VTT.f90, every term written here with einsum on the unpacked
antisymmetrised spin-orbital integrals of test_oracle_sigma_dense.Dense.  J0 is
the dense CIS(Doo) Jacobian (its columns pinned in test_oracle_sigma_dense);
the equations are solved with numpy.linalg.solve.  A planted mistake in the
oracle (a sign of a calL2_VTT term, the effectT2 type) must fail.
"""
import numpy as np
import pytest

import oracle as O
import oracle_perturb as OP
from test_oracle_sigma_dense import SYSTEMS, _dense_J, _oracle, dense

TIGHT = dict(diis_dim=8, diis_cyc=300, diis_max_err=1e-13)


def P_ab(X):
    return X - X.transpose(1, 0, 2, 3)


def P_ij(X):
    return X - X.transpose(0, 1, 3, 2)


class DenseGR:
    def __init__(self, D, s, op):
        self.D, self.s, self.op = D, s, op
        g = D.g
        self.vvoo, self.ooov, self.vvvo = g("vvoo"), g("ooov"), g("vvvo")
        self.ovvo, self.oooo, self.vvvv = g("ovvo"), g("oooo"), g("vvvv")
        self.J0 = _dense_J(D, s, op, "cisdoo")
        self.p1 = tuple(O.dpd_index(s, op.dR1)[:2])
        self.p2 = tuple(O.dpd_index(s, op.dR2))

    def pack(self, c1, c2):
        return np.concatenate([c1[self.p1], c2[self.p2]])

    def unpack(self, x):
        n1 = self.op.l1
        return (O.unpack(self.s, self.op.dR1, x[:n1])[:, :, 0, 0], O.unpack(self.s, self.op.dR2, x[n1:]))

    def jdd(self, t):
        ring = np.einsum("aeim,mbej->abij", t, self.ovvo, optimize=True)
        return (P_ij(P_ab(ring)) + 0.5 * np.einsum("abmn,mnij->abij", t, self.oooo, optimize=True)
                + 0.5 * np.einsum("abef,efij->abij", self.vvvv, t, optimize=True))

    def d1(self, r):
        """MP2 dressing of the CIS(Doo) singles block (J0 minus the SF-CIS block)"""
        t, v = self.D.mp2(), self.vvoo
        fvv = -0.5 * np.einsum("afmn,efmn->ae", t, v)
        foo = 0.5 * np.einsum("efin,efmn->mi", t, v)
        w = np.einsum("bfjn,fenm->embj", t, v)
        return np.einsum("ae,ei->ai", fvv, r) - np.einsum("am,mi->ai", r, foo) + np.einsum("fnai,fn->ai", w, r)

    def l1_vtt(self, r, t):
        ooov, vvvo, v = self.ooov, self.vvvo, self.vvoo
        gli = -(np.einsum("mnie,en->mi", ooov, r) + 0.5 * np.einsum("efin,efmn->mi", t, v))
        gad = -np.einsum("efam,fm->ae", vvvo, r) + 0.5 * np.einsum("afmn,efmn->ae", t, v)
        gld = np.einsum("efmn,fn->em", v, r)
        return (np.einsum("am,mi->ai", r, gli) - np.einsum("ae,ei->ai", gad, r)
                + np.einsum("aeim,em->ai", t, gld))

    def l2_vtt(self, r, t):
        ooov, vvvo, v, ovvo = self.ooov, self.vvvo, self.vvoo, self.ovvo

        def c2_r1ijka(w, fact):  # -fact P(ab) w(ij,mb) r(a,m)
            return -fact * P_ab(np.einsum("ijmb,am->abij", w, r))

        def c2_r1abci(w, fact):  # fact P(ij) w(ab,ej) r(e,i)
            return fact * P_ij(np.einsum("abej,ei->abij", w, r))

        c = c2_r1ijka(-P_ab(np.einsum("kbcj,ci->ijkb", ovvo, r)), -1.0)        # Wkbij_r1iabj (P on ij)
        c += c2_r1abci(np.einsum("abdc,cj->abdj", self.vvvv, r), 0.5)           # Wabdj_r1abcd
        c += c2_r1ijka(-np.einsum("klij,bl->ijkb", self.oooo, r), 0.5)         # Wkbij_r1Uklij
        foo = np.einsum("mnie,en->mi", ooov, r) + 0.5 * np.einsum("efin,efmn->mi", t, v)
        c += -P_ij(np.einsum("abim,mj->abij", t, foo))                          # c2_r2Fmi
        fvv = np.einsum("efam,fm->ae", vvvo, r) - 0.5 * np.einsum("afmn,efmn->ae", t, v)
        c += P_ab(np.einsum("aeij,be->abij", t, fvv))                           # c2_r2Fae
        w = (-np.einsum("nmje,bn->embj", ooov, r) + np.einsum("febm,fj->embj", vvvo, r)
             + 0.5 * np.einsum("fenm,bfjn->embj", v, t))
        c += P_ij(P_ab(np.einsum("aeim,embj->abij", t, w, optimize=True)))    # c2_r2Wmbej
        c += c2_r1ijka(-0.5 * np.einsum("cdij,cdbk->ijkb", t, vvvo), 1.0)       # Wkbij_r2abci
        wkl = 0.5 * P_ij(np.einsum("mnie,ej->mnij", ooov, r)) + 0.25 * np.einsum("efmn,efij->mnij", v, t)
        c += np.einsum("abmn,mnij->abij", t, wkl, optimize=True)                # c2_r2Wmnij (x2 on m<n)
        return c

    def solve(self, c1, c2):
        return self.unpack(np.linalg.solve(self.J0, -self.pack(c1, c2)))

    def ground_state(self):
        D = self.D
        z1 = np.zeros((D.nv, D.no))
        T1, T2, E = {}, {}, {}
        T1[1], T2[1] = self.solve(z1, self.vvoo)
        E[2] = 0.25 * np.sum(self.vvoo * T2[1])
        T1[2], T2[2] = self.solve(z1, self.jdd(T2[1]))
        tau = P_ij(np.einsum("ai,bj->abij", T1[1], T1[1]))
        E[3] = 0.25 * np.sum(self.vvoo * (T2[2] + tau))
        c1 = -self.d1(T1[1]) + self.l1_vtt(T1[1], T2[1])
        c2 = self.jdd(T2[2]) + self.l2_vtt(T1[1], T2[1])
        T1[3], T2[3] = self.solve(c1, c2)
        tau = P_ab(P_ij(np.einsum("ai,bj->abij", T1[2], T1[1])))
        E[4] = 0.25 * np.sum(self.vvoo * (T2[3] + tau))
        return dict(T1=T1, T2=T2, E=E)


_cache = {}


def _setup(name):
    if name not in _cache:
        D = dense(name)
        s, op = _oracle(SYSTEMS[name], "sf-cpsd", 1)
        _cache[name] = (D, s, op, DenseGR(D, s, op).ground_state())
    return _cache[name]


@pytest.mark.parametrize("name", ["c2v", "d2h", "c2v-b"])
def test_gr_ground_state_equals_dense(name):
    D, s, op, ref = _setup(name)
    pt = OP.Perturb(op, **TIGHT)
    gs = pt.gr_ground_state()
    for n in (1, 2, 3):
        np.testing.assert_allclose(O.unpack(s, pt.dT2, gs["T2"][n].arr), ref["T2"][n], atol=1e-11, err_msg=str(n))
        np.testing.assert_allclose(O.unpack(s, pt.dT1, gs["T1"][n].arr)[:, :, 0, 0], ref["T1"][n], atol=1e-11,
                                   err_msg=str(n))
        assert gs["iterations"][n] > 1
    for k in (2, 3, 4):
        assert gs["E"][k] == pytest.approx(ref["E"][k], rel=1e-9, abs=1e-14), k
    # the CIS(Doo) dressing couples T1 into order 1 (CPS(D) has T1^(1) = 0)
    assert np.abs(ref["T1"][1]).max() > 1e-6


def test_gr_planted_mistakes_are_caught(monkeypatch):
    name = "c2v"
    D, s, op, ref = _setup(name)

    def run():
        pt = OP.Perturb(op, **TIGHT)
        return pt, pt.gr_ground_state()

    def differs(pt, gs):
        for n in (1, 2, 3):
            if np.abs(O.unpack(s, pt.dT2, gs["T2"][n].arr) - ref["T2"][n]).max() > 1e-9:
                return True
            if np.abs(O.unpack(s, pt.dT1, gs["T1"][n].arr)[:, :, 0, 0] - ref["T1"][n]).max() > 1e-9:
                return True
        return any(abs(gs["E"][k] - ref["E"][k]) > 1e-12 for k in (2, 3, 4))

    assert not differs(*run())
    plants = [("effectT2", lambda f: (lambda t2, a, b, fact, kind: f(t2, a, b, fact, 3 - kind))),
              ("Wkbij_r1Uklij", lambda f: (lambda w, r1, u, fact: f(w, r1, u, -fact))),
              ("c1_r2Fme", lambda f: (lambda c1, r2, g, fact: f(c1, r2, g, -fact))),
              ("Wabdj_r1abcd", lambda f: (lambda w, r1, v, fact: f(w, r1, v, 2.0 * fact))),
              ("Fme_r1abij", lambda f: (lambda g, r1, v, fact: f(g, r1, v, -fact)))]
    for name_, plant in plants:
        real = getattr(OP, name_)
        monkeypatch.setattr(OP, name_, plant(real))
        assert differs(*run()), name_
        monkeypatch.setattr(OP, name_, real)
