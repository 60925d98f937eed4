"""Oracle of the perturbation drivers (TEST INFRASTRUCTURE ONLY).

Restates, on top of the C oracle's contract surface (oracle.py) and the
problem of oracle_sigma.py:
  * libsrint/dosodiis2_d.f:1-118 + dodiis.f:1-25   DIIS extrapolation (the
    history lives in memory instead of the DIISTMP2 / TMPAMPTUDE files)
  * CPSd.SF/run1.f90, run2.f90, run3.f90 (+ calC1.f90, calC2.f90, calC2_VR.f90,
    calC2_VTT.f90, calC1_VR2.f90, calC2_VR1.f90, calC2_VR2.f90, calEcorr.f90)
    ground-state CPS(D): T^(1), T^(2), T^(3) and E^(2), E^(3), E^(4)
  * EE-CPSd.SF/run.f90, run_order.f90, calJ.f90 (calJ2, calJ3), calE_C.f90,
    calV_C.f90, calC_JR.f90, calC_VTR.f90, calC1_VR2.f90, calC2_VR1.f90,
    calC2_VR2.f90, solver.f90, orthogonalR2L.f90: EE-CPS(D) excitation
    energies order by order, w = w^(0) + w^(2) + w^(3) per CIS root
  * grCPSd.SF/run.f90, run1-3.f90, solver.f90, calA0R1.f90, calA0R2.f90,
    calA1R1.f90 (calA2R1), calA1R2.f90, calC1.f90, calC2.f90, calL1_VTT.f90,
    calL2_VTT.f90, calEcorr.f90 (+ lib_cc/effectT2.f90): grCPS(D) ground state,
    the CIS(Doo) Jacobian as zeroth order, T^(1..3) and E^(2..4)
  * the named contractions they call (lib_cc/*.f90, one function each).
Only tests/ import this module.
"""
from __future__ import annotations

import numpy as np

import oracle as O
from oracle_sigma import OracleProblem

Z = O.ZORB


# ---------------------------------------------------------------------------
# DIIS (libsrint/dosodiis2_d.f, dodiis.f)
# ---------------------------------------------------------------------------
def dodiis(B, n):
    """dodiis.f:1-25: coefficients c (sum 1) from the error matrix B(n,n); the
    Lagrange row and column carry -1e-4 (the reference's preconditioner)"""
    A = np.zeros((n + 1, n + 1))
    A[1:, 1:] = B[:n, :n]
    A[0, 1:] = -1.0e-4
    A[1:, 0] = -1.0e-4
    b = np.zeros(n + 1)
    b[0] = -1.0e-4
    return np.linalg.solve(A, b)[1:]


class Diis:
    """dosodiis2_d.f: call step(iter, x) once per iteration with the new
    vector; returns the extrapolated vector and the signed largest element of
    the extrapolated residual (1.0 on the first call)"""

    def __init__(self, maxdim):
        self.maxdim = int(maxdim)
        self.err = [None] * self.maxdim  # records 1..maxdim
        self.vec = [None] * self.maxdim  # records maxdim+1..2*maxdim
        self.prev = None                 # unit 21 (TMPAMPTUDE)
        self.ndimr = 1
        self.idimr = 1

    def step(self, it, x):
        x = np.array(x, dtype=np.float64)
        if it == 1:
            self.idimr = 1
            self.ndimr = 1
            self.prev = x.copy()
            return x, 1.0
        k = self.idimr - 1
        self.vec[k] = self.prev.copy()
        e = x - self.prev
        self.err[k] = e.copy()
        n = self.ndimr
        B = np.zeros((n, n))
        for i in range(n):
            for j in range(i + 1):
                B[i, j] = B[j, i] = float(np.dot(self.err[i], self.err[j]))
        c = dodiis(B, n)
        out = np.zeros_like(x)
        for i in range(n):
            out += c[i] * self.err[i]
        resm = float(out[int(np.argmax(np.abs(out)))]) if out.size else 0.0
        for i in range(n):
            out += c[i] * self.vec[i]
        self.prev = out.copy()
        self.ndimr = min(self.maxdim, n + 1)
        self.idimr += 1
        if self.idimr > self.maxdim:
            self.idimr = 1
        return out, resm


# ---------------------------------------------------------------------------
# named contractions (lib_cc/*.f90); packed (iop 0) pairs as the drivers use
# ---------------------------------------------------------------------------
def c1_r2abci(c1, r2, abci, fact):      # c1_r2abci.f90
    O.vv2v_442(abci, r2, c1, fact if r2.desc.iop[0] == 0 else 0.5 * fact, "feam", "feim", "ai")


def c1_r2ijka(c1, r2, ijka, fact):      # c1_r2ijka.f90
    O.vv2v_442(r2, ijka, c1, -fact if r2.desc.iop[1] == 0 else -0.5 * fact, "aemn", "mnie", "ai")


def c2_r2Wmbej(c2, r2, w, fact):        # c2_r2Wmbej.f90
    O.vv2v_444(r2, w, c2, fact, "aeim", "embj", "abij", 3)


def c2_r2Wmnij(c2, r2, w, fact):        # c2_r2Wmnij.f90
    O.vv2v_444(r2, w, c2, fact if r2.desc.iop[0] == 0 else 0.5 * fact, "abmn", "mnij", "abij", 0)


def c2_r2abcd(c2, r2, abcd, fact):      # c2_r2abcd.f90
    O.vv2v_444(abcd, r2, c2, 0.5 * fact if r2.desc.iop[0] == 1 else fact, "abef", "efij", "abij", 0)


def c1_r1iabj(c1, r1, w, fact):         # c1_r1iabj.f90
    O.vv2v_422(w, r1, c1, fact, "fnai", "fn", "ai")


def c2_r1ijka(c2, r1, w, fact):         # c2_r1ijka.f90
    O.vv2v_424(w, r1, c2, -fact, "ijmb", "am", "abij", 1)


def c2_r1abci(c2, r1, abci, fact):      # c2_r1abci.f90
    O.vv2v_424(abci, r1, c2, fact, "abej", "ei", "abij", 2)


def c1_r1Fmi(c1, r1, f, fact):          # c1_r1Fmi.f90
    O.vv2v_222(r1, f, c1, -fact, "am", "mi", "ai")


def c1_r1Fae(c1, r1, f, fact):          # c1_r1Fae.f90
    O.vv2v_222(f, r1, c1, fact, "ae", "ei", "ai")


def c2_r2Fmi(c2, r2, f, fact):          # c2_r2Fmi.f90
    O.vv2v_424(r2, f, c2, -fact, "abim", "mj", "abij", 2)


def c2_r2Fae(c2, r2, f, fact):          # c2_r2Fae.f90
    O.vv2v_424(r2, f, c2, fact, "aeij", "be", "abij", 1)


def Fmi_r2abij(f, r2, abij, fact):      # Fmi_r2abij.f90
    O.vv2v_442(abij, r2, f, 0.5 * fact if r2.desc.iop[0] == 1 else fact, "efmn", "efin", "mi")


def Fae_r2abij(f, r2, abij, fact):      # Fae_r2abij.f90
    O.vv2v_442(r2, abij, f, -fact if r2.desc.iop[1] == 0 else -0.5 * fact, "afmn", "efmn", "ae")


def Wmbej_r2abij(w, r2, abij, fact):    # Wmbej_r2abij.f90
    O.vv2v_444(abij, r2, w, 0.5 * fact, "fenm", "bfjn", "embj", 0)


def Wmnij_r2abij(w, r2, abij, fact):    # Wmnij_r2abij.f90
    O.vv2v_444(abij, r2, w, fact if r2.desc.iop[0] == 0 else 0.5 * fact, "efmn", "efij", "mnij", 0)


def Fmi_r1ijka(f, r1, ijka, fact):      # Fmi_r1ijka.f90
    O.vv2v_422(ijka, r1, f, fact, "mnie", "en", "mi")


def Fae_r1abci(f, r1, abci, fact):      # Fae_r1abci.f90
    O.vv2v_422(abci, r1, f, fact, "efam", "fm", "ae")


def Wkbij_r2ijka(w, r2, ijka, fact):    # Wkbij_r2ijka.f90
    O.vv2v_444(ijka, r2, w, fact, "klic", "bcjl", "ijkb", 1)


def Wkbij_r2abci(w, r2, abci, fact):    # Wkbij_r2abci.f90: into (ij,bk), pair swap, add
    d = w.desc
    tmp = O.Opt(O.Desc([d.pops[0], d.pops[1], d.pops[3], d.pops[2]], (d.iop[0], 1), d.sym))
    O.vv2v_444(r2, abci, tmp, -fact if r2.desc.iop[0] == 0 else -0.5 * fact, "cdij", "cdbk", "ijbk", 0)
    sw = O.isymtrso(tmp.arr, d.pops[0], d.pops[1], d.iop[0], 0, d.pops[3], d.pops[2], 1, 1, d.sym)
    w.arr += sw[: w.arr.size]


def Wmbej_r1ijka(w, r1, ijka, fact):    # Wmbej_r1ijka.f90
    O.vv2v_424(ijka, r1, w, -fact, "nmje", "bn", "embj", 0)


def Wmbej_r1abci(w, r1, abci, fact):    # Wmbej_r1abci.f90
    O.vv2v_424(abci, r1, w, fact, "febm", "fj", "embj", 0)


def Wmnij_r1ijka(w, r1, ijka, fact):    # Wmnij_r1ijka.f90
    O.vv2v_424(ijka, r1, w, fact, "mnie", "ej", "mnij", 2)


def Wkbij_r1iabj(w, r1, iabj, fact):    # Wkbij_r1iabj.f90
    O.vv2v_424(iabj, r1, w, -fact, "ckbj", "ci", "ijkb", 1)


def Wabdj_r1abcd(w, r1, abcd, fact):    # Wabdj_r1abcd.f90 (both in core)
    O.vv2v_424(abcd, r1, w, fact, "abdc", "cj", "abdj", 0)


def Wkbij_r1Uklij(w, r1, u, fact):      # Wkbij_r1Uklij.f90
    O.vv2v_424(u, r1, w, -fact, "klij", "bl", "ijkb", 0)


def Fme_r1abij(f, r1, abij, fact):      # Fme_r1abij.f90
    O.vv2v_422(abij, r1, f, fact, "efmn", "fn", "em")


def c1_r2Fme(c1, r2, f, fact):          # c1_r2Fme.f90
    O.vv2v_422(r2, f, c1, fact, "aeim", "em", "ai")


def effectT2(t2, t1a, t1b, fact, kind):  # effectT2.f90 (p_effectT2 types 1 and 2)
    """t2(ab,ij) += fact * [t1a(a,i) t1b(b,j)] antisymmetrised in ij (type 1,
    t1a = t1b) or in ab and ij (type 2), kept in t2's packing"""
    O.vv2v_224(t1a, t1b, t2, fact, "ai", "bj", "abij", 2 if kind == 1 else 3)


# ---------------------------------------------------------------------------
# CPS(D) ground state (CPSd.SF)
# ---------------------------------------------------------------------------
class Perturb:
    """Ground-state CPS(D) and EE-CPS(D) on an oracle problem (method sf-cpsd
    or cisdoo: the integrals and orbital energies are what is used)."""

    def __init__(self, op: OracleProblem, diis_dim=10, diis_cyc=50, diis_max_err=1.0e-7):
        self.op = op
        self.diis_dim, self.diis_cyc, self.diis_max_err = diis_dim, diis_cyc, diis_max_err
        v, o = op.popv, op.popo
        self.v, self.o = v, o
        self.dT1 = O.Desc([v, o, Z, Z], (1, 1), 1)
        self.dT2 = O.Desc([v, v, o, o], (0, 0), 1)
        self.lib = O.lib()

    def _opt(self, name):
        return self.op.opt(name)

    def denomT1(self, r1: O.Opt, sym):
        """denomT.f90:30-76: t(a,i) = r(a,i) / (e_i - e_a)"""
        d = np.zeros(max(r1.desc.length(), 1))
        self.lib.orc_eorb_diff1(O._dp(d), O._dp(self.op.eorb), O._lp(self.v)[1], O._lp(self.o)[1], O.i64(sym))
        return r1.arr / (-d[: r1.desc.length()])

    def denomT2(self, r2: O.Opt, sym):
        """denomT.f90:94-249: t(ab,ij) = r(ab,ij) / (e_i + e_j - e_a - e_b)"""
        t = np.zeros(max(r2.desc.length(), 1))
        self.lib.orc_denomT2(O._dp(t), O._dp(np.ascontiguousarray(r2.arr)), O._dp(self.op.eorb),
                             O._lp(self.v)[1], O._lp(self.o)[1], O.i64(sym), O.i64(r2.desc.iop[0]))
        return t[: r2.desc.length()]

    def oed2(self, sym):
        """Eorb_diff2: e(ab,ij) = e_a + e_b - e_i - e_j in the R2 layout"""
        d = O.Desc([self.v, self.v, self.o, self.o], (0, 0), sym)
        x = np.zeros(max(d.length(), 1))
        self.lib.orc_eorb_diff2(O._dp(x), O._dp(self.op.eorb), O._lp(self.v)[1], O._lp(self.o)[1], O.i64(sym),
                                O.i64(0))
        return x[: d.length()]

    def _t1_diis(self, C1):
        """CPSd.SF/run2.f90:43-71 (run3.f90 the same): -e(a,i) T1 = V T1 + C1 by
        Jacobi steps with DIIS; returns T1 and the number of iterations"""
        diis = Diis(self.diis_dim)
        R1 = O.Opt(self.dT1, C1.arr.copy())
        T1 = O.Opt(self.dT1)
        it = 0
        while True:
            it += 1
            T1.arr = self.denomT1(R1, 1)
            T1.arr, err = diis.step(it, T1.arr)
            if it > 1 and abs(err) < self.diis_max_err:
                break
            if it > self.diis_cyc:
                break
            R1 = O.Opt(self.dT1, C1.arr.copy())
            c1_r1iabj(R1, T1, self._opt("iabj"), 1.0)
        return T1, it

    def ground_state(self):
        """CPSd.SF/run.f90: orders 1-3.  Returns T1/T2 per order (index 1..3), the
        correlation energies E2, E3, E4 and the DIIS iteration counts."""
        abij = self._opt("abij")
        T1 = {1: O.Opt(self.dT1)}
        # run1.f90: T1^(1) = 0, T2^(1) = <ab||ij>/e(ij,ab), E^(2) = <T2^(1)|abij>
        T2 = {1: O.Opt(self.dT2, self.denomT2(abij, 1))}
        E = {2: float(np.dot(T2[1].arr, abij.arr))}
        iters = {}
        for n in (2, 3):
            # calC1.f90: C1^(n) = <a,i|[V,T2^(n-1)]|0> (calC1_VR2)
            C1 = O.Opt(self.dT1)
            c1_r2abci(C1, T2[n - 1], self._opt("abci"), 1.0)
            c1_r2ijka(C1, T2[n - 1], self._opt("ijka"), 1.0)
            # calC2.f90 / calC2_VR.f90: C2^(n) = <ab,ij|[V,T^(n-1)]|0> (+ VTT at n = 3)
            C2 = O.Opt(self.dT2)
            if n > 2:
                c2_r1ijka(C2, T1[n - 1], self._opt("ijka"), 1.0)
                c2_r1abci(C2, T1[n - 1], self._opt("abci"), 1.0)
            c2_r2Wmbej(C2, T2[n - 1], self._opt("iabj"), 1.0)
            c2_r2Wmnij(C2, T2[n - 1], self._opt("ijkl"), 1.0)
            c2_r2abcd(C2, T2[n - 1], self._opt("abcd"), 1.0)
            if n == 3:
                self._c2_vtt(C2, T2[1], 1.0)
            T2[n] = O.Opt(self.dT2, self.denomT2(C2, 1))
            T1[n], iters[n] = self._t1_diis(C1)
            E[n + 1] = float(np.dot(T2[n].arr, abij.arr))  # calEcorr.f90
        return dict(T1=T1, T2=T2, E=E, iterations=iters)

    def _c2_vtt(self, C2, T2, fact):
        """CPSd.SF/calC2_VTT.f90 (T1^(1) = 0)"""
        v, o = self.v, self.o
        abij = self._opt("abij")
        Fli = O.Opt(O.Desc([o, o, Z, Z], (1, 1), 1))
        Fmi_r2abij(Fli, T2, abij, 1.0)
        c2_r2Fmi(C2, T2, Fli, fact)
        Fad = O.Opt(O.Desc([v, v, Z, Z], (1, 1), 1))
        Fae_r2abij(Fad, T2, abij, 1.0)
        c2_r2Fae(C2, T2, Fad, fact)
        W = O.Opt(O.Desc([v, o, v, o], (1, 1), 1))
        Wmbej_r2abij(W, T2, abij, 1.0)
        c2_r2Wmbej(C2, T2, W, fact)
        Wkl = O.Opt(O.Desc([o, o, o, o], (0, 0), 1))
        Wmnij_r2abij(Wkl, T2, abij, 0.5)
        c2_r2Wmnij(C2, T2, Wkl, 2.0 * fact)

    # -----------------------------------------------------------------------
    # grCPS(D) ground state (grCPSd.SF): zeroth-order Jacobian = CIS(Doo)
    # -----------------------------------------------------------------------
    def _gr_intermediates(self):
        """grCPSd.SF/run.f90:29-47: MP2t, Fli, Fad and Wkaci (iabj layout)"""
        v, o = self.v, self.o
        abij = self._opt("abij")
        X = {"MP2t": O.Opt(self.dT2, self.denomT2(abij, 1))}
        X["Fli"] = O.Opt(O.Desc([o, o, Z, Z], (1, 1), 1))
        Fmi_r2abij(X["Fli"], X["MP2t"], abij, -1.0)
        X["Fad"] = O.Opt(O.Desc([v, v, Z, Z], (1, 1), 1))
        Fae_r2abij(X["Fad"], X["MP2t"], abij, -1.0)
        X["Wkaci"] = O.Opt(O.Desc([v, o, v, o], (1, 1), 1), self._opt("iabj").arr.copy())
        Wmbej_r2abij(X["Wkaci"], X["MP2t"], abij, 2.0)
        return X

    def _gr_a0(self, X, T1, T2, R1, R2):
        """calA0R1.f90 + calA0R2.f90: R += A0' T (A0 without its orbital-energy diagonal)"""
        c1_r1iabj(R1, T1, X["Wkaci"], 1.0)
        c1_r1Fmi(R1, T1, X["Fli"], -1.0)
        c1_r1Fae(R1, T1, X["Fad"], -1.0)
        c1_r2ijka(R1, T2, self._opt("ijka"), 1.0)
        c1_r2abci(R1, T2, self._opt("abci"), 1.0)
        c2_r1ijka(R2, T1, self._opt("ijka"), 1.0)
        c2_r1abci(R2, T1, self._opt("abci"), 1.0)

    def _gr_ecorr(self, order, T2, T1s):
        """calEcorr.f90: E^(k+1) = <abij| T2^(k) + [T1 T1]^(k)>"""
        t = O.Opt(self.dT2, T2.arr.copy())
        if order == 2:
            effectT2(t, T1s[1], T1s[1], 1.0, 1)
        elif order == 3:
            effectT2(t, T1s[2], T1s[1], 1.0, 2)
        return float(np.dot(t.arr, self._opt("abij").arr))

    def _gr_solver(self, X, C1, C2, order, T1s):
        """solver.f90: T = (C + A0' T) / D by Jacobi steps with DIIS on (T1 | T2)"""
        l1 = self.dT1.length()
        diis = Diis(self.diis_dim)
        R1 = O.Opt(self.dT1, C1.arr.copy())
        R2 = O.Opt(self.dT2, C2.arr.copy())
        it = 0
        E = 0.0
        while True:
            it += 1
            x = np.concatenate([self.denomT1(R1, 1), self.denomT2(R2, 1)])
            x, err = diis.step(it, x)
            T1 = O.Opt(self.dT1, x[:l1].copy())
            T2 = O.Opt(self.dT2, x[l1:].copy())
            E = self._gr_ecorr(order, T2, T1s)
            if it > 1 and abs(err) < self.diis_max_err:
                break
            if it > self.diis_cyc:
                break
            R1 = O.Opt(self.dT1, C1.arr.copy())
            R2 = O.Opt(self.dT2, C2.arr.copy())
            self._gr_a0(X, T1, T2, R1, R2)
        return T1, T2, E, it

    def _gr_l1_vtt(self, L1, T1, T2):
        """calL1_VTT.f90 with T^(1)"""
        v, o = self.v, self.o
        abij, ijka, abci = self._opt("abij"), self._opt("ijka"), self._opt("abci")
        Gli = O.Opt(O.Desc([o, o, Z, Z], (1, 1), 1))
        Fmi_r1ijka(Gli, T1, ijka, -1.0)
        Fmi_r2abij(Gli, T2, abij, -1.0)
        Gad = O.Opt(O.Desc([v, v, Z, Z], (1, 1), 1))
        Fae_r1abci(Gad, T1, abci, -1.0)
        Fae_r2abij(Gad, T2, abij, -1.0)
        Gld = O.Opt(O.Desc([v, o, Z, Z], (1, 1), 1))
        Fme_r1abij(Gld, T1, abij, 1.0)
        c1_r1Fmi(L1, T1, Gli, -1.0)
        c1_r1Fae(L1, T1, Gad, -1.0)
        c1_r2Fme(L1, T2, Gld, 1.0)

    def _gr_l2_vtt(self, L2, T1, T2):
        """calL2_VTT.f90 with T^(1)"""
        v, o = self.v, self.o
        abij, ijka, abci, iabj = self._opt("abij"), self._opt("ijka"), self._opt("abci"), self._opt("iabj")
        d_ooov = O.Desc([o, o, o, v], (0, 1), 1)
        tmp = O.Opt(d_ooov)
        Wkbij_r1iabj(tmp, T1, iabj, 1.0)
        c2_r1ijka(L2, T1, tmp, -1.0)
        tmp = O.Opt(O.Desc([v, v, v, o], (0, 1), 1))
        Wabdj_r1abcd(tmp, T1, self._opt("abcd"), 1.0)
        c2_r1abci(L2, T1, tmp, 0.5)
        tmp = O.Opt(d_ooov)
        Wkbij_r1Uklij(tmp, T1, self._opt("ijkl"), 1.0)
        c2_r1ijka(L2, T1, tmp, 0.5)
        tmp = O.Opt(O.Desc([o, o, Z, Z], (1, 1), 1))
        Fmi_r1ijka(tmp, T1, ijka, 1.0)
        Fmi_r2abij(tmp, T2, abij, 1.0)
        c2_r2Fmi(L2, T2, tmp, 1.0)
        tmp = O.Opt(O.Desc([v, v, Z, Z], (1, 1), 1))
        Fae_r1abci(tmp, T1, abci, 1.0)
        Fae_r2abij(tmp, T2, abij, 1.0)
        c2_r2Fae(L2, T2, tmp, 1.0)
        tmp = O.Opt(O.Desc([v, o, v, o], (1, 1), 1))
        Wmbej_r1ijka(tmp, T1, ijka, 1.0)
        Wmbej_r1abci(tmp, T1, abci, 1.0)
        Wmbej_r2abij(tmp, T2, abij, 1.0)
        c2_r2Wmbej(L2, T2, tmp, 1.0)
        tmp = O.Opt(d_ooov)
        Wkbij_r2abci(tmp, T2, abci, 1.0)
        c2_r1ijka(L2, T1, tmp, 1.0)
        tmp = O.Opt(O.Desc([o, o, o, o], (0, 0), 1))
        Wmnij_r1ijka(tmp, T1, ijka, 0.5)
        Wmnij_r2abij(tmp, T2, abij, 0.5)
        c2_r2Wmnij(L2, T2, tmp, 2.0)

    def gr_ground_state(self):
        """grCPSd.SF/run.f90 (run1.f90, run2.f90, run3.f90, calC1.f90, calC2.f90,
        calA1R1.f90 (calA2R1), calA1R2.f90, calL1_VTT.f90, calL2_VTT.f90):
        orders 1-3 of A0 T^(n) + C^(n) = 0 with the CIS(Doo) Jacobian as A0.
        Returns T1/T2 per order, E2..E4 and the DIIS iteration counts."""
        X = self._gr_intermediates()
        T1, T2, E, iters = {}, {}, {}, {}
        for n in (1, 2, 3):
            C1 = O.Opt(self.dT1)
            C2 = O.Opt(self.dT2)
            if n == 1:
                C2.arr = self._opt("abij").arr.copy()  # run1.f90:18-19
            else:
                if n == 3:
                    # calA2R1: -[c1_r1iabj(Wkaci - iabj) - Fli - Fad] on T1^(1)
                    W = O.Opt(X["Wkaci"].desc, X["Wkaci"].arr - self._opt("iabj").arr)
                    c1_r1iabj(C1, T1[1], W, -1.0)
                    c1_r1Fmi(C1, T1[1], X["Fli"], 1.0)
                    c1_r1Fae(C1, T1[1], X["Fad"], 1.0)
                    self._gr_l1_vtt(C1, T1[1], T2[1])
                # calA1R2: ring + hh + pp ladder on T2^(n-1)
                c2_r2Wmbej(C2, T2[n - 1], self._opt("iabj"), 1.0)
                c2_r2Wmnij(C2, T2[n - 1], self._opt("ijkl"), 1.0)
                c2_r2abcd(C2, T2[n - 1], self._opt("abcd"), 1.0)
                if n == 3:
                    self._gr_l2_vtt(C2, T1[1], T2[1])
            T1[n], T2[n], E[n + 1], iters[n] = self._gr_solver(X, C1, C2, n, T1)
        return dict(T1=T1, T2=T2, E=E, iterations=iters)

    # -----------------------------------------------------------------------
    # EE-CPS(D) (EE-CPSd.SF)
    # -----------------------------------------------------------------------
    def intermediates(self, gs):
        """calJ.f90: J^(2) from T^(1) (calJ2) and J^(3) from T^(2) (calJ3)"""
        v, o = self.v, self.o
        abij, ijka, abci = self._opt("abij"), self._opt("ijka"), self._opt("abci")
        J = {}
        T2 = gs["T2"][1]
        J["J2_Foo"] = O.Opt(O.Desc([o, o, Z, Z], (1, 1), 1))
        Fmi_r2abij(J["J2_Foo"], T2, abij, 1.0)
        J["J2_Fvv"] = O.Opt(O.Desc([v, v, Z, Z], (1, 1), 1))
        Fae_r2abij(J["J2_Fvv"], T2, abij, 1.0)
        J["J2_Wovoo"] = O.Opt(O.Desc([o, o, o, v], (0, 1), 1))
        Wkbij_r2ijka(J["J2_Wovoo"], T2, ijka, 1.0)
        Wkbij_r2abci(J["J2_Wovoo"], T2, abci, 1.0)
        J["J2_Wovvo"] = O.Opt(O.Desc([v, o, v, o], (1, 1), 1))
        Wmbej_r2abij(J["J2_Wovvo"], T2, abij, 2.0)
        T1, T2 = gs["T1"][2], gs["T2"][2]
        J["J3_Foo"] = O.Opt(O.Desc([o, o, Z, Z], (1, 1), 1))
        Fmi_r1ijka(J["J3_Foo"], T1, ijka, 1.0)
        Fmi_r2abij(J["J3_Foo"], T2, abij, 1.0)
        J["J3_Fvv"] = O.Opt(O.Desc([v, v, Z, Z], (1, 1), 1))
        Fae_r1abci(J["J3_Fvv"], T1, abci, 1.0)
        Fae_r2abij(J["J3_Fvv"], T2, abij, 1.0)
        J["J3_Wovvo"] = O.Opt(O.Desc([v, o, v, o], (1, 1), 1))
        Wmbej_r1ijka(J["J3_Wovvo"], T1, ijka, 1.0)
        Wmbej_r1abci(J["J3_Wovvo"], T1, abci, 1.0)
        Wmbej_r2abij(J["J3_Wovvo"], T2, abij, 2.0)
        return J

    def _orthogonal(self, R1, L):
        """orthogonalR2L.f90: R -= (<L|R> / |L|) L (|L| = 2-norm, as written)"""
        nrm = float(np.linalg.norm(L))
        R1 -= (float(np.dot(L, R1)) / nrm) * L

    def _solver(self, C1t, w0, L, sym):
        """solver.f90: (J_CIS - w0) R1 = C1t by Jacobi steps on the orbital-energy
        denominators with DIIS, R1 kept orthogonal to L"""
        dR1 = O.Desc([self.v, self.o, Z, Z], (1, 1), sym)
        C1 = -C1t
        pR1 = O.Opt(dR1, C1.copy())
        R1 = O.Opt(dR1)
        diis = Diis(self.diis_dim)
        it = 0
        while True:
            it += 1
            R1.arr = self.denomT1(pR1, sym)
            self._orthogonal(R1.arr, L)
            R1.arr, err = diis.step(it, R1.arr)
            self._orthogonal(R1.arr, L)
            if it > 1 and abs(err) < self.diis_max_err:
                break
            if it > self.diis_cyc:
                break
            pR1 = O.Opt(dR1, C1.copy() - w0 * R1.arr)
            c1_r1iabj(pR1, R1, self._opt("iabj"), 1.0)
        return R1.arr, it

    def _calE_C(self, order, R, J, gs, sym):
        """calE_C.f90: C^(n) = sum_k J^(k) R^(n-k) (the terms CPS(D-3) keeps)"""
        v, o = self.v, self.o
        dR1 = O.Desc([v, o, Z, Z], (1, 1), sym)
        dR2 = O.Desc([v, v, o, o], (0, 0), sym)
        C1, C2 = O.Opt(dR1), O.Opt(dR2)
        abci, ijka, iabj, ijkl, abcd = (self._opt(k) for k in ("abci", "ijka", "iabj", "ijkl", "abcd"))
        r1, r2 = O.Opt(dR1, R[order - 1][0].copy()), O.Opt(dR2, R[order - 1][1].copy())
        if order != 1:  # calC1_VR2 (R2^(0) = 0)
            c1_r2abci(C1, r2, abci, 1.0)
            c1_r2ijka(C1, r2, ijka, 1.0)
        if order != 3:
            if order != 2:  # calC2_VR1 (R1^(1) = 0)
                c2_r1ijka(C2, r1, ijka, 1.0)
                c2_r1abci(C2, r1, abci, 1.0)
            if order != 1:  # calC2_VR2
                c2_r2Wmbej(C2, r2, iabj, 1.0)
                c2_r2Wmnij(C2, r2, ijkl, 1.0)
                c2_r2abcd(C2, r2, abcd, 1.0)
        if order == 1:
            return C1, C2
        if order != 3:
            r1, r2 = O.Opt(dR1, R[order - 2][0].copy()), O.Opt(dR2, R[order - 2][1].copy())
            # calC_J2R
            c2_r1ijka(C2, r1, J["J2_Wovoo"], 1.0)
            c1_r1iabj(C1, r1, J["J2_Wovvo"], 1.0)
            c1_r1Fmi(C1, r1, J["J2_Foo"], 1.0)
            c1_r1Fae(C1, r1, J["J2_Fvv"], 1.0)
            # calC_VTR with T^(1) (T1^(1) = 0, R2^(0) = 0)
            self._c_vtr(C2, gs["T2"][1], r1, sym, 1.0)
        if order == 2:
            return C1, C2
        r1 = O.Opt(dR1, R[order - 3][0].copy())
        # calC_J3R
        c1_r1iabj(C1, r1, J["J3_Wovvo"], 1.0)
        c1_r1Fmi(C1, r1, J["J3_Foo"], 1.0)
        c1_r1Fae(C1, r1, J["J3_Fvv"], 1.0)
        return C1, C2

    def _c_vtr(self, C2, T2, r1, sym, fact):
        """calC_VTR.f90: the terms with R1 (R2^(0) = 0) and T2^(1) (T1^(1) = 0)"""
        v, o = self.v, self.o
        ijka, abci = self._opt("ijka"), self._opt("abci")
        W = O.Opt(O.Desc([o, o, o, o], (0, 0), sym))
        Wmnij_r1ijka(W, r1, ijka, 1.0)
        c2_r2Wmnij(C2, T2, W, fact)
        W = O.Opt(O.Desc([v, o, v, o], (1, 1), sym))
        Wmbej_r1abci(W, r1, abci, 1.0)
        c2_r2Wmbej(C2, T2, W, fact)
        G = O.Opt(O.Desc([o, o, Z, Z], (1, 1), sym))
        Fmi_r1ijka(G, r1, ijka, 1.0)
        c2_r2Fmi(C2, T2, G, fact)
        G = O.Opt(O.Desc([v, v, Z, Z], (1, 1), sym))
        Fae_r1abci(G, r1, abci, 1.0)
        c2_r2Fae(C2, T2, G, fact)

    def excited_state(self, gs, J, w0, L0, sym):
        """EE-CPSd.SF/run.f90 + run_order.f90 for one CIS root (w0, L0 = R1^(0)):
        returns w^(1..3), the vectors R^(0..2) and the solver iterations"""
        v, o = self.v, self.o
        L0 = np.asarray(L0, dtype=np.float64)
        R0 = L0.copy()
        nrm = float(np.linalg.norm(R0))
        if abs(nrm - 1.0) > 1.0e-10:  # run.f90:57-60 (scales all but the last element)
            R0[:-1] *= 1.0 / nrm
        l2 = O.Desc([v, v, o, o], (0, 0), sym).length()
        R = {0: (R0, np.zeros(l2))}
        w = {}
        C = None
        oed2 = self.oed2(sym)
        iters = {}
        for order in (1, 2, 3):
            if order != 1:
                # calV_C: Ctilde^(n-1) = sum_{k=2}^{n-1} w^(k) R^(n-1-k) - C^(n-1)
                n1 = order - 1
                Ct1, Ct2 = -C[0].arr.copy(), -C[1].arr.copy()
                for k in range(2, n1 + 1):
                    Ct1 += w[k] * R[n1 - k][0]
                    Ct2 += w[k] * R[n1 - k][1]
                R2 = Ct2 / (oed2 - w0)
                if order != 2:
                    R1, iters[order] = self._solver(Ct1, w0, L0, sym)
                else:
                    R1 = np.zeros_like(R0)
                R[order - 1] = (R1, R2)
            C1, C2 = self._calE_C(order, R, J, gs, sym)
            w[order] = float(np.dot(L0, C1.arr))
            C = (C1, C2)
        return dict(w=w, R=R, iterations=iters)
