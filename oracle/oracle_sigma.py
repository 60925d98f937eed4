"""Oracle sigma builders and Davidson solver (TEST INFRASTRUCTURE ONLY).

Restates, on top of the C oracle's contract surface:
  * EE-CIS.SF/calAR.f90:6-34                  (method "sf-cis")
  * CISdoo.SD.SF/run.f90:31-49 intermediates and CISdoo.SD.SF/calAR.f90:9-79
                                              (method "cisdoo")
  * the above plus EE-CPSd.SF/calC2_VR2.f90:4-47 (method "sf-cpsd")
  * lib_cc/preDavidson.f90:120-223 spin-flip (alpha->beta) guesses
  * lib_cc/mod_davidson.f90:40-726 (Davidson, initDavidson, sort, Schmidt_2D,
    SchmidtS) with numpy.linalg.eig standing in for LAPACK dgeev.
Integrals are inputs (numpy arrays in DPD storage); this module never builds them.
"""
from __future__ import annotations

import numpy as np

import oracle as O

Z = O.ZORB


class OracleProblem:
    def __init__(self, sym: O.Sym, popo, popv, popa, popb, vrta, vrtb, eorb, ints: dict, method="sf-cpsd", symR=1):
        self.sym = sym
        self.popo, self.popv = np.asarray(popo), np.asarray(popv)
        self.popa, self.popb = np.asarray(popa), np.asarray(popb)
        self.vrta, self.vrtb = np.asarray(vrta), np.asarray(vrtb)
        self.eorb = np.asarray(eorb, dtype=np.float64)
        self.ints = ints
        self.method = method
        self.symR = symR
        sym.activate()
        v, o = self.popv, self.popo
        self.dR1 = O.Desc([v, o, Z, Z], (1, 1), symR)
        self.dR2 = O.Desc([v, v, o, o], (0, 0), symR)
        self.l1 = self.dR1.length()
        self.l2 = self.dR2.length() if method != "sf-cis" else 0
        self.dim = self.l1 + self.l2
        self.d = {
            "iabj": O.Desc([v, o, v, o], (1, 1), 1), "ijka": O.Desc([o, o, o, v], (0, 1), 1),
            "ijkl": O.Desc([o, o, o, o], (0, 0), 1), "abij": O.Desc([v, v, o, o], (0, 0), 1),
            "abci": O.Desc([v, v, v, o], (0, 1), 1), "abcd": O.Desc([v, v, v, v], (0, 0), 1),
            "fab": O.Desc([v, v, Z, Z], (1, 1), 1), "fij": O.Desc([o, o, Z, Z], (1, 1), 1),
        }
        self.oed1 = np.zeros(max(self.l1, 1))
        lib = O.lib()
        pv, _ = O._lp(v)
        po, _ = O._lp(o)
        lib.orc_eorb_diff1(O._dp(self.oed1), O._dp(self.eorb), O._lp(v)[1], O._lp(o)[1], O.i64(symR))
        self.oed1 = self.oed1[: self.l1]
        if method != "sf-cis":
            self.oed2 = np.zeros(max(self.l2, 1))
            lib.orc_eorb_diff2(O._dp(self.oed2), O._dp(self.eorb), O._lp(v)[1], O._lp(o)[1], O.i64(symR), O.i64(0))
            self.oed2 = self.oed2[: self.l2]
            self._intermediates()

    def opt(self, name, arr=None):
        return O.Opt(self.d[name], arr if arr is not None else self.ints[name])

    def _mp2_amplitudes(self):
        """MP2t(ab,ij) = -<ab||ij>/e(ab,ij) (denomT.f90:94-249, called at
        CISdoo.SD.SF/run.f90:32), in the layout of abij"""
        v, o = self.popv, self.popo
        dT = self.d["abij"]
        t = np.zeros(max(dT.length(), 1))
        O.lib().orc_denomT2(O._dp(t), O._dp(self.ints["abij"]), O._dp(self.eorb), O._lp(v)[1], O._lp(o)[1],
                            O.i64(1), O.i64(0))
        return t[: dT.length()]

    def _intermediates(self):
        """CISdoo.SD.SF/run.f90:31-49"""
        v, o = self.popv, self.popo
        lib = O.lib()
        self.mp2t = O.Opt(self.d["abij"], self._mp2_amplitudes())
        abij = self.opt("abij")
        # Fac = f_vv + Fae_r2abij(MP2t, abij, 1.0)
        fac = np.zeros(max(self.d["fab"].length(), 1))
        lib.orc_putDiagFock(O._dp(fac), O._dp(self.eorb), O._lp(v)[1], O._lp(o)[1], b"V"[0], O.ctypes.c_double(1.0))
        self.Fac = O.Opt(self.d["fab"], fac[: self.d["fab"].length()].copy())
        O.vv2v_442(self.mp2t, abij, self.Fac, -1.0, "afmn", "efmn", "ae")
        # Fki = f_oo + Fmi_r2abij(MP2t, abij, 1.0)
        fki = np.zeros(max(self.d["fij"].length(), 1))
        lib.orc_putDiagFock(O._dp(fki), O._dp(self.eorb), O._lp(v)[1], O._lp(o)[1], b"O"[0], O.ctypes.c_double(1.0))
        self.Fki = O.Opt(self.d["fij"], fki[: self.d["fij"].length()].copy())
        O.vv2v_442(abij, self.mp2t, self.Fki, 1.0, "efmn", "efin", "mi")
        # Wkaci = iabj + Wmbej_r2abij(MP2t, abij, 2.0)
        self.Wkaci = O.Opt(self.d["iabj"], self.ints["iabj"].copy())
        O.vv2v_444(abij, self.mp2t, self.Wkaci, 0.5 * 2.0, "fenm", "bfjn", "embj", 0)

    def sigma(self, R):
        """One vector R = (R1, R2) -> J R."""
        R = np.ascontiguousarray(R, dtype=np.float64)
        out = np.zeros(self.dim)
        R1 = O.Opt(self.dR1, R[: self.l1].copy())
        C1 = O.Opt(self.dR1, np.zeros(self.l1))
        if self.method == "sf-cis":
            # EE-CIS.SF/calAR.f90:28-32
            C1.arr[:] = self.oed1 * R1.arr
            O.vv2v_422(self.opt("iabj"), R1, C1, 1.0, "fnai", "fn", "ai")
            return C1.arr.copy()
        R2 = O.Opt(self.dR2, R[self.l1:].copy())
        C2 = O.Opt(self.dR2, np.zeros(self.l2))
        # CISdoo.SD.SF/calAR.f90:41-56
        O.vv2v_222(self.Fac, R1, C1, 1.0, "ae", "ei", "ai")              # c1_r1Fae
        O.vv2v_222(R1, self.Fki, C1, -1.0, "am", "mi", "ai")             # c1_r1Fmi
        O.vv2v_422(self.Wkaci, R1, C1, 1.0, "fnai", "fn", "ai")          # c1_r1iabj
        O.vv2v_442(self.opt("abci"), R2, C1, 1.0, "feam", "feim", "ai")  # c1_r2abci
        O.vv2v_442(R2, self.opt("ijka"), C1, -1.0, "aemn", "mnie", "ai")  # c1_r2ijka
        # CISdoo.SD.SF/calAR.f90:63-77
        C2.arr[:] = self.oed2 * R2.arr
        O.vv2v_424(self.opt("abci"), R1, C2, 1.0, "abej", "ei", "abij", 2)  # c2_r1abci
        O.vv2v_424(self.opt("ijka"), R1, C2, -1.0, "ijmb", "am", "abij", 1)  # c2_r1ijka
        if self.method == "sf-cpsd":
            # EE-CPSd.SF/calC2_VR2.f90:29-41
            O.vv2v_444(R2, self.opt("iabj"), C2, 1.0, "aeim", "embj", "abij", 3)  # c2_r2Wmbej
            O.vv2v_444(R2, self.opt("ijkl"), C2, 1.0, "abmn", "mnij", "abij", 0)  # c2_r2Wmnij
            O.vv2v_444(self.opt("abcd"), R2, C2, 1.0, "abef", "efij", "abij", 0)  # c2_r2abcd
        out[: self.l1] = C1.arr
        out[self.l1:] = C2.arr
        return out

    def diag(self):
        if self.method == "sf-cis":
            return self.oed1.copy()
        return np.concatenate([self.oed1, self.oed2])

    def guess(self, nroot):
        """lib_cc/preDavidson.f90:120-223 (spin_flip=ab)"""
        s = self.sym
        n = s.nirrep
        offo = np.concatenate([[0], np.cumsum(self.popo[:n])])
        offv = offo[-1] + np.concatenate([[0], np.cumsum(self.popv[:n])])
        cand = []
        for hi in range(1, n + 1):
            ha = s.dp(hi, self.symR)
            for i in range(int(self.popa[hi - 1])):
                for a in range(int(self.vrtb[ha - 1])):
                    e = self.eorb[offv[ha - 1] + self.vrta[ha - 1] + a] - self.eorb[offo[hi - 1] + i]
                    cand.append((e, len(cand), ha, hi, a, i))
        cand.sort(key=lambda c: (c[0], c[1]))
        R = np.zeros((nroot, self.dim))
        E = []
        for r in range(nroot):
            e, _, ha, hi, a, i = cand[r]
            idx = O.putidx([ha, hi, 1, 1, int(self.vrta[ha - 1]) + a + 1, i + 1, 1, 1], self.dR1)
            assert idx > 0
            R[r, idx - 1] = 1.0
            E.append(e)
        return R, np.array(E)


def _select_roots(vecs, vals, dimIter, nroot, lowest):
    """mod_davidson.f90:224-251: sort (exchange sort, :614-635), then pick by overlap."""
    idx = list(range(dimIter))
    for i in range(dimIter - 1):
        for j in range(i + 1, dimIter):
            if vals[idx[j]] < vals[idx[i]]:
                idx[i], idx[j] = idx[j], idx[i]
    if not lowest:
        for iroot in range(nroot):
            tmp1 = 0.0
            ii = 0
            for j in range(dimIter):
                tmp2 = abs(vecs[iroot, idx[j]])
                if tmp2 > tmp1:
                    tmp1 = tmp2
                    ii = j
            sel = idx[ii]
            # (iroot,...,i,...) -> (i,iroot,...)
            for j in range(1, ii - iroot + 1):
                idx[ii - j + 1] = idx[ii - j]
            idx[iroot] = sel
    return idx


def _schmidt_2d(M):
    """mod_davidson.f90:640-679"""
    out = M.copy()
    for i in range(1, out.shape[1]):
        for j in range(i):
            ov = out[:, j] @ out[:, i]
            out[:, i] -= ov * out[:, j]
        out[:, i] /= np.linalg.norm(out[:, i])
    return out


def davidson(sigma_fn, diagA, guesses, nroot, conv=1e-8, max_dim=20, max_cyc=50, lowest=False, final_check=True,
             log=None):
    """lib_cc/mod_davidson.f90:40-561 restated on in-memory arrays (the
    reference streams b and Ab from disk; the arithmetic is the same)."""
    dimA = len(diagA)
    dimAbar = max_dim
    dimDel = dimAbar - nroot - 5
    if nroot >= dimAbar or nroot == 0:
        raise ValueError("too much roots required")
    maxError = conv
    nsig = 0
    nskip = 0
    b = np.zeros((dimAbar, dimA))
    Ab = np.zeros((dimAbar, dimA))
    Abar = np.zeros((dimAbar, dimAbar))
    vecNeed = np.zeros((dimAbar, nroot))
    eigenVal = np.zeros(nroot)
    # initDavidson :563-610
    for i in range(nroot):
        b[i] = guesses[i]
        Ab[i] = sigma_fn(b[i])
        nsig += 1
        for j in range(nroot):
            Abar[j, i] = guesses[j] @ Ab[i]
        vecNeed[i, i] = 1.0
    iterations = 1
    dimIter = nroot
    dimFree = dimAbar - dimIter
    dim_exceed = False
    ndimIter = ndimFree = 0
    R = np.zeros((nroot, dimA))
    converged = False
    res = np.zeros(nroot)
    history = []
    while True:
        if iterations > max_cyc:
            break
        w, V = np.linalg.eig(Abar[:dimIter, :dimIter])
        valR = w.real.copy()
        vecR = V.real.copy()
        for i in range(dimIter):
            nrm = np.linalg.norm(vecR[:, i])
            if abs(nrm - 1.0) > 1e-10:
                vecR[:, i] /= nrm
        lowestE = _select_roots(vecR, valR, dimIter, nroot, lowest)
        for iroot in range(nroot):
            vecNeed[:dimIter, iroot] = vecR[:, lowestE[iroot]]
            eigenVal[iroot] = valR[lowestE[iroot]]
        history.append(eigenVal.copy())
        if dim_exceed:
            # ndimIter = nroot + 5: the reference reads past the subspace when
            # davidson_dim < nroot + 5; treated as an input error here
            if ndimIter > dimIter:
                raise ValueError("subspace collapse needs davidson_dim >= nroot + 5")
            old = np.stack([vecR[:, lowestE[i]] for i in range(ndimIter)], axis=1)
            new = _schmidt_2d(old)
            nb = new.T @ b[:dimIter]
            nAb = new.T @ Ab[:dimIter]
            # the reference accumulates with daxpy in i order; same sums
            nb = np.zeros((ndimIter, dimA))
            nAb = np.zeros((ndimIter, dimA))
            for r in range(ndimIter):
                for i in range(dimIter):
                    nAb[r] += new[i, r] * Ab[i]
                    nb[r] += new[i, r] * b[i]
            b[:ndimIter] = nb
            Ab[:ndimIter] = nAb
            dimFree = ndimFree
            dimIter = ndimIter
            for i in range(dimIter):
                for j in range(dimIter):
                    Abar[j, i] = Ab[i] @ b[j]
        nNewBase = 0
        conv_ctrl = np.zeros(nroot, dtype=int)
        for iroot in range(nroot):
            if dim_exceed:
                Aq = Ab[iroot].copy()
                q = b[iroot].copy()
            else:
                Aq = np.zeros(dimA)
                q = np.zeros(dimA)
                for i in range(dimIter):
                    Aq += vecNeed[i, iroot] * Ab[i]
                    q += vecNeed[i, iroot] * b[i]
            t = np.linalg.norm(q)
            if abs(t - 1.0) > 1e-10:
                q /= t
                Aq /= t
            R[iroot] = q
            e = Aq - eigenVal[iroot] * q
            k = int(np.argmax(np.abs(e)))
            res[iroot] = abs(e[k])
            if abs(e[k]) < maxError:
                conv_ctrl[iroot] = 1
                continue
            if dimFree == 0:
                continue
            den = eigenVal[iroot] - diagA
            small = np.abs(den) < 1e-4
            den = np.where(small, np.where(den < 0.0, -1e-4, 1e-4), den)
            p = e / den
            # SchmidtS :687-726
            p = p / np.linalg.norm(p)
            for i in range(dimIter + nNewBase):
                ov = b[i] @ p
                p = p - ov * b[i]
            nrm = np.linalg.norm(p)
            if nrm < 1e-6:
                nskip += 1
                continue
            p = p / nrm
            b[dimIter + nNewBase] = p
            Ab[dimIter + nNewBase] = sigma_fn(p)
            nsig += 1
            dimFree -= 1
            nNewBase += 1
        if log is not None:
            log(iterations, eigenVal.copy(), res.copy())
        if conv_ctrl.sum() == nroot:
            converged = True
            break
        if dim_exceed:
            dim_exceed = False
        if dimFree == 0:
            dim_exceed = True
            ndimIter = dimAbar - dimDel
            if ndimIter < nroot:
                ndimIter = nroot
                ndimFree = dimAbar - ndimIter
            else:
                ndimFree = dimDel
        for j in range(dimIter):
            for i in range(dimIter, dimIter + nNewBase):
                Abar[i, j] = b[i] @ Ab[j]
        for j in range(dimIter, dimIter + nNewBase):
            for i in range(dimIter + nNewBase):
                Abar[i, j] = b[i] @ Ab[j]
        dimIter += nNewBase
        iterations += 1
    final = np.zeros(nroot)
    if final_check:
        for iroot in range(nroot):
            e = sigma_fn(R[iroot]) - eigenVal[iroot] * R[iroot]
            nsig += 1
            final[iroot] = np.max(np.abs(e))
    return {"w": eigenVal.copy(), "R": R, "iterations": iterations, "converged": converged, "sigma_calls": nsig,
            "skipped": nskip, "residual": res, "final_residual": final, "history": history}
