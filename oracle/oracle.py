"""CPU oracle bindings (TEST INFRASTRUCTURE ONLY).

Python side of the CPU restatement of the reference's DPD engine path.  The C
engine (dpd_oracle.c) restates the libsrint metadata, sort kernels and the
vv2v_dpd_* engines plus the lib_cc mod_contract surface; this module adds

  * ctypes bindings for it,
  * a numpy dense-tensor view of DPD storage (unpack/pack) used as an
    independent brute-force check of the layout and of every contraction,
  * the named contractions and sigma builders (restated from lib_cc/c1_*.f90,
    c2_*.f90, Fae_r2abij.f90, Fmi_r2abij.f90, Wmbej_r2abij.f90,
    EE-CIS.SF/calAR.f90, CISdoo.SD.SF/{run,calAR}.f90, EE-CPSd.SF/calC2_VR2.f90),
  * the Davidson solver restated from lib_cc/mod_davidson.f90:40-726.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module; it is the checker, never the thing measured or shipped.
"""
from __future__ import annotations

import ctypes
import glob
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

i64 = ctypes.c_int64
dptr = ctypes.POINTER(ctypes.c_double)
lptr = ctypes.POINTER(ctypes.c_int64)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class OracleError(RuntimeError):
    pass


class _Desc(ctypes.Structure):
    _fields_ = [("pop", (i64 * 8) * 4), ("iop", i64 * 2), ("sym", i64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_last_error.restype = ctypes.c_char_p
        for name in ("orc_irpdso", "orc_idsymsoc", "orc_isymoffso", "orc_desc_len", "orc_putidx"):
            getattr(L, name).restype = i64
        _lib = L
        _try_blas(L)
    return _lib


def _try_blas(L):
    """Hand OpenBLAS's ILP64 dgemm (bundled with numpy) to the oracle, as the
    reference links OpenBLAS-ILP64 (make.config:99-114)."""
    import numpy
    pats = glob.glob(os.path.join(os.path.dirname(numpy.__file__), "..", "numpy.libs", "libscipy_openblas64_*.so"))
    if not pats:
        return False
    try:
        ob = ctypes.CDLL(pats[0])
        fn = ob.scipy_dgemm_64_
    except (OSError, AttributeError):
        return False
    global _openblas
    _openblas = ob
    L.orc_set_dgemm(ctypes.cast(fn, ctypes.c_void_p))
    try:
        L.orc_set_blas_threads_fn(ctypes.cast(ob.scipy_openblas_set_num_threads64_, ctypes.c_void_p))
    except AttributeError:
        pass
    return True


_openblas = None


def blas_name():
    lib()
    return "openblas(scipy_openblas64)" if lib().orc_has_blas() else "naive"


def set_threads(n):
    lib().orc_set_threads(int(n))
    if _openblas is not None:
        try:
            _openblas.scipy_openblas_set_num_threads64_(ctypes.c_int(int(n)))
        except AttributeError:
            pass


def _check(rc):
    if rc != 0:
        raise OracleError(lib().orc_last_error().decode())


def _lp(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(lptr)


def _dp(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(dptr)


# ---------------------------------------------------------------------------
# symmetry
# ---------------------------------------------------------------------------
def xor_dirprd(nirrep):
    """Abelian direct-product table: XOR of 0-based irrep ids, 1-based values."""
    d = np.ones((8, 8), dtype=np.int64)
    for i in range(nirrep):
        for j in range(nirrep):
            d[i, j] = (i ^ j) + 1
    return d


class Sym:
    def __init__(self, nirrep):
        self.nirrep = nirrep
        self.dirprd = xor_dirprd(nirrep)

    def activate(self):
        d = np.ascontiguousarray(self.dirprd.reshape(-1))
        lib().orc_set_symmetry(i64(self.nirrep), d.ctypes.data_as(lptr))

    def dp(self, a, b):
        return int(self.dirprd[a - 1, b - 1])


ZORB = np.array([1, 0, 0, 0, 0, 0, 0, 0], dtype=np.int64)


@dataclass
class Desc:
    """Mirror of `type opt` (lib_cc/mod_operator.f90:8-25) without the storage."""
    pops: list  # 4 arrays of 8 ints
    iop: tuple = (1, 1)
    sym: int = 1

    def c(self):
        d = _Desc()
        for k in range(4):
            for h in range(8):
                d.pop[k][h] = int(self.pops[k][h])
        d.iop[0], d.iop[1] = int(self.iop[0]), int(self.iop[1])
        d.sym = int(self.sym)
        return d

    def length(self):
        d = self.c()
        return int(lib().orc_desc_len(ctypes.byref(d)))


def desc2(popp, popq, iop=1, sym=1):
    """2-index tensor as used by vv2v_424/442 operands (pop[2..3] unused)."""
    return Desc([np.asarray(popp), np.asarray(popq), ZORB, ZORB], (iop, 1), sym)


def irpdso(sym, popp, popq, iop):
    _, a = _lp(popp)
    pa, _ = None, None
    ap, app = _lp(popp)
    aq, aqp = _lp(popq)
    return int(lib().orc_irpdso(i64(sym), app, aqp, i64(iop)))


def idsymsoc(irrepx, popp, popq, ioppq, popr, pops, ioprs):
    arrs = [_lp(x) for x in (popp, popq, popr, pops)]
    return int(lib().orc_idsymsoc(i64(irrepx), arrs[0][1], arrs[1][1], i64(ioppq), arrs[2][1], arrs[3][1],
                                  i64(ioprs)))


def isymoffso(irrepq1, irrep, popp, popq, iop):
    a, b = _lp(popp), _lp(popq)
    return int(lib().orc_isymoffso(i64(irrepq1), i64(irrep), a[1], b[1], i64(iop)))


def getidx(idx, desc):
    info = np.zeros(8, dtype=np.int64)
    d = desc.c()
    _check(lib().orc_getidx(i64(idx), info.ctypes.data_as(lptr), ctypes.byref(d)))
    return info


def putidx(info8, desc):
    a = np.ascontiguousarray(info8, dtype=np.int64)
    d = desc.c()
    return int(lib().orc_putidx(a.ctypes.data_as(lptr), ctypes.byref(d)))


# ---------------------------------------------------------------------------
# sorts
# ---------------------------------------------------------------------------
def iexpso(x, popp, popq, ioppq, iexppq, popr, pops, ioprs, iexprs, sym):
    out_len = idsymsoc(sym, popp, popq, 1 if iexppq else ioppq, popr, pops, 1 if iexprs else ioprs)
    out = np.zeros(max(out_len, 1))
    ps = [_lp(p) for p in (popp, popq, popr, pops)]
    _check(lib().orc_iexpso(_dp(x), _dp(out), ps[0][1], ps[1][1], i64(ioppq), i64(iexppq), ps[2][1], ps[3][1],
                            i64(ioprs), i64(iexprs), i64(sym)))
    return out[:out_len]


def icompso(x, popp, popq, ioppq, icpq, popr, pops, ioprs, icrs, sym):
    out_len = idsymsoc(sym, popp, popq, icpq if icpq != 1 else ioppq, popr, pops, icrs if icrs != 1 else ioprs)
    out = np.zeros(max(out_len, 1))
    ps = [_lp(p) for p in (popp, popq, popr, pops)]
    _check(lib().orc_icompso(_dp(x), _dp(out), ps[0][1], ps[1][1], i64(ioppq), i64(icpq), ps[2][1], ps[3][1],
                             i64(ioprs), i64(icrs), i64(sym)))
    return out[:out_len]


def isymtrso(x, popp, popq, ioppq, icpq, popr, pops, ioprs, icrs, sym):
    out = np.zeros(max(len(x), 1))
    ps = [_lp(p) for p in (popp, popq, popr, pops)]
    _check(lib().orc_isymtrso(_dp(x), _dp(out), ps[0][1], ps[1][1], i64(ioppq), i64(icpq), ps[2][1], ps[3][1],
                              i64(ioprs), i64(icrs), i64(sym)))
    return out[:len(x)]


def itranspso(x, popp, popq, ioppq, popr, pops, ioprs, sym):
    out = np.zeros(max(len(x), 1))
    ps = [_lp(p) for p in (popp, popq, popr, pops)]
    _check(lib().orc_itranspso(_dp(x), _dp(out), ps[0][1], ps[1][1], i64(ioppq), ps[2][1], ps[3][1], i64(ioprs),
                               i64(sym)))
    return out[:len(x)]


def iassymso(x, sym, popp, popq, popr, pops, iop):
    """libsrint/iassymso.f:1-142"""
    io = {0: (0, 1, 0, 0), 1: (1, 1, 1, 0), 2: (1, 1, 1, 1), 3: (1, 1, 1, 1), 4: (1, 0, 0, 0), 5: (1, 1, 0, 1)}[iop]
    out_len = idsymsoc(sym, popp, popq, io[2], popr, pops, io[3])
    out = np.zeros(max(out_len, 1))
    ps = [_lp(p) for p in (popp, popq, popr, pops)]
    _check(lib().orc_iassymso(_dp(x), _dp(out), i64(sym), ps[0][1], ps[1][1], ps[2][1], ps[3][1], i64(iop)))
    return out[:out_len]


def sstgenso(x, pop1, pop2, pop3, pop4, irrepx, type4):
    out = np.zeros(max(len(x), 1))
    ps = [_lp(p) for p in (pop1, pop2, pop3, pop4)]
    _check(lib().orc_sstgenso(_dp(x), _dp(out), ps[0][1], ps[1][1], ps[2][1], ps[3][1], i64(irrepx),
                              type4.encode()))
    return out[:len(x)]


# ---------------------------------------------------------------------------
# engines and contract surface
# ---------------------------------------------------------------------------
def _eng(fn, v1, d1, t1, v2, d2, t2, v3, d3, t3, f1, f2):
    c1, c2, c3 = d1.c(), d2.c(), d3.c()
    _check(fn(_dp(v1), ctypes.byref(c1), t1.encode(), _dp(v2), ctypes.byref(c2), t2.encode(), _dp(v3),
              ctypes.byref(c3), t3.encode(), ctypes.c_double(f1), ctypes.c_double(f2)))


def vv2v_dpd_444(v1, d1, t1, v2, d2, t2, v3, d3, t3, f1, f2):
    _eng(lib().orc_vv2v_dpd_444, v1, d1, t1, v2, d2, t2, v3, d3, t3, f1, f2)


def vv2v_dpd_424(v1, d1, t1, v2, d2, t2, v3, d3, t3, f1, f2):
    _eng(lib().orc_vv2v_dpd_424, v1, d1, t1, v2, d2, t2, v3, d3, t3, f1, f2)


def vv2v_dpd_442(v1, d1, t1, v2, d2, t2, v3, d3, t3, f1, f2):
    _eng(lib().orc_vv2v_dpd_442, v1, d1, t1, v2, d2, t2, v3, d3, t3, f1, f2)


def vv2v_dpd_323(v1, d1, t1, v2, d2, t2, v3, d3, t3, f1, f2):
    """libsrint/vv2v_dpd_323.f90:22-145: v3(pqr) = f1 * v1(pqm) v2(mr) + f2 * v3.
    d1/d3 describe the rank-3 tensors (pops [P, Q, M, ZORB], iop (pair, 1)),
    d2 the 2-index v2 (pops [M, R, ZORB, ZORB], iop (1, 1)).  The wrapper's
    checks (:55-88), then the 444 call with the single indices paired with
    1-sized dummies named by the first two free letters (:90-145)."""
    cc = t1[2]
    if cc in t3:
        raise OracleError("vv2v_dpd_323: the third index of v1 must be contracted")
    intact = t1[:2] == t3[:2]
    if d2.iop[0] != 1:
        raise OracleError("vv2v_dpd_323: v2 must have iop 1")
    if not intact and (d1.iop[0] != 1 or d3.iop[0] != 1):
        raise OracleError("vv2v_dpd_323: reordered pair needs iop 1")
    free = [ch for ch in "abcdefghijklmnopqrstuvwxyz" if ch not in t1 + t2 + t3][:2]
    c1, c2 = free
    t1x, t3x = t1 + c1, t3 + c2
    if t2[0] == cc:
        t2x = t2[0] + c1 + t2[1] + c2
    elif t2[1] == cc:
        t2x = t2[0] + c2 + t2[1] + c1
    else:
        raise OracleError("vv2v_dpd_323: contracted index not in v2")
    d2x = Desc([d2.pops[0], ZORB, d2.pops[1], ZORB], (1, 1), d2.sym)
    vv2v_dpd_444(v1, d1, t1x, v2, d2x, t2x, v3, d3, t3x, f1, f2)


def contract_general(v1, d1, t1, v2, d2, t2, v3, d3, t3, f1, f2):
    c1, c2, c3 = d1.c(), d2.c(), d3.c()
    _check(lib().orc_contract_general(_dp(v1), ctypes.byref(c1), t1.encode(), ctypes.c_int(len(t1)), _dp(v2),
                                      ctypes.byref(c2), t2.encode(), ctypes.c_int(len(t2)), _dp(v3),
                                      ctypes.byref(c3), t3.encode(), ctypes.c_int(len(t3)), ctypes.c_double(f1),
                                      ctypes.c_double(f2)))


@dataclass
class Opt:
    """Tensor = descriptor + FP64 storage (mirror of `type opt` with incore=T)."""
    desc: Desc
    arr: np.ndarray = None

    def __post_init__(self):
        if self.arr is None:
            self.arr = np.zeros(self.desc.length())

    @property
    def len(self):
        return self.desc.length()

    def norm(self):
        return float(np.linalg.norm(self.arr))


def vv2v_444(v1: Opt, v2: Opt, v3: Opt, fact, t1, t2, t3, anti=0):
    c1, c2, c3 = v1.desc.c(), v2.desc.c(), v3.desc.c()
    _check(lib().orc_vv2v_444(_dp(v1.arr), ctypes.byref(c1), _dp(v2.arr), ctypes.byref(c2), _dp(v3.arr),
                              ctypes.byref(c3), ctypes.c_double(fact), t1.encode(), t2.encode(), t3.encode(),
                              i64(anti)))


def vv2v_222(v1, v2, v3, fact, t1, t2, t3):
    """mod_contract.f90:166-180"""
    vv2v_444(v1, v2, v3, fact, t1 + "=+", t2 + "=-", t3 + "+-", 0)


def vv2v_422(v1, v2, v3, fact, t1, t2, t3):
    """mod_contract.f90:184-199"""
    vv2v_444(v1, v2, v3, fact, t1, t2 + "+-", t3 + "+-", 0)


def vv2v_224(v1, v2, v3, fact, t1, t2, t3, anti=0):
    """mod_contract.f90:203-219"""
    vv2v_444(v1, v2, v3, fact, t1 + "+-", t2 + "+-", t3, anti)


def vv2v_424(v1: Opt, v2: Opt, v3: Opt, fact, t1, t2, t3, anti=0):
    c1, c2, c3 = v1.desc.c(), v2.desc.c(), v3.desc.c()
    _check(lib().orc_vv2v_424(_dp(v1.arr), ctypes.byref(c1), _dp(v2.arr), ctypes.byref(c2), _dp(v3.arr),
                              ctypes.byref(c3), ctypes.c_double(fact), t1.encode(), t2.encode(), t3.encode(),
                              i64(anti)))


def vv2v_442(v1: Opt, v2: Opt, v3: Opt, fact, t1, t2, t3):
    c1, c2, c3 = v1.desc.c(), v2.desc.c(), v3.desc.c()
    _check(lib().orc_vv2v_442(_dp(v1.arr), ctypes.byref(c1), _dp(v2.arr), ctypes.byref(c2), _dp(v3.arr),
                              ctypes.byref(c3), ctypes.c_double(fact), t1.encode(), t2.encode(), t3.encode()))


# ---------------------------------------------------------------------------
# numpy dense view of DPD storage (independent brute-force layout)
# ---------------------------------------------------------------------------
def _pair_elems(sym: Sym, popp, popq, iop, X):
    """(p, q) global indices of the pair block of irrep X in storage order.

    Order: sub-blocks by irrep of q ascending, p fastest inside
    (build_dpd_tables.f90:28-43); iop=0 keeps p<q only (q outer), iop=2 p<=q."""
    offp = np.concatenate([[0], np.cumsum(popp)])
    offq = np.concatenate([[0], np.cumsum(popq)])
    P, Q = [], []
    for hq in range(1, sym.nirrep + 1):
        hp = sym.dp(hq, X)
        if iop != 1 and hp > hq:
            continue
        npp, nq = int(popp[hp - 1]), int(popq[hq - 1])
        for q in range(nq):
            if iop != 1 and hp == hq:
                ps = range(q) if iop == 0 else range(q + 1)
            else:
                ps = range(npp)
            for p in ps:
                P.append(offp[hp - 1] + p)
                Q.append(offq[hq - 1] + q)
    return np.array(P, dtype=np.int64), np.array(Q, dtype=np.int64)


def dpd_index(sym: Sym, desc: Desc):
    """Global (p,q,r,s) index arrays of every stored element, storage order."""
    out = [[], [], [], []]
    for xr in range(1, sym.nirrep + 1):
        xl = sym.dp(xr, desc.sym)
        P, Q = _pair_elems(sym, desc.pops[0], desc.pops[1], desc.iop[0], xl)
        R, S = _pair_elems(sym, desc.pops[2], desc.pops[3], desc.iop[1], xr)
        if len(P) == 0 or len(R) == 0:
            continue
        out[0].append(np.tile(P, len(R)))
        out[1].append(np.tile(Q, len(R)))
        out[2].append(np.repeat(R, len(P)))
        out[3].append(np.repeat(S, len(P)))
    return [np.concatenate(o) if o else np.zeros(0, np.int64) for o in out]


def dims(desc):
    return tuple(int(np.sum(p)) for p in desc.pops)


def unpack(sym: Sym, desc: Desc, arr):
    """DPD storage -> dense 4-d array; packed pairs are mirrored (iop 0 with a
    minus sign and zero diagonal, iop 2 symmetric)."""
    D = np.zeros(dims(desc))
    p, q, r, s = dpd_index(sym, desc)
    D[p, q, r, s] = arr
    if desc.iop[0] != 1:
        sg = -1.0 if desc.iop[0] == 0 else 1.0
        m = p != q
        D[q[m], p[m], r[m], s[m]] = sg * arr[m]
    if desc.iop[1] != 1:
        sg = -1.0 if desc.iop[1] == 0 else 1.0
        p2, q2, r2, s2, a2 = p, q, r, s, arr
        if desc.iop[0] != 1:
            sg0 = -1.0 if desc.iop[0] == 0 else 1.0
            m = p != q
            p2 = np.concatenate([p, q[m]]); q2 = np.concatenate([q, p[m]])
            r2 = np.concatenate([r, r[m]]); s2 = np.concatenate([s, s[m]])
            a2 = np.concatenate([arr, sg0 * arr[m]])
        m = r2 != s2
        D[p2[m], q2[m], s2[m], r2[m]] = sg * a2[m]
    return D


def pack(sym: Sym, desc: Desc, D):
    p, q, r, s = dpd_index(sym, desc)
    return np.ascontiguousarray(D[p, q, r, s])


def irrep_of(pops):
    """irrep (1-based) of every global index of a space"""
    return np.repeat(np.arange(1, 9), pops)


def sym_mask(sym: Sym, desc: Desc):
    """Dense boolean mask of symmetry-allowed elements."""
    h = [irrep_of(p) for p in desc.pops]
    d = sym.dirprd
    H = d[d[h[0][:, None] - 1, h[1][None, :] - 1][:, :, None] - 1, h[2][None, None, :] - 1]
    H = d[H[..., None] - 1, h[3][None, None, None, :] - 1]
    return H == desc.sym


def random_dpd(sym: Sym, desc: Desc, rng, antisym=True):
    """Random tensor consistent with its packing: full dense random masked by
    symmetry, (anti)symmetrised on packed pairs, packed."""
    D = rng.standard_normal(dims(desc)) * sym_mask(sym, desc)
    if desc.iop[0] != 1:
        D = D - D.transpose(1, 0, 2, 3) if desc.iop[0] == 0 else D + D.transpose(1, 0, 2, 3)
    if desc.iop[1] != 1:
        D = D - D.transpose(0, 1, 3, 2) if desc.iop[1] == 0 else D + D.transpose(0, 1, 3, 2)
    return pack(sym, desc, D)
