"""dpdhip -- Python front end of the MI355X DPD contraction engine.

Mirrors the reference's lib_cc interface (lib_cc/mod_operator.f90 `type opt`,
lib_cc/mod_contract.f90 vv2v_444/222/422/224/424/442, the libsrint sort kernels)
over the C-ABI in include/dpd_hip.h.  Device memory comes from torch (HIP);
every computation runs in the HIP library.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import DpdError, check, lib

__all__ = ["Context", "Opt", "Spec", "Problem", "DpdError", "xor_dirprd", "ZORB", "sigma_group", "davidson_batch"]

ZORB = np.array([1, 0, 0, 0, 0, 0, 0, 0], dtype=np.int64)


def _torch():
    import torch
    return torch


def xor_dirprd(nirrep):
    """Abelian direct-product table (XOR of 0-based irrep ids), 1-based values."""
    d = np.ones((8, 8), dtype=np.int64)
    for i in range(nirrep):
        for j in range(nirrep):
            d[i, j] = (i ^ j) + 1
    return d


def _i64(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int64))
    return a, a.ctypes.data_as(_lib.c_i64_p)


class Context:
    """Replaces the reference's global symmetry/IOP state (mod_symm, mod_iop)."""

    def __init__(self, nirrep, device=0, dirprd=None):
        torch = _torch()
        self.nirrep = nirrep
        self.device = device
        self.dirprd = xor_dirprd(nirrep) if dirprd is None else np.asarray(dirprd, dtype=np.int64)
        torch.cuda.init()
        h = ctypes.c_void_p()
        d, dp = _i64(self.dirprd.reshape(-1))
        check(lib().dpd_ctx_create(ctypes.byref(h), nirrep, dp, device))
        self.h = h
        self.data_prec = 8
        self.use_current_stream()

    def use_current_stream(self):
        torch = _torch()
        s = torch.cuda.current_stream(self.device).cuda_stream
        check(lib().dpd_ctx_set_stream(self.h, ctypes.c_void_p(s)))

    def synchronize(self):
        check(lib().dpd_ctx_synchronize(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().dpd_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_gemm_config(self, cfg=-1, packing=1):
        """GEMM kernels of later plans.  cfg: -1 automatic (tile shape timed on a
        plan's first call), -2 automatic by cost model, 0 first-generation kernel,
        1..7 tiles 64x64, 128x64, 64x128, 128x128, 32x32, 64x192 (packed kernel), 7 the
        BF16x3 128x64 tile that reads A straight into registers (128x64 in the other precisions).
        packing: 0 operands gathered inside the GEMM, 1 packed panels when an
        operand is a problem constant (packed once), 2 packed panels always."""
        check(lib().dpd_ctx_set_gemm_config(self.h, int(cfg), int(packing)))

    def set_data_precision(self, prec):
        """the reference's mod_iop prec (libsrint/readinp.f90:120-127): 8 FP64
        tensors, 4 FP32 tensors for the engines, sorts and xgemm (FP32 MFMA, the
        result update formed in FP64 and rounded once); the contract surface
        stays FP64"""
        check(lib().dpd_ctx_set_data_precision(self.h, int(prec)))
        self.data_prec = int(prec)

    def dtype(self):
        return _torch().float32 if self.data_prec == 4 else _torch().float64

    def set_split_k(self, mode=-1):
        """split-K of the packed-panel GEMMs: -1 automatic, 0 off, 1 on"""
        check(lib().dpd_ctx_set_split_k(self.h, mode))

    def profile(self, on=True, max_records=100000):
        """Enable/disable per-launch HIP-event timing of the tagged GEMM stages."""
        check(lib().dpd_profile_enable(self.h, int(on), int(max_records)))

    def profile_report(self):
        import json
        buf = ctypes.create_string_buffer(1 << 18)
        check(lib().dpd_profile_report(self.h, buf, len(buf)))
        return json.loads(buf.value.decode())

    # metadata ---------------------------------------------------------------
    def irpdso(self, sym, popp, popq, iop):
        a, ap = _i64(popp)
        b, bp = _i64(popq)
        return int(lib().dpd_irpdso(self.h, sym, ap, bp, iop))

    def idsymsoc(self, irrepx, popp, popq, ioppq, popr, pops, ioprs):
        ps = [_i64(x) for x in (popp, popq, popr, pops)]
        return int(lib().dpd_idsymsoc(self.h, irrepx, ps[0][1], ps[1][1], ioppq, ps[2][1], ps[3][1], ioprs))

    def isymoffso(self, irrepq1, irrep, popp, popq, iop):
        a, ap = _i64(popp)
        b, bp = _i64(popq)
        return int(lib().dpd_isymoffso(self.h, irrepq1, irrep, ap, bp, iop))


class Opt:
    """Mirror of `type opt` (lib_cc/mod_operator.f90:8-25): populations of the
    four indices, pair packing (ioppq, ioprs), symmetry and a device array (FP64,
    or FP32 when the context's data precision is 4)."""

    def __init__(self, ctx: Context, popp, popq, popr=ZORB, pops=ZORB, ioppq=1, ioprs=1, sym=1, arr=None):
        torch = _torch()
        self.ctx = ctx
        self.pops = [np.asarray(p, dtype=np.int64) for p in (popp, popq, popr, pops)]
        self.ioppq, self.ioprs, self.sym = int(ioppq), int(ioprs), int(sym)
        t = self._struct()
        self.len = int(lib().dpd_tensor_len(ctx.h, ctypes.byref(t)))
        if arr is None:
            arr = torch.zeros(max(self.len, 1), dtype=ctx.dtype(), device=f"cuda:{ctx.device}")
        self.arr = arr

    def _struct(self, arr=None):
        t = _lib.Tensor_()
        for k in range(4):
            for h in range(8):
                t.pop[k][h] = int(self.pops[k][h])
        t.iop[0], t.iop[1], t.sym = self.ioppq, self.ioprs, self.sym
        a = arr if arr is not None else getattr(self, "arr", None)
        t.data = _data(self.ctx, a, self.len) if a is not None else None
        return t

    def norm(self):
        return float(self.arr[: self.len].norm())

    def numpy(self):
        return self.arr[: self.len].cpu().numpy()


def _data(ctx, x, n=None):
    """device pointer of a tensor the library reads or writes as elements of the
    context's data precision (FP64, or FP32 under prec 4); anything else is
    refused before the pointer is handed over"""
    torch = _torch()
    if not isinstance(x, torch.Tensor) or x.device.type != "cuda" or x.device.index != ctx.device:
        raise ValueError(f"tensor data must be a tensor on cuda:{ctx.device}")
    if x.dtype != ctx.dtype():
        raise ValueError(f"tensor data must be {ctx.dtype()} (data precision {ctx.data_prec}), got {x.dtype}")
    if not x.is_contiguous():
        raise ValueError("tensor data must be contiguous")
    if n is not None and x.numel() < n:
        raise ValueError(f"tensor data holds {x.numel()} elements, the layout needs {n}")
    return x.data_ptr()


def _lab(s):
    return s.encode()


def vv2v_dpd_444(v1: Opt, t1, v2: Opt, t2, v3: Opt, t3, fact1, fact2):
    """libsrint/vv2v_dpd_444.f90: v3 = fact1*v1(x)v2 + fact2*v3"""
    a, b, c = v1._struct(), v2._struct(), v3._struct()
    check(lib().dpd_vv2v_dpd_444(v1.ctx.h, ctypes.byref(a), _lab(t1), ctypes.byref(b), _lab(t2), ctypes.byref(c),
                                 _lab(t3), fact1, fact2))


def vv2v_dpd_323(v1: Opt, t1, v2: Opt, t2, v3: Opt, t3, fact1, fact2):
    """libsrint/vv2v_dpd_323.f90: v3(pqr) = fact1*v1(pqm)v2(mr) + fact2*v3; the
    rank-3 Opts use popp, popq, popr and ioppq, v2 popp, popq"""
    a, b, c = v1._struct(), v2._struct(), v3._struct()
    check(lib().dpd_vv2v_dpd_323(v1.ctx.h, ctypes.byref(a), _lab(t1), ctypes.byref(b), _lab(t2), ctypes.byref(c),
                                 _lab(t3), fact1, fact2))


def vv2v_dpd_424(v1: Opt, t1, v2: Opt, t2, v3: Opt, t3, fact1, fact2):
    a, b, c = v1._struct(), v2._struct(), v3._struct()
    check(lib().dpd_vv2v_dpd_424(v1.ctx.h, ctypes.byref(a), _lab(t1), ctypes.byref(b), _lab(t2), ctypes.byref(c),
                                 _lab(t3), fact1, fact2))


def vv2v_dpd_442(v1: Opt, t1, v2: Opt, t2, v3: Opt, t3, fact1, fact2):
    a, b, c = v1._struct(), v2._struct(), v3._struct()
    check(lib().dpd_vv2v_dpd_442(v1.ctx.h, ctypes.byref(a), _lab(t1), ctypes.byref(b), _lab(t2), ctypes.byref(c),
                                 _lab(t3), fact1, fact2))


def _contract(fn, v1, v2, v3, fact, t1, t2, t3, *extra):
    a, b, c = v1._struct(), v2._struct(), v3._struct()
    check(fn(v1.ctx.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), fact, _lab(t1), _lab(t2), _lab(t3),
             *extra))


def vv2v_444(v1, v2, v3, fact, t1, t2, t3, anti=0):
    """lib_cc/mod_contract.f90:10-162: v3 += fact * P{v1 (x) v2}"""
    _contract(lib().dpd_vv2v_444, v1, v2, v3, fact, t1, t2, t3, anti)


def vv2v_222(v1, v2, v3, fact, t1, t2, t3):
    _contract(lib().dpd_vv2v_222, v1, v2, v3, fact, t1, t2, t3)


def vv2v_422(v1, v2, v3, fact, t1, t2, t3):
    _contract(lib().dpd_vv2v_422, v1, v2, v3, fact, t1, t2, t3)


def vv2v_224(v1, v2, v3, fact, t1, t2, t3, anti=0):
    _contract(lib().dpd_vv2v_224, v1, v2, v3, fact, t1, t2, t3, anti)


def vv2v_424(v1, v2, v3, fact, t1, t2, t3, anti=0):
    _contract(lib().dpd_vv2v_424, v1, v2, v3, fact, t1, t2, t3, anti)


def vv2v_442(v1, v2, v3, fact, t1, t2, t3):
    _contract(lib().dpd_vv2v_442, v1, v2, v3, fact, t1, t2, t3)


# ---- sort kernels (device tensors in and out) ---------------------------------
def iexpso(ctx, x, out, popp, popq, ioppq, iexppq, popr, pops, ioprs, iexprs, sym):
    ps = [_i64(p) for p in (popp, popq, popr, pops)]
    check(lib().dpd_iexpso(ctx.h, _data(ctx, x), _data(ctx, out), ps[0][1], ps[1][1], ioppq, iexppq, ps[2][1],
                           ps[3][1], ioprs, iexprs, sym))


def icompso(ctx, x, out, popp, popq, ioppq, icpq, popr, pops, ioprs, icrs, sym):
    ps = [_i64(p) for p in (popp, popq, popr, pops)]
    check(lib().dpd_icompso(ctx.h, _data(ctx, x), _data(ctx, out), ps[0][1], ps[1][1], ioppq, icpq, ps[2][1],
                            ps[3][1], ioprs, icrs, sym))


def isymtrso(ctx, x, out, popp, popq, ioppq, icpq, popr, pops, ioprs, icrs, sym):
    ps = [_i64(p) for p in (popp, popq, popr, pops)]
    check(lib().dpd_isymtrso(ctx.h, _data(ctx, x), _data(ctx, out), ps[0][1], ps[1][1], ioppq, icpq, ps[2][1],
                             ps[3][1], ioprs, icrs, sym))


def itranspso(ctx, x, out, popp, popq, ioppq, popr, pops, ioprs, sym):
    ps = [_i64(p) for p in (popp, popq, popr, pops)]
    check(lib().dpd_itranspso(ctx.h, _data(ctx, x), _data(ctx, out), ps[0][1], ps[1][1], ioppq, ps[2][1], ps[3][1],
                              ioprs, sym))


def iassymso(ctx, x, out, sym, popp, popq, popr, pops, iop):
    """libsrint/iassymso.f:1-142 (pair antisymmetriser / symmetriser)"""
    ps = [_i64(p) for p in (popp, popq, popr, pops)]
    check(lib().dpd_iassymso(ctx.h, _data(ctx, x), _data(ctx, out), sym, ps[0][1], ps[1][1], ps[2][1], ps[3][1], iop))


def sstgenso(ctx, x, out, pop1, pop2, pop3, pop4, irrepx, type4):
    ps = [_i64(p) for p in (pop1, pop2, pop3, pop4)]
    check(lib().dpd_sstgenso(ctx.h, _data(ctx, x), _data(ctx, out), ps[0][1], ps[1][1], ps[2][1], ps[3][1], irrepx,
                             type4.encode()))


# ---- dense GEMM (libsrint/xgemmso.f router) ----------------------------------
def xgemm(ctx, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, prec="fp64"):
    check(lib().dpd_xgemm(ctx.h, transa.encode(), transb.encode(), m, n, k, alpha, _data(ctx, A), lda,
                          _data(ctx, B), ldb, beta, _data(ctx, C), ldc, PRECISIONS[prec]))


# ---- vector primitives ------------------------------------------------------
def dot(ctx, x, y):
    r = ctypes.c_double()
    check(lib().dpd_dot(ctx.h, x.numel(), x.data_ptr(), y.data_ptr(), ctypes.byref(r)))
    return r.value


def nrm2(ctx, x):
    r = ctypes.c_double()
    check(lib().dpd_nrm2(ctx.h, x.numel(), x.data_ptr(), ctypes.byref(r)))
    return r.value


def amax(ctx, x):
    i = ctypes.c_int64()
    v = ctypes.c_double()
    check(lib().dpd_amax(ctx.h, x.numel(), x.data_ptr(), ctypes.byref(i), ctypes.byref(v)))
    return i.value, v.value


def axpy(ctx, a, x, y):
    check(lib().dpd_axpy(ctx.h, x.numel(), a, x.data_ptr(), y.data_ptr()))


def scal(ctx, a, x):
    check(lib().dpd_scal(ctx.h, x.numel(), a, x.data_ptr()))


def precondition(ctx, w, diag, e, p):
    check(lib().dpd_precondition(ctx.h, e.numel(), w, diag.data_ptr(), e.data_ptr(), p.data_ptr()))


# ---- synthetic Hamiltonians ---------------------------------------------------
@dataclass
class Spec:
    """Seeded synthetic spin-free Hamiltonian (see csrc/dpd_synth.cpp)."""
    nso: tuple
    nalpha: int
    nbeta: int
    seed: int = 0

    @property
    def nirrep(self):
        return len(self.nso)

    def c(self):
        s = _lib.SynthSpec_()
        s.nirrep = len(self.nso)
        for h, n in enumerate(self.nso):
            s.nso[h] = n
        s.nalpha, s.nbeta, s.seed = self.nalpha, self.nbeta, self.seed
        check(lib().dpd_synth_layout(ctypes.byref(s)))
        return s

    def layout(self):
        s = self.c()
        g = lambda f: np.array(list(getattr(s, f)), dtype=np.int64)
        return {k: g(k) for k in ("popa", "popb", "vrta", "vrtb", "popo", "popv")}

    def fill(self, which, iop, length):
        s = self.c()
        out = np.zeros(max(int(length), 1))
        check(lib().dpd_synth_fill(ctypes.byref(s), which.encode(), iop, out.ctypes.data_as(_lib.c_double_p),
                                   int(length)))
        return out[: int(length)]

    def eri(self, idx4):
        s = self.c()
        idx = np.ascontiguousarray(np.asarray(idx4, dtype=np.int64).reshape(-1, 4))
        out = np.zeros(len(idx))
        check(lib().dpd_synth_eri(ctypes.byref(s), len(idx), idx.ctypes.data_as(_lib.c_i64_p),
                                  out.ctypes.data_as(_lib.c_double_p)))
        return out


def small_eig(A):
    """Subspace eigensolve (host FP64): eigenvalues and unit-norm right
    eigenvectors of a real nonsymmetric matrix."""
    A = np.ascontiguousarray(np.asarray(A, dtype=np.float64))
    n = A.shape[0]
    wr, wi = np.zeros(n), np.zeros(n)
    V = np.zeros(n * n)
    P = lambda a: a.ctypes.data_as(_lib.c_double_p)
    check(lib().dpd_small_eig(n, P(A), P(wr), P(wi), P(V)))
    return wr, wi, V.reshape(n, n, order="F")


METHODS = {"sf-cis": 0, "cisdoo": 1, "sf-cpsd": 2}
PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x3": 2}


def _check_vectors(x, what, dim, device, nvec=None):
    """x must be a contiguous float64 tensor on cuda:<device> holding a whole
    number (nvec, when given) of dim-long vectors; returns the vector count."""
    torch = _torch()
    if not isinstance(x, torch.Tensor) or x.dtype != torch.float64 or not x.is_cuda \
            or x.device.index != device or not x.is_contiguous():
        raise ValueError(f"{what}: expected a contiguous float64 tensor on cuda:{device}")
    n = x.numel()
    if dim <= 0 or n == 0 or n % dim:
        raise ValueError(f"{what}: {n} elements is not a whole number of vectors of length {dim}")
    if nvec is not None and n != nvec * dim:
        raise ValueError(f"{what}: expected {nvec} vector(s) of length {dim}, got {n} elements")
    return n // dim


def sigma_group(problems, Rs, outs=None):
    """sigma of several problems (same context, method, precision) with one
    grouped GEMM launch per contraction"""
    torch = _torch()
    n = len(problems)
    if len(Rs) != n or (outs is not None and len(outs) != n):
        raise ValueError("one R (and one out) per problem")
    nv = [p._vectors(R, f"Rs[{i}]") for i, (p, R) in enumerate(zip(problems, Rs))]
    outs = outs if outs is not None else [torch.empty_like(R) for R in Rs]
    for i, (p, o) in enumerate(zip(problems, outs)):
        p._vectors(o, f"outs[{i}]", nv[i])
    P = (ctypes.c_void_p * n)(*[p.h.value for p in problems])
    X = (ctypes.c_void_p * n)(*[R.data_ptr() for R in Rs])
    Y = (ctypes.c_void_p * n)(*[o.data_ptr() for o in outs])
    nv = np.array([R.numel() // p.dim for p, R in zip(problems, Rs)], dtype=np.int64)
    check(lib().dpd_problem_sigma_group(P, X, Y, nv.ctypes.data_as(_lib.c_i64_p), n))
    return outs


def davidson_batch(problems, nroot, conv=1e-8, max_dim=20, max_cyc=50, lowest=False, final_check=True, Rs=None,
                   grouped=True, graph=True):
    """Davidson solves of several problems in lockstep; their sigma steps are
    applied together (grouped GEMM launches, replayed from a hipGraph)."""
    n = len(problems)
    if Rs is None:
        Rs = [p.guess(nroot)[0] for p in problems]
    if len(Rs) != n:
        raise ValueError("one R per problem")
    for i, (p, R) in enumerate(zip(problems, Rs)):
        p._vectors(R, f"Rs[{i}]", nroot)
    opts = (_lib.DavidsonOpts_ * n)()
    for o in opts:
        o.nroot, o.max_dim, o.max_cyc, o.conv = nroot, max_dim, max_cyc, conv
        o.lowest, o.final_check, o.transpose = int(lowest), int(final_check), 0
    ws = [np.zeros(nroot) for _ in range(n)]
    P = (ctypes.c_void_p * n)(*[p.h.value for p in problems])
    R = (ctypes.c_void_p * n)(*[r.data_ptr() for r in Rs])
    W = (ctypes.c_void_p * n)(*[w.ctypes.data for w in ws])
    flags = (1 if grouped else 0) | (2 if graph else 0)
    stats = np.zeros(4, dtype=np.int64)
    rc = lib().dpd_davidson_batch(P, n, opts, R, W, flags, stats.ctypes.data_as(_lib.c_i64_p))
    if rc not in (0, _lib.DPD_ENOTCONV):
        check(rc)
    infos = [{"converged": o.iterations <= max_cyc and all(r < conv for r in o.residual[:nroot]),
              "iterations": o.iterations, "sigma_calls": o.sigma_calls, "skipped": o.skipped,
              "residual": list(o.residual[:nroot]), "final_residual": list(o.final_residual[:nroot])}
             for o in opts]
    for i in infos:
        i["batch"] = {"steps": int(stats[0]), "graph_replays": int(stats[1]), "graphs": int(stats[2]),
                      "captures_abandoned": int(stats[3])}
    return ws, Rs, infos


class Problem:
    """Device-resident operator (integrals + intermediates) of one method and
    excited-state symmetry; sigma = J R for batches of vectors."""

    def __init__(self, ctx: Context, spec: Spec, method="sf-cpsd", symR=1):
        self.ctx = ctx
        self.spec = spec
        self.method = method
        self.symR = symR
        s = spec.c()
        h = ctypes.c_void_p()
        check(lib().dpd_problem_create(ctx.h, ctypes.byref(s), METHODS[method], symR, ctypes.byref(h)))
        self.h = h
        self.dim = int(lib().dpd_problem_dim(h))
        self.precision = "fp64"
        self.version = 0  # moves on with every set() and refresh()

    def set_precision(self, prec):
        check(lib().dpd_problem_set_precision(self.h, PRECISIONS[prec]))
        self.precision = prec

    def _vectors(self, x, what, nvec=None):
        """Number of vectors in x, which the library reads or writes as nvec * dim
        contiguous doubles on this problem's device; anything else is refused
        here, before a device pointer is handed over."""
        return _check_vectors(x, what, self.dim, self.ctx.device, nvec)

    def sigma(self, R, out=None):
        torch = _torch()
        nvec = self._vectors(R, "R")
        if out is None:
            out = torch.empty_like(R)
        self._vectors(out, "out", nvec)
        check(lib().dpd_problem_sigma(self.h, R.data_ptr(), out.data_ptr(), nvec))
        return out

    def sigma_t(self, L, out=None):
        """t = J^T L (the transposed operator, same terms with operand roles exchanged)."""
        torch = _torch()
        nvec = self._vectors(L, "L")
        if out is None:
            out = torch.empty_like(L)
        self._vectors(out, "out", nvec)
        check(lib().dpd_problem_sigma_t(self.h, L.data_ptr(), out.data_ptr(), nvec))
        return out

    INPUTS = ("eorb", "iabj", "ijka", "ijkl", "abij", "abci", "abcd")

    def tensor_len(self, name):
        n = int(lib().dpd_problem_tensor_len(self.h, name.encode()))
        if n < 0:  # unknown name, or a tensor that is not on the device (abcd out of core)
            raise _lib.DpdError(_lib.DPD_EINVAL, f"{name}: {lib().dpd_last_error().decode()}")
        return n

    def get(self, name):
        """Device copy of a named input or intermediate (float64, flat)."""
        torch = _torch()
        out = torch.empty(self.tensor_len(name), dtype=torch.float64, device=f"cuda:{self.ctx.device}")
        if out.numel():
            check(lib().dpd_problem_get(self.h, name.encode(), out.data_ptr()))
        self.ctx.synchronize()
        return out

    def set(self, name, x):
        """Overwrite an input with a device tensor of its length; call refresh() after."""
        n = self.tensor_len(name)
        if x.numel() != n or x.dtype != _torch().float64 or not x.is_cuda or x.device.index != self.ctx.device:
            raise ValueError(f"{name}: expected a float64 tensor of {n} elements on cuda:{self.ctx.device}")
        x = x.contiguous()
        check(lib().dpd_problem_set(self.h, name.encode(), x.data_ptr()))
        self.ctx.synchronize()
        self.version += 1

    def refresh(self):
        check(lib().dpd_problem_refresh(self.h))
        self.version += 1

    def grad(self, L, R, scale=1.0, out=None):
        """scale * d(L . J R)/d(inputs), accumulated into out (dict name -> tensor,
        zero-initialised when not given); returns the dict."""
        torch = _torch()
        if out is None:
            out = {}
            for k in self.INPUTS:
                n = self.tensor_len(k)
                if n:
                    out[k] = torch.zeros(n, dtype=torch.float64, device=f"cuda:{self.ctx.device}")
        self._vectors(L, "L", 1)
        self._vectors(R, "R", 1)
        for k in self.INPUTS:
            if k in out:
                _check_vectors(out[k], f"out[{k!r}]", self.tensor_len(k), self.ctx.device, 1)
        ptr = [out[k].data_ptr() if k in out else None for k in self.INPUTS]
        check(lib().dpd_problem_grad(self.h, L.data_ptr(), R.data_ptr(), float(scale), *ptr))
        self.ctx.synchronize()
        return out

    def diag(self):
        torch = _torch()
        d = torch.empty(self.dim, dtype=torch.float64, device=f"cuda:{self.ctx.device}")
        check(lib().dpd_problem_diag(self.h, d.data_ptr()))
        return d

    def guess(self, nroot):
        torch = _torch()
        R = torch.zeros(nroot * self.dim, dtype=torch.float64, device=f"cuda:{self.ctx.device}")
        e = np.zeros(nroot)
        check(lib().dpd_problem_guess(self.h, nroot, R.data_ptr(), e.ctypes.data_as(_lib.c_double_p)))
        return R, e

    def davidson(self, nroot, conv=1e-8, max_dim=20, max_cyc=50, lowest=False, final_check=True, R=None,
                 transpose=False, graph=False):
        """Non-Hermitian Davidson (lib_cc/mod_davidson.f90:40-561); transpose=True
        solves with J^T (left eigenvectors); graph=True replays each iteration's
        sigma step from a hipGraph."""
        if R is None:
            R, _ = self.guess(nroot)
        self._vectors(R, "R", nroot)
        o = _lib.DavidsonOpts_()
        o.nroot, o.max_dim, o.max_cyc, o.conv = nroot, max_dim, max_cyc, conv
        o.lowest, o.final_check, o.transpose = int(lowest), int(final_check), int(transpose)
        o.graph = int(graph)
        w = np.zeros(nroot)
        rc = lib().dpd_davidson(self.h, ctypes.byref(o), R.data_ptr(), w.ctypes.data_as(_lib.c_double_p))
        if rc not in (0, _lib.DPD_ENOTCONV):
            check(rc)
        info = {"converged": rc == 0, "iterations": o.iterations, "sigma_calls": o.sigma_calls,
                "skipped": o.skipped, "graph_replays": o.graph_replays,
                "residual": list(o.residual[:nroot]), "final_residual": list(o.final_residual[:nroot])}
        return w, R, info

    def set_abcd_out_of_core(self, on=True):
        """keep <ab||ef> in pinned host memory and stream it block by block into
        the pp-ladder (the reference's out-of-core c2_r2abcd, c2_r2abcd.f90:29-104);
        on=False brings it back to the device"""
        check(lib().dpd_problem_set_abcd_out_of_core(self.h, int(bool(on))))

    def cpsd(self, diis_dim=10, diis_cyc=50, diis_max_err=1e-7):
        """Ground-state CPS(D) (CPSd.SF/run.f90): returns (T, info) with T the
        amplitudes of orders 1-3 as a (3, dpd_problem_cpsd_len) device tensor,
        each row (T1 | T2(a<b,i<j)) of the totally symmetric irrep, and info the
        correlation energies E^(2), E^(3), E^(4) and the DIIS iteration counts."""
        torch = _torch()
        n = int(lib().dpd_problem_cpsd_len(self.h))
        T = torch.empty(3 * n, dtype=torch.float64, device=f"cuda:{self.ctx.device}")
        o = _lib.CpsdOpts_()
        o.diis_dim, o.diis_cyc, o.diis_max_err = diis_dim, diis_cyc, diis_max_err
        check(lib().dpd_problem_cpsd(self.h, ctypes.byref(o), T.data_ptr()))
        return T.view(3, n), {"energy": list(o.energy), "iterations": list(o.iterations)}

    def grcpsd(self, diis_dim=10, diis_cyc=50, diis_max_err=1e-7):
        """Ground-state grCPS(D) (grCPSd.SF/run.f90): the CIS(Doo) Jacobian as
        zeroth order.  Returns (T, info) as cpsd() does; info["iterations"] holds
        the DIIS iterations of orders 1, 2, 3."""
        torch = _torch()
        n = int(lib().dpd_problem_cpsd_len(self.h))
        T = torch.empty(3 * n, dtype=torch.float64, device=f"cuda:{self.ctx.device}")
        o = _lib.CpsdOpts_()
        o.diis_dim, o.diis_cyc, o.diis_max_err = diis_dim, diis_cyc, diis_max_err
        it = np.zeros(3, dtype=np.int64)
        check(lib().dpd_problem_grcpsd(self.h, ctypes.byref(o), T.data_ptr(), it.ctypes.data_as(_lib.c_i64_p)))
        return T.view(3, n), {"energy": list(o.energy), "iterations": [int(x) for x in it]}

    def eecpsd(self, T, w0, L, diis_dim=10, diis_cyc=50, diis_max_err=1e-7):
        """EE-CPS(D) excitation energies (EE-CPSd.SF/run_order.f90) of CIS roots
        of this problem's symmetry: w0 (nroot CIS energies), L (nroot CIS vectors
        of length |R1|, device), T from cpsd().  Returns (w, iterations) with w of
        shape (nroot, 3): w^(1), w^(2), w^(3); the total is w0 + w^(2) + w^(3)."""
        torch = _torch()
        w0 = np.ascontiguousarray(np.atleast_1d(np.asarray(w0, dtype=np.float64)))
        nroot = len(w0)
        lay = self.spec.layout()
        l1 = self.ctx.irpdso(self.symR, lay["popv"], lay["popo"], 1)
        _check_vectors(L, "L", l1, self.ctx.device, nroot)
        n = int(lib().dpd_problem_cpsd_len(self.h))
        _check_vectors(T, "T", 3 * n, self.ctx.device, 1)
        o = _lib.CpsdOpts_()
        o.diis_dim, o.diis_cyc, o.diis_max_err = diis_dim, diis_cyc, diis_max_err
        w = np.zeros(3 * nroot)
        its = np.zeros(nroot, dtype=np.int64)
        check(lib().dpd_problem_eecpsd(self.h, ctypes.byref(o), T.data_ptr(), nroot,
                                       w0.ctypes.data_as(_lib.c_double_p), L.data_ptr(),
                                       w.ctypes.data_as(_lib.c_double_p), its.ctypes.data_as(_lib.c_i64_p)))
        return w.reshape(nroot, 3), its

    def close(self):
        # a problem lives in its context: once the context was closed explicitly
        # the library must not be asked to destroy it (its buffers stay until exit)
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            lib().dpd_problem_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
