"""The reference's on-disk formats through the C-ABI (csrc/dpd_io.cpp):

* node_* -- the pyscf2cc "interface" node file (pyscf2cc/mod_interface.f90),
* moio_* -- symmetry-packed moio lists and their ".c" control files
  (libsrint/clmoioso0.f, getallso_d.f),
* stream_* / eigvec_* -- IO_stream vector files (lib_cc/IO.f90:87-154) and the
  Davidson eigenvector file (lib_cc/mod_davidson.f90:388-390, :588-589).

Tensor data may be numpy arrays (host) or float64 CUDA tensors (device).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib
from ._lib import check, lib

_TYPES = {"i": (np.int64, b"i"), "d": (np.float64, b"d"), "s": (np.float32, b"s")}


def _ptr(x):
    if isinstance(x, np.ndarray):
        return x.ctypes.data_as(ctypes.c_void_p)
    return ctypes.c_void_p(x.data_ptr())


def _path(p):
    return os.fsencode(p)


def node_init(path):
    check(lib().dpd_node_init(_path(path)))


def node_write(path, name, data):
    """write an int64 / float64 / float32 array (or a str) as node `name`"""
    if isinstance(data, str):
        b = data.encode()
        buf = ctypes.create_string_buffer(b, len(b))
        check(lib().dpd_node_write(_path(path), name.encode(), b"c", ctypes.cast(buf, ctypes.c_void_p), len(b)))
        return
    a = np.asarray(data)
    kind = "i" if a.dtype.kind in "iu" else ("s" if a.dtype == np.float32 else "d")
    a = np.ascontiguousarray(a.ravel(order="F"), dtype=_TYPES[kind][0])
    check(lib().dpd_node_write(_path(path), name.encode(), _TYPES[kind][1], _ptr(a), a.size))


def node_size(path, name):
    """bytes held by node `name` (-1: absent)"""
    return int(lib().dpd_node_size(_path(path), name.encode()))


def node_read(path, name, kind="d", shape=None):
    """node `name` as a numpy array of int64 ('i'), float64 ('d'), float32 ('s')
    or a str ('c'); 2-d shapes are Fortran ordered as the reference writes them"""
    n = node_size(path, name)
    if n < 0:
        raise KeyError(name)
    if kind == "c":
        buf = ctypes.create_string_buffer(n)
        check(lib().dpd_node_read(_path(path), name.encode(), b"c", ctypes.cast(buf, ctypes.c_void_p), n))
        return buf.raw.decode()
    dt, code = _TYPES[kind]
    a = np.zeros(n // np.dtype(dt).itemsize, dtype=dt)
    check(lib().dpd_node_read(_path(path), name.encode(), code, _ptr(a), a.size))
    return a.reshape(shape, order="F") if shape is not None else a


def _pops(p):
    a = np.zeros(8, dtype=np.int64)
    p = np.asarray(p, dtype=np.int64).ravel()
    a[: p.size] = p
    return a


def moio_write(ctx, path, irrepx, popp, popq, ioppq, popr, pops, ioprs, data):
    ps = [_pops(p) for p in (popp, popq, popr, pops)]
    cp = [p.ctypes.data_as(_lib.c_i64_p) for p in ps]
    check(lib().dpd_moio_write(ctx.h, _path(path), irrepx, cp[0], cp[1], ioppq, cp[2], cp[3], ioprs, _ptr(data)))


def moio_info(path):
    irr = np.zeros(1, dtype=np.int64)
    pops = np.zeros(32, dtype=np.int64)
    iop = np.zeros(2, dtype=np.int64)
    n = np.zeros(1, dtype=np.int64)
    check(lib().dpd_moio_info(_path(path), *(a.ctypes.data_as(_lib.c_i64_p) for a in (irr, pops, iop, n))))
    return {"irrepx": int(irr[0]), "pops": pops.reshape(4, 8), "iop": (int(iop[0]), int(iop[1])),
            "length": int(n[0])}


def moio_read(path, out=None):
    """the whole list (getallso): into `out` (numpy or CUDA float64) or a new numpy array"""
    n = moio_info(path)["length"]
    if out is None:
        out = np.zeros(n)
    check(lib().dpd_moio_read(_path(path), _ptr(out), n))
    return out


def stream_write(path, idx, val=0.0, vec=None, mode=1, replace=False):
    n = 0 if vec is None else (vec.size if isinstance(vec, np.ndarray) else vec.numel())
    p = _ptr(vec) if vec is not None else None
    check(lib().dpd_stream_write(_path(path), idx, float(val), p, n, mode, int(replace)))


def stream_read(path, idx, n=0, mode=1, out=None):
    val = ctypes.c_double(0.0)
    if out is None and mode != 2:
        out = np.zeros(n)
    check(lib().dpd_stream_read(_path(path), idx, ctypes.byref(val), _ptr(out) if mode != 2 else None, n, mode))
    return val.value, out


def eigvec_write(path, w, R, dim, istate=1, replace=True):
    w = np.ascontiguousarray(np.atleast_1d(np.asarray(w, dtype=np.float64)))
    check(lib().dpd_eigvec_write(_path(path), istate, len(w), dim, w.ctypes.data_as(_lib.c_double_p), _ptr(R),
                                 int(replace)))


def eigvec_read(path, nroot, dim, istate=1, out=None):
    w = np.zeros(nroot)
    if out is None:
        out = np.zeros(nroot * dim)
    check(lib().dpd_eigvec_read(_path(path), istate, nroot, dim, w.ctypes.data_as(_lib.c_double_p), _ptr(out)))
    return w, out


# ---- a problem's integrals as the reference's moio lists ------------------------------
# file names of EE-CPSd.SF/mod_variable.f90:31-47 (CISdoo.SD.SF the same)
INTEGRAL_FILES = {"ijkl": "ijkl_ns_", "ijka": "ijka_ns_", "abij": "abij_ns_", "iabj": "iabj_ns_",
                  "abci": "abci_ns_", "abcd": "abcd_ns_"}
# their layouts (allocate_mem.f90 with iop = 0): spaces of the four indices, (ioppq, ioprs)
_LAYOUT = {"ijkl": ("oooo", (0, 0)), "ijka": ("ooov", (0, 1)), "abij": ("vvoo", (0, 0)),
           "iabj": ("vovo", (1, 1)), "abci": ("vvvo", (0, 1)), "abcd": ("vvvv", (0, 0))}


def _problem_pops(prob, name):
    lay = prob.spec.layout()
    sp, iop = _LAYOUT[name]
    return [lay["popo"] if c == "o" else lay["popv"] for c in sp], iop


def save_integrals(prob, directory):
    """every integral block of the problem as a moio list under its reference name"""
    os.makedirs(directory, exist_ok=True)
    for name, fname in INTEGRAL_FILES.items():
        if prob.tensor_len(name) <= 0:
            continue
        pops, iop = _problem_pops(prob, name)
        x = prob.get(name)
        moio_write(prob.ctx, os.path.join(directory, fname), 1, pops[0], pops[1], iop[0], pops[2], pops[3], iop[1], x)


def load_integrals(prob, directory):
    """read the moio lists into the problem (layouts must match) and refresh it"""
    import torch
    for name, fname in INTEGRAL_FILES.items():
        path = os.path.join(directory, fname)
        n = prob.tensor_len(name)
        if n <= 0 or not os.path.exists(path):
            continue
        info = moio_info(path)
        pops, iop = _problem_pops(prob, name)
        if info["irrepx"] != 1 or info["iop"] != tuple(iop) or info["length"] != n or \
                any(not np.array_equal(info["pops"][k], _pops(pops[k])) for k in range(4)):
            raise ValueError(f"{path}: layout differs from the problem's {name}")
        x = torch.empty(n, dtype=torch.float64, device=f"cuda:{prob.ctx.device}")
        moio_read(path, x)
        prob.set(name, x)
    prob.refresh()


def write_amplitudes(path, T):
    """the f_T file of CPSd.SF (IO_stream, order n at element (n-1)*len + 1)"""
    n = T.shape[-1]
    for k in range(T.shape[0]):
        stream_write(path, k * n + 1, vec=T[k].contiguous() if hasattr(T, "contiguous") else T[k], mode=3,
                     replace=(k == 0))


def read_amplitudes(path, n, orders=3, out=None):
    import torch
    if out is None:
        out = torch.empty(orders * n, dtype=torch.float64, device="cuda:0")
    for k in range(orders):
        stream_read(path, k * n + 1, n=n, mode=3, out=out[k * n:(k + 1) * n])
    return out.view(orders, n)
