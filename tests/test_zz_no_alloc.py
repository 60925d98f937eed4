"""dpd_xgemm_noalloc (the allocation-free variant of dpd_xgemm) against
dpd_xgemm, called through the C-ABI on two calls in a row, the second with a
smaller input.  FP64 and FP32 run the same gather kernel with the same tile in
both entry points (bit-identical); BF16x3 packs into the static scratch and
never splits K, so it agrees with dpd_xgemm to FP32-accumulation accuracy.
Also checks that the context's workspace stays unallocated, the refusal of a
call whose panels exceed the scratch, and the empty-C quick return."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PRECS = {"fp64": 0, "fp32": 1, "bf16x3": 2}
# (m, n, k, transa, transb): the first call, then a smaller one
CALLS = [(300, 170, 257, "T", "N"), (70, 45, 37, "N", "T")]


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _call(lib, fn, ctx, shape, A, lda, B, ldb, C, alpha=1.5, beta=-0.5, prec=0):
    m, n, k, ta, tb = shape
    return getattr(lib, fn)(ctx.h, ta.encode(), tb.encode(), m, n, k, alpha, A.data_ptr(), lda, B.data_ptr(), ldb,
                            beta, C.data_ptr(), m, prec)


@pytest.mark.parametrize("prec", list(PRECS))
def test_noalloc_matches_xgemm_two_calls(torch, prec):
    import dpdhip
    from dpdhip._lib import lib
    L = lib()
    ref_ctx = dpdhip.Context(4, 0)
    na_ctx = dpdhip.Context(4, 0)
    for call, shape in enumerate(CALLS):
        m, n, k, ta, tb = shape
        print(f"[no_alloc] {prec} call {call} m={m} n={n} k={k} {ta}{tb}", flush=True)
        rng = np.random.default_rng(100 * call + PRECS[prec])
        A = rng.standard_normal((m, k) if ta == "N" else (k, m))
        B = rng.standard_normal((k, n) if tb == "N" else (n, k))
        C = rng.standard_normal((m, n))
        gA, gB = _dev(torch, A.flatten(order="F")), _dev(torch, B.flatten(order="F"))
        c_ref, c_na = _dev(torch, C.flatten(order="F")), _dev(torch, C.flatten(order="F"))
        assert _call(L, "dpd_xgemm", ref_ctx, shape, gA, A.shape[0], gB, B.shape[0], c_ref, prec=PRECS[prec]) == 0
        rc = _call(L, "dpd_xgemm_noalloc", na_ctx, shape, gA, A.shape[0], gB, B.shape[0], c_na, prec=PRECS[prec])
        assert rc == 0, L.dpd_last_error().decode()
        torch.cuda.synchronize()
        got, ref = c_na.cpu().numpy(), c_ref.cpu().numpy()
        if prec == "bf16x3":
            opA = A if ta == "N" else A.T
            opB = B if tb == "N" else B.T
            mag = 1.5 * (np.abs(opA) @ np.abs(opB)).flatten(order="F")
            assert np.all(np.abs(got - ref) <= 2 * max(k, 32) * 2.0 ** -24 * mag + 1e-12), np.abs(got - ref).max()
        else:
            assert np.array_equal(got, ref), np.abs(got - ref).max()
        # nothing went through the context's grow-only workspace
        assert L.dpd_ctx_workspace_bytes(na_ctx.h) == 0
    ref_ctx.close()
    na_ctx.close()


def test_noalloc_refuses_oversized_and_returns_on_empty(torch):
    import dpdhip
    from dpdhip._lib import lib
    L = lib()
    ctx = dpdhip.Context(4, 0)
    m = n = k = 4096  # BF16x3 panels: 2 x 4096 x 4096 x 4 bytes = 128 MiB > 64 MiB
    A = torch.zeros(m * k, dtype=torch.float64, device="cuda:0")
    B = torch.zeros(k * n, dtype=torch.float64, device="cuda:0")
    C = torch.ones(m * n, dtype=torch.float64, device="cuda:0")
    print("[no_alloc] oversized bf16x3 call (refused on the host)", flush=True)
    rc = _call(L, "dpd_xgemm_noalloc", ctx, (m, n, k, "N", "N"), A, m, B, k, C, prec=2)
    assert rc != 0
    assert "scratch" in L.dpd_last_error().decode()
    assert float(C.min()) == 1.0 and float(C.max()) == 1.0  # C untouched
    print("[no_alloc] empty C", flush=True)
    assert _call(L, "dpd_xgemm_noalloc", ctx, (0, 5, 7, "N", "N"), A, 1, B, 7, C, prec=0) == 0
    ctx.close()
