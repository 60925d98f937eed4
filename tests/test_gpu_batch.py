"""Batches of independent Hamiltonians: grouped sigma (one GEMM launch per
contraction for all problems) and the lockstep Davidson with and without the
hipGraph replay, against the same problems solved one at a time (which the
other GPU tests pin to the oracle).  FP64: relative 1e-11 for sigma, 1e-9 Eh
and equal iteration counts for the solves; FP32 + FP64 Rayleigh quotients
within 1e-6 Eh."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def batch_specs(n, seed0=100, lo=10, hi=20):
    """seeded C2v Hamiltonians: lo..hi spatial orbitals with a random irrep
    split, n_alpha in 3..5, n_beta = n_alpha - 2"""
    import dpdhip
    out = []
    for s in range(seed0, seed0 + n):
        rng = np.random.default_rng(s)
        norb = int(rng.integers(lo, hi + 1))
        w = rng.dirichlet([3.0] * 4)
        nso = tuple(int(x) + 1 for x in rng.multinomial(norb - 4, w))
        na = int(rng.integers(3, 6))
        out.append(dpdhip.Spec(nso, na, na - 2, s))
    return out


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


@pytest.fixture(scope="module")
def ctx(torch):
    import dpdhip
    return dpdhip.Context(4, 0)


CFGS = [(-1, 1), (1, 2), (2, 2), (3, 2), (4, 2), (5, 2), (6, 2)]


@pytest.mark.parametrize("cfg", CFGS, ids=["auto", "pk64x64", "pk128x64", "pk64x128", "pk128x128", "pk32x32",
                                           "pk64x192"])
@pytest.mark.parametrize("method", ["cisdoo", "sf-cpsd"])
@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_sigma_group_matches_single(torch, ctx, cfg, method, prec):
    import dpdhip
    ctx.set_gemm_config(*cfg)
    try:
        probs = [dpdhip.Problem(ctx, sp, method, sym) for sp, sym in zip(batch_specs(5), (1, 2, 1, 3, 4))]
        for p in probs:
            p.set_precision(prec)
        print(f"[batch] sigma {method} {prec} cfg={cfg} dims={[p.dim for p in probs]}", flush=True)
        g = torch.Generator(device="cuda:0").manual_seed(7)
        Rs = [torch.randn(3 * p.dim, dtype=torch.float64, device="cuda:0", generator=g) for p in probs]
        single = [p.sigma(R) for p, R in zip(probs, Rs)]
        for rep in range(2):
            grouped = dpdhip.sigma_group(probs, Rs)
            for a, b in zip(grouped, single):
                tol = {"fp64": 1e-11, "fp32": 1e-5, "bf16x3": 1e-4}[prec]
                assert float((a - b).abs().max()) <= tol * float(b.abs().max()), (rep, prec)
    finally:
        ctx.set_gemm_config(-1, 1)


@pytest.mark.parametrize("mode", ["grouped+graph", "grouped", "sequential"])
def test_davidson_batch_matches_single(torch, ctx, mode):
    import dpdhip
    probs = [dpdhip.Problem(ctx, sp, "sf-cpsd", 1) for sp in batch_specs(4)]
    print(f"[batch] davidson {mode} dims={[p.dim for p in probs]}", flush=True)
    ref = [p.davidson(3, conv=1e-8, max_dim=24, max_cyc=100, lowest=True) for p in probs]
    ws, Rs, infos = dpdhip.davidson_batch(probs, 3, conv=1e-8, max_dim=24, max_cyc=100, lowest=True,
                                          grouped=mode != "sequential", graph=mode == "grouped+graph")
    print(f"[batch] {infos[0]['batch']}", flush=True)
    if mode == "grouped+graph":
        assert infos[0]["batch"]["graph_replays"] > 0 and infos[0]["batch"]["captures_abandoned"] == 0
    for (w0, R0, i0), w, i in zip(ref, ws, infos):
        assert i0["converged"] and i["converged"]
        assert i["iterations"] == i0["iterations"]
        assert np.max(np.abs(w - w0)) <= 1e-9, (w, w0)
        assert max(i["final_residual"]) < 1e-6


def test_davidson_batch_fp32_refined(torch, ctx):
    import dpdhip
    probs = [dpdhip.Problem(ctx, sp, "sf-cpsd", 1) for sp in batch_specs(6, seed0=300)]
    ref = [p.davidson(3, conv=1e-8, max_dim=24, max_cyc=100, lowest=True)[0] for p in probs]
    for p in probs:
        p.set_precision("fp32")
    ws, Rs, infos = dpdhip.davidson_batch(probs, 3, conv=1e-5, max_dim=24, max_cyc=100, lowest=True,
                                          final_check=False)
    for p in probs:
        p.set_precision("fp64")
    S = dpdhip.sigma_group(probs, Rs)
    for p, R, Sv, w0, i in zip(probs, Rs, S, ref, infos):
        assert i["converged"]
        Rm, Sm = R.view(3, -1), Sv.view(3, -1)
        w = ((Rm * Sm).sum(1) / (Rm * Rm).sum(1)).cpu().numpy()
        assert np.max(np.abs(np.sort(w) - np.sort(w0))) <= 1e-6, (w, w0)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_sigma_group_same_shape_members(torch, ctx, prec):
    """members with the same orbital structure (one shared plan per contraction)
    but different integrals: each keeps its own packed integral panels"""
    import dpdhip
    base = batch_specs(3, seed0=400)
    specs = base + [dpdhip.Spec(sp.nso, sp.nalpha, sp.nbeta, sp.seed + 1000) for sp in base[:2]]
    probs = [dpdhip.Problem(ctx, sp, "sf-cpsd", 1) for sp in specs]
    for p in probs:
        p.set_precision(prec)
    print(f"[batch] same-shape sigma {prec} dims={[p.dim for p in probs]}", flush=True)
    g = torch.Generator(device="cuda:0").manual_seed(11)
    Rs = [torch.randn(2 * p.dim, dtype=torch.float64, device="cuda:0", generator=g) for p in probs]
    single = [p.sigma(R) for p, R in zip(probs, Rs)]
    tol = {"fp64": 1e-11, "fp32": 1e-5, "bf16x3": 1e-4}[prec]
    for rep in range(2):
        grouped = dpdhip.sigma_group(probs, Rs)
        for i, (a, b) in enumerate(zip(grouped, single)):
            assert float((a - b).abs().max()) <= tol * float(b.abs().max()), (rep, i)


def test_davidson_batch_graph_shrinking_member_set(torch, ctx):
    """8 members that converge at different iterations: the captured grouped
    sigma covers every root of its members, so it is replayed while members
    converge (a smaller member set is captured once the requests fall below
    a quarter of the covered work); every member reproduces its one-at-a-time solve"""
    import dpdhip
    probs = [dpdhip.Problem(ctx, sp, "sf-cpsd", 1) for sp in batch_specs(8, seed0=500, lo=10, hi=24)]
    print(f"[batch] shrinking set dims={[p.dim for p in probs]}", flush=True)
    ref = [p.davidson(3, conv=1e-8, max_dim=24, max_cyc=100, lowest=True) for p in probs]
    ws, Rs, infos = dpdhip.davidson_batch(probs, 3, conv=1e-8, max_dim=24, max_cyc=100, lowest=True, grouped=True,
                                          graph=True)
    print(f"[batch] {infos[0]['batch']} iterations={[i['iterations'] for i in infos]}", flush=True)
    assert len({i["iterations"] for i in infos}) > 1  # the member set does shrink
    # a graph covers all roots of its members, so converging members keep it alive
    b = infos[0]["batch"]
    assert b["captures_abandoned"] == 0 and b["graph_replays"] >= 0.6 * b["steps"], b
    for (w0, R0, i0), w, i in zip(ref, ws, infos):
        assert i0["converged"] and i["converged"]
        assert i["iterations"] == i0["iterations"]
        assert np.max(np.abs(w - w0)) <= 1e-9, (w, w0)
