"""GPU parity of the sigma builders, diagonal and initial guesses against the
oracle on the same seeded synthetic Hamiltonian (C2v 24 orbitals [11,2,4,7],
n_alpha=6, n_beta=4, and a small D2h system), FP64: relative 1e-10."""
import numpy as np
import pytest

import oracle as O
import oracle_sigma as OS

pytestmark = pytest.mark.gpu

C2V24 = dict(nso=(11, 2, 4, 7), nalpha=6, nbeta=4, seed=1)
SMALL_D2H = dict(nso=(5, 2, 3, 3, 4, 1, 2, 2), nalpha=5, nbeta=3, seed=5)


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


_ctx = {}


def ctx(n):
    import dpdhip
    if n not in _ctx:
        _ctx[n] = dpdhip.Context(n, 0)
    return _ctx[n]


def oracle_problem(cfg, method, symR):
    import dpdhip
    sp = dpdhip.Spec(cfg["nso"], cfg["nalpha"], cfg["nbeta"], cfg["seed"])
    lay = sp.layout()
    n = len(cfg["nso"])
    s = O.Sym(n)
    s.activate()
    v, o = lay["popv"], lay["popo"]
    Z = O.ZORB
    descs = {"iabj": O.Desc([v, o, v, o], (1, 1), 1), "ijka": O.Desc([o, o, o, v], (0, 1), 1),
             "ijkl": O.Desc([o, o, o, o], (0, 0), 1), "abij": O.Desc([v, v, o, o], (0, 0), 1),
             "abci": O.Desc([v, v, v, o], (0, 1), 1), "abcd": O.Desc([v, v, v, v], (0, 0), 1)}
    ints = {k: sp.fill(k, 0, d.length()) for k, d in descs.items()}
    eorb = sp.fill("eorb", 0, int(o.sum() + v.sum()))
    return sp, OS.OracleProblem(s, o, v, lay["popa"], lay["popb"], lay["vrta"], lay["vrtb"], eorb, ints, method,
                                symR)


@pytest.mark.parametrize("cfg,symR", [(C2V24, 1), (C2V24, 2), (SMALL_D2H, 1), (SMALL_D2H, 5)],
                         ids=["c2v-A1", "c2v-A2", "d2h-Ag", "d2h-B1u"])
@pytest.mark.parametrize("method", ["sf-cis", "cisdoo", "sf-cpsd"])
def test_sigma_matches_oracle(torch, cfg, symR, method):
    import dpdhip
    sp, op = oracle_problem(cfg, method, symR)
    gp = dpdhip.Problem(ctx(len(cfg["nso"])), sp, method, symR)
    assert gp.dim == op.dim
    np.testing.assert_allclose(gp.diag().cpu().numpy(), op.diag(), rtol=0, atol=1e-15)
    rng = np.random.default_rng(42)
    nvec = 3
    R = rng.standard_normal((nvec, op.dim))
    got = gp.sigma(torch.from_numpy(R.reshape(-1).copy()).to("cuda:0")).cpu().numpy().reshape(nvec, op.dim)
    for v in range(nvec):
        ref = op.sigma(R[v])
        np.testing.assert_allclose(got[v], ref, rtol=1e-10, atol=1e-10 * np.abs(ref).max())
    # a single vector gives the same result as inside a batch
    one = gp.sigma(torch.from_numpy(R[1].copy()).to("cuda:0")).cpu().numpy()
    np.testing.assert_allclose(one, got[1], rtol=1e-12, atol=1e-13)
    # spin-flip guesses (preDavidson.f90:120-223)
    gR, ge = gp.guess(3)
    oR, oe = op.guess(3)
    assert np.array_equal(gR.cpu().numpy().reshape(3, -1), oR)
    np.testing.assert_allclose(ge, oe, rtol=0, atol=0)


@pytest.mark.parametrize("prec,bound", [("fp32", 1e-5), ("bf16x3", 2e-4)])
def test_sigma_mixed_close_to_fp64(torch, prec, bound):
    import dpdhip
    sp = dpdhip.Spec(C2V24["nso"], C2V24["nalpha"], C2V24["nbeta"], C2V24["seed"])
    gp = dpdhip.Problem(ctx(4), sp, "sf-cpsd", 1)
    rng = np.random.default_rng(3)
    R = torch.from_numpy(rng.standard_normal(2 * gp.dim)).to("cuda:0")
    s64 = gp.sigma(R).cpu().numpy()
    gp.set_precision(prec)
    s32 = gp.sigma(R).cpu().numpy()
    err = np.abs(s32 - s64).max() / np.abs(s64).max()
    assert 1e-12 < err < bound, err


@pytest.mark.parametrize("packing", [0, 2])
def test_sigma_kernel_paths_agree(torch, packing):
    """the same FP64 sigma through the gathering kernel (packing 0) and with
    every operand packed (packing 2) as through the default path"""
    import dpdhip
    sp = dpdhip.Spec(C2V24["nso"], C2V24["nalpha"], C2V24["nbeta"], C2V24["seed"])
    c = ctx(4)
    gp = dpdhip.Problem(c, sp, "sf-cpsd", 1)
    rng = np.random.default_rng(4)
    R = torch.from_numpy(rng.standard_normal(3 * gp.dim)).to("cuda:0")
    ref = gp.sigma(R).cpu().numpy()
    try:
        c.set_gemm_config(-1, packing)
        got = gp.sigma(R).cpu().numpy()
    finally:
        c.set_gemm_config(-1, 1)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_sigma_split_k_off_with_autotuned_tiles(torch, prec):
    """split-K off together with the autotuned tile shape: the tuner packs the
    operand panels in whichever of its passes runs first, and grouped launches
    honour the setting; one problem and a group of two equal the results with
    split-K automatic."""
    import dpdhip
    sp1 = dpdhip.Spec(SMALL_D2H["nso"], SMALL_D2H["nalpha"], SMALL_D2H["nbeta"], SMALL_D2H["seed"])
    sp2 = dpdhip.Spec((4, 3, 3, 2, 3, 2, 2, 1), 4, 2, 9)
    out = {}
    for mode in (-1, 0):
        print(f"[split-k] prec={prec} mode={mode}", flush=True)
        c = dpdhip.Context(8, 0)
        c.set_gemm_config(-1, 1)
        c.set_split_k(mode)
        p1 = dpdhip.Problem(c, sp1, "sf-cpsd", 1)
        p2 = dpdhip.Problem(c, sp2, "sf-cpsd", 1)
        for p in (p1, p2):
            p.set_precision(prec)
        rng = np.random.default_rng(5)
        R1 = torch.from_numpy(rng.standard_normal(3 * p1.dim)).cuda()
        R2 = torch.from_numpy(rng.standard_normal(2 * p2.dim)).cuda()
        single = p1.sigma(R1).cpu().numpy()
        grouped = [x.cpu().numpy() for x in dpdhip.sigma_group([p1, p2], [R1, R2])]
        out[mode] = (single, grouped)
        p1.close()
        p2.close()
        c.close()
    tol = 1e-11 if prec == "fp64" else 1e-5
    ref_single, ref_grouped = out[-1]
    got_single, got_grouped = out[0]
    scale = np.abs(ref_single).max()
    np.testing.assert_allclose(got_single, ref_single, rtol=0, atol=tol * scale)
    for g, r in zip(got_grouped, ref_grouped):
        np.testing.assert_allclose(g, r, rtol=0, atol=tol * np.abs(r).max())
    np.testing.assert_allclose(got_grouped[0], ref_single, rtol=0, atol=tol * scale)


@pytest.mark.parametrize("split", [0, -1])
def test_bf16x3_direct_a_tile_equals_staged_tile(torch, split):
    """The BF16x3 tile that loads A straight into registers (GEMM config 7,
    128x64) accumulates every output element over the same K steps in the same
    MFMA order as the LDS-staged 128x64 tile (config 2): with split-K off the
    sigma vectors are bit-identical, single and grouped; with split-K automatic
    they agree to FP32 accumulation accuracy."""
    import dpdhip
    sp1 = dpdhip.Spec(SMALL_D2H["nso"], SMALL_D2H["nalpha"], SMALL_D2H["nbeta"], SMALL_D2H["seed"])
    sp2 = dpdhip.Spec((4, 3, 3, 2, 3, 2, 2, 1), 4, 2, 9)
    out = {}
    for cfg in (2, 7):
        print(f"[direct-a] split={split} cfg={cfg}", flush=True)
        c = dpdhip.Context(8, 0)
        c.set_gemm_config(cfg, 1)
        c.set_split_k(split)
        p1 = dpdhip.Problem(c, sp1, "sf-cpsd", 1)
        p2 = dpdhip.Problem(c, sp2, "sf-cpsd", 1)
        for p in (p1, p2):
            p.set_precision("bf16x3")
        rng = np.random.default_rng(11)
        R1 = torch.from_numpy(rng.standard_normal(5 * p1.dim)).cuda()
        R2 = torch.from_numpy(rng.standard_normal(2 * p2.dim)).cuda()
        single = p1.sigma(R1).cpu().numpy()
        grouped = [x.cpu().numpy() for x in dpdhip.sigma_group([p1, p2], [R1, R2])]
        out[cfg] = (single, grouped)
        p1.close()
        p2.close()
        c.close()
    (s2, g2), (s7, g7) = out[2], out[7]
    if split == 0:
        assert np.array_equal(s7, s2)
        for a, b in zip(g7, g2):
            assert np.array_equal(a, b)
    else:
        np.testing.assert_allclose(s7, s2, rtol=0, atol=1e-5 * np.abs(s2).max())
        for a, b in zip(g7, g2):
            np.testing.assert_allclose(a, b, rtol=0, atol=1e-5 * np.abs(b).max())
