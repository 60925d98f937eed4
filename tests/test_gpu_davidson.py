"""GPU Davidson against the oracle's restatement of lib_cc/mod_davidson.f90 on the
same seeded Hamiltonians and guesses: same iteration count, eigenvalues within
1e-8 Eh (the FP64 bar), and the converged roots are eigenvalues of the dense
operator (numpy eig of J built column by column from the oracle sigma)."""
import numpy as np
import pytest

import oracle as O
import oracle_sigma as OS

pytestmark = pytest.mark.gpu

C2V24 = dict(nso=(11, 2, 4, 7), nalpha=6, nbeta=4, seed=1)
TINY = dict(nso=(4, 1, 2, 3), nalpha=3, nbeta=1, seed=3)


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


def make(cfg, method, symR):
    import dpdhip
    from test_gpu_sigma import oracle_problem
    sp, op = oracle_problem(cfg, method, symR)
    gp = dpdhip.Problem(ctx(len(cfg["nso"])), sp, method, symR)
    return sp, op, gp


@pytest.mark.parametrize("method,max_dim", [("sf-cis", 20), ("cisdoo", 20), ("sf-cpsd", 20), ("sf-cpsd", 12)],
                         ids=["sf-cis", "cisdoo", "sf-cpsd", "sf-cpsd-collapse"])
def test_davidson_matches_oracle(torch, method, max_dim):
    sp, op, gp = make(C2V24, method, 1)
    nroot = 3
    guesses, _ = op.guess(nroot)
    ref = OS.davidson(op.sigma, op.diag(), guesses, nroot, conv=1e-7, max_dim=max_dim, max_cyc=60)
    w, R, info = gp.davidson(nroot, conv=1e-7, max_dim=max_dim, max_cyc=60)
    assert ref["converged"] and info["converged"]
    assert info["iterations"] == ref["iterations"], (info["iterations"], ref["iterations"])
    assert np.max(np.abs(w - ref["w"])) <= 1e-8, (w, ref["w"])
    assert max(info["final_residual"]) < 1e-6
    # eigenvectors agree up to sign
    Rg = R.cpu().numpy().reshape(nroot, -1)
    for k in range(nroot):
        s = np.sign(Rg[k] @ ref["R"][k])
        assert np.max(np.abs(s * Rg[k] - ref["R"][k])) < 1e-6


@pytest.mark.parametrize("method", ["sf-cis", "cisdoo", "sf-cpsd"])
def test_davidson_roots_are_dense_eigenvalues(torch, method):
    """Tiny system: build J densely from the oracle sigma, its exact spectrum
    must contain every converged Davidson root (1e-8)."""
    sp, op, gp = make(TINY, method, 1)
    n = op.dim
    J = np.stack([op.sigma(np.eye(n)[k]) for k in range(n)], axis=1)
    ev = np.linalg.eigvals(J)
    w, R, info = gp.davidson(2, conv=1e-9, max_dim=12, max_cyc=100)
    assert info["converged"]
    for x in w:
        assert np.min(np.abs(ev - x)) < 1e-8, (x, np.sort(ev.real)[:6])
    if method == "sf-cis":
        # the SF-CIS block is symmetric in this spin-orbital basis
        assert np.allclose(J, J.T, atol=1e-12)


def test_davidson_not_converged_path(torch):
    """max_dim=10 with 3 roots collapses to 8 vectors every other iteration and
    does not converge in 60 cycles: both report it (mod_davidson.f90:168-178)
    after the same number of iterations with the same eigenvalues."""
    sp, op, gp = make(C2V24, "sf-cpsd", 1)
    guesses, _ = op.guess(3)
    ref = OS.davidson(op.sigma, op.diag(), guesses, 3, conv=1e-7, max_dim=10, max_cyc=60)
    w, R, info = gp.davidson(3, conv=1e-7, max_dim=10, max_cyc=60)
    assert not ref["converged"] and not info["converged"]
    assert info["iterations"] == ref["iterations"] == 61
    assert np.max(np.abs(w - ref["w"])) <= 1e-8


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_mixed_precision_davidson_energies(torch, prec):
    """mixed-precision contractions (FP64 vectors, orthogonalisation and
    subspace) converged to 1e-5, then FP64 Rayleigh quotients of the vectors:
    energies within 1e-6 Eh of the FP64 Davidson at 1e-7"""
    sp, op, gp = make(C2V24, "sf-cpsd", 1)
    nroot = 3
    w64, _, info64 = gp.davidson(nroot, conv=1e-7, max_dim=20, max_cyc=60)
    gp.set_precision(prec)
    try:
        wm, R, infom = gp.davidson(nroot, conv=1e-5, max_dim=20, max_cyc=60)
    finally:
        gp.set_precision("fp64")
    assert info64["converged"] and infom["converged"]
    S = gp.sigma(R).view(nroot, -1)
    Rm = R.view(nroot, -1)
    wr = ((Rm * S).sum(1) / (Rm * Rm).sum(1)).cpu().numpy()
    assert np.max(np.abs(wr - w64)) < 1e-6, (wr, w64)


@pytest.mark.parametrize("cfg,method,nroot,max_dim", [
    (dict(nso=(4, 1, 2, 3), nalpha=3, nbeta=1, seed=3), "sf-cis", 3, 12),
    (dict(nso=(3, 1, 1, 2), nalpha=2, nbeta=1, seed=5), "sf-cis", 3, 12),
    (dict(nso=(3, 1, 1, 2), nalpha=2, nbeta=1, seed=5), "sf-cpsd", 4, 20)], ids=["tiny-cis", "small-cis", "small-cpsd"])
def test_davidson_skipped_directions(torch, cfg, method, nroot, max_dim):
    """subspaces that fill the space: directions with remaining norm < 1e-6
    are dropped (:437) -- the batched orthogonalisation redoes the roots after a
    dropped one; the same directions are dropped as in the oracle"""
    sp, op, gp = make(cfg, method, 1)
    guesses, _ = op.guess(nroot)
    ref = OS.davidson(op.sigma, op.diag(), guesses, nroot, conv=1e-9, max_dim=max_dim, max_cyc=60)
    w, R, info = gp.davidson(nroot, conv=1e-9, max_dim=max_dim, max_cyc=60)
    assert ref["skipped"] > 0
    assert info["skipped"] == ref["skipped"], (info["skipped"], ref["skipped"])
    assert info["iterations"] == ref["iterations"]
    assert info["converged"] == ref["converged"]
    assert np.max(np.abs(w - ref["w"])) <= 1e-8, (w, ref["w"])


def test_davidson_collapse_needs_room(torch):
    """a collapse keeps nroot + 5 vectors: davidson_dim below that is rejected
    (the reference reads past its subspace there)"""
    import dpdhip
    cfg = dict(nso=(4, 1, 2, 3), nalpha=3, nbeta=1, seed=3)
    sp, op, gp = make(cfg, "sf-cis", 1)
    with pytest.raises(ValueError):
        OS.davidson(op.sigma, op.diag(), op.guess(4)[0], 4, conv=1e-9, max_dim=8, max_cyc=60)
    with pytest.raises(dpdhip._lib.DpdError, match="nroot \\+ 5"):
        gp.davidson(4, conv=1e-9, max_dim=8, max_cyc=60)


def test_block_davidson_16_roots(torch):
    """block Davidson with 16 roots (the 16 lowest, davidson_dim 40): all new
    directions of an iteration go through one batched sigma; iteration count
    and roots as in the oracle"""
    cfg = dict(nso=(8, 2, 3, 5), nalpha=5, nbeta=3, seed=11)  # dim 2674
    sp, op, gp = make(cfg, "sf-cpsd", 1)
    nroot = 16
    guesses, _ = op.guess(nroot)
    ref = OS.davidson(op.sigma, op.diag(), guesses, nroot, conv=1e-7, max_dim=40, max_cyc=150, lowest=True)
    w, R, info = gp.davidson(nroot, conv=1e-7, max_dim=40, max_cyc=150, lowest=True)
    assert ref["converged"] and info["converged"]
    assert info["iterations"] == ref["iterations"]
    assert np.max(np.abs(w - ref["w"])) <= 1e-8
    assert max(info["final_residual"]) < 1e-6


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_davidson_graph_replay_equals_eager(torch, prec):
    """graph=True replays each iteration's sigma step from a hipGraph: the same
    vectors go through the same kernels, so eigenvalues, iteration count and
    eigenvectors equal the eager solve's, and most steps are replays"""
    sp, op, gp = make(C2V24, "sf-cpsd", 1)
    gp.set_precision(prec)
    try:
        kw = dict(conv=1e-7 if prec == "fp64" else 1e-5, max_dim=20, max_cyc=100, lowest=True, final_check=False)
        g, _ = gp.guess(3)
        w0, R0, i0 = gp.davidson(3, R=g.clone(), **kw)
        print(f"[graph] {prec} eager {i0['iterations']} iterations", flush=True)
        w1, R1, i1 = gp.davidson(3, R=g.clone(), graph=True, **kw)
        print(f"[graph] {prec} graph {i1['iterations']} iterations, {i1['graph_replays']} replays", flush=True)
    finally:
        gp.set_precision("fp64")
    assert i0["converged"] and i1["converged"]
    assert i1["iterations"] == i0["iterations"]
    assert np.array_equal(w1, w0)
    assert torch.equal(R1, R0)
    assert i1["graph_replays"] >= i1["iterations"] // 2
    gp.close()
