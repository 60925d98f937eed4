"""GPU perturbation drivers (dpd_problem_cpsd, dpd_problem_grcpsd, dpd_problem_eecpsd) against the
oracle (oracle/oracle_perturb.py, pinned by a dense restatement) on the same
seeded Hamiltonians: ground-state CPS(D) amplitudes of orders 1-3 and the
energies E^(2), E^(3), E^(4); EE-CPS(D) w^(1..3) of the lowest spin-flip CIS
roots (found by the GPU Davidson on the SF-CIS operator and handed to both
sides).  FP64; the DIIS solves converge to 1e-12, so the two sides agree to
1e-10 (amplitudes relative to their largest element, energies absolute)."""
import numpy as np
import pytest

import oracle as O
import oracle_perturb as OP

pytestmark = pytest.mark.gpu

DIIS = dict(diis_dim=8, diis_cyc=200, diis_max_err=1e-12)
SYSTEMS = {
    "c2v-small": dict(nso=(4, 2, 2, 3), nalpha=4, nbeta=2, seed=3),
    "d2h-small": dict(nso=(2, 1, 1, 2, 1, 1, 1, 1), nalpha=3, nbeta=1, seed=11),
    "c2v-24": dict(nso=(11, 2, 4, 7), nalpha=6, nbeta=4, seed=1),
}


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


def _oracle(cfg, method, symR):
    from test_oracle_sigma_dense import _oracle as build
    return build(cfg, method, symR)[1]


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_cpsd_ground_state_matches_oracle(torch, name):
    import dpdhip
    cfg = SYSTEMS[name]
    print(f"[perturb] ground state {name}", flush=True)
    ctx = dpdhip.Context(len(cfg["nso"]), 0)
    prob = dpdhip.Problem(ctx, dpdhip.Spec(cfg["nso"], cfg["nalpha"], cfg["nbeta"], cfg["seed"]), "sf-cpsd", 1)
    T, info = prob.cpsd(**DIIS)
    op = _oracle(cfg, "sf-cpsd", 1)
    pt = OP.Perturb(op, **DIIS)
    gs = pt.ground_state()
    Th = T.cpu().numpy()
    l1 = pt.dT1.length()
    for n in (1, 2, 3):
        ref = np.concatenate([gs["T1"][n].arr, gs["T2"][n].arr])
        scale = max(np.abs(ref).max(), 1e-300)
        assert np.abs(Th[n - 1] - ref).max() <= 1e-10 * scale, (n, np.abs(Th[n - 1] - ref).max())
        assert Th[n - 1].shape[0] == l1 + pt.dT2.length()
    for k, e in zip((2, 3, 4), info["energy"]):
        assert e == pytest.approx(gs["E"][k], abs=1e-11), k
    print(f"[perturb] {name}: E2..E4 = {info['energy']}, DIIS iterations {info['iterations']}", flush=True)
    prob.close()
    ctx.close()


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_grcpsd_ground_state_matches_oracle(torch, name):
    """grCPS(D) (grCPSd.SF/run.f90): the CIS(Doo) Jacobian as zeroth order; the
    oracle's gr_ground_state is pinned by tests/test_oracle_grcpsd_dense.py"""
    import dpdhip
    cfg = SYSTEMS[name]
    print(f"[perturb] grCPS(D) ground state {name}", flush=True)
    ctx = dpdhip.Context(len(cfg["nso"]), 0)
    prob = dpdhip.Problem(ctx, dpdhip.Spec(cfg["nso"], cfg["nalpha"], cfg["nbeta"], cfg["seed"]), "sf-cpsd", 1)
    T, info = prob.grcpsd(**DIIS)
    op = _oracle(cfg, "sf-cpsd", 1)
    pt = OP.Perturb(op, **DIIS)
    gs = pt.gr_ground_state()
    Th = T.cpu().numpy()
    for n in (1, 2, 3):
        ref = np.concatenate([gs["T1"][n].arr, gs["T2"][n].arr])
        scale = max(np.abs(ref).max(), 1e-300)
        assert np.abs(Th[n - 1] - ref).max() <= 1e-10 * scale, (n, np.abs(Th[n - 1] - ref).max())
        assert abs(info["iterations"][n - 1] - gs["iterations"][n]) <= 1, (n, info["iterations"], gs["iterations"])
    for k, e in zip((2, 3, 4), info["energy"]):
        assert e == pytest.approx(gs["E"][k], abs=1e-11), k
    # T1^(1) is coupled in by the CIS(Doo) dressing (zero in CPS(D))
    assert np.abs(Th[0][: pt.dT1.length()]).max() > 0.0
    print(f"[perturb] {name}: grCPS(D) E2..E4 = {info['energy']}, DIIS iterations {info['iterations']}", flush=True)
    prob.close()
    ctx.close()


@pytest.mark.parametrize("name,symR", [("c2v-small", 1), ("c2v-small", 2), ("d2h-small", 3), ("c2v-24", 1)])
def test_eecpsd_matches_oracle(torch, name, symR):
    import dpdhip
    cfg = SYSTEMS[name]
    spec = dpdhip.Spec(cfg["nso"], cfg["nalpha"], cfg["nbeta"], cfg["seed"])
    ctx = dpdhip.Context(len(cfg["nso"]), 0)
    cis = dpdhip.Problem(ctx, spec, "sf-cis", symR)
    nroot = 2
    print(f"[perturb] CIS roots {name} symR={symR}", flush=True)
    w0, L, cinfo = cis.davidson(nroot, conv=1e-10, max_dim=20, max_cyc=100, lowest=True, final_check=False)
    assert cinfo["converged"]
    prob = dpdhip.Problem(ctx, spec, "sf-cpsd", symR)
    g1 = dpdhip.Problem(ctx, spec, "sf-cpsd", 1)
    T, _ = g1.cpsd(**DIIS)
    print(f"[perturb] EE-CPS(D) {name} symR={symR} w0={w0}", flush=True)
    w, its = prob.eecpsd(T.reshape(-1), w0, L, **DIIS)
    op = _oracle(cfg, "sf-cpsd", symR)
    op1 = _oracle(cfg, "sf-cpsd", 1)
    pt1 = OP.Perturb(op1, **DIIS)
    gs = pt1.ground_state()
    J = pt1.intermediates(gs)
    pt = OP.Perturb(op, **DIIS)
    Lh = L.cpu().numpy().reshape(nroot, -1)
    for r in range(nroot):
        ref = pt.excited_state(gs, J, float(w0[r]), Lh[r], symR)
        assert w[r, 0] == pytest.approx(0.0, abs=1e-14)
        for n in (2, 3):
            assert w[r, n - 1] == pytest.approx(ref["w"][n], abs=1e-10), (r, n)
        print(f"[perturb] root {r}: w0={w0[r]:.10f} w2={w[r, 1]:.3e} w3={w[r, 2]:.3e} "
              f"(oracle {ref['w'][2]:.3e}, {ref['w'][3]:.3e}) DIIS {its[r]}", flush=True)
    for p in (cis, prob, g1):
        p.close()
    ctx.close()


def test_perturb_needs_every_integral_block(torch):
    import dpdhip
    cfg = SYSTEMS["c2v-small"]
    ctx = dpdhip.Context(4, 0)
    cis = dpdhip.Problem(ctx, dpdhip.Spec(cfg["nso"], cfg["nalpha"], cfg["nbeta"], cfg["seed"]), "sf-cis", 1)
    with pytest.raises(dpdhip.DpdError):
        cis.cpsd()
    cis.close()
    ctx.close()
