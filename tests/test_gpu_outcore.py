"""Out-of-core <ab||ef> (dpd_problem_set_abcd_out_of_core; the reference's
c2_r2abcd.f90:29-104 out-of-core ladder): with abcd streamed from pinned host
memory block by block, sigma equals the in-core sigma (FP64, relative 1e-12:
only the GEMM kernel and so the summation order differ), the Davidson roots and
the CPS(D) / EE-CPS(D) energies equal the in-core ones, and switching back
restores the device copy bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SYSTEMS = [dict(nso=(5, 2, 3, 3, 4, 1, 2, 2), nalpha=5, nbeta=3, seed=5, n=8),
           dict(nso=(11, 2, 4, 7), nalpha=6, nbeta=4, seed=1, n=4)]


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


@pytest.mark.parametrize("cfg", SYSTEMS, ids=["d2h", "c2v-24"])
@pytest.mark.parametrize("symR", [1, 2])
def test_sigma_out_of_core_equals_in_core(torch, cfg, symR):
    import dpdhip
    ctx = dpdhip.Context(cfg["n"], 0)
    p = dpdhip.Problem(ctx, dpdhip.Spec(cfg["nso"], cfg["nalpha"], cfg["nbeta"], cfg["seed"]), "sf-cpsd", symR)
    R = torch.randn(3 * p.dim, dtype=torch.float64, device="cuda:0", generator=torch.Generator("cuda:0").manual_seed(7))
    abcd = p.get("abcd").clone()
    ref = p.sigma(R)
    print(f"[ooc] sigma nirrep={cfg['n']} symR={symR}", flush=True)
    p.set_abcd_out_of_core(True)
    got = p.sigma(R)
    scale = float(ref.abs().max())
    assert float((got - ref).abs().max()) <= 1e-12 * scale
    with pytest.raises(dpdhip.DpdError):
        p.get("abcd")
    p.set_precision("fp32")
    g32 = p.sigma(R)
    assert float((g32 - ref).abs().max()) <= 1e-5 * scale
    p.set_precision("fp64")
    p.set_abcd_out_of_core(False)
    assert torch.equal(p.get("abcd"), abcd)
    # panels repacked and the pp-ladder's tile timed again: the same sums, maybe in another order
    assert float((p.sigma(R) - ref).abs().max()) <= 1e-12 * scale
    p.close()
    ctx.close()


def test_davidson_and_perturbation_out_of_core(torch):
    import dpdhip
    cfg = SYSTEMS[1]
    ctx = dpdhip.Context(cfg["n"], 0)
    p = dpdhip.Problem(ctx, dpdhip.Spec(cfg["nso"], cfg["nalpha"], cfg["nbeta"], cfg["seed"]), "sf-cpsd", 1)
    kw = dict(conv=1e-8, max_dim=20, max_cyc=100, lowest=True, final_check=False)
    w_in, _, i_in = p.davidson(3, **kw)
    T_in, e_in = p.cpsd()
    p.set_abcd_out_of_core(True)
    print("[ooc] davidson + cpsd out of core", flush=True)
    w_out, _, i_out = p.davidson(3, **kw)
    T_out, e_out = p.cpsd()
    assert i_in["iterations"] == i_out["iterations"]
    np.testing.assert_allclose(w_out, w_in, atol=1e-10)
    np.testing.assert_allclose(e_out["energy"], e_in["energy"], rtol=1e-10, atol=1e-14)
    assert float((T_out - T_in).abs().max()) <= 1e-10 * float(T_in.abs().max())
    p.close()
    ctx.close()
