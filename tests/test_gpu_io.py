"""The reference's file formats on the GPU data path (dpdhip.io over
csrc/dpd_io.cpp): a problem's integral blocks written as moio lists under the
reference's names and loaded into another problem give the same operator bit
for bit; a Davidson restarted from the eigenvector file it wrote
(mod_davidson.f90:388-390 writes, :588-589 reads) converges at once to the
same roots; ground-state amplitudes round-trip through the f_T stream file."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CFG = dict(nso=(4, 2, 2, 3), nalpha=4, nbeta=2)


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


def test_integrals_through_moio_lists(torch, tmp_path):
    import dpdhip
    from dpdhip import io as dio
    ctx = dpdhip.Context(4, 0)
    a = dpdhip.Problem(ctx, dpdhip.Spec(CFG["nso"], CFG["nalpha"], CFG["nbeta"], 1), "sf-cpsd", 1)
    b = dpdhip.Problem(ctx, dpdhip.Spec(CFG["nso"], CFG["nalpha"], CFG["nbeta"], 2), "sf-cpsd", 1)
    print("[io] save / load integrals", flush=True)
    dio.save_integrals(a, str(tmp_path))
    # b has other integrals until it loads a's (the orbital energies are not in the lists)
    b.set("eorb", a.get("eorb"))
    b.refresh()
    R = torch.randn(2 * a.dim, dtype=torch.float64, device="cuda:0", generator=torch.Generator("cuda:0").manual_seed(5))
    assert not torch.equal(a.sigma(R), b.sigma(R))
    dio.load_integrals(b, str(tmp_path))
    assert torch.equal(a.sigma(R), b.sigma(R))
    for p in (a, b):
        p.close()
    ctx.close()


def test_davidson_restart_from_eigenvector_file(torch, tmp_path):
    """initDavidson (mod_davidson.f90:563-610) takes the file's vectors as an
    orthonormal basis as they stand, so the restart is exact for a symmetric J:
    the SF-CIS roots that the EE programs read as their guess."""
    import dpdhip
    from dpdhip import io as dio
    ctx = dpdhip.Context(4, 0)
    p = dpdhip.Problem(ctx, dpdhip.Spec(CFG["nso"], CFG["nalpha"], CFG["nbeta"], 3), "sf-cis", 1)
    w, R, info = p.davidson(2, conv=1e-10, max_dim=20, max_cyc=100, lowest=True)
    assert info["converged"]
    f = str(tmp_path / "CIS_sym1")
    dio.eigvec_write(f, w, R, p.dim)
    w2, R2 = dio.eigvec_read(f, 2, p.dim, out=torch.empty_like(R))
    assert np.array_equal(w2, w) and torch.equal(R2, R)
    print("[io] restarted Davidson", flush=True)
    w3, _, info3 = p.davidson(2, conv=1e-8, max_dim=20, max_cyc=100, lowest=True, R=R2.clone())
    assert info3["converged"] and info3["iterations"] == 1
    np.testing.assert_allclose(w3, w, atol=1e-10)
    p.close()
    ctx.close()


def test_amplitudes_through_stream_file(torch, tmp_path):
    import dpdhip
    from dpdhip import io as dio
    ctx = dpdhip.Context(4, 0)
    p = dpdhip.Problem(ctx, dpdhip.Spec(CFG["nso"], CFG["nalpha"], CFG["nbeta"], 4), "sf-cpsd", 1)
    T, _ = p.cpsd()
    f = str(tmp_path / "CPSd_T__")
    dio.write_amplitudes(f, T)
    T2 = dio.read_amplitudes(f, T.shape[1])
    assert torch.equal(T, T2)
    raw = np.fromfile(f, "<f8")
    np.testing.assert_array_equal(raw, T.cpu().numpy().ravel())
    p.close()
    ctx.close()
