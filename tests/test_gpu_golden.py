"""FP64 GPU Davidson against the committed CPU-oracle results
(tests/golden/*_sfcpsd_oracle.json, written by tests/golden/make_oracle_references.py
with the C engines + OpenBLAS and the numpy restatement of lib_cc/mod_davidson.f90):
same iteration count and every root within 1e-8 Eh."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _golden(name):
    path = os.path.join(HERE, "golden", f"{name}_sfcpsd_oracle.json")
    if not os.path.exists(path):
        pytest.skip(f"{path} not generated")
    return json.load(open(path))


@pytest.mark.parametrize("name", ["c2v-24", "d2h-120"])
def test_fp64_davidson_matches_cpu_oracle(name):
    import torch

    import dpdhip
    g = _golden(name)
    ctx = dpdhip.Context(len(g["nso"]), 0)
    prob = dpdhip.Problem(ctx, dpdhip.Spec(tuple(g["nso"]), g["nalpha"], g["nbeta"], g["seed"]), g["method"],
                          g["symR"])
    assert prob.dim == g["dim"]
    guess, e0 = prob.guess(g["nroot"])
    np.testing.assert_allclose(e0, g["guess_energies"], rtol=0, atol=1e-12)
    w, R, info = prob.davidson(g["nroot"], conv=g["conv"], max_dim=g.get("max_dim", 20),
                               max_cyc=g.get("max_cyc", 50), lowest=g.get("lowest", False), R=guess)
    torch.cuda.synchronize()
    assert info["converged"] == g["converged"]
    assert info["iterations"] == g["iterations"], (info["iterations"], g["iterations"])
    dw = np.abs(np.asarray(w) - np.asarray(g["w"]))
    assert dw.max() <= 1e-8, dw
