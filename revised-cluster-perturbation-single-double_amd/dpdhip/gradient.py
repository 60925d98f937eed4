"""Gradient of one excitation energy with respect to the orbital energies and
the MO integral blocks, as a ``torch.autograd.Function`` over the device solver.

For the non-symmetric operator J(theta) of a method, with right eigenvector R
and left eigenvector L of the eigenvalue w,

    dw/dtheta = L^T (dJ/dtheta) R / (L^T R)

(first-order perturbation theory of a simple eigenvalue).  R comes from the
Davidson solve with J, L from the solve with J^T (``transpose=True``), and the
contraction L^T (dJ/dtheta) R for every stored element of every input is
``dpd_problem_grad``: each sigma term contributes the product of the two vectors
contracted into the shape of its integral block, and the MP2 amplitudes and the
CIS(Doo) intermediates are recomputed from the inputs and differentiated
through.  The reference solves for energies only (lib_cc/mod_davidson.f90); it
has no gradient.

theta are the stored elements of the inputs: the orbital energies ("eorb",
occupied irreps then virtual) and the integral blocks "iabj", "ijka", "ijkl",
"abij", "abci", "abcd" in their packed layouts, each block an independent
input (the reference keeps them as separate arrays, CISdoo.SD.SF/allocate_mem.f90).
Elements that vanish by spin conservation (an alpha and a beta orbital in
positions the integral cannot connect) are zero in the inputs; the gradient
has entries there too, but a change of such an element couples the spin-flip
sector to others, so differences along it follow another state.
"""
from __future__ import annotations

import torch

from . import Problem


def _pick(w_ref, ws):
    k = min(range(len(ws)), key=lambda i: abs(ws[i] - w_ref))
    return k, abs(ws[k] - w_ref)


class ExcitationEnergy(torch.autograd.Function):
    """w = ExcitationEnergy.apply(problem, solve, names, *tensors)

    ``solve`` is a dict of Davidson options: nroot, root (index among the lowest
    nroot), conv, max_dim, max_cyc.  The tensors (float64, on the problem's
    device) are written into the problem before the solve, so the function is
    the map from inputs to the eigenvalue."""

    @staticmethod
    def forward(ctx, problem: Problem, solve: dict, names, *tensors):
        for n, t in zip(names, tensors):
            problem.set(n, t.detach())
        problem.refresh()
        nroot = int(solve.get("nroot", 1))
        root = int(solve.get("root", 0))
        kw = dict(conv=solve.get("conv", 1e-9), max_dim=solve.get("max_dim", 20),
                  max_cyc=solve.get("max_cyc", 100), lowest=True)
        # warm starts are copied: the solver overwrites its guesses in place
        R0, L0 = solve.get("R0"), solve.get("L0")
        wr, R, ir = problem.davidson(nroot, R=None if R0 is None else R0.clone(), **kw)
        if not ir["converged"]:
            raise RuntimeError(f"right eigenproblem did not converge: {ir['residual']}")
        wl, L, il = problem.davidson(nroot, R=None if L0 is None else L0.clone(), transpose=True, **kw)
        if not il["converged"]:
            raise RuntimeError(f"left eigenproblem did not converge: {il['residual']}")
        k, dw = _pick(wr[root], wl)
        if dw > max(1e-6, 100 * kw["conv"]):
            raise RuntimeError(f"left/right eigenvalues differ: {wr[root]} vs {wl[k]}")
        n = problem.dim
        Rv = R[root * n:(root + 1) * n].clone()
        Lv = L[k * n:(k + 1) * n].clone()
        # two-sided Rayleigh quotient: its error is the product of the left and
        # right residuals, so it is the eigenvalue to far below the tolerance
        lr = torch.dot(Lv, Rv)
        w = torch.dot(Lv, problem.sigma(Rv)) / lr
        ctx.problem = problem
        ctx.names = names
        # another forward on the same problem may overwrite its inputs before
        # this backward runs: the inputs are kept and the problem's version
        # recorded, so that backward evaluates the gradient at this theta
        ctx.version = problem.version
        ctx.save_for_backward(Lv, Rv, *tensors)
        ctx.info = {"right": ir, "left": il, "w_right": float(wr[root]), "w_left": float(wl[k]), "R": Rv, "L": Lv}
        if solve.get("info") is not None:
            solve["info"].update(ctx.info)
        return w

    @staticmethod
    def backward(ctx, gw):
        Lv, Rv, *tensors = ctx.saved_tensors
        lr = float(torch.dot(Lv, Rv))
        if abs(lr) < 1e-12:
            raise RuntimeError("left and right eigenvectors are orthogonal: the eigenvalue is not simple")
        problem = ctx.problem
        if problem.version != ctx.version:
            for n, t in zip(ctx.names, tensors):
                problem.set(n, t.detach())
            problem.refresh()
        g = problem.grad(Lv, Rv, scale=1.0 / lr)
        out = [gw * g[n] if n in g else None for n in ctx.names]
        return (None, None, None, *out)


def excitation_energy(problem: Problem, inputs: dict, info: dict | None = None, **solve):
    """Eigenvalue of root ``solve['root']`` (default the lowest) as a
    differentiable function of the named input tensors; ``info`` (a dict, when
    given) receives the Davidson reports of the right and left solves."""
    names = tuple(inputs)
    solve["info"] = info
    return ExcitationEnergy.apply(problem, solve, names, *(inputs[n] for n in names))
