"""The adjoint gradient formula on the CPU alone (no GPU): dJ/deps_e =
-2 Re(lambda^T dA_e/deps E), A lambda = conj(E) on the probes, evaluated with
direct sparse solves of the globally assembled oracle system, against central
finite differences of J = sum |E_probe|^2.  The relative step is 1e-4: its
truncation error (~h^2) and the rounding noise of the difference (~1/h) both
stay below 1e-7 of the gradient; at 1e-6 the difference of two J values
cancels to solve rounding and carries ~1e-5 relative noise."""
import numpy as np
import pytest

from tests.test_gpu_adjoint import CpuSystem, element_dA, interior_probes, probe_edges


@pytest.mark.parametrize("pml", ["sigma", "stretch"])
def test_adjoint_formula_matches_finite_differences(pml, small_model, small_sigma_model):
    model = small_sigma_model if pml == "sigma" else small_model
    cpu = CpuSystem(model)
    pe = probe_edges(model, interior_probes(model))
    E = cpu.solve()
    g = np.zeros(model.n_edges, complex)
    g[pe] = np.conj(E[pe])
    lam = cpu.solve(rhs=g)
    rng = np.random.default_rng(5)
    els = rng.choice(len(model.eps), 6, replace=False)
    grads, fds = [], []
    for el in els:
        dA = element_dA(model, el)
        e = cpu.edges[el]
        grads.append(-2.0 * np.real(lam[e] @ dA @ E[e]))
        h = 1e-4 * model.eps[el]
        P = cpu.perturbation(el, dA)
        fds.append((np.sum(np.abs(cpu.solve(h * P)[pe]) ** 2) - np.sum(np.abs(cpu.solve(-h * P)[pe]) ** 2)) / (2 * h))
    grads, fds = np.array(grads), np.array(fds)
    assert np.abs(grads - fds).max() <= 1e-6 * np.abs(grads).max()
    # a sign or transpose mistake is caught: the conjugated adjoint fails
    wrong = np.array([-2.0 * np.real(np.conj(lam[cpu.edges[el]]) @ element_dA(model, el) @ E[cpu.edges[el]])
                      for el in els])
    assert np.abs(wrong - fds).max() > 1e-3 * np.abs(grads).max()
