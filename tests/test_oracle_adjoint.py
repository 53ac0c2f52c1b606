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


def test_oracle_load_override_matches_direct_adjoint(small_sigma_model):
    """fwo_hddm_solve_load: (i) the assembled load passed explicitly gives the
    same bits as the plain solve; (ii) an adjoint load conj(E) on the probes
    gives the direct solution of A lambda = g (A = A^T) to the stop error."""
    from oracle import fw_oracle as fo
    from tests.helpers import to_global
    model = small_sigma_model
    om = fo.OracleModel(model)
    f = np.concatenate([om.get(s)["f"] for s in range(model.nsub)])
    r0 = fo.hddm_solve(model, eps=1e-12, max_loop=5000)
    r1 = fo.hddm_solve(model, eps=1e-12, max_loop=5000, load=f)
    assert np.array_equal(r0["x_flat"], r1["x_flat"]) and r0["stats"]["iters"] == r1["stats"]["iters"]
    cpu = CpuSystem(model)
    probes = interior_probes(model)
    pe = probe_edges(model, probes)
    E = r0["x_flat"]
    g = np.zeros_like(E)
    g[probes] = np.conj(E[probes])
    ra = fo.hddm_solve(model, eps=1e-12, max_loop=5000, load=g)
    gg = np.zeros(model.n_edges, complex)
    gg[pe] = np.conj(E[probes])
    lam = cpu.solve(rhs=gg)
    offs = np.concatenate([[0], np.cumsum(model.nmp)])
    lam_h = to_global(model, [ra["x_flat"][offs[s]:offs[s + 1]] for s in range(model.nsub)])
    assert np.linalg.norm(lam_h - lam) <= 1e-9 * np.linalg.norm(lam)
    # a load placed on the wrong copy is caught: shifting it by one DOF fails
    ra2 = fo.hddm_solve(model, eps=1e-12, max_loop=5000, load=np.roll(g, 1))
    lam2 = to_global(model, [ra2["x_flat"][offs[s]:offs[s + 1]] for s in range(model.nsub)])
    assert np.linalg.norm(lam2 - lam) > 1e-3 * np.linalg.norm(lam)


def test_adjoint_matches_perturbation_differences_at_small_step(small_sigma_model):
    """The adjoint gradient against central differences at h = 1e-6 in eps_r
    (the step of the large-model GPU check), each side through oracle HDDM
    solves of the perturbed system written as x0 + d with
    A(eps +- h) d = -dA x0: per element within 1e-8 relative."""
    import dataclasses
    from oracle import fw_oracle as fo
    from tests.test_gpu_cfg5 import element_dofs
    rng = np.random.default_rng(1)
    m = dataclasses.replace(small_sigma_model,
                            eps=np.asarray(small_sigma_model.eps) * rng.uniform(1, 4, len(small_sigma_model.eps)))
    probes = interior_probes(m)
    x0 = fo.hddm_solve(m, eps=1e-12, max_loop=5000)["x_flat"]
    g = np.zeros_like(x0)
    g[probes] = np.conj(x0[probes])
    lam = fo.hddm_solve(m, eps=1e-12, max_loop=5000, load=g)["x_flat"]
    EPS0, H = 8.854e-12, 1e-6
    for el in rng.choice(len(m.eps), 4, replace=False):
        d = element_dofs(m, el)
        dA = element_dA(m, el)
        grad = -2.0 * np.real(lam[d] @ dA @ x0[d])
        dj = []
        for sgn in (1, -1):
            e = np.asarray(m.eps, float).copy()
            e[el] += sgn * H * EPS0
            load = np.zeros_like(x0)
            load[d] = -(sgn * H * EPS0) * (dA @ x0[d])
            dx = fo.hddm_solve(dataclasses.replace(m, eps=e), eps=1e-12, max_loop=5000, load=load)["x_flat"][probes]
            dj.append(np.sum(2.0 * np.real(np.conj(x0[probes]) * dx) + np.abs(dx) ** 2))
        fd = (dj[0] - dj[1]) / (2 * H * EPS0)
        assert abs(grad - fd) <= 1e-8 * abs(fd), (el, grad, fd)
