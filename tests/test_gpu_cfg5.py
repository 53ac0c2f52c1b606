"""cfg 5 at full size (run with -m gpu): the cfg-2 mesh (40^3 cells, 384,000
tetrahedra, 1,100 subdomains) with per-element eps_r ~ U[1,4],
reference-mode absorber.

Adjoint gradient of J = sum |E_k|^2 over 16 interior probe edges,
dJ/deps_e for every element from one forward and one adjoint solve on the
same factors (fwh_set_load, fwh_grad_eps), checked on 64 random elements
(56 drawn from all elements, 8 from those holding a probe edge, so that the
sample includes gradients of every size):
  * against the reference gradient from CPU reference solves: the oracle's
    forward and adjoint HDDM solves (both to 1e-12) with the adjoint formula
    -2 Re(lambda_e^T dA_e/deps E_e) evaluated on the CPU;
  * against central finite differences of J at h = 1e-6 in eps_r: J at
    eps_r +- h of one element, each through a GPU solve of the perturbed
    system (eight systems per batch load).  The perturbed solution is
    computed as x0 + d with A(eps +- h) d = -dA x0 (x0 the nominal solution,
    dA = +-h eps0 dA_e/deps on the element's six edges: the same solution,
    exact algebra), so each solve's stop error is relative to d, which is
    O(h), and not to x: solving for x itself, the two sides' stop and
    rounding errors (~5e-13 of J) divided by 2h swamp every gradient below
    1e-5 of the largest.  J(x0 + d) - J(x0) = sum_p 2 Re(conj(x0_p) d_p) +
    |d_p|^2 over the probes, without cancellation.
Bars: per element, |g - g_ref| <= 1e-6 |g_ref| + 1e-9 max |g| (the second
term for gradients that vanish to rounding).

And the 8-frequency batch, omega0 (0.9 + 0.2 k / 7), k = 0..7, as one
multi-system load and solve against eight separate solves: bit for bit when
the load-time kernel choices agree, else rel. L2 1e-9 at a 1e-12 stop."""
import dataclasses
import os
import sys

import numpy as np
import pytest

from oracle import fw_oracle as fo
from tests.test_gpu_adjoint import element_dA, interior_probes

pytestmark = pytest.mark.gpu

EPS = 1e-12
H = 1e-6          # central-difference step in eps_r
REL_BAR = 1e-6    # relative, per element
FLOOR = 1e-9      # of the largest gradient entry
CHOICES = ("smv_waves", "smv_split", "solve_depth", "solve_waves", "unit_depth", "solve_bucketed", "schur_gemm",
           "solve_persist", "solve_nlarge")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cfg5():
    from adventure_fullwave_amd import mesh
    say("[gpu] cfg5 model")
    return mesh.make_config(5, pml_mode="sigma")


def element_dofs(model, el):
    """the element's six DOFs in the per-subdomain solution layout"""
    off = np.concatenate([[0], np.cumsum(np.asarray(model.nmp, np.int64))])
    s = int(np.searchsorted(model.el_off, el, side="right") - 1)
    return off[s] + np.asarray(model.nop).reshape(-1, 10)[el, 4:]


def sample_elements(model, probes, seed=55):
    rng = np.random.default_rng(seed)
    nel = len(model.eps)
    nop = np.asarray(model.nop).reshape(-1, 10)
    off = np.concatenate([[0], np.cumsum(np.asarray(model.nmp, np.int64))])
    sub_of_el = np.repeat(np.arange(model.nsub), np.diff(model.el_off))
    dofs = off[sub_of_el][:, None] + nop[:, 4:]
    near = np.nonzero(np.isin(dofs, probes).any(axis=1))[0]
    pick_near = rng.choice(near, 8, replace=False)
    rest = np.setdiff1d(np.arange(nel), pick_near)
    return np.sort(np.concatenate([rng.choice(rest, 56, replace=False), pick_near]))


@pytest.fixture(scope="module")
def gpu_grad(cfg5):
    from adventure_fullwave_amd import fwhip
    probes = interior_probes(cfg5, 16, seed=5)
    say("[gpu] cfg5 forward + adjoint solves and gradient")
    sol = fwhip.Solver(cfg5)
    st = sol.solve(eps=EPS, max_loop=40000, graph_iters=16, schur=True)
    assert st["converged"] == 1
    E = sol.x()
    g = np.zeros_like(E)
    g[probes] = np.conj(E[probes])
    sol.set_load(g)
    st_a = sol.solve(eps=EPS, max_loop=40000, graph_iters=16, schur=True)
    assert st_a["converged"] == 1
    lam = sol.x()
    sol.set_load(None)
    grad = sol.grad_eps(E, lam)
    sol.close()
    J = float(np.sum(np.abs(E[probes]) ** 2))
    say(f"[gpu] cfg5 forward {st['iters']} / adjoint {st_a['iters']} iterations, J = {J:.6e}")
    return dict(probes=probes, E=E, lam=lam, grad=grad, J=J, iters=st["iters"])


def test_cfg5_gradient_matches_cpu_reference(cfg5, gpu_grad):
    threads = max(1, min(16, os.cpu_count() or 1))
    probes = gpu_grad["probes"]
    say(f"[gpu] cfg5 oracle forward solve to {EPS:g} on {threads} host threads")
    fwd = fo.hddm_solve(cfg5, eps=EPS, max_loop=40000, threads=threads)
    E = fwd["x_flat"]
    g = np.zeros_like(E)
    g[probes] = np.conj(E[probes])
    say("[gpu] cfg5 oracle adjoint solve")
    adj = fo.hddm_solve(cfg5, eps=EPS, max_loop=40000, threads=threads, load=g)
    assert fwd["stats"]["converged"] == 1 and adj["stats"]["converged"] == 1
    lam = adj["x_flat"]
    els = sample_elements(cfg5, probes)
    ref = np.array([-2.0 * np.real(lam[element_dofs(cfg5, el)] @ element_dA(cfg5, el) @ E[element_dofs(cfg5, el)])
                    for el in els])
    got = gpu_grad["grad"][els]
    scale = np.abs(gpu_grad["grad"]).max()
    rel = np.abs(got - ref) / (np.abs(ref) + FLOOR / REL_BAR * scale)
    say(f"[gpu] cfg5 gradient vs CPU reference on 64 elements: max rel. error {rel.max():.3e} "
        f"(|grad| {np.abs(ref).min():.2e}..{np.abs(ref).max():.2e}, max over the mesh {scale:.3e})")
    assert np.all(np.abs(got - ref) <= REL_BAR * np.abs(ref) + FLOOR * scale)
    # the forward fields agree too
    assert np.linalg.norm(gpu_grad["E"] - E) <= 1e-9 * np.linalg.norm(E)


def test_cfg5_gradient_matches_finite_differences(cfg5, gpu_grad):
    from adventure_fullwave_amd import fwhip
    probes = gpu_grad["probes"]
    x0 = gpu_grad["E"]
    els = sample_elements(cfg5, probes)
    eps0 = np.asarray(cfg5.eps, np.float64)
    EPS0 = 8.854e-12
    dJ = {}
    jobs = [(el, sgn) for el in els for sgn in (1, -1)]
    for b in range(0, len(jobs), 8):
        batch = jobs[b:b + 8]
        models = []
        for el, sgn in batch:
            e = eps0.copy()
            e[el] += sgn * H * EPS0
            models.append(dataclasses.replace(cfg5, eps=e))
        sol = fwhip.Solver(models)
        for k, (el, sgn) in enumerate(batch):
            d = element_dofs(cfg5, el)
            load = np.zeros_like(x0)
            load[d] = -(sgn * H * EPS0) * (element_dA(cfg5, el) @ x0[d])
            sol.select(k).set_load(load)
        sol.solve(eps=EPS, max_loop=40000, graph_iters=16, schur=True)
        for k, (el, sgn) in enumerate(batch):
            st = sol.select(k).stats()
            assert st["converged"] == 1
            dx = sol.x()[probes]
            dJ[(el, sgn)] = float(np.sum(2.0 * np.real(np.conj(x0[probes]) * dx) + np.abs(dx) ** 2))
        sol.close()
        say(f"[gpu] cfg5 finite differences: {b + len(batch)} of {len(jobs)} solves")
    scale = np.abs(gpu_grad["grad"]).max()
    fd = np.array([(dJ[(el, 1)] - dJ[(el, -1)]) / (2 * H * EPS0) for el in els])
    got = gpu_grad["grad"][els]
    rel = np.abs(got - fd) / (np.abs(fd) + FLOOR / REL_BAR * scale)
    say(f"[gpu] cfg5 gradient vs central differences on 64 elements: max rel. error {rel.max():.3e} "
        f"(|fd| {np.abs(fd).min():.2e}..{np.abs(fd).max():.2e})")
    assert np.all(np.abs(got - fd) <= REL_BAR * np.abs(fd) + FLOOR * scale)


def test_cfg5_frequency_batch(cfg5):
    from adventure_fullwave_amd import fwhip
    fac = [0.9 + 0.2 * k / 7 for k in range(8)]
    models = [dataclasses.replace(cfg5, omega=cfg5.omega * f) for f in fac]
    say("[gpu] cfg5 batch of 8 frequencies")
    b = fwhip.Solver(models)
    assert b.info["nsys"] == 8
    b.solve(eps=EPS, max_loop=40000, graph_iters=16, schur=True)
    stats_b = b.stats_all()
    for k, m in enumerate(models):
        s1 = fwhip.Solver(m)
        st1 = s1.solve(eps=EPS, max_loop=40000, graph_iters=16, schur=True)
        b.select(k)
        xb, x1 = b.x(), s1.x()
        same = all(b.info[c] == s1.info[c] for c in CHOICES)
        say(f"[gpu] frequency {k}: iters batch={stats_b[k]['iters']} single={st1['iters']} same_choices={same} "
            f"{ {c: (b.info[c], s1.info[c]) for c in CHOICES} }")
        assert stats_b[k]["converged"] == 1 and st1["converged"] == 1
        if same:
            assert stats_b[k]["iters"] == st1["iters"] and np.array_equal(xb, x1)
        else:
            assert np.linalg.norm(xb - x1) <= 1e-9 * np.linalg.norm(x1)
        s1.close()
    b.close()
