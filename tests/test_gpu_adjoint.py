"""Adjoint permittivity gradient (run with -m gpu).

Objective J(eps) = sum over probe edges of |E_k|^2 with A(eps) E = f.  A is
complex-symmetric, so dJ/deps_e = -2 Re(lambda^T dA/deps_e E) with
A lambda = conj(E) on the probes: one more interface solve on the same
factors (fwh_set_load) and a per-element kernel (fwh_grad_eps).

Checked against (i) the same adjoint formula evaluated on the CPU with direct
sparse solves of the globally assembled system (element matrices from the
oracle), (ii) central finite differences of J with direct solves, and (iii)
a torch.autograd.Function around the GPU path, differentiated by torch and by
central differences of GPU solves.  Bar: 1e-6 relative to the largest
gradient entry.  Finite-difference steps are 1e-4 relative (truncation and
rounding noise both below 1e-7; tests/test_oracle_adjoint.py).
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import fw_oracle as fo
from tests.helpers import global_system, to_global

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def interior_probes(model, n=16, seed=0):
    """per-DOF indices (fwh_get_x layout) of n edges that are neither interface
    nor PEC edges"""
    off = np.concatenate([[0], np.cumsum(np.asarray(model.nmp, np.int64))])
    cand = []
    for s in range(model.nsub):
        bad = np.zeros(model.nmp[s], bool)
        bad[model.ib_nfid[model.ib_off[s]:model.ib_off[s + 1]]] = True
        bad[model.ibc[model.bc_off[s]:model.bc_off[s + 1]]] = True
        cand.append(off[s] + np.nonzero(~bad)[0])
    cand = np.concatenate(cand)
    return np.sort(np.random.default_rng(seed).choice(cand, n, replace=False))


def element_edges(model):
    """global edge ids of every element's 6 edges"""
    nop = np.asarray(model.nop).reshape(-1, 10)
    off = np.concatenate([[0], np.cumsum(np.asarray(model.nmp, np.int64))])
    out = np.zeros((nop.shape[0], 6), np.int64)
    for s in range(model.nsub):
        e0, e1 = model.el_off[s], model.el_off[s + 1]
        out[e0:e1] = model.ledge_global[off[s] + nop[e0:e1, 4:]]
    return out


def element_dA(model, el):
    """dA_e/deps (6x6) from the oracle element at eps = 2 eps_e and 0 (exact:
    the element matrix is affine in eps)"""
    nop = np.asarray(model.nop).reshape(-1, 10)
    crd = np.asarray(model.crd).reshape(-1, 3)
    s = int(np.searchsorted(model.el_off, el, side="right") - 1)
    v = crd[model.ap_off[s] + nop[el, :4]]
    pml = None if model.pml is None else np.asarray(model.pml).reshape(-1, 6)[el]
    e = model.eps[el]
    a2, _ = fo.element(v, 2 * e, model.sigma[el], model.omega, pml=pml)
    a0, _ = fo.element(v, 0.0, model.sigma[el], model.omega, pml=pml)
    return (a2 - a0) / (2 * e)


class CpuSystem:
    """direct solves of the globally assembled system (oracle matrices)"""

    def __init__(self, model):
        self.model = model
        om = fo.OracleModel(model)
        self.A, self.F, self.free = global_system(model, om)
        self.idx = np.nonzero(self.free)[0]
        self.Af = self.A[self.idx][:, self.idx].tocsc()
        self.edges = element_edges(model)

    def solve(self, A_extra=None, rhs=None):
        A = self.Af if A_extra is None else (self.Af + A_extra).tocsc()
        b = (self.F if rhs is None else rhs)[self.idx]
        x = np.zeros(self.model.n_edges, complex)
        x[self.idx] = spla.spsolve(A, b)
        return x

    def perturbation(self, el, dA):
        """sparse dA of element el on the free DOFs"""
        g = self.edges[el]
        pos = -np.ones(self.model.n_edges, np.int64)
        pos[self.idx] = np.arange(self.idx.size)
        p = pos[g]
        keep = p >= 0
        r, c = np.meshgrid(p[keep], p[keep], indexing="ij")
        return sp.coo_matrix((dA[np.ix_(keep, keep)].ravel(), (r.ravel(), c.ravel())),
                             shape=(self.idx.size, self.idx.size)).tocsr()


def probe_edges(model, probes):
    off = np.concatenate([[0], np.cumsum(np.asarray(model.nmp, np.int64))])
    return np.asarray(model.ledge_global)[probes]


@pytest.fixture(scope="module", params=["sigma", "stretch"])
def case(request, small_model, small_sigma_model):
    model = small_sigma_model if request.param == "sigma" else small_model
    probes = interior_probes(model)
    cpu = CpuSystem(model)
    pe = probe_edges(model, probes)
    E = cpu.solve()
    g = np.zeros(model.n_edges, complex)
    g[pe] = np.conj(E[pe])
    lam = cpu.solve(rhs=g)
    return dict(model=model, probes=probes, cpu=cpu, pe=pe, E=E, lam=lam, pml=request.param)


def gpu_gradient(fw, model, probes, eps=1e-11):
    sol = fw.Solver(model)
    st = sol.solve(eps=eps, max_loop=20000, graph_iters=16, schur=True)
    assert st["converged"] == 1
    E = sol.x()
    f = np.zeros_like(E)
    f[probes] = np.conj(E[probes])
    sol.set_load(f)
    st = sol.solve(eps=eps, max_loop=20000, graph_iters=16, schur=True)
    assert st["converged"] == 1
    lam = sol.x()
    sol.set_load(None)
    grad = sol.grad_eps(E, lam)
    J = float(np.sum(np.abs(E[probes]) ** 2))
    sol.close()
    return J, grad, E


def test_gradient_matches_cpu_adjoint_and_fd(case):
    from adventure_fullwave_amd import fwhip
    model, cpu = case["model"], case["cpu"]
    say(f"[gpu] adjoint gradient pml={case['pml']}")
    J, grad, E = gpu_gradient(fwhip, model, case["probes"])
    # CPU adjoint gradient at every element
    Eg, lg = case["E"], case["lam"]
    ref = np.zeros(len(grad))
    for el in range(len(grad)):
        g = cpu.edges[el]
        ref[el] = -2.0 * np.real(lg[g] @ element_dA(model, el) @ Eg[g])
    scale = np.abs(ref).max()
    assert np.abs(grad - ref).max() <= 1e-6 * scale
    # the forward GPU solution is the direct one
    x = to_global(model, [E[o:o + n] for o, n in zip(np.concatenate([[0], np.cumsum(model.nmp)[:-1]]), model.nmp)])
    assert np.linalg.norm(x - Eg) <= 1e-8 * np.linalg.norm(Eg)
    # central differences of J with direct solves on the 8 elements of largest |grad|
    pe = case["pe"]
    for el in np.argsort(-np.abs(ref))[:8]:
        h = 1e-4 * model.eps[el]
        dA = cpu.perturbation(el, element_dA(model, el))
        Jp = np.sum(np.abs(cpu.solve(h * dA)[pe]) ** 2)
        Jm = np.sum(np.abs(cpu.solve(-h * dA)[pe]) ** 2)
        fd = (Jp - Jm) / (2 * h)
        assert abs(fd - grad[el]) <= 1e-6 * scale, (el, fd, grad[el], ref[el])


def test_torch_autograd_harness(case):
    """torch.autograd.Function around the GPU solves: backward = adjoint
    gradient, checked against central differences of the same GPU path."""
    import torch
    from adventure_fullwave_amd import fwhip
    model, probes = case["model"], case["probes"]
    say(f"[gpu] torch autograd harness pml={case['pml']}")

    def with_eps(eps):
        import dataclasses
        return dataclasses.replace(model, eps=np.ascontiguousarray(eps, np.float64))

    class FwObjective(torch.autograd.Function):
        @staticmethod
        def forward(ctx, eps):
            J, grad, _ = gpu_gradient(fwhip, with_eps(eps.detach().cpu().numpy()), probes, eps=1e-12)
            ctx.save_for_backward(torch.from_numpy(grad))
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            (g,) = ctx.saved_tensors
            return gout * g

    eps0 = torch.tensor(np.asarray(model.eps, np.float64), requires_grad=True)
    J = FwObjective.apply(eps0)
    J.backward()
    g = eps0.grad.numpy()
    scale = np.abs(g).max()
    for el in np.argsort(-np.abs(g))[:4]:
        h = 1e-4 * float(model.eps[el])
        ep = eps0.detach().clone()
        em = eps0.detach().clone()
        ep[el] += h
        em[el] -= h
        fd = (FwObjective.apply(ep).item() - FwObjective.apply(em).item()) / (2 * h)
        assert abs(fd - g[el]) <= 1e-6 * scale, (el, fd, g[el])
