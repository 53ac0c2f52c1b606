"""Factors released after the Schur formation and recomputed in chunks (run
with -m gpu).

fwh_release_factors frees the skyline factors and block inverses once the
explicit S_s are formed; the explicit iteration needs neither, and the solves
that do (first and last steps, right-hand-side solves, the adjoint's solves)
recompute them chunk by chunk within a byte budget with the load's own kernels.
The recomputed factors are the load's, so every result must equal the resident
path's bit for bit: a solve, an adjoint solve on the same context, and a batch of
right-hand sides, with budgets that cut the model into several chunks.  The
implicit Schur step is refused afterwards, as are a budget smaller than one
subdomain's factors and a release before the formation.  A load under a
budget below the model's factors (fwh_set_factor_budget) never holds them all:
its S_s, solutions and adjoint gradient equal the resident load's bit for bit.
"""
import dataclasses
import sys

import numpy as np
import pytest

from tests.helpers import pinned_env

pytestmark = pytest.mark.gpu

# the LDL^T form pinned too (two waves, depth 16: the chunked solves run the
# two-wave kernel), so that two loads make the same choices
PINS = {"FWH_SOLVE_PERSIST": "0", "FWH_SMV_WAVES": "4", "FWH_SMV_SPLIT": "0", "FWH_SOLVE_DEPTH": "16",
        "FWH_SOLVE_WAVES": "2"}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def formed(model):
    from adventure_fullwave_amd import fwhip
    with pinned_env(PINS):
        sol = fwhip.Solver(model)
        sol.schur_form()
    return sol


@pytest.mark.parametrize("nchunk", [3, 7])
def test_released_factors_same_bits(small_sigma_model, cfg1_sigma_model, nchunk):
    from adventure_fullwave_amd import fwhip
    for model in (small_sigma_model, cfg1_sigma_model):
        a = formed(model)
        b = formed(model)
        # same load-time choices, so the resident path's bits are the reference
        for k in ("solve_depth", "solve_waves", "solve_persist", "smv_waves", "smv_split"):
            assert a.info[k] == b.info[k], k
        n = int(a.info["n_dof"])
        budget = int(max(b.info["factor_need_max"], b.info["factor_need_total"] / nchunk))
        b.release_factors(budget)
        say(f"[gpu] {model.name}: released {b.info['factor_bytes_released'] / 1e6:.2f} MB, "
            f"{b.info['factor_chunks']} chunks of at most {b.info['factor_chunk_bytes'] / 1e6:.2f} MB")
        assert b.info["factors_resident"] == 0 and b.info["factor_chunks"] >= 2
        assert b.info["factor_chunk_bytes"] <= budget
        st_a = a.solve(eps=1e-8, max_loop=20000, graph_iters=16, schur=True)
        st_b = b.solve(eps=1e-8, max_loop=20000, graph_iters=16, schur=True)
        assert st_a["iters"] == st_b["iters"] and st_b["converged"] == 1
        assert np.array_equal(a.x(), b.x())
        # adjoint-style solve: a new load on the same factors
        rng = np.random.default_rng(5)
        f = np.zeros(n, np.complex128)
        pick = rng.choice(n, 16, replace=False)
        f[pick] = rng.standard_normal(16) + 1j * rng.standard_normal(16)
        for s in (a, b):
            s.set_load(f)
        ya = a.solve(eps=1e-10, max_loop=20000, graph_iters=16, schur=True)
        yb = b.solve(eps=1e-10, max_loop=20000, graph_iters=16, schur=True)
        assert ya["iters"] == yb["iters"] and np.array_equal(a.x(), b.x())
        rhs = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        assert np.array_equal(a.ldl_solve_all(rhs), b.ldl_solve_all(rhs))
        assert np.array_equal(a.ldl_solve_all(rhs, plain=True), b.ldl_solve_all(rhs, plain=True))
        with pytest.raises(fwhip.FwhError, match="FWH_E_STATE"):
            b.solve(eps=1e-8, max_loop=100, graph_iters=0, schur=False)
        with pytest.raises(fwhip.FwhError, match="FWH_E_STATE"):
            b.schur_apply(np.zeros(int(b.info["n_iface"]), np.complex128))
        a.close()
        b.close()


def test_released_factors_batch(small_sigma_model):
    m = small_sigma_model
    systems = [m, dataclasses.replace(m, omega=1.05 * m.omega)]
    a = formed(systems)
    b = formed(systems)
    b.release_factors(int(max(b.info["factor_need_max"], b.info["factor_need_total"] / 3)))
    for s in (a, b):
        s.solve(eps=1e-8, max_loop=5000, graph_iters=8, schur=True)
    for k in range(2):
        assert a.select(k).stats()["iters"] == b.select(k).stats()["iters"]
        assert np.array_equal(a.x(), b.x())
    a.close()
    b.close()


def test_release_refusals(small_sigma_model):
    from adventure_fullwave_amd import fwhip
    with pinned_env(PINS):
        sol = fwhip.Solver(small_sigma_model)
    with pytest.raises(fwhip.FwhError, match="FWH_E_STATE"):
        sol.release_factors(1 << 30)  # before the formation
    sol.schur_form()
    with pytest.raises(fwhip.FwhError, match="FWH_E_LIMIT"):
        sol.release_factors(int(sol.info["factor_need_max"]) - 16)  # below one subdomain's factors
    assert sol.info["factors_resident"] == 1
    sol.release_factors(1 << 30)
    with pytest.raises(fwhip.FwhError, match="FWH_E_STATE"):
        sol.release_factors(1 << 30)
    sol.close()


@pytest.mark.parametrize("nchunk", [2, 5])
def test_streamed_load_same_bits(small_sigma_model, cfg1_sigma_model, nchunk):
    """A model whose factors exceed the budget is loaded streamed: its S_s are
    formed chunk by chunk and the factors are never all resident.  With the
    LDL^T kernel form a streamed load uses (depth 16, one wave) pinned on the
    resident load, every result is the same bits."""
    from adventure_fullwave_amd import fwhip
    pins = dict(PINS, FWH_SOLVE_DEPTH="16", FWH_SOLVE_WAVES="1")
    for model in (small_sigma_model, cfg1_sigma_model):
        with pinned_env(pins):
            a = fwhip.Solver(model)
            a.schur_form()
            total = a.info["factor_need_total"]
            budget = int(max(a.info["factor_need_max"], total / nchunk))
            b = fwhip.Solver(model, factor_budget=budget)
            b.schur_form()
        say(f"[gpu] {model.name} streamed: {b.info['factor_chunks']} chunks of at most "
            f"{b.info['factor_chunk_bytes'] / 1e6:.2f} MB (factors {total / 1e6:.2f} MB)")
        assert a.info["factors_resident"] == 1 and b.info["factors_resident"] == 0
        assert b.info["factor_chunks"] >= 2 and b.info["schur_gemm"] == 1
        assert b.info["factor_chunk_bytes"] <= budget
        rng = np.random.default_rng(9)
        ng = int(a.info["n_iface"])
        p = rng.standard_normal(ng) + 1j * rng.standard_normal(ng)
        assert np.array_equal(a.schur_apply(p, explicit=True), b.schur_apply(p, explicit=True))
        st_a = a.solve(eps=1e-8, max_loop=20000, graph_iters=16, schur=True)
        st_b = b.solve(eps=1e-8, max_loop=20000, graph_iters=16, schur=True)
        assert st_a["iters"] == st_b["iters"] and np.array_equal(a.x(), b.x())
        # the adjoint gradient of J = sum |E|^2 over 8 edges, on the same contexts
        E = a.x()
        probes = rng.choice(E.size, 8, replace=False)
        f = np.zeros_like(E)
        f[probes] = np.conj(E[probes])
        grads = []
        for sol in (a, b):
            sol.set_load(f)
            sol.solve(eps=1e-10, max_loop=20000, graph_iters=16, schur=True)
            grads.append(sol.grad_eps(E, sol.x()))
            sol.set_load(None)
        assert np.array_equal(grads[0], grads[1]) and np.abs(grads[0]).max() > 0
        with pytest.raises(fwhip.FwhError, match="FWH_E_STATE"):
            b.solve(eps=1e-8, max_loop=100, graph_iters=0, schur=False)
        a.close()
        b.close()
