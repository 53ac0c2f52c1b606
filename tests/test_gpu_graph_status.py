"""The status step of explicit interface iterations inside graphs (run with
-m gpu).

Launched one kernel at a time, an iteration is the Schur matvec, k_iter_ab
(alpha and the vector updates) and k_iter_c (beta, stop test, status
machine; hddm_solver_s.h:311-350, hddm_solver.h:125-143).  In a graph of
explicit iterations k_iter_c runs only after the last one: the Schur matvec
of the next iteration takes the status step from k_iter_ab's partials (every
workgroup, k_iter_c's reduction order) and the next k_iter_ab commits it.
Both forms must give the same
bits: solution, iteration count, loop count and residual history, for graphs
whose length divides the count or not, for single systems and batches, and
at max_loop."""
import dataclasses
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def fw():
    from adventure_fullwave_amd import fwhip
    return fwhip


def run(sol, graph_iters, **kw):
    st = sol.solve(eps=kw.get("eps", 1e-8), max_loop=kw.get("max_loop", 5000), graph_iters=graph_iters,
                   schur=True)
    return st, sol.x().copy(), sol.resid(st["loops"]).copy()


@pytest.mark.parametrize("pml", ["sigma", "stretch"])
def test_graph_status_step_bit_identical(fw, small_model, small_sigma_model, pml):
    model = small_sigma_model if pml == "sigma" else small_model
    sol = fw.Solver(model)
    st0, x0, r0 = run(sol, 0)
    say(f"[gpu] explicit {pml}: {st0['iters']} iterations one kernel at a time")
    assert st0["converged"] == 1
    for g in (16, 5, 1):
        st, x, r = run(sol, g)
        say(f"[gpu] explicit {pml} graph {g}: {st['iters']} iterations")
        assert st["iters"] == st0["iters"] and st["loops"] == st0["loops"] and st["converged"] == 1
        assert np.array_equal(x, x0)
        assert np.array_equal(r, r0)
    sol.close()


def test_graph_status_step_batch(fw, small_sigma_model):
    m = small_sigma_model
    models = [dataclasses.replace(m, omega=0.9 * m.omega), m, dataclasses.replace(m, omega=1.1 * m.omega)]
    sol = fw.Solver(models)
    out = {}
    for g in (0, 16, 7):
        sol.solve(eps=1e-8, max_loop=5000, graph_iters=g, schur=True)
        out[g] = []
        for k in range(len(models)):
            sol.select(k)
            st = sol.stats()
            out[g].append((st["iters"], st["loops"], st["converged"], sol.x().copy(), sol.resid(st["loops"]).copy()))
        say(f"[gpu] batch graph {g}: iterations {[o[0] for o in out[g]]}")
    for g in (16, 7):
        for a, b in zip(out[0], out[g]):
            assert a[:3] == b[:3] and a[2] == 1
            assert np.array_equal(a[3], b[3])
            assert np.array_equal(a[4], b[4])
    sol.close()


@pytest.mark.parametrize("graph_iters", [4, 3, 0])
def test_graph_status_step_max_loop(fw, small_model, graph_iters):
    """max_loop = 7 ends the loop with the status machine's count (7
    iterations, 8 loops, not converged) whatever the graph length; the last
    solve uses the search direction, as one kernel at a time."""
    sol = fw.Solver(small_model)
    st, x, r = run(sol, graph_iters, eps=1e-30, max_loop=7)
    say(f"[gpu] explicit max_loop 7 graph {graph_iters}: {st['iters']} iterations, {st['loops']} loops")
    assert st["iters"] == 7 and st["loops"] == 8 and st["converged"] == 0
    st0, x0, r0 = run(sol, 0, eps=1e-30, max_loop=7)
    assert np.array_equal(x, x0) and np.array_equal(r, r0)
    sol.close()
