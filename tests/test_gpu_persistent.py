"""Persistent forms of the batched kernels (run with -m gpu).

The explicit Schur matvec and the batched LDL^T kernel can run as
persistent grids (as many workgroups as are resident at once) whose
workgroups claim work items from a queue in device memory until none is
left; the workgroup that ends last puts the queue back to zero for the next
launch.  The load-time timing picks them where they pay; here they are
forced (FWH_SMV_PERSIST, FWH_SOLVE_PERSIST, FWH_SOLVE_NLARGE, FWH_SOLVE_DEPTH,
read at load or when the S_s are formed).

- Matvec: every item writes its own outputs, so the persistent grid must give
  the same bits as one workgroup per item: Schur applies (repeated, so a queue
  left dirty by one launch would show in the next), solves with and without
  graphs (finished systems leave their launches early, before any claim), and
  a batch of two systems finishing at different iterations.
- LDL^T kernel: form 1 (one wave per subdomain) must give the results of
  the one-wave kernel at the same depth, subdomain by subdomain.  The
  persistent kernel runs the same steps in the same order, but the compiler
  may fuse a complex product's multiplies and adds differently in another
  kernel, so agreement is to rounding (1e-10 of the subdomain's largest
  entry, the bar of the other solve comparisons; bit-identical where the
  fusion agrees).  Repeated launches agree bit for bit (the queue is back to
  zero).  Solves through it match the oracle as the static forms do (count
  within 2 %, solution rel. L2 1e-7).  Form 2 (two size-bucketed queues) is
  switched off and refused when forced.
"""
import dataclasses
import sys

import numpy as np
import pytest

from oracle import fw_oracle as fo
from tests.helpers import pinned_env

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# the LDL^T form of the first and last steps, which the load otherwise picks
# by timing (two loads may pick forms that round differently)
LDL_PINS = {"FWH_SOLVE_DEPTH": 16, "FWH_SOLVE_WAVES": 1}


def formed(model, **env):
    from adventure_fullwave_amd import fwhip
    with pinned_env({k: str(v) for k, v in env.items()}):
        sol = fwhip.Solver(model)
        sol.schur_form()
    return sol


def loaded(model, **env):
    from adventure_fullwave_amd import fwhip
    with pinned_env({k: str(v) for k, v in env.items()}):
        return fwhip.Solver(model)


@pytest.mark.parametrize("split", [0, 2])
def test_smv_persistent_same_bits(small_sigma_model, cfg1_sigma_model, split):
    for model in (small_sigma_model, cfg1_sigma_model):
        out = []
        for persist in (0, 1):
            sol = formed(model, FWH_SMV_PERSIST=persist, FWH_SMV_SPLIT=split, FWH_SMV_WAVES=4, **LDL_PINS)
            say(f"[gpu] {model.name} matvec split={split} persist={persist}: grid {sol.info['smv_pgrid']} "
                f"for {sol.info['smv_nwork']} items")
            assert sol.info["smv_persist"] == persist
            rng = np.random.default_rng(11)
            ng = int(sol.info["n_iface"])
            p = rng.standard_normal(ng) + 1j * rng.standard_normal(ng)
            qs = [sol.schur_apply(p, explicit=True) for _ in range(3)]
            for q in qs[1:]:
                assert np.array_equal(q, qs[0])
            runs = []
            for graph in (16, 0):
                st = sol.solve(eps=1e-8, max_loop=20000, graph_iters=graph, schur=True)
                runs.append((st["iters"], sol.x().copy(), sol.resid(st["loops"]).copy()))
            assert runs[0][0] == runs[1][0] and np.array_equal(runs[0][1], runs[1][1])
            q_after = sol.schur_apply(p, explicit=True)
            assert np.array_equal(q_after, qs[0])
            out.append((qs[0], runs[0]))
            sol.close()
        (q0, r0), (q1, r1) = out
        assert np.array_equal(q0, q1)
        assert r0[0] == r1[0] and np.array_equal(r0[1], r1[1]) and np.array_equal(r0[2], r1[2])


def test_smv_persistent_batch(small_sigma_model):
    m = small_sigma_model
    systems = [m, dataclasses.replace(m, omega=1.05 * m.omega)]
    res = []
    for persist in (0, 1):
        b = formed(systems, FWH_SMV_PERSIST=persist, FWH_SMV_SPLIT=2, FWH_SMV_WAVES=4, **LDL_PINS)
        assert b.info["smv_persist"] == persist
        b.solve(eps=1e-8, max_loop=5000, graph_iters=8, schur=True)
        st = b.stats_all()
        say(f"[gpu] matvec batch persist={persist}: iterations {[s['iters'] for s in st]}")
        res.append([(st[k]["iters"], b.select(k).x().copy()) for k in range(2)])
        b.close()
    for (i0, x0), (i1, x1) in zip(*res):
        assert i0 == i1 and np.array_equal(x0, x1)


def sub_ranges(sol, nsub):
    off, out = 0, []
    for s in range(nsub):
        n = sol.sub(s)["n"]
        out.append((off, off + n))
        off += n
    return out


@pytest.mark.parametrize("depth", [16, 8])
def test_solve_persistent_forms_same_bits(small_sigma_model, cfg1_sigma_model, depth):
    for model in (small_sigma_model, cfg1_sigma_model):
        ref1 = loaded(model, FWH_SOLVE_DEPTH=depth, FWH_SOLVE_WAVES=1)
        f1 = loaded(model, FWH_SOLVE_PERSIST=1, FWH_SOLVE_DEPTH=depth)
        if depth == 8 and ref1.info["solve_depth"] != 8:
            pytest.skip("a profile row is wider than the short blocks")
        say(f"[gpu] {model.name} depth {depth}: form 1 grid {f1.info['solve_pgrid']}")
        assert f1.info["solve_persist"] == 1 and f1.info["solve_depth"] == depth
        rng = np.random.default_rng(depth)
        n = int(ref1.info["n_dof"])
        rhs = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        x1 = ref1.ldl_solve_all(rhs)
        y1 = f1.ldl_solve_all(rhs)
        y1b = f1.ldl_solve_all(rhs)  # the queue is back to zero
        assert np.array_equal(y1b, y1)
        same, worst = 0, 0.0
        for s, (a, b) in enumerate(sub_ranges(ref1, model.nsub)):
            err = np.abs(y1[a:b] - x1[a:b]).max() / np.abs(x1[a:b]).max()
            worst = max(worst, err)
            assert err <= 1e-10, f"subdomain {s}: {err:.3e}"
            same += int(np.array_equal(y1[a:b], x1[a:b]))
        say(f"[gpu] bit-identical subdomains: {same} of {model.nsub}; largest difference {worst:.2e}")
        for sol in (ref1, f1):
            sol.close()


def test_solve_persistent_form2_is_off(small_sigma_model):
    """The size-bucketed form is switched off: forcing it is refused."""
    from adventure_fullwave_amd import fwhip
    with pytest.raises(fwhip.FwhError, match="switched off"):
        loaded(small_sigma_model, FWH_SOLVE_PERSIST=2, FWH_SOLVE_NLARGE=2)


@pytest.mark.parametrize("form", [1, pytest.param(2, marks=pytest.mark.skip(
    reason="the size-bucketed persistent form is switched off: an iterative solve through it went silent "
           "on the GPU, cause not established"))])
def test_solve_persistent_matches_oracle(cfg1_sigma_model, form):
    model = cfg1_sigma_model
    env = dict(FWH_SOLVE_PERSIST=form)
    if form == 2:
        env["FWH_SOLVE_NLARGE"] = model.nsub // 4
    sol = loaded(model, **env)
    assert sol.info["solve_persist"] == form
    ref = fo.hddm_solve(model, eps=1e-8, max_loop=20000)
    it_ref = ref["stats"]["iters"]
    for graph in (16, 0):
        st = sol.solve(eps=1e-8, max_loop=20000, graph_iters=graph)
        say(f"[gpu] cfg1 implicit, persistent form {form}, graph {graph}: iters {st['iters']} oracle {it_ref}")
        assert st["converged"] == 1 and abs(st["iters"] - it_ref) <= max(1, 0.02 * it_ref)
        assert np.linalg.norm(sol.x() - ref["x_flat"]) <= 1e-7 * np.linalg.norm(ref["x_flat"])
    sol.close()
