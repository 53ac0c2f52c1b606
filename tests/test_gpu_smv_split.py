"""Explicit Schur matvec with split subdomains (run with -m gpu).

The matvec can split the largest subdomains' Schur complements into work
items over contiguous tile ranges (each writing its own slot copies and
(p,q) part, summed by the iteration's gather).  The load-time timing keeps a
split only where it pays; here it is forced (FWH_SMV_SPLIT, read when the
S_s are formed) on models where it cuts every subdomain into several items:
the Schur apply must match the unsplit one to rounding (1e-13 relative), the
HDDM solve (batched with graphs, and plain: one launch per subdomain with
its items) must match the oracle as the unsplit form does (count within 2 %,
solution rel. L2 1e-7), and a batch of two systems must split both alike."""
import dataclasses
import os
import sys

import numpy as np
import pytest

from oracle import fw_oracle as fo

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def formed(model, split):
    from adventure_fullwave_amd import fwhip
    keep = os.environ.get("FWH_SMV_SPLIT")
    os.environ["FWH_SMV_SPLIT"] = str(split)
    try:
        sol = fwhip.Solver(model)
        sol.schur_form()
    finally:
        if keep is None:
            del os.environ["FWH_SMV_SPLIT"]
        else:
            os.environ["FWH_SMV_SPLIT"] = keep
    return sol


@pytest.mark.parametrize("split", [2, 8])
def test_split_schur_apply_and_solve(small_sigma_model, cfg1_sigma_model, split):
    for model in (small_sigma_model, cfg1_sigma_model):
        s0 = formed(model, 0)
        s1 = formed(model, split)
        say(f"[gpu] {model.name} split m={split}: {s1.info['smv_nwork']} items for {model.nsub} subdomains")
        assert s0.info["smv_split"] == 0 and s0.info["smv_nwork"] == model.nsub
        assert s1.info["smv_split"] == split and s1.info["smv_nwork"] > model.nsub
        rng = np.random.default_rng(split)
        ng = int(s0.info["n_iface"])
        p = rng.standard_normal(ng) + 1j * rng.standard_normal(ng)
        q0 = s0.schur_apply(p, explicit=True)
        q1 = s1.schur_apply(p, explicit=True)
        assert np.abs(q1 - q0).max() <= 1e-13 * np.abs(q0).max()
        q1p = s1.schur_apply(p, explicit=True, plain=True)
        assert np.abs(q1p - q1).max() <= 1e-13 * np.abs(q0).max()
        ref = fo.hddm_solve(model, eps=1e-8, max_loop=20000)
        it_ref = ref["stats"]["iters"]
        for plain, graph in ((False, 16), (True, 0)):
            st = s1.solve(eps=1e-8, max_loop=20000, graph_iters=graph, schur=True, plain=plain)
            say(f"[gpu] {model.name} split m={split} plain={plain}: iters {st['iters']} oracle {it_ref}")
            assert st["converged"] == 1 and abs(st["iters"] - it_ref) <= max(1, 0.02 * it_ref)
            x = s1.x()
            assert np.linalg.norm(x - ref["x_flat"]) <= 1e-7 * np.linalg.norm(ref["x_flat"])
        s0.close()
        s1.close()


def test_split_batch(small_sigma_model):
    m = small_sigma_model
    systems = [m, dataclasses.replace(m, omega=1.05 * m.omega)]
    b = formed(systems, 4)
    assert b.info["smv_split"] == 4
    b.solve(eps=1e-8, max_loop=5000, graph_iters=8, schur=True)
    for k, mk in enumerate(systems):
        s1 = formed(mk, 4)
        st1 = s1.solve(eps=1e-8, max_loop=5000, graph_iters=8, schur=True)
        stb = b.select(k).stats()
        say(f"[gpu] split batch system {k}: iters batch={stb['iters']} single={st1['iters']}")
        assert stb["converged"] == 1 and abs(stb["iters"] - st1["iters"]) <= max(1, 0.02 * st1["iters"])
        assert np.linalg.norm(b.x() - s1.x()) <= 1e-7 * np.linalg.norm(s1.x())
        s1.close()
    b.close()


def test_split_independent_of_work_order(small_sigma_model):
    """The matvec's work order (picked at formation by timing) only
    schedules: with split items, their slot copies and (p,q) parts are
    numbered in subdomain order, so every order gives the same bits."""
    from adventure_fullwave_amd import fwhip
    out = []
    for order in (0, 1, 2):
        keep = os.environ.get("FWH_SMV_ORDER")
        os.environ["FWH_SMV_ORDER"] = str(order)
        try:
            sol = formed(small_sigma_model, 2)
        finally:
            if keep is None:
                del os.environ["FWH_SMV_ORDER"]
            else:
                os.environ["FWH_SMV_ORDER"] = keep
        assert sol.info["smv_order"] == order and sol.info["smv_split"] == 2
        rng = np.random.default_rng(3)
        ng = int(sol.info["n_iface"])
        p = rng.standard_normal(ng) + 1j * rng.standard_normal(ng)
        q = sol.schur_apply(p, explicit=True)
        st = sol.solve(eps=1e-8, max_loop=5000, graph_iters=16, schur=True)
        say(f"[gpu] split m=2 order {order}: {st['iters']} iterations")
        out.append((q, st["iters"], sol.x().copy(), sol.resid(st["loops"]).copy()))
        sol.close()
    for q, it, x, r in out[1:]:
        assert np.array_equal(q, out[0][0]) and it == out[0][1]
        assert np.array_equal(x, out[0][2]) and np.array_equal(r, out[0][3])
