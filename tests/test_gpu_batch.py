"""Batches of systems on one mesh (fwh_load_batch; run with -m gpu).

Three systems share the small model's mesh and decomposition: two
frequencies of the reference-mode absorber model and the delta-gap variant
at a third frequency.  The batch solve (every launch serving all three) must
reproduce three separate single-system solves: bit for bit when the
load-time kernel choices agree (matvec waves and split, LDL^T prefetch
depth, size buckets: they set summation orders), else with counts and
solutions within twice the system's oracle bars (each of the two meets
them; with other rounding the 1e-8 stop comes some iterations earlier or
later); and each system must match the oracle (count and solution bars per
system from measured spreads, ORACLE_BARS).  Per-system accessors (select)
and the structure check are covered too."""
import dataclasses
import sys

import numpy as np
import pytest

from oracle import fw_oracle as fo
from tests.helpers import pinned_env, with_delta_gap

pytestmark = pytest.mark.gpu

CHOICES = ("smv_waves", "smv_split", "solve_depth", "solve_waves", "unit_depth", "solve_bucketed", "schur_gemm",
           "solve_persist", "solve_nlarge")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def systems(small_sigma_model):
    m = small_sigma_model
    return [dataclasses.replace(m, omega=0.9 * m.omega), m,
            dataclasses.replace(with_delta_gap(m), omega=1.05 * m.omega)]


@pytest.mark.parametrize("schur", [False, True])
def test_batch_matches_single_solves(systems, schur):
    from adventure_fullwave_amd import fwhip
    say(f"[gpu] batch of {len(systems)} systems, explicit={schur}")
    b = fwhip.Solver(systems)
    assert b.info["nsys"] == len(systems)
    b.solve(eps=1e-8, max_loop=5000, graph_iters=8, schur=schur)
    stats_b = b.stats_all()
    for k, m in enumerate(systems):
        s1 = fwhip.Solver(m)
        st1 = s1.solve(eps=1e-8, max_loop=5000, graph_iters=8, schur=schur)
        b.select(k)
        xb, x1 = b.x(), s1.x()
        rb, r1 = b.resid(st1["loops"]), s1.resid(st1["loops"])
        same = all(b.info[c] == s1.info[c] for c in CHOICES)
        say(f"[gpu] system {k}: iters batch={stats_b[k]['iters']} single={st1['iters']} same_choices={same} "
            f"{ {c: (b.info[c], s1.info[c]) for c in CHOICES} }")
        assert stats_b[k]["converged"] == 1 and st1["converged"] == 1
        if same:
            assert stats_b[k]["iters"] == st1["iters"]
            assert np.array_equal(xb, x1)
            assert np.array_equal(rb, r1)
        else:
            assert abs(stats_b[k]["iters"] - st1["iters"]) <= max(1, 2 * ORACLE_BARS[k][0] * st1["iters"])
            assert np.linalg.norm(xb - x1) <= 2 * ORACLE_BARS[k][1] * np.linalg.norm(x1)
        s1.close()
    b.close()


# Per-system bars against the oracle (count, solution rel. L2).  The
# delta-gap system at 1.05 omega has a residual plateau near 5e-8 before the
# 1e-8 stop, so its count is bimodal: over 12 few-ulp permittivity
# perturbations both GPU paths stop at 124, 125 or 129 iterations (the
# oracle's own count moves 124-125 over 18), and its 1e-8-stopped solution
# lies 1.5e-7 from the fully converged one (the other two: 7e-9 and 2.7e-8),
# so two correct stops can differ by twice that.
ORACLE_BARS = [(0.02, 1e-7), (0.02, 1e-7), (0.05, 3e-7)]


def test_batch_systems_match_oracle(systems):
    from adventure_fullwave_amd import fwhip
    say("[gpu] batch systems vs oracle")
    b = fwhip.Solver(systems)
    b.solve(eps=1e-8, max_loop=5000, graph_iters=8, schur=True)
    for k, m in enumerate(systems):
        ref = fo.hddm_solve(m, eps=1e-8, max_loop=5000)
        st = b.select(k).stats()
        it_ref = ref["stats"]["iters"]
        say(f"[gpu] system {k}: iters gpu={st['iters']} oracle={it_ref}")
        count_bar, x_bar = ORACLE_BARS[k]
        assert abs(st["iters"] - it_ref) <= max(1, count_bar * it_ref)
        x = b.x()
        assert np.linalg.norm(x - ref["x_flat"]) / np.linalg.norm(ref["x_flat"]) <= x_bar
    b.close()


def test_batch_per_system_operators(systems):
    """post, schur_apply and get_sub address the selected system (the
    matvec choices pinned, so that batch and single solves round alike)"""
    from adventure_fullwave_amd import fwhip
    say("[gpu] batch per-system operators")
    with pinned_env():
        b = fwhip.Solver(systems)
        b.solve(eps=1e-8, max_loop=5000, graph_iters=8, schur=True)
    rng = np.random.default_rng(5)
    ng = int(b.info["n_iface"])
    p = rng.standard_normal(ng) + 1j * rng.standard_normal(ng)
    for k in (2, 0):
        m = systems[k]
        with pinned_env():
            s1 = fwhip.Solver(m)
            s1.solve(eps=1e-8, max_loop=5000, graph_iters=8, schur=True)
        b.select(k)
        for fb, f1 in zip(b.post(), s1.post()):
            assert np.array_equal(fb, f1)
        qb, q1 = b.schur_apply(p), s1.schur_apply(p)
        assert np.array_equal(qb, q1)
        for s in (0, m.nsub - 1):
            gb, g1 = b.sub(s), s1.sub(s)
            for key in ("val", "f", "factor"):
                assert np.array_equal(gb[key], g1[key])
        s1.close()
    b.close()


def test_batch_rejects_other_mesh(small_sigma_model, cfg1_sigma_model):
    from adventure_fullwave_amd import fwhip
    say("[gpu] batch structure check")
    with pytest.raises(fwhip.FwhError, match="differs from system 0"):
        fwhip.Solver([small_sigma_model, cfg1_sigma_model])
    b = fwhip.Solver([small_sigma_model, small_sigma_model])
    with pytest.raises(fwhip.FwhError, match="out of range"):
        b.select(2)
    b.close()
