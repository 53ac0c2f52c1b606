"""Every form of the batched LDL^T kernel (run with -m gpu).

The load-time tuner picks the prefetch depth (8 or 16 loads per lane, with
a row-block partition for each) and the waves per subdomain (1, 2 or 4,
with the same row blocks; four waves also over wide blocks of up to 2,048
entries, depth 32) by timing, so a given run may exercise only one form
in the iteration sweeps.  Here each form is forced (FWH_SOLVE_DEPTH for
every launch kind, FWH_SOLVE_WAVES outside the formation, read at load) and
checked against the oracle: the subdomain LDL^T solves (1e-10 relative; one
launch per subdomain gives the same bits), the implicit Schur step and the
explicit S_s formed with it (1e-10), and the full HDDM solve (iteration
count within 2 %, solution rel. L2 1e-7)."""
import os
import sys

import numpy as np
import pytest

from oracle import fw_oracle as fo
from tests.helpers import pinned_env

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sigma_ref(small_sigma_model):
    return fo.hddm_solve(small_sigma_model, eps=1e-8, max_loop=5000)


@pytest.mark.parametrize("depth,waves", [(16, 1), (8, 1), (16, 2), (8, 2), (16, 4), (8, 4), (32, 4)])
def test_solve_variant(small_sigma_model, sigma_ref, depth, waves):
    from adventure_fullwave_amd import fwhip
    model = small_sigma_model
    say(f"[gpu] LDL^T variant depth={depth} waves={waves}")
    with pinned_env({"FWH_SOLVE_DEPTH": str(depth), "FWH_SOLVE_WAVES": str(waves)}):
        sol = fwhip.Solver(model)
    # (depth 32, the wide blocks, is a form of the iteration sweeps only)
    unit_ok = sol.info["unit_depth"] in (8, 16) if depth == 32 else sol.info["unit_depth"] == depth
    if sol.info["solve_depth"] != depth or not unit_ok or sol.info["solve_waves"] != waves:
        sol.close()
        pytest.skip(f"depth {depth} with {waves} waves not available on this model")
    om = fo.OracleModel(model)
    # subdomain LDL^T solves against the oracle's factors
    rng = np.random.default_rng(depth)
    nd = int(sol.info["n_dof"])
    rhs = rng.standard_normal(nd) + 1j * rng.standard_normal(nd)
    x = sol.ldl_solve_all(rhs.copy())
    assert np.array_equal(sol.ldl_solve_all(rhs.copy(), plain=True), x)
    off = 0
    for s in range(model.nsub):
        g = om.get(s)
        n = g["n"]
        xo = fo.sky_solve(g["idx1"], g["idx2"], g["factor"], rhs[off:off + n])
        assert np.abs(x[off:off + n] - xo).max() <= 1e-10 * np.abs(xo).max()
        off += n
    # implicit Schur step, then the explicit S_s formed with this variant
    ng = int(sol.info["n_iface"])
    p = rng.standard_normal(ng) + 1j * rng.standard_normal(ng)
    qi = sol.schur_apply(p)
    qe = sol.schur_apply(p, explicit=True)
    assert np.abs(qe - qi).max() <= 1e-10 * np.abs(qi).max()
    # full solve
    st = sol.solve(eps=1e-8, max_loop=5000, graph_iters=8)
    it_ref = sigma_ref["stats"]["iters"]
    say(f"[gpu] iters gpu={st['iters']} oracle={it_ref}")
    assert st["converged"] == 1
    assert abs(st["iters"] - it_ref) <= max(1, 0.02 * it_ref)
    xs = sol.x()
    assert np.linalg.norm(xs - sigma_ref["x_flat"]) / np.linalg.norm(sigma_ref["x_flat"]) <= 1e-7
    sol.close()


def test_wide_blocks_batch_and_released(small_sigma_model):
    """The four-wave form over the wide row blocks (forced: FWH_SOLVE_DEPTH=32,
    FWH_SOLVE_WAVES=4) where the systems of a batch and the chunked
    recomputation address the blocks: a batch of two systems (two
    frequencies) gives each system's single-solve bits, implicit and
    explicit, and with the factors released in three chunks the explicit
    solve and the right-hand-side solves (batched and one launch per
    subdomain) give the resident bits.  The matvec's own load-time choices
    are pinned too, for the whole test: they are read when the S_s are
    formed, at the first explicit solve."""
    import dataclasses
    from adventure_fullwave_amd import fwhip
    pins = {"FWH_SOLVE_DEPTH": "32", "FWH_SOLVE_WAVES": "4", "FWH_SOLVE_PERSIST": "0", "FWH_SMV_WAVES": "4",
            "FWH_SMV_SPLIT": "0", "FWH_SMV_PERSIST": "0"}
    m = small_sigma_model
    systems = [dataclasses.replace(m, omega=0.95 * m.omega), m]
    with pinned_env(pins):
        b = fwhip.Solver(systems)
        singles = [fwhip.Solver(x) for x in systems]
        assert b.info["solve_depth"] == 32 and b.info["solve_waves"] == 4
        for schur in (False, True):
            b.solve(eps=1e-8, max_loop=5000, graph_iters=8, schur=schur)
            stats = b.stats_all()
            for k, s1 in enumerate(singles):
                assert s1.info["solve_depth"] == 32
                st1 = s1.solve(eps=1e-8, max_loop=5000, graph_iters=8, schur=schur)
                b.select(k)
                say(f"[gpu] wide blocks, batch system {k}, explicit={schur}: "
                    f"{stats[k]['iters']} / {st1['iters']} iterations")
                assert stats[k]["iters"] == st1["iters"] and st1["converged"] == 1
                assert np.array_equal(b.x(), s1.x())
        b.close()
        a = singles[1]
        r = fwhip.Solver(m)
        r.schur_form()
        budget = int(max(r.info["factor_need_max"], r.info["factor_need_total"] / 3))
        r.release_factors(budget)
        assert r.info["factors_resident"] == 0 and r.info["factor_chunks"] >= 2
        st_a = a.solve(eps=1e-8, max_loop=5000, graph_iters=8, schur=True)
        st_r = r.solve(eps=1e-8, max_loop=5000, graph_iters=8, schur=True)
    say(f"[gpu] wide blocks, released in {r.info['factor_chunks']} chunks: {st_r['iters']} / {st_a['iters']} iterations")
    assert st_a["iters"] == st_r["iters"] and np.array_equal(a.x(), r.x())
    rhs = np.random.default_rng(9).standard_normal(int(a.info["n_dof"])) * (1 + 0.5j)
    assert np.array_equal(a.ldl_solve_all(rhs), r.ldl_solve_all(rhs))
    assert np.array_equal(a.ldl_solve_all(rhs, plain=True), r.ldl_solve_all(rhs, plain=True))
    for s1 in singles:
        s1.close()
    r.close()


def test_schur_formation_paths(small_sigma_model):
    """S_s formed on the matrix cores (A_GG - W^T D^-1 W, W by forward
    sweeps) and by the Schur step on every interface unit vector give the same
    explicit operator (1e-12 relative) and both match the implicit one."""
    from adventure_fullwave_amd import fwhip
    say("[gpu] Schur formation: GEMM vs unit-vector steps")
    model = small_sigma_model
    rng = np.random.default_rng(11)
    q = {}
    for form in ("gemm", "solve"):
        if form == "solve":
            os.environ["FWH_SCHUR_FORM"] = "solve"
        try:
            sol = fwhip.Solver(model)
            sol.schur_form()
        finally:
            os.environ.pop("FWH_SCHUR_FORM", None)
        assert sol.info["schur_gemm"] == (1 if form == "gemm" else 0)
        ng = int(sol.info["n_iface"])
        p = np.random.default_rng(11).standard_normal(ng) + 1j * np.random.default_rng(12).standard_normal(ng)
        q[form] = sol.schur_apply(p, explicit=True)
        q["implicit"] = sol.schur_apply(p)
        sol.close()
    scale = np.abs(q["implicit"]).max()
    assert np.abs(q["gemm"] - q["solve"]).max() <= 1e-12 * scale
    assert np.abs(q["gemm"] - q["implicit"]).max() <= 1e-10 * scale


@pytest.mark.parametrize("model_name", ["small_sigma_model", "cfg1_sigma_model"])
def test_factor_forms(request, model_name):
    """The blocked factorisation (16-row blocks on the matrix cores, the
    default) and the row-by-row one (FWH_FACTOR=row, the reference's loop
    order) give the same factors of every subdomain to the factor bar of the
    oracle parity test, 1e-10 relative to the subdomain's largest entry (the
    summation orders differ; measured 2.9e-12 on cfg 1, whose indefinite
    pivots reach 6e5), and solutions of the same accuracy: the
    iteration counts within 2 % and the solutions within rel. L2 1e-7 (the
    oracle bar; COCG carries the factors' last-bit differences into the
    iterates)."""
    from adventure_fullwave_amd import fwhip
    model = request.getfixturevalue(model_name)
    say(f"[gpu] factor forms: blocked vs row ({model_name})")
    fac, res = {}, {}
    for form in ("blocked", "row"):
        if form == "row":
            os.environ["FWH_FACTOR"] = "row"
        try:
            sol = fwhip.Solver(model)
        finally:
            os.environ.pop("FWH_FACTOR", None)
        assert sol.info["factor_blocked"] == (1 if form == "blocked" else 0)
        fac[form] = [sol.sub(s)["factor"] for s in range(model.nsub)]
        st = sol.solve(eps=1e-8, max_loop=20000, graph_iters=16, schur=True)
        res[form] = (st["iters"], sol.x())
        sol.close()
    for a, b in zip(fac["blocked"], fac["row"]):
        assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()
    (ib, xb), (ir, xr) = res["blocked"], res["row"]
    say(f"[gpu] iterations blocked={ib} row={ir}")
    assert abs(ib - ir) <= max(1, 0.02 * ir)
    assert np.linalg.norm(xb - xr) <= 1e-7 * np.linalg.norm(xr)


@pytest.mark.parametrize("schur", [True, False])
def test_gather_forms(small_sigma_model, schur):
    """An interface unknown's slot values are summed in the reference's order
    (hddm_solver_sp.h:41-54, hddm_solver_s.h:404-424) from a four-wide table
    of slot ids when it lies in at most four subdomains, else from the CSR.
    The box partitions here put every interface edge in 2-4 subdomains, so
    FWH_GATHER=csr sends all of them through the CSR loop: both must give
    the same bits (iterations, residual history, solution), in the explicit
    and the implicit form (the explicit one with split matvec items, whose
    slot copies lengthen the lists; the matvec choices that set summation
    orders are pinned for both)."""
    from adventure_fullwave_amd import fwhip
    out = {}
    with pinned_env():
        for form in ("table", "csr"):
            if form == "csr":
                os.environ["FWH_GATHER"] = "csr"
            try:
                sol = fwhip.Solver(small_sigma_model)
            finally:
                os.environ.pop("FWH_GATHER", None)
            st = sol.solve(eps=1e-8, max_loop=20000, graph_iters=16, schur=schur)
            assert not schur or sol.info["smv_split"] == 2
            out[form] = (st["iters"], st["loops"], st["converged"], sol.x().copy(), sol.resid(st["loops"]).copy())
            sol.close()
            say(f"[gpu] gather {form} explicit={schur}: {st['iters']} iterations")
    a, b = out["table"], out["csr"]
    assert a[:3] == b[:3] and a[2] == 1
    assert np.array_equal(a[3], b[3])
    assert np.array_equal(a[4], b[4])
