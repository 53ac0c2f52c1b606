"""The reference-shaped interface on the GPU (run with -m gpu).

hddm_solver_CS_hip, called by the C driver tests/refdrv.c with the
reference's own structs (PartMesh / DomMesh / MTRL / OPinfo / Inbc /
DirichletBC / ANTENNA, lib/std/types.h) as hddm_Init leaves them, must give
the same bits as fwh_solve on the flat description: the per-subdomain
solutions written to DomMesh.x_c, the element fields written to
DomMesh.rvalue (post_sp), and the residual history written to
HDDM_Opt.fp in the reference's log format (hddm_solver_s.h:205,321).  The
load-time kernel choices that set summation orders are pinned for both
(FWH_SMV_WAVES, FWH_SMV_SPLIT, FWH_SOLVE_DEPTH).  advlas_skyline_ldl_decomposite / _solve
with DMatrix arguments against the oracle's restatement of
advlas_skyline.c:79-167."""
import os
import sys

import numpy as np
import pytest

from oracle import fw_oracle as fo
from tests import refapi
from tests.helpers import with_delta_gap
from tests.test_oracle_skyline import dense_from_sky, random_profile
from tests.test_refapi import write_volt

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture()
def pinned(tmp_path):
    """pinned kernel choices and a working directory with MODEL/ (volt.dat)"""
    keep = {k: os.environ.get(k) for k in ("FWH_SMV_WAVES", "FWH_SMV_SPLIT", "FWH_SOLVE_DEPTH", "FWH_SOLVE_WAVES")}
    os.environ["FWH_SMV_WAVES"] = "4"
    os.environ["FWH_SMV_SPLIT"] = "2"
    os.environ["FWH_SOLVE_DEPTH"] = "16"
    os.environ["FWH_SOLVE_WAVES"] = "1"
    (tmp_path / "MODEL").mkdir()
    old = os.getcwd()
    os.chdir(tmp_path)
    yield tmp_path
    os.chdir(old)
    for k, v in keep.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


def check_same_bits(model, tmp_path, eps=1e-8, max_loop=20000):
    from adventure_fullwave_amd import fwhip
    logpath = str(tmp_path / "calc_log")
    rc, st, x, B, E, J = refapi.solve(model, eps=eps, max_loop=max_loop, logpath=logpath)
    assert rc == 0, fwhip.lib().fwh_last_error()
    sol = fwhip.Solver(model)
    st1 = sol.solve(eps=eps, max_loop=max_loop, graph_iters=16, schur=True)
    say(f"[gpu] {model.name}: hddm_solver_CS_hip {st['iters']} iterations, fwh_solve {st1['iters']}")
    assert st["iters"] == st1["iters"] and st["converged"] == 1
    assert np.array_equal(x, sol.x())
    for a, b in zip((B, E, J), sol.post()):
        assert np.array_equal(a, b)
    r = sol.resid(st1["loops"])
    lines = open(logpath).read().splitlines()
    assert lines[0] == f"# DOF on interface {int(sol.info['n_iface'])}"
    assert lines[1] == f"0 {1.0:e}  # Norm {st1['r0']:e}"
    assert len(lines) == st1["loops"] + 1
    for k in (1, st1["loops"] // 2, st1["loops"] - 1):
        loop, val = lines[k + 1].split()
        assert int(loop) == k and float(val) == float(f"{r[k]:e}")
    sol.close()


def test_hddm_solver_cs_hip_same_bits(small_sigma_model, cfg1_sigma_model, pinned):
    write_volt(pinned, [])
    for model in (small_sigma_model, cfg1_sigma_model):
        check_same_bits(model, pinned)


def test_hddm_solver_cs_hip_delta_gap(small_sigma_model, pinned):
    gap = with_delta_gap(small_sigma_model)
    h, c = 1.0 / 8, 0.5
    write_volt(pinned, [{"lo": np.array([c - 0.6 * h, c - 0.6 * h, c - 1.1 * h]),
                         "hi": np.array([c + 0.6 * h, c + 0.6 * h, c + 1.1 * h]), "dV": 1.0},
                        {"lo": np.array([c - 0.6 * h, c - 0.6 * h, c - 0.1 * h]),
                         "hi": np.array([c + 0.6 * h, c + 0.6 * h, c + 1.1 * h]), "dV": 2.0}])
    check_same_bits(gap, pinned)


def check_hooks_same_bits(model, tmp_path, eps=1e-8, max_loop=20000):
    """the reference's driver loop over set_hddm_func_fullwave_hip against
    the whole-call entry point: the same main-loop count, solution, element
    fields and residual log"""
    from adventure_fullwave_amd import fwhip
    log_call, log_hooks = str(tmp_path / "calc_log_call"), str(tmp_path / "calc_log_hooks")
    rc, st, x, B, E, J = refapi.solve(model, eps=eps, max_loop=max_loop, logpath=log_call)
    assert rc == 0, fwhip.lib().fwh_last_error()
    iters, xh, Bh, Eh, Jh = refapi.solve_hooks(model, eps=eps, max_loop=max_loop, logpath=log_hooks)
    say(f"[gpu] {model.name} (max_loop {max_loop}): hook table {iters} loops, hddm_solver_CS_hip {st['iters']}")
    assert iters == st["iters"]
    assert np.array_equal(x, xh)
    for a, b in zip((B, E, J), (Bh, Eh, Jh)):
        assert np.array_equal(a, b)
    assert open(log_call).read() == open(log_hooks).read()


def test_hook_table_same_bits(small_sigma_model, cfg1_sigma_model, pinned):
    """set_hddm_func_fullwave_hip under the reference's parent loop
    (hddm_solver.h:59-183, restated in tests/refdrv.c): the per-subdomain
    hooks queue, the iteration hooks run the GPU ahead in graphs and set
    hddm_sw.status to LAST at the loop where the GPU's stop test fired; the
    results equal hddm_solver_CS_hip's bit for bit, for a converged solve
    (small model, cfg 1) and for one stopped by hddm_max_loop at a loop
    that is not a multiple of the graph length."""
    write_volt(pinned, [])
    for model in (small_sigma_model, cfg1_sigma_model):
        check_hooks_same_bits(model, pinned)
    check_hooks_same_bits(small_sigma_model, pinned, max_loop=37)


def test_hddm_solver_cs_hip_not_executed(small_sigma_model, pinned):
    """hddm_max_loop = 0: "HDDM Solver was not executed!!!", no results
    (hddm_solver.h:103-105)"""
    write_volt(pinned, [])
    rc, st, x, B, E, J = refapi.solve(small_sigma_model, max_loop=0)
    assert rc == 0 and st["iters"] == 0 and not x.any()


@pytest.mark.parametrize("n,maxw,seed", [(1, 0, 0), (17, 5, 2), (200, 40, 4), (1500, 120, 5)])
def test_advlas_skyline_on_gpu(n, maxw, seed):
    say(f"[gpu] advlas skyline n={n}")
    rng = np.random.default_rng(seed)
    idx1, idx2, nlen = random_profile(rng, n, maxw)
    mat = rng.standard_normal(nlen) + 1j * rng.standard_normal(nlen)
    A = dense_from_sky(idx1, idx2, mat)
    A += np.diag(2.0 * np.abs(A).sum(axis=1) * np.exp(1j * rng.uniform(0, np.pi, n)))
    mat = np.array([A[i, j] for i in range(n) for j in range(idx1[i], i + 1)])
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    fac, work, x = refapi.advlas(idx1, idx2, mat, b)
    ref = fo.sky_factor(idx1, idx2, mat)
    assert np.abs(fac - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(work - ref[idx2 + np.arange(n)]).max() <= 1e-12 * np.abs(work).max()
    xr = fo.sky_solve(idx1, idx2, ref, b)
    assert np.abs(x - xr).max() <= 1e-12 * np.abs(xr).max()
    assert np.abs(A @ x - b).max() <= 1e-11 * np.abs(b).max()


def test_advlas_on_subdomain_factors(small_sigma_model):
    """the GPU skyline calls on every subdomain's Dirichlet matrix of the
    small model (profiles after SetBC, as the reference factors them)"""
    om = fo.OracleModel(small_sigma_model)
    rng = np.random.default_rng(8)
    for s in range(small_sigma_model.nsub):
        g = om.get(s)
        n = g["n"]
        # the unfactored skyline: rebuilt from the factor L D L^T
        L = np.eye(n, dtype=complex)
        dinv = np.zeros(n, complex)
        for i in range(n):
            for j in range(g["idx1"][i], i):
                L[i, j] = g["factor"][g["idx2"][i] + j]
            dinv[i] = g["factor"][g["idx2"][i] + i]
        A = L @ np.diag(1.0 / dinv) @ L.T
        mat = np.array([A[i, j] for i in range(n) for j in range(g["idx1"][i], i + 1)])
        b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        fac, _, x = refapi.advlas(g["idx1"], g["idx2"], mat, b)
        ref = fo.sky_factor(g["idx1"], g["idx2"], mat)
        assert np.abs(fac - ref).max() <= 1e-10 * np.abs(ref).max()
        xr = fo.sky_solve(g["idx1"], g["idx2"], ref, b)
        assert np.abs(x - xr).max() <= 1e-10 * np.abs(xr).max()


@pytest.mark.parametrize("pc", [3, 0])
def test_mymt_cocg_on_subdomain_matrices(small_sigma_model, pc):
    """mymt_CS_AIJ_Solver_COCG on the reference's AIJ_MAT (filled by the C
    driver) for every subdomain's Dirichlet matrix of the small model (the
    reference's SetBC + DeleteZero + Assem storage), with IC(0) and with no
    preconditioner: iteration counts against the oracle's COCG within 6 %
    (the CPU spread under summation-order permutations, as for the batched
    subdomain COCG), solutions against a dense solve (rel. L2 1e-7 at a 1e-9
    stop), and an unknown preconditioner returns SOLVER_STATUS_UNKNOWN_PC."""
    from tests.helpers import dense_lower
    from tests.test_gpu_parity import dirichlet_csr
    om = fo.OracleModel(small_sigma_model)
    rng = np.random.default_rng(12)
    for s in range(small_sigma_model.nsub):
        rp, cl, vl = dirichlet_csr(small_sigma_model, om, s)
        n = len(rp) - 1
        b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        st, x = refapi.mymt_cocg(rp, cl, vl, b, pc=pc, eps=1e-9, ic_shift=1.2)
        sto, xo, _ = fo.cocg_icc(rp, cl, vl, b, eps=1e-9, ic_shift=1.2, use_icc=pc == 3)
        say(f"[gpu] mymt COCG pc={pc} subdomain {s}: {st} iterations, oracle {sto}")
        assert st > 0 and sto > 0
        assert abs(st - sto) <= max(2, 0.06 * sto)
        xd = np.linalg.solve(dense_lower(rp, cl, vl, n), b)
        assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-7
    st, _ = refapi.mymt_cocg(rp, cl, vl, b, pc=4)
    assert st == -4


@pytest.mark.parametrize("pc", [3, 0])
def test_mymt_cocg_nmpcm_reuses_the_factor(small_sigma_model, pc):
    """The two subdomain solver hooks in the reference's order
    (HDDM_FUNC.dom_solver, then dom_solver_nmpcm on the same AIJ_MAT):
    mymt_CS_AIJ_Solver_COCG leaves the IC(0) factor in aij_mat->pc_c in the
    reference's layout (L in the matrix's pattern, then d = 1/diag(L)), which
    matches the oracle's make step (fwo_icc_make, aij_pc_func.h:36-102) to
    1e-10 relative, L and d each; mymt_CS_AIJ_Solver_COCG_nmpcm with it on
    another right-hand side returns the same value and bits as a
    make-and-solve call on that right-hand side (same factor, same kernel);
    with no preconditioner nothing is left and the same holds."""
    from tests.test_gpu_parity import dirichlet_csr
    om = fo.OracleModel(small_sigma_model)
    rng = np.random.default_rng(21)
    for s in (0, small_sigma_model.nsub - 1):
        rp, cl, vl = dirichlet_csr(small_sigma_model, om, s)
        n = len(rp) - 1
        b1 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        b2 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        st1, x1, st2, x2, pcm = refapi.mymt_cocg_nmpcm(rp, cl, vl, b1, b2, pc=pc, eps=1e-9, ic_shift=1.2)
        say(f"[gpu] mymt COCG then _nmpcm, pc={pc}, subdomain {s}: {st1} and {st2} iterations")
        assert st1 > 0 and st2 > 0
        if pc == 3:
            ref = fo.icc_make(rp, cl, vl, ic_shift=1.2)
            nnz = int(rp[n])
            assert pcm is not None and len(pcm) == len(ref)
            for part in (slice(0, nnz), slice(nnz, nnz + n)):
                err = np.abs(pcm[part] - ref[part]).max() / np.abs(ref[part]).max()
                say(f"[gpu]   IC(0) factor part vs oracle: {err:.2e}")
                assert err <= 1e-10
        else:
            assert pcm is None
        st_a, x_a = refapi.mymt_cocg(rp, cl, vl, b1, pc=pc, eps=1e-9, ic_shift=1.2)
        st_b, x_b = refapi.mymt_cocg(rp, cl, vl, b2, pc=pc, eps=1e-9, ic_shift=1.2)
        assert (st1, st2) == (st_a, st_b)
        assert np.array_equal(x1, x_a) and np.array_equal(x2, x_b)
