"""GPU parity of libfwhip against the oracle (run with -m gpu).

Through the C ABI: element kernel, assembled subdomain matrices, Dirichlet
profile (bit-exact), skyline factor, batched and plain LDL^T solves, the
Schur step (implicit and explicit), the full HDDM solve (iteration count,
residual history, solution) on the same seeded models, and the element
post-processing.  Tolerances: element/assembly/post 1e-12 relative (different
but exact-in-math formulas and FMA contraction), factor 1e-10 relative,
solves and Schur steps 1e-10 relative; HDDM iteration count within 2 % with
the reference-mode absorber (4 % / 8 % for the rounding-chaotic
stretched-PML small model / cfg 1, whose counts themselves spread
3.4 % / 7.2 % under few-ulp perturbations), solution within rel. L2 1e-7 of
the oracle."""
import sys

import numpy as np
import pytest

from oracle import fw_oracle as fo

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def fw():
    from adventure_fullwave_amd import fwhip
    return fwhip


@pytest.fixture(scope="module")
def small(small_model, fw):
    say("[gpu] load small model")
    return fw.Solver(small_model)


@pytest.fixture(scope="module")
def small_oracle(small_model):
    return fo.OracleModel(small_model)


def test_element_batch(fw):
    say("[gpu] element batch")
    from tests.test_oracle_element import random_tet
    rng = np.random.default_rng(3)
    nel = 64
    xyz = np.stack([random_tet(rng, scale=1 + (k % 5)) for k in range(nel)])
    eps = rng.uniform(1, 10, nel) * 8.854e-12
    sig = rng.uniform(0, 0.1, nel)
    pml = np.tile([1.0, 0.0, 1.0, -0.3, 1.0, -0.1], (nel, 1))
    pml[::2] = [1, 0, 1, 0, 1, 0]
    src = (np.arange(nel) % 3 == 0).astype(np.int32)
    joe = rng.standard_normal((nel, 24))
    ctx = fw.Context(0)
    ae, be = ctx.element_batch(xyz, eps, sig, 2.1e9, pml=pml, src=src, joe=joe)
    for e in range(nel):
        a0, b0 = fo.element(xyz[e], eps[e], sig[e], 2.1e9, pml=pml[e], src=src[e], joe=joe[e])
        assert np.abs(ae[e] - a0).max() <= 1e-12 * np.abs(a0).max()
        assert np.abs(be[e] - b0).max() <= 1e-12 * max(np.abs(b0).max(), 1e-300)
    ctx.close()


def test_assembly_profile_factor(small, small_oracle, small_model):
    say("[gpu] assembly/profile/factor")
    for s in range(small_model.nsub):
        g = small.sub(s)
        o = small_oracle.get(s)
        assert np.array_equal(g["rowptr"], o["rowptr"]) and np.array_equal(g["col"], o["col"])
        assert np.abs(g["val"] - o["val"]).max() <= 1e-12 * np.abs(o["val"]).max()
        assert np.abs(g["f"] - o["f"]).max() <= 1e-12 * max(np.abs(o["f"]).max(), 1e-300)
        assert np.array_equal(g["idx1"], o["idx1"]) and np.array_equal(g["idx2"], o["idx2"])
        assert g["nlen"] == o["nlen"]
        assert np.abs(g["factor"] - o["factor"]).max() <= 1e-10 * np.abs(o["factor"]).max()


@pytest.mark.parametrize("plain", [False, True])
def test_ldl_solve_all(small, small_oracle, small_model, plain):
    say(f"[gpu] ldl solve all plain={plain}")
    rng = np.random.default_rng(4)
    n = int(small.info["n_dof"])
    rhs = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x = small.ldl_solve_all(rhs, plain=plain)
    off = 0
    for s in range(small_model.nsub):
        o = small_oracle.get(s)
        ns = o["n"]
        ref = fo.sky_solve(o["idx1"], o["idx2"], o["factor"], rhs[off:off + ns])
        assert np.abs(x[off:off + ns] - ref).max() <= 1e-10 * np.abs(ref).max()
        off += ns


def oracle_schur(model, om, p):
    """q = sum_s R_s^T (A x_s)|_G with p given per owned unknown (part-major)."""
    goff = np.concatenate([[0], np.cumsum(model.infree)])
    # map (part, ln) -> global unknown via ownership and op lists
    npart = model.npart
    lnmap = []
    for pp in range(npart):
        tf = model.infree[pp] + model.outfree[pp]
        m = -np.ones(tf, np.int64)
        m[model.outfree[pp]:] = goff[pp] + np.arange(model.infree[pp])
        lnmap.append(m)
    for pp in range(npart):
        for o in range(pp):
            a = model.op_ibid[model.op_off[pp * npart + o]:model.op_off[pp * npart + o + 1]]
            b = model.op_ibid[model.op_off[o * npart + pp]:model.op_off[o * npart + pp + 1]]
            for lp, lo in zip(a, b):
                if lnmap[pp][lp] < 0 and lo >= model.outfree[o]:
                    lnmap[pp][lp] = lnmap[o][lo]
    q = np.zeros_like(p)
    for s in range(model.nsub):
        pp = model.sub_part[s]
        ln = model.ib_ln[model.ib_off[s]:model.ib_off[s + 1]]
        nf = model.ib_nfid[model.ib_off[s]:model.ib_off[s + 1]]
        gid = lnmap[pp][ln]
        _, y = om.schur_sub(s, p[gid])
        np.add.at(q, gid, y[nf])
    return q


@pytest.mark.parametrize("explicit", [False, True])
@pytest.mark.parametrize("plain", [False, True])
def test_schur_apply(small, small_oracle, small_model, plain, explicit):
    say(f"[gpu] schur apply plain={plain} explicit={explicit}")
    rng = np.random.default_rng(5)
    ng = int(small.info["n_iface"])
    p = rng.standard_normal(ng) + 1j * rng.standard_normal(ng)
    q = small.schur_apply(p, plain=plain, explicit=explicit)
    ref = oracle_schur(small_model, small_oracle, p)
    assert np.abs(q - ref).max() <= 1e-10 * np.abs(ref).max()


def test_schur_explicit_symmetric_linear(small, small_model):
    """The explicit S (formed from unit-vector Schur steps, lower half kept)
    is complex-symmetric and linear: x^T S y = y^T S x, S(a x + y) = a Sx + Sy."""
    say("[gpu] explicit schur symmetry/linearity")
    rng = np.random.default_rng(7)
    ng = int(small.info["n_iface"])
    x = rng.standard_normal(ng) + 1j * rng.standard_normal(ng)
    y = rng.standard_normal(ng) + 1j * rng.standard_normal(ng)
    sx = small.schur_apply(x, explicit=True)
    sy = small.schur_apply(y, explicit=True)
    assert abs(x @ sy - y @ sx) <= 1e-12 * abs(x @ sy)
    a = 0.3 - 1.7j
    sxy = small.schur_apply(a * x + y, explicit=True)
    assert np.abs(sxy - (a * sx + sy)).max() <= 1e-12 * np.abs(sxy).max()
    assert small.info["ms_schur_form"] > 0 and small.info["smat_bytes"] > 0


@pytest.fixture(scope="module")
def small_refs(small_model, small_sigma_model):
    return {"stretch": (small_model, fo.hddm_solve(small_model, eps=1e-8, max_loop=5000)),
            "sigma": (small_sigma_model, fo.hddm_solve(small_sigma_model, eps=1e-8, max_loop=5000))}


# Iteration-count bars.  Reference-mode absorber: 2 % (the oracle's own count
# moves 213-216 over 24 few-ulp load perturbations).  Stretched PML: 4 %: over
# 16 few-ulp perturbations of eps the oracle's count moves 457-465, the GPU
# implicit path's 456-462 and the GPU explicit path's (S_s formed as
# A_GG - W^T D^-1 W, exactly symmetric by construction) 450-466, 3.4 % of the
# oracle's nominal 464; the unperturbed explicit run gives 450.
COUNT_BAR = {"sigma": 0.02, "stretch": 0.04}


@pytest.mark.parametrize("pml", ["sigma", "stretch"])
@pytest.mark.parametrize("mode", ["batched_graph", "batched", "plain", "explicit_graph", "explicit_plain"])
def test_hddm_solve_small(fw, small_refs, mode, pml):
    say(f"[gpu] hddm solve small {mode} pml={pml}")
    model, ref = small_refs[pml]
    sol = fw.Solver(model)
    st = sol.solve(eps=1e-8, max_loop=5000, plain=mode.endswith("plain"),
                   graph_iters=(8 if mode.endswith("graph") else 0), schur=mode.startswith("explicit"))
    it_ref = ref["stats"]["iters"]
    say(f"[gpu] iters gpu={st['iters']} oracle={it_ref}")
    assert st["converged"] == 1
    assert abs(st["iters"] - it_ref) <= max(1, COUNT_BAR[pml] * it_ref)
    r = sol.resid(st["loops"])
    k = min(30, len(r), len(ref["resid"]))
    assert np.allclose(r[:k], ref["resid"][:k], rtol=1e-6)
    x = sol.x()
    xo = ref["x_flat"]
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-7
    sol.close()


@pytest.fixture(scope="module")
def cfg1_ref(cfg1_model):
    return fo.hddm_solve(cfg1_model, eps=1e-8, max_loop=20000)


@pytest.fixture(scope="module")
def cfg1_sigma_ref(cfg1_sigma_model):
    return fo.hddm_solve(cfg1_sigma_model, eps=1e-8, max_loop=20000)


@pytest.mark.parametrize("schur", [False, True])
def test_hddm_cfg1_matches_oracle(fw, cfg1_sigma_model, cfg1_sigma_ref, schur):
    """cfg 1 with the reference-mode absorber: iteration count within 2 % and
    solution within rel. L2 1e-7 of the oracle (the oracle's own count is 417
    for 17 of 18 few-ulp load perturbations, 423 for one)."""
    say(f"[gpu] hddm cfg1 sigma explicit={schur}")
    ref = cfg1_sigma_ref
    sol = fw.Solver(cfg1_sigma_model)
    st = sol.solve(eps=1e-8, max_loop=20000, graph_iters=16, schur=schur)
    it_ref = ref["stats"]["iters"]
    say(f"[gpu] cfg1 sigma iters gpu={st['iters']} oracle={it_ref}")
    assert st["converged"] == 1
    assert abs(st["iters"] - it_ref) <= max(1, 0.02 * it_ref)
    x = sol.x()
    assert np.linalg.norm(x - ref["x_flat"]) / np.linalg.norm(ref["x_flat"]) <= 1e-7
    sol.close()


@pytest.mark.parametrize("schur", [False, True])
def test_hddm_cfg1_stretch(fw, cfg1_model, cfg1_ref, schur):
    """cfg 1 with the complex-stretched PML.  The COCG iteration count of this
    case is rounding-chaotic: the oracle's own count spans 3150-3380 (7.2 % of
    its nominal 3199) over 18 few-ulp load perturbations, so the count bar is
    8 %; the solution bar stays rel. L2 1e-7, and the explicit and implicit
    Schur operators must agree to 1e-10."""
    say(f"[gpu] hddm cfg1 stretch explicit={schur}")
    ref = cfg1_ref
    sol = fw.Solver(cfg1_model)
    st = sol.solve(eps=1e-8, max_loop=20000, graph_iters=16, schur=schur)
    it_ref = ref["stats"]["iters"]
    say(f"[gpu] cfg1 stretch iters gpu={st['iters']} oracle={it_ref}")
    assert st["converged"] == 1
    assert abs(st["iters"] - it_ref) <= 0.08 * it_ref
    x = sol.x()
    assert np.linalg.norm(x - ref["x_flat"]) / np.linalg.norm(ref["x_flat"]) <= 1e-7
    if schur:
        rng = np.random.default_rng(9)
        ng = int(sol.info["n_iface"])
        p = rng.standard_normal(ng) + 1j * rng.standard_normal(ng)
        qi = sol.schur_apply(p)
        qe = sol.schur_apply(p, explicit=True)
        assert np.abs(qe - qi).max() <= 1e-10 * np.abs(qi).max()
    sol.close()


def test_max_loop_semantics(fw, small_model):
    say("[gpu] max loop semantics")
    sol = fw.Solver(small_model)
    st = sol.solve(eps=1e-30, max_loop=7, graph_iters=4)
    assert st["iters"] == 7 and st["converged"] == 0 and st["loops"] == 8
    ref = fo.hddm_solve(small_model, eps=1e-30, max_loop=7)
    x = sol.x()
    assert np.linalg.norm(x - ref["x_flat"]) / np.linalg.norm(ref["x_flat"]) <= 1e-9
    st0 = sol.solve(eps=1e-8, max_loop=0)
    assert st0["iters"] == 0
    sol.close()


def dirichlet_csr(model, om, s):
    """The reference's AIJ_MAT after SetBC + DeleteZero + Assem."""
    g = om.get(s)
    n = g["n"]
    bc = np.zeros(n, bool)
    bc[model.ib_nfid[model.ib_off[s]:model.ib_off[s + 1]]] = True
    bc[model.ibc[model.bc_off[s]:model.bc_off[s + 1]]] = True
    rp, cl, vl = [0], [], []
    for i in range(n):
        if bc[i]:
            cl.append(i)
            vl.append(1.0 + 0j)
        else:
            for k in range(g["rowptr"][i], g["rowptr"][i + 1]):
                j = g["col"][k]
                if not bc[j] and g["val"][k] != 0:
                    cl.append(j)
                    vl.append(g["val"][k])
        rp.append(len(cl))
    return np.array(rp, np.int32), np.array(cl, np.int32), np.array(vl)


def test_cocg_icc_all(small, small_oracle, small_model):
    say("[gpu] subdomain cocg+icc")
    rng = np.random.default_rng(6)
    n = int(small.info["n_dof"])
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    from adventure_fullwave_amd import fwhip
    x = np.zeros(n, np.complex128)
    st = np.zeros(small_model.nsub, np.int32)
    bb = np.ascontiguousarray(b)
    fwhip._check(fwhip.lib().fwh_cocg_icc_all(small.ctx.h, bb.ctypes.data, x.ctypes.data, st.ctypes.data,
                                               1e-9, 1.2), "fwh_cocg_icc_all")
    off = 0
    for s in range(small_model.nsub):
        rp, cl, vl = dirichlet_csr(small_model, small_oracle, s)
        ns = len(rp) - 1
        sto, xo, _ = fo.cocg_icc(rp, cl, vl, b[off:off + ns], eps=1e-9, ic_shift=1.2)
        assert sto > 0 and st[s] > 0
        # subdomain COCG counts at eps 1e-9 move with the summation order of
        # the products: a CPU COCG with permuted dot-product order gives
        # 157-200 iterations where the oracle gives 164-201 on these
        # subdomains (up to 9 of 179), so the bar is 6 %
        assert abs(int(st[s]) - sto) <= max(2, 0.06 * sto)
        from tests.helpers import dense_lower
        xd = np.linalg.solve(dense_lower(rp, cl, vl, ns), b[off:off + ns])
        assert np.linalg.norm(x[off:off + ns] - xd) / np.linalg.norm(xd) < 1e-7
        off += ns


def test_post_matches_oracle(fw, small_model):
    """fwh_post (HDDM_post_em_E for every element) against the oracle on the
    GPU's own solution; elements without a listed material give zeros."""
    say("[gpu] element post-processing")
    sol = fw.Solver(small_model)
    st = sol.solve(eps=1e-8, max_loop=5000, graph_iters=8)
    assert st["converged"] == 1
    x = sol.x()
    got = sol.post()
    ref = fo.post_model(small_model, x)
    for g, r in zip(got, ref):
        assert np.abs(g - r).max() <= 1e-12 * np.abs(r).max()
    nel = int(sol.info["n_elem"])
    nu = np.full(nel, 7.957747e+05)
    nu[::3] = 0.0
    got2 = sol.post(nu)
    for g2, g in zip(got2, got):
        assert np.all(g2[::3] == 0)
        keep = np.arange(nel) % 3 != 0
        assert np.array_equal(g2[keep], g[keep])
    sol.close()
