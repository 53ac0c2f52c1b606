"""Delta-gap source (volt.dat with dg > 0,
lib/fem/make_elem_em.c:320-396).

The load is be_i = sigma*omega * sum_j gs_ij * (i dV_j) with
gs_ij = (grad lambda_j, N_i).  Pinned here without retyping the oracle's
formula:
  (i) the edge functions are N_k = ls_k w_ab (Whitney's lambda_a grad lambda_b
      - lambda_b grad lambda_a, edge k from a to b), so
      gs_ij = V ls_i (grad lambda_b - grad lambda_a) . grad lambda_j / 4
      (N_i is affine: its integral is V N_i(centroid), lambda = 1/4 there);
 (ii) Whitney's exact sequence grad lambda_j = sum_k inc_kj w_k gives
      gs = SS diag(1/ls) inc with the oracle's own mass matrix SS (taken
      from the imaginary part of the element matrix, i*sigma*SS);
(iii) sum_j gs_ij = 0 (the lambdas sum to 1): a uniform gap voltage loads
      nothing.
"""
import numpy as np
import pytest

from oracle import fw_oracle as fo

EDGE = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]


def random_tet(rng):
    while True:
        v = rng.normal(size=(4, 3))
        D = np.linalg.det(np.stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]]))
        if D > 0.05:
            return v


def signed_lengths(v):
    ls = np.zeros(6)
    for k, (a, b) in enumerate(EDGE):
        d = v[b] - v[a]
        nz = d[np.nonzero(d)[0][0]]
        ls[k] = np.linalg.norm(d) * np.sign(nz)
    return ls


def grads(v):
    """grad lambda_j (4 x 3) from the barycentric map"""
    T = np.vstack([np.ones(4), v.T])          # lambda = T^-1 [1, x, y, z]
    return np.linalg.inv(T)[:, 1:]


@pytest.mark.parametrize("seed", range(6))
def test_dg_load_closed_form(seed):
    rng = np.random.default_rng(1000 + seed)
    v = random_tet(rng)
    sigma, omega = rng.uniform(0.1, 3.0), rng.uniform(1e8, 1e10)
    dv = rng.normal(size=4)
    be = fo.dg_load(v, sigma, omega, dv)
    ls = signed_lengths(v)
    g = grads(v)
    V = np.linalg.det(np.stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]])) / 6.0
    gs = np.zeros((6, 4))
    for i, (a, b) in enumerate(EDGE):
        gs[i] = V * ls[i] * (g[b] - g[a]) @ g.T / 4.0
    ref = 1j * sigma * omega * gs @ dv
    assert np.abs(be - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("seed", range(6))
def test_dg_load_exact_sequence(seed):
    rng = np.random.default_rng(2000 + seed)
    v = random_tet(rng)
    sigma, omega = rng.uniform(0.1, 3.0), rng.uniform(1e8, 1e10)
    ae, _ = fo.element(v, 0.0, sigma, omega)
    SS = ae.imag / sigma
    inc = np.zeros((6, 4))
    for k, (a, b) in enumerate(EDGE):
        inc[k, b], inc[k, a] = 1.0, -1.0
    gs = SS @ np.diag(1.0 / signed_lengths(v)) @ inc
    for j in range(4):
        dv = np.zeros(4)
        dv[j] = 1.0
        be = fo.dg_load(v, sigma, omega, dv)
        ref = 1j * sigma * omega * gs[:, j]
        assert np.abs(be - ref).max() <= 1e-12 * np.abs(gs).max() * sigma * omega


def test_dg_uniform_voltage_loads_nothing():
    rng = np.random.default_rng(7)
    v = random_tet(rng)
    be = fo.dg_load(v, 1.0, 1e9, np.full(4, 2.5))
    scale = np.abs(fo.dg_load(v, 1.0, 1e9, np.array([2.5, 0, 0, 0]))).max()
    assert np.abs(be).max() <= 1e-13 * scale


def test_dg_zero_conductivity_loads_nothing():
    rng = np.random.default_rng(8)
    v = random_tet(rng)
    assert np.all(fo.dg_load(v, 0.0, 1e9, rng.normal(size=4)) == 0)


def test_model_assembly_uses_gap_load(small_sigma_model):
    """With dgv set, every subdomain's assembled load is the sum of the gap
    loads of its elements (the current source is not used), and the gap
    vertices hold the voltages of the last box containing them."""
    from tests.helpers import with_delta_gap
    m = with_delta_gap(small_sigma_model)
    dgv = np.asarray(m.dgv)
    assert sorted(np.unique(dgv)) == [0.0, 1.0, 2.0]
    crd = np.asarray(m.crd).reshape(-1, 3)
    assert np.all(crd[dgv == 2.0][:, 2] >= 0.5 - 1e-12)
    assert np.all(crd[dgv == 1.0][:, 2] < 0.5)
    om = fo.OracleModel(m)
    nop = np.asarray(m.nop).reshape(-1, 10)
    total = 0.0
    for s in range(m.nsub):
        f = np.zeros(int(m.nmp[s]), complex)
        for el in range(m.el_off[s], m.el_off[s + 1]):
            vi = m.ap_off[s] + nop[el, :4]
            be = fo.dg_load(crd[vi], m.sigma[el], m.omega, dgv[vi])
            np.add.at(f, nop[el, 4:], be)
        fo_s = om.get(s)["f"]
        assert np.abs(fo_s - f).max() <= 1e-13 * max(np.abs(f).max(), 1e-300)
        total += np.abs(f).sum()
    assert total > 0


def test_read_volt(tmp_path):
    """volt.dat: dg, then x/y/z ranges and dV per box; the reference's
    sample file (dg = 0) selects the current source."""
    from adventure_fullwave_amd import advio
    p = tmp_path / "volt.dat"
    p.write_text("2\n-0.1 0.1\n-0.2 0.2\n0.0 1.0\n5.0\n0 0\n0 0\n0.5 0.5\n-3.0\n")
    boxes = advio.read_volt(str(p))
    assert len(boxes) == 2
    assert np.array_equal(boxes[0]["lo"], [-0.1, -0.2, 0.0]) and np.array_equal(boxes[0]["hi"], [0.1, 0.2, 1.0])
    assert boxes[0]["dV"] == 5.0 and boxes[1]["dV"] == -3.0
    x = np.array([[0, 0, 0.5], [0, 0, 0.0], [0.1, 0.2, 1.0], [0.11, 0, 0.5], [0, 0, 0.25]])
    # (0,0,0.5) lies in both boxes: the last one wins; bounds are inclusive
    assert np.array_equal(advio.gap_voltage(x, boxes), [-3.0, 5.0, 5.0, 0.0, 5.0])
    p.write_text("0\n")
    assert advio.read_volt(str(p)) == []
