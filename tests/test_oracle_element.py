"""Pins the oracle's element kernel (tet_MakeElement_EM_E, lib/fem/
make_elem_em.c:14-400) against independent mathematics:

* Whitney edge functions built from numpy barycentric gradients
  (N_k = l_k si_k (lam_a grad lam_b - lam_b grad lam_a)), integrated with an
  exact degree-2 quadrature rule (mass) and exact constant curls (curl-curl);
* the discrete exact-sequence identity: the curl-curl matrix annihilates
  every discrete gradient;
* the source term against quadrature of N_i . J with J linear;
* the reference's 24-term six-times-volume formula (tetrahedral.c:44-61).
"""
import numpy as np
import pytest

from oracle import fw_oracle as fo

EDGE = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
NU = 7.957747e5


def random_tet(rng, scale=1.0, shift=0.0):
    while True:
        v = rng.standard_normal((4, 3)) * scale + shift
        d = np.dot(v[1] - v[0], np.cross(v[2] - v[0], v[3] - v[0]))
        if abs(d) > 0.05 * scale ** 3:
            if d < 0:
                v[[1, 2]] = v[[2, 1]]
            return v


def whitney(v):
    """grad lambda (4x3), lengths*signs (6), function N(r) -> (6,3)."""
    M = np.hstack([np.ones((4, 1)), v])           # lambda = M^-T [1, r]
    C = np.linalg.inv(M)                          # columns: coefficients of lambda_i
    grads = C[1:, :].T                            # (4,3)
    ls = []
    for a, b in EDGE:
        d = v[b] - v[a]
        si = next((np.sign(x) for x in d if x != 0), 0.0)
        ls.append(np.linalg.norm(d) * si)
    ls = np.array(ls)

    def lam(r):
        return C.T @ np.concatenate([[1.0], r])

    def N(r):
        L = lam(r)
        return np.array([ls[k] * (L[a] * grads[b] - L[b] * grads[a]) for k, (a, b) in enumerate(EDGE)])
    return grads, ls, lam, N


# exact for polynomials of degree 2 on a tetrahedron
QA, QB = 0.5854101966249685, 0.1381966011250105
QBARY = np.array([[QA, QB, QB, QB], [QB, QA, QB, QB], [QB, QB, QA, QB], [QB, QB, QB, QA]])


def quad(v, fn):
    vol = abs(np.dot(v[1] - v[0], np.cross(v[2] - v[0], v[3] - v[0]))) / 6.0
    return sum(0.25 * vol * fn(QBARY[q] @ v) for q in range(4))


def reference_matrices(v, eps, sigma, omega):
    grads, ls, lam, N = whitney(v)
    vol = np.dot(v[1] - v[0], np.cross(v[2] - v[0], v[3] - v[0])) / 6.0
    curls = np.array([2 * ls[k] * np.cross(grads[a], grads[b]) for k, (a, b) in enumerate(EDGE)])
    RR = vol * curls @ curls.T
    SS = quad(v, lambda r: N(r) @ N(r).T)
    return RR, SS, (NU / omega) * RR - eps * omega * SS + 1j * sigma * SS


@pytest.mark.parametrize("seed", range(6))
def test_element_matrix_vs_quadrature(seed):
    rng = np.random.default_rng(seed)
    v = random_tet(rng, scale=0.1 + seed, shift=seed * 0.3)
    eps, sigma, omega = 8.854e-12 * (1 + seed), 0.01 * seed, 2.2e9
    ae, be = fo.element(v, eps, sigma, omega)
    RR, SS, A = reference_matrices(v, eps, sigma, omega)
    assert np.allclose(ae, ae.T, rtol=1e-13, atol=0)
    scale = np.abs(A).max()
    assert np.abs(ae - A).max() < 1e-11 * scale
    assert np.abs(be).max() == 0.0


def test_curlcurl_kills_gradients():
    rng = np.random.default_rng(7)
    v = random_tet(rng)
    # only the curl-curl part: eps = sigma = 0, omega = NU so ae = RR
    ae, _ = fo.element(v, 0.0, 0.0, NU)
    _, ls, _, _ = whitney(v)
    for phi in np.eye(4):
        c = np.array([(phi[b] - phi[a]) / ls[k] for k, (a, b) in enumerate(EDGE)])
        assert np.abs(ae.real @ c).max() < 1e-12 * np.abs(ae).max()


def test_mass_positive_definite():
    rng = np.random.default_rng(8)
    v = random_tet(rng)
    ae, _ = fo.element(v, -1.0, 0.0, 1.0e30)   # -(eps w) SS with eps=-1, w huge -> ~w SS
    w = np.linalg.eigvalsh(ae.real)
    assert w.min() > 0


def test_source_vs_quadrature():
    rng = np.random.default_rng(9)
    v = random_tet(rng)
    J = rng.standard_normal((4, 3)) + 1j * rng.standard_normal((4, 3))
    joe = np.stack([J.real, J.imag], axis=2).reshape(24)
    _, be = fo.element(v, 1e-11, 0.0, 2e9, src=1, joe=joe)
    _, _, lam, N = whitney(v)
    ref = quad(v, lambda r: N(r) @ (lam(r) @ J))
    assert np.abs(be - ref).max() < 1e-12 * np.abs(ref).max()


def test_stretch_one_is_reference():
    rng = np.random.default_rng(10)
    v = random_tet(rng)
    a0, _ = fo.element(v, 3e-11, 0.05, 1.3e9)
    a1, _ = fo.element(v, 3e-11, 0.05, 1.3e9, pml=np.array([1.0, 0, 1.0, 0, 1.0, 0]))
    assert np.abs(a0 - a1).max() < 1e-13 * np.abs(a0).max()


def test_stretch_tensor_vs_quadrature():
    """Uniaxial stretch: mu^-1 -> diag(1/Lx,1/Ly,1/Lz), eps -> diag(Lx,Ly,Lz)."""
    rng = np.random.default_rng(11)
    v = random_tet(rng)
    s = np.array([1 - 0.7j, 1.0 + 0j, 1 - 0.2j])
    lam_ = np.array([s[1] * s[2] / s[0], s[0] * s[2] / s[1], s[0] * s[1] / s[2]])
    eps, sigma, omega = 2e-11, 0.03, 1.7e9
    pml = np.stack([s.real, s.imag], axis=1).reshape(6)
    ae, _ = fo.element(v, eps, sigma, omega, pml=pml)
    grads, ls, lam, N = whitney(v)
    vol = np.dot(v[1] - v[0], np.cross(v[2] - v[0], v[3] - v[0])) / 6.0
    curls = np.array([2 * ls[k] * np.cross(grads[a], grads[b]) for k, (a, b) in enumerate(EDGE)])
    RR = vol * (curls / lam_) @ curls.T
    SS = quad(v, lambda r: (N(r) * lam_) @ N(r).T)
    A = (NU / omega) * RR + (-eps * omega + 1j * sigma) * SS
    assert np.abs(ae - A).max() < 1e-11 * np.abs(A).max()


def test_volume_formula_matches_reference_expression():
    """D of the oracle equals the reference's 24-term expression."""
    rng = np.random.default_rng(12)
    v = random_tet(rng)
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    D24 = (-x[0]*y[1]*z[2] + x[0]*y[1]*z[3] + x[0]*y[2]*z[1] - x[0]*y[2]*z[3] - x[0]*y[3]*z[1]
           + x[0]*y[3]*z[2] + x[1]*y[0]*z[2] - x[1]*y[0]*z[3] - x[1]*y[2]*z[0] + x[1]*y[2]*z[3]
           + x[1]*y[3]*z[0] - x[1]*y[3]*z[2] - x[2]*y[0]*z[1] + x[2]*y[0]*z[3] + x[2]*y[1]*z[0]
           - x[2]*y[1]*z[3] - x[2]*y[3]*z[0] + x[2]*y[3]*z[1] + x[3]*y[0]*z[1] - x[3]*y[0]*z[2]
           - x[3]*y[1]*z[0] + x[3]*y[1]*z[2] + x[3]*y[2]*z[0] - x[3]*y[2]*z[1])
    D = np.dot(v[1] - v[0], np.cross(v[2] - v[0], v[3] - v[0]))
    assert abs(D24 - D) < 1e-12 * abs(D)
    # the element kernel accepts exactly the positively oriented tets
    fo.element(v, 1e-11, 0, 1e9)
    with pytest.raises(ValueError):
        fo.element(v[[0, 2, 1, 3]], 1e-11, 0, 1e9)


def test_mutation_dropped_mass_term_is_caught():
    """A plausible mistake (mass matrix without the centroid term) fails."""
    rng = np.random.default_rng(13)
    v = random_tet(rng)
    _, SS, _ = reference_matrices(v, 0, 0, 1.0)
    grads, ls, lam, N = whitney(v)
    g = v.mean(axis=0)
    bad = quad(v, lambda r: (N(r) - N(g)) @ (N(r) - N(g)).T)   # drops the F F^T part
    assert np.abs(bad - SS).max() > 1e-3 * np.abs(SS).max()
