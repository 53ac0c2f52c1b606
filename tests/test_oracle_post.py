"""Oracle element post-processing (HDDM_post_em_E, main/fullwave/fullwave.c:640-763)
pinned by closed forms.

The lowest-order edge space holds every field E(r) = alpha + beta x r exactly,
and its edge DOF is the tangential component along the edge oriented by
coordinate order (tetNedelec_Direction, lib/fem/tetrahedral.c:225-263), which
is constant along the edge for such a field.  With s = nu/omega the reference's
outputs are then
    rotE = i s curl E = 2 i s beta,
    E    = 2 s E(centroid)        (the reference scales by 2/D, not 1/D),
    Je   = sigma s E(centroid).
"""
import numpy as np
import pytest

from oracle import fw_oracle as fo
from tests.test_oracle_element import random_tet

NU = 7.957747e+05
EDGES = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]


def tangent_dofs(v, field, edges=EDGES):
    out = np.zeros(6, np.complex128)
    for k, (a, b) in enumerate(edges):
        d = v[b] - v[a]
        si = next((np.sign(c) for c in d if c != 0), 0)
        t = si * d / np.linalg.norm(d)
        out[k] = field(v[a]) @ t
    return out


def linear_field(rng):
    al = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    be = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    return al, be, (lambda r: al + np.cross(be, r))


@pytest.mark.parametrize("seed", range(8))
def test_post_linear_field_closed_form(seed):
    rng = np.random.default_rng(100 + seed)
    v = random_tet(rng, scale=1 + seed).reshape(4, 3)
    al, be, f = linear_field(rng)
    xe = tangent_dofs(v, f)
    omega, sigma = 2.3e9, 0.37
    rot, E, J = fo.post_element(v, xe, NU, sigma, omega)
    s = NU / omega
    g = v.mean(0)
    Eg = f(g)
    scale = np.abs(s * Eg).max() + np.abs(s * be).max()
    assert np.abs(rot - 2j * s * be).max() <= 1e-12 * scale
    assert np.abs(E - 2 * s * Eg).max() <= 1e-12 * scale
    assert np.abs(J - sigma * s * Eg).max() <= 1e-12 * scale


def test_post_without_material_is_zero():
    """nu = 0 for an element whose flag is not a listed material (:670-675)."""
    rng = np.random.default_rng(7)
    v = random_tet(rng).reshape(4, 3)
    xe = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    for o in fo.post_element(v, xe, 0.0, 0.5, 1e9):
        assert np.all(o == 0)


def test_post_detects_edge_order_mistake():
    """The closed form is sensitive to the mid-node order: DOFs in the file
    order before the 8<->9 swap (edges (1,3),(2,3) exchanged) fail it."""
    rng = np.random.default_rng(11)
    v = random_tet(rng).reshape(4, 3)
    _, be, f = linear_field(rng)
    wrong = [(0, 1), (0, 2), (0, 3), (1, 2), (2, 3), (1, 3)]
    xe = tangent_dofs(v, f, wrong)
    rot, E, _ = fo.post_element(v, xe, NU, 0.0, 1e9)
    s = NU / 1e9
    assert np.abs(E - 2 * s * f(v.mean(0))).max() > 1e-3 * np.abs(s * f(v.mean(0))).max()
