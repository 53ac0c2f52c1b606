"""Invariants of the synthetic decomposed models (the data the reference
derives in lib/hddm/make_PartMesh.c): orientation, ownership, part-pair
lists, Dirichlet lists and the configuration sizes quoted for the workloads."""
import numpy as np
import pytest

from adventure_fullwave_amd import mesh


def test_kuhn_box_counts():
    for n in (1, 3, 12):
        crd, tets, cell, dims = mesh.kuhn_box(*(np.linspace(0, 1, n + 1),) * 3)
        assert len(tets) == 6 * n ** 3
        _, ev = mesh.edges_of(tets, len(crd))
        nx = ny = nz = n
        axis = nx * (ny + 1) * (nz + 1) * 3
        face = nx * ny * (nz + 1) * 3
        assert len(ev) == axis + face + nx * ny * nz
        v = crd[tets]
        d = np.einsum("ij,ij->i", v[:, 1] - v[:, 0], np.cross(v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]))
        assert (d > 0).all()


def test_cfg1_sizes(cfg1_model):
    m = cfg1_model
    assert m.n_edges == 13428 and m.n_tets == 10368 and m.nsub == 64 and m.npart == 4


def test_ownership_and_lists(small_model):
    m = small_model
    np_ = m.npart
    tf = m.infree + m.outfree
    for p in range(np_):
        for q in range(np_):
            a = m.op_ibid[m.op_off[p * np_ + q]:m.op_off[p * np_ + q + 1]]
            b = m.op_ibid[m.op_off[q * np_ + p]:m.op_off[q * np_ + p + 1]]
            assert len(a) == len(b)
            if p < q:
                # the lower part owns every DOF it shares with a higher part
                # unless a still lower part owns it
                assert (a < tf[p]).all() and (b < m.outfree[q]).all()
    # every part-level interface index is referenced by some subdomain
    for s in range(m.nsub):
        ln = m.ib_ln[m.ib_off[s]:m.ib_off[s + 1]]
        nf = m.ib_nfid[m.ib_off[s]:m.ib_off[s + 1]]
        assert (np.diff(nf) > 0).all()          # local-id order (sort_inbc)
        assert (ln < tf[m.sub_part[s]]).all()
        ess = m.ibc[m.bc_off[s]:m.bc_off[s + 1]]
        assert len(np.intersect1d(ess, nf)) == 0


@pytest.mark.parametrize("order", ["rcm", "none"])
def test_orders_give_same_sets(order):
    m = mesh.make_config(1, small=(4, (2, 1, 1), (1, 1, 1)), order=order)
    assert m.nsub == 2 and m.n_interface > 0
