"""Partition files and the reference's interface reclassification on the
CPU (no GPU).

* Round trip: a synthetic model with PEC (essential) edges, written as
  reference partition files (advio.write_partition: node-level interface of
  vertices and mid-nodes, DirichletBCs_Axn0 with the essential mid-nodes and
  their end vertices) and read back through the reference's rules
  (advio.reclassify_interface, make_PartMesh.c:158-858, and the
  make_PartMesh_FW_TetQ numbering): the same essential edges, the same
  interface edges and count, the same interface numbering as the synthetic
  generator's own, an equivalent system (oracle solutions agree to the stop
  error, essential edges exactly zero), sorted Dirichlet and interface
  lists, consistent part-pair lists.  A loader that ignores the file's
  Dirichlet list fails.
* On the reference's own sample partition (when the reference tree is
  present): the rules leave exactly the mid-node interface (20,264 unknowns,
  the count the reference logged), renumber it copies-first in file order,
  and pair the part lists node for node."""
import os

import numpy as np
import pytest

from adventure_fullwave_amd import advio
from oracle import fw_oracle as fo
from tests.helpers import to_global

REF_MODEL = "/root/reference/main/fullwave/MODEL"


def materials(model):
    key = np.stack([model.eps, model.sigma, model.src], 1)
    uq, inv = np.unique(key, axis=0, return_inverse=True)
    mt = {"eps": {i + 1: uq[i, 0] for i in range(len(uq))}, "sigma": {i + 1: uq[i, 1] for i in range(len(uq))},
          "antenna": {}, "omega": model.omega}
    # the source: a Parallelpiped around the source edge's two vertices
    joe = np.asarray(model.joe).reshape(-1, 6)
    crd = np.asarray(model.crd).reshape(-1, 3)
    on = np.abs(joe).sum(1) > 0
    sv = np.unique(np.round(crd[on], 12), axis=0)
    lo, hi = sv.min(0) - 1e-3, sv.max(0) + 1e-3
    ap = np.array([lo, [hi[0], lo[1], lo[2]], [lo[0], hi[1], lo[2]], [lo[0], lo[1], hi[2]]])
    j = joe[on][0]
    pp = {"ap": ap, "Jor": -j[1::2].copy(), "Joi": -j[0::2].copy()}
    ants = {i + 1: [pp] for i in range(len(uq)) if uq[i, 2]}
    return inv.ravel() + 1, mt, ants


def global_sets(model, node_of_edge=None):
    offs = np.concatenate([[0], np.cumsum(model.nmp)])
    g = np.asarray(model.ledge_global) if node_of_edge is None else node_of_edge(np.asarray(model.ledge_global))
    ess, ifc = set(), set()
    for s in range(model.nsub):
        ess.update(g[offs[s] + model.ibc[model.bc_off[s]:model.bc_off[s + 1]]].tolist())
        ifc.update(g[offs[s] + model.ib_nfid[model.ib_off[s]:model.ib_off[s + 1]]].tolist())
    return ess, ifc


def pair_lists_consistent(model):
    offs = np.concatenate([[0], np.cumsum(model.nmp)])
    gl = [dict() for _ in range(model.npart)]
    for s in range(model.nsub):
        p = model.sub_part[s]
        for nf, ln in zip(model.ib_nfid[model.ib_off[s]:model.ib_off[s + 1]],
                          model.ib_ln[model.ib_off[s]:model.ib_off[s + 1]]):
            g = int(model.ledge_global[offs[s] + nf])
            if gl[p].setdefault(int(ln), g) != g:
                return False
    n = model.npart
    for p in range(n):
        for q in range(n):
            if p == q:
                continue
            a = model.op_ibid[model.op_off[p * n + q]:model.op_off[p * n + q + 1]]
            b = model.op_ibid[model.op_off[q * n + p]:model.op_off[q * n + p + 1]]
            if len(a) != len(b) or any(gl[p][int(x)] != gl[q][int(y)] for x, y in zip(a, b)):
                return False
    return True


@pytest.fixture(scope="module")
def round_trip(small_sigma_model, tmp_path_factory):
    m = small_sigma_model
    flags, mt, ants = materials(m)
    d = str(tmp_path_factory.mktemp("partition"))
    advio.write_partition(m, d, flags)
    parts = [advio.read_partition(os.path.join(d, f"advhddm_in_{p}.adv")) for p in range(m.npart)]
    return m, parts, mt, ants, advio.build_full_wave_model(parts, mt, ants)


def test_round_trip_structure(round_trip):
    m, parts, mt, ants, m2 = round_trip
    assert all(P["dbc"] is not None for P in parts)
    nv = int(np.unique(np.round(np.asarray(m.crd).reshape(-1, 3), 12), axis=0).shape[0])
    ess1, if1 = global_sets(m)
    ess2, if2 = global_sets(m2, lambda g: g - nv)   # mid-node id = vertices + edge id
    assert len(ess1) > 0 and ess1 == ess2
    assert if1 == if2 and not (ess2 & if2)
    assert m2.n_interface == m.n_interface and m2.nsub == m.nsub
    for s in range(m2.nsub):
        ibc = m2.ibc[m2.bc_off[s]:m2.bc_off[s + 1]]
        nf = m2.ib_nfid[m2.ib_off[s]:m2.ib_off[s + 1]]
        assert np.all(np.diff(ibc) > 0) and np.all(np.diff(nf) > 0)   # sort_i / sort_inbc
    assert pair_lists_consistent(m2)
    # the synthetic generator numbers its interface as the reference's rules
    # number the partition files written from it: the same ln_part of every
    # interface edge in every part, the same part-pair lists
    assert ln_maps(m) == ln_maps(m2, nv)
    assert np.array_equal(m.op_off, m2.op_off) and np.array_equal(m.op_ibid, m2.op_ibid)


def ln_maps(model, shift=0):
    offs = np.concatenate([[0], np.cumsum(model.nmp)])
    out = [dict() for _ in range(model.npart)]
    for s in range(model.nsub):
        p = model.sub_part[s]
        for nf, ln in zip(model.ib_nfid[model.ib_off[s]:model.ib_off[s + 1]],
                          model.ib_ln[model.ib_off[s]:model.ib_off[s + 1]]):
            out[p][int(model.ledge_global[offs[s] + nf]) - shift] = int(ln)
    return out


def test_round_trip_same_system(round_trip):
    m, parts, mt, ants, m2 = round_trip
    r1 = fo.hddm_solve(m, eps=1e-12, max_loop=5000)
    r2 = fo.hddm_solve(m2, eps=1e-12, max_loop=5000)
    offs1 = np.concatenate([[0], np.cumsum(m.nmp)])
    x1 = to_global(m, [r1["x_flat"][offs1[s]:offs1[s + 1]] for s in range(m.nsub)])
    x2 = np.zeros(int(m2.n_edges), complex)
    x2[m2.ledge_global] = r2["x_flat"]
    nv = int(m2.n_edges) - m.n_edges
    assert np.linalg.norm(x2[nv:] - x1) <= 1e-10 * np.linalg.norm(x1)
    ess2, _ = global_sets(m2)
    assert np.all(x2[list(ess2)] == 0)
    # a loader that ignored the file's Dirichlet list solves another system
    bad = advio.build_full_wave_model([dict(P, dbc=None) for P in parts], mt, ants)
    assert bad.n_interface > m2.n_interface and len(bad.ibc) == 0
    rb = fo.hddm_solve(bad, eps=1e-12, max_loop=5000)
    xb = np.zeros(int(bad.n_edges), complex)
    xb[bad.ledge_global] = rb["x_flat"]
    assert np.linalg.norm(xb[nv:] - x1) > 1e-3 * np.linalg.norm(x1)


@pytest.mark.skipif(not os.path.isdir(REF_MODEL), reason="reference tree not present")
def test_reclassification_rules_on_sample():
    files = sorted(os.listdir(os.path.join(REF_MODEL, "model")))
    parts = [advio.read_partition(os.path.join(REF_MODEL, "model", f"advhddm_in_{p}.adv"))
             for p in range(len(files))]
    out = advio.reclassify_interface(parts)
    assert sum(o["infree"] for o in out) == 20264
    for p, (P, o) in enumerate(zip(parts, out)):
        # every surviving entry is a mid-node interface node of its subdomain
        ln_old, ln_new, charge = {}, {}, {}
        for d in range(P["nsub"]):
            mids = set(P["nop"][d][:, 4:].ravel().tolist())
            kept = {int(a) for a in o["inbc"][d][:, 0]}
            assert kept == {int(r[0]) for r in P["inbc"][d] if int(r[0]) in mids}
            for r, new in zip([r for r in P["inbc"][d] if int(r[0]) in mids], o["inbc"][d][:, 1]):
                ln_old[int(r[4])] = int(r[4])
                ln_new[int(r[4])] = int(new)
                charge[int(r[4])] = int(r[2])
        old = sorted(ln_old)
        other = [k for k in old if charge[k] != p]
        own = [k for k in old if charge[k] == p]
        # copies charged elsewhere first, then the owned ones, both in file order
        assert [ln_new[k] for k in other + own] == list(range(len(old)))
        assert o["outfree"] == len(other) and o["infree"] == len(own)
    m = advio.load_reference_model(REF_MODEL)
    assert pair_lists_consistent(m)
