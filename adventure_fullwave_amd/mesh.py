"""Synthetic decomposed models for the HDDM full-wave path.

The reference reads an ADVENTURE partition (model/advhddm_in_P.adv, written by
ADVENTURE_Metis) and derives, per part and subdomain, the data that its solver
loop consumes (lib/hddm/make_PartMesh.c:58-1707): element connectivity with
local vertex and edge ids, Dirichlet lists, the interface map
(Inbc.nfid -> ln_part) and the part-pair sharing lists (OPinfo.ibid).  This
module builds exactly that data model for box meshes (6-tetrahedron Kuhn
split of a hex grid), so that the oracle and the HIP library see identical
inputs.  Conventions kept from the reference:

* edge k of a tetrahedron joins vertices (0,1),(0,2),(0,3),(1,2),(1,3),(2,3)
  (lib/fem/tetrahedral.h:19-24 after the mid-node swap of
  lib/fem/make_MESH.c:573-584), six-times volume positive;
* an interface edge is owned by the lowest-index part that holds it; in each
  part the non-owned copies come first (ln_part < outfree) and the owned ones
  after (lib/hddm/make_PartMesh.c:752-858);
* each subdomain's Dirichlet list is [essential edges | interface edges]
  with the interface edges in local-id order (make_PartMesh.c:1082-1194).

Edges that carry an essential (PEC) condition never become interface DOFs.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
from typing import Dict, List, Optional, Sequence

import numpy as np

EDGE = np.array([[0, 1], [0, 2], [0, 3], [1, 2], [1, 3], [2, 3]], dtype=np.int64)
EPS0 = 8.854e-12          # vacuum permittivity as written in the reference's mtrl.dat
C0 = 299792458.0
ETA0 = 376.730313668
# cfg 3: radius of the sphere scatterer (box of unit side), and the share of
# a subdomain's edge DOFs its cells' own edges make up (the rest lie on the
# faces it shares; more of them in small subdomains) at the smallest and
# the largest target size, so that subdomain sizes follow their targets
CFG3_RADIUS = 0.3
CFG3_FILL = (0.9, 0.78)


# --------------------------------------------------------------------------
# geometry
# --------------------------------------------------------------------------
def kuhn_box(xs: np.ndarray, ys: np.ndarray, zs: np.ndarray, orient: bool = True):
    """Vertices and positively oriented tetrahedra of a (possibly graded)
    tensor grid, six tetrahedra per hex cell around the cell diagonal.
    orient=False keeps every tetrahedron in path order (cell corner 0, then
    one, two and three unit steps to the opposite corner), the order the
    bisection refinement needs.

    Returns (crd[nv,3], tets[ne,4], cell_of_tet[ne], (nx,ny,nz))."""
    nx, ny, nz = len(xs) - 1, len(ys) - 1, len(zs) - 1
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    # vertex id = i + (nx+1)*(j + (ny+1)*k)
    crd = np.stack([X.transpose(2, 1, 0).ravel(), Y.transpose(2, 1, 0).ravel(),
                    Z.transpose(2, 1, 0).ravel()], axis=1)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = (a.transpose(2, 1, 0).ravel() for a in (i, j, k))  # cell = i + nx*(j + ny*k)

    def vid(di, dj, dk):
        return (i + di) + (nx + 1) * ((j + dj) + (ny + 1) * (k + dk))

    corner = {b: vid(b & 1, (b >> 1) & 1, (b >> 2) & 1) for b in range(8)}
    tets = []
    for perm in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
        b1 = 1 << perm[0]
        b2 = b1 | (1 << perm[1])
        tets.append(np.stack([corner[0], corner[b1], corner[b2], corner[7]], axis=1))
    ncell = nx * ny * nz
    tets = np.stack(tets, axis=1).reshape(ncell * 6, 4).astype(np.int64)
    cell = np.repeat(np.arange(ncell), 6)
    if orient:
        tets = orient_tets(crd, tets)
    return crd, tets, cell, (nx, ny, nz)


def orient_tets(crd: np.ndarray, tets: np.ndarray) -> np.ndarray:
    """The tetrahedra with vertices 1 and 2 swapped where the six-times
    volume would be negative."""
    tets = tets.copy()
    v = crd[tets]
    d = np.einsum("ij,ij->i", v[:, 1] - v[:, 0], np.cross(v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]))
    neg = d < 0
    tets[neg, 1], tets[neg, 2] = tets[neg, 2].copy(), tets[neg, 1].copy()
    return tets


def refine_ball(crd: np.ndarray, tets: np.ndarray, cell: np.ndarray, center, radius: float, active_radius: float,
                generations: int = 3):
    """Conforming local refinement of a Kuhn mesh (tetrahedra in path order)
    by newest-vertex bisection: every tetrahedron whose centroid lies within
    `radius` of `center` is bisected until it is `generations` bisections
    below the base mesh (three halve the mesh size).

    Bisection follows Maubach's rule for tagged simplices (J. M. Maubach,
    SIAM J. Sci. Comput. 16 (1995) 210-227): a tetrahedron (x0, x1, x2, x3)
    with tag k is cut at the midpoint z of its refinement edge (x0, xk) into
    (x0, .., x_{k-1}, z, x_{k+1}, .., x3) and (x1, .., xk, z, x_{k+1}, .., x3),
    both with tag k - 1 (3 after 1); the base tetrahedra have tag 3 (their
    refinement edge is the cell diagonal).  Conformity: before the
    tetrahedra around an edge are cut, every one of them whose refinement
    edge differs is refined first (recursively), so all of them share the
    edge as refinement edge and are cut together at the same midpoint.  On
    Kuhn meshes this closure terminates.

    Only the tetrahedra with centroids within `active_radius` take part; a
    closure that would reach past them raises.  Returns (crd, tets in path
    order, base cell of every tetrahedron, bisections of every tetrahedron)."""
    center = np.asarray(center, dtype=float)
    cent = crd[tets].mean(axis=1)
    act = np.linalg.norm(cent - center, axis=1) < active_radius
    frozen = tets[~act]
    verts = [tuple(c) for c in crd]
    T: Dict[int, list] = {}
    E: Dict[tuple, set] = {}
    mids: Dict[tuple, int] = {}

    pairs = tuple((int(i), int(j)) for i, j in EDGE)

    def ekey(a, b):
        return (a, b) if a < b else (b, a)

    def add(t, rec):
        T[t] = rec
        x = rec[0]
        for i, j in pairs:
            a, b = x[i], x[j]
            key = (a, b) if a < b else (b, a)
            got = E.get(key)
            if got is None:
                E[key] = {t}
            else:
                got.add(t)

    def remove(t):
        rec = T.pop(t)
        x = rec[0]
        for i, j in pairs:
            a, b = x[i], x[j]
            E[(a, b) if a < b else (b, a)].discard(t)
        return rec

    nid = [0]
    for t, c in zip(np.nonzero(act)[0], cell[act]):
        add(nid[0], [tuple(int(v) for v in tets[t]), 3, 0, int(c)])
        nid[0] += 1

    def ref_edge(rec):
        x, k = rec[0], rec[1]
        return ekey(x[0], x[k])

    def bisect(t):
        x, k, g, c = remove(t)
        e = ekey(x[0], x[k])
        z = mids.get(e)
        if z is None:
            a, b = verts[e[0]], verts[e[1]]
            z = len(verts)
            verts.append(((a[0] + b[0]) * 0.5, (a[1] + b[1]) * 0.5, (a[2] + b[2]) * 0.5))
            mids[e] = z
        k2 = k - 1 if k > 1 else 3
        add(nid[0], [x[:k] + (z,) + x[k + 1:], k2, g + 1, c])
        add(nid[0] + 1, [x[1:k + 1] + (z,) + x[k + 1:], k2, g + 1, c])
        nid[0] += 2

    def refine(t):
        stack = [t]
        while stack:
            u = stack[-1]
            if u not in T:
                stack.pop()
                continue
            e = ref_edge(T[u])
            around = E[e]
            blocker = next((w for w in around if ref_edge(T[w]) != e), None)
            if blocker is not None:
                stack.append(blocker)
                continue
            for w in list(around):
                bisect(w)
            stack.pop()

    c2 = center * 1.0
    r2 = radius * radius
    changed = True
    while changed:
        changed = False
        for t in list(T):
            rec = T.get(t)
            if rec is None or rec[2] >= generations:
                continue
            x = rec[0]
            cx = sum(verts[v][0] for v in x) * 0.25 - c2[0]
            cy = sum(verts[v][1] for v in x) * 0.25 - c2[1]
            cz = sum(verts[v][2] for v in x) * 0.25 - c2[2]
            if cx * cx + cy * cy + cz * cz < r2:
                refine(t)
                changed = True
    # the frozen tetrahedra must not have an edge that was cut
    if mids:
        cut = np.array(sorted(mids), dtype=np.int64)
        key_cut = cut[:, 0] * len(verts) + cut[:, 1]
        a = frozen[:, EDGE[:, 0]]
        b = frozen[:, EDGE[:, 1]]
        key = np.minimum(a, b) * len(verts) + np.maximum(a, b)
        if np.isin(key, key_cut).any():
            raise ValueError("refinement closure reached past the active region")
    recs = list(T.values())
    new_tets = np.array([r[0] for r in recs], dtype=np.int64).reshape(-1, 4)
    out_tets = np.concatenate([frozen, new_tets])
    out_cell = np.concatenate([cell[~act], np.array([r[3] for r in recs], dtype=np.int64)])
    out_gen = np.concatenate([np.zeros(len(frozen), np.int64), np.array([r[2] for r in recs], dtype=np.int64)])
    return np.array(verts, dtype=float), out_tets, out_cell, out_gen


def edges_of(tets: np.ndarray, nv: int):
    """Global edge ids (sorted by (min vertex, max vertex)) of every tet edge."""
    a = tets[:, EDGE[:, 0]]
    b = tets[:, EDGE[:, 1]]
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    key = lo * nv + hi
    ukey, inv = np.unique(key.ravel(), return_inverse=True)
    return inv.reshape(-1, 6).astype(np.int64), np.stack([ukey // nv, ukey % nv], axis=1)


def split_sizes(n: int, pieces: int) -> np.ndarray:
    base = n // pieces
    sizes = np.full(pieces, base, dtype=np.int64)
    sizes[: n - base * pieces] += 1
    return sizes


def rcm_order(n: int, rows: np.ndarray, cols: np.ndarray) -> np.ndarray:
    """Reverse Cuthill-McKee order of a symmetric pattern (profile reducer)."""
    if n == 0:
        return np.zeros(0, dtype=np.int64)
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    m = coo_matrix((np.ones(len(rows)), (rows, cols)), shape=(n, n)).tocsr()
    m = m + m.T
    return np.asarray(reverse_cuthill_mckee(m.tocsr(), symmetric_mode=True), dtype=np.int64)


# --------------------------------------------------------------------------
# decomposed model
# --------------------------------------------------------------------------
@dataclasses.dataclass
class Model:
    """Flat arrays of the fwh_mesh_desc C struct (include/fwh_mesh.h)."""
    npart: int
    nsub: int
    omega: float
    sub_part: np.ndarray
    el_off: np.ndarray
    nop: np.ndarray
    eps: np.ndarray
    sigma: np.ndarray
    src: np.ndarray
    pml: Optional[np.ndarray]
    ap_off: np.ndarray
    crd: np.ndarray
    joe: np.ndarray
    nmp: np.ndarray
    bc_off: np.ndarray
    ibc: np.ndarray
    ib_off: np.ndarray
    ib_nfid: np.ndarray
    ib_ln: np.ndarray
    infree: np.ndarray
    outfree: np.ndarray
    op_off: np.ndarray
    op_ibid: np.ndarray
    dgv: Optional[np.ndarray] = None     # [nap] delta-gap voltage per local vertex, or None
    # bookkeeping (not part of the C struct)
    ledge_global: np.ndarray = None      # [sum nmp] global edge id of each local edge
    mp2sn: np.ndarray = None             # [sum nmp] subdomain node id of each mid-node
                                         # (dm->mpid2snid; reference models only)
    n_edges: int = 0
    n_tets: int = 0
    name: str = ""
    meta: dict = dataclasses.field(default_factory=dict)

    @property
    def n_interface(self) -> int:
        return int(self.infree.sum())

    def desc(self):
        """ctypes fwh_mesh_desc pointing into this model's arrays."""
        return make_desc(self)


class FwhMeshDesc(ctypes.Structure):
    _fields_ = [
        ("npart", ctypes.c_int), ("nsub", ctypes.c_int), ("omega", ctypes.c_double),
        ("sub_part", ctypes.c_void_p), ("el_off", ctypes.c_void_p), ("nop", ctypes.c_void_p),
        ("eps", ctypes.c_void_p), ("sigma", ctypes.c_void_p), ("src", ctypes.c_void_p),
        ("pml", ctypes.c_void_p), ("ap_off", ctypes.c_void_p), ("crd", ctypes.c_void_p),
        ("joe", ctypes.c_void_p), ("nmp", ctypes.c_void_p), ("bc_off", ctypes.c_void_p),
        ("ibc", ctypes.c_void_p), ("ib_off", ctypes.c_void_p), ("ib_nfid", ctypes.c_void_p),
        ("ib_ln", ctypes.c_void_p), ("infree", ctypes.c_void_p), ("outfree", ctypes.c_void_p),
        ("op_off", ctypes.c_void_p), ("op_ibid", ctypes.c_void_p), ("dgv", ctypes.c_void_p),
    ]


_INT_FIELDS = ("sub_part", "el_off", "nop", "src", "ap_off", "nmp", "bc_off", "ibc", "ib_off",
               "ib_nfid", "ib_ln", "infree", "outfree", "op_off", "op_ibid")
_DBL_FIELDS = ("eps", "sigma", "pml", "crd", "joe", "dgv")


def make_desc(m: Model) -> FwhMeshDesc:
    d = FwhMeshDesc()
    d.npart, d.nsub, d.omega = int(m.npart), int(m.nsub), float(m.omega)
    keep = []
    for f in _INT_FIELDS:
        a = getattr(m, f)
        a = np.ascontiguousarray(a, dtype=np.int32)
        setattr(m, f, a)
        keep.append(a)
        if a.size == 0:
            a = np.zeros(1, dtype=np.int32)
            keep.append(a)
        setattr(d, f, a.ctypes.data)
    for f in _DBL_FIELDS:
        a = getattr(m, f)
        if a is None:
            setattr(d, f, None)
            continue
        a = np.ascontiguousarray(a, dtype=np.float64)
        setattr(m, f, a)
        keep.append(a)
        setattr(d, f, a.ctypes.data)
    d._keep = keep
    return d


def decompose(crd, tets, cell, sub_of_cell, part_of_sub, eps_e, sigma_e, src_e, joe_v,
              pml_e=None, pec_edge_mask=None, omega=1.0, order="rcm", name="") -> Model:
    """Build the per-part / per-subdomain data model from a global mesh.

    sub_of_cell: subdomain id of every hex cell (ids must be part-major:
    part_of_sub non-decreasing).  joe_v: [nv,6] complex source per vertex."""
    nv = len(crd)
    ne = len(tets)
    tet_edges, edge_vtx = edges_of(tets, nv)
    nedge = len(edge_vtx)
    sub_of_tet = sub_of_cell[cell]
    nsub = int(part_of_sub.shape[0])
    npart = int(part_of_sub.max()) + 1
    assert np.all(np.diff(part_of_sub) >= 0), "subdomains must be part-major"

    pec = np.zeros(nedge, dtype=bool) if pec_edge_mask is None else pec_edge_mask.astype(bool)

    # which subdomains / parts hold each edge
    es = np.unique(sub_of_tet[:, None].repeat(6, 1).ravel() * nedge + tet_edges.ravel())
    es_sub, es_edge = es // nedge, es % nedge
    nsub_of_edge = np.bincount(es_edge, minlength=nedge)
    is_iface = (nsub_of_edge >= 2) & ~pec
    ep = np.unique(part_of_sub[es_sub] * nedge + es_edge)
    ep_part, ep_edge = ep // nedge, ep % nedge
    owner = np.full(nedge, np.iinfo(np.int64).max)
    np.minimum.at(owner, ep_edge, ep_part)

    # part-level interface numbering: non-owned first, then owned (global id order)
    infree = np.zeros(npart, dtype=np.int64)
    outfree = np.zeros(npart, dtype=np.int64)
    ln_maps: List[Dict[int, int]] = []
    part_iface: List[np.ndarray] = []
    for p in range(npart):
        e = ep_edge[(ep_part == p) & is_iface[ep_edge]]
        e.sort()
        nonown = e[owner[e] != p]
        own = e[owner[e] == p]
        outfree[p], infree[p] = len(nonown), len(own)
        order_p = np.concatenate([nonown, own])
        part_iface.append(order_p)
        ln_maps.append(order_p)

    # op lists: edges shared by parts p and q, ln ids in p, global-id order
    op_lists = []
    for p in range(npart):
        lnp = np.empty(nedge, dtype=np.int64)
        lnp[:] = -1
        lnp[ln_maps[p]] = np.arange(len(ln_maps[p]))
        row = []
        for q in range(npart):
            if q == p:
                row.append(np.zeros(0, dtype=np.int64))
                continue
            shared = np.intersect1d(ln_maps[p], ln_maps[q])
            row.append(lnp[shared])
        op_lists.append(row)
    op_off = [0]
    op_ibid = []
    for p in range(npart):
        for q in range(npart):
            op_ibid.append(op_lists[p][q])
            op_off.append(op_off[-1] + len(op_lists[p][q]))
    op_ibid = np.concatenate(op_ibid) if op_ibid else np.zeros(0, np.int64)

    # per subdomain local data
    order_tets = np.argsort(sub_of_tet, kind="stable")
    counts = np.bincount(sub_of_tet, minlength=nsub)
    el_off = np.concatenate([[0], np.cumsum(counts)])
    nop_all, ap_off, crd_all, joe_all, nmp = [], [0], [], [], []
    bc_off, ibc_all, ib_off, ib_nfid_all, ib_ln_all, ledge_all = [0], [], [0], [], [], []
    eps_all, sig_all, src_all, pml_all = [], [], [], []
    for s in range(nsub):
        p = int(part_of_sub[s])
        tl = order_tets[el_off[s]:el_off[s + 1]]
        tv = tets[tl]
        te = tet_edges[tl]
        verts, vinv = np.unique(tv.ravel(), return_inverse=True)
        edges, einv = np.unique(te.ravel(), return_inverse=True)
        einv = einv.reshape(-1, 6)
        n = len(edges)
        # local edge order: interior edges (profile-reduced), then interface
        # edges, then essential edges
        kind = np.where(pec[edges], 2, np.where(is_iface[edges], 1, 0))
        inter = np.nonzero(kind == 0)[0]
        if order == "rcm" and len(inter) > 1:
            # pattern among interior edges
            pos = np.full(n, -1)
            pos[inter] = np.arange(len(inter))
            a = np.repeat(einv, 6, axis=1).ravel()
            b = np.tile(einv, (1, 6)).ravel()
            keep = (pos[a] >= 0) & (pos[b] >= 0)
            perm = rcm_order(len(inter), pos[a[keep]], pos[b[keep]])
            inter = inter[perm]
        newold = np.concatenate([inter, np.nonzero(kind == 1)[0], np.nonzero(kind == 2)[0]])
        old2new = np.empty(n, dtype=np.int64)
        old2new[newold] = np.arange(n)
        edges_new = edges[newold]
        einv = old2new[einv]
        nop = np.concatenate([vinv.reshape(-1, 4), einv], axis=1)
        nop_all.append(nop.ravel())
        crd_all.append(crd[verts])
        joe_all.append(joe_v[verts])
        ap_off.append(ap_off[-1] + len(verts))
        nmp.append(n)
        ledge_all.append(edges_new)
        ess = np.nonzero(pec[edges_new])[0]
        ibc_all.append(ess)
        bc_off.append(bc_off[-1] + len(ess))
        ifc = np.nonzero(is_iface[edges_new])[0]          # local id order
        lnp = np.empty(nedge, dtype=np.int64)
        lnp[ln_maps[p]] = np.arange(len(ln_maps[p]))
        ib_nfid_all.append(ifc)
        ib_ln_all.append(lnp[edges_new[ifc]])
        ib_off.append(ib_off[-1] + len(ifc))
        eps_all.append(eps_e[tl])
        sig_all.append(sigma_e[tl])
        src_all.append(src_e[tl])
        if pml_e is not None:
            pml_all.append(pml_e[tl])

    cat = lambda xs, dt: (np.concatenate(xs).astype(dt) if len(xs) else np.zeros(0, dt))
    return Model(
        npart=npart, nsub=nsub, omega=float(omega),
        sub_part=part_of_sub.astype(np.int32), el_off=np.asarray(el_off, np.int32),
        nop=cat(nop_all, np.int32), eps=cat(eps_all, np.float64), sigma=cat(sig_all, np.float64),
        src=cat(src_all, np.int32), pml=(cat(pml_all, np.float64).reshape(-1) if pml_e is not None else None),
        ap_off=np.asarray(ap_off, np.int32), crd=cat(crd_all, np.float64).reshape(-1),
        joe=cat(joe_all, np.float64).reshape(-1), nmp=np.asarray(nmp, np.int32),
        bc_off=np.asarray(bc_off, np.int32), ibc=cat(ibc_all, np.int32),
        ib_off=np.asarray(ib_off, np.int32), ib_nfid=cat(ib_nfid_all, np.int32),
        ib_ln=cat(ib_ln_all, np.int32), infree=infree.astype(np.int32),
        outfree=outfree.astype(np.int32), op_off=np.asarray(op_off, np.int32),
        op_ibid=op_ibid.astype(np.int32), ledge_global=cat(ledge_all, np.int64),
        n_edges=nedge, n_tets=ne, name=name,
    )


# --------------------------------------------------------------------------
# the five workload configurations
# --------------------------------------------------------------------------
def _block_ids(n_cells: Sequence[int], pieces: Sequence[int]):
    """Per-axis block index of every cell index along each axis."""
    out = []
    for n, p in zip(n_cells, pieces):
        sz = split_sizes(n, p)
        out.append(np.repeat(np.arange(p), sz))
    return out


def block_partition(dims, sub_grid, part_grid):
    """Subdomain of every cell for a block decomposition; subdomain ids are
    part-major.  sub_grid: subdomain blocks per axis; part_grid: parts per
    axis (each part a block of subdomain blocks)."""
    nx, ny, nz = dims
    bx, by, bz = _block_ids((nx, ny, nz), sub_grid)
    px, py, pz = _block_ids(sub_grid, part_grid)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = (a.transpose(2, 1, 0).ravel() for a in (i, j, k))
    sbx, sby, sbz = bx[i], by[j], bz[k]
    part = px[sbx] + part_grid[0] * (py[sby] + part_grid[1] * pz[sbz])
    # rank subdomain blocks part-major, then block (z, y, x) order
    blk = sbx + sub_grid[0] * (sby + sub_grid[1] * sbz)
    nblk = sub_grid[0] * sub_grid[1] * sub_grid[2]
    bxs, bys, bzs = np.meshgrid(np.arange(sub_grid[0]), np.arange(sub_grid[1]), np.arange(sub_grid[2]),
                                indexing="ij")
    bxs, bys, bzs = (a.transpose(2, 1, 0).ravel() for a in (bxs, bys, bzs))
    bpart = px[bxs] + part_grid[0] * (py[bys] + part_grid[1] * pz[bzs])
    rank = np.lexsort((np.arange(nblk), bpart))
    new_id = np.empty(nblk, dtype=np.int64)
    new_id[rank] = np.arange(nblk)
    return new_id[blk], bpart[rank]


def morton_order(dims) -> np.ndarray:
    """The cells (id i + nx (j + ny k)) in Morton (Z-curve) order."""
    nx, ny, nz = dims
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = (a.transpose(2, 1, 0).ravel() for a in (i, j, k))

    def spread(v):
        v = v.astype(np.uint64)
        out = np.zeros_like(v)
        for b in range(11):
            out |= ((v >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b)
        return out
    code = spread(i) | (spread(j) << np.uint64(1)) | (spread(k) << np.uint64(2))
    return np.argsort(code, kind="stable")


def weighted_sfc_partition(dims, cell_weight: np.ndarray, target_sizes: np.ndarray, nparts: int):
    """Uneven subdomains of a locally refined grid: the Morton-ordered cells
    cut where the running weight (edge DOFs of the cell's tetrahedra) passes
    the running sum of the target sizes; parts group consecutive
    subdomains."""
    order = morton_order(dims)
    ncell = len(order)
    run = np.cumsum(cell_weight[order])
    bounds = np.cumsum(target_sizes)
    bounds = bounds[bounds < run[-1]]
    sub_seq = np.searchsorted(bounds, run - 0.5 * cell_weight[order], side="right")
    nsub = int(sub_seq.max()) + 1
    # drop empty ids (a cell heavier than a target)
    used, sub_seq = np.unique(sub_seq, return_inverse=True)
    nsub = len(used)
    sub_of_cell = np.empty(ncell, dtype=np.int64)
    sub_of_cell[order] = sub_seq
    part_of_sub = (np.arange(nsub) * nparts) // nsub
    return sub_of_cell, part_of_sub


def sfc_partition(dims, target_sizes_cells: np.ndarray, nparts: int):
    """Uneven subdomains: cut the Morton-ordered cell sequence into chunks of
    the given cell counts; parts group consecutive subdomains."""
    nx, ny, nz = dims
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = (a.transpose(2, 1, 0).ravel() for a in (i, j, k))

    def spread(v):
        v = v.astype(np.uint64)
        out = np.zeros_like(v)
        for b in range(11):
            out |= ((v >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b)
        return out
    code = spread(i) | (spread(j) << np.uint64(1)) | (spread(k) << np.uint64(2))
    order = np.argsort(code, kind="stable")
    ncell = nx * ny * nz
    bounds = np.minimum(np.cumsum(target_sizes_cells), ncell)
    bounds = bounds[bounds < ncell]
    bounds = np.concatenate([bounds, [ncell]])
    nsub = len(bounds)
    sub_seq = np.searchsorted(bounds, np.arange(ncell), side="right")
    sub_of_cell = np.empty(ncell, dtype=np.int64)
    sub_of_cell[order] = sub_seq
    part_of_sub = (np.arange(nsub) * nparts) // nsub
    return sub_of_cell, part_of_sub


def pml_profile(cent, lo, hi, d, sigma_max):
    """Quadratic conductivity profile per axis at element centroids."""
    depth = np.maximum(np.maximum(lo + d - cent, cent - (hi - d)), 0.0)
    return sigma_max * (np.minimum(depth, d) / d) ** 2, depth


def make_config(cfg: int, seed: Optional[int] = None, pml_mode: str = "stretch",
                scale: int = 1, order: str = "rcm", small: Optional[tuple] = None) -> Model:
    """Synthetic workload `cfg` (1..5).  Seeds default to 20261203 + cfg.

    pml_mode: "stretch" = complex coordinate stretching (new feature);
    "sigma" = the reference-expressible absorber (graded conductivity
    layers); "none" = no absorber.  small=(n, sub, part) shrinks cfg 1-2
    style boxes for tests."""
    if seed is None:
        seed = 20261203 + cfg
    rng = np.random.default_rng(seed)
    eps_r_elem = None
    pec_box = None
    gap_box = None
    if cfg in (1, 2, 5) or small is not None:
        if small is not None:
            n, sub, part = small
            npml = max(1, n // 4)
        elif cfg == 1:
            n, sub, part, npml = 12, (4, 4, 4), (2, 2, 1), 3
        else:
            n, sub, part, npml = 40, (11, 10, 10), (2, 2, 2), 3
        dims = (n, n, n)
        h = 1.0 / n
        xs = ys = zs = np.linspace(0.0, 1.0, n + 1)
        eps_r = 1.0 if (cfg == 1 or small is not None) else float(rng.uniform(1.0, 4.0))
        if cfg == 5:
            eps_r = 1.0
    elif cfg == 3:
        # 48^3 base grid; the tetrahedra inside a seeded sphere scatterer are
        # refined 2x (three bisections each, conforming closure around them)
        n = 48
        dims = (n, n, n)
        h = 1.0 / n
        xs = ys = zs = np.linspace(0.0, 1.0, n + 1)
        center = rng.uniform(0.4, 0.6, size=3)
        radius = CFG3_RADIUS
        npml = 3
        eps_r = 1.0
        sub = part = None
    elif cfg == 4:
        dims = (64, 64, 48)
        h = 1.0 / 64
        xs = np.linspace(0.0, 1.0, 65)
        ys = np.linspace(0.0, 1.0, 65)
        zs = np.linspace(0.0, 0.75, 49)
        npml = 4
        sub, part = (16, 16, 14), (2, 2, 2)
        eps_r = 1.0
    else:
        raise ValueError(cfg)

    if cfg == 3:
        crd, tets, cell, dims = kuhn_box(xs, ys, zs, orient=False)
        crd, tets, cell, _ = refine_ball(crd, tets, cell, center, radius, radius + 4.0 * h)
        tets = orient_tets(crd, tets)
    else:
        crd, tets, cell, dims = kuhn_box(xs, ys, zs)
    nv, ne = len(crd), len(tets)
    cent = crd[tets].mean(axis=1)
    lo = np.array([xs[0], ys[0], zs[0]])
    hi = np.array([xs[-1], ys[-1], zs[-1]])
    dpml = npml * h

    # materials
    eps_e = np.full(ne, eps_r * EPS0)
    sigma_e = np.zeros(ne)
    if cfg == 3:
        inside = np.linalg.norm(cent - center, axis=1) < radius
        eps_e[inside] = 6.0 * EPS0
        eps_r_max = 6.0
    elif cfg == 4:
        layer = np.minimum((cent[:, 2] / (0.75 / 8)).astype(int), 7)
        er = rng.uniform(1.0, 10.0, size=8)
        tand = rng.uniform(0.0, 0.02, size=8)
        eps_e = er[layer] * EPS0
        eps_r_max = float(er.max())
    elif cfg == 5:
        eps_e = rng.uniform(1.0, 4.0, size=ne) * EPS0
        eps_r_max = 4.0
    else:
        eps_r_max = eps_r
    # 10 cells per wavelength at the highest permittivity
    f = C0 / (10.0 * h * math.sqrt(eps_r_max))
    omega = 2.0 * math.pi * f
    if cfg == 3:
        sigma_e[inside] = 0.5 * omega * EPS0          # eps_r = 6 - 0.5j
    if cfg == 4:
        sigma_e = omega * eps_e * tand[layer]

    # absorber
    eta = ETA0 / math.sqrt(eps_r if cfg not in (3, 4, 5) else 1.0)
    sig_max = 3.0 * math.log(1.0e6) / (2.0 * eta * dpml)
    sprof, depth = pml_profile(cent, lo, hi, dpml, sig_max)
    pml_e = None
    if pml_mode == "stretch":
        epsb = eps_e[:, None]
        s = 1.0 - 1j * sprof / (omega * epsb)
        pml_e = np.stack([s.real, s.imag], axis=2).reshape(ne, 6)
    elif pml_mode == "sigma":
        layer_k = np.ceil(np.max(depth, axis=1) / h - 1e-9).astype(int)
        in_pml = layer_k > 0
        sigma_e[in_pml] += sig_max * ((layer_k[in_pml] - 0.5) / npml) ** 2
    elif pml_mode != "none":
        raise ValueError(pml_mode)

    # source: unit z current on the edge through the box centre
    tet_edges, edge_vtx = edges_of(tets, nv)
    mid = crd[edge_vtx].mean(axis=1)
    dz = crd[edge_vtx[:, 1]] - crd[edge_vtx[:, 0]]
    zedge = (np.abs(dz[:, 0]) < 1e-12) & (np.abs(dz[:, 1]) < 1e-12)
    cbox = 0.5 * (lo + hi)
    dist = np.linalg.norm(mid - cbox, axis=1) + np.where(zedge, 0.0, 1e9)
    src_edge = int(np.argmin(dist))
    src_e = np.any(tet_edges == src_edge, axis=1).astype(np.int32)
    joe_v = np.zeros((nv, 6))
    # Joe = -(Jo + i Jor) with Jor = z: per vertex (re,im) x (x,y,z)
    for vtx in edge_vtx[src_edge]:
        joe_v[vtx, 5] = -1.0

    pec = np.zeros(len(edge_vtx), dtype=bool)
    # PEC outer boundary: edges lying in a face of the box
    ev0, ev1 = crd[edge_vtx[:, 0]], crd[edge_vtx[:, 1]]
    for a in range(3):
        for val in (lo[a], hi[a]):
            pec |= (np.abs(ev0[:, a] - val) < 1e-12) & (np.abs(ev1[:, a] - val) < 1e-12)
    if cfg == 4:
        # PEC strip dipole in the middle layer plane, fed by the centre edge gap
        zc = zs[len(zs) // 2]
        strip = ((np.abs(ev0[:, 2] - zc) < 1e-12) & (np.abs(ev1[:, 2] - zc) < 1e-12) &
                 (np.abs(mid[:, 0] - 0.5) < 1.5 * h) & (np.abs(mid[:, 1] - 0.5) < 0.25) &
                 (np.abs(mid[:, 1] - 0.5) > 1.5 * h))
        pec |= strip
        src_e = np.any(tet_edges == src_edge, axis=1).astype(np.int32)

    # decomposition
    if cfg == 3:
        # lognormal subdomain sizes (edge DOFs) clipped to [150, 4000]; a cell
        # weighs the edges of its tetrahedra (7/6 per tetrahedron), and the
        # targets are scaled by CFG3_FILL for the edges a subdomain shares
        ncell = dims[0] * dims[1] * dims[2]
        sizes = np.exp(rng.normal(math.log(600.0), 0.7, size=20000)).clip(150, 4000)
        weight = np.bincount(cell, minlength=ncell) * (7.0 / 6.0)
        u = np.log(sizes / 150.0) / math.log(4000.0 / 150.0)
        fill = CFG3_FILL[0] + (CFG3_FILL[1] - CFG3_FILL[0]) * u
        sub_of_cell, part_of_sub = weighted_sfc_partition(dims, weight, fill * sizes, 8)
        part_of_sub = part_of_sub.astype(np.int64)
    else:
        sub_of_cell, part_of_sub = block_partition(dims, sub, part)

    model = decompose(crd, tets, cell, sub_of_cell, part_of_sub, eps_e, sigma_e, src_e,
                      joe_v, pml_e=pml_e, pec_edge_mask=pec, omega=omega, order=order,
                      name=f"cfg{cfg}" if small is None else f"box{small[0]}")
    model.meta = dict(cfg=cfg, seed=seed, dims=dims, pml_mode=pml_mode, f=f, npml=npml,
                      eps_r=eps_r, src_edge=src_edge)
    return model
