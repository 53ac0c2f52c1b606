"""Reader for the reference's ADVENTURE on-disk model and the sample-model
builder.

An ADV file is a concatenation of documents; each document is
``"AdvDoc\\0" | u64 data_len | data | u64 prop_len | props`` with
``props = u32 id_len, id, u32 nprops, (u32 klen, key, u32 vlen, value)*``
(little-endian).  Documents are found by their ``content_type`` or ``label``
property, as adv_dbox_find_by_property does for the reference
(lib/std/fileio.c:760-990 reads the partition files this way).

`read_partition` returns the raw per-part data of ``model/advhddm_in_P.adv``
(`write_partition` writes it); `build_full_wave_model` turns the parts into
the solver's data model with the reference's full-wave rules:

* local vertex ids follow subdomain-node order over the element vertices,
  local edge (mid-node) ids the remaining nodes in subdomain-node order
  (make_MESH_make_ap / make_MESH_change_ap, lib/fem/make_MESH.c:237-272,
  539-553), and the 8<->9 mid-node swap of make_MESH_change_nop
  (make_MESH.c:557-585) makes edge k join tet vertices tet_n[k];
* only mid-nodes carry DOFs: the file's node-level interface is
  reclassified as the reference does (reclassify_interface: essential
  nodes and vertex nodes leave the interface, the mid-nodes keep the charge
  part the file gives them and are renumbered per part, copies charged
  elsewhere first, in file order; lib/hddm/make_PartMesh.c:158-858), and a
  subdomain's Dirichlet list keeps its essential mid-nodes, sorted
  (make_PartMesh.c:1082-1094);
* the current source is the antenna's Parallelpiped definition
  (read_Jodefine_file_pp, fileio.c:1226-1293; make_Jo_from_definition_pp,
  fileio.c:1556-1594) on the vertices of antenna-flagged elements, with the
  two correct-Jo passes of main.c:92-98 leaving Jor = the file's Jor and
  Jo = the file's Joi, so Joe = -(Jo + i Jor) (make_elem_em.c:291-296);
* a volt.dat with dg > 0 replaces the current source by the delta-gap
  voltages of its boxes on the vertices inside them (make_elem_em.c:320-396).
"""
from __future__ import annotations

import os
import struct
from typing import Dict, List

import numpy as np

from . import mesh


def read_adv(path: str) -> List[dict]:
    """All documents of an ADV file: [{'props': {...}, 'data': bytes}]."""
    data = open(path, "rb").read()
    docs = []
    off = 0
    while off < len(data):
        if data[off:off + 7] != b"AdvDoc\0":
            raise ValueError(f"{path}: bad document header at {off}")
        off += 7
        (dlen,) = struct.unpack_from("<Q", data, off)
        off += 8
        payload = data[off:off + dlen]
        off += dlen
        (plen,) = struct.unpack_from("<Q", data, off)
        off += 8
        props = data[off:off + plen]
        off += plen
        p = 0
        (idl,) = struct.unpack_from("<I", props, p)
        p += 4 + idl
        (npr,) = struct.unpack_from("<I", props, p)
        p += 4
        kv = {}
        for _ in range(npr):
            (kl,) = struct.unpack_from("<I", props, p)
            p += 4
            key = props[p:p + kl].decode()
            p += kl
            (vl,) = struct.unpack_from("<I", props, p)
            p += 4
            kv[key] = props[p:p + vl].decode(errors="replace")
            p += vl
        docs.append({"props": kv, "data": payload})
    return docs


def _find(docs, key, value):
    for d in docs:
        if d["props"].get(key) == value:
            return d
    raise KeyError(f"document {key}={value} not found")


def _per_sub_int(buf: bytes, nsub: int, width: int = 1):
    """Per-subdomain int32 records: [n, n*width ints] x nsub."""
    a = np.frombuffer(buf, dtype="<i4")
    out, p = [], 0
    for _ in range(nsub):
        n = int(a[p])
        p += 1
        out.append(a[p:p + n * width].copy())
        p += n * width
    return out


def read_partition(path: str) -> dict:
    """Raw data of one ``advhddm_in_P.adv`` (fileio.c:760-990)."""
    docs = read_adv(path)
    opt = _find(docs, "label", "Options")["props"]
    node = _find(docs, "content_type", "Node")
    nnd = int(node["props"]["num_items"])
    crd = np.frombuffer(node["data"], dtype="<f8")[: 3 * nnd].reshape(nnd, 3).copy()
    p2g = np.frombuffer(_find(docs, "label", "NodeIndex_PartToGlobal")["data"], dtype="<i4")[:nnd].copy()
    ifd = _find(docs, "content_type", "HDDM_InterfaceDOF")
    npart = int(ifd["props"]["num_parts"])
    a = np.frombuffer(ifd["data"], dtype="<i4")
    op, p = [], 0
    for _ in range(npart):
        n = int(a[p])
        op.append(a[p + 1:p + 1 + n].copy())
        p += 1 + n
    el = _find(docs, "content_type", "HDDM_Element")
    nsub = int(el["props"]["num_subdomains"])
    nd_elem = int(el["props"]["num_nodes_per_element"])
    nops = [x.reshape(-1, nd_elem) for x in _per_sub_int(el["data"], nsub, nd_elem)]
    s2p = _per_sub_int(_find(docs, "label", "NodeIndex_SubdomainToPart")["data"], nsub)
    inbc = [x.reshape(-1, 5) for x in _per_sub_int(_find(docs, "label", "InterfaceDOF")["data"], nsub, 5)]
    flags = _per_sub_int(_find(docs, "label", "Flag")["data"], nsub)
    dbc = None
    if opt.get("DirichletBCs_Axn0") == "NEED":
        dbc = _per_sub_int(_find(docs, "label", "DirichletBCs_Axn0")["data"], nsub)
    return dict(part=int(ifd["props"]["part_number"]), npart=npart, nsub=nsub,
                nnd_all=int(node["props"]["num_items_orig"]), crd=crd, pnid2gnid=p2g, op=op,
                infree=int(ifd["props"]["total_num_infree"]), outfree=int(ifd["props"]["total_num_outfree"]),
                nop=nops, snid2pnid=s2p, inbc=inbc, flag=flags, dbc=dbc,
                elem_type=el["props"]["element_type"])


def read_mtrl(path: str) -> dict:
    """mtrl.dat (read_mtrldat_file, fileio.c:3477-3572)."""
    tok = open(path).read().split()
    out = {"eps": {}, "sigma": {}, "antenna": {}, "omega": None}
    i = 0
    while i < len(tok):
        t = tok[i]
        if t == "Permittivity":
            n = int(tok[i + 1])
            for k in range(n):
                out["eps"][int(tok[i + 2 + 2 * k])] = float(tok[i + 3 + 2 * k])
            i += 2 + 2 * n
        elif t == "Conductor":
            n = int(tok[i + 1])
            for k in range(n):
                out["sigma"][int(tok[i + 2 + 2 * k])] = float(tok[i + 3 + 2 * k])
            i += 2 + 2 * n
        elif t == "Antenna":
            n = int(tok[i + 1])
            for k in range(n):
                out["antenna"][int(tok[i + 2 + 3 * k])] = (tok[i + 3 + 3 * k], tok[i + 4 + 3 * k])
            i += 2 + 3 * n
        elif t == "AntennaOmega":
            out["omega"] = float(tok[i + 1])
            i += 2
        else:
            i += 1
    return out


def read_parallelpiped(path: str) -> List[dict]:
    """Parallelpiped blocks of a Jo definition file: 4 points, Jor, Joi."""
    tok = open(path).read().split()
    out = []
    i = 0
    while i < len(tok):
        if tok[i] == "Parallelpiped":
            v = [float(x) for x in tok[i + 1:i + 19]]
            out.append({"ap": np.array(v[:12]).reshape(4, 3), "Jor": np.array(v[12:15]),
                        "Joi": np.array(v[15:18])})
            i += 19
        else:
            i += 1
    return out


def in_parallelpiped(x: np.ndarray, ap: np.ndarray) -> np.ndarray:
    """Points inside (or on) the parallelepiped p0 + sum t_k (p_k - p0),
    t in [0,1]; boundary tolerance 1e-5 of the shortest edge, as the
    reference's pp.tolerance."""
    vec = ap[1:4] - ap[0]
    t = np.linalg.solve(vec.T, (x - ap[0]).T).T
    tol = 1e-5 * np.linalg.norm(vec, axis=1).min() / np.linalg.norm(vec, axis=1)
    return np.all((t >= -tol) & (t <= 1 + tol), axis=1)


def read_volt(path: str) -> List[dict]:
    """Delta-gap definition ``volt.dat`` (read per element by
    tet_MakeElement_EM_E, lib/fem/make_elem_em.c:72-96,333-347): a first
    line ``dg``, then per gap box four lines ``x0 x1`` / ``y0 y1`` /
    ``z0 z1`` / ``dV``.  dg = 0 (no boxes) selects the current source."""
    lines = open(path).read().splitlines()
    dg = int(lines[0].split()[0])
    boxes = []
    for i in range(dg):
        x0, x1 = (float(t) for t in lines[1 + 4 * i].split()[:2])
        y0, y1 = (float(t) for t in lines[2 + 4 * i].split()[:2])
        z0, z1 = (float(t) for t in lines[3 + 4 * i].split()[:2])
        dv = float(lines[4 + 4 * i].split()[0])
        boxes.append({"lo": np.array([x0, y0, z0]), "hi": np.array([x1, y1, z1]), "dV": dv})
    return boxes


def gap_voltage(x: np.ndarray, boxes: List[dict]) -> np.ndarray:
    """Per point the dV of the last gap box containing it (bounds included,
    make_elem_em.c:350-361), 0 outside every box."""
    x = np.asarray(x, np.float64).reshape(-1, 3)
    out = np.zeros(len(x))
    for b in boxes:
        inside = np.all((b["lo"] <= x) & (x <= b["hi"]), axis=1)
        out[inside] = b["dV"]
    return out


# Pinbc bc_flag values and the node-change marks of the reference's interface
# reclassification (lib/hddm/libhddm.h:25-34)
NOT_CHANGE_NODE = -1          # __ADVFW_HDDM_CHANGE_INNER_BC_NOT_CHANGE__: a mid-node (edge DOF)
FLAG_ESSENTIAL = -2
FLAG_NOT_CHANGE = -1
FLAG_NO_DOF = 0


def reclassify_interface(parts: List[dict]) -> List[dict]:
    """The E-method interface of every part, from the partition file's own
    node interface (make_PartMesh_change_innerbc_func,
    lib/hddm/make_PartMesh.c:158-248, with the s-mode exchange of
    make_PartMesh_s.c:51-128):

    * 1st step (:393-422): an interface node in a subdomain's own Dirichlet
      list marks its part-level entry essential;
    * per subdomain the mid-nodes are "not change" nodes, the vertices carry
      no edge DOF (make_nf_A_Method, :1600-1624); a part-level entry takes
      the charge part of the inbc records that name it as a mid-node
      (add_nf, :374-389), an essential entry no DOF;
    * 3rd step (:491-586): essential stays, "not change" entries keep their
      DOF, the rest (vertex nodes) become NO_DOF;
    * 4th step (:590-640): the surviving entries are renumbered in their
      file order, copies owned by another part (charge part != this part)
      first, then the owned ones after them;
    * change_Inbc / change_OPinfo (:707-858): every subdomain's inbc list
      and every part-pair list keep their surviving entries, in file order,
      with the new numbers.

    Returns per part {"infree", "outfree", "op": [list per part],
    "inbc": [(snid, ln_part) arrays per subdomain]}.  The flags that only
    the magnetic-field formulations produce (ONLY_ONE, ONLY_THIS_PART,
    INFREE, OUTFREE) cannot arise here."""
    out = []
    for p, P in enumerate(parts):
        tfree = int(P["infree"]) + int(P["outfree"])
        nf = np.zeros(tfree, np.int64)
        flag = np.zeros(tfree, np.int64)
        charge = np.zeros(tfree, np.int64)
        if P["dbc"] is not None:                                  # 1st step
            for d in range(P["nsub"]):
                ib = P["inbc"][d]
                hit = np.isin(ib[:, 0], P["dbc"][d])
                flag[ib[hit, 4]] = FLAG_ESSENTIAL
        node_nf = []
        for d in range(P["nsub"]):                                # add_nf
            nnd = len(P["snid2pnid"][d])
            nfd = np.zeros(nnd, np.int64)
            nfd[P["nop"][d][:, 4:].ravel()] = NOT_CHANGE_NODE
            node_nf.append(nfd)
            for snid, _coord, cp, _lnc, lnp in P["inbc"][d]:
                if flag[lnp] == FLAG_ESSENTIAL:
                    nf[lnp] = 0
                elif nfd[snid] == NOT_CHANGE_NODE:
                    nf[lnp] = NOT_CHANGE_NODE
                    charge[lnp] = cp
        keep = flag != FLAG_ESSENTIAL                             # 3rd step
        flag[keep & (nf == NOT_CHANGE_NODE)] = FLAG_NOT_CHANGE
        flag[keep & (nf == 0)] = FLAG_NO_DOF
        if np.any(keep & (nf > 0)):
            raise ValueError(f"part {p}: an interface entry of a magnetic-field formulation in an E-method model")
        alive = flag == FLAG_NOT_CHANGE                           # 4th step
        own = alive & (charge == p)
        other = alive & (charge != p)
        new_ln = np.full(tfree, -1, np.int64)
        new_ln[other] = np.arange(int(other.sum()))
        new_ln[own] = int(other.sum()) + np.arange(int(own.sum()))
        inbc = []
        for d in range(P["nsub"]):                                # change_Inbc
            ib = P["inbc"][d]
            ok = alive[ib[:, 4]] & (node_nf[d][ib[:, 0]] != 0)
            inbc.append(np.stack([ib[ok, 0], new_ln[ib[ok, 4]]], axis=1))
        op = [new_ln[o[alive[o]]] for o in P["op"]]               # change_OPinfo
        out.append({"infree": int(own.sum()), "outfree": int(other.sum()), "op": op, "inbc": inbc})
    return out


def build_full_wave_model(parts: List[dict], mtrl: dict, antenna_pp: Dict[int, List[dict]],
                          order: str = "reference", name: str = "sample",
                          volt: List[dict] = None) -> mesh.Model:
    """Solver data model of a partitioned quadratic-tet model (full-wave
    E method), numbered as the reference numbers it: the interface of
    reclassify_interface, and per subdomain (make_PartMesh_FW_TetQ,
    make_PartMesh.c:1628-1707) vertices and mid-nodes in subdomain-node
    order, the 8<->9 mid-node swap, the essential mid-nodes of the
    subdomain's Dirichlet list sorted (make_PartMesh.c:1082-1094) and the
    interface mid-nodes sorted by local id (sort_inbc, :1167-1194).
    volt: gap boxes of a volt.dat with dg > 0 (delta-gap source), or
    None/empty for the current source."""
    npart = len(parts)
    # global node data
    nglob = parts[0]["nnd_all"]
    gcrd = np.zeros((nglob, 3))
    for P in parts:
        gcrd[P["pnid2gnid"]] = P["crd"]
    # global element list in part-major / subdomain order, global node ids
    subs = []   # (part, sub, gnop[nel,10] in post-swap order, flags, gnode order)
    for p, P in enumerate(parts):
        for d in range(P["nsub"]):
            snid2gnid = P["pnid2gnid"][P["snid2pnid"][d]]
            nop = P["nop"][d].copy()
            nop[:, [8, 9]] = nop[:, [9, 8]]           # edge k joins tet_n[k]
            subs.append((p, d, snid2gnid[nop], P["flag"][d], snid2gnid))
    nsub = len(subs)
    iface = reclassify_interface(parts)
    infree = np.array([x["infree"] for x in iface], np.int64)
    outfree = np.array([x["outfree"] for x in iface], np.int64)
    op_off, op_ibid = [0], []
    for p in range(npart):
        for q in range(npart):
            lst = iface[p]["op"][q] if q != p else np.zeros(0, np.int64)
            op_ibid.append(lst)
            op_off.append(op_off[-1] + len(lst))
    # global mid-nodes: every node that is a mid-node of some element
    is_mid = np.zeros(nglob, bool)
    for (p, d, gn, fl, s2g) in subs:
        is_mid[gn[:, 4:].ravel()] = True

    # antenna vertices with the Parallelpiped current
    ant_flags = sorted(antenna_pp)
    Jor_v = np.zeros((nglob, 3))
    Joi_v = np.zeros((nglob, 3))
    ant_vertex = np.zeros(nglob, bool)
    for (p, d, gn, fl, s2g) in subs:
        for f in ant_flags:
            ant_vertex[np.unique(gn[fl == f][:, :4])] = True
    for f in ant_flags:
        for pp in antenna_pp[f]:
            idx = np.nonzero(ant_vertex)[0]
            inside = idx[in_parallelpiped(gcrd[idx], pp["ap"])]
            Jor_v[inside] = pp["Jor"]
            Joi_v[inside] = pp["Joi"]
    Jc = -(Joi_v + 1j * Jor_v)          # Joe = -(Jo + i Jor), Jo = file Joi, Jor = file Jor
    joe_g = np.stack([Jc.real, Jc.imag], axis=2).reshape(nglob, 6)

    nop_all, ap_off, crd_all, joe_all, nmp = [], [0], [], [], []
    bc_off, ib_off, ib_nfid_all, ib_ln_all, ledge, mp2sn, ibc_all = [0], [0], [], [], [], [], []
    eps_all, sig_all, src_all, el_off = [], [], [], [0]
    for (p, d, gn, fl, s2g) in subs:
        vert = np.zeros(len(s2g), bool)
        g2s = {int(g): k for k, g in enumerate(s2g)}
        sl = np.vectorize(g2s.get)(gn)
        vert[sl[:, :4].ravel()] = True
        ap = np.full(len(s2g), -1, np.int64)
        ap[vert] = np.arange(vert.sum())
        mp = np.full(len(s2g), -1, np.int64)
        mp[~vert] = np.arange((~vert).sum())
        nop = np.concatenate([ap[sl[:, :4]], mp[sl[:, 4:]]], axis=1)
        assert (nop >= 0).all()
        nop_all.append(nop.ravel())
        crd_all.append(gcrd[s2g[vert]])
        joe_all.append(joe_g[s2g[vert]])
        ap_off.append(ap_off[-1] + int(vert.sum()))
        mid_g = s2g[~vert]
        nmp.append(len(mid_g))
        ledge.append(mid_g)
        mp2sn.append(np.nonzero(~vert)[0])
        # essential mid-nodes of the subdomain's Dirichlet list, sorted
        # (make_MESH_change_dbc_Axn0_1st/2nd + sort_i; vertex entries go to
        # the antenna conditions, make_MESH.c:409-432)
        dbc = parts[p]["dbc"]
        ess = np.zeros(0, np.int64)
        if dbc is not None:
            raw = np.asarray(dbc[d], np.int64)
            ess = np.sort(mp[raw[~vert[raw]]])
        ibc_all.append(ess)
        bc_off.append(bc_off[-1] + len(ess))
        # interface mid-nodes sorted by local id, with their new ln_part
        ib = iface[p]["inbc"][d]
        nfid = mp[ib[:, 0]]
        srt = np.argsort(nfid, kind="stable")
        ib_nfid_all.append(nfid[srt])
        ib_ln_all.append(ib[srt, 1])
        ib_off.append(ib_off[-1] + len(nfid))
        eps_all.append(np.array([mtrl["eps"][int(f)] for f in fl]))
        sig_all.append(np.array([mtrl["sigma"].get(int(f), 0.0) for f in fl]))
        src_all.append(np.isin(fl, ant_flags).astype(np.int32))
        el_off.append(el_off[-1] + len(fl))

    cat = np.concatenate
    crd_cat = cat(crd_all).reshape(-1)
    dgv = gap_voltage(crd_cat, volt) if volt else None
    return mesh.Model(
        npart=npart, nsub=nsub, omega=float(mtrl["omega"]), dgv=dgv,
        sub_part=np.array([s[0] for s in subs], np.int32), el_off=np.array(el_off, np.int32),
        nop=cat(nop_all).astype(np.int32), eps=cat(eps_all), sigma=cat(sig_all),
        src=cat(src_all).astype(np.int32), pml=None, ap_off=np.array(ap_off, np.int32),
        crd=crd_cat, joe=cat(joe_all).reshape(-1), nmp=np.array(nmp, np.int32),
        bc_off=np.array(bc_off, np.int32), ibc=cat(ibc_all).astype(np.int32), ib_off=np.array(ib_off, np.int32),
        ib_nfid=cat(ib_nfid_all).astype(np.int32), ib_ln=cat(ib_ln_all).astype(np.int32),
        infree=infree.astype(np.int32), outfree=outfree.astype(np.int32),
        op_off=np.array(op_off, np.int32), op_ibid=cat(op_ibid).astype(np.int32),
        ledge_global=cat(ledge).astype(np.int64), n_edges=int(nglob), n_tets=int(el_off[-1]),
        mp2sn=cat(mp2sn).astype(np.int32), name=name)


def load_npz_model(path: str) -> mesh.Model:
    """A data model saved with numpy.savez (e.g. the sample-model fixture)."""
    z = np.load(path, allow_pickle=False)
    kw = {k: z[k] for k in ("sub_part", "el_off", "nop", "eps", "sigma", "src", "ap_off", "crd", "joe", "nmp",
                            "bc_off", "ibc", "ib_off", "ib_nfid", "ib_ln", "infree", "outfree", "op_off",
                            "op_ibid", "ledge_global")}
    for k in ("mp2sn", "dgv"):
        if k in z.files:
            kw[k] = z[k]
    return mesh.Model(npart=int(z["npart"]), nsub=int(z["nsub"]), omega=float(z["omega"]), pml=None,
                      n_edges=int(z["n_edges"]), n_tets=int(z["n_tets"]), name=os.path.basename(path), **kw)


def load_reference_model(model_dir: str) -> mesh.Model:
    """The full-wave model of a reference MODEL directory (model/, mtrl.dat,
    antenna definitions, volt.dat)."""
    files = sorted(f for f in os.listdir(os.path.join(model_dir, "model")) if f.startswith("advhddm_in_"))
    parts = [read_partition(os.path.join(model_dir, "model", f))
             for f in sorted(files, key=lambda s: int(s.split("_")[-1].split(".")[0]))]
    mtrl = read_mtrl(os.path.join(model_dir, "mtrl.dat"))
    pp = {}
    for flag, (mode, fname) in mtrl["antenna"].items():
        pp[flag] = read_parallelpiped(os.path.join(model_dir, fname))
    vpath = os.path.join(model_dir, "volt.dat")
    volt = read_volt(vpath) if os.path.exists(vpath) else None
    return build_full_wave_model(parts, mtrl, pp, volt=volt)


# ---------------------------------------------------------------------------
# partition files (the inverse of read_partition)
# ---------------------------------------------------------------------------
def _int_doc(records, width: int = 1) -> bytes:
    """int32 records [n, n*width values] one after another"""
    out = []
    for r in records:
        r = np.asarray(r, "<i4").ravel()
        out.append(np.array([len(r) // width], "<i4").tobytes())
        out.append(r.tobytes())
    return b"".join(out)


def write_partition(model, out_dir: str, flags: np.ndarray, essential: bool = True) -> List[str]:
    """Writes `model` as reference partition files model/advhddm_in_P.adv
    (the layout read_partition reads, fileio.c:760-990): node-level data as
    ADVENTURE_Metis leaves it, before the full-wave reclassification.

    Global nodes: the vertices (told apart by coordinates), then one
    mid-node per global edge.  Per part the nodes of its subdomains in
    global order; per subdomain its nodes in global order, elements with the
    file's mid-node order (edges (1,3) and (2,3) in the file's positions 9
    and 8).  Interface nodes (vertices and mid-nodes held by two or more
    subdomains) are charged to the lowest part holding them and numbered per
    part, copies charged elsewhere first, each group in global order; the
    part-pair lists in global order.  flags: per element material flag.
    essential: write DirichletBCs_Axn0 (Options NEED) with the essential
    edges' mid-nodes and their end vertices."""
    os.makedirs(out_dir, exist_ok=True)
    nop = np.asarray(model.nop, np.int64).reshape(-1, 10)
    crd = np.asarray(model.crd, np.float64).reshape(-1, 3)
    nmp = np.asarray(model.nmp, np.int64)
    dof0 = np.concatenate([[0], np.cumsum(nmp)])
    lg = np.asarray(model.ledge_global, np.int64)
    # global vertex ids by coordinates
    uq, vinv = np.unique(np.round(crd, 12), axis=0, return_inverse=True)
    vinv = vinv.ravel()
    nv = len(uq)
    ne_glob = int(lg.max()) + 1
    gcrd = np.zeros((nv + ne_glob, 3))
    gcrd[:nv] = uq
    sub_nodes, sub_gnop = [], []
    ess_nodes = []
    for s in range(model.nsub):
        e0, e1 = model.el_off[s], model.el_off[s + 1]
        a0 = model.ap_off[s]
        gv = vinv[a0 + nop[e0:e1, :4]]
        ge = nv + lg[dof0[s] + nop[e0:e1, 4:]]
        gn = np.concatenate([gv, ge], axis=1)
        vtx_of_edge = gv[:, [0, 0, 0, 1, 1, 2]], gv[:, [1, 2, 3, 2, 3, 3]]
        gcrd[ge] = 0.5 * (gcrd[vtx_of_edge[0]] + gcrd[vtx_of_edge[1]])
        sub_gnop.append(gn)
        sub_nodes.append(np.unique(gn))
        ibc = np.asarray(model.ibc[model.bc_off[s]:model.bc_off[s + 1]], np.int64)
        eg = nv + lg[dof0[s] + ibc]
        ends = set()
        for k in range(6):
            hit = np.isin(ge[:, k], eg)
            ends.update(vtx_of_edge[0][hit, k].tolist())
            ends.update(vtx_of_edge[1][hit, k].tolist())
        ess_nodes.append(np.union1d(eg, np.array(sorted(ends), np.int64)))
    nglob = nv + ne_glob
    holders = np.zeros(nglob, np.int64)
    owner = np.full(nglob, np.iinfo(np.int64).max)
    for s in range(model.nsub):
        holders[sub_nodes[s]] += 1
        owner[sub_nodes[s]] = np.minimum(owner[sub_nodes[s]], model.sub_part[s])
    is_if = holders >= 2
    npart = model.npart
    part_nodes = [np.unique(np.concatenate([sub_nodes[s] for s in range(model.nsub) if model.sub_part[s] == p]))
                  for p in range(npart)]
    lnmap = []
    for p in range(npart):
        ifn = part_nodes[p][is_if[part_nodes[p]]]
        order = np.concatenate([ifn[owner[ifn] != p], ifn[owner[ifn] == p]])
        m = np.full(nglob, -1, np.int64)
        m[order] = np.arange(len(order))
        lnmap.append((m, int((owner[ifn] != p).sum()), int((owner[ifn] == p).sum())))
    paths = []
    nel_all = int(model.el_off[-1])
    for p in range(npart):
        subs = [s for s in range(model.nsub) if model.sub_part[s] == p]
        pn = part_nodes[p]
        g2p = np.full(nglob, -1, np.int64)
        g2p[pn] = np.arange(len(pn))
        m, outf, inf = lnmap[p]
        docs = []
        docs.append(("Node", [("content_type", "Node"), ("num_items", len(pn)), ("num_items_orig", nglob),
                              ("format", "f8f8f8")], gcrd[pn].astype("<f8").tobytes()))
        docs.append(("p2g", [("label", "NodeIndex_PartToGlobal"), ("num_items", len(pn))],
                     pn.astype("<i4").tobytes()))
        opdata = []
        for q in range(npart):
            if q == p:
                opdata.append(np.zeros(0, np.int64))
                continue
            shared = np.intersect1d(pn[is_if[pn]], part_nodes[q])
            opdata.append(m[shared])
        docs.append(("ifd", [("content_type", "HDDM_InterfaceDOF"), ("num_parts", npart), ("part_number", p),
                             ("num_subdomains", len(subs)), ("total_num_infree", inf),
                             ("total_num_outfree", outf)], _int_doc(opdata)))
        els, s2p, inbc, dbc, flg, e2g = [], [], [], [], [], []
        for s in subs:
            sn = sub_nodes[s]
            g2s = np.full(nglob, -1, np.int64)
            g2s[sn] = np.arange(len(sn))
            raw = g2s[sub_gnop[s]]
            raw[:, [8, 9]] = raw[:, [9, 8]]          # the file's mid-node order
            els.append(raw)
            s2p.append(g2p[sn])
            ifs = sn[is_if[sn]]
            rows = np.stack([g2s[ifs], np.zeros(len(ifs), np.int64), owner[ifs],
                             np.array([lnmap[owner[g]][0][g] for g in ifs], np.int64), m[ifs]], axis=1)
            inbc.append(rows)
            dbc.append(g2s[np.intersect1d(ess_nodes[s], sn)])
            flg.append(flags[model.el_off[s]:model.el_off[s + 1]])
            e2g.append(np.arange(model.el_off[s], model.el_off[s + 1]))
        docs.append(("el", [("content_type", "HDDM_Element"), ("element_type", "3DQuadraticTetrahedron"),
                            ("num_nodes_per_element", 10), ("num_subdomains", len(subs)),
                            ("num_items_orig", nel_all)], _int_doc(els, 10)))
        docs.append(("s2p", [("label", "NodeIndex_SubdomainToPart")], _int_doc(s2p)))
        docs.append(("inbc", [("label", "InterfaceDOF"), ("format", "i4i4i4i4")], _int_doc(inbc, 5)))
        if essential:
            docs.append(("dbc", [("label", "DirichletBCs_Axn0")], _int_doc(dbc)))
        docs.append(("flag", [("label", "Flag")], _int_doc(flg)))
        docs.append(("e2g", [("label", "ElementIndex_SubdomainToGlobal")], _int_doc(e2g)))
        docs.append(("opt", [("label", "Options"), ("DirichletBCs_Axn0", "NEED" if essential else "NO_NEED")],
                     b""))
        path = os.path.join(out_dir, f"advhddm_in_{p}.adv")
        with open(path, "wb") as f:
            for k, (tag, props, payload) in enumerate(docs):
                f.write(_doc_bytes(f"fwhip-part{p}-{tag}-{k}", props, payload))
        paths.append(path)
    return paths


# ---------------------------------------------------------------------------
# result files (write_outdata_file_HDDM_FW, lib/std/fileio.c:2930-3224)
# ---------------------------------------------------------------------------
RESULT_UNITS = {"MagneticField": "T", "ElectricField": "T", "CurrentDensity": "A/m^2"}


def _doc_bytes(doc_id: str, props: List[tuple], payload: bytes) -> bytes:
    """One AdvDoc document; props in the order given."""
    pb = [struct.pack("<I", len(doc_id)), doc_id.encode(), struct.pack("<I", len(props))]
    for k, v in props:
        kb, vb = k.encode(), str(v).encode()
        pb += [struct.pack("<I", len(kb)), kb, struct.pack("<I", len(vb)), vb]
    pbytes = b"".join(pb)
    return b"AdvDoc\0" + struct.pack("<Q", len(payload)) + payload + struct.pack("<Q", len(pbytes)) + pbytes


def result_documents(model, x, B, E, J, part: int, id_prefix: str = "fwhip"):
    """The documents of part `part`'s result file, as (props, payload) pairs
    in the reference's order: MagneticVectorPotential (per subdomain nmp,
    then (mpid2snid, re, im) per mid-node), then the real parts of
    MagneticField, ElectricField, CurrentDensity, then their imaginary parts
    (per subdomain nel, then 3 values per element; fileio.c:2962-3219).
    x: per-subdomain solutions concatenated; B, E, J: (nel, 3) complex."""
    if model.mp2sn is None:
        raise ValueError("the model has no mid-node -> subdomain node map (mp2sn)")
    subs = np.nonzero(np.asarray(model.sub_part) == part)[0]
    dof0 = np.concatenate([[0], np.cumsum(np.asarray(model.nmp, np.int64))])
    rec = np.dtype([("id", "<i4"), ("re", "<f8"), ("im", "<f8")])
    out = []
    chunks = []
    for s in subs:
        n = int(model.nmp[s])
        r = np.empty(n, rec)
        r["id"] = model.mp2sn[dof0[s]:dof0[s + 1]]
        xs = np.asarray(x[dof0[s]:dof0[s + 1]])
        r["re"], r["im"] = xs.real, xs.imag
        chunks += [struct.pack("<i", n), r.tobytes()]
    out.append(([("index_byte", 4), ("label", "MagneticVectorPotential"), ("format", "f8f8"),
                 ("fega_type", "NodeVariable"), ("num_subdomains", len(subs)),
                 ("content_type", "HDDM_FEGenericAttribute")], b"".join(chunks)))
    nel_all = int(model.el_off[-1])
    for comp, tag in ((np.real, "Real"), (np.imag, "Imaginary")):
        for name, F in (("MagneticField", B), ("ElectricField", E), ("CurrentDensity", J)):
            chunks = []
            for s in subs:
                e0, e1 = int(model.el_off[s]), int(model.el_off[s + 1])
                chunks += [struct.pack("<i", e1 - e0), np.ascontiguousarray(comp(F[e0:e1]), "<f8").tobytes()]
            out.append(([("index_byte", 4), ("unit", RESULT_UNITS[name]), ("label", name + tag),
                         ("format", "f8f8f8"), ("fega_type", "AllElementVariable"),
                         ("num_items_orig", nel_all), ("num_subdomains", len(subs)),
                         ("content_type", "HDDM_FEGenericAttribute")], b"".join(chunks)))
    return out


def write_result(out_dir: str, model, x, B, E, J, id_prefix: str = "fwhip"):
    """advhddm_out_P.adv for every part plus the options file advhddm_out.adv
    (write_outdata_file_HDDM_FW and write_outdata_file_options_FW,
    fileio.c:2930-3224).  Document ids are this writer's own (the reference
    draws unique ids from ADVENTURE_IO)."""
    os.makedirs(out_dir, exist_ok=True)
    paths = []
    opts = [("Work", "HDDM_ANALYSIS"), ("Formulation", "EMETHOD"), ("Analysis", "Mag_Full_Wavecurrent"),
            ("index_byte", 4), ("label", "Options"), ("fega_type", "AllElementConstant"), ("format", ""),
            ("num_items", 0), ("content_type", "FEGenericAttribute")]
    p = os.path.join(out_dir, "advhddm_out.adv")
    with open(p, "wb") as f:
        f.write(_doc_bytes(f"{id_prefix}-options", opts, b""))
    paths.append(p)
    for part in range(model.npart):
        p = os.path.join(out_dir, f"advhddm_out_{part}.adv")
        with open(p, "wb") as f:
            for k, (props, payload) in enumerate(result_documents(model, x, B, E, J, part)):
                f.write(_doc_bytes(f"{id_prefix}-{part}-{k}", props, payload))
        paths.append(p)
    return paths


def read_result(path: str) -> Dict[str, list]:
    """Decode a part result file: {'x': [complex per subdomain], 'MagneticField':
    [(nel,3) complex per subdomain], 'ElectricField': ..., 'CurrentDensity': ...,
    'ids': [mpid2snid per subdomain], 'props': {label: props}}."""
    docs = {d["props"]["label"]: d for d in read_adv(path)}
    out = {"props": {k: d["props"] for k, d in docs.items()}}
    d = docs["MagneticVectorPotential"]["data"]
    rec = np.dtype([("id", "<i4"), ("re", "<f8"), ("im", "<f8")])
    off, xs, ids = 0, [], []
    for _ in range(int(docs["MagneticVectorPotential"]["props"]["num_subdomains"])):
        (n,) = struct.unpack_from("<i", d, off)
        off += 4
        r = np.frombuffer(d, rec, count=n, offset=off)
        off += rec.itemsize * n
        xs.append(_cplx(r["re"], r["im"]))
        ids.append(r["id"].copy())
    out["x"], out["ids"] = xs, ids
    for name in ("MagneticField", "ElectricField", "CurrentDensity"):
        parts = []
        for tag in ("Real", "Imaginary"):
            d = docs[name + tag]["data"]
            off, vals = 0, []
            for _ in range(int(docs[name + tag]["props"]["num_subdomains"])):
                (n,) = struct.unpack_from("<i", d, off)
                off += 4
                vals.append(np.frombuffer(d, "<f8", count=3 * n, offset=off).reshape(n, 3))
                off += 24 * n
            parts.append(vals)
        out[name] = [_cplx(re, im) for re, im in zip(*parts)]
    return out


def _cplx(re, im):
    """complex array from its parts, keeping signed zeros (re + 1j*im would
    turn -0.0 into +0.0)"""
    out = np.empty(np.shape(re), np.complex128)
    out.real = re
    out.imag = im
    return out

