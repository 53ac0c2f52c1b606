"""Reader for the reference's ADVENTURE on-disk model and the sample-model
builder.

An ADV file is a concatenation of documents; each document is
``"AdvDoc\\0" | u64 data_len | data | u64 prop_len | props`` with
``props = u32 id_len, id, u32 nprops, (u32 klen, key, u32 vlen, value)*``
(little-endian).  Documents are found by their ``content_type`` or ``label``
property, as adv_dbox_find_by_property does for the reference
(lib/std/fileio.c:760-990 reads the partition files this way).

`read_partition` returns the raw per-part data of ``model/advhddm_in_P.adv``;
`build_full_wave_model` turns the parts into the solver's data model with the
reference's full-wave rules:

* local vertex ids follow subdomain-node order over the element vertices,
  local edge (mid-node) ids the remaining nodes in subdomain-node order
  (make_MESH_make_ap / make_MESH_change_ap, lib/fem/make_MESH.c:237-272,
  539-553), and the 8<->9 mid-node swap of make_MESH_change_nop
  (make_MESH.c:557-585) makes edge k join tet vertices tet_n[k];
* only mid-nodes carry DOFs: an edge is an interface DOF when more than one
  subdomain holds it (vertex interface nodes are dropped,
  lib/hddm/make_PartMesh.c:1603-1624); it is owned by the lowest part
  holding it;
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


def build_full_wave_model(parts: List[dict], mtrl: dict, antenna_pp: Dict[int, List[dict]],
                          order: str = "reference", name: str = "sample",
                          volt: List[dict] = None) -> mesh.Model:
    """Solver data model of a partitioned quadratic-tet model (full-wave
    E method).  order="reference" keeps the reference's local numbering.
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
    # which subdomains / parts hold each mid-node
    mid_sets: Dict[int, set] = {}
    holders_sub = np.zeros(nglob, np.int64)
    part_min = np.full(nglob, npart, np.int64)
    part_sets = [set() for _ in range(nglob)] if False else None
    seen = np.full(nglob, -1, np.int64)
    for s, (p, d, gn, fl, s2g) in enumerate(subs):
        mids = np.unique(gn[:, 4:])
        holders_sub[mids] += 1
        part_min[mids] = np.minimum(part_min[mids], p)
    is_iface = holders_sub >= 2
    # parts holding each interface mid-node
    part_has = np.zeros((npart, nglob), bool)
    for (p, d, gn, fl, s2g) in subs:
        part_has[p, np.unique(gn[:, 4:])] = True
    ln_maps = []
    infree = np.zeros(npart, np.int64)
    outfree = np.zeros(npart, np.int64)
    for p in range(npart):
        e = np.nonzero(part_has[p] & is_iface)[0]
        nonown = e[part_min[e] != p]
        own = e[part_min[e] == p]
        outfree[p], infree[p] = len(nonown), len(own)
        ln_maps.append(np.concatenate([nonown, own]))
    lnp_all = []
    for p in range(npart):
        lnp = np.full(nglob, -1, np.int64)
        lnp[ln_maps[p]] = np.arange(len(ln_maps[p]))
        lnp_all.append(lnp)
    op_off, op_ibid = [0], []
    for p in range(npart):
        for q in range(npart):
            if p == q:
                lst = np.zeros(0, np.int64)
            else:
                shared = np.intersect1d(ln_maps[p], ln_maps[q])
                lst = lnp_all[p][shared]
            op_ibid.append(lst)
            op_off.append(op_off[-1] + len(lst))

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
    bc_off, ib_off, ib_nfid_all, ib_ln_all, ledge, mp2sn = [0], [0], [], [], [], []
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
        bc_off.append(bc_off[-1])
        ifc = np.nonzero(is_iface[mid_g])[0]
        ib_nfid_all.append(ifc)
        ib_ln_all.append(lnp_all[p][mid_g[ifc]])
        ib_off.append(ib_off[-1] + len(ifc))
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
        bc_off=np.array(bc_off, np.int32), ibc=np.zeros(0, np.int32), ib_off=np.array(ib_off, np.int32),
        ib_nfid=cat(ib_nfid_all).astype(np.int32), ib_ln=cat(ib_ln_all).astype(np.int32),
        infree=infree.astype(np.int32), outfree=outfree.astype(np.int32),
        op_off=np.array(op_off, np.int32), op_ibid=cat(op_ibid).astype(np.int32),
        ledge_global=cat(ledge).astype(np.int64), n_edges=int(is_iface.size), n_tets=int(el_off[-1]),
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

