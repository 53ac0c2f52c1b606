"""Writes the sample-model fixtures from the reference's shipped example.

Reads (data only) /root/reference/main/fullwave/MODEL: the six partition
files, mtrl.dat, antenna.dat, calc_log/log_g_HDDM_Full_Wave and the result
files result/advhddm_out_P.adv.  Writes
  sample_model.npz     the solver data model (advio.load_reference_model),
  sample_log.json      the full-wave residual history the reference logged,
  sample_result.npz    the reference's own solution (complex64): the edge
                       values of every subdomain in model order, and the
                       element fields B, E, J at 4096 seeded elements,
  sample_result.json   the documents of each result file: label, properties
                       (document id left out) and payload length.
Run from the repository root: python tests/golden/make_sample_fixture.py
"""
import json
import os
import re
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from adventure_fullwave_amd import advio  # noqa: E402

MODEL = "/root/reference/main/fullwave/MODEL"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    m = advio.load_reference_model(MODEL)
    fields = {k: getattr(m, k) for k in ("sub_part", "el_off", "nop", "eps", "sigma", "src", "ap_off", "crd",
                                         "joe", "nmp", "bc_off", "ibc", "ib_off", "ib_nfid", "ib_ln", "infree",
                                         "outfree", "op_off", "op_ibid", "ledge_global", "mp2sn")}
    np.savez_compressed(os.path.join(HERE, "sample_model.npz"), npart=m.npart, nsub=m.nsub, omega=m.omega,
                        n_edges=m.n_edges, n_tets=m.n_tets, **fields)
    txt = open(os.path.join(MODEL, "calc_log", "log_g_HDDM_Full_Wave")).read()
    fw = txt.split("# Full Wave Analysis")[1]
    dof = int(re.search(r"# DOF on interface (\d+)", fw).group(1))
    norm = float(re.search(r"# Norm ([0-9.e+-]+)", fw).group(1))
    resid = [float(x.group(2)) for x in re.finditer(r"^(\d+) ([0-9.e+-]+)", fw, re.M)]
    json.dump({"source": "main/fullwave/MODEL/calc_log/log_g_HDDM_Full_Wave (Full Wave Analysis block)",
               "hddm_eps": 1e-3, "dof_on_interface": dof, "r0_norm": norm, "resid": resid},
              open(os.path.join(HERE, "sample_log.json"), "w"))
    print("wrote", os.path.join(HERE, "sample_model.npz"), "iterations", len(resid) - 1)

    # the reference's result files
    xs, ids, F = [], [], {k: [] for k in ("MagneticField", "ElectricField", "CurrentDensity")}
    layout = []
    for part in range(m.npart):
        path = os.path.join(MODEL, "result", f"advhddm_out_{part}.adv")
        r = advio.read_result(path)
        xs += r["x"]
        ids += r["ids"]
        for k in F:
            F[k] += r[k]
        layout.append([{"label": d["props"]["label"], "props": d["props"], "payload_bytes": len(d["data"])}
                       for d in advio.read_adv(path)])
    opts = advio.read_adv(os.path.join(MODEL, "result", "advhddm_out.adv"))
    x = np.concatenate(xs)
    assert np.array_equal(np.concatenate(ids), m.mp2sn), "mid-node order differs from the reference's result"
    nel = int(m.el_off[-1])
    sel = np.sort(np.random.default_rng(20261203).choice(nel, 4096, replace=False))
    np.savez_compressed(os.path.join(HERE, "sample_result.npz"), x=x.astype(np.complex64), elements=sel,
                        **{k: np.concatenate(v)[sel].astype(np.complex64) for k, v in F.items()})
    json.dump({"source": "main/fullwave/MODEL/result/advhddm_out*.adv (hddm_eps 1e-3)",
               "options": [{"props": d["props"], "payload_bytes": len(d["data"])} for d in opts],
               "parts": layout}, open(os.path.join(HERE, "sample_result.json"), "w"), indent=0)
    print("wrote sample_result.npz / .json:", x.size, "edge values,", sel.size, "elements")


if __name__ == "__main__":
    main()
