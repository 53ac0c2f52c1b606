"""Result files (write_outdata_file_HDDM_FW, lib/std/fileio.c:2930-3224).

The writer is checked against the reference's own shipped result files of the
sample MODEL: the layout recorded in tests/golden/sample_result.json (document
order, properties without the document id, payload sizes) always, and the
payload bytes themselves when the reference tree is present.
"""
import json
import os

import numpy as np
import pytest

from adventure_fullwave_amd import advio

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
REF_RESULT = "/root/reference/main/fullwave/MODEL/result"


@pytest.fixture(scope="module")
def sample():
    return advio.load_npz_model(os.path.join(GOLD, "sample_model.npz"))


@pytest.fixture(scope="module")
def ref_values():
    z = np.load(os.path.join(GOLD, "sample_result.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


def fields_for(model, seed=0):
    rng = np.random.default_rng(seed)
    nel = int(model.el_off[-1])
    return [rng.standard_normal((nel, 3)) + 1j * rng.standard_normal((nel, 3)) for _ in range(3)]


def test_writer_layout_matches_reference(tmp_path, sample, ref_values):
    lay = json.load(open(os.path.join(GOLD, "sample_result.json")))
    x = ref_values["x"].astype(np.complex128)
    B, E, J = fields_for(sample)
    paths = advio.write_result(str(tmp_path), sample, x, B, E, J)
    opts = advio.read_adv(paths[0])
    assert [d["props"] for d in opts] == [o["props"] for o in lay["options"]]
    for part, p in enumerate(paths[1:]):
        docs = advio.read_adv(p)
        want = lay["parts"][part]
        assert [d["props"]["label"] for d in docs] == [w["label"] for w in want]
        for d, w in zip(docs, want):
            assert d["props"] == w["props"]
            assert len(d["data"]) == w["payload_bytes"]


def test_writer_roundtrip(tmp_path, sample):
    rng = np.random.default_rng(1)
    n = int(np.sum(sample.nmp))
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    B, E, J = fields_for(sample, 2)
    paths = advio.write_result(str(tmp_path), sample, x, B, E, J)
    xs, Fs = [], {"MagneticField": [], "ElectricField": [], "CurrentDensity": []}
    for p in paths[1:]:
        r = advio.read_result(p)
        xs += r["x"]
        for k in Fs:
            Fs[k] += r[k]
    assert np.array_equal(np.concatenate(xs), x)
    for k, F in zip(("MagneticField", "ElectricField", "CurrentDensity"), (B, E, J)):
        assert np.array_equal(np.concatenate(Fs[k]), F)


@pytest.mark.skipif(not os.path.isdir(REF_RESULT), reason="reference tree not present")
def test_writer_payload_bytes_equal_reference(tmp_path, sample):
    """Decode the reference's result files and write them again: every
    document's payload is byte-identical and its properties equal."""
    xs, F = [], {"MagneticField": [], "ElectricField": [], "CurrentDensity": []}
    for part in range(sample.npart):
        r = advio.read_result(os.path.join(REF_RESULT, f"advhddm_out_{part}.adv"))
        xs += r["x"]
        for k in F:
            F[k] += r[k]
    x = np.concatenate(xs)
    B, E, J = (np.concatenate(F[k]) for k in ("MagneticField", "ElectricField", "CurrentDensity"))
    paths = advio.write_result(str(tmp_path), sample, x, B, E, J)
    for part, p in enumerate(paths[1:]):
        ours = advio.read_adv(p)
        ref = advio.read_adv(os.path.join(REF_RESULT, f"advhddm_out_{part}.adv"))
        assert len(ours) == len(ref)
        for a, b in zip(ours, ref):
            assert a["props"] == b["props"]
            assert a["data"] == b["data"]
