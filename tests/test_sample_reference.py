"""Known answer from the reference itself: its shipped sample MODEL
(main/fullwave/MODEL, 6 parts x 73 subdomains, 43,587 quadratic tets) and
the residual history it logged (calc_log/log_g_HDDM_Full_Wave).  The
fixtures under tests/golden/ were written from those files by
tests/golden/make_sample_fixture.py.

The oracle reproduces the interface size, ||r0||inf and the first ten
residual ratios to the printed 7 digits.  Beyond ~15 iterations the max-norm
COCG history is chaotic under rounding and the histories separate."""
import json
import os

import numpy as np
import pytest

from adventure_fullwave_amd import advio
from oracle import fw_oracle as fo

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_MODEL = "/root/reference/main/fullwave/MODEL"


@pytest.fixture(scope="module")
def sample():
    return advio.load_npz_model(os.path.join(HERE, "sample_model.npz"))


@pytest.fixture(scope="module")
def log():
    return json.load(open(os.path.join(HERE, "sample_log.json")))


def test_interface_size_matches_reference(sample, log):
    assert sample.n_interface == log["dof_on_interface"] == 20264
    assert sample.nsub == 438 and sample.npart == 6


def test_oracle_reproduces_reference_log(sample, log):
    res = fo.hddm_solve(sample, eps=log["hddm_eps"], max_loop=12)
    assert res["stats"]["r0"] == pytest.approx(log["r0_norm"], rel=2e-6)
    ref = np.array(log["resid"][:11])
    got = res["resid"][:11]
    assert np.allclose(got, ref, rtol=2e-6, atol=0)


def test_mutation_missing_midnode_swap_is_caught(sample, log):
    """Without the reference's 8<->9 mid-node swap (make_MESH.c:573-584)
    edges (1,3) and (2,3) trade places and the history no longer matches."""
    bad = advio.load_npz_model(os.path.join(HERE, "sample_model.npz"))
    nop = bad.nop.reshape(-1, 10).copy()
    nop[:, [8, 9]] = nop[:, [9, 8]]
    bad.nop = nop.reshape(-1)
    res = fo.hddm_solve(bad, eps=log["hddm_eps"], max_loop=3)
    assert not np.allclose(res["resid"][1:3], log["resid"][1:3], rtol=1e-3)


@pytest.mark.skipif(not os.path.isdir(REF_MODEL), reason="reference tree not present")
def test_loader_matches_fixture(sample):
    m = advio.load_reference_model(REF_MODEL)
    for k in ("nop", "ib_nfid", "ib_ln", "nmp", "el_off", "op_ibid"):
        assert np.array_equal(getattr(m, k), getattr(sample, k)), k
    assert np.allclose(m.crd, sample.crd) and np.allclose(m.joe, sample.joe)
