"""The allocation-free element batch (fwh_element_batch_static, run with -m gpu).

Through the C ABI: two calls in a row, the second with fewer elements (the
shared device array keeps the first call's contents), each checked against
fwh_element_batch on the same inputs (the same element kernel: bit for
bit), with and without the optional PML / source arrays; a call beyond the
size limit is refused with FWH_E_LIMIT and leaves the next call correct."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def random_batch(rng, nel):
    from tests.test_oracle_element import random_tet
    xyz = np.stack([random_tet(rng, scale=1 + (k % 5)) for k in range(nel)]).reshape(nel, 12)
    eps = rng.uniform(1, 10, nel) * 8.854e-12
    sig = rng.uniform(0, 0.1, nel)
    pml = np.tile([1.0, 0.0, 1.0, -0.3, 1.0, -0.1], (nel, 1))
    pml[::2] = [1, 0, 1, 0, 1, 0]
    src = (np.arange(nel) % 3 == 0).astype(np.int32)
    joe = rng.standard_normal((nel, 24))
    return xyz, eps, sig, pml, src, joe


def call(L, h, name, nel, xyz, eps, sig, omega, pml, src, joe):
    ae = np.full((nel, 72), np.nan)
    be = np.full((nel, 12), np.nan)
    p = lambda a: None if a is None else np.ascontiguousarray(a).ctypes.data
    rc = getattr(L, name)(h, nel, p(xyz), p(eps), p(sig), ctypes.c_double(omega), p(pml), p(src), p(joe),
                          p(ae), p(be))
    return rc, ae, be


def test_static_element_batch_matches_allocating_one():
    from adventure_fullwave_amd import fwhip
    L = fwhip.lib()
    L.fwh_element_batch_static.argtypes = L.fwh_element_batch.argtypes
    L.fwh_element_batch_static.restype = ctypes.c_int
    ctx = fwhip.Context(0)
    rng = np.random.default_rng(11)
    omega = 2.1e9
    for nel, optional in ((3000, True), (257, True), (100, False)):
        say(f"[gpu] static element batch: {nel} elements, optional arrays {optional}")
        xyz, eps, sig, pml, src, joe = random_batch(rng, nel)
        if not optional:
            pml = src = joe = None
        rc1, ae1, be1 = call(L, ctx.h, "fwh_element_batch", nel, xyz, eps, sig, omega, pml, src, joe)
        rc2, ae2, be2 = call(L, ctx.h, "fwh_element_batch_static", nel, xyz, eps, sig, omega, pml, src, joe)
        assert rc1 == 0 and rc2 == 0, L.fwh_last_error()
        assert np.array_equal(ae1, ae2) and np.array_equal(be1, be2)
    xyz, eps, sig, pml, src, joe = random_batch(rng, 8)
    big = 4096 + 1  # FWH_STATIC_MAX_ELEM + 1
    rc, _, _ = call(L, ctx.h, "fwh_element_batch_static", big, np.zeros((big, 12)), np.ones(big), np.zeros(big),
                    omega, None, None, None)
    assert rc == -5 and b"exceed" in L.fwh_last_error()
    rc1, ae1, be1 = call(L, ctx.h, "fwh_element_batch", 8, xyz, eps, sig, omega, pml, src, joe)
    rc2, ae2, be2 = call(L, ctx.h, "fwh_element_batch_static", 8, xyz, eps, sig, omega, pml, src, joe)
    assert rc1 == 0 and rc2 == 0 and np.array_equal(ae1, ae2) and np.array_equal(be1, be2)
    ctx.close()
