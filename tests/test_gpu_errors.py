"""Error behaviour of the C ABI (run with -m gpu).

The reference reports bad input and numerical trouble through emessage ->
COM_Exit (lib/std/errormessage.h:13): print and exit.  libfwhip returns a
negative FWH_E_* code with a message instead (the binding turns it into
emessage); these tests check that every kind of bad input
and every out-of-order call is refused with the right code before any
kernel runs on it, and that a refused load leaves the context usable."""
import dataclasses
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def fw():
    from adventure_fullwave_amd import fwhip
    return fwhip


def code_of(fw, exc):
    for code, name in fw.FWH_E.items():
        if name in str(exc.value):
            return name
    return None


def load_error(fw, model):
    with pytest.raises(fw.FwhError) as e:
        fw.Solver(model)
    return code_of(fw, e), str(e.value)


def test_bad_descriptors_are_refused(fw, small_model):
    say("[gpu] bad descriptors")
    m = small_model
    # omega <= 0
    assert load_error(fw, dataclasses.replace(m, omega=0.0))[0] == "FWH_E_ARG"
    # interface id out of range
    nf = np.array(m.ib_nfid, copy=True)
    nf[0] = m.nmp[0] + 5
    assert load_error(fw, dataclasses.replace(m, ib_nfid=nf))[0] == "FWH_E_ARG"
    # an element edge id out of range
    nop = np.array(m.nop, copy=True)
    nop[4] = m.nmp[0] + 1
    assert load_error(fw, dataclasses.replace(m, nop=nop))[0] == "FWH_E_ARG"
    # op lists of a part pair differing in length
    if m.npart > 1:
        off = np.array(m.op_off, copy=True)
        off[2:] += 1
        ibid = np.concatenate([np.asarray(m.op_ibid)[:off[1]], [0], np.asarray(m.op_ibid)[off[1]:]]).astype(np.int32)
        assert load_error(fw, dataclasses.replace(m, op_off=off, op_ibid=ibid))[0] == "FWH_E_ARG"
    # subdomains not part-major
    if m.npart > 1:
        sp = np.array(m.sub_part, copy=True)[::-1].copy()
        assert load_error(fw, dataclasses.replace(m, sub_part=sp))[0] == "FWH_E_ARG"


def test_degenerate_element_is_refused(fw, small_model):
    say("[gpu] degenerate element")
    m = small_model
    crd = np.array(m.crd, copy=True).reshape(-1, 3)
    nop = np.asarray(m.nop).reshape(-1, 10)
    v = m.ap_off[0] + nop[0, :4]
    crd[v[3]] = crd[v[0]]            # collapse element 0
    code, msg = load_error(fw, dataclasses.replace(m, crd=crd.reshape(-1)))
    assert code == "FWH_E_MESH" and "degenerate" in msg


def test_calls_out_of_order(fw, small_model):
    say("[gpu] calls out of order")
    ctx = fw.Context(0)
    L = fw.lib()
    import ctypes
    out = np.zeros(4)
    # nothing loaded yet
    rc = L.fwh_get_x(ctx.h, out.ctypes.data)
    assert rc == -4  # FWH_E_STATE
    sol = fw.Solver(small_model, ctx=ctx)
    # a second load into the same context
    with pytest.raises(fw.FwhError, match="FWH_E_STATE"):
        L2 = fw.lib()
        d = small_model.desc()
        fw._check(L2.fwh_load(ctx.h, ctypes.byref(d)), "fwh_load")
    # results before a solve, iterations before the first step
    with pytest.raises(fw.FwhError, match="FWH_E_STATE"):
        sol.x()
    with pytest.raises(fw.FwhError, match="FWH_E_STATE"):
        sol.last_opts = sol.opts()
        sol.iterate(1)
    with pytest.raises(fw.FwhError, match="FWH_E_ARG"):
        sol.select(1)
    # negative max_loop
    with pytest.raises(fw.FwhError, match="FWH_E_ARG"):
        sol.first(max_loop=-1)
    # the context still solves after the refusals
    st = sol.solve(eps=1e-8, max_loop=5000, graph_iters=8)
    assert st["converged"] == 1
    sol.close()


def test_refused_load_leaves_context_usable(fw, small_model):
    say("[gpu] refused load then a good one")
    ctx = fw.Context(0)
    bad = dataclasses.replace(small_model, omega=-1.0)
    with pytest.raises(fw.FwhError):
        fw.Solver(bad, ctx=ctx)
    sol = fw.Solver(small_model, ctx=ctx)
    st = sol.solve(eps=1e-8, max_loop=5000, graph_iters=8)
    assert st["converged"] == 1
    sol.close()
