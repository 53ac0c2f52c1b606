"""End-to-end parity at the benchmarked size (run with -m gpu).

cfg 2 (40^3 cells, 462,520 edge DOFs, 1,100 subdomains in 8 parts,
reference-mode absorber) solved to ||r||_inf / ||r0||_inf < 1e-8 on the GPU
in both Schur forms (explicit S_s on the matrix cores, implicit LDL^T steps)
and by the oracle with its subdomain Schur steps on host threads (the same
sums as its sequential loop, fw_oracle.c).  Bars: iteration count within
2 % (the reference-mode bar of the smaller cases; measured: oracle 1,816,
GPU 1,812 explicit and 1,816-1,821 implicit) and solution rel. L2 X_BAR,
from the measured 1e-8-stop error of this case (below).  The oracle run
takes about 20 s on 16 host threads."""
import os
import sys

import numpy as np
import pytest

from oracle import fw_oracle as fo

pytestmark = pytest.mark.gpu

# the oracle's 1e-8-stopped solution lies 2.3e-7 (rel. L2) from its
# 1e-12-stopped one on this case (1,816 and 2,984 iterations); two correct
# 1e-8 stops may differ by twice that
X_BAR = 5e-7


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cfg2():
    from adventure_fullwave_amd import mesh
    return mesh.make_config(2, pml_mode="sigma")


@pytest.fixture(scope="module")
def cfg2_ref(cfg2):
    threads = max(1, min(16, os.cpu_count() or 1))
    say(f"[gpu] cfg2 oracle solve on {threads} host threads (about 20 s on 16)")
    return fo.hddm_solve(cfg2, eps=1e-8, max_loop=4000, threads=threads)


@pytest.mark.parametrize("schur", [True, False])
def test_cfg2_matches_oracle(cfg2, cfg2_ref, schur):
    from adventure_fullwave_amd import fwhip
    say(f"[gpu] cfg2 solve explicit={schur}")
    sol = fwhip.Solver(cfg2)
    st = sol.solve(eps=1e-8, max_loop=4000, graph_iters=16, schur=schur)
    ref = cfg2_ref
    it_ref = ref["stats"]["iters"]
    say(f"[gpu] cfg2 iters gpu={st['iters']} oracle={it_ref}")
    assert ref["stats"]["converged"] == 1 and st["converged"] == 1
    assert abs(st["iters"] - it_ref) <= 0.02 * it_ref
    r = sol.resid(st["loops"])
    k = min(30, len(r), len(ref["resid"]))
    assert np.allclose(r[:k], ref["resid"][:k], rtol=1e-6)
    x = sol.x()
    rel = np.linalg.norm(x - ref["x_flat"]) / np.linalg.norm(ref["x_flat"])
    say(f"[gpu] cfg2 solution rel. L2 to the oracle {rel:.3e}")
    assert rel <= X_BAR
    sol.close()
