"""The stated configurations at their stated horizons on the device.

C2, elastic BCC impact (10^3 cells, 32,000 elements): 20,000 LeapFrog steps
against the oracle fixture written by tests/golden/make_c2_fixture.py (the
all-cores oracle, same seeded lattice): the bottom reaction history of
every step within 1e-9 and the displacements and velocities of 256 seeded
nodes within 1e-8 (relative, infinity norm).  The C3 crush at a mid-crush
state and the C4 sweep are in test_gpu_fullsize.py and test_gpu_batch.py."""
import os

import numpy as np
import pytest

from dynlattice_amd import lattice as L
from tests.conftest import rel_inf

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_c2_20000_steps_match_oracle_fixture():
    from dynlattice_amd.gpu import GpuModel
    z = np.load(os.path.join(GOLDEN, "c2_bcc10_elastic.npz"))
    p = L.bcc_c2()
    assert p.n_elems == 32_000 and float(z["dt"]) == p.dt
    n = int(z["nsteps"])
    g = GpuModel(p)
    g.set_reaction_group(p.groups["bottom"], 2)
    g.set_graph(32)
    g.step(n)
    h = g.history(n)["reaction"]
    d, v = g.get_state()
    g.close()
    nodes = z["nodes"]
    print("C2:", n, "steps; reaction", rel_inf(h, z["reaction"]), "disp",
          rel_inf(d[nodes], z["disp"]), "velo", rel_inf(v[nodes], z["velo"]))
    assert rel_inf(h, z["reaction"]) <= 1e-9
    assert rel_inf(d[nodes], z["disp"]) <= 1e-8
    assert rel_inf(v[nodes], z["velo"]) <= 1e-8
