"""The stated configurations at their stated horizons on the device.

C1, the plain cantilever (32 elements, tip force and torque): 20,000
LeapFrog steps against the oracle run alongside, in SO(3) mode: the tip
trajectory, sampled every 100 steps, within 1e-8 (relative, infinity
norm); the internal force at twenty of the trajectory's states within 32 E A
eps, and the clamped node's reaction (z force, y moment) of every step
within 1e-7 of its range: at these ~1e-7 strains the force is E A times a
difference of numbers near 1, so sub-ulp changes of the node positions
move it by ~1e-7 of its size (measured on the oracle alone).

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


@pytest.mark.parametrize("dof", [2, 4])
def test_c1_20000_steps_match_oracle(dof):
    from dynlattice_amd.gpu import GpuModel
    from oracle.ref_cpu import RefModel
    p = L.cantilever_c1()
    r, g = RefModel(p), GpuModel(p)
    clamp = np.array([0], dtype=np.int32)
    g.set_reaction_group(clamp, dof)
    g.set_graph(100)
    n, chunk = 20_000, 100
    hr, hg, tr, tg, fe = [], [], [], [], []
    for k in range(n // chunk):
        hr.append(r.leapfrog(chunk, group=clamp, group_dof=dof)["reaction"])
        g.step(chunk)
        hg.append(g.history(chunk)["reaction"])
        d, v = g.get_state()
        tr.append(r.disp.reshape(-1, 6)[-1].copy())
        tg.append(d[-1].copy())
        if k % 10 == 9:
            # the internal force at the trajectory's own states (the GPU's),
            # both sides from the same inputs
            fe.append(np.abs(g.int_vector() - r.int_vector(d, v)).max())
    g.close()
    hr, hg, tr, tg = map(np.array, (hr, hg, tr, tg))
    ea_eps = p.sections.area[0] * p.sections.young[0] * np.finfo(float).eps
    print("C1:", n, "steps; reaction", rel_inf(hg, hr), "tip", rel_inf(tg[:, :3], tr[:, :3]),
          rel_inf(tg[:, 3:], tr[:, 3:]), "f_int at the states", max(fe) / ea_eps, "x EA eps",
          "tip z range", tr[:, 2].min(), tr[:, 2].max())
    # the tip moves (bends under the load, twists under the torque)
    assert tr[:, 2].min() < -3e-3 and np.abs(tr[:, 3]).max() > 5e-4
    assert rel_inf(tg[:, :3], tr[:, :3]) <= 1e-8
    assert rel_inf(tg[:, 3:], tr[:, 3:]) <= 1e-8
    # The strains here are ~1e-7, so the internal force is E A times a
    # difference of numbers near 1: moving the nodes by 1e-16 m (half an ulp
    # of a 1 m coordinate) changes the oracle's own force by 9e-8 of its
    # 11 N maximum.  Two implementations of the same force therefore agree to
    # a few ulps of the unit-length strain, E A eps = 1.5e-8 N, and not to
    # 1e-11 of the force; the bar is 32 of those.
    assert max(fe) <= 32 * ea_eps
    # The reaction is the clamped element's force, so the histories inherit
    # the same sensitivity to the ulp-level differences the two runs' node
    # positions pick up over the steps: held to 1e-7 of their range.
    assert rel_inf(hg, hr) <= 1e-7


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
