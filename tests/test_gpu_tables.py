"""GET_TABLE observables on the GPU (dl_get_table, dl_get_energies) against
the oracle at seeded random states: element strain / stress tables (spatial
and material) within 1e-11 relative, elastoplastic stress tables after a
crush has yielded (the return map from the committed state, which the table
leaves unchanged), node potential and kinetic energies, masses, and totals."""
import numpy as np
import pytest

from oracle.ref_cpu import RefModel
from tests.conftest import rel_inf
from tests.helpers import random_state, small_bcc

pytestmark = pytest.mark.gpu

TABLES = ("strain", "mat_strain", "stress", "mat_stress")


def _gpu(p, **kw):
    from dynlattice_amd.gpu import GpuModel
    return GpuModel(p, **kw)


@pytest.mark.parametrize("so3", [True, False])
def test_elastic_tables_match_oracle(so3):
    p = small_bcc()
    p.so3 = so3
    r, g = RefModel(p), _gpu(p)
    d, v = random_state(p, seed=11, du=2e-5, dth=0.05, dv=0.4)
    g.set_state(d, v)
    for name in TABLES:
        t_g, t_r = g.table(name), r.elem_table(name, d)
        assert rel_inf(t_g, t_r) <= 1e-11, (name, rel_inf(t_g, t_r))
    ep_r = r.elem_potential_energy(d)
    node_ep = np.zeros(p.n_nodes)
    np.add.at(node_ep, p.elem_nodes[:, 0], 0.5 * ep_r)
    np.add.at(node_ep, p.elem_nodes[:, 1], 0.5 * ep_r)
    assert rel_inf(g.table("potentialEnergy"), node_ep) <= 1e-11
    ek_r = r.kinetic_energy(d, v)
    assert rel_inf(g.table("kineticEnergy"), ek_r) <= 1e-12
    mt = r.mass_blocks(d, lumped=True).reshape(-1, 36)[:, 0]
    assert rel_inf(g.table("mass"), mt) <= 1e-14
    e = g.energies()
    assert abs(e["potential"] - ep_r.sum()) <= 1e-11 * ep_r.sum()
    assert abs(e["kinetic"] - ek_r.sum()) <= 1e-12 * ek_r.sum()
    assert abs(e["mass"] - mt.sum()) <= 1e-13 * mt.sum()
    g.close()


def test_plastic_tables_after_yielding_match_oracle():
    """one step from a yielding state commits plastic strain on both sides;
    the tables are then evaluated at another state (a return map from that
    committed state)"""
    p = small_bcc(plastic=True, exponent=2.0)
    r, g = RefModel(p), _gpu(p)
    d, v = random_state(p, seed=4, du=4e-5, dth=0.08)
    r.disp = d.reshape(-1).copy()
    r.velo = v.reshape(-1).copy()
    r.leapfrog(1)
    g.set_state(d, v)
    g.step(1)
    assert g.plastic_active().sum() > 0
    assert rel_inf(g.plastic_state()[0], r.plastic_state()[0]) <= 1e-11
    committed = [x.copy() for x in g.plastic_state()]
    d2, v2 = random_state(p, seed=12, du=5e-5, dth=0.1)
    g.set_state(d2, v2)
    for name in TABLES:
        t_g, t_r = g.table(name), r.elem_table(name, d2)
        assert rel_inf(t_g, t_r) <= 1e-11, (name, rel_inf(t_g, t_r))
    ep_r = r.elem_potential_energy(d2)
    assert abs(g.energies()["potential"] - ep_r.sum()) <= 1e-11 * ep_r.sum()
    for a, b in zip(committed, g.plastic_state()):
        assert np.array_equal(a, b)
    g.close()
