"""updateWhen of the explicit modules on the GPU (dl_set_mass_update) against
the oracle.

ExplicitModule::advance (ExplicitModule.cpp:256-266) re-assembles the mass
only at the steps whose start satisfies updateWhen(t, i), and after INIT;
the solve and the gyroscopic term use the last assembled mass in between.
The device keeps that mass as the packed spatial inertia of the last update.
Checks: a free rod tumbling about a non-principal axis (where a stale
inertia changes the motion) under several schedules, LeapFrog and
MilneDevice (adaptive: the same accept/reject sequence; the clock moves on
accepted attempts only), the SO(3) and the additive update, an
elastoplastic crush with contact; "every step" equals the default bit for
bit; graph replay equals direct launches; set_state makes the next step
assemble."""
import numpy as np
import pytest

from oracle.ref_cpu import RefModel
from oracle.ref_cpu import milne_opts as ref_opts
from tests.conftest import rel_inf
from tests.helpers import small_bcc
from tests.test_oracle_pins import _tumbling_rod

pytestmark = pytest.mark.gpu

CONDS = ("i % 5 < 1", "0", "t < 3e-5 || i % 17 == 3")


def _gpu(p, cond, **kw):
    from dynlattice_amd.gpu import GpuModel
    g = GpuModel(p, **kw)
    g.set_mass_update(cond)
    return g


def _ref(p, cond):
    r = RefModel(p)
    r.set_mass_update(cond)
    return r


@pytest.mark.parametrize("so3", [True, False])
@pytest.mark.parametrize("cond", CONDS)
def test_leapfrog_schedule_matches_oracle(so3, cond):
    p = _tumbling_rod()
    p.so3 = so3
    print(f"tumbling rod, so3={so3}, updateWhen {cond!r}", flush=True)
    g, r = _gpu(p, cond), _ref(p, cond)
    for _ in range(3):
        g.step(100)
        r.leapfrog(100)
        d, v = g.get_state()
        assert rel_inf(d, r.disp.reshape(-1, 6)) <= 1e-10
        assert rel_inf(v, r.velo.reshape(-1, 6)) <= 1e-9
    # the schedule matters
    h = _gpu(p, None)
    h.step(300)
    assert np.abs(h.get_state()[1] - v).max() > 1e-8 * np.abs(v).max()
    g.close()
    h.close()


def test_every_step_is_the_default_and_graphs_replay():
    p = _tumbling_rod()
    a, b = _gpu(p, None), _gpu(p, "1")
    a.step(200)
    b.step(200)
    assert all(np.array_equal(x, y) for x, y in zip(a.get_state(), b.get_state()))
    c, dd = _gpu(p, "i % 5 < 1"), _gpu(p, "i % 5 < 1")
    c.set_graph(7)
    c.step(200)
    dd.step(200)
    assert all(np.array_equal(x, y) for x, y in zip(c.get_state(), dd.get_state()))
    # set_state: the next step assembles at the new state; with a schedule
    # that never updates, a restart equals a fresh context
    e, f = _gpu(p, "0"), _gpu(p, "0")
    e.step(50)
    e.set_state(np.zeros((p.n_nodes, 6)), p.velo0)
    e.step(30)
    f.step(30)
    assert all(np.array_equal(x, y) for x, y in zip(e.get_state(), f.get_state()))
    for m in (a, b, c, dd, e, f):
        m.close()


def test_milne_schedule_matches_oracle():
    from dynlattice_amd.gpu import milne_opts
    p = _tumbling_rod()
    cond = "i % 4 < 1"
    g, r = _gpu(p, cond), _ref(p, cond)
    q = g.milne(300, milne_opts(p.dt, 1e-7, dt_min=p.dt / 8, dt_max=4 * p.dt))
    o = r.milne(300, ref_opts(p.dt, 1e-7, dt_min=p.dt / 8, dt_max=4 * p.dt))
    assert np.array_equal(q["accepted"], o["accepted"]) and 0 < q["accepted"].sum() < 300
    assert np.abs(q["dt"] - o["dt"]).max() <= 1e-8 * o["dt"].max()
    d, v = g.get_state()
    assert rel_inf(d, r.disp.reshape(-1, 6)) <= 1e-8
    assert g.get_time()[1] == r.get_time()[1] == o["accepted"].sum()
    g.close()


def test_crush_with_schedule_matches_oracle():
    p = small_bcc(n=2, m=3, plastic=True, contact=True, exponent=2.0)
    p.velo0[p.groups["top"], 2] = -40.0
    cond = "i % 20 < 1"
    g, r = _gpu(p, cond), _ref(p, cond)
    g.set_reaction_group(p.groups["top"], 2)
    g.step(300)
    o = r.leapfrog(300, group=p.groups["top"], group_dof=2)
    d, v = g.get_state()
    assert rel_inf(d, r.disp.reshape(-1, 6)) <= 1e-8
    assert rel_inf(g.history(300)["reaction"], o["reaction"]) <= 1e-9
    assert np.array_equal(g.plastic_active(), r.plastic_state()[3])
    g.close()


def test_schedule_and_node_time_steps_exclude_each_other():
    """A schedule runs on the context's clock, which per-node time steps
    (batched sweeps) do not define: either order of the two calls is refused."""
    from dynlattice_amd.gpu import DL_ERR_STATE, DLError, GpuModel
    p = _tumbling_rod()
    g = GpuModel(p)
    g.set_mass_update("i % 5 < 1")
    with pytest.raises(DLError) as e:
        g.set_node_dt(np.full(p.n_nodes, p.dt))
    assert e.value.code == DL_ERR_STATE
    g.set_mass_update(None)
    g.set_node_dt(np.full(p.n_nodes, p.dt))
    with pytest.raises(DLError) as e:
        g.set_mass_update("i % 5 < 1")
    assert e.value.code == DL_ERR_STATE
    g.close()
