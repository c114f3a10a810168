"""MilneDevice on the GPU (dl_milne_step) against the oracle's ref_milne.

* Fixed step (dt_min = dt_max = dt, every attempt accepted): the state after
  every attempt chunk within 1e-8 relative (displacement bar), error
  estimates within 1e-6 relative.  The error estimate is an RMS of
  differences of nearly equal states (u_pre - u_new ~ dt^2 a), so the
  rounding of the two states (the GPU fuses multiply-adds, the oracle does
  not) is amplified by |u| / |dt^2 a| in it.
* Adaptive step: the same accept/reject sequence, step sizes within 1e-8
  relative (they follow the error estimate: measured 3e-10 on the rod,
  1e-11 on the crush), reactions at the start states and contact pair
  counts; final state within 1e-8 relative, the plastic active set of the
  last evaluation identical.
* Device-side control: attempts replayed from a captured graph, and a run
  split over several calls, equal one direct run bit for bit.
* The swinging rod of Lang et al. 2011 on the GPU within the reference's
  tolerance of transient test 2.
"""
import numpy as np
import pytest

from dynlattice_amd import lattice as L
from oracle.ref_cpu import RefModel
from oracle.ref_cpu import milne_opts as ref_opts
from tests.conftest import rel_inf
from tests.helpers import small_bcc

pytestmark = pytest.mark.gpu

U_TOL = 1e-8


def _gpu(p, **kw):
    from dynlattice_amd.gpu import GpuModel
    return GpuModel(p, **kw)


def _opts(dt, prec, **kw):
    from dynlattice_amd.gpu import milne_opts
    return milne_opts(dt, prec, **kw)


def _crush_case(so3=True):
    p = small_bcc(n=2, m=3, plastic=True, contact=True, exponent=2.0)
    p.so3 = so3
    p.velo0[p.groups["top"], 2] = -40.0
    return p


def _state(m):
    d, v = m.get_state()
    return d.reshape(-1, 6), v.reshape(-1, 6)


@pytest.mark.parametrize("case", ["rod", "crush", "crush_additive"])
def test_fixed_step_matches_oracle(case):
    if case == "rod":
        p = L.swinging_rod()
        n, chunk, dt = 600, 200, 2e-5
        group = None
    else:
        p = _crush_case(so3=(case == "crush"))
        # a tenth of the LeapFrog limit: the predictor-corrector pair
        # amplifies the highest modes by sqrt(1 + (w dt)^4 / 4) per step
        n, chunk, dt = 300, 100, p.dt / 10
        group = p.groups["bottom"]
    r = RefModel(p, grid=True)
    g = _gpu(p)
    if group is not None:
        g.set_reaction_group(group, 2)
    ro = ref_opts(dt, 1e-6, dt_min=dt, dt_max=dt)
    go = _opts(dt, 1e-6, dt_min=dt, dt_max=dt)
    ra, rb = [], []
    for k in range(n // chunk):
        a = r.milne(chunk, ro, group=group)
        b = g.milne(chunk, go)
        ra.append(a["reaction"])
        rb.append(b["reaction"])
        assert a["accepted"].all() and b["accepted"].all()
        assert np.all(b["dt"] == dt) and np.all(b["dt_next"] == dt)
        assert rel_inf(b["err"], a["err"]) <= 1e-6, (k, rel_inf(b["err"], a["err"]))
        dg, vg = _state(g)
        dr, vr = r.disp.reshape(-1, 6), r.velo.reshape(-1, 6)
        print(case, k, "disp", rel_inf(dg, dr), "velo", rel_inf(vg, vr))
        assert rel_inf(dg, dr) <= U_TOL
        assert rel_inf(vg, vr) <= 1e-7
        if group is not None:
            assert np.array_equal(b["n_contacts"], a["n_contacts"])
    if group is not None:
        # over the whole history: the first rows, before the stress wave
        # reaches the supports, are rounding noise of cancelling element forces
        print(case, "reaction", rel_inf(np.concatenate(rb), np.concatenate(ra)))
        assert rel_inf(np.concatenate(rb), np.concatenate(ra)) <= 1e-8
    if case != "rod":
        assert np.array_equal(g.plastic_active().astype(bool), r.plastic_state()[3].astype(bool))
        assert rel_inf(g.plastic_state()[0], r.plastic_state()[0]) <= 1e-7
    g.close()


@pytest.mark.parametrize("case", ["rod", "crush"])
def test_adaptive_step_matches_oracle(case):
    if case == "rod":
        p = L.swinging_rod()
        n, prec, group = 3000, 1e-6, None
    else:
        # precision 1e-4: at 1e-3 this crush is ill-conditioned in the
        # reference algorithm itself (the oracle's serial and all-cores builds,
        # which differ only in summation order, drift apart to 6e-4 in dt over
        # 400 attempts); at 1e-4 they agree to 1e-11
        p = _crush_case()
        n, prec, group = 400, 1e-4, p.groups["bottom"]
    r = RefModel(p, grid=True)
    g = _gpu(p)
    if group is not None:
        g.set_reaction_group(group, 2)
    a = r.milne(n, ref_opts(p.dt, prec), group=group)
    b = g.milne(n, _opts(p.dt, prec))
    assert np.array_equal(b["accepted"], a["accepted"])
    if case == "crush":
        assert (~a["accepted"]).sum() >= 5  # the controller rejects and recovers
    assert rel_inf(b["dt"], a["dt"]) <= 1e-8, rel_inf(b["dt"], a["dt"])
    assert rel_inf(b["dt_next"], a["dt_next"]) <= 1e-8
    assert b["n_accepted"] == int(a["accepted"].sum())
    assert abs(b["time"] - a["dt"][a["accepted"]].sum()) <= 1e-8 * b["time"]
    dg, vg = _state(g)
    dr, vr = r.disp.reshape(-1, 6), r.velo.reshape(-1, 6)
    print(case, "disp", rel_inf(dg, dr), "velo", rel_inf(vg, vr), "dt", rel_inf(b["dt"], a["dt"]))
    assert rel_inf(dg, dr) <= U_TOL
    if group is not None:
        assert np.array_equal(b["n_contacts"], a["n_contacts"])
        assert rel_inf(b["reaction"], a["reaction"]) <= 1e-8
        assert np.array_equal(g.plastic_active().astype(bool), r.plastic_state()[3].astype(bool))
    g.close()


def test_graph_replay_and_split_runs_are_bitwise_equal():
    p = _crush_case()
    outs = []
    for mode in ("direct", "graph", "split"):
        g = _gpu(p)
        g.set_reaction_group(p.groups["bottom"], 2)
        if mode == "graph":
            g.set_graph(16)
        o = _opts(p.dt, 1e-3)
        if mode == "split":
            parts = []
            for n in (37, 63, 100):
                q = g.milne(n, o)
                parts.append(q)
                o = _opts(float(q["dt_next"][-1]), 1e-3, dt_min=p.dt / 100, dt_max=p.dt * 100)
            res = {k: np.concatenate([q[k] for q in parts])
                   for k in ("dt", "err", "accepted", "reaction", "n_contacts")}
        else:
            q = g.milne(200, o)
            res = {k: q[k] for k in ("dt", "err", "accepted", "reaction", "n_contacts")}
        d, v = g.get_state()
        outs.append((res, d, v, g.plastic_state()[0]))
        g.close()
    base = outs[0]
    for res, d, v, ep in outs[1:]:
        for k in res:
            assert np.array_equal(res[k], base[0][k]), k
        assert np.array_equal(d, base[1]) and np.array_equal(v, base[2])
        assert np.array_equal(ep, base[3])


def test_swinging_rod_matches_lang_et_al_on_gpu():
    from tests.test_oracle_milne import curve_distance_2d, lang_snapshots
    p = L.swinging_rod()
    g = _gpu(p)
    g.set_graph(50)
    o = _opts(p.dt, 1e-4)
    t, ts, ds = 0.0, [0.0], [np.zeros((p.n_nodes, 6))]
    while t < 1.0:
        q = g.milne(500, o)
        t += q["time"]
        o = _opts(float(q["dt_next"][-1]), 1e-4, dt_min=p.dt / 100, dt_max=p.dt * 100)
        ts.append(t)
        ds.append(_state(g)[0].copy())
    ts = np.array(ts)
    dists = []
    for tt, pts in lang_snapshots().items():
        k = min(np.searchsorted(ts, tt), len(ts) - 1)
        if ts[k] == tt or k == 0:
            d = ds[k]
        else:
            w = (tt - ts[k - 1]) / (ts[k] - ts[k - 1])
            d = (1 - w) * ds[k - 1] + w * ds[k]
        sim = np.column_stack([p.coords[:, 0] + d[:, 0], p.coords[:, 2] + d[:, 2]])
        dists.append(curve_distance_2d(sim, pts))
    print("snapshot distances", np.round(dists, 4))
    assert max(dists) <= 0.05
    g.close()


def test_milne_input_errors():
    from dynlattice_amd.gpu import DL_ERR_INPUT, DLError
    p = L.swinging_rod()
    g = _gpu(p)
    for bad in (_opts(0.0, 1e-6), _opts(1e-6, 0.0), _opts(1e-6, 1e-6, dt_min=1e-5, dt_max=1e-6)):
        with pytest.raises(DLError) as e:
            g.milne(1, bad)
        assert e.value.code == DL_ERR_INPUT
    g.close()


def test_spin_up_case_matches_oracle():
    """the reference's transient test 1 as read from its case files (a
    program-driven rotational acceleration of a clamped hub, additive
    update): 60,000 adaptive attempts (t = 3.95 s) at the default precision,
    identical accept/reject, step sizes and states.  Beyond about 4.2 s the
    reference algorithm is unstable against out-of-plane perturbations of
    this planar motion (the oracle collapses to dt_min from a 1e-15 rad/s
    seed), and the device's rounding differs from the oracle's exact zeros
    out of plane, so the comparison stops before."""
    import os

    from dynlattice_amd import case as CS
    here = os.path.dirname(os.path.abspath(__file__))
    c = CS.load(os.path.join(here, "golden", "transient", "test1.pro"),
                case_name="tests/transient/test1")
    p = c.problem
    r, g = RefModel(p), _gpu(p)
    n, prec = 60000, CS.DEFAULT_PRECISION
    a = r.milne(n, ref_opts(c.dt, prec))
    b = g.milne(n, _opts(c.dt, prec))
    assert np.array_equal(b["accepted"], a["accepted"])
    t_a = a["dt"][a["accepted"]].sum()
    print("t", t_a, "dt", rel_inf(b["dt"], a["dt"]))
    assert t_a > 3.9 and abs(b["time"] - t_a) <= 1e-9 * t_a
    assert rel_inf(b["dt"], a["dt"]) <= 1e-6
    dg, vg = _state(g)
    assert rel_inf(dg, r.disp.reshape(-1, 6)) <= U_TOL
    assert rel_inf(vg, r.velo.reshape(-1, 6)) <= U_TOL
    g.close()
