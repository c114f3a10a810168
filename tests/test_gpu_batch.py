"""Batched sweep and captured-graph stepping.

* A batch of independent instances in one context reproduces every instance
  run alone (bitwise: the kernels never mix instances; per-node dt, per-group
  reaction sums in the same order).
* Replaying steps from a captured hipGraph gives the same bits as direct
  launches, for several replays and a remainder, histories included.
* The first instance of the batch also matches the oracle.
"""
import numpy as np
import pytest

from dynlattice_amd import batch as BT
from dynlattice_amd import lattice as L
from oracle.ref_cpu import RefModel
from tests.conftest import rel_inf

pytestmark = pytest.mark.gpu


def _instances():
    ps = []
    for i, (n, m, r, sy, v) in enumerate([(2, 2, 3e-4, 150e6, -20.0), (3, 3, 5e-4, 250e6, -10.0),
                                          (2, 5, 7e-4, 300e6, -30.0), (1, 4, 4e-4, 200e6, -5.0)]):
        R = 8 * n ** 3
        p = L.bcc_lattice(n, m, 10e-3, np.full(R, r), 70e9, 0.3, 2700.0)
        p.plastic = L.Plastic.resultant_capacities(np.full(R, r), sy, kin_ratio=0.01, exponent=2.0)
        L.crush_bcs(p, v)
        p.dt = L.stable_dt(p, 0.5)
        ps.append(p)
    return ps


def _standalone(batch, b, ps):
    """instance b of the batch on its own, with the batch's (translated) coordinates"""
    p = ps[b]
    lo, hi = batch.node_off[b], batch.node_off[b + 1]
    q = L.Problem(coords=batch.problem.coords[lo:hi].copy(), elem_nodes=p.elem_nodes,
                  elem_strut=p.elem_strut, n_struts=p.n_struts, sections=p.sections,
                  plastic=p.plastic, so3=p.so3, cmask=p.cmask, cval=p.cval, fext=p.fext,
                  disp0=p.disp0, velo0=p.velo0, dt=p.dt, groups=p.groups)
    return q.finalize()


def test_batch_equals_instances_run_alone():
    from dynlattice_amd.gpu import GpuModel
    ps = _instances()
    bt = BT.concat(ps)
    g = BT.make_model(bt)
    n = 150
    g.step(n, out=False)
    d, v = g.get_state()
    h = g.group_history(n)
    for b in range(len(ps)):
        q = _standalone(bt, b, ps)
        s = GpuModel(q)
        s.set_reaction_group(q.groups["bottom"], 2)
        s.step(n, q.dt, out=False)
        ds, vs = s.get_state()
        hs = s.history(n)["reaction"]
        lo, hi = bt.node_off[b], bt.node_off[b + 1]
        assert np.array_equal(d[lo:hi], ds), b
        assert np.array_equal(v[lo:hi], vs), b
        assert np.array_equal(h[:, b], hs), b
        s.close()


def test_batch_instance_matches_oracle():
    ps = _instances()
    bt = BT.concat(ps)
    g = BT.make_model(bt)
    n = 150
    g.step(n, out=False)
    d, _ = g.get_state()
    q = _standalone(bt, 1, ps)
    r = RefModel(q)
    r.leapfrog(n)
    lo, hi = bt.node_off[1], bt.node_off[2]
    assert rel_inf(d[lo:hi], r.disp.reshape(-1, 6)) <= 1e-8


@pytest.mark.parametrize("K", [1, 7, 32])
def test_graph_replay_equals_direct(K):
    from tests.helpers import random_segments
    for p in (_instances()[1], random_segments(seed=31, kS=1e4, kN=1e4)):
        if p.contact is not None:
            c = p.coords.mean(0)
            p.velo0[:, :3] = -(p.coords + p.disp0[:, :3] - c) * (0.8 / (300 * p.dt))
        outs = []
        for graph in (0, K):
            from dynlattice_amd.gpu import GpuModel
            g = GpuModel(p, flags=1)
            if "bottom" in p.groups:
                g.set_reaction_group(p.groups["bottom"], 2)
            g.set_graph(graph)
            n = 3 * K + 5
            o1 = g.step(n)
            o2 = g.step(n)  # a second call replays the cached graph
            d, v = g.get_state()
            outs.append((d, v, g.history(n), o1.e_diss, o2.e_diss))
            g.close()
        (d0, v0, h0, a0, b0), (d1, v1, h1, a1, b1) = outs
        assert np.array_equal(d0, d1) and np.array_equal(v0, v1)
        for k in h0:
            assert np.array_equal(h0[k], h1[k]), k
        assert a0 == a1 and b0 == b1


def test_c4_sweep_of_256_instances():
    """Configuration C4: the 256 seeded instances (n in 4..16 cells, 2..6
    elements per strut, per-instance dt) in one context replaying captured
    graphs, 100 steps: every instance equals its own context bit for bit
    (state and reaction history), and the smallest, median and largest
    instances match the oracle (displacements within 1e-8, bottom reaction
    within 1e-9; in 100 steps the stress wave reaches the bottom face of the
    smallest instances only, so the larger ones check a reaction history
    that is still exactly zero)."""
    from dynlattice_amd.gpu import GpuModel
    probs = [L.c4_instance(i) for i in range(256)]
    bt = BT.concat(probs)
    g = BT.make_model(bt, graph_steps=10)
    n = 100
    g.step(n, out=False)
    d, v = g.get_state()
    h = g.group_history(n)
    g.close()
    sizes = np.array([p.n_elems for p in probs])
    print("instances", len(probs), "elements", sizes.sum(), "range", sizes.min(), sizes.max())
    for b, p in enumerate(probs):
        q = _standalone(bt, b, probs)
        s = GpuModel(q)
        s.set_reaction_group(q.groups["bottom"], 2)
        s.step(n, q.dt, out=False)
        ds, vs = s.get_state()
        hs = s.history(n)["reaction"]
        s.close()
        lo, hi = bt.node_off[b], bt.node_off[b + 1]
        assert np.array_equal(d[lo:hi], ds), b
        assert np.array_equal(v[lo:hi], vs), b
        assert np.array_equal(h[:, b], hs), b
    order = np.argsort(sizes, kind="stable")
    for b in (order[0], order[len(order) // 2], order[-1]):
        q = _standalone(bt, b, probs)
        r = RefModel(q, omp=True)
        o = r.leapfrog(n, group=q.groups["bottom"], group_dof=2)
        lo, hi = bt.node_off[b], bt.node_off[b + 1]
        print("instance", b, q.n_elems, "elements: disp", rel_inf(d[lo:hi], r.disp.reshape(-1, 6)),
              "reaction", rel_inf(h[:, b], o["reaction"]))
        assert rel_inf(d[lo:hi], r.disp.reshape(-1, 6)) <= 1e-8
        assert rel_inf(h[:, b], o["reaction"]) <= 1e-9


def test_autotune_schedules_give_the_same_bits():
    """every schedule the sweep autotuner can choose (instance order, graph
    length, assembly layout) computes the same per-instance results: the
    chosen one only changes the speed"""
    ps = _instances()
    order, K, layout, times = BT.autotune(ps, graph_lengths=(0, 7), steps=20, flags=1)
    assert len(times) == len(BT.ORDERS) * 2 + (len(BT.LAYOUTS) - 1) * 2
    assert layout in BT.LAYOUTS and order in BT.ORDERS
    ref = None
    for order in BT.ORDERS:
        perm = BT.instance_order(ps, order)
        inv = np.argsort(perm)
        for lay, fl in BT.LAYOUTS.items():
            for k in (0, 7):
                bt = BT.ordered_batch(ps, order)
                g = BT.make_model(bt, flags=1 | fl, graph_steps=k)
                g.step(60, out=False)
                d, v = g.get_state()
                h = g.group_history(60)
                g.close()
                # per instance, in the given order
                per = [(d[bt.node_off[b]:bt.node_off[b + 1]], h[:, b]) for b in inv]
                if ref is None:
                    ref = per
                else:
                    for (da, ha), (db, hb) in zip(per, ref):
                        assert np.array_equal(da, db) and np.array_equal(ha, hb), (order, lay, k)
