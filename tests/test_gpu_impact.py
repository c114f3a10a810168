"""ImpactBC on the GPU: the impact-driven oscillators' closed forms, and an
elastoplastic BCC crush with contact whose top plate is an impactor (group
weight, initial velocity) against the oracle with LeapFrog and MilneDevice;
graph replay bitwise."""
import math

import numpy as np
import pytest

from oracle.ref_cpu import RefModel
from oracle.ref_cpu import milne_opts as ref_opts
from tests.conftest import rel_inf
from tests.helpers import small_bcc
from tests.test_impact import E_, LEN, _expected, _rods

pytestmark = pytest.mark.gpu


def _gpu(p, **kw):
    from dynlattice_amd.gpu import GpuModel
    return GpuModel(p, **kw)


def test_impact_oscillators_closed_form():
    radii = [1.5e-3, 2.5e-3]
    p = _rods(radii)
    W = 0.08
    p.impact = [(np.array([1, 3]), 0, W)]
    p.velo0[[1, 3], 0] = -2.0
    g = _gpu(p)
    dt, K = 2e-6, 300
    g.step(K, dt)
    d, _ = g.get_state()
    u, _ = _expected(sum(E_ * math.pi * x ** 2 / LEN for x in radii), W, -2.0, dt, K)
    d = d.reshape(-1, 6)
    assert abs(d[1, 0] - u) <= 1e-10 * abs(u) and abs(d[3, 0] - u) <= 1e-10 * abs(u)
    g.close()


def _impactor_crush():
    p = small_bcc(n=2, m=3, plastic=True, contact=True, exponent=2.0)
    top = p.groups["top"]
    p.cmask[top, 2] = 0          # the plate's z motion follows the impact model
    p.velo0[top, 2] = -20.0
    p.impact = [(top, 2, 2e-3)]
    return p


def test_impactor_leapfrog_matches_oracle():
    p = _impactor_crush()
    r, g = RefModel(p, grid=True), _gpu(p)
    g.set_reaction_group(p.groups["bottom"], 2)
    n = 300
    out = r.leapfrog(n, group=p.groups["bottom"])
    g.step(n)
    d, v = g.get_state()
    h = g.history(n)
    top = p.groups["top"]
    print("disp", rel_inf(d, r.disp.reshape(-1, 6)), "plate v", v.reshape(-1, 6)[top[0], 2])
    assert rel_inf(d, r.disp.reshape(-1, 6)) <= 1e-8
    assert rel_inf(v, r.velo.reshape(-1, 6)) <= 1e-7
    assert rel_inf(h["reaction"], out["reaction"]) <= 1e-9
    # the plate is slowed by the lattice
    assert v.reshape(-1, 6)[top, 2].max() > -20.0
    g.close()


def test_impactor_milne_matches_oracle():
    from dynlattice_amd.gpu import milne_opts
    p = _impactor_crush()
    r, g = RefModel(p, grid=True), _gpu(p)
    n, prec = 200, 1e-4
    a = r.milne(n, ref_opts(p.dt, prec))
    b = g.milne(n, milne_opts(p.dt, prec))
    assert np.array_equal(a["accepted"], b["accepted"])
    assert rel_inf(b["dt"], a["dt"]) <= 1e-8
    d, _ = g.get_state()
    print("disp", rel_inf(d, r.disp.reshape(-1, 6)))
    assert rel_inf(d, r.disp.reshape(-1, 6)) <= 1e-8
    g.close()


def test_impactor_graph_replay_is_bitwise():
    p = _impactor_crush()
    outs = []
    for graph in (0, 16):
        g = _gpu(p)
        g.set_graph(graph)
        g.step(96)
        outs.append(g.get_state())
        g.close()
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
