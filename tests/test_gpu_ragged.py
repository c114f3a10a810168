"""Ragged meshes on the GPU against the oracle: struts of 1, 2, 3, 63, 64,
65, 130 and 200 elements in one mesh (chains longer than a 64-lane
wavefront are split into pieces that meet at a shared node; short chains
share wavefronts with padding lanes), three of them meeting at a joint, and
a single-element mesh.  Internal force at a seeded state, LeapFrog and
MilneDevice runs, both rotation updates."""
import math

import numpy as np
import pytest

from dynlattice_amd import lattice as L
from oracle.ref_cpu import RefModel
from oracle.ref_cpu import milne_opts as ref_opts
from tests.conftest import rel_inf
from tests.helpers import random_state

pytestmark = pytest.mark.gpu

COUNTS = (1, 2, 3, 63, 64, 65, 130, 200)


def ragged():
    """one strut per count, each a straight line of its own direction; the
    struts of 63, 64 and 65 elements share their first node (a joint)"""
    coords, en, es = [np.zeros(3)], [], []
    rng = np.random.default_rng(3)
    for s, m in enumerate(COUNTS):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        start = 0 if m in (63, 64, 65) else None
        if start is None:
            coords.append(np.array([0.3 * s, 0.2, 0.1 * s]))
            start = len(coords) - 1
        a = start
        for k in range(1, m + 1):
            coords.append(coords[start] + d * (0.5 * k / m))
            b = len(coords) - 1
            en.append((a, b))
            es.append(s)
            a = b
    R = len(COUNTS)
    r = 1e-3 * (1.0 + 0.1 * np.arange(R))
    p = L.Problem(coords=np.array(coords), elem_nodes=np.array(en, np.int32),
                  elem_strut=np.array(es, np.int32), n_struts=R,
                  sections=L.Sections.circle(r, 2e9, 0.3, 1200.0)).finalize()
    p.cmask[0, :] = 1
    p.dt = L.stable_dt(p, 0.5)
    return p


def _gpu(p, **kw):
    from dynlattice_amd.gpu import GpuModel
    return GpuModel(p, **kw)


@pytest.mark.parametrize("so3", [True, False])
def test_ragged_chains_match_oracle(so3):
    p = ragged()
    p.so3 = so3
    assert p.n_elems == sum(COUNTS)
    d, v = random_state(p, seed=17, du=2e-6, dth=2e-3, dv=0.05)
    d[0] = 0.0
    v[0] = 0.0
    r, g = RefModel(p), _gpu(p)
    g.set_state(d, v)
    assert rel_inf(g.int_vector(), r.int_vector(d, v)) <= 1e-11
    r.disp, r.velo = d.reshape(-1).copy(), v.reshape(-1).copy()
    r.leapfrog(300)
    g.step(300)
    dg, vg = g.get_state()
    assert rel_inf(dg, r.disp.reshape(-1, 6)) <= 1e-8
    assert rel_inf(vg, r.velo.reshape(-1, 6)) <= 1e-7
    a = r.milne(100, ref_opts(p.dt, 1e-6))
    b = g.milne(100, _milne(p.dt, 1e-6))
    assert np.array_equal(a["accepted"], b["accepted"])
    assert rel_inf(b["dt"], a["dt"]) <= 1e-8
    dg, _ = g.get_state()
    assert rel_inf(dg, r.disp.reshape(-1, 6)) <= 1e-8
    g.close()


def _milne(dt, prec):
    from dynlattice_amd.gpu import milne_opts
    return milne_opts(dt, prec)


def test_single_element():
    p = L.straight_rod(1, 0.2, 2e-3, 2e9, 0.3, 1200.0)
    p.cmask[0, :] = 1
    p.fext[1] = [0.0, 0.0, -0.05, 1e-7, 0.0, 0.0]
    p.dt = L.stable_dt(p, 0.5)
    r, g = RefModel(p), _gpu(p)
    r.leapfrog(500)
    g.step(500)
    dg, _ = g.get_state()
    assert rel_inf(dg, r.disp.reshape(-1, 6)) <= 1e-8
    assert np.abs(dg[1]).max() > 0
    g.close()
