"""Pins of the oracle's rod-rod contact (RodContactModel.cpp:245-1208).

* The uniform-grid broadphase yields exactly the candidate set of the
  reference's O(R^2) rod/element AABB loops.
* Brute force on tiny inputs: every candidate pair's true minimum segment
  distance (dense sampling) decides the STS branch the same way as the
  closed-form closest points; pairs farther apart than r_A + r_B are never
  active.
* Force balance: contact forces sum to zero (action = reaction)."""
import numpy as np
import pytest

from oracle.ref_cpu import RefModel
from tests.helpers import random_segments


def _state(p, seed, du=1e-4):
    rng = np.random.default_rng(seed)
    d = np.zeros((p.n_nodes, 6))
    d[:, :3] = rng.normal(size=(p.n_nodes, 3)) * du
    return d


@pytest.mark.parametrize("seed", [11, 12, 13, 14])
def test_grid_broadphase_equals_brute_force(seed):
    p = random_segments(seed=seed)
    d = _state(p, seed)
    a, b = RefModel(p).contact_candidates(d)
    ag, bg = RefModel(p, grid=True).contact_candidates(d)
    assert len(a) > 50
    assert set(zip(a.tolist(), b.tolist())) == set(zip(ag.tolist(), bg.tolist()))


def _seg_dist(pa, pb, n=401):
    u = np.linspace(0.0, 1.0, n)
    A = pa[0][None, :] + u[:, None] * (pa[1] - pa[0])[None, :]
    B = pb[0][None, :] + u[:, None] * (pb[1] - pb[0])[None, :]
    D = np.sqrt(((A[:, None, :] - B[None, :, :]) ** 2).sum(-1))
    i, j = np.unravel_index(np.argmin(D), D.shape)
    return D[i, j], 2 * u[i] - 1, 2 * u[j] - 1


@pytest.mark.parametrize("seed", [11, 12])
def test_active_pairs_against_brute_force_distance(seed):
    p = random_segments(seed=seed)
    d = _state(p, seed)
    r = RefModel(p)
    ca, cb = r.contact_candidates(d)
    r.contact_force(d)
    aa, ab = r.contact_active()
    active = set(zip(aa.tolist(), ab.tolist()))
    black = set(zip(*[x.tolist() for x in r.blacklist()]))
    x = p.coords + d[:, :3]
    rad = p.contact.radius
    h = 0.05 / 400 * 3  # sampling resolution of the distance
    n_sts = 0
    for ea, eb in zip(ca.tolist(), cb.tolist()):
        if (ea, eb) in black or (eb, ea) in black:
            continue
        pa = x[p.elem_nodes[ea]]
        pb = x[p.elem_nodes[eb]]
        R = rad[p.elem_strut[ea]] + rad[p.elem_strut[eb]]
        dist, ua, ub = _seg_dist(pa, pb)
        if dist > R + h:
            assert (ea, eb) not in active
        if abs(ua) < 0.9 and abs(ub) < 0.9 and dist < R - h:
            # interior closest points within reach: segment-to-segment contact
            assert (ea, eb) in active
            n_sts += 1
    assert n_sts > 3


def test_contact_forces_balance():
    p = random_segments(seed=13)
    d = _state(p, 13)
    f, n = RefModel(p).contact_force(d)
    assert n > 10
    assert np.abs(f[:, :3].sum(0)).max() < 1e-10 * np.abs(f).max()


def test_blacklist_removes_initially_touching_pairs():
    p = random_segments(seed=21, spread=1.0)
    r = RefModel(p)
    ba, bb = r.blacklist()
    assert len(ba) > 10
    f, n = r.contact_force(p.disp0)
    assert n == 0
    assert np.all(f == 0.0)
