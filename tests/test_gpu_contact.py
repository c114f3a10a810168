"""GPU parity of rod-rod contact against the oracle: candidate, active and
blacklist pair sets bit-exact; contact forces within 1e-11 relative; a
colliding LeapFrog run within 1e-8 relative with equal per-step active counts."""
import numpy as np
import pytest

from oracle.ref_cpu import RefModel
from tests.conftest import rel_inf
from tests.helpers import random_segments

pytestmark = pytest.mark.gpu


def _sets(a, b):
    return set(zip(a.tolist(), b.tolist()))


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_contact_pair_sets_and_forces(seed):
    from dynlattice_amd.gpu import GpuModel
    p = random_segments(seed=seed)
    r = RefModel(p)
    g = GpuModel(p)
    rng = np.random.default_rng(seed)
    d = np.zeros((p.n_nodes, 6))
    d[:, :3] = rng.normal(size=(p.n_nodes, 3)) * 1e-4
    d[:, 3:] = rng.normal(size=(p.n_nodes, 3)) * 1e-3
    g.set_state(d)
    fr = r.int_vector(d)
    fg = g.int_vector()
    ca, cb = r.contact_candidates(d)
    aa, ab = r.contact_active()
    ga, gb = g.contact_candidates()
    gaa, gab = g.contact_pairs()
    print("seed", seed, "candidates", len(ca), "active", len(aa))
    assert len(aa) > 10
    assert _sets(ga, gb) == _sets(ca, cb)
    assert _sets(gaa, gab) == _sets(aa, ab)
    assert rel_inf(fg, fr) <= 1e-11


def test_contact_blacklist_excludes_initial_touching():
    from dynlattice_amd.gpu import GpuModel
    p = random_segments(seed=21, spread=1.0)  # INIT in the packed state
    r = RefModel(p)
    ba, bb = r.blacklist()
    assert len(ba) > 10
    g = GpuModel(p)
    g.set_state(p.disp0)
    fr = r.int_vector(p.disp0)
    fg = g.int_vector()
    ga, gb = g.contact_pairs()
    aa, ab = r.contact_active()
    assert _sets(ga, gb) == _sets(aa, ab)
    assert not (_sets(ga, gb) & _sets(ba, bb))
    assert rel_inf(fg, fr) <= 1e-11 or np.abs(fr).max() == 0.0


def test_colliding_segments_leapfrog():
    from dynlattice_amd.gpu import GpuModel
    K = 300
    p = random_segments(seed=31, kS=1e4, kN=1e4)
    c = p.coords.mean(0)
    mid = p.coords + p.disp0[:, :3]
    p.velo0[:, :3] = -(mid - c) * (0.8 / (K * p.dt))
    r = RefModel(p)
    g = GpuModel(p)
    out = r.leapfrog(K)
    g.step(K)
    h = g.history(K)
    d, _ = g.get_state()
    assert out["n_contacts"].max() > 5
    assert np.array_equal(h["n_contacts"], out["n_contacts"])
    assert rel_inf(d, r.disp.reshape(-1, 6)) <= 1e-8


def _run_with_skin(p, skin, K, max_pairs=0, flags=0):
    import copy

    from dynlattice_amd.gpu import GpuModel
    q = copy.deepcopy(p)
    q.contact.skin = skin
    g = GpuModel(q, flags=flags, max_pairs=max_pairs)
    if "top" in q.groups:
        g.set_reaction_group(q.groups["top"], 2)
    o = g.step(K)
    h = g.history(K)
    d, v = g.get_state()
    g.close()
    return d, v, h, o.n_candidates


def _cases():
    from dynlattice_amd import lattice as L
    p = random_segments(seed=31, kS=1e4, kN=1e4)
    c = p.coords.mean(0)
    K = 300
    p.velo0[:, :3] = -(p.coords + p.disp0[:, :3] - c) * (0.8 / (K * p.dt))
    return [(p, K), (L.bcc_adjoint_check(), 1600)]


@pytest.mark.parametrize("case", [0, 1])
def test_step_candidate_lists_equal_full_broadphase(case):
    """The step's skin lists find exactly the broadphase's pairs: per-step
    active counts and the last step's candidate count equal those of the full
    cell scan; the lists are sorted by partner id, so any skin (rebuilt at
    different steps) gives the same bits."""
    p, K = _cases()[case]
    r = float(p.contact.radius.max())
    d0, v0, h0, c0 = _run_with_skin(p, -1.0, K)
    cap = 400 * p.n_elems  # room for the widest skin
    runs = [_run_with_skin(p, s, K, cap) for s in (0.05 * r, 0.0, 2.0 * r)]
    assert h0["n_contacts"].max() > 5
    for d, v, h, cand in runs:
        assert np.array_equal(h["n_contacts"], h0["n_contacts"])
        assert cand == c0
        assert rel_inf(d, d0) <= 1e-8
    for d, v, h, cand in runs[1:]:
        assert np.array_equal(d, runs[0][0]) and np.array_equal(v, runs[0][1])
        assert np.array_equal(h["reaction"], runs[0][2]["reaction"])


def test_step_candidate_list_capacity_is_checked():
    from dynlattice_amd.gpu import DL_ERR_CAPACITY, DLError
    p, K = _cases()[0]
    with pytest.raises(DLError) as e:
        _run_with_skin(p, 0.0, 5, max_pairs=4)
    assert e.value.code == DL_ERR_CAPACITY


def test_crowded_joints_pair_sets_and_lists():
    """Joints shared by twelve struts and a strut with five joint nodes (the
    strut filter's lists longer than the scan keeps in registers): candidate
    and active sets equal the oracle's, forces within 1e-11, and the step's
    candidate lists find the full broadphase's pairs every step."""
    from dynlattice_amd.gpu import GpuModel
    from tests.helpers import hub_struts
    p = hub_struts()
    r = RefModel(p)
    g = GpuModel(p)
    rng = np.random.default_rng(41)
    d = np.zeros((p.n_nodes, 6))
    d[:, :3] = rng.normal(size=(p.n_nodes, 3)) * 1e-4
    d[:, 3:] = rng.normal(size=(p.n_nodes, 3)) * 1e-3
    g.set_state(d)
    fr = r.int_vector(d)
    fg = g.int_vector()
    ca, cb = r.contact_candidates(d)
    aa, ab = r.contact_active()
    ga, gb = g.contact_candidates()
    gaa, gab = g.contact_pairs()
    g.close()
    print("candidates", len(ca), "active", len(aa))
    assert len(ca) > 100 and len(aa) > 0
    assert _sets(ga, gb) == _sets(ca, cb)
    assert _sets(gaa, gab) == _sets(aa, ab)
    assert rel_inf(fg, fr) <= 1e-11
    # the struts pulled towards the hub collide while stepping
    K = 200
    c = p.coords.mean(0)
    p.velo0[:, :3] = -(p.coords - c) * (0.3 / (K * p.dt))
    d0, v0, h0, c0 = _run_with_skin(p, -1.0, K)
    d1, v1, h1, c1 = _run_with_skin(p, 0.5 * float(p.contact.radius.max()), K, 400 * p.n_elems)
    out = RefModel(p).leapfrog(K)
    print("active per step (max)", h0["n_contacts"].max(), "candidates", c0)
    assert h0["n_contacts"].max() > 5
    assert np.array_equal(h0["n_contacts"], out["n_contacts"])
    assert np.array_equal(h1["n_contacts"], h0["n_contacts"]) and c1 == c0
    assert rel_inf(d1, d0) <= 1e-8


def test_capacity_overflow_reported_on_every_call():
    """Truncated candidate lists are never used silently: after an overflow
    the lists stay stale, so the next call rebuilds them and raises again."""
    import copy

    from dynlattice_amd.gpu import DL_ERR_CAPACITY, DLError, GpuModel
    p, K = _cases()[0]
    q = copy.deepcopy(p)
    q.contact.skin = 0.0
    g = GpuModel(q, max_pairs=4)
    for _ in range(3):
        with pytest.raises(DLError) as e:
            g.step(5)
        assert e.value.code == DL_ERR_CAPACITY
    g.close()


def test_reinit_under_graph_replay_uses_the_new_blacklist():
    """INIT after a change of contact radius refreshes the blacklist of a
    context that replays captured step graphs: the run equals, bit for bit, a
    fresh context with the new radius and direct launches."""
    import copy

    from dynlattice_amd.gpu import GpuModel
    p = random_segments(seed=21, spread=1.0)  # INIT in the packed state
    p.contact.skin = 0.5 * float(p.contact.radius.max())
    K = 16
    g = GpuModel(p)
    g.set_graph(8)
    g.step(K)
    d1, v1 = g.get_state()
    r2 = 1.3 * p.contact.radius
    g.set_sections(contact_radius=r2)
    g.set_state(d1, v1)
    g.init()
    g.step(K)
    da, va = g.get_state()
    g.close()
    q = copy.deepcopy(p)
    q.contact.radius = r2.copy()
    f = GpuModel(q, init=False)
    f.set_state(d1, v1)
    f.init()
    f.step(K)
    db, vb = f.get_state()
    f.close()
    # the radius change matters: INIT at d1 blacklists more pairs at the
    # larger radius, so a stale blacklist would change the run
    nb = []
    for pr in (p, q):
        ref = RefModel(pr, init=False)
        ref.disp[:] = d1.reshape(-1)
        ref.init()
        nb.append(len(ref.blacklist()[0]))
    print("blacklisted pairs at d1: radius r", nb[0], "radius 1.3 r", nb[1])
    assert nb[1] > nb[0]
    assert np.array_equal(da, db) and np.array_equal(va, vb)
