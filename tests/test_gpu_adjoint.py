"""Adjoint design gradients on the GPU (dl_adjoint_forward / dl_adjoint_backward).

* The checkpointed forward run is the explicit step: same reaction history
  and final state, bit for bit; backward restores that final state and is
  deterministic.
* Section gradients (area, moments, moduli, density, contact radius) equal
  central finite differences of the GPU forward run itself, with and
  without contact, for an objective linear in the reaction history.
* Radius gradients through the torch.autograd.Function equal central finite
  differences of the oracle (fixture tests/golden/adjoint_fd_bcc2.json, made
  by tests/golden/make_adjoint_fd.py) on all 64 struts of the contact case,
  for J = absorbed energy + peak plate force.
* Elastoplastic material: the checkpointed forward is the explicit step
  (plastic state included) and backward restores it; section and
  yield-weight gradients equal central differences of the GPU forward
  (1e-5); radius and yield-stress gradients equal the oracle's central
  differences (fixture tests/golden/adjoint_fd_bcc2_plastic.json, made by
  tests/golden/make_adjoint_fd_plastic.py) within 1e-4, without contact on
  every strut, with contact on the struts whose differences at h and 2h
  agree (the others straddle a jump of the reference's contact branches).
"""
import json
import os

import numpy as np
import pytest

from dynlattice_amd import lattice as L
from dynlattice_amd.gpu import SECTION_FIELDS

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _sections(p):
    return {f: np.array(getattr(p.sections, f), dtype=np.float64) for f in SECTION_FIELDS}


def _no_contact_case():
    R = 8 * 2 ** 3
    r = 0.5e-3 * (1.0 + 0.1 * (2.0 * L.uniform01(9, R) - 1.0))
    p = L.bcc_lattice(2, 2, 10e-3, r, 2e9, 0.3, 1200.0)
    L.crush_bcs(p, -5.0)
    p.dt = L.stable_dt(p, 0.5)
    return p


def _sim(p, nsteps, every, max_pairs=0):
    from dynlattice_amd.adjoint import ImpactSimulation
    return ImpactSimulation(p, nsteps, every, max_pairs=max_pairs)


def test_checkpointed_forward_is_the_explicit_step():
    from dynlattice_amd.gpu import GpuModel
    p = L.bcc_adjoint_check()
    n = 700
    sim = _sim(p, n, 64)
    r = sim.run(_sections(p), p.contact.radius)
    d_fw, v_fw = sim.model.get_state()
    g = GpuModel(p)
    g.set_reaction_group(p.groups["top"], 2)
    g.step(n)
    d, v = g.get_state()
    assert np.array_equal(r[:, 0], g.history(n)["reaction"])
    assert np.array_equal(d, d_fw) and np.array_equal(v, v_fw)
    # backward restores the final state and repeats bit for bit
    w = np.random.default_rng(0).normal(size=r.shape)
    g1 = sim.gradients(w)
    d2, v2 = sim.model.get_state()
    assert np.array_equal(d2, d_fw) and np.array_equal(v2, v_fw)
    g2 = sim.gradients(w)
    for k in g1:
        assert np.array_equal(g1[k], g2[k]), k
    g.close()
    sim.close()


def test_gradients_do_not_depend_on_the_list_skin():
    """The reverse sweep rebuilds the contact candidate lists only at the
    forward run's logged builds, from the taped state of the build in effect.
    The lists fix the pairs, not the build time, so a small skin (builds
    inside the segments, and the reverse crossing back over them) and a wide
    one (a build per segment) give the same gradient bits."""
    import copy
    p0 = L.bcc_adjoint_check()
    r = float(p0.contact.radius.max())
    n, every = 1600, 100
    grads, builds = [], []
    w = np.random.default_rng(4).normal(size=(n, 1))
    for skin in (0.02 * r, 2.0 * r):
        p = copy.deepcopy(p0)
        p.contact.skin = skin
        sim = _sim(p, n, every, max_pairs=400 * p.n_elems)
        sim.run(_sections(p), p.contact.radius)
        builds.append(sim.model.contact_list_stats()[1])
        grads.append(sim.gradients(w))
        sim.close()
    print("list builds per run:", builds)
    nseg = n // every
    assert builds[0] > 2 * nseg  # builds inside the segments
    for k in grads[0]:
        assert np.array_equal(grads[0][k], grads[1][k]), k


def test_backward_requires_a_forward_run():
    from dynlattice_amd.gpu import DL_ERR_STATE, DLError
    p = _no_contact_case()
    sim = _sim(p, 10, 4)
    sim.run(_sections(p))
    sim.model.step(1)  # stepping discards the checkpoints
    with pytest.raises(DLError) as e:
        sim.gradients(np.zeros((10, 1)))
    assert e.value.code == DL_ERR_STATE
    sim.close()


def _fd_check(p, nsteps, every, picks, contact, rtol, max_pairs=0):
    rng = np.random.default_rng(3)
    sim = _sim(p, nsteps, every, max_pairs)
    base = _sections(p)
    cr = p.contact.radius.copy() if contact else None
    r = sim.run(base, cr)
    w = rng.normal(size=r.shape)
    g = sim.gradients(w)

    def J(sec, crad):
        return float((w * sim.run(sec, crad)).sum())

    # finite-difference noise floor: rounding of J over a step of 1e-6 x0
    J_abs = float(np.abs(w * r).sum())

    checked = 0
    for field, s in picks:
        if field == "contact_radius":
            x0 = cr[s]
        else:
            x0 = base[field][s]
        h = 1e-6 * abs(x0)
        vals = []
        for sgn in (1, -1):
            sec = {k: v.copy() for k, v in base.items()}
            c2 = None if cr is None else cr.copy()
            if field == "contact_radius":
                c2[s] += sgn * h
            else:
                sec[field][s] += sgn * h
            vals.append(J(sec, c2))
        fd = (vals[0] - vals[1]) / (2 * h)
        ad = g[field][s]
        assert abs(fd - ad) <= rtol * abs(fd) + 1e-7 * J_abs / abs(x0), (field, s, fd, ad)
        checked += 1
    sim.close()
    return checked


def test_section_gradients_match_gpu_finite_differences():
    p = _no_contact_case()
    picks = [("area", 3), ("I1", 5), ("I2", 5), ("J", 11), ("young", 17), ("shear_modulus", 2),
             ("shear_factor", 40), ("density", 23), ("density", 0), ("area", 63)]
    assert _fd_check(p, 300, 64, picks, contact=False, rtol=1e-5) == len(picks)


def test_section_gradients_match_gpu_finite_differences_with_contact():
    p = L.bcc_adjoint_check()
    picks = [("area", 3), ("I1", 21), ("young", 17), ("density", 40), ("contact_radius", 7),
             ("contact_radius", 33), ("contact_radius", 50)]
    assert _fd_check(p, 1600, 100, picks, contact=True, rtol=1e-4) == len(picks)


def test_radius_gradient_matches_oracle_finite_differences():
    """torch.autograd path: radius -> sections (+ contact radius) -> GPU run
    -> J; d J / d r against the oracle's central differences (h = 1e-6 r)
    on all 64 struts"""
    import torch
    from dynlattice_amd.adjoint import impact_objective, response_from_radius
    fx = json.load(open(os.path.join(HERE, "golden", "adjoint_fd_bcc2.json")))
    p = L.bcc_adjoint_check()
    assert abs(p.dt - fx["dt"]) == 0.0
    sim = _sim(p, fx["nsteps"], 100)
    r = torch.tensor(p.contact.radius, dtype=torch.float64, requires_grad=True)
    reaction = response_from_radius(sim, r, 70e9, 0.3, 2700.0)
    J = impact_objective(reaction, p.dt, fx["v_plate"])
    assert abs(J.item() - fx["objective"]) <= 1e-8 * abs(fx["objective"])
    J.backward()
    ad = r.grad.numpy()[fx["struts"]]
    fd = np.array(fx["dJ_dr"])
    err = np.abs(ad - fd) / np.abs(fd)
    assert err.max() <= 1e-4, (int(np.argmax(err)), float(err.max()))
    sim.close()


def test_gradients_with_init_blacklist_and_crowded_joints():
    """Struts around a 12-strut joint with pairs blacklisted at INIT (kept in
    the candidate lists as marked entries, which the pair VJPs skip like the
    forward pass), pulled towards the joint so that they collide: section and
    contact-radius gradients of the joint's reaction against central
    differences of the GPU forward."""
    from tests.helpers import hub_struts
    p = hub_struts()
    K = 200
    c = p.coords.mean(0)
    p.velo0[:, :3] = -(p.coords - c) * (0.3 / (K * p.dt))
    p.groups["top"] = np.array([0], dtype=np.int32)
    picks = [("area", 3), ("young", 18), ("density", 20), ("contact_radius", 5),
             ("contact_radius", 18)]
    assert _fd_check(p, K, 50, picks, contact=True, rtol=1e-4,
                     max_pairs=400 * p.n_elems) == len(picks)


# ------------------------------------------------------------------ plastic
# The elastoplastic material: the reverse sweep runs through the cutting-plane
# return map (every correction it took), the committed plastic state's
# adjoint travels back step by step, and the reaction rows (GroupOutput's
# output load case: the elastic predictor) are reversed as such.
SIGMA_Y = 250e6


def _plastic_case(contact=True, precision=None):
    p = L.bcc_adjoint_check(sigma_y=SIGMA_Y)
    if not contact:
        p.contact = None
    if precision is not None:
        p.plastic.precision = precision
    return p


def test_plastic_checkpointed_forward_is_the_explicit_step():
    """reaction history, final state and plastic state of the checkpointed
    forward equal the explicit step's bit for bit; backward restores them
    (plastic state and active set included) and repeats bit for bit"""
    from dynlattice_amd.gpu import GpuModel
    p = _plastic_case()
    n = 700
    sim = _sim(p, n, 64)
    r = sim.run(_sections(p), p.contact.radius)
    d_fw, v_fw = sim.model.get_state()
    ep_fw, al_fw = sim.model.plastic_state()
    act_fw = sim.model.plastic_active()
    assert act_fw.mean() > 0.2
    g = GpuModel(p)
    g.set_reaction_group(p.groups["top"], 2)
    g.step(n)
    d, v = g.get_state()
    assert np.array_equal(r[:, 0], g.history(n)["reaction"])
    assert np.array_equal(d, d_fw) and np.array_equal(v, v_fw)
    ep, al = g.plastic_state()
    assert np.array_equal(ep, ep_fw) and np.array_equal(al, al_fw)
    w = np.random.default_rng(0).normal(size=r.shape)
    g1 = sim.gradients(w)
    d2, v2 = sim.model.get_state()
    assert np.array_equal(d2, d_fw) and np.array_equal(v2, v_fw)
    ep2, al2 = sim.model.plastic_state()
    assert np.array_equal(ep2, ep_fw) and np.array_equal(al2, al_fw)
    assert np.array_equal(sim.model.plastic_active(), act_fw)
    g2 = sim.gradients(w)
    for k in g1:
        assert np.array_equal(g1[k], g2[k]), k
    assert np.abs(g1["yield_weight"]).max() > 0
    # a second forward run starts again from the zero plastic state
    r2 = sim.run(_sections(p), p.contact.radius)
    assert np.array_equal(r, r2)
    g.close()
    sim.close()


def test_plastic_section_gradients_match_gpu_finite_differences():
    """section and yield-weight gradients of a yielding crush (no contact)
    against central differences of the GPU forward; the return map's
    precision property at 1e-12, so that its stopping test does not make J
    jump inside the difference"""
    p = _plastic_case(contact=False, precision=1e-12)
    nsteps, every = 800, 64
    rng = np.random.default_rng(4)
    sim = _sim(p, nsteps, every)
    base = _sections(p)
    yw0 = p.plastic.weights.copy()
    r = sim.run(base, None, yw0)
    assert sim.model.plastic_active().mean() > 0.2
    w = rng.normal(size=r.shape)
    g = sim.gradients(w)
    J_abs = float(np.abs(w * r).sum())
    picks = [("area", 3), ("I1", 21), ("J", 11), ("young", 17), ("shear_modulus", 40),
             ("density", 8), ("yield_weight", (5, 2)), ("yield_weight", (21, 3)),
             ("yield_weight", (40, 0)), ("yield_weight", (13, 5))]
    for field, s in picks:
        x0 = yw0[s] if field == "yield_weight" else base[field][s]
        h = 1e-6 * abs(x0)
        vals = []
        for sgn in (1, -1):
            sec = {k: v.copy() for k, v in base.items()}
            yw = yw0.copy()
            if field == "yield_weight":
                yw[s] += sgn * h
            else:
                sec[field][s] += sgn * h
            vals.append(float((w * sim.run(sec, None, yw)).sum()))
        fd = (vals[0] - vals[1]) / (2 * h)
        ad = g[field][s]
        assert abs(fd - ad) <= 1e-5 * abs(fd) + 1e-7 * J_abs / abs(x0), (field, s, fd, ad)
    sim.close()


def _oracle_fd_check(name, contact):
    """radius and yield-stress gradients through the torch path against the
    oracle's central differences, on the variables whose differences at h
    and 2h agree within 1e-4 (the others straddle a jump of J)"""
    import torch
    from dynlattice_amd.adjoint import impact_objective, response_from_radius
    fx = json.load(open(os.path.join(HERE, "golden", "adjoint_fd_bcc2_plastic.json")))
    case = fx[name]
    p = _plastic_case(contact=contact)
    assert abs(p.dt - fx["dt"]) == 0.0
    sim = _sim(p, fx["nsteps"], 100)
    r = torch.tensor(L.bcc_adjoint_check().contact.radius, dtype=torch.float64,
                     requires_grad=True)
    sy = torch.full_like(r, fx["sigma_y"]).requires_grad_(True)
    reaction = response_from_radius(sim, r, 70e9, 0.3, 2700.0, sigma_y=sy)
    J = impact_objective(reaction, p.dt, fx["v_plate"])
    assert abs(J.item() - case["objective"]) <= 1e-8 * abs(case["objective"])
    J.backward()
    out = {}
    x0 = {"dJ_dr": r.detach().numpy(), "dJ_dsy": sy.detach().numpy()}
    for key, grad, idx in (("dJ_dr", r.grad.numpy(), fx["struts"]),
                           ("dJ_dsy", sy.grad.numpy(), fx["sy_struts"])):
        f1, f2 = (np.array(v) for v in case[key])
        ad = grad[idx]
        # rounding floor of a central difference of J over a step h x0
        noise = 1e-13 * abs(case["objective"]) / (fx["h_rel"][0] * x0[key][idx])
        clean = np.abs(f1 - f2) <= 1e-4 * np.abs(f1) + noise
        # (a strut that never yields has no yield-stress derivative: both 0)
        err = np.abs(ad - f1) / (1e-4 * np.abs(f1) + noise)
        assert err[clean].max() <= 1.0, (key, np.array(idx)[clean][np.argmax(err[clean])],
                                         float(err[clean].max()))
        out[key] = int(clean.sum())
    sim.close()
    return out


def test_plastic_gradients_match_oracle_finite_differences():
    n = _oracle_fd_check("nocontact", contact=False)
    assert n["dJ_dr"] >= 50 and n["dJ_dsy"] >= 12, n


def test_plastic_contact_gradients_match_oracle_finite_differences():
    n = _oracle_fd_check("contact", contact=True)
    assert n["dJ_dr"] >= 12, n
