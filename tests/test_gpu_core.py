"""GPU parity of the element force, mass, gyro and LeapFrog step against the oracle.

Tolerances: nodal internal forces within 1e-11
relative (infinity norm over the vector) per evaluation; nodal displacements
within 1e-8 relative after the stated horizon; plastic active sets
bit-exact except points within 1e-12 of the yield surface.
"""
import numpy as np
import pytest

from dynlattice_amd import lattice as L
from oracle.ref_cpu import RefModel
from tests.conftest import rel_inf
from tests.helpers import random_state, small_bcc

pytestmark = pytest.mark.gpu

F_TOL = 1e-11
U_TOL = 1e-8


@pytest.fixture(scope="module")
def gpu():
    from dynlattice_amd import gpu as G
    G.load()
    return G


def _pair(p, **kw):
    from dynlattice_amd.gpu import GpuModel
    return RefModel(p), GpuModel(p, **kw)


@pytest.mark.parametrize("case", ["cantilever", "bcc"])
def test_rest_state_zero_force(gpu, case):
    p = L.cantilever_c1() if case == "cantilever" else small_bcc()
    _, g = _pair(p)
    f = g.int_vector()
    assert np.all(f == 0.0)


@pytest.mark.parametrize("case", ["cantilever", "bcc", "bcc_plastic"])
def test_int_vector_matches_oracle(gpu, case):
    if case == "cantilever":
        p = L.cantilever_c1()
    else:
        p = small_bcc(plastic=(case == "bcc_plastic"))
    r, g = _pair(p)
    d, v = random_state(p, seed=3, du=2e-5 if case != "cantilever" else 1e-3, dth=0.05)
    g.set_state(d, v)
    fr = r.int_vector(d, v)
    fg = g.int_vector()
    print(case, "rel", rel_inf(fg, fr), "max", np.abs(fr).max())
    assert rel_inf(fg, fr) <= F_TOL


def test_int_vector_output_load_case_plastic(gpu):
    p = small_bcc(plastic=True)
    r, g = _pair(p)
    d, v = random_state(p, seed=4, du=4e-5, dth=0.08)
    g.set_state(d, v)
    fr = r.int_vector(d, v, inelastic=False)
    fg = g.int_vector(inelastic=False)
    assert rel_inf(fg, fr) <= F_TOL


def test_gyro_matches_oracle(gpu):
    p = small_bcc()
    r, g = _pair(p)
    d, v = random_state(p, seed=5, du=1e-5, dth=0.3, dv=30.0)
    g.set_state(d, v)
    fr = r.int_vector(d, v, gyro=True)
    fg = g.int_vector(gyro=True)
    assert rel_inf(fg, fr) <= F_TOL


@pytest.mark.parametrize("lumped", [False, True])
def test_mass_blocks_match_oracle(gpu, lumped):
    p = small_bcc()
    r, g = _pair(p)
    d, _ = random_state(p, seed=6, dth=0.4)
    g.set_state(d)
    br = r.mass_blocks(d, lumped=lumped)
    bg = g.mass_blocks(lumped=lumped)
    assert rel_inf(bg, br) <= 1e-12


def test_state_roundtrip(gpu):
    p = small_bcc()
    _, g = _pair(p)
    d, v = random_state(p, seed=8, dth=0.7, dv=1.0)
    g.set_state(d, v)
    d2, v2 = g.get_state()
    assert np.array_equal(v2, v)
    assert np.array_equal(d2[:, :3], d[:, :3])
    assert np.abs(d2[:, 3:] - d[:, 3:]).max() < 1e-14


def test_leapfrog_cantilever(gpu):
    """Plain cantilever (32 elements, tip force and torque), SO(3) mode."""
    p = L.cantilever_c1()
    r, g = _pair(p)
    n = 2000
    r.leapfrog(n)
    g.step(n)
    d, v = g.get_state()
    dr = r.disp.reshape(-1, 6)
    print("tip", d[-1], dr[-1])
    assert rel_inf(d[:, :3], dr[:, :3]) <= U_TOL
    assert rel_inf(d[:, 3:], dr[:, 3:]) <= U_TOL


@pytest.mark.parametrize("so3", [True, False])
def test_leapfrog_bcc_reaction(gpu, so3):
    p = small_bcc()
    p.so3 = so3
    r, g = _pair(p)
    g.set_reaction_group(p.groups["bottom"], 2)
    n = 400
    out = r.leapfrog(n, group=p.groups["bottom"], group_dof=2)
    g.step(n)
    h = g.history(n)
    d, v = g.get_state()
    dr = r.disp.reshape(-1, 6)
    assert rel_inf(d, dr) <= U_TOL
    assert rel_inf(v, r.velo.reshape(-1, 6)) <= 1e-7
    assert rel_inf(h["reaction"], out["reaction"]) <= 1e-9


def test_plastic_single_step_sum_abs(gpu):
    """Sum-abs surface (p = 1): one step from a yielding state.  Over long
    horizons this case is ill-conditioned in the reference algorithm itself
    (a 1e-16 m perturbation of the oracle's own start state moves a 300-step
    crush by 3 %: the cutting plane's path across surface corners depends on
    the last bits), so only short-horizon parity is meaningful for p = 1."""
    p = small_bcc(plastic=True, exponent=1.0)
    r, g = _pair(p)
    d, v = random_state(p, seed=3, du=4e-5, dth=0.08, dv=0.5)
    g.set_state(d, v)
    r.disp[:] = d.reshape(-1)
    r.velo[:] = v.reshape(-1)
    r.leapfrog(1)
    g.step(1)
    dg, vg = g.get_state()
    assert rel_inf(dg, r.disp.reshape(-1, 6)) <= 1e-10
    act_r = r.plastic_state()[3]
    assert act_r.sum() > 10
    assert np.array_equal(g.plastic_active(), act_r)


def test_plastic_step_active_set(gpu):
    """Quadratic interaction surface (p = 2), 300-step crush at -40 m/s."""
    p = small_bcc(plastic=True, exponent=2.0)
    p.velo0[p.groups["top"], 2] = -40.0
    r, g = _pair(p, flags=1)
    g.set_reaction_group(p.groups["bottom"], 2)
    n = 300
    out = r.leapfrog(n, group=p.groups["bottom"], group_dof=2, energy=True)
    o = g.step(n)
    h = g.history(n)
    d, _ = g.get_state()
    assert rel_inf(d, r.disp.reshape(-1, 6)) <= U_TOL
    ep_r, al_r, _, act_r = r.plastic_state()
    ep_g, al_g = g.plastic_state()
    act_g = g.plastic_active()
    assert act_r.sum() > 0, "case must yield"
    assert np.array_equal(act_g, act_r)
    assert rel_inf(ep_g, ep_r) <= 1e-8
    ediss_r = out["energies"][:, 1]
    assert ediss_r[-1] > 0
    assert abs(o.e_diss - ediss_r[-1]) <= 1e-8 * ediss_r[-1]


def test_step_deterministic(gpu):
    from dynlattice_amd.gpu import GpuModel
    p = small_bcc(plastic=True)
    p.velo0[p.groups["top"], 2] = -40.0
    res = []
    for _ in range(2):
        g = GpuModel(p)
        g.step(200)
        res.append(g.get_state()[0])
    assert np.array_equal(res[0], res[1])


@pytest.mark.parametrize("m,so3", [(4, True), (3, True), (4, False)])
def test_shared_end_frames_equal_per_element_frames(gpu, m, so3):
    """k_strut takes an inner element's end-b frame from the next lane (the
    strut's frame at their common node) and reads only chain ends' frames;
    with DL_FRAMES_PER_ELEMENT it reads both frames of every element.  Same
    frames, so the runs agree bit for bit: 4-element struts (identity lane
    packing) and 3-element struts (padded wavefronts), LeapFrog with
    plasticity and contact, and MilneDevice attempts."""
    from dynlattice_amd.gpu import DL_FRAMES_PER_ELEMENT, GpuModel, milne_opts
    p = small_bcc(m=m, plastic=True, contact=True, exponent=2.0)
    p.so3 = so3
    p.velo0[p.groups["top"], 2] = -20.0
    res = []
    for flags in (0, DL_FRAMES_PER_ELEMENT):
        g = GpuModel(p, flags=flags)
        g.step(150)
        g.milne(40, milne_opts(p.dt, 1e-4))
        res.append(g.get_state() + (g.plastic_active(),))
        g.close()
    for a, b in zip(res[0], res[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("m,so3", [(4, True), (3, True), (4, False)])
def test_wavefront_joints_equal_joint_kernel(gpu, m, so3):
    """A BCC cell centre's eight chains sit in one wavefront: k_strut sums
    that joint by shuffles in slot order and updates it; with
    DL_JOINTS_IN_MEMORY every joint goes through the slot buffer and the
    joint kernel.  The same additions in the same order, so LeapFrog with
    plasticity and contact, fixed-step MilneDevice attempts and the action
    path's internal force agree bit for bit.  The MilneDevice error estimate
    is a sum of per-block partials, and the joint's node moves to another
    block, so with an adaptive step only the accept/reject sequence is equal
    and the states agree to rounding."""
    from dynlattice_amd.gpu import DL_JOINTS_IN_MEMORY, GpuModel, milne_opts
    p = small_bcc(n=3, m=m, plastic=True, contact=True, exponent=2.0)
    p.so3 = so3
    p.velo0[p.groups["top"], 2] = -20.0
    dtm = p.dt / 10
    res, ada = [], []
    for flags in (0, DL_JOINTS_IN_MEMORY):
        g = GpuModel(p, flags=flags)
        g.set_reaction_group(p.groups["bottom"], 2)
        g.step(150)
        h = g.history(150)["reaction"]
        g.milne(40, milne_opts(dtm, 1e-4, dt_min=dtm, dt_max=dtm))
        res.append(g.get_state() + (g.plastic_active(), g.int_vector(), h))
        q = g.milne(40, milne_opts(dtm, 1e-4))
        ada.append((g.get_state()[0], q["accepted"]))
        g.close()
    for a, b in zip(res[0], res[1]):
        assert np.array_equal(a, b)
    assert np.array_equal(ada[0][1], ada[1][1])
    assert rel_inf(ada[0][0], ada[1][0]) <= 1e-12


@pytest.mark.parametrize("so3", [True, False])
def test_static_rollup_closed_form(gpu, so3):
    """The reference's static test 1 (two full turns under an end moment
    8 pi): at the closed-form static solution the device's internal force is
    zero at the inner nodes and balances the end moment, as the oracle's."""
    import math

    from dynlattice_amd.gpu import GpuModel
    from tests.test_oracle_element import rollup_problem, rollup_state
    p = rollup_problem(so3)
    d = rollup_state(p)
    g = GpuModel(p)
    g.set_state(d, np.zeros_like(d))
    f = g.int_vector()
    g.close()
    M = 8 * math.pi
    assert np.allclose(f[1:-1], 0.0, atol=1e-12 * M)
    assert np.allclose(f[-1], [0, 0, 0, 0, 0, M], atol=1e-12 * M)
    assert np.allclose(f[0], [0, 0, 0, 0, 0, -M], atol=1e-12 * M)
    assert rel_inf(f, RefModel(p).int_vector(d)) <= 1e-11


def test_commit_energy_matches_oracle(gpu):
    p = small_bcc(plastic=True)
    r, g = _pair(p)
    d, v = random_state(p, seed=9, du=4e-5, dth=0.08)
    g.set_state(d, v)
    r.int_vector(d, v)
    g.int_vector()
    ep_r, ed_r = r.commit()
    ep_g, ed_g = g.commit()
    assert abs(ep_g - ep_r) <= 1e-10 * abs(ep_r)
    assert abs(ed_g - ed_r) <= 1e-10 * max(abs(ed_r), 1e-30)


def test_zero_mass_is_an_error(gpu):
    from dynlattice_amd.gpu import DLError, GpuModel, DL_ERR_ZERO_MASS
    p = L.cantilever_c1()
    p.sections.density[:] = 0.0
    with pytest.raises(DLError) as ei:
        GpuModel(p)
    assert ei.value.code == DL_ERR_ZERO_MASS


def test_plastic_iteration_cap_is_an_error(gpu):
    from dynlattice_amd.gpu import DLError, DL_ERR_PLASTIC_ITER
    p = small_bcc(plastic=True)
    p.plastic.max_iter = 0
    _, g = _pair(p)
    d, v = random_state(p, seed=3, du=4e-5, dth=0.08)
    g.set_state(d, v)
    with pytest.raises(DLError) as ei:
        g.int_vector()
    assert ei.value.code == DL_ERR_PLASTIC_ITER


@pytest.mark.parametrize("contact", [False, True])
def test_groupoutput_reaction_plastic_crush(gpu, contact):
    """The reaction history is GroupOutput's (output load case at the state
    after each step, elastic predictor from the committed plastic state,
    contact forces of that state): equal to the oracle's within 1e-9 on a
    yielding crush; a run split into three calls (whose last rows come from
    the closing evaluation) equals one call bit for bit."""
    from dynlattice_amd.gpu import GpuModel
    p = small_bcc(plastic=True, exponent=2.0, contact=contact)
    p.velo0[p.groups["top"], 2] = -40.0
    K = 300
    out = RefModel(p).leapfrog(K, group=p.groups["top"], group_dof=2)
    runs = []
    for parts in ((K,), (100, 37, 163)):
        g = GpuModel(p)
        g.set_reaction_group(p.groups["top"], 2)
        rows = []
        for n in parts:
            g.step(n)
            rows.append(g.history(n)["reaction"])
        runs.append((np.concatenate(rows), g.plastic_active()))
        g.close()
    h, act = runs[0]
    print("contact", contact, "reaction rel", rel_inf(h, out["reaction"]), "active", act.sum())
    assert act.sum() > 0, "case must yield"
    assert rel_inf(h, out["reaction"]) <= 1e-9
    assert np.array_equal(runs[1][0], h)
