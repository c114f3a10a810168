"""Pins of the oracle's SO(3) maps, rod element and mass against closed forms
and invariants (helpers.cpp:99-197, Line3D.cpp:102-273,
SpecialCosseratRodModel.cpp:536-1031).  Each check is chosen so that a
dropped term, a wrong sign or a transposed operand fails it."""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from dynlattice_amd import lattice as L
from oracle.ref_cpu import RefModel, expvec, expvecp, logmat
from tests.conftest import rel_inf
from tests.helpers import random_state, small_bcc


# ------------------------------------------------------------------ SO(3)
def test_expvec_is_rodrigues():
    rng = np.random.default_rng(0)
    for _ in range(50):
        v = rng.normal(size=3) * rng.uniform(0.01, 3.0)
        R = expvec(v)
        assert np.abs(R - Rotation.from_rotvec(v).as_matrix()).max() < 1e-14
    assert np.array_equal(expvec(np.zeros(3)), np.eye(3))


def test_logmat_inverts_expvec():
    rng = np.random.default_rng(1)
    for _ in range(50):
        v = rng.normal(size=3)
        v *= rng.uniform(1e-6, math.pi - 0.1) / np.linalg.norm(v)
        assert np.abs(logmat(expvec(v)) - v).max() < 1e-12


def test_expvecp_is_the_derivative():
    rng = np.random.default_rng(2)
    for _ in range(20):
        psi = rng.normal(size=3) * 0.7
        dpsi = rng.normal(size=3)
        h = 1e-6
        fd = (expvec(psi + h * dpsi) - expvec(psi - h * dpsi)) / (2 * h)
        assert np.abs(expvecp(psi, dpsi) - fd).max() < 1e-8
    # |psi| < TINY branch: derivative of exp at the identity is skew(psi')
    d = np.array([0.3, -0.2, 0.1])
    S = np.array([[0, -d[2], d[1]], [d[2], 0, -d[0]], [-d[1], d[0], 0]])
    assert np.array_equal(expvecp(np.zeros(3), d), S)


# ------------------------------------------------------------------ element
def _rod(direction=(1.0, 2.0, 2.0), n_elem=1, length=0.3):
    return L.straight_rod(n_elem=n_elem, length=length, radius=0.01, young=210e9,
                          poisson=0.3, density=7850.0, direction=direction)


def test_rest_state_gives_exactly_zero():
    p = small_bcc()
    r = RefModel(p)
    assert np.all(r.int_vector(np.zeros((p.n_nodes, 6))) == 0.0)


def test_frames_map_e3_to_the_tangent():
    p = _rod()
    r = RefModel(p)
    lam, s0 = r.frames()
    t = np.array([1.0, 2.0, 2.0]) / 3.0
    assert np.abs(lam[0, 0] @ np.array([0, 0, 1.0]) - t).max() < 1e-15
    assert np.abs(lam[0, 0] @ lam[0, 0].T - np.eye(3)).max() < 1e-15
    assert np.abs(s0[0] - [0, 0, 1, 0, 0, 0]).max() < 1e-15


def test_pure_stretch():
    p = _rod()
    r = RefModel(p)
    eps = 1e-3
    t = np.array([1.0, 2.0, 2.0]) / 3.0
    d = np.zeros((2, 6))
    d[1, :3] = eps * 0.3 * t
    f = r.int_vector(d)
    EA = p.sections.young[0] * p.sections.area[0]
    assert np.abs(f[1, :3] - EA * eps * t).max() < 1e-9 * EA * eps
    assert np.abs(f[0, :3] + EA * eps * t).max() < 1e-9 * EA * eps
    assert np.abs(f[:, 3:]).max() < 1e-9 * EA * eps * 0.3


def test_pure_twist():
    p = _rod()
    r = RefModel(p)
    k = 0.05
    t = np.array([1.0, 2.0, 2.0]) / 3.0
    d = np.zeros((2, 6))
    d[1, 3:] = k * 0.3 * t
    f = r.int_vector(d)
    GJ = p.sections.shear_modulus[0] * p.sections.J[0]
    assert np.abs(f[1, 3:] - GJ * k * t).max() < 1e-9 * GJ * k
    assert np.abs(f[0, 3:] + GJ * k * t).max() < 1e-9 * GJ * k
    assert np.abs(f[:, :3]).max() < 1e-9 * GJ * k


def test_pure_bending_of_one_element():
    """Node rotations 0 and k L about the material x axis, chord along the
    mid-frame axis with its rest length: strain (0, 0, 0, k, 0, 0), so the end
    moments are -+ EI1 k Lambda_r e1 and the forces vanish."""
    p = _rod(direction=(0.0, 0.0, 1.0), length=0.2)
    r = RefModel(p)
    lam, _ = r.frames()
    LN = lam[0, 0]
    k, Lr = 0.3, 0.2
    ax = LN @ np.array([1.0, 0, 0])
    Rm = Rotation.from_rotvec(0.5 * k * Lr * ax).as_matrix()
    chord = Lr * (Rm @ LN @ np.array([0, 0, 1.0]))
    d = np.zeros((2, 6))
    d[1, :3] = chord - (p.coords[1] - p.coords[0])
    d[1, 3:] = k * Lr * ax
    f = r.int_vector(d)
    EI = p.sections.young[0] * p.sections.I1[0]
    m = EI * k * (Rm @ ax)
    assert np.abs(f[1, 3:] - m).max() < 1e-8 * EI * k
    assert np.abs(f[0, 3:] + m).max() < 1e-8 * EI * k
    assert np.abs(f[:, :3]).max() < 1e-8 * EI * k


def test_pure_shear_of_one_element():
    """Transverse offset delta of node b, no rotations: shear strain delta/L0
    in material x, force kGA delta/L0 along Lambda_N e1."""
    p = _rod(direction=(0.0, 0.0, 1.0), length=0.2)
    r = RefModel(p)
    lam, _ = r.frames()
    e1 = lam[0, 0] @ np.array([1.0, 0, 0])
    delta = 1e-5
    d = np.zeros((2, 6))
    d[1, :3] = delta * e1
    f = r.int_vector(d)
    kGA = p.sections.shear_factor[0] * p.sections.shear_modulus[0] * p.sections.area[0]
    n = kGA * delta / 0.2 * e1
    assert np.abs(f[1, :3] - n).max() < 1e-9 * np.abs(n).max()
    assert np.abs(f[0, :3] + n).max() < 1e-9 * np.abs(n).max()


def rollup_problem(so3):
    """The reference's static test 1 (tests/beam/test1.pro, Simo & Vu-Quoc
    Ex. 7.1): a rod along y, L = 1, five elements, EA = GA = EI = GJ = 2
    (young = shear_modulus = 2, area = area_moment = polar_moment = 1,
    shear_correction = 1), clamped at y = 0, end moment 8 pi about z."""
    p = L.straight_rod(5, 1.0, direction=(0.0, 1.0, 0.0))
    one = np.ones(1)
    p.sections = L.Sections(area=one, I1=one.copy(), I2=one.copy(), J=one.copy(),
                            shear_factor=one.copy(), young=2 * one, shear_modulus=2 * one,
                            density=one.copy())
    p.so3 = so3
    return p


def rollup_state(p):
    """The static solution of 2-node elements under a uniform moment M:
    every element bends by the same angle D = M L0 / EI and carries no force,
    i.e. node k rotated by k D about z, chords of length L0 along the mean
    rotation of their ends.  M = 8 pi: D = 4 pi / 5, two full turns, the tip
    back at the root (tests/beam/test1.py: displacement (0, -1, 0, 0, 0,
    4 pi), response 8 pi)."""
    n, L0, D = p.n_nodes, 0.2, 4 * math.pi / 5
    x = np.zeros((n, 3))
    for k in range(n - 1):
        a = (k + 0.5) * D
        x[k + 1] = x[k] + L0 * np.array([-math.sin(a), math.cos(a), 0.0])
    d = np.zeros((n, 6))
    d[:, :3] = x - p.coords
    d[:, 5] = D * np.arange(n)
    return d


@pytest.mark.parametrize("so3", [True, False])
def test_static_rollup_closed_form(so3):
    p = rollup_problem(so3)
    d = rollup_state(p)
    assert np.allclose(d[-1], [0.0, -1.0, 0.0, 0.0, 0.0, 4 * math.pi], atol=1e-12)
    f = RefModel(p).int_vector(d)
    M = 8 * math.pi
    # the internal force balances the end moment: zero at the inner nodes,
    # M at the tip, the reaction -M at the clamp
    assert np.allclose(f[1:-1], 0.0, atol=1e-12 * M)
    assert np.allclose(f[-1], [0, 0, 0, 0, 0, M], atol=1e-12 * M)
    assert np.allclose(f[0], [0, 0, 0, 0, 0, -M], atol=1e-12 * M)


@pytest.mark.parametrize("seed", [0, 1])
def test_force_and_moment_balance(seed):
    p = small_bcc()
    r = RefModel(p)
    d, _ = random_state(p, seed=seed, du=3e-5, dth=0.1)
    f = r.int_vector(d)
    x = p.coords + d[:, :3]
    scale = np.abs(f).max()
    assert np.abs(f[:, :3].sum(0)).max() < 1e-10 * scale
    mom = np.cross(x, f[:, :3]).sum(0) + f[:, 3:].sum(0)
    assert np.abs(mom).max() < 1e-10 * scale * np.abs(x).max()


def test_rigid_motion_gives_no_force():
    p = small_bcc()
    r = RefModel(p)
    Q = Rotation.from_rotvec([0.4, -0.7, 0.25]).as_matrix()
    tr = np.array([1e-3, -2e-3, 5e-4])
    d = np.zeros((p.n_nodes, 6))
    d[:, :3] = p.coords @ Q.T - p.coords + tr
    d[:, 3:] = Rotation.from_matrix(Q).as_rotvec()
    f = r.int_vector(d)
    EA = (p.sections.young * p.sections.area).max()
    assert np.abs(f).max() < 1e-9 * EA


def test_objectivity():
    p = small_bcc()
    r = RefModel(p)
    d, _ = random_state(p, seed=4, du=3e-5, dth=0.1)
    f = r.int_vector(d)
    Q = Rotation.from_rotvec([-0.3, 0.5, 0.9])
    x = p.coords + d[:, :3]
    d2 = np.zeros_like(d)
    d2[:, :3] = x @ Q.as_matrix().T - p.coords
    d2[:, 3:] = (Q * Rotation.from_rotvec(d[:, 3:])).as_rotvec()
    f2 = r.int_vector(d2)
    fq = np.concatenate([f[:, :3] @ Q.as_matrix().T, f[:, 3:] @ Q.as_matrix().T], 1)
    assert rel_inf(f2, fq) < 1e-9


# ------------------------------------------------------------------ mass
def test_translational_mass_sums_to_rho_A_L():
    p = small_bcc()
    r = RefModel(p)
    B = r.mass_blocks(np.zeros((p.n_nodes, 6)))
    x = p.coords[p.elem_nodes[:, 1]] - p.coords[p.elem_nodes[:, 0]]
    Le = np.sqrt((x * x).sum(1))
    s = p.elem_strut
    total = (p.sections.density[s] * p.sections.area[s] * Le).sum()
    for i in range(3):
        assert abs(B[:, i, i].sum() - total) < 1e-12 * total


def test_rotational_block_closed_form():
    """One element along x: node block = 1/2 Lambda diag(rho I1 L + rho A L^3/12,
    rho I2 L + rho A L^3/12, rho J L) Lambda^T with Lambda e3 = e_x."""
    p = _rod(direction=(1.0, 0.0, 0.0), length=0.25)
    r = RefModel(p)
    B = r.mass_blocks(np.zeros((2, 6)))
    s = p.sections
    Lr = 0.25
    rho, A = s.density[0], s.area[0]
    Jxx = 0.5 * rho * s.J[0] * Lr
    Jyy = 0.5 * (rho * s.I1[0] * Lr + rho * A * Lr ** 3 / 12)
    expect = np.diag([Jxx, Jyy, Jyy])
    for n in range(2):
        assert np.abs(B[n, 3:, 3:] - expect).max() < 1e-14 * Jyy
        assert np.abs(B[n, :3, :3] - 0.5 * rho * A * Lr * np.eye(3)).max() < 1e-14 * rho * A * Lr
        assert np.all(B[n, :3, 3:] == 0.0)


def test_lumped_mass_is_the_diagonal():
    p = small_bcc()
    r = RefModel(p)
    d, _ = random_state(p, seed=5, dth=0.5)
    Bc = r.mass_blocks(d)
    Bl = r.mass_blocks(d, lumped=True)
    for i in range(6):
        assert np.allclose(Bl[:, i, i], Bc[:, i, i], rtol=1e-15, atol=0)
    off = Bl.copy()
    for i in range(6):
        off[:, i, i] = 0
    assert np.all(off == 0.0)


# ------------------------------------------------------------------ dynamics
def test_torque_free_spin_follows_exp():
    """A free rod spinning about its axis (SO(3) mode): the rotation after n
    steps is exp(n dt w) exactly as the update composes exp(dt w) (A14)."""
    p = _rod(direction=(0.0, 0.0, 1.0), n_elem=4, length=0.2)
    w = 30.0
    p.velo0[:, 5] = w
    p.dt = 1e-6
    r = RefModel(p)
    r.leapfrog(200)
    th = r.disp.reshape(-1, 6)[:, 3:]
    assert np.abs(th[:, 2] - 200 * 1e-6 * w).max() < 1e-12
    assert np.abs(th[:, :2]).max() < 1e-14


def test_energy_conserved_after_unloading():
    """tests/transient/test3.py:53 criterion: after the load is removed the
    total energy must not rise by more than 1e-5 relative (here: a plucked
    cantilever released from a static offset, no load)."""
    p = L.cantilever_c1()
    p.fext[:] = 0.0
    x = p.coords[:, 0]
    p.disp0[:, 2] = 1e-3 * (x / 1.0) ** 2
    p.disp0[:, 4] = -2e-3 * x
    r = RefModel(p)
    out = r.leapfrog(3000, energy=True)
    E = out["energies"][:, 0] + out["energies"][:, 2]
    rise = (E.max() - E[0]) / E[0]
    assert rise < 1e-5
    # symplectic update: bounded oscillation, no secular drift
    assert (E.max() - E.min()) / E[0] < 1e-2
