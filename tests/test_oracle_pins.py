"""Pins of the oracle's contact penalty forces and gyroscopic term by what the
mathematics fixes, and proof that each pin bites: every check runs on the
oracle and on builds of it with one planted mistake (REF_PLANT in
oracle/ref_cpu.c), and must pass on the first and fail on the second.

Contact (RodContactModel.cpp:1031-1208; forces :1147, :1204):
* segment-to-segment, two crossing segments at separation d < R = r_A + r_B
  with interior closest points u_A, u_B: node I of A gets the internal force
  -k_S (d - R) N_I(u_A) n, node J of B +k_S (d - R) N_J(u_B) n, n the unit
  vector from A's closest point to B's (repulsive: -f_int pushes them apart);
* node-to-segment, a node s of B over the interior of A at height d: s gets
  k_N (d - R) n, A's nodes -k_N (d - R) N_I(u) n, n from the foot point to s;
* moment balance: sum_I x_I x f_I = 0 over every pair in every branch (both
  forms put the force on the line through the two contact points, so any
  wrong weight moves the force off it), here over random segment sets.

Gyroscopic term (SpecialCosseratRodModel.cpp:951-967; ExplicitModule.cpp:
411-412): it does no work, so energies cannot see it; what it must do is
keep the spatial angular momentum sum_n (R J0 R^T w + m x x v) of a free rod
tumbling about a non-principal axis constant -- to first order in dt -- and
it is added once per rod model holding the node (a reference quirk: joints
get it k times)."""
import numpy as np
import pytest

from dynlattice_amd import lattice as L
from oracle.ref_cpu import RefModel, build_planted
from tests.helpers import random_segments

KS, KN = 1e6, 2e6


def _two_segments(pa, pb, radius=1e-3, lift=0.02):
    """one-element struts A (pa) and B (pb), B lifted by `lift` in the rest
    configuration (so INIT blacklists nothing); returns the problem and the
    displacement that puts B at pb"""
    coords = np.array([pa[0], pa[1], pb[0], pb[1]], np.float64)
    coords[2:, 2] += lift
    p = L.Problem(coords=coords, elem_nodes=np.array([[0, 1], [2, 3]]),
                  elem_strut=np.array([0, 1]), n_struts=2,
                  sections=L.Sections.circle(np.full(2, 0.5e-3), 2e9, 0.3, 1200.0))
    p.contact = L.Contact(radius=np.full(2, radius), penalty_sts=KS, penalty_nts=KN)
    p.finalize()
    d = np.zeros((4, 6))
    d[2:, 2] = -lift
    return p, d


def check_sts(libpath=None):
    """crossing segments, interior closest points: relative error of the
    contact force against the closed form"""
    Lh, a, dz = 0.01, 0.003, 1.5e-3
    pa = [(-Lh, 0, 0), (Lh, 0, 0)]
    pb = [(a, -0.6 * Lh, dz), (a, 1.4 * Lh, dz)]
    p, d = _two_segments(pa, pb)
    r = RefModel(p, libpath=libpath)
    f, n = r.contact_force(d)
    assert n == 1
    uA, uB = a / Lh, -0.4
    NA = np.array([0.5 * (1 - uA), 0.5 * (1 + uA)])
    NB = np.array([0.5 * (1 - uB), 0.5 * (1 + uB)])
    nrm = np.array([0.0, 0.0, 1.0])
    sc = KS * (dz - 2e-3)
    fe = np.zeros((4, 3))
    fe[0], fe[1] = -sc * NA[0] * nrm, -sc * NA[1] * nrm
    fe[2], fe[3] = sc * NB[0] * nrm, sc * NB[1] * nrm
    # repulsive: the force on A (-f_int) points away from B
    assert -f[0, 2] - f[1, 2] < 0.0
    return np.abs(f[:, :3] - fe).max() / np.abs(fe).max()


def check_nts(libpath=None):
    """the end node of a vertical segment B at height d over the interior of
    A (B's closest point on the infinite lines lies outside B): relative
    error of the node-to-segment force against the closed form"""
    Lh, b, dz = 0.01, -0.25 * 0.01, 1.2e-3
    pa = [(-Lh, 0, 0), (Lh, 0, 0)]
    pb = [(b, 0, dz), (b, 0, dz + 2 * Lh)]
    p, d = _two_segments(pa, pb)
    r = RefModel(p, libpath=libpath)
    f, n = r.contact_force(d)
    assert n == 1
    u = b / Lh
    nrm = np.array([0.0, 0.0, 1.0])
    sc = KN * (dz - 2e-3)
    fe = np.zeros((4, 3))
    fe[2] = sc * nrm
    fe[0] = sc * (-0.5 * (1 - u)) * nrm
    fe[1] = sc * (-0.5 * (1 + u)) * nrm
    assert f[2, 2] < 0.0  # -f_int on the node pushes it up, away from A
    return np.abs(f[:, :3] - fe).max() / np.abs(fe).max()


def check_moment_balance(libpath=None):
    """max over random segment sets of |sum_I x_I x f_I| / sum_I |x_I| |f_I|"""
    worst = 0.0
    for seed in (11, 12, 13, 14):
        p = random_segments(seed=seed, kS=KS, kN=KN)
        rng = np.random.default_rng(seed)
        d = np.zeros((p.n_nodes, 6))
        d[:, :3] = rng.normal(size=(p.n_nodes, 3)) * 1e-4
        f, n = RefModel(p, libpath=libpath).contact_force(d)
        assert n > 10
        x = p.coords + d[:, :3]
        M = np.cross(x, f[:, :3]).sum(0)
        scale = (np.linalg.norm(x, axis=1) * np.linalg.norm(f[:, :3], axis=1)).sum()
        worst = max(worst, float(np.abs(M).max() / scale))
    return worst


def _tumbling_rod(cfl=0.5, m=4, length=0.1, radius=5e-3, W=50.0):
    """a free strut along x, rotating rigidly about (0.3, 0.5, 1) (not a
    principal axis of the rod)"""
    x = np.linspace(0.0, length, m + 1)
    coords = np.stack([x, np.zeros_like(x), np.zeros_like(x)], 1)
    p = L.Problem(coords=coords, elem_nodes=np.stack([np.arange(m), np.arange(1, m + 1)], 1),
                  elem_strut=np.zeros(m, np.int32), n_struts=1,
                  sections=L.Sections.circle(np.array([radius]), 2e9, 0.3, 1200.0))
    p.finalize()
    w = W * np.array([0.3, 0.5, 1.0])
    p.velo0[:, :3] = np.cross(w[None, :], coords - coords.mean(0))
    p.velo0[:, 3:] = w
    p.dt = L.stable_dt(p, cfl)
    return p


def _angular_momentum(r, p):
    B = r.mass_blocks()
    d = r.disp.reshape(-1, 6)
    v = r.velo.reshape(-1, 6)
    x = p.coords + d[:, :3]
    H = np.zeros(3)
    for n in range(p.n_nodes):
        H += B[n, 3:, 3:] @ v[n, 3:] + B[n, 0, 0] * np.cross(x[n], v[n, :3])
    return H


def angular_momentum_drift(libpath=None, cfl=0.5, t_end=0.0387):
    """largest relative change of the total angular momentum of the tumbling
    rod over t_end (about 1.3 rad of rotation)"""
    p = _tumbling_rod(cfl)
    r = RefModel(p, libpath=libpath)
    H0 = _angular_momentum(r, p)
    K = int(round(t_end / p.dt))
    worst = 0.0
    for _ in range(10):
        r.leapfrog(K // 10)
        worst = max(worst, float(np.abs(_angular_momentum(r, p) - H0).max() / np.abs(H0).max()))
    return worst


def check_gyro_multiplicity(libpath=None):
    """two struts sharing a node: the gyroscopic force (GET_INT_VECTOR with
    and without MATRIX2) is k w x (J w) with k the number of rod models
    holding the node and J the node's assembled rotational inertia; relative
    error over all nodes (at the rest configuration, where the elastic force
    is exactly zero and the difference carries no cancellation)"""
    X = np.array([[0, 0, 0], [0.01, 0, 0], [0.02, 0, 0], [0.02, 0.01, 0.005], [0.02, 0.02, 0.01]],
                 np.float64)
    p = L.Problem(coords=X, elem_nodes=np.array([[0, 1], [1, 2], [2, 3], [3, 4]]),
                  elem_strut=np.array([0, 0, 1, 1]), n_struts=2,
                  sections=L.Sections.circle(np.array([1e-3, 1.5e-3]), 2e9, 0.3, 1200.0))
    p.finalize()
    rng = np.random.default_rng(5)
    d = np.zeros((5, 6))
    v = rng.normal(size=(5, 6)) * 30.0
    r = RefModel(p, libpath=libpath)
    g = r.int_vector(d, v, gyro=True) - r.int_vector(d, v, gyro=False)
    B = r.mass_blocks(d)
    k = np.array([1, 1, 2, 1, 1])
    ge = np.array([k[n] * np.cross(v[n, 3:], B[n, 3:, 3:] @ v[n, 3:]) for n in range(5)])
    assert np.abs(g[:, :3]).max() == 0.0
    return np.abs(g[:, 3:] - ge).max() / np.abs(ge).max()


def test_sts_force_closed_form():
    assert check_sts() <= 1e-13


def test_nts_force_closed_form():
    assert check_nts() <= 1e-13


def test_contact_moment_balance():
    assert check_moment_balance() <= 1e-13


def test_gyro_keeps_angular_momentum():
    d1 = angular_momentum_drift(cfl=0.5)
    d2 = angular_momentum_drift(cfl=0.25)
    print("angular momentum drift", d1, d2)
    assert d1 <= 2e-5
    # first order in dt: the drift halves with the step
    assert 1.6 <= d1 / d2 <= 2.4


def test_gyro_once_per_rod_model():
    assert check_gyro_multiplicity() <= 1e-13


CHECKS = {
    1: (check_sts, 1e-13), 2: (check_sts, 1e-13),
    3: (check_nts, 1e-13), 4: (check_nts, 1e-13),
    5: (angular_momentum_drift, 2e-5), 6: (check_gyro_multiplicity, 1e-13),
}
ALSO = {2: (check_moment_balance, 1e-13), 4: (check_moment_balance, 1e-13)}


@pytest.mark.parametrize("plant", sorted(CHECKS))
def test_pins_fail_on_planted_mistakes(plant, tmp_path):
    lib = build_planted(plant, str(tmp_path))
    for fn, tol in [CHECKS[plant]] + ([ALSO[plant]] if plant in ALSO else []):
        try:
            err = fn(lib)
        except AssertionError:
            continue  # a structural check inside the pin already failed
        print("plant", plant, fn.__name__, err)
        assert err > 100 * tol, (plant, fn.__name__, err)
