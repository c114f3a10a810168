"""Two reference quirks the oracle reproduces, pinned by closed forms.

Q4 (RodContactModel.cpp:974-975): for exactly parallel segments the closed-
form closest points are 0/0.  The NaN parameters are "outside" both segments
(containsLocalPoint fails for NaN), so the pair goes to the speculative
branch: every end node that projects into the other segment is a node-to-
segment contact (:660-720).

Q5 (SpecialCosseratRodModel.cpp:651-654): a node whose tangent is -e3 gets
the frame -I ("180 deg turn == point mirroring"), which is no rotation.  With
Lambda = -I the material curvature is unchanged but the spatial moment is
-Lambda sigma, so such a strut resists twist with the opposite sign:
f_rz = -G J phi / L at the twisted node, where a strut along +e3 gives
+G J phi / L."""
import numpy as np

from dynlattice_amd import lattice as L
from oracle.ref_cpu import RefModel

KN = 2e6


def parallel_segments(lift=0.25):
    """A = [0, 1/16] and B = [1/32, 3/32] on the x axis, B at height 2^-10
    (exact binary fractions: the closest-point denominator is exactly 0);
    B lifted for INIT.  Returns the problem and the displacement into place."""
    h = 2.0 ** -10
    X = np.array([[0, 0, 0], [1 / 16, 0, 0], [1 / 32, 0, h + lift], [3 / 32, 0, h + lift]])
    p = L.Problem(coords=X, elem_nodes=np.array([[0, 1], [2, 3]]), elem_strut=np.array([0, 1]),
                  n_struts=2, sections=L.Sections.circle(np.full(2, 0.5e-3), 2e9, 0.3, 1200.0))
    p.contact = L.Contact(radius=np.full(2, 1e-3), penalty_sts=1e6, penalty_nts=KN)
    p.finalize()
    d = np.zeros((4, 6))
    d[2:, 2] = -lift
    return p, d, h


def parallel_closed_form(h):
    """A's end node 1 over B's middle and B's end node 0 over A's middle:
    two node-to-segment contacts at distance h, u = 0 on the master"""
    sc = KN * (h - 2e-3)
    f = np.zeros((4, 3))
    f[0, 2] = -0.5 * sc
    f[1, 2] = -1.5 * sc
    f[2, 2] = 1.5 * sc
    f[3, 2] = 0.5 * sc
    return f


def test_q4_parallel_segments_take_the_end_node_branch():
    p, d, h = parallel_segments()
    r = RefModel(p)
    assert len(r.blacklist()[0]) == 0
    f, n = r.contact_force(d)
    assert n == 1
    fe = parallel_closed_form(h)
    assert np.abs(f[:, :3] - fe).max() <= 1e-13 * np.abs(fe).max()


def twisted_strut(down, phi=1e-3, length=0.05):
    X = np.array([[0, 0, length], [0, 0, 0]] if down else [[0, 0, 0], [0, 0, length]], np.float64)
    p = L.Problem(coords=X, elem_nodes=np.array([[0, 1]]), elem_strut=np.array([0]), n_struts=1,
                  sections=L.Sections.circle(np.array([1e-3]), 2e9, 0.3, 1200.0)).finalize()
    top = 0 if down else 1
    d = np.zeros((2, 6))
    d[top, 5] = phi
    return p, d, top


def test_q5_strut_along_minus_e3_gets_minus_identity():
    for down, sign in ((True, -1.0), (False, 1.0)):
        p, d, top = twisted_strut(down)
        r = RefModel(p)
        lam, _ = r.frames()
        assert np.array_equal(lam.reshape(-1, 3, 3)[0], sign * np.eye(3))
        f = r.int_vector(d)
        GJ = p.sections.shear_modulus[0] * p.sections.J[0]
        fe = sign * GJ * 1e-3 / 0.05
        assert abs(f[top, 5] - fe) <= 1e-9 * abs(fe)
        assert abs(f[1 - top, 5] + fe) <= 1e-9 * abs(fe)
