"""GPU parity on the reference's edge cases and the contact closed forms.

* Q4, exactly parallel segments (RodContactModel.cpp:974-975): the 0/0
  closest points send the pair to the end-node branch; pair set and forces
  equal the oracle's and the closed form.
* Q5, struts along -e3 get the frame -I (SpecialCosseratRodModel.cpp:651-
  654): the device keeps identity quaternions plus a flag (Lambda = -R);
  internal force, tables and a LeapFrog run equal the oracle's; an element
  between a -I frame and a rotation frame is refused.
* The segment-to-segment and node-to-segment closed forms that pin the
  oracle's contact forces, evaluated on the device."""
import numpy as np
import pytest

from dynlattice_amd import lattice as L
from oracle.ref_cpu import RefModel
from tests.conftest import rel_inf
from tests.helpers import random_state
from tests.test_oracle_pins import _two_segments
from tests.test_oracle_quirks import parallel_closed_form, parallel_segments, twisted_strut

pytestmark = pytest.mark.gpu


def test_q4_parallel_segments():
    from dynlattice_amd.gpu import GpuModel
    p, d, h = parallel_segments()
    g = GpuModel(p)
    g.set_state(d)
    f = g.int_vector()
    a, b = g.contact_pairs()
    g.close()
    assert len(a) == 1
    fe = parallel_closed_form(h)
    fr, _ = RefModel(p).contact_force(d)
    assert np.abs(f[:, :3] - fe).max() <= 1e-12 * np.abs(fe).max()
    assert rel_inf(f, fr) <= 1e-11


def test_contact_closed_forms_on_device():
    from dynlattice_amd.gpu import GpuModel
    KS, KN = 1e6, 2e6
    cases = []
    Lh = 0.01
    # segment-to-segment: interior closest points u_A = 0.3, u_B = -0.4
    cases.append(([(-Lh, 0, 0), (Lh, 0, 0)], [(0.003, -0.6 * Lh, 1.5e-3), (0.003, 1.4 * Lh, 1.5e-3)]))
    # node-to-segment: the end node of a vertical B over A at u = -0.25
    cases.append(([(-Lh, 0, 0), (Lh, 0, 0)], [(-0.25 * Lh, 0, 1.2e-3), (-0.25 * Lh, 0, 1.2e-3 + 2 * Lh)]))
    for k, (pa, pb) in enumerate(cases):
        p, d = _two_segments(pa, pb)
        g = GpuModel(p)
        g.set_state(d)
        f = g.int_vector()
        g.close()
        fe = np.zeros((4, 3))
        nz = np.array([0.0, 0.0, 1.0])
        if k == 0:
            sc = KS * (1.5e-3 - 2e-3)
            NA, NB = (0.35, 0.65), (0.7, 0.3)
            fe[0], fe[1] = -sc * NA[0] * nz, -sc * NA[1] * nz
            fe[2], fe[3] = sc * NB[0] * nz, sc * NB[1] * nz
        else:
            sc = KN * (1.2e-3 - 2e-3)
            u = -0.25
            fe[2] = sc * nz
            fe[0], fe[1] = sc * (-0.5 * (1 - u)) * nz, sc * (-0.5 * (1 + u)) * nz
        assert np.abs(f[:, :3] - fe).max() <= 1e-12 * np.abs(fe).max(), k


def _lattice_with_down_struts():
    """a 2x2x2 BCC lattice plus four vertical struts hanging from the top
    face's corners, meshed from top to bottom (tangent -e3)"""
    R0 = 64
    r = 0.5e-3 * (1.0 + 0.05 * L.normal(3, R0))
    p = L.bcc_lattice(2, 3, 10e-3, r, 2e9, 0.3, 1200.0)
    X = list(p.coords)
    en = [tuple(e) for e in p.elem_nodes]
    es = list(p.elem_strut)
    top = p.groups["top"]
    s = R0
    for n0 in top[:4]:
        prev = int(n0)
        for k in range(1, 4):
            X.append(p.coords[n0] + np.array([0.0, 0.0, -4e-3 * k]))
            cur = len(X) - 1
            en.append((prev, cur))
            es.append(s)
            prev = cur
        s += 1
    q = L.Problem(coords=np.array(X), elem_nodes=np.array(en), elem_strut=np.array(es), n_struts=s,
                  sections=L.Sections.circle(np.r_[r, np.full(s - R0, 0.5e-3)], 2e9, 0.3, 1200.0))
    q.groups = dict(p.groups)
    q.finalize()
    L.crush_bcs(q, -5.0)
    q.dt = L.stable_dt(q, 0.5)
    return q


@pytest.mark.parametrize("so3", [True, False])
def test_q5_minus_e3_struts_match_oracle(so3):
    from dynlattice_amd.gpu import GpuModel
    p = _lattice_with_down_struts()
    p.so3 = so3
    r = RefModel(p)
    lam, _ = r.frames()
    # every end frame of the hanging struts' elements is -I
    assert (np.abs(lam.reshape(-1, 3, 3) + np.eye(3)).max(axis=(1, 2)) == 0).sum() >= 8
    g = GpuModel(p)
    d, v = random_state(p, seed=9, du=2e-5, dth=0.05)
    g.set_state(d, v)
    fr = r.int_vector(d, v)
    fg = g.int_vector()
    assert rel_inf(fg, fr) <= 1e-11
    for name, kind in (("strain", 0), ("mat_strain", 1), ("stress", 2), ("mat_stress", 3)):
        tr = r.elem_table(name, d)
        assert rel_inf(g.table(name), tr) <= 1e-11, name
    # a short LeapFrog run from the same state: the -I struts resist twist
    # and bending with the wrong sign, so their highest modes grow by about
    # e per step in both the oracle and the device, and only a few steps
    # stay well conditioned
    g2 = GpuModel(p)
    g2.set_state(d, v)
    r2 = RefModel(p)
    r2.disp[:] = d.reshape(-1)
    r2.velo[:] = v.reshape(-1)
    r2.leapfrog(8)
    g2.step(8)
    dg, _ = g2.get_state()
    assert rel_inf(dg, r2.disp.reshape(-1, 6)) <= 1e-8
    g.close()
    g2.close()


def test_q5_twist_sign_on_device():
    from dynlattice_amd.gpu import GpuModel
    for down, sign in ((True, -1.0), (False, 1.0)):
        p, d, top = twisted_strut(down)
        g = GpuModel(p)
        g.set_state(d)
        f = g.int_vector()
        g.close()
        GJ = p.sections.shear_modulus[0] * p.sections.J[0]
        fe = sign * GJ * 1e-3 / 0.05
        assert abs(f[top, 5] - fe) <= 1e-9 * abs(fe)


def test_q5_mixed_frames_refused():
    from dynlattice_amd.gpu import DL_ERR_INPUT, DLError, GpuModel
    X = np.array([[0, 0, 0], [0, 0, -0.01], [0.01, 0, -0.02]], np.float64)
    p = L.Problem(coords=X, elem_nodes=np.array([[0, 1], [1, 2]]), elem_strut=np.array([0, 0]),
                  n_struts=1, sections=L.Sections.circle(np.array([1e-3]), 2e9, 0.3, 1200.0))
    p.finalize()
    with pytest.raises(DLError) as e:
        GpuModel(p)
    assert e.value.code == DL_ERR_INPUT


def test_adjoint_refuses_minus_identity_frames():
    from dynlattice_amd.gpu import DL_ERR_STATE, DLError, GpuModel
    p, d, top = twisted_strut(True)
    g = GpuModel(p)
    g.set_reaction_group(np.array([1], np.int32), 2)
    with pytest.raises(DLError) as e:
        g.adjoint_forward(4)
    assert e.value.code == DL_ERR_STATE
    g.close()
