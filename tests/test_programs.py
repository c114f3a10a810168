"""Boundary-data time programs and point masses (CPU): the oracle and the
device evaluator against closed forms.

* Expressions: the oracle's recursive-descent evaluator (oracle/ref_cpu.c)
  and the library's compiler + postfix evaluator (the device code, run on
  the host through dl_program_eval) equal the same expressions written out
  in Python, on the reference's scaleFunc strings (tests/transient/test1.pro,
  test3.pro, test2.pro's initLoad) and on operator-precedence cases.
* Dynamics: a free rigid rod under loads proportional to its lumped masses
  translates rigidly, so the LeapFrog velocity after K steps is a closed-form
  sum of the program's scales at the step END times t_k + dt (the reference
  evaluates boundary data after ExplicitModule::advance), and of the STEP
  scales initLoad + (k+1) loadIncr (NeumannModel) and initDisp + k dispIncr
  (DirichletModel, as a prescribed acceleration).
* MilneDevice moves the program clock only on accepted attempts.
* Point masses: the same rigid translation with point masses loaded in
  proportion, and a rigid spin about the rod axis (no gyroscopic moment about
  a principal axis) with point inertias, both at exactly the prescribed rate.
"""
import math

import numpy as np
import pytest

from dynlattice_amd import lattice as L
from oracle.ref_cpu import RefModel, expr_eval, milne_opts

# (expression, python function of t and i)
EXPRS = [
    ("if (t<15, 6/15 * (1 - cos(2*PI/15 * t)), 0)",
     lambda t, i: 6 / 15 * (1 - math.cos(2 * math.pi / 15 * t)) if t < 15 else 0.0),
    ("if (t<=2, if (t<=1, t*50, (2-t)*50), 0)",
     lambda t, i: (t * 50 if t <= 1 else (2 - t) * 50) if t <= 2 else 0.0),
    ("PI * (5e-3)^2 * 0.1 * 1.1e3 * 9.81", lambda t, i: math.pi * 5e-3 ** 2 * 0.1 * 1.1e3 * 9.81),
    ("-2^2 + 2^-1 - -t", lambda t, i: -(2 ** 2) + 2 ** -1 + t),
    ("2^3^2", lambda t, i: 2.0 ** 9),
    ("t % 0.01 < 0.003", lambda t, i: float(math.fmod(t, 0.01) < 0.003)),
    ("!(t > 1) && t >= 0.5 || i == 3", lambda t, i: float((not t > 1 and t >= 0.5) or i == 3)),
    ("max(t, 2) - min(t, 2) + abs(-t) + sqrt(t) + exp(-t) + log(t)",
     lambda t, i: max(t, 2) - min(t, 2) + abs(-t) + math.sqrt(t) + math.exp(-t) + math.log(t)),
    ("floor(t * 7) + ceil(t) + pow(t, 1.5) + tan(t) + atan(t) + asin(t / 17) + acos(t / 17)",
     lambda t, i: (math.floor(t * 7) + math.ceil(t) + t ** 1.5 + math.tan(t) + math.atan(t)
                   + math.asin(t / 17) + math.acos(t / 17))),
    ("1 + 2 * 3 - 4 / 8 * i", lambda t, i: 1 + 2 * 3 - 4 / 8 * i),
    ("(1 + 2) * (3 - t) != 0", lambda t, i: float((1 + 2) * (3 - t) != 0)),
]


@pytest.mark.parametrize("k", range(len(EXPRS)))
def test_expressions_match_python(k):
    from dynlattice_amd.gpu import program_eval
    e, f = EXPRS[k]
    for t in (0.25, 0.8123, 1.0, 1.7, 3.3, 14.9, 16.0):
        for i in (0.0, 3.0):
            want = f(t, i)
            a = expr_eval(e, t, i)
            b = program_eval(e, t, i)
            assert a == pytest.approx(want, rel=1e-14, abs=1e-15), (e, t)
            assert b == pytest.approx(want, rel=1e-14, abs=1e-15), (e, t)


@pytest.mark.parametrize("bad", ["1 +", "sin(1", "foo(2)", "if(1, 2)", "2 $ 3", ""])
def test_bad_expressions_are_rejected(bad):
    from dynlattice_amd.gpu import DLError, program_eval
    with pytest.raises(ValueError):
        expr_eval(bad)
    with pytest.raises(DLError):
        program_eval(bad)


def _free_rod(n_elem=4):
    p = L.straight_rod(n_elem, 1.0, 5e-3, 5e6, 0.3, 1.1e3)
    mt = RefModel(p).mass_blocks(lumped=True).reshape(-1, 36)[:, 0]
    return p, mt


def test_leapfrog_expression_load_at_step_end_times():
    """f_n = m_n * 50 t along z: v_z after K steps = sum_k dt * 50 (t0 + (k+1) dt)"""
    p, mt = _free_rod()
    p.fext[:, 2] = mt
    L.add_program(p, L.Program("expr", expr="50 * t"), np.arange(p.n_nodes), [2])
    p.t0 = 0.3
    r = RefModel(p)
    dt, K = 1e-4, 40
    r.leapfrog(K, dt)
    want = sum(dt * 50 * (p.t0 + (k + 1) * dt) for k in range(K))
    v = r.velo.reshape(-1, 6)
    assert np.allclose(v[:, 2], want, rtol=1e-12, atol=0)
    assert np.allclose(v[:, :2], 0.0) and np.allclose(v[:, 3:], 0.0, atol=1e-12)
    t, k = r.get_time()
    assert k == K and t == pytest.approx(p.t0 + K * dt, rel=1e-14)


def test_leapfrog_step_programs_neumann_and_dirichlet():
    """Neumann initLoad a + (k+1) loadIncr b on dz; Dirichlet initDisp a' +
    k dispIncr b' as the prescribed acceleration of dx (all nodes)"""
    p, mt = _free_rod()
    p.fext[:, 2] = mt
    p.cmask[:, 0] = 1
    p.cval[:, 0] = 1.0
    L.add_program(p, L.Program("step", a=2.0, b=0.5, c=1.0), np.arange(p.n_nodes), [2])
    L.add_program(p, L.Program("step", a=-1.0, b=0.25, c=0.0), np.arange(p.n_nodes), [0])
    r = RefModel(p)
    dt, K = 1e-4, 25
    r.leapfrog(K, dt)
    vz = dt * sum(2.0 + 0.5 * (k + 1) for k in range(K))
    vx = dt * sum(-1.0 + 0.25 * k for k in range(K))
    v = r.velo.reshape(-1, 6)
    assert np.allclose(v[:, 2], vz, rtol=1e-12, atol=0)
    assert np.allclose(v[:, 0], vx, rtol=1e-12, atol=0)


def test_milne_clock_moves_on_accepted_attempts_only():
    p, mt = _free_rod()
    p.fext[:, 2] = mt
    L.add_program(p, L.Program("expr", expr="50 * t"), np.arange(p.n_nodes), [2])
    r = RefModel(p)
    o = r.milne(40, milne_opts(2e-3, 1e-9, dt_min=1e-7, dt_max=1.0))
    assert (~o["accepted"]).sum() >= 1
    t, k = r.get_time()
    assert k == int(o["accepted"].sum())
    assert t == pytest.approx(o["dt"][o["accepted"]].sum(), rel=1e-13)


def test_point_masses_rigid_translation_and_spin():
    p, mt = _free_rod()
    N = p.n_nodes
    # the rod's own axial inertia (x: a principal axis of every node's inertia)
    Jx = RefModel(p).mass_blocks(lumped=True).reshape(-1, 6, 6)[:, 3, 3]
    pm = np.linspace(0.01, 0.05, N)
    pi = np.linspace(1e-6, 3e-6, N)
    p.point_mass = pm
    p.point_inertia = pi
    # loads in proportion to the rod + point masses and axial inertias
    p.fext[:, 2] = 3.0 * (mt + pm)
    p.fext[:, 3] = 7.0 * (Jx + pi)
    p.velo0[:, 3] = 0.2
    r = RefModel(p)
    dt, K = 1e-4, 30
    r.leapfrog(K, dt)
    v = r.velo.reshape(-1, 6)
    assert np.allclose(v[:, 2], 3.0 * dt * K, rtol=1e-12, atol=0)
    assert np.allclose(v[:, 3], 0.2 + 7.0 * dt * K, rtol=1e-10, atol=0)
