"""updateWhen of the explicit modules on the oracle (CPU).

ExplicitModule::advance (ExplicitModule.cpp:256-266) re-assembles the mass
matrix (UPD_MATRIX2 -> assembleM_ at the current rotations) only when its
updateWhen condition of (t, i) holds, before the step's time advance; the
solve and the gyroscopic term use the last assembled mass in between; INIT
assembles it once (ExplicitModule::init ends with updateMass_).  The checks:
the mass in use after a run is the mass assembled (independently, by the
same assembly the step uses) at the state of the last step whose start
satisfied the condition; "always" reproduces the default run bit for bit;
a stale mass changes a tumbling rod's motion; MilneDevice moves the clock
only on accepted attempts."""
import numpy as np

from dynlattice_amd import lattice as L
from oracle.ref_cpu import RefModel
from tests.test_oracle_pins import _tumbling_rod


def _run(p, cond, chunks):
    r = RefModel(p)
    if cond is not None:
        r.set_mass_update(cond)
    states = []
    for k in chunks:
        r.leapfrog(k)
        states.append(r.disp.copy())
    return r, states


def test_always_equals_the_default():
    p = _tumbling_rod()
    a, _ = _run(p, None, [40])
    b, _ = _run(p, "1", [40])
    assert np.array_equal(a.disp, b.disp) and np.array_equal(a.velo, b.velo)


def test_mass_in_use_is_the_mass_of_the_last_scheduled_step():
    p = _tumbling_rod()
    # steps start at i = 0, 1, ..., 12; "i % 5 < 1" updates at the starts of
    # steps i = 0, 5, 10, i.e. at the states after 0, 5 and 10 steps
    r, states = _run(p, "i % 5 < 1", [10, 3])
    B = r.mass_in_use()
    assert np.array_equal(B, r.mass_blocks(states[0]))
    assert not np.allclose(B, r.mass_blocks(states[1]), rtol=1e-6, atol=0)
    # never after INIT: the initial state's mass
    r0, _ = _run(p, "0", [13])
    assert np.array_equal(r0.mass_in_use(), r0.mass_blocks(np.zeros(6 * p.n_nodes)))
    # the schedule matters: a stale inertia turns the tumbling rod differently
    rd, _ = _run(p, None, [13])
    assert np.abs(r.velo - rd.velo).max() > 1e-9 * np.abs(rd.velo).max()


def test_milne_clock_moves_on_accepted_attempts_only():
    from oracle.ref_cpu import milne_opts
    p = _tumbling_rod()
    r = RefModel(p)
    r.set_mass_update("i % 4 < 1")
    # a precision no attempt meets at dt_min < dt: rejections, then dt_min,
    # which is always accepted; each rejection keeps the clock and the state
    o = r.milne(12, milne_opts(p.dt, 1e-30, dt_min=p.dt / 4, dt_max=p.dt))
    acc = o["accepted"].astype(bool)
    assert 0 < acc.sum() < 12
    t, k = r.get_time()
    assert k == acc.sum() and abs(t - o["dt"][acc].sum()) <= 1e-15 * t
