"""The reference's transient tests end to end on the GPU.

Test 2: the case files
(tests/golden/transient/test2.pro + includes + test2.geo) are read, the
swinging rod runs with MilneDevice to t = 1 s as the case prescribes
(additive rotation update, sampling "t % 1e-4 < deltaTime"), disp.gz is
written in the CSVOutput layout, and the reference's validator
(tests/transient/test2.py, restated below on the same pandas calls) accepts
it: every snapshot of Lang et al. 2011 within curve distance 0.05.

The case leaves the precision at JIVE's library default, which is not in
the reference's sources; the run here uses 1e-4 (the oracle shows the
comparison unchanged from 1e-4 to 1e-6).

Test 1 (spin-up, Simo & Vu-Quoc Ex. 5.1) and test 3 (out-of-plane elbow,
Ex. 5.2) run from their case files as they are: second-order meshes as
2-node elements on the same nodes, Sample outputs (disp.gz rows of the
data sets, energy.csv with its header), checked with their validators'
criteria restated (tests/transient/test1.py, test3.py)."""
import os

import numpy as np
import pytest

from tests.test_oracle_milne import curve_distance_2d

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _validate(disp_path, ref_path):
    """tests/transient/test2.py's reading and comparison"""
    import pandas as pd
    data = pd.read_csv(disp_path, index_col=["time", "label"]).xs("disp", level="label")
    data.columns = pd.MultiIndex.from_tuples(
        [(n[:n.find("[")], int(n[n.find("[") + 1:n.find("]")])) for n in data.columns],
        names=["dof", "node"])
    dx = data.xs("dx", axis=1, level="dof").copy()
    dz = data.xs("dz", axis=1, level="dof").copy()
    end_dx, end_dz = dx[1].copy(), dz[1].copy()
    nmax = data.columns.unique("node").max()
    for i in range(1, nmax):
        dx[i] = dx[i + 1]
        dz[i] = dz[i + 1]
    dx[nmax] = end_dx
    dz[nmax] = end_dz
    lines = open(ref_path).read().splitlines()
    times = lines[0].split(",")
    rows = [ln.split(",") for ln in lines[2:]]
    dists = []
    n_nodes = len(dx.columns)
    for j in range(0, len(times), 2):
        t_val = float(times[j][1:])
        pts = np.array([(float(r[j]), float(r[j + 1])) for r in rows
                        if j + 1 < len(r) and r[j] != "" and r[j + 1] != ""])
        if t_val == 0.0:
            dx.loc[0.0] = [0.0] * len(dx.columns)
            dz.loc[0.0] = [0.0] * len(dz.columns)
        if t_val not in dx.index:
            dx.loc[t_val] = np.nan
            dx = dx.sort_index().interpolate()
        if t_val not in dz.index:
            dz.loc[t_val] = np.nan
            dz = dz.sort_index().interpolate()
        sim_x = (dx.loc[t_val] + 1 / (n_nodes - 1) * np.array(dx.columns.unique())).values
        sim_z = dz.loc[t_val].values
        dists.append(curve_distance_2d(np.column_stack([sim_x, sim_z]), pts))
    return dists, float(max(data.index))


def test_transient_test2_case_on_gpu(tmp_path):
    from dynlattice_amd import case as CS
    c = CS.load(os.path.join(HERE, "golden", "transient", "test2.pro"),
                case_name="tests/transient/test2")
    paths, t_end, steps = CS.run(c, out_dir=str(tmp_path), precision=1e-4)
    print("t_end", t_end, "steps", steps)
    assert t_end > 1.0
    dists, last = _validate(paths["disp"], os.path.join(HERE, "golden", "transient_test2_ref.csv"))
    print("snapshot distances", np.round(dists, 4))
    assert last > 0.9 and len(dists) == 11
    assert max(dists) <= 0.05


def _progress(t, steps):
    print(f"  t = {t:.4f} s after {steps} steps", flush=True)


def test_transient_test1_case_on_gpu(tmp_path):
    """the spin-up from its case files to t = 4 s (the literature comparison
    to 30 s runs on the oracle: test_oracle_spin_up_matches_the_literature;
    beyond about 4.2 s this planar motion is unstable against out-of-plane
    rounding in the reference algorithm, see test_spin_up_case_matches_oracle):
    one Sample row (8 data sets) per step, and the hub angle, which
    integrates the program-driven acceleration, on its closed form
    psi(t) = 6/15 (t^2/2 + (15/2pi)^2 (cos(2 pi t/15) - 1)) within 2e-4
    (the corrector integrates a prescribed acceleration to first order in
    dt: 6.4e-5 on the oracle)"""
    import math

    from dynlattice_amd import case as CS
    c = CS.load(os.path.join(HERE, "golden", "transient", "test1.pro"),
                case_name="tests/transient/test1")
    c.run_while = "t <= 4"
    paths, t_end, steps = CS.run(c, out_dir=str(tmp_path), progress=_progress)
    print("t_end", t_end, "steps", steps)
    sim = np.loadtxt(paths["disp"], delimiter=",")
    assert sim.shape == (steps, 8) and t_end > 4.0
    assert np.array_equal(sim[:, -1], np.sort(sim[:, -1]))
    t = sim[:, -1]
    w = 2 * math.pi / 15
    psi = 6 / 15 * (t ** 2 / 2 + (np.cos(w * t) - 1) / w ** 2)
    err = np.abs(sim[:, -2] - psi).max() / np.abs(psi).max()
    print("hub angle vs closed form", err)
    assert err <= 2e-4
    # still planar at 4 s
    assert np.abs(sim[:, 2:5]).max() <= 1e-9


def test_transient_test3_case_on_gpu(tmp_path):
    """the elbow to t = 30 s (SO(3) update, precision 1e-4 as the case sets):
    energy.csv and disp.gz as the validator reads them; after unloading the
    total energy never rises by 1e-5 between samples (the enforced part of
    tests/transient/test3.py); the out-of-plane displacements' distance to
    Simo & Vu-Quoc's curves is printed (the reference's validator does not
    enforce it)"""
    import pandas as pd

    from dynlattice_amd import case as CS
    from tests.helpers import energy_rise, interp_on_reference, relative_L2
    c = CS.load(os.path.join(HERE, "golden", "transient", "test3.pro"),
                case_name="tests/transient/test3")
    paths, t_end, steps = CS.run(c, out_dir=str(tmp_path), progress=_progress)
    print("t_end", t_end, "steps", steps)
    energy = pd.read_csv(paths["enSample"], index_col="time")
    disp = pd.read_csv(paths["disp"], index_col="time")
    assert max(energy.index) > 0.9 and len(energy) > 25000
    rise = energy_rise(energy.index.values, energy["load"].values, energy["E_tot"].values)
    ref = pd.read_csv(os.path.join(HERE, "golden", "transient_test3_ref.csv"), header=[0, 1])
    t_sim = disp.index.values.astype(float)
    e_elbow = relative_L2(interp_on_reference(t_sim, disp["dz[1]"].values, ref["Ellbow"]["X"].values),
                          ref["Ellbow"]["Y"].values)
    tip = ref["Tip"].dropna()
    e_tip = relative_L2(interp_on_reference(t_sim, disp["dz[2]"].values, tip["X"].values),
                        tip["Y"].values)
    print("max energy rise", rise, "literature distance elbow", e_elbow, "tip", e_tip)
    assert rise < 1e-5


def test_crush_case_on_gpu_matches_oracle(tmp_path):
    """tests/golden/crush/crush.pro (ElastoPlasticRod with the quadratic
    capacity interaction, RodContact, LeapFrog, SO(3)) run by case.run for
    its 400 steps: the Sample rows' bottom.resp.dz (GroupOutput: output load
    case, contact included) and dissipatedEnergy against the oracle on the
    case's own problem"""
    from dynlattice_amd import case as CS
    from oracle.ref_cpu import RefModel
    c = CS.load(os.path.join(HERE, "golden", "crush", "crush.pro"))
    print(f"crush case: {c.problem.n_elems} elements", flush=True)
    paths, t, i = CS.run(c, str(tmp_path))
    assert i == 400
    rows = np.loadtxt(paths["resp"], delimiter=",", ndmin=2)
    assert rows.shape == (400, 5)
    assert np.array_equal(rows[:, 0], np.arange(1, 401))
    ref = RefModel(c.problem, grid=True)
    o = ref.leapfrog(400, group=c.problem.groups["bottom"], group_dof=2, energy=True)
    r = o["reaction"]
    assert np.abs(rows[:, 2] - r).max() <= 1e-9 * np.abs(r).max()
    ed = o["energies"][:, 1]  # the oracle's rows: potential, dissipated, kinetic
    assert ed[-1] > 0
    assert np.abs(rows[:, 4] - ed).max() <= 1e-8 * ed[-1]
    assert ref.plastic_state()[3].mean() > 0.05
