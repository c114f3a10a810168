"""Reading the reference's transient case files (CPU): the swinging-rod case
of the reference's tests/transient (test2.pro with its includes input.pro,
model.pro, output.pro and test2.geo, copied to tests/golden/transient as
input fixtures) becomes the same problem as the hand-built
lattice.swinging_rod() (mesh, groups, supports, gravity loads, material,
integrator settings, output), and the CSV writer produces what the
reference's validator (tests/transient/test2.py) reads."""
import os

import numpy as np
import pytest

from dynlattice_amd import case as CS
from dynlattice_amd import geo as G
from dynlattice_amd import jemprops as JP
from dynlattice_amd import lattice as L

HERE = os.path.dirname(os.path.abspath(__file__))
T2 = os.path.join(HERE, "golden", "transient", "test2.pro")


def test_property_reader():
    props = JP.parse(T2, case_name="tests/transient/test2")
    assert JP.get(props, "Solver.integrator.type") == "MilneDevice"
    assert JP.get(props, "Solver.integrator.deltaTime") == 1e-6
    # the override in test2.pro replaces model.pro's translational + rotational list
    assert JP.get(props, "model.model.fixed.dofs") == ["dx", "dy", "dz"]
    # += on a property set merges params.rod_details into the rod child
    assert JP.get(props, "model.model.rodMesh.child.material.radius") == 5e-3
    # $(...) substitution of an earlier property
    assert JP.get(props, "model.model.force.initLoad") == \
        "PI * (5e-3)^2 * 0.1 * 1.1e3 * 9.81"
    assert JP.get(props, "Output.disp.file") == "tests/transient/test2/disp.gz"
    assert JP.get(props, "Output.modules") == ["loadextent", "disp", "paraview"]


def test_geo_reader_and_line_mesher():
    pts, lines = G.read(os.path.join(HERE, "golden", "transient", "test2.geo"), {"size": 0.25})
    assert pts[1][3] == 0.25 and lines == [(1, 1, 2)]
    x, en, es, names = G.mesh(pts, lines)
    # 1 / 0.25 = 4 elements; points first, then the line's interior nodes
    assert np.allclose(x[:, 0], [0.0, 1.0, 0.25, 0.5, 0.75])
    assert en.tolist() == [[0, 2], [2, 3], [3, 4], [4, 1]]
    assert names == ["beam_1"]


def test_swinging_rod_case_equals_the_hand_built_problem():
    c = CS.load(T2, case_name="tests/transient/test2")
    p, q = c.problem, L.swinging_rod()
    assert c.integrator == "MilneDevice" and c.dt == q.dt and c.run_while == "t <= 1"
    assert not p.so3  # no dofs_SO3 in test2.pro: additive rotation update
    # same geometry up to node numbering
    order = np.argsort(p.coords[:, 0])
    assert np.allclose(p.coords[order], q.coords, atol=1e-15)
    perm = np.empty_like(order)
    perm[order] = np.arange(len(order))
    e_p = np.sort(perm[p.elem_nodes], axis=1)
    assert np.array_equal(e_p[np.argsort(e_p[:, 0])], q.elem_nodes)
    assert np.array_equal(p.cmask[order], q.cmask)
    assert np.allclose(p.fext[order][:, 2][1:], q.fext[1:, 2], rtol=1e-15)  # node 0 is fixed
    for f in ("area", "I1", "J", "shear_factor", "young", "shear_modulus", "density"):
        assert np.allclose(getattr(p.sections, f), getattr(q.sections, f), rtol=1e-15), f
    assert p.groups["fixed"].tolist() == [0] and p.groups["free"].tolist() == [1]
    assert c.outputs[0].file == "tests/transient/test2/disp.gz"
    assert c.outputs[0].save_when == "t % 1e-4 < deltaTime"


def test_csv_layout_reads_like_the_reference_validator(tmp_path):
    pd = pytest.importorskip("pandas")
    rows = [(0.5, "disp", np.arange(12.0)), (1.0, "disp", -np.arange(12.0))]
    path = str(tmp_path / "case" / "disp.gz")
    CS.write_csv(path, CS.DOF_NAMES, 2, rows)
    # as tests/transient/test2.py reads it
    data = pd.read_csv(path, index_col=["time", "label"]).xs("disp", level="label")
    data.columns = pd.MultiIndex.from_tuples(
        [(n[:n.find("[")], int(n[n.find("[") + 1:n.find("]")])) for n in data.columns],
        names=["dof", "node"])
    dz = data.xs("dz", axis=1, level="dof")
    assert dz.loc[0.5, 1] == 8.0 and dz.loc[1.0, 0] == -2.0
    assert list(data.index) == [0.5, 1.0]


def test_unsupported_inputs_are_rejected(tmp_path):
    with pytest.raises(G.GeoError):
        G.mesh({1: (0, 0, 0, 0.1), 2: (1, 0, 0, 0.1)}, [(1, 1, 2)], order=3)
    with pytest.raises(G.GeoError):
        G.mesh({1: (0, 0, 0, 0.1), 2: (1, 0, 0, 0.2)}, [(1, 1, 2)])


T1 = os.path.join(HERE, "golden", "transient", "test1.pro")
T3 = os.path.join(HERE, "golden", "transient", "test3.pro")


def test_spin_up_case_reads_as_the_reference_builds_it():
    """tests/transient/test1.pro (Simo & Vu-Quoc Ex. 5.1): the second-order
    mesh of 4 elements becomes 8 two-node elements on the same 9 nodes; the
    hub is clamped by model.pro's Dirichlet model and its rz then follows
    the LoadScale'd Dirichlet model listed after it (the later constraint
    replaces the earlier); the Sample output's data sets"""
    c = CS.load(T1, case_name="tests/transient/test1")
    p = c.problem
    assert p.n_nodes == 9 and p.n_elems == 8 and not p.so3
    assert np.allclose(np.sort(p.coords[:, 0]), np.arange(9) * 1.25)
    hub, tip = p.groups["fixed"], p.groups["free"]
    assert hub.tolist() == [0] and tip.tolist() == [1]
    assert p.cmask[hub].all() and np.array_equal(p.cval[hub][0], [0, 0, 0, 0, 0, 1.0])
    assert p.dof_prog[hub][0].tolist() == [-1, -1, -1, -1, -1, 0]
    assert p.programs[0].expr == "if (t<15, 6/15 * (1 - cos(2*PI/15 * t)), 0)"
    s = p.sections
    assert np.allclose([s.area[0], s.I1[0], s.young[0], s.shear_modulus[0], s.density[0]],
                       [6e-3, 3e-6, 5.6e10 / 12, 2e9, 200.0], rtol=1e-14)
    o = c.outputs[0]
    assert o.label == "" and o.separator == "," and o.save_when == "1"
    assert o.data_sets == ["free.disp.dx", "free.disp.dy", "free.disp.dz", "free.disp.rx",
                           "free.disp.ry", "free.disp.rz", "fixed.disp.rz", "t"]


def test_elbow_case_reads_as_the_reference_builds_it():
    """tests/transient/test3.pro (Simo & Vu-Quoc Ex. 5.2): SO(3) update,
    precision 1e-4, the elbow's dz load program, shear correction 2 and
    inertia correction 1e4, the energy Sample with its header"""
    c = CS.load(T3, case_name="tests/transient/test3")
    p = c.problem
    assert p.n_nodes == 41 and p.n_elems == 40 and p.so3
    assert {k: v.tolist() for k, v in p.groups.items() if k != "all"} == \
        {"fixed": [0], "free": [2], "elbow": [1]}
    assert p.cmask[0].all() and not p.cmask[1:].any()
    assert p.fext[1].tolist() == [0, 0, 1, 0, 0, 0] and p.dof_prog[1, 2] == 0
    assert c.milne["precision"] == 1e-4 and c.dt == 1e-5
    assert p.sections.shear_factor[0] == 2.0 and p.sections.inertia_correct == 1e4
    en = [o for o in c.outputs if o.name == "enSample"][0]
    assert en.header == "step,time,time_step,run_time,load,E_kin,E_pot,E_tot"
    assert en.save_when == "t % 0.001 < deltaTime" and en.data_sets[-1] == \
        "kineticEnergy+potentialEnergy"


def test_sample_writer_layout(tmp_path):
    path = str(tmp_path / "c" / "e.csv")
    CS.write_sample(path, "a,b", ",", [[1.0, 2.5], [3.0, -4.0]])
    lines = open(path).read().splitlines()
    assert lines == ["a,b", "1.0,2.5", "3.0,-4.0"]
    assert np.loadtxt(path, delimiter=",", skiprows=1).tolist() == [[1.0, 2.5], [3.0, -4.0]]


def test_oracle_spin_up_matches_the_literature():
    """The oracle's MilneDevice on the spin-up case (as read from the case
    files) against Simo & Vu-Quoc's curves with tests/transient/test1.py's
    criterion: relative L2 <= 5 % for the hub angle, the rotational
    deviation and the tip's axial and tangential deviations, t <= 30 s at
    precision 1e-6 (the library default here; JIVE's default is not in the
    reference's sources).  Sampled every 200 attempts (every 10 gives the
    same errors to three digits)."""
    from oracle.ref_cpu import RefModel, milne_opts
    from tests.helpers import transient_test1_errors
    c = CS.load(T1, case_name="tests/transient/test1")
    p, m = c.problem, c.milne
    r = RefModel(p)
    hub, tip = p.groups["fixed"][0], p.groups["free"][0]
    t, dt, rows = 0.0, c.dt, []
    while t <= 30.0:
        o = r.milne(200, milne_opts(dt, CS.DEFAULT_PRECISION, dt_min=m["dt_min"],
                                    dt_max=m["dt_max"]))
        t += float(o["dt"][o["accepted"]].sum())
        dt = float(o["dt_next"][-1])
        d = r.disp.reshape(-1, 6)
        rows.append(list(d[tip]) + [d[hub, 5], t])
    err = transient_test1_errors(np.array(rows), os.path.join(HERE, "golden"))
    print(err)
    assert max(err.values()) <= 0.05, err


def test_oracle_elbow_energy_criterion():
    """The oracle's MilneDevice on the elbow case (as read from the case
    files, SO(3) update, precision 1e-4) to t = 30 s with
    tests/transient/test3.py's criterion: after the load is removed (t > 2)
    the total energy never rises by 1e-5 or more between samples; it
    decays slowly (the corrector's numerical dissipation)."""
    from oracle.ref_cpu import RefModel, milne_opts
    from tests.helpers import energy_rise
    c = CS.load(T3, case_name="tests/transient/test3")
    p, m = c.problem, c.milne
    r = RefModel(p)
    t, dt, rows = 0.0, c.dt, []
    while t <= 30.0:
        o = r.milne(40, milne_opts(dt, m["precision"], dt_min=m["dt_min"], dt_max=m["dt_max"]))
        t += float(o["dt"][o["accepted"]].sum())
        dt = float(o["dt_next"][-1])
        load = CS._sample_value(p.programs[0].expr, t, 0, lambda n: None)
        ek, ep = r.kinetic_energy().sum(), r.elem_potential_energy().sum()
        rows.append((t, load, ek + ep))
    a = np.array(rows)
    assert a[-1, 0] > 30.0
    rise = energy_rise(a[:, 0], a[:, 1], a[:, 2])
    after = a[a[:, 1] == 0, 2]
    print("max rise", rise, "E_tot after unloading", after[0], after[-1])
    assert rise < 1e-5
    assert 0.0 < (after[0] - after[-1]) / after[0] < 0.05


# ------------------------------------------------------- ElastoPlasticRod, RodContact
# the yieldCond strings of the reference's plastic cases (data, quoted from
# tests/plastic/test1.pro:22-28, test2a.pro:22, test2b.pro:22, test3.pro:22-28)
YIELD_T1 = ("  abs(dx) + abs(dy) + abs(dz+0*h_dz) + abs(rx/0.12) + abs(ry/0.12) + abs(rz/0.12) "
            "- 10 * (1+0*h_0)")
YIELD_T2A = "abs(dz+0*h_dz) - 10 * (1+1*h_0)"
YIELD_T2B = "abs(dz+20*h_dz) - 10 * (1+0*h_0)"
YIELD_T3 = ("  abs(dx/( 700-h_dx))^2.04 + abs(dy/( 700-h_dy))^2.04 + abs(dz/(1470-h_dz))^1.76 "
            "+ abs(rx/(0.62-h_rx))^2.09 + abs(ry/(0.62-h_ry))^2.09 + abs(rz/(0.56-h_rz))^1.73 - 1")


def test_reference_yield_conditions_read_into_the_family():
    from dynlattice_amd.yieldfn import YieldError, to_family
    f = to_family(YIELD_T1, True, True)
    assert np.allclose(f["weights"], [1, 1, 1, 1 / 0.12, 1 / 0.12, 1 / 0.12], rtol=1e-12)
    assert np.array_equal(f["exponent"], np.ones(6))
    assert abs(f["yield_level"] - 10.0) < 1e-12 and abs(f["iso_coupling"]) < 1e-12
    assert np.abs(f["kin_coupling"]).max() < 1e-12
    f = to_family(YIELD_T2A, True, True)
    assert np.allclose(f["weights"], [0, 0, 1, 0, 0, 0]) and abs(f["iso_coupling"] - 1) < 1e-12
    f = to_family(YIELD_T2B, True, True)
    assert np.allclose(f["kin_coupling"], [0, 0, 20, 0, 0, 0], rtol=1e-12)
    # hardening inside a denominator is no power-sum term: refused
    with pytest.raises(YieldError):
        to_family(YIELD_T3, False, True)
    # general exponents, isotropic coupling of either sign
    f = to_family("(dx/2)^2 + (dz/3)^2 + abs(rz+0.5*h_rz)^1.5 - 4*(1-0.2*h_0)", True, True)
    assert np.allclose(f["weights"], [0.5, 0, 1 / 3, 0, 0, 1], rtol=1e-12)
    assert np.allclose(f["exponent"], [2, 1, 2, 1, 1, 1.5], rtol=1e-12)
    assert abs(f["iso_coupling"] + 0.2) < 1e-12 and abs(f["kin_coupling"][5] - 0.5) < 1e-12
    with pytest.raises(YieldError):
        to_family("dx^2 + dy - 1", False, False)  # a linear term without abs
    with pytest.raises(YieldError):
        to_family("abs(dx) - 1 + h_0", False, False)  # h_0 without isotropic hardening


CRUSH = os.path.join(HERE, "golden", "crush", "crush.pro")


def _crush_by_hand():
    R = 64
    r = np.full(R, 1.2e-3)
    p = L.bcc_lattice(2, 3, 10e-3, r, 70e9, 0.3, 2700.0)
    p.plastic = L.Plastic.resultant_capacities(r, 250e6, kin_ratio=0.01, exponent=2.0)
    p.contact = L.Contact(radius=r.copy(), penalty_sts=1e6, penalty_nts=1e6)
    L.crush_bcs(p, -20.0)
    p.dt = 2.5e-7
    return p


def test_crush_case_reads_material_and_contact():
    """tests/golden/crush/crush.pro (an explicit crush in the reference's
    property layout, Debug / Matrix / Multi): ElastoPlasticRod with the
    quadratic capacity interaction and a diagonal kinematicTensor, a
    RodContact child, LeapFrog with the SO(3) update"""
    c = CS.load(CRUSH)
    p, h = c.problem, _crush_by_hand()
    assert c.integrator == "LeapFrog" and p.so3 and p.dt == 2.5e-7
    assert (p.n_elems, p.n_nodes, p.n_struts) == (h.n_elems, h.n_nodes, h.n_struts)
    pl = p.plastic
    assert np.allclose(pl.weights, h.plastic.weights, rtol=1e-12, atol=0)
    assert pl.exponent == (2.0,) * 6 and pl.has_kin and not pl.has_iso
    assert np.allclose(pl.kin_coupling, 1.0, rtol=1e-12) and abs(pl.yield_level - 1) < 1e-12
    Cd = np.array([0.9 * p.sections.shear_modulus[0] * p.sections.area[0]] * 2 +
                  [p.sections.young[0] * p.sections.area[0], p.sections.young[0] * p.sections.I1[0],
                   p.sections.young[0] * p.sections.I2[0], p.sections.shear_modulus[0] * p.sections.J[0]])
    assert np.allclose(pl.kin_const, Cd / 100.0, rtol=1e-12) and pl.kin_ratio == 0.0
    assert np.array_equal(p.contact.radius, h.contact.radius)
    assert p.contact.penalty_sts == p.contact.penalty_nts == 1e6
    assert len(p.groups["bottom"]) == len(p.groups["top"]) == 9
    assert (p.velo0[p.groups["top"], 2] == -20.0).all()


def test_crush_case_runs_like_the_hand_built_problem_on_the_oracle():
    """the case's problem and the hand-built one (other node numbering, same
    physics) give the same bottom reaction and dissipation on the oracle"""
    from oracle.ref_cpu import RefModel
    c = CS.load(CRUSH)
    outs = []
    for p in (c.problem, _crush_by_hand()):
        ref = RefModel(p, grid=True)
        o = ref.leapfrog(300, group=p.groups["bottom"], group_dof=2)
        outs.append((o["reaction"], ref.plastic_state()[3].mean()))
    (ra, aa), (rb, ab) = outs
    assert aa > 0.05 and aa == ab
    assert np.abs(ra - rb).max() <= 1e-9 * np.abs(rb).max()


def test_crush_case_refuses_what_the_device_cannot_do(tmp_path):
    import shutil
    for name in ("crush.pro", "crush.geo"):
        shutil.copy(os.path.join(HERE, "golden", "crush", name), tmp_path / name)
    base = (tmp_path / "crush.pro").read_text()
    for extra, msg in (
            ('model.model.model.lattice.child.material.yieldCond = "abs(dx/(1-h_dx))^2 - 1";',
             "yield condition"),
            ("model.model.model.lattice.child.material.kinematicTensor = [ 1., 1., 0., 0., 0., 0., "
             "0., 1., 0., 0., 0., 0., 0., 0., 1., 0., 0., 0., 0., 0., 0., 1., 0., 0., 0., 0., 0., "
             "0., 1., 0., 0., 0., 0., 0., 0., 1. ];", "diagonal"),
            ('model.model.model.lattice.contact.updateWhen = "i%10 < 1";', "updateWhen")):
        (tmp_path / "crush.pro").write_text(base + extra + "\n")
        with pytest.raises(CS.CaseError, match=msg):
            CS.load(str(tmp_path / "crush.pro"))
