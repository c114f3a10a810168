"""Transient cases from the reference's input files, run on the GPU, with
output the reference's validators read.

:func:`load` reads a .pro case (:mod:`.jemprops`) and its .geo geometry
(:mod:`.geo`) and builds the :class:`~.lattice.Problem` the models of the
case describe:

* ``Input``: GMSHInput (file, order 1 or 2 -- a second-order mesh becomes
  2-node elements on GMSH's second-order node set, see :mod:`.geo` --,
  onelab constants) and GroupInput node
  groups (``xtype``/``ytype``/``ztype`` min | max | mid, ``xval``, ``nodes``,
  ``all``; GroupInputModule.cpp:167-330), plus the built-in group ``all``;
* ``model.model``: the Lattice rod mesh with its ElasticRod material
  (cross sections circle | square | rectangle | explicit area and moments,
  ElasticRodMaterial.cpp:65-190) or ElastoPlasticRod material (the same
  sections; yieldCond read into the device's power-sum family by
  :mod:`.yieldfn`, isotropicCoefficient, a diagonal kinematicTensor,
  maxIter, precision; ElastoPlasticRodMaterial.cpp:55-147), its RodContact
  child ``contact`` (penaltySTS / penaltyNTS or penalty, radius;
  RodContactModel.cpp:60-98), Dirichlet (INCREMENT or under LoadScale;
  the value is a prescribed acceleration in dynamics), Neumann (loadIncr /
  initLoad, or under LoadScale), LoadScale (scaleFunc of t), InitLoad
  (initial displacements and velocities), Mass (point masses), ImpactBC
  (group-reaction-driven accelerations) and None;
  time dependence becomes device programs (:class:`~.lattice.Program`);
  a later Dirichlet model's constraint on a dof replaces an earlier one
  (JIVE's Constraints keep one constraint per dof, the last added: a
  library behaviour outside the reference's sources);
* ``Solver.integrator``: MilneDevice or LeapFrog with deltaTime, precision,
  min/maxDtime, stepSaftey, increase/decreaseFactor, lengthScale; the
  SO(3) update when ``dofs_SO3`` is given (ExplicitModule.cpp:94-106);
  updateWhen, the condition under which a step re-assembles the mass
  (ExplicitModule.cpp:119-126, 256-266);
* ``control.runWhile`` ("t <= T", "t < T", "i < N");
* ``Output``: CSVOutput modules of the state vectors with saveWhen
  (CSVDataPrinter.cpp: header "time,label,<dof>[<node>]...", one row per
  sample); Sample modules (JIVE's SampleModule: one row of dataSets
  expressions per sample, joined by the separator, after the optional
  header; sampleWhen, default every step) over the variables t, i,
  deltaTime, runtime, the energies kineticEnergy, potentialEnergy,
  dissipatedEnergy and mass, and GroupOutput's group variables
  <group>.disp.<dof> and <group>.velo.<dof> (group averages) and
  <group>.resp.<dof> (group sums of the internal force, output load case;
  GroupOutputModule.cpp:119-160).

:func:`run` integrates the case on the GPU (GpuModel) and writes the CSV
files.  The MilneDevice precision default of the reference is JIVE's
``Solver::PRECISION``, a library constant outside its sources; here it is
``DEFAULT_PRECISION`` unless the case sets ``precision``.
"""
from __future__ import annotations

import gzip
import math
import os
import re
from dataclasses import dataclass, field

import numpy as np

from . import geo as G
from . import jemprops as P
from . import lattice as L

DEFAULT_PRECISION = 1e-6
DOF_NAMES = ("dx", "dy", "dz", "rx", "ry", "rz")


class CaseError(ValueError):
    pass


@dataclass
class Output:
    name: str
    file: str
    label: str          # vector label in the CSV ("disp" or "velo"); "" for a Sample module
    save_when: str      # expression of t, i, deltaTime
    data_sets: list = field(default_factory=list)  # Sample: expressions, one column each
    separator: str = " "
    header: str | None = None


@dataclass
class Case:
    name: str
    problem: L.Problem
    integrator: str
    dt: float
    milne: dict
    run_while: str
    outputs: list = field(default_factory=list)
    dof_names: tuple = DOF_NAMES
    mass_update: str | None = None  # updateWhen of the integrator (None: every step)


def _groups(props, coords, en):
    groups = {"all": np.arange(coords.shape[0], dtype=np.int32)}
    gi = P.get(props, "Input.groupInput", {}) or {}
    for name in gi.get("nodeGroups", []) or []:
        spec = gi.get(name, {}) or {}
        if "nodes" in spec:
            groups[name] = np.array(spec["nodes"], np.int32)
            continue
        if spec.get("all", False):
            groups[name] = np.arange(coords.shape[0], dtype=np.int32)
            continue
        eps = float(spec.get("eps", 1e-5))
        sel = np.ones(coords.shape[0], bool)
        for ax, xn in enumerate("xyz"):
            t = spec.get(f"{xn}type")
            if t is not None:
                lo, hi = coords[:, ax].min(), coords[:, ax].max()
                if t == "min":
                    val = lo
                elif t == "max":
                    val = hi
                elif t == "mid":
                    val = 0.5 * (hi + lo)
                elif t == "firstHalf":
                    sel &= (coords[:, ax] > lo - eps) & (coords[:, ax] < lo - eps + (hi - lo) / 2)
                    continue
                elif t == "secondHalf":
                    sel &= (coords[:, ax] > lo + eps + (hi - lo) / 2) & (coords[:, ax] < hi + eps)
                    continue
                else:
                    raise CaseError(f"group {name}: unknown {xn}type {t}")
                sel &= np.abs(coords[:, ax] - val) < eps
            elif f"{xn}val" in spec:
                sel &= np.abs(coords[:, ax] - P.number(spec[f"{xn}val"])) < eps
            elif f"{xn}vals" in spec:
                vals = [P.number(v) for v in spec[f"{xn}vals"]]
                sel &= np.any([np.abs(coords[:, ax] - v) < eps for v in vals], axis=0)
            elif f"{xn}bounds" in spec:
                b = [P.number(v) for v in spec[f"{xn}bounds"]]
                sel &= (coords[:, ax] > b[0] - eps) & (coords[:, ax] < b[1] + eps)
        groups[name] = np.nonzero(sel)[0].astype(np.int32)
    return groups


def _sections(mat, R):
    young = P.number(mat["young"])
    if "shear_modulus" in mat:
        G_ = P.number(mat["shear_modulus"])
    else:
        nu = P.number(mat["poisson_ratio"])
        G_ = young / 2.0 / (nu + 1.0)
    cs = mat.get("cross_section")
    if cs is None:
        A = P.number(mat["area"])
        am = mat["area_moment"]
        am = [P.number(v) for v in (am if isinstance(am, list) else [am])]
        I1, I2 = am[0], am[-1]
        kappa = 5.0 / 6.0
    elif cs == "square":
        s = mat["side_length"]
        s = P.number(s[0] if isinstance(s, list) else s)
        A, I1 = s ** 2, s ** 4 / 12.0
        I2 = I1
        kappa = 5.0 / 6.0
    elif cs == "circle":
        r = P.number(mat["radius"])
        A, I1 = math.pi * r ** 2, math.pi * r ** 4 / 4.0
        I2 = I1
        kappa = 9.0 / 10.0
    elif cs == "rectangle":
        a, b = [P.number(v) for v in mat["side_length"]]
        A = a * b
        I1, I2 = b ** 3 * a / 12.0, a ** 3 * b / 12.0
        kappa = 5.0 / 6.0
    else:
        raise CaseError(f"unknown cross section {cs}")
    J = I1 + I2
    kappa = P.number(mat.get("shear_correction", kappa))
    J = P.number(mat.get("polar_moment", J))
    rho = P.number(mat.get("density", 0.0))
    ones = np.ones(R)
    return L.Sections(area=A * ones, I1=I1 * ones, I2=I2 * ones, J=J * ones,
                      shear_factor=kappa * ones, young=young * ones, shear_modulus=G_ * ones,
                      density=rho * ones,
                      inertia_correct=P.number(mat.get("inertia_correct", 1.0)))


def _plastic(mat, R):
    """ElastoPlasticRod (ElastoPlasticRodMaterial.cpp:55-147): the yield
    condition as the device's power-sum family, the hardening moduli
    (isotropicCoefficient; kinematicTensor, 36 values row by row, diagonal
    only), maxIter and precision (defaults 20 and 1e-5, :38-46)"""
    from .yieldfn import YieldError, to_family
    has_iso = "isotropicCoefficient" in mat
    kin = mat.get("kinematicTensor")
    has_kin = kin is not None
    if "yieldCond" not in mat:
        raise CaseError("Expected a yield function for an elasto-plastic material!")
    if "yieldDeriv" in mat:
        raise CaseError("yieldDeriv: the device differentiates the yield family analytically; "
                        "give yieldCond only")
    try:
        fam = to_family(mat["yieldCond"], has_iso, has_kin)
    except YieldError as e:
        raise CaseError(str(e)) from None
    kc = np.zeros(6)
    if has_kin:
        H = np.array([P.number(x) for x in (kin if isinstance(kin, list) else [kin])])
        if H.size != 36:
            raise CaseError("kinematicTensor: 36 values (6 x 6, row by row) expected")
        H = H.reshape(6, 6)
        if np.abs(H - np.diag(np.diag(H))).max() > 0.0:
            raise CaseError("kinematicTensor: only a diagonal hardening tensor is supported")
        kc = np.diag(H).copy()
    return L.Plastic(weights=np.tile(fam["weights"], (R, 1)), yield_level=fam["yield_level"],
                     kin_coupling=tuple(float(x) for x in fam["kin_coupling"]),
                     iso_coupling=float(fam["iso_coupling"]), has_iso=has_iso,
                     iso_modulus=P.number(mat["isotropicCoefficient"]) if has_iso else 0.0,
                     has_kin=has_kin, kin_const=tuple(float(x) for x in kc), kin_ratio=0.0,
                     precision=P.number(mat.get("precision", 1e-5)),
                     max_iter=int(P.number(mat.get("maxIter", 20))),
                     exponent=tuple(float(x) for x in fam["exponent"]))


def _contact(spec, R):
    """RodContact (RodContactModel.cpp:60-98): penaltySTS and penaltyNTS, or
    penalty for both; the scalar radius (pairs close at twice it); every rod
    of the lattice.  Its updateWhen only lets the output load case reuse the
    last pairs; the device searches at every evaluation, i.e. the default
    condition, and other conditions are refused."""
    if spec.get("type") != "RodContact":
        raise CaseError(f"contact model {spec.get('type')} is not supported")
    if "penaltySTS" in spec and "penaltyNTS" in spec:
        ks, kn = P.number(spec["penaltySTS"]), P.number(spec["penaltyNTS"])
    elif "penalty" in spec:
        ks = kn = P.number(spec["penalty"])
    else:
        raise CaseError("RodContact: penalty (or penaltySTS and penaltyNTS) required")
    if "radius" not in spec:
        raise CaseError("RodContact: radius required")
    uw = spec.get("updateWhen", True)
    if not (uw is True or (isinstance(uw, str) and uw.strip() in ("true", "1"))):
        raise CaseError("RodContact updateWhen: only the default (every evaluation) is supported")
    return L.Contact(radius=np.full(R, P.number(spec["radius"])), penalty_sts=ks,
                     penalty_nts=kn)


def _dof_index(names_t, names_r):
    idx = {n: i for i, n in enumerate(names_t)}
    idx.update({n: 3 + i for i, n in enumerate(names_r)})
    return idx


def _model_root(props) -> str:
    """path of the Multi model that holds the Lattice and the boundary
    models: below ``model``, through Debug / Matrix wrappers (the transient
    cases use model.model, the plastic and contact ones model.model.model)"""
    path = "model"
    for _ in range(8):
        node = P.get(props, path, {}) or {}
        if node.get("type") == "Multi":
            return path
        if node.get("type") not in ("Debug", "Matrix") or "model" not in node:
            break
        path += ".model"
    raise CaseError("no Multi model below 'model' (through Debug / Matrix wrappers)")


def _apply_models(props, p, dofidx, mdl_root="model.model"):
    """boundary models -> cmask/cval/fext, time programs, initial state,
    point masses.  A dof's value is (sum of its factors) x (one program's
    scale); constant scales are folded into the factors."""
    root = P.get(props, mdl_root, {})
    names = root.get("models", [])
    N = p.n_nodes
    # program of each dof's force / acceleration value: -1 constant, -2 unset
    pf = np.full((N, 6), -2, np.int64)
    pa = np.full((N, 6), -2, np.int64)

    def contribute(prog, nodes, dof, factor, force):
        """adds factor to the dof's base force (force) or sets its
        prescribed acceleration, under prog (None: constant 1)"""
        nodes = np.asarray(nodes)
        if prog is not None and prog.kind == "step" and prog.b == 0.0:
            factor, prog = factor * prog.a, None   # constant scale
        if prog is None:
            q = -1
        else:
            found = [j for j, x in enumerate(p.programs) if x is prog]
            if found:
                q = found[0]
            else:
                p.programs.append(prog)
                q = len(p.programs) - 1
        own = pf if force else pa
        cur = own[nodes, dof]
        if force and np.any((cur != -2) & (cur != q)):
            raise CaseError("a dof takes forces from two different time dependences")
        own[nodes, dof] = q
        if force:
            np.add.at(p.fext[:, dof], nodes, factor)
        else:
            p.cmask[nodes, dof] = 1
            p.cval[nodes, dof] = factor

    def model(spec, scale_expr=None):
        typ = spec.get("type", "None")
        if typ in ("None", "Lattice"):
            return
        if typ == "LoadScale":
            model(spec["model"], spec["scaleFunc"])
            return
        if typ == "Dirichlet":
            groups, dofs = spec["nodeGroups"], spec["dofs"]
            fac = spec.get("factors")
            if fac is None:
                fac = [0.0] * len(groups)
                if "loaded" in spec:
                    fac[int(spec["loaded"])] = 1.0
            if "dispRate" in spec:
                raise CaseError("Dirichlet dispRate: the explicit modules never set its step size")
            if scale_expr is not None:
                prog = L.Program("expr", expr=scale_expr)
            elif "dispIncr" in spec:
                prog = L.Program("step", a=P.number(spec.get("initDisp", 0.0)),
                                 b=P.number(spec["dispIncr"]), c=0.0)
            else:
                prog = None  # load-scale method without a parent: unit scale
            for g, d, f in zip(groups, dofs, fac):
                contribute(prog, p.groups[g], dofidx[d], P.number(f), force=False)
            return
        if typ == "Neumann":
            groups, dofs, fac = spec["nodeGroups"], spec["dofs"], spec["factors"]
            if scale_expr is not None:
                prog = L.Program("expr", expr=scale_expr)
            elif "loadIncr" in spec:
                prog = L.Program("step", a=P.number(spec.get("initLoad", 0.0)),
                                 b=P.number(spec["loadIncr"]), c=1.0)
            else:
                raise CaseError("Neumann without loadIncr needs a LoadScale parent")
            for g, d, f in zip(groups, dofs, fac):
                contribute(prog, p.groups[g], dofidx[d], P.number(f), force=True)
            return
        if typ == "InitLoad":
            for key, arr in (("disp", p.disp0), ("velo", p.velo0)):
                gs = spec.get(f"{key}Groups", [])
                for g, d, v in zip(gs, spec.get(f"{key}Dofs", []), spec.get(f"{key}Vals", [])):
                    arr[p.groups[g], dofidx[d]] = P.number(v)
            return
        if typ == "Mass":
            tm = spec["totalMass"]
            tm = tm if isinstance(tm, list) else [tm]
            dset = set(dofidx[d] for d in spec["dofs"])
            trans, rot = dset & {0, 1, 2}, dset & {3, 4, 5}
            if trans not in (set(), {0, 1, 2}) or rot not in (set(), {3, 4, 5}):
                raise CaseError("Mass: all or none of the translational / rotational dofs")
            if p.point_mass is None:
                p.point_mass = np.zeros(N)
                p.point_inertia = np.zeros(N)
            for g, m in zip(spec["nodeGroups"], tm):
                nodes = p.groups[g]
                share = P.number(m) / len(nodes)
                if trans:
                    p.point_mass[nodes] += share
                if rot:
                    p.point_inertia[nodes] += share
            return
        if typ == "ImpactBC":
            for g, d, w in zip(spec["nodeGroups"], spec["dofs"], spec["weights"]):
                p.impact.append((np.asarray(p.groups[g], np.int32), dofidx[d], P.number(w)))
            return
        raise CaseError(f"model type {typ} is not supported")

    for n in names:
        model(root[n])
    if p.programs:
        # a constrained dof's value is its acceleration, a free dof's its force
        q = np.where(p.cmask.astype(bool), pa, pf)
        p.dof_prog = np.where(q >= 0, q, -1).astype(np.int32)


def load(pro_path: str, case_name: str | None = None) -> Case:
    props = P.parse(pro_path, case_name=case_name)
    name = props["CASE_NAME"]
    base = os.path.dirname(pro_path)
    inp = P.get(props, "Input.input", {})
    if inp.get("type") != "GMSHInput":
        raise CaseError("only GMSHInput geometry is supported")
    geo_file = inp["file"]
    geo_path = geo_file if os.path.isabs(geo_file) else os.path.join(base, os.path.basename(geo_file))
    onelab = {k: P.number(v) for k, v in (inp.get("onelab") or {}).items()}
    pts, lines = G.read(geo_path, onelab)
    coords, en, es, _ = G.mesh(pts, lines, int(inp.get("order", 1)))
    # the rod model and its material
    mroot = _model_root(props)
    root = P.get(props, mroot, {})
    rod = next((root[m] for m in root.get("models", []) if root[m].get("type") == "Lattice"), None)
    if rod is None:
        raise CaseError("no Lattice rod model")
    child = rod["child"]
    mat = child["material"]
    if mat.get("type") not in ("ElasticRod", "ElastoPlasticRod"):
        raise CaseError(f"material {mat.get('type')} is not supported by the case reader")
    R = len(lines)
    p = L.Problem(coords=coords, elem_nodes=en, elem_strut=es, n_struts=R,
                  sections=_sections(mat, R)).finalize()
    if mat.get("type") == "ElastoPlasticRod":
        p.plastic = _plastic(mat, R)
    if "contact" in rod:
        p.contact = _contact(rod["contact"], R)
    p.groups = _groups(props, coords, en)
    dofidx = _dof_index(child.get("dofNamesTrans", ["dx", "dy", "dz"]),
                        child.get("dofNamesRot", ["rx", "ry", "rz"]))
    _apply_models(props, p, dofidx, mroot)
    integ = P.get(props, "Solver.integrator", {})
    kind = integ.get("type", "MilneDevice")
    if kind not in ("MilneDevice", "LeapFrog"):
        raise CaseError(f"integrator {kind} is not supported")
    p.so3 = "dofs_SO3" in integ
    dt = P.number(integ["deltaTime"])
    p.dt = dt
    p.name = name
    milne = {"precision": P.number(integ.get("precision", DEFAULT_PRECISION)),
             "dt_min": P.number(integ.get("minDtime", dt / 100.0)),
             "dt_max": P.number(integ.get("maxDtime", dt * 100.0)),
             "safety": P.number(integ.get("stepSaftey", 0.9)),
             "incr": P.number(integ.get("increaseFactor", 1.2)),
             "decr": P.number(integ.get("decreaseFactor", 0.8)),
             "len_scale": P.number(integ.get("lengthScale", 1e-3))}
    outs = []
    out = P.get(props, "Output", {}) or {}
    for m in out.get("modules", []) or []:
        spec = out.get(m, {}) or {}
        if spec.get("type") == "Sample":
            ds = spec.get("dataSets", [])
            ds = ds if isinstance(ds, list) else [ds]
            outs.append(Output(m, spec["file"], "", spec.get("sampleWhen", "1"),
                               data_sets=[str(x) for x in ds],
                               separator=spec.get("separator", " "), header=spec.get("header")))
            continue
        if spec.get("type") != "CSVOutput":
            continue
        vecs = spec.get("vectors", ["state = disp"])
        for v in vecs:
            lhs, _, rhs = v.partition("=")
            label = rhs.strip() or lhs.strip()
            if label not in ("disp", "velo"):
                raise CaseError(f"CSVOutput vector {v!r} is not supported")
            outs.append(Output(m, spec["file"], label, spec.get("saveWhen", "1")))
    uw = integ.get("updateWhen")
    if isinstance(uw, bool) or uw is None:
        if uw is False:
            raise CaseError("updateWhen = false: the mass would never be assembled")
        uw = None
    elif isinstance(uw, (int, float)):
        uw = None if uw != 0 else "0"
    return Case(name=name, problem=p, integrator=kind, dt=dt, milne=milne,
                mass_update=uw,
                run_while=P.get(props, "control.runWhile", "t <= 0"), outputs=outs,
                dof_names=tuple(list(child.get("dofNamesTrans", ["dx", "dy", "dz"])) +
                                list(child.get("dofNamesRot", ["rx", "ry", "rz"]))))


def _cond(expr: str, t: float, i: int, dtime: float) -> bool:
    from .gpu import program_eval
    e = re.sub(r"\bdeltaTime\b", f"({dtime!r})", expr)
    return program_eval(e, t, float(i)) != 0.0


def write_csv(path: str, dof_names, n_nodes: int, rows):
    """CSVDataPrinter layout (CSVDataPrinter.cpp: startVectors_, writePrefix_,
    writeVector_): header time,label,<dof>[<node>]... node-major, then one
    row per sample (t, label, values)"""
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "wt") as f:
        f.write("time,label," + ",".join(f"{d}[{n}]" for n in range(n_nodes) for d in dof_names)
                + "\n")
        for t, label, vec in rows:
            f.write(repr(float(t)) + "," + label + "," + ",".join(repr(float(x)) for x in vec) + "\n")


_NAME = re.compile(r"[A-Za-z_][A-Za-z0-9_]*(?:\.[A-Za-z_][A-Za-z0-9_]*)*")


def _sample_value(expr: str, t: float, i: int, var) -> float:
    """value of a Sample data set: names other than t, i, PI and functions
    are looked up with var(name) and substituted, then the expression is
    evaluated by the program compiler"""
    from .gpu import program_eval

    def rep(m):
        v = var(m.group(0))
        return m.group(0) if v is None else f"({float(v)!r})"
    return program_eval(_NAME.sub(rep, expr), t, float(i))


def write_sample(path: str, header, separator: str, rows):
    """SampleModule layout: the header line if any, then one row per sample"""
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "wt") as f:
        if header:
            f.write(header + "\n")
        for r in rows:
            f.write(separator.join(repr(float(x)) for x in r) + "\n")


def run(case: Case, out_dir: str = ".", precision: float | None = None, chunk: int = 1,
        device: int = 0, progress=None, progress_every: float = 10.0):
    """Integrates the case on the GPU while runWhile holds; samples every
    output whenever its saveWhen / sampleWhen holds after a taken step (the
    reference's output modules run after each accepted step) and writes the
    files under out_dir.  progress(t, steps), if given, is called about
    every progress_every seconds of wall time.  Returns {output name: path},
    the final time and the number of steps taken."""
    import time as _time

    from .gpu import DL_TRACK_DISSIPATION, GpuModel, milne_opts
    p = case.problem
    # dissipatedEnergy (the material's energyDiss_ summed at COMMIT) needs the
    # step's dissipation sums
    g = GpuModel(p, device=device, flags=DL_TRACK_DISSIPATION if p.plastic is not None else 0)
    if case.mass_update is not None:
        g.set_mass_update(case.mass_update)
    rows = {o.name: [] for o in case.outputs}
    dof_of = {n: k for k, n in enumerate(case.dof_names)}
    t_start = _time.perf_counter()

    def sample(t, i, dt_next):
        cache = {}

        def state():
            if "state" not in cache:
                cache["state"] = g.get_state()
            return cache["state"]

        def var(name):
            if name in ("t", "i", "PI"):
                return None
            if name == "deltaTime":
                return dt_next
            if name == "runtime":
                return _time.perf_counter() - t_start
            if name in ("kineticEnergy", "potentialEnergy", "dissipatedEnergy", "mass"):
                if "en" not in cache:
                    cache["en"] = g.energies()
                key = {"kineticEnergy": "kinetic", "potentialEnergy": "potential",
                       "dissipatedEnergy": "dissipated", "mass": "mass"}[name]
                return cache["en"][key]
            parts = name.split(".")
            if len(parts) == 3 and parts[0] in p.groups and parts[2] in dof_of:
                nodes, k = p.groups[parts[0]], dof_of[parts[2]]
                if parts[1] in ("disp", "velo"):
                    d, v = state()
                    a = d if parts[1] == "disp" else v
                    return float(a[nodes, k].sum() / len(nodes))
                if parts[1] == "resp":
                    if "resp" not in cache:
                        cache["resp"] = g.int_vector(inelastic=False)
                    return float(cache["resp"][nodes, k].sum())
            return None

        for o in case.outputs:
            if not _cond(o.save_when, t, i, dt_next):
                continue
            if o.label:
                d, v = state()
                rows[o.name].append((t, o.label, (d if o.label == "disp" else v).reshape(-1)))
            else:
                rows[o.name].append([_sample_value(e, t, i, var) for e in o.data_sets])

    t, i = 0.0, 0
    m = case.milne
    prec = m["precision"] if precision is None else precision
    dt = case.dt
    t_report = _time.perf_counter()
    try:
        while _cond(case.run_while, t, i, dt):
            if progress is not None and _time.perf_counter() - t_report > progress_every:
                progress(t, i)
                t_report = _time.perf_counter()
            if case.integrator == "LeapFrog":
                g.step(1, dt, out=False)
                t += dt
                i += 1
                sample(t, i, dt)
                continue
            q = g.milne(chunk, milne_opts(dt, prec, dt_min=m["dt_min"], dt_max=m["dt_max"],
                                          safety=m["safety"], incr=m["incr"], decr=m["decr"],
                                          len_scale=m["len_scale"]))
            for k in range(len(q["dt"])):
                if q["accepted"][k]:
                    t += float(q["dt"][k])
                    i += 1
                    if chunk == 1:
                        sample(t, i, float(q["dt_next"][k]))
            if chunk > 1:
                sample(t, i, float(q["dt_next"][-1]))
            dt = float(q["dt_next"][-1])
    finally:
        g.close()
    paths = {}
    for o in case.outputs:
        path = os.path.join(out_dir, P._subst(o.file, {"CASE_NAME": case.name}))
        if o.label:
            write_csv(path, case.dof_names, p.n_nodes, rows[o.name])
        else:
            write_sample(path, o.header, o.separator, rows[o.name])
        paths[o.name] = path
    return paths, t, i
