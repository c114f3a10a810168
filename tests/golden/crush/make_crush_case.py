"""Writes the explicit crush case crush.pro / crush.geo next to this script.

    python tests/golden/crush/make_crush_case.py

A 2 x 2 x 2 BCC lattice of 10 mm cells (every cell centre joined to its 8
corners: 64 struts, lines of GMSH points) with the ElastoPlasticRod material
(solid circle r = 1.2 mm, E = 70 GPa, nu = 0.3, rho = 2700; yieldCond the
quadratic interaction of the resultant capacities at s_y = 250 MPa, written
out as a JEM expression; kinematicTensor = diag(C) / 100), a RodContact
child (penalty 1e6 N/m, radius 1.2 mm), bottom face clamped, top face moving
at -20 m/s (InitLoad velocity, zero prescribed acceleration), LeapFrog with
the SO(3) update.  The property layout is the reference's (model.pro of its
contact and plastic tests: Debug / Matrix / Multi with a Lattice model).
"""
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
CELL, N, R = 10e-3, 2, 1.2e-3
E, NU, RHO, SY = 70e9, 0.3, 2700.0, 250e6
DT = 2.5e-7


def geo():
    pts, lines = {}, []

    def pid(x, y, z):
        key = (round(x, 9), round(y, 9), round(z, 9))
        if key not in pts:
            pts[key] = len(pts) + 1
        return pts[key]

    for k in range(N):
        for j in range(N):
            for i in range(N):
                c = pid((i + 0.5) * CELL, (j + 0.5) * CELL, (k + 0.5) * CELL)
                for dk in (0, 1):
                    for dj in (0, 1):
                        for di in (0, 1):
                            lines.append((c, pid((i + di) * CELL, (j + dj) * CELL,
                                                 (k + dk) * CELL)))
    lc = math.sqrt(3.0) / 2.0 * CELL / 3.0 * 1.01  # three elements per strut
    out = ["// 2 x 2 x 2 BCC lattice, 10 mm cells (make_crush_case.py)"]
    for key, t in sorted(pts.items(), key=lambda kv: kv[1]):
        out.append(f"Point({t}) = {{ {key[0]!r}, {key[1]!r}, {key[2]!r}, {lc!r} }};")
    for n, (a, b) in enumerate(lines, 1):
        out.append(f"Line({n}) = {{ {a}, {b} }};")
    return "\n".join(out) + "\n"


def pro():
    A = math.pi * R ** 2
    Np, Vp = SY * A, SY * A / math.sqrt(3.0)
    Mp, Tp = SY * 4.0 * R ** 3 / 3.0, SY * 2.0 * math.pi * R ** 3 / (3.0 * math.sqrt(3.0))
    G = E / 2.0 / (1.0 + NU)
    I = math.pi * R ** 4 / 4.0
    C = [0.9 * G * A, 0.9 * G * A, E * A, E * I, E * I, G * 2 * I]
    H = [[C[i] / 100.0 if i == j else 0.0 for j in range(6)] for i in range(6)]
    yc = " + ".join(f"(({d}+{h})/{cap!r})^2" for d, h, cap in
                    zip(("dx", "dy", "dz", "rx", "ry", "rz"),
                        ("h_dx", "h_dy", "h_dz", "h_rx", "h_ry", "h_rz"),
                        (Vp, Vp, Np, Mp, Mp, Tp)))
    return f"""// explicit crush of a 2 x 2 x 2 BCC lattice (make_crush_case.py)
control.runWhile = "i < 400";

Solver.modules = [ "integrator" ];
Solver.integrator.type = "LeapFrog";
Solver.integrator.deltaTime = {DT!r};
Solver.integrator.dofs_SO3 = [ "rx", "ry", "rz" ];

Input.modules = [ "input", "groupInput" ];
Input.input.type = "GMSHInput";
Input.input.file = "$(CASE_NAME).geo";
Input.groupInput.type = "GroupInput";
Input.groupInput.nodeGroups = [ "bottom", "top" ];
Input.groupInput.bottom.ztype = "min";
Input.groupInput.top.ztype = "max";

params.material.type = "ElastoPlasticRod";
params.material.young = {E!r};
params.material.poisson_ratio = {NU!r};
params.material.density = {RHO!r};
params.material.shear_correction = 0.9;
params.material.cross_section = "circle";
params.material.radius = {R!r};
params.material.yieldCond = "{yc} - 1";
params.material.kinematicTensor = [ {", ".join(repr(x) for row in H for x in row)} ];

model.type = "Debug";
model.model.type = "Matrix";
model.model.model.type = "Multi";
model.model.model.models = [ "lattice", "bottom", "top", "impact" ];
model.model.model.lattice.type = "Lattice";
model.model.model.lattice.prefix = "beam_";
model.model.model.lattice.child.type = "specialCosseratRod";
model.model.model.lattice.child.dofNamesTrans = ["dx", "dy", "dz"];
model.model.model.lattice.child.dofNamesRot = ["rx", "ry", "rz"];
model.model.model.lattice.child.material = params.material;
model.model.model.lattice.contact.type = "RodContact";
model.model.model.lattice.contact.penalty = 1e6;
model.model.model.lattice.contact.radius = {R!r};

model.model.model.bottom.type = "Dirichlet";
model.model.model.bottom.dispIncr = 0.;
model.model.model.bottom.nodeGroups = [ "bottom", "bottom", "bottom", "bottom", "bottom", "bottom" ];
model.model.model.bottom.dofs = [ "dx", "dy", "dz", "rx", "ry", "rz" ];
model.model.model.bottom.factors = [ 0., 0., 0., 0., 0., 0. ];

model.model.model.top.type = "Dirichlet";
model.model.model.top.dispIncr = 0.;
model.model.model.top.nodeGroups = [ "top" ];
model.model.model.top.dofs = [ "dz" ];
model.model.model.top.factors = [ 0. ];

model.model.model.impact.type = "InitLoad";
model.model.model.impact.veloGroups = [ "top" ];
model.model.model.impact.veloDofs = [ "dz" ];
model.model.model.impact.veloVals = [ -20. ];

Output.modules = [ "resp" ];
Output.resp.type = "Sample";
Output.resp.file = "$(CASE_NAME)/resp.csv";
Output.resp.dataSets = [ "i", "t", "bottom.resp.dz", "top.resp.dz", "dissipatedEnergy" ];
Output.resp.separator = ",";
"""


if __name__ == "__main__":
    with open(os.path.join(HERE, "crush.geo"), "w") as f:
        f.write(geo())
    with open(os.path.join(HERE, "crush.pro"), "w") as f:
        f.write(pro())
