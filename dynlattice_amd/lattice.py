"""Lattice and rod generators: the mesh-side data format of the hot path.

The reference reads a GMSH ``.geo`` file and turns every curve into one
element group ``beam_<k>`` ("one curve = one strut",
src/modules/GMSHInputModule.cpp:308-415); a LatticeModel then creates one
rod model per group (src/models/LatticeModel.cpp:61-74).  This module
produces the same structure directly as arrays:

* ``coords``      [N, 3] float64   node coordinates
* ``elem_nodes``  [E, 2] int32     2-node line elements (Line2)
* ``elem_strut``  [E]    int32     strut (element group) of each element;
  the elements of one strut are contiguous and ordered along the strut
* per-strut section arrays (area, moments, moduli, density)

Random per-strut radii come from a counter-based splitmix64 stream with a
Box-Muller transform, so every seed reproduces the same lattice on every
machine.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n outputs of splitmix64 started at ``seed`` (state advanced by gamma)."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int) -> np.ndarray:
    """Uniform doubles in (0, 1): 53 high bits of splitmix64, half-ulp offset."""
    z = splitmix64(seed, n)
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (2.0 ** -53)


def normal(seed: int, n: int) -> np.ndarray:
    """Standard normal deviates by Box-Muller on consecutive uniform pairs."""
    u = uniform01(seed, 2 * n)
    u1, u2 = u[0::2], u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


@dataclass
class Sections:
    """Per-strut cross-section and elastic constants (ElasticRodMaterial.cpp:65-228)."""
    area: np.ndarray
    I1: np.ndarray
    I2: np.ndarray
    J: np.ndarray
    shear_factor: np.ndarray
    young: np.ndarray
    shear_modulus: np.ndarray
    density: np.ndarray
    inertia_correct: float = 1.0

    @staticmethod
    def circle(radius, young, poisson, density, n=None):
        """Solid circle: A = pi r^2, I = pi r^4/4, J = I1 + I2, kappa = 9/10
        (ElasticRodMaterial.cpp:158-164,183), G = E / 2 / (nu + 1) (:102)."""
        r = np.asarray(radius, dtype=np.float64)
        if n is not None and r.ndim == 0:
            r = np.full(n, float(r))
        R = r.shape[0]
        E = np.broadcast_to(np.asarray(young, np.float64), (R,)).copy()
        nu = np.broadcast_to(np.asarray(poisson, np.float64), (R,))
        rho = np.broadcast_to(np.asarray(density, np.float64), (R,)).copy()
        A = math.pi * r ** 2
        I = math.pi * r ** 4 / 4.0
        return Sections(area=A, I1=I.copy(), I2=I.copy(), J=I + I,
                        shear_factor=np.full(R, 9.0 / 10.0),
                        young=E, shear_modulus=E / 2.0 / (nu + 1.0), density=rho)


@dataclass
class Plastic:
    """Power-sum yield family  f = sum_i |w_i (s_i + ck_i q_i)|^p_i - Y (1 + ci q_0),
    q = -H alpha; p_i = 1 is the sum-abs shape of tests/plastic/test1.pro:22-28,
    p_i ~ 2 the shape of tests/plastic/test3.pro:22-28.  ``weights``
    is [R, 6]; component order (dx, dy, dz, rx, ry, rz) = (V1, V2, N, M1, M2, T)
    in the material frame (rod axis = material z)."""
    weights: np.ndarray
    yield_level: float = 1.0
    kin_coupling: tuple = (1.0,) * 6
    iso_coupling: float = 0.0
    has_iso: bool = False
    iso_modulus: float = 0.0
    has_kin: bool = True
    kin_const: tuple = (0.0,) * 6
    kin_ratio: float = 0.01
    precision: float = 1e-5
    max_iter: int = 20
    exponent: tuple = (1.0,) * 6

    @staticmethod
    def resultant_capacities(radius, sigma_y, kin_ratio=0.01, exponent=1.0):
        """Plastic capacities of a solid circle: N_p = s_y A, V_p = s_y A/sqrt3,
        M_p = s_y 4 r^3/3, T_p = s_y 2 pi r^3/(3 sqrt3); H_kin = kin_ratio * C.
        exponent 1: sum-abs interaction; 2: quadratic interaction."""
        r = np.asarray(radius, np.float64)
        sy = np.broadcast_to(np.asarray(sigma_y, np.float64), r.shape)
        A = math.pi * r ** 2
        Np = sy * A
        Vp = sy * A / math.sqrt(3.0)
        Mp = sy * 4.0 * r ** 3 / 3.0
        Tp = sy * 2.0 * math.pi * r ** 3 / (3.0 * math.sqrt(3.0))
        w = np.stack([1 / Vp, 1 / Vp, 1 / Np, 1 / Mp, 1 / Mp, 1 / Tp], axis=1)
        return Plastic(weights=np.ascontiguousarray(w), kin_ratio=kin_ratio,
                       exponent=(float(exponent),) * 6)


@dataclass
class Contact:
    """Rod-rod penalty contact (RodContactModel.cpp:60-98).  ``radius`` is per
    strut; the pair threshold is r_A + r_B (2 r for the reference's scalar r)."""
    radius: np.ndarray
    penalty_sts: float = 1e6
    penalty_nts: float = 1e6
    skin: float = 0.0   # step candidate-list skin (0: half the largest radius, < 0: off)


@dataclass
class Program:
    """Time program of boundary values: the value of a dof that follows it
    is (its cval / fext entry) x scale, evaluated at the end time of every
    step.  kind "const": scale = a; "step": scale = a + b (k + c), k = steps
    taken (Neumann initLoad/loadIncr: c = 1; Dirichlet initDisp/dispIncr:
    c = 0); "expr": scale = expr(t, i) in JEM function syntax (LoadScale
    scaleFunc)."""
    kind: str = "const"
    a: float = 1.0
    b: float = 0.0
    c: float = 0.0
    expr: str = ""

    KINDS = {"const": 0, "step": 1, "expr": 2}


def add_program(p: "Problem", prog: Program, nodes, dofs) -> int:
    """attaches prog to the given nodes' dofs; returns its index"""
    if p.dof_prog is None:
        p.dof_prog = np.full((p.n_nodes, 6), -1, np.int32)
    p.programs.append(prog)
    q = len(p.programs) - 1
    for d in np.atleast_1d(dofs):
        p.dof_prog[np.asarray(nodes), int(d)] = q
    return q


@dataclass
class Problem:
    """One explicit-dynamics case: mesh, materials, contact, boundary data."""
    coords: np.ndarray
    elem_nodes: np.ndarray
    elem_strut: np.ndarray
    n_struts: int
    sections: Sections
    plastic: Optional[Plastic] = None
    contact: Optional[Contact] = None
    so3: bool = True
    # boundary data, DOF-major per node: [N, 6]
    cmask: Optional[np.ndarray] = None   # uint8, 1 = constrained dof
    cval: Optional[np.ndarray] = None    # prescribed acceleration
    fext: Optional[np.ndarray] = None    # external nodal force
    disp0: Optional[np.ndarray] = None   # [N, 6]
    velo0: Optional[np.ndarray] = None   # [N, 6]
    dt: float = 0.0
    groups: Dict[str, np.ndarray] = field(default_factory=dict)
    name: str = ""
    # time programs of the boundary values (see Program); dof_prog [N, 6]
    # holds each dof's program index or -1
    programs: list = field(default_factory=list)
    dof_prog: Optional[np.ndarray] = None
    t0: float = 0.0
    # point masses per node (MassModel): translational mass, rotational inertia
    point_mass: Optional[np.ndarray] = None
    point_inertia: Optional[np.ndarray] = None
    # ImpactBC groups: (nodes, dof, weight)
    impact: list = field(default_factory=list)

    @property
    def n_nodes(self):
        return int(self.coords.shape[0])

    @property
    def n_elems(self):
        return int(self.elem_nodes.shape[0])

    def zeros6(self):
        return np.zeros((self.n_nodes, 6), dtype=np.float64)

    def finalize(self):
        N = self.n_nodes
        if self.cmask is None:
            self.cmask = np.zeros((N, 6), dtype=np.uint8)
        if self.cval is None:
            self.cval = np.zeros((N, 6))
        if self.fext is None:
            self.fext = np.zeros((N, 6))
        if self.disp0 is None:
            self.disp0 = np.zeros((N, 6))
        if self.velo0 is None:
            self.velo0 = np.zeros((N, 6))
        self.coords = np.ascontiguousarray(self.coords, dtype=np.float64)
        self.elem_nodes = np.ascontiguousarray(self.elem_nodes, dtype=np.int32)
        self.elem_strut = np.ascontiguousarray(self.elem_strut, dtype=np.int32)
        return self


def stable_dt(p: Problem, cfl: float = 0.5) -> float:
    """dt = cfl * min_e L_e / c_e with c = sqrt(E / rho): the axial CFL bound of
    a 2-node element with lumped mass (omega_max = 2 c / L)."""
    x = p.coords[p.elem_nodes[:, 1]] - p.coords[p.elem_nodes[:, 0]]
    L = np.sqrt((x * x).sum(1))
    s = p.elem_strut
    c = np.sqrt(p.sections.young[s] / p.sections.density[s])
    return float(cfl * np.min(L / c))


def straight_rod(n_elem=32, length=1.0, radius=0.01, young=210e9, poisson=0.3,
                 density=7850.0, direction=(1.0, 0.0, 0.0)) -> Problem:
    """Single straight rod from the origin along ``direction`` (one strut)."""
    d = np.asarray(direction, np.float64)
    d = d / np.linalg.norm(d)
    s = np.linspace(0.0, length, n_elem + 1)
    coords = s[:, None] * d[None, :]
    en = np.stack([np.arange(n_elem), np.arange(1, n_elem + 1)], axis=1).astype(np.int32)
    p = Problem(coords=coords, elem_nodes=en, elem_strut=np.zeros(n_elem, np.int32),
                n_struts=1, sections=Sections.circle(radius, young, poisson, density, n=1))
    p.groups["fixed"] = np.array([0], np.int32)
    p.groups["tip"] = np.array([n_elem], np.int32)
    return p.finalize()


def cantilever_c1() -> Problem:
    """Plain cantilever: 32 elements along x, L = 1 m, r = 0.01 m, steel;
    node 0 clamped, tip force (0, 0, -10 N) and torque 0.5 N m about x."""
    p = straight_rod()
    p.cmask[0, :] = 1
    p.fext[32, 2] = -10.0
    p.fext[32, 3] = 0.5
    p.dt = stable_dt(p, 0.5)
    p.name = "cantilever-32el"
    return p


def swinging_rod(n_elem: int = 10) -> Problem:
    """Swinging flexible rod under gravity (Lang, Linn, Arnold 2011,
    Example 1; the reference's tests/transient/test2.pro): 1 m along x,
    circle r = 5 mm, E = 5e6, nu = 0.4999999, shear correction 1,
    rho = 1100; translations of the node at x = 0 fixed, rotations free;
    gravity as nodal loads rho A L_e g along -z, half at both ends
    (the Neumann factors -1 on all nodes, +0.5 on the two ends); at rest at
    t = 0; first time step 1e-6 s (MilneDevice, adaptive)."""
    p = straight_rod(n_elem, 1.0, 5e-3, 5e6, 0.4999999, 1.1e3)
    p.sections.shear_factor[:] = 1.0
    fixed, free = int(p.groups["fixed"][0]), int(p.groups["tip"][0])
    p.cmask[fixed, :3] = 1
    w = math.pi * (5e-3) ** 2 * (1.0 / n_elem) * 1.1e3 * 9.81
    p.fext[:, 2] = -w
    p.fext[fixed, 2] += 0.5 * w
    p.fext[free, 2] += 0.5 * w
    p.dt = 1e-6
    p.name = f"swinging-rod-{n_elem}el"
    return p


def bcc_lattice(n: int, m: int, cell: float, radii: np.ndarray, young, poisson,
                density) -> Problem:
    """BCC lattice of n^3 cells; 8 struts per cell from the cell centre to its
    corners, m elements per strut.

    Node numbering: corner joints ((n+1)^3, x fastest), then cell centres
    (n^3), then the m-1 interior nodes of every strut, strut by strut.  Strut
    s = 8 * cell + corner; its elements run from the centre outwards.
    E = 8 m n^3, N = (n+1)^3 + n^3 + 8 (m-1) n^3."""
    n1 = n + 1
    nc = n1 ** 3
    ncell = n ** 3
    R = 8 * ncell
    g = np.arange(n1, dtype=np.float64) * cell
    cz, cy, cx = np.meshgrid(g, g, g, indexing="ij")
    corners = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1)
    h = (np.arange(n, dtype=np.float64) + 0.5) * cell
    kz, ky, kx = np.meshgrid(h, h, h, indexing="ij")
    centres = np.stack([kx.ravel(), ky.ravel(), kz.ravel()], axis=1)
    iz, iy, ix = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
    offs = np.array([[dx, dy, dz] for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)])
    # corner joint index of strut (cell, k)
    cj = ((iz[:, None] + offs[None, :, 2]) * n1 + (iy[:, None] + offs[None, :, 1])) * n1 \
        + (ix[:, None] + offs[None, :, 0])
    cj = cj.reshape(-1)                      # [R]
    centre_idx = nc + np.repeat(np.arange(ncell), 8)   # [R]
    a = centres[np.repeat(np.arange(ncell), 8)]
    b = corners[cj]
    t = np.arange(1, m, dtype=np.float64) / m           # interior parameters
    interior = a[:, None, :] + t[None, :, None] * (b - a)[:, None, :]   # [R, m-1, 3]
    coords = np.concatenate([corners, centres, interior.reshape(-1, 3)], axis=0)
    base = nc + ncell
    chain = np.empty((R, m + 1), dtype=np.int64)
    chain[:, 0] = centre_idx
    chain[:, m] = cj
    if m > 1:
        chain[:, 1:m] = base + np.arange(R)[:, None] * (m - 1) + np.arange(m - 1)[None, :]
    en = np.stack([chain[:, :-1].reshape(-1), chain[:, 1:].reshape(-1)], axis=1).astype(np.int32)
    es = np.repeat(np.arange(R, dtype=np.int32), m)
    secs = Sections.circle(np.asarray(radii, np.float64), young, poisson, density)
    p = Problem(coords=coords, elem_nodes=en, elem_strut=es, n_struts=R, sections=secs)
    H = n * cell
    z = coords[:, 2]
    tol = 1e-9 * cell
    p.groups["bottom"] = np.nonzero(np.abs(z) < tol)[0].astype(np.int32)
    p.groups["top"] = np.nonzero(np.abs(z - H) < tol)[0].astype(np.int32)
    return p.finalize()


def crush_bcs(p: Problem, v_top: float):
    """Bottom face clamped (a = v = 0 on all DOFs); top face moves with the
    prescribed velocity v_top along z (initial velocity, a_z = 0)."""
    b, t = p.groups["bottom"], p.groups["top"]
    p.cmask[b, :] = 1
    p.cval[b, :] = 0.0
    p.cmask[t, 2] = 1
    p.cval[t, 2] = 0.0
    p.velo0[t, 2] = v_top
    return p


def bcc_c2() -> Problem:
    """Elastic BCC impact: 10^3 cells of 10 mm, 4 elements per strut,
    r = 0.5 mm (1 + 0.05 N(0,1)), seed 2, E = 2 GPa, nu = 0.3, rho = 1200,
    bottom clamped, top at -5 m/s, dt = 0.25 x the stable estimate.

    20,000 steps at that dt take the top face to 42 % nominal strain.  At
    0.5 x the estimate the same 20,000 steps would reach 84 %, and the
    explicit SO(3) update (the oracle's as much as the device's) overflows
    near step 13,500 (57 % strain) of this lattice without contact; at 0.25
    a 1e-15 relative perturbation of the initial velocities stays below
    3e-12 in the state and the reaction over the whole run."""
    n, m, cell = 10, 4, 10e-3
    R = 8 * n ** 3
    r = 0.5e-3 * (1.0 + 0.05 * normal(2, R))
    p = bcc_lattice(n, m, cell, r, 2e9, 0.3, 1200.0)
    crush_bcs(p, -5.0)
    p.dt = stable_dt(p, 0.25)
    p.name = "bcc-10x10x10-m4-elastic"
    return p


def bcc_c3(n: int = 40, contact: bool = True) -> Problem:
    """Elastoplastic BCC crush with contact: n^3 cells (n = 40: 2,048,000
    elements), 4 elements per strut, r = 0.4 mm exp(0.1 N(0,1) - 0.005)
    (lognormal, mean 0.4 mm), seed 3, E = 70 GPa, nu = 0.3, rho = 2700,
    yield 250 MPa (quadratic interaction of the resultants, p = 2), kinematic
    hardening H = C/100,
    penalty contact 1e6 N/m; top plate at -20 m/s.  dt = 0.25 x the axial
    CFL bound: at 0.5 x CFL the SO(3) explicit update of the reference
    algorithm diverges near 35 % nominal strain (nodal spins grow to 1e6
    rad/s and the return map stops converging; the oracle agrees with the
    device there), at 0.25 x CFL the crush runs to 40 %."""
    m, cell = 4, 10e-3
    R = 8 * n ** 3
    r = 0.4e-3 * np.exp(0.1 * normal(3, R) - 0.005)
    p = bcc_lattice(n, m, cell, r, 70e9, 0.3, 2700.0)
    p.plastic = Plastic.resultant_capacities(r, 250e6, kin_ratio=0.01, exponent=2.0)
    if contact:
        p.contact = Contact(radius=r.copy(), penalty_sts=1e6, penalty_nts=1e6)
    crush_bcs(p, -20.0)
    p.dt = stable_dt(p, 0.25)
    p.name = f"bcc-{n}x{n}x{n}-m4-elastoplastic-contact" if contact else \
        f"bcc-{n}x{n}x{n}-m4-elastoplastic"
    return p


def bcc_c5(n: int = 20, sigma_y=None) -> Problem:
    """Adjoint case: n^3 cells (n = 20: 256,000 elements), elastic with
    contact, r = 0.5 mm (1 + 0.1 U[-1, 1]), seed 5.  With sigma_y: the
    elastoplastic variant (quadratic interaction of the resultant
    capacities, kinematic hardening H = C/100)."""
    m, cell = 4, 10e-3
    R = 8 * n ** 3
    r = 0.5e-3 * (1.0 + 0.1 * (2.0 * uniform01(5, R) - 1.0))
    p = bcc_lattice(n, m, cell, r, 70e9, 0.3, 2700.0)
    p.contact = Contact(radius=r.copy(), penalty_sts=1e6, penalty_nts=1e6)
    crush_bcs(p, -20.0)
    p.dt = stable_dt(p, 0.5)
    p.name = f"bcc-{n}x{n}x{n}-m4-elastic-contact"
    if sigma_y is not None:
        p.plastic = Plastic.resultant_capacities(r, sigma_y, kin_ratio=0.01, exponent=2.0)
        p.name = f"bcc-{n}x{n}x{n}-m4-elastoplastic-contact"
    return p


def bcc_adjoint_check(radius=None, dt=None, sigma_y=None) -> Problem:
    """Small adjoint check case: 2^3 BCC cells of 10 mm, 3 elements per
    strut (64 struts, 192 elements), r = 1.2 mm (1 + 0.1 U[-1, 1]), seed 5
    (or the given per-strut radii), E = 70 GPa, nu = 0.3, rho = 2700,
    elastic, penalty contact 1e6 N/m with contact radius = strut radius,
    bottom clamped, top at -20 m/s.  Within 1600 steps about 30 contact pairs
    close.  With sigma_y (a scalar or per strut): elastoplastic, quadratic
    interaction of the resultant capacities, kinematic hardening H = C/100."""
    n, m, cell = 2, 3, 10e-3
    R = 8 * n ** 3
    r = 1.2e-3 * (1.0 + 0.1 * (2.0 * uniform01(5, R) - 1.0)) if radius is None else \
        np.asarray(radius, np.float64)
    p = bcc_lattice(n, m, cell, r, 70e9, 0.3, 2700.0)
    p.contact = Contact(radius=r.copy(), penalty_sts=1e6, penalty_nts=1e6)
    crush_bcs(p, -20.0)
    p.dt = stable_dt(p, 0.5) if dt is None else dt
    p.name = "bcc-2x2x2-m3-elastic-contact"
    if sigma_y is not None:
        p.plastic = Plastic.resultant_capacities(r, sigma_y, kin_ratio=0.01, exponent=2.0)
        p.name = "bcc-2x2x2-m3-elastoplastic-contact"
    return p


C4_STEPS = 10_000   # steps of every design-sweep instance
C4_STRAIN = 0.2     # nominal strain every instance reaches in those steps


def c4_instance(i: int) -> Problem:
    """Design-sweep instance i (seed 1000 + i): n ~ U{4..16}, m ~ U{2..6},
    r ~ U[0.2, 0.8] mm, yield ~ U[150, 350] MPa; elastoplastic BCC crush at
    its own stable dt (0.5 x the estimate), the top face moving so that the
    C4_STEPS steps take every instance to C4_STRAIN nominal strain (a fixed
    plate speed would crush the small lattices flat long before 10,000 steps
    and leave the large ones barely loaded)."""
    u = uniform01(1000 + i, 4)
    n = 4 + int(u[0] * 13)
    m = 2 + int(u[1] * 5)
    r = 0.2e-3 + 0.6e-3 * u[2]
    sy = 150e6 + 200e6 * u[3]
    R = 8 * n ** 3
    cell = 10e-3
    p = bcc_lattice(n, m, cell, np.full(R, r), 70e9, 0.3, 2700.0)
    p.plastic = Plastic.resultant_capacities(np.full(R, r), sy, kin_ratio=0.01, exponent=2.0)
    p.dt = stable_dt(p, 0.5)
    crush_bcs(p, -C4_STRAIN * n * cell / (C4_STEPS * p.dt))
    p.name = f"sweep-{i}-n{n}-m{m}"
    return p
