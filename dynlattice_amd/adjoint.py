"""Design gradients through the explicit lattice dynamics (torch.autograd).

:class:`ImpactSimulation` owns one device context (:class:`~.gpu.GpuModel`)
for a lattice problem.  :class:`LatticeResponse` is a
``torch.autograd.Function`` that maps the per-strut section data and contact
radii to the reaction history of the top plate, [nsteps, n_groups]:

* forward  = ``dl_adjoint_forward``: the explicit steps on the GPU, with a
  checkpoint of the state every ``checkpoint_every`` steps;
* backward = ``dl_adjoint_backward``: the hand-written reverse sweep
  (node update, element internal force, contact pair forces), recomputing
  each segment from its checkpoint.

Any objective composed in torch from the reaction history (absorbed energy,
peak force, ...) is then differentiable with respect to e.g. strut radii,
through :func:`circle_sections` (radius -> area, moments) also written in
torch.  With the elastoplastic material the yield weights are an input too
(:func:`capacity_weights`: radius and yield stress -> the resultant
capacities' weights), and the reverse sweep runs through the return map.  Nothing here computes on the CPU beyond those per-strut maps: the
dynamics and the adjoint run only in libdlgpu.so.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import lattice as L
from .gpu import SECTION_FIELDS, GpuModel

GRAD_FIELDS = SECTION_FIELDS + ("contact_radius",)


def circle_sections(radius, young, poisson, density):
    """Solid circular sections as torch tensors (the map of
    :meth:`lattice.Sections.circle`): A = pi r^2, I1 = I2 = pi r^4 / 4,
    J = I1 + I2, kappa = 9/10, G = E / 2 / (nu + 1)."""
    r = radius
    young = torch.as_tensor(young, dtype=r.dtype).expand_as(r)
    density = torch.as_tensor(density, dtype=r.dtype).expand_as(r)
    A = math.pi * r ** 2
    inertia = math.pi * r ** 4 / 4.0
    return {"area": A, "I1": inertia, "I2": inertia, "J": inertia + inertia,
            "shear_factor": torch.full_like(r, 9.0 / 10.0), "young": young,
            "shear_modulus": young / 2.0 / (poisson + 1.0), "density": density}


def capacity_weights(radius, sigma_y):
    """Yield weights of solid circles from their plastic capacities (the map
    of :meth:`lattice.Plastic.resultant_capacities`) as a torch tensor
    [n_struts, 6]: 1/V_p, 1/V_p, 1/N_p, 1/M_p, 1/M_p, 1/T_p with
    N_p = s_y A, V_p = s_y A / sqrt3, M_p = s_y 4 r^3 / 3,
    T_p = s_y 2 pi r^3 / (3 sqrt3)."""
    r = radius
    sy = torch.as_tensor(sigma_y, dtype=r.dtype).expand_as(r)
    A = math.pi * r ** 2
    Np = sy * A
    Vp = sy * A / math.sqrt(3.0)
    Mp = sy * 4.0 * r ** 3 / 3.0
    Tp = sy * 2.0 * math.pi * r ** 3 / (3.0 * math.sqrt(3.0))
    return torch.stack([1 / Vp, 1 / Vp, 1 / Np, 1 / Mp, 1 / Mp, 1 / Tp], dim=1)


class ImpactSimulation:
    """A lattice problem on the GPU, re-run from its initial state (zero
    plastic state for the elastoplastic material) for every set of section
    data (SO(3) update)."""

    def __init__(self, problem: L.Problem, nsteps: int, checkpoint_every: int = 100,
                 group: str = "top", dof: int = 2, device: int = 0, max_pairs: int = 0):
        self.p = problem
        self.nsteps = int(nsteps)
        self.every = int(checkpoint_every)
        # max_pairs: contact candidate-list capacity (0: the library default)
        self.model = GpuModel(problem, device=device, init=False, max_pairs=max_pairs)
        self.model.set_reaction_group(problem.groups[group], dof)

    def run(self, sections: dict, contact_radius=None, yield_weights=None) -> np.ndarray:
        """reaction history [nsteps, 1] for the given per-strut arrays (and
        yield weights [n_struts, 6] of the elastoplastic material; None keeps
        the problem's)"""
        sec = L.Sections(**{f: np.asarray(sections[f], np.float64) for f in SECTION_FIELDS},
                         inertia_correct=self.p.sections.inertia_correct)
        m = self.model
        m.set_sections(sec, contact_radius)
        if self.p.plastic is not None:
            w = self.p.plastic.weights if yield_weights is None else yield_weights
            m.set_yield_weights(np.asarray(w, np.float64))
            E = self.p.n_elems
            m.set_plastic_state(np.zeros((E, 6)), np.zeros((E, 7)))
        m.set_state(self.p.disp0, self.p.velo0)
        m.init()
        return m.adjoint_forward(self.nsteps, self.p.dt, self.every)

    def gradients(self, reaction_bar: np.ndarray) -> dict:
        return self.model.adjoint_backward(reaction_bar)

    def close(self):
        self.model.close()


class LatticeResponse(torch.autograd.Function):
    """reaction = LatticeResponse.apply(sim, area, I1, I2, J, shear_factor,
    young, shear_modulus, density, contact_radius[, yield_weight]) ->
    [nsteps, n_groups] (float64, CPU).  contact_radius may be None for a
    problem without contact; yield_weight [n_struts, 6] (elastoplastic
    material) may be None to keep the problem's."""

    @staticmethod
    def forward(ctx, sim, area, I1, I2, J, shear_factor, young, shear_modulus, density,
                contact_radius, yield_weight=None):
        vals = dict(zip(SECTION_FIELDS, (area, I1, I2, J, shear_factor, young, shear_modulus,
                                          density)))
        sec = {k: v.detach().cpu().numpy().astype(np.float64) for k, v in vals.items()}
        cr = None if contact_radius is None else contact_radius.detach().cpu().numpy()
        yw = None if yield_weight is None else yield_weight.detach().cpu().numpy()
        reaction = sim.run(sec, cr, yw)
        ctx.sim = sim
        ctx.has_cr = contact_radius is not None
        ctx.has_yw = yield_weight is not None
        return torch.from_numpy(reaction)

    @staticmethod
    def backward(ctx, grad_reaction):
        g = ctx.sim.gradients(grad_reaction.detach().cpu().numpy().astype(np.float64))
        out = [torch.from_numpy(g[f]) for f in SECTION_FIELDS]
        out.append(torch.from_numpy(g["contact_radius"]) if ctx.has_cr else None)
        out.append(torch.from_numpy(g["yield_weight"]) if ctx.has_yw else None)
        return (None, *out)


def response_from_radius(sim: ImpactSimulation, radius: torch.Tensor, young, poisson, density,
                         contact_from_radius: bool = True, sigma_y=None) -> torch.Tensor:
    """reaction history as a differentiable function of per-strut radii (the
    contact radius follows the section radius when contact_from_radius; with
    the elastoplastic material and sigma_y given, the yield weights follow
    the radius and sigma_y -- a tensor that may require grad too -- through
    :func:`capacity_weights`)"""
    s = circle_sections(radius, young, poisson, density)
    cr = radius if (contact_from_radius and sim.p.contact is not None) else None
    yw = None
    if sim.p.plastic is not None and sigma_y is not None:
        yw = capacity_weights(radius, sigma_y)
    return LatticeResponse.apply(sim, *[s[f] for f in SECTION_FIELDS], cr, yw)


def impact_objective(reaction: torch.Tensor, dt: float, v_plate: float,
                     peak_weight: float = 1.0) -> torch.Tensor:
    """Energy absorbed through the moving plate plus the peak plate force:
    J = sum_n dt v F_n + w max_n F_n, where F_n = reaction_n * sign(v) is the
    force the plate pushes with and reaction_n is the internal-force sum of
    the plate nodes along the plate's motion (first group)."""
    F = reaction[:, 0] * float(np.sign(v_plate))
    return dt * abs(v_plate) * F.sum() + peak_weight * F.max()
