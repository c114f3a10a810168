"""Batched design sweeps: many independent lattices in one device context.

The reference runs one model per process.  A design sweep (many lattices of
different size, strut count and material) is expressed here as ONE context
whose mesh is the disjoint union of all instances:

* per-instance time steps via a per-node dt (each instance keeps its own
  stable dt, dl_set_node_dt);
* one reaction group per instance (dl_set_reaction_groups);
* instances are translated apart along x, so that even with contact enabled
  no two instances can touch;
* the step sequence is replayed from captured hipGraphs (dl_set_graph); the
  graph length, the order of the instances in the union and the assembly
  layout are chosen by :func:`autotune` on the real workload.

A persistent kernel looping over the steps with grid-wide barriers between
the element and joint phases is not used: a cooperative grid barrier costs
7.5 us with 64 resident blocks, 25 us with 256 and 78 us with 768 on an
MI355X, against about 1.5 us per kernel node of a replayed graph.

Every instance computes exactly what it would compute alone: the kernels
never mix data of two instances (nodes, elements and struts are disjoint).
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import List

import numpy as np

from . import lattice as L


@dataclass
class Batch:
    problem: L.Problem
    node_off: np.ndarray     # [B+1]
    elem_off: np.ndarray     # [B+1]
    strut_off: np.ndarray    # [B+1]
    node_dt: np.ndarray      # [N]
    group_nodes: np.ndarray  # reaction nodes, concatenated by instance
    group_ptr: np.ndarray    # [B+1]
    dts: np.ndarray          # [B]


def x_positions(problems: List[L.Problem], gap: float = 0.1) -> np.ndarray:
    """x at which each instance starts when placed side by side in the given
    order, gap apart"""
    x, out = 0.0, []
    for p in problems:
        out.append(x)
        x += p.coords[:, 0].max() - p.coords[:, 0].min() + gap
    return np.array(out)


def concat(problems: List[L.Problem], gap: float = 0.1, group: str = "bottom",
           x0s=None) -> Batch:
    """Disjoint union of problems (all elastic or all elastoplastic with the
    same global yield constants; per-strut data may differ).  x0s: where each
    instance starts along x (default: side by side in this order); giving
    every instance the same place whatever the order of the union keeps its
    coordinates, hence its results, the same bits."""
    B = len(problems)
    nn = np.array([p.n_nodes for p in problems])
    ne = np.array([p.n_elems for p in problems])
    ns = np.array([p.n_struts for p in problems])
    node_off = np.r_[0, np.cumsum(nn)]
    elem_off = np.r_[0, np.cumsum(ne)]
    strut_off = np.r_[0, np.cumsum(ns)]
    coords, en, es = [], [], []
    xs = x_positions(problems, gap) if x0s is None else np.asarray(x0s, np.float64)
    for b, p in enumerate(problems):
        c = p.coords.copy()
        c[:, 0] += xs[b] - c[:, 0].min()
        coords.append(c)
        en.append(p.elem_nodes + node_off[b])
        es.append(p.elem_strut + strut_off[b])
    fields = ("area", "I1", "I2", "J", "shear_factor", "young", "shear_modulus", "density")
    sec = L.Sections(*[np.concatenate([getattr(p.sections, f) for p in problems]) for f in fields],
                     inertia_correct=problems[0].sections.inertia_correct)
    P = L.Problem(coords=np.concatenate(coords), elem_nodes=np.concatenate(en),
                  elem_strut=np.concatenate(es), n_struts=int(strut_off[-1]), sections=sec,
                  so3=problems[0].so3)
    P.finalize()
    for name in ("cmask", "cval", "fext", "disp0", "velo0"):
        setattr(P, name, np.concatenate([getattr(p, name) for p in problems]))
    if problems[0].plastic is not None:
        q = problems[0].plastic
        w = np.concatenate([p.plastic.weights for p in problems])
        P.plastic = L.Plastic(weights=w, yield_level=q.yield_level, kin_coupling=q.kin_coupling,
                              iso_coupling=q.iso_coupling, has_iso=q.has_iso,
                              iso_modulus=q.iso_modulus, has_kin=q.has_kin,
                              kin_const=q.kin_const, kin_ratio=q.kin_ratio,
                              precision=q.precision, max_iter=q.max_iter, exponent=q.exponent)
    if problems[0].contact is not None:
        r = np.concatenate([p.contact.radius for p in problems])
        P.contact = L.Contact(radius=r, penalty_sts=problems[0].contact.penalty_sts,
                              penalty_nts=problems[0].contact.penalty_nts)
    dts = np.array([p.dt for p in problems])
    node_dt = np.repeat(dts, nn)
    gn = [p.groups[group] + node_off[b] for b, p in enumerate(problems)]
    gp = np.r_[0, np.cumsum([len(g) for g in gn])].astype(np.int32)
    P.dt = float(dts.min())
    P.name = f"batch-{B}"
    return Batch(P, node_off, elem_off, strut_off, node_dt, np.concatenate(gn).astype(np.int32),
                 gp, dts)


def make_model(batch: Batch, flags=0, graph_steps=0):
    """GpuModel over the batch with per-node dt and per-instance reaction groups."""
    from .gpu import GpuModel
    g = GpuModel(batch.problem, flags=flags)
    g.set_node_dt(batch.node_dt)
    g.set_reaction_groups(batch.group_nodes, batch.group_ptr, 2)
    if graph_steps:
        g.set_graph(graph_steps)
    return g


def autotune_graph(model, candidates=(0, 8, 32, 128), steps=256, dt=None):
    """Pick the captured-graph length with the lowest measured time per step.

    Each candidate runs ``steps`` steps once to warm up (and capture), then
    once timed; the state advances during tuning, which is fine for a
    throughput choice.  Returns (best, {candidate: seconds_per_step})."""
    dt = model.p.dt if dt is None else dt
    times = {}
    for k in candidates:
        model.set_graph(k)
        model.step(steps, dt, out=False)
        model.sync()
        t0 = time.perf_counter()
        model.step(steps, dt, out=False)
        model.sync()
        times[k] = (time.perf_counter() - t0) / steps
    best = min(times, key=times.get)
    model.set_graph(best)
    return best, times


# orders of the instances in the union that autotune tries: as given, by
# decreasing size, and grouped by elements per strut (chains of one length
# pack the same way into 64-lane wavefronts) and then by size
ORDERS = ("given", "size-descending", "chain-length")


def ordered_batch(problems: List[L.Problem], order: str) -> Batch:
    """the union in the given order; every instance at its place of the
    given order (so its results do not depend on the order)"""
    perm = instance_order(problems, order)
    xs = x_positions(problems)
    return concat([problems[i] for i in perm], x0s=xs[perm])


def instance_order(problems: List[L.Problem], order: str) -> np.ndarray:
    n = len(problems)
    if order == "given":
        return np.arange(n)
    size = np.array([p.n_elems for p in problems])
    if order == "size-descending":
        return np.argsort(-size, kind="stable")
    if order == "chain-length":
        per = np.array([p.n_elems // max(p.n_struts, 1) for p in problems])
        return np.lexsort((-size, per))
    raise ValueError(f"unknown order {order}")


# assembly layouts that autotune tries (dl_options.flags; every one gives the
# same bits): joints whose chains all sit in one wavefront summed in LDS
# inside the strut kernel, or every shared node through the slot buffer and
# the joint kernel; inner elements' end-b frames from the next lane, or both
# end frames read per element
LAYOUTS = {"wavefront-joints": 0, "joints-in-memory": 4, "frames-per-element": 2}


def _time_graphs(g, graph_lengths, steps, label, times, log):
    for k in graph_lengths:
        g.set_graph(k)
        g.step(max(steps, k), out=False)
        g.sync()
        t0 = time.perf_counter()
        g.step(max(steps, k), out=False)
        g.sync()
        times[label + (k,)] = (time.perf_counter() - t0) / max(steps, k)
        if log:
            print(f"autotune: {' '.join(str(x) for x in label)}, graph {k}: "
                  f"{1e3 * times[label + (k,)]:.4f} ms/step", file=log, flush=True)


def autotune(problems: List[L.Problem], orders=ORDERS, graph_lengths=(0, 10, 100), steps=100,
             flags=0, log=None, layouts=tuple(LAYOUTS)):
    """Schedule of a batched sweep, chosen by measurement in two stages.
    (1) For every order of the instances in the union (a different mesh
    numbering, hence lane packing, memory layout and block -> instance
    mapping) and every captured-graph length, ``steps`` steps after one
    warm-up pass of the same length (which also captures the graph).
    (2) At the best order, every assembly layout of :data:`LAYOUTS` with
    every graph length.  The instances advance during tuning, from their
    initial state for every context.  Returns (order, graph length, layout,
    {(order, layout, length): seconds per step})."""
    times = {}
    base = layouts[0]
    for order in orders:
        bt = ordered_batch(problems, order)
        g = make_model(bt, flags=flags | LAYOUTS[base])
        _time_graphs(g, graph_lengths, steps, (order, base), times, log)
        g.sync()  # raises a pending device error
        g.close()
    best_order = min(times, key=times.get)[0]
    bt = ordered_batch(problems, best_order)
    for lay in layouts[1:]:
        g = make_model(bt, flags=flags | LAYOUTS[lay])
        _time_graphs(g, graph_lengths, steps, (best_order, lay), times, log)
        g.sync()
        g.close()
    best = min(times, key=times.get)
    return best[0], best[2], best[1], times


def c4_batch(n_instances=256):
    """The design sweep of 256 seeded elastoplastic BCC lattices."""
    return concat([L.c4_instance(i) for i in range(n_instances)])
