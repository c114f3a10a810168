#!/usr/bin/env python3
"""Benchmark of the explicit rod-lattice step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A "step" is one LeapFrog explicit step (element Gauss-point loop, material
update, nodal assembly, central-difference SO(3) update[, contact]) over the
whole lattice, with all state resident in HBM.  The metric is element-GP
updates per second = elements x steps / wall time of the timed region.

Prints one JSON line with the metric, a roofline object for the dominant
kernel (k_strut; achieved = algorithmic bytes per launch / average launch
duration from HIP events on the library's stream) and the CPU baseline (the
oracle's C restatement timed on a bounded sample on this host).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 vector
# algorithmic bytes (per element-GP update / per node), FP64, each array once per step
B_ELEM = 136.0              # conn 8 + LambdaN 64 + eps0 48 + L0 8 + section 8
B_ELEM_PLASTIC = 208.0      # eps_p, alpha(6) read+write + 16
B_NODE = 248.0              # read x, theta, v, M^-1; write x, theta, v


def workload(name):
    from dynlattice_amd import lattice as L
    if name == "bcc10-elastic":
        return L.bcc_c2()
    if name == "bcc40-elastoplastic-contact":
        return L.bcc_c3(40, contact=True)
    if name == "bcc40-elastoplastic":
        return L.bcc_c3(40, contact=False)
    if name == "bcc20-elastic-contact":
        return L.bcc_c5(20)
    if name.startswith("bcc") and name.endswith("-elastic"):
        n = int(name[3:-8])
        from dynlattice_amd.lattice import bcc_lattice, crush_bcs, normal, stable_dt
        R = 8 * n ** 3
        p = bcc_lattice(n, 4, 10e-3, 0.5e-3 * (1.0 + 0.05 * normal(2, R)), 2e9, 0.3, 1200.0)
        crush_bcs(p, -5.0)
        p.dt = stable_dt(p, 0.5)
        p.name = name
        return p
    raise SystemExit(f"unknown workload {name}")


def yield_family(p):
    """the yield condition's family: power-sum of the weighted resultants"""
    if p.plastic is None:
        return None
    e = set(float(x) for x in p.plastic.exponent)
    if e == {1.0}:
        return "sum-abs resultants (p=1)"
    if e == {2.0}:
        return "quadratic resultant interaction (p=2)"
    return "power-sum resultants (p=" + ",".join(f"{x:g}" for x in p.plastic.exponent) + ")"


def private_node_count(p):
    """nodes held by exactly one strut-chain position (updated inside k_strut)"""
    # chain positions: element starts + last node of every strut
    last = np.r_[p.elem_strut[1:] != p.elem_strut[:-1], True]
    cnt = np.bincount(p.elem_nodes[:, 0], minlength=p.n_nodes) + \
        np.bincount(p.elem_nodes[last, 1], minlength=p.n_nodes)
    return int((cnt == 1).sum())


def cpu_threads():
    """host threads the all-cores oracle uses (OpenMP's OMP_NUM_THREADS)"""
    try:
        return int(os.environ["OMP_NUM_THREADS"])
    except (KeyError, ValueError):
        return os.cpu_count() or 1


def cpu_baseline(p, budget_s, steps_gpu, omp, snap=None):
    """The oracle (C restatement, oracle/ref_cpu.c) on this host, bounded
    sample: the same lattice, K steps with K chosen so the timed part takes
    about budget_s.  omp: the all-cores build (struts and pairs threaded
    with OpenMP) on cpu_threads() threads; else the serial build.  snap: a
    state snapshot (displacements, velocities, plastic state) the oracle
    starts from, i.e. the state the GPU steps were timed at."""
    t_all = time.perf_counter()
    ref = oracle_from(p, snap, omp)
    state = "the start of the run" if snap is None else "the GPU's timed state"
    t0 = time.perf_counter()
    ref.leapfrog(1)
    t1 = time.perf_counter() - t0
    k = int(max(1, min(steps_gpu, budget_s / max(t1, 1e-9))))
    t0 = time.perf_counter()
    ref.leapfrog(k)
    dt = time.perf_counter() - t0
    cores = cpu_threads() if omp else 1
    build = "OpenMP over struts and contact pairs" if omp else "1 thread"
    return {"value": p.n_elems * k / dt, "unit": "element-GP updates/s", "cores": cores,
            "kind": "port", "wall_s": round(time.perf_counter() - t_all, 2),
            "sample": f"oracle (oracle/ref_cpu.c, gcc -O3 -ffp-contract=off, {build}, "
                      f"{cores} thread(s)) on the same {p.name} lattice from {state}, "
                      f"{k} LeapFrog steps after 1 warm-up step, {dt:.1f} s"
                      + ("; contact broadphase on the oracle's uniform grid (candidate sets "
                         "equal to the reference's all-pairs loops, which cannot run at this "
                         "size)" if p.contact is not None else "")}


PROFILE = os.path.join(ROOT, "profiles", "bcc40_elastoplastic_contact.md")


def profiled_counters(workload):
    """Figures of the k_strut launch that rocprofv3 --pmc runs measured and
    the profile summary of this workload records: HBM bytes per launch
    (FETCH_SIZE, WRITE_SIZE, scaled by the copy calibration) as "traffic",
    FP64 operations per launch (the FP64 VALU instruction counters) as
    "fp64_flops"; {} when there is no summary."""
    out = {}
    if workload != "bcc-40x40x40-m4-elastoplastic-contact" or not os.path.exists(PROFILE):
        return out
    for line in open(PROFILE):
        if line.startswith("k_strut HBM bytes per launch:"):
            out["traffic"] = int(line.split(":")[1])
        if line.startswith("k_strut FP64 operations per launch:"):
            out["fp64_flops"] = float(line.split(":")[1])
    return out


def copy_bandwidth(nbytes=1 << 30, reps=5):
    """Device-to-device copy of nbytes on the library's device, timed with
    HIP events: bytes read + written per second (GB/s), best of reps."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    e0, e1 = C.c_void_p(), C.c_void_p()
    hip.hipEventCreate(C.byref(e0))
    hip.hipEventCreate(C.byref(e1))
    best = 0.0
    for _ in range(reps):
        hip.hipEventRecord(e0, None)
        assert hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), 3, None) == 0
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = C.c_float()
        hip.hipEventElapsedTime(C.byref(ms), e0, e1)
        best = max(best, 2.0 * nbytes / (ms.value * 1e-3) / 1e9)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    return best


def calibrate_copy(nbytes=1 << 30):
    """A device-to-device copy of known size (the runtime's copy kernel), used
    to scale the FETCH_SIZE / WRITE_SIZE counters of rocprofv3 --pmc runs."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # DeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"calibration_copy_bytes": nbytes, "copies": 3}), flush=True)


def sweep(args):
    print(json.dumps(sweep_measure(args.sweep_instances, args.sweep_steps,
                                   args.sweep_tune_steps, args.sweep_seq_steps, args.warmup,
                                   0.0 if args.no_cpu else args.cpu_budget)), flush=True)


def sweep_measure(n_inst=256, nsteps=10_000, tune_steps=50, seq_steps=200, warmup=20,
                  cpu_budget=15.0):
    """Configuration C4: the 256 seeded elastoplastic BCC lattices (per-instance
    dt, each crushed to 20 % nominal strain in its 10,000 steps) as one
    batched context whose steps replay captured hipGraphs, with the schedule
    (instance order x graph length, then the assembly layout) chosen by
    batch.autotune; nsteps steps of the sweep are timed from the initial
    state (10,000: the whole sweep).  Against it: the same instances one
    after another in their own contexts (seq_steps steps each), and, with a
    cpu_budget, the oracle on the median instance (all cores, one core)."""
    from dynlattice_amd import batch as BT
    from dynlattice_amd import lattice as L
    from dynlattice_amd.gpu import DL_TRACK_DISSIPATION, GpuModel
    probs = [L.c4_instance(i) for i in range(n_inst)]
    flags = DL_TRACK_DISSIPATION
    E = sum(p.n_elems for p in probs)
    print(f"sweep: {n_inst} instances, {E} elements", file=sys.stderr, flush=True)
    order, K, layout, tune = BT.autotune(probs, steps=tune_steps, flags=flags, log=sys.stderr)
    bt = BT.ordered_batch(probs, order)
    g = BT.make_model(bt, flags=flags | BT.LAYOUTS[layout], graph_steps=K)
    t0 = time.perf_counter()
    g.step(nsteps, out=False)
    g.sync()
    t_batch = time.perf_counter() - t0
    hist = g.group_history(nsteps)  # (dl_sync above raised any device error)
    # per-step kernel phases of the batched step (event-timed, direct launches)
    g.set_graph(0)
    g.set_timing(True)
    g.step(20, out=False)
    g.sync()
    phases = {}
    for ph in ("step", "strut", "joint"):
        ms, cnt = g.phase_time(ph)
        phases[ph] = ms / max(cnt, 1)
    g.set_timing(False)
    g.close()
    npriv = private_node_count(bt.problem)
    b_strut = (B_ELEM + B_ELEM_PLASTIC) * E + B_NODE * npriv
    a_strut = b_strut / (phases["strut"] * 1e-3) / 1e9
    # the same instances one by one: their own contexts, direct launches, all
    # on ONE stream so that no two instances overlap on the device
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    stream = C.c_void_p()
    assert hip.hipStreamCreateWithFlags(C.byref(stream), 1) == 0
    k_seq = seq_steps
    singles = []
    for p in probs:
        m = GpuModel(p, stream=stream, flags=flags)
        m.set_reaction_group(p.groups["bottom"], 2)
        m.step(warmup, p.dt, out=False)
        singles.append(m)
    for m in singles:
        m.sync()
    t0 = time.perf_counter()
    for m, p in zip(singles, probs):
        m.step(k_seq, p.dt, out=False)
    for m in singles:
        m.sync()
    t_seq = time.perf_counter() - t0
    for m in singles:
        m.close()
    hip.hipStreamDestroy(stream)
    batched = E * nsteps / t_batch
    sequential = E * k_seq / t_seq
    res = {"metric": "batched design-sweep element-GP updates/s (FP64)",
           "value": batched, "unit": "element-GP updates/s", "n_gpus": 1,
           "steps": nsteps, "warmup": 0, "ms_per_step": 1e3 * t_batch / nsteps,
           "higher_is_better": True, "dtype": "f64",
           "data": "synthetic (seeded BCC lattice generator per instance)",
           "config": {"workload": f"sweep-{n_inst}-bcc-elastoplastic", "instances": n_inst,
                      "elements": E, "nodes": bt.problem.n_nodes,
                      "dt_range": [float(bt.dts.min()), float(bt.dts.max())],
                      "steps_per_instance": nsteps, "nominal_strain": L.C4_STRAIN,
                      "yield": yield_family(bt.problem)},
           "schedule": {"instance_order": order, "graph_steps": K, "layout": layout,
                        "autotune_ms_per_step": {f"{o}/{lay}/{k}": round(1e3 * t, 5)
                                                 for (o, lay, k), t in tune.items()}},
           "wall_s": t_batch,
           "phase_ms": {k: round(v, 5) for k, v in phases.items()},
           "roofline": {"bound": "hbm", "kernel": "k_strut", "achieved": round(a_strut, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(a_strut / HBM_PEAK_GBS, 4),
                        "algorithmic_bytes_per_launch": b_strut,
                        "avg_launch_ms": round(phases["strut"], 5)},
           "reaction_peak_range_N": [float(np.abs(hist).max(0).min()),
                                     float(np.abs(hist).max(0).max())],
           "sequential_value": sequential, "sequential_steps_per_instance": k_seq,
           "batched_over_sequential": batched / sequential}
    if cpu_budget > 0:
        from oracle.ref_cpu import RefModel
        b = int(np.argsort([p.n_elems for p in probs])[n_inst // 2])
        q = probs[b]
        for omp, key in ((True, "cpu_baseline"), (False, "cpu_baseline_serial")):
            ref = RefModel(q, omp=omp)
            t1 = time.perf_counter()
            ref.leapfrog(1)
            t1 = time.perf_counter() - t1
            k = int(max(1, min(nsteps, cpu_budget / max(t1, 1e-9))))
            t1 = time.perf_counter()
            ref.leapfrog(k)
            t1 = time.perf_counter() - t1
            cores = cpu_threads() if omp else 1
            res[key] = {"value": q.n_elems * k / t1, "unit": "element-GP updates/s",
                        "cores": cores, "kind": "port",
                        "sample": f"oracle (oracle/ref_cpu.c, {cores} thread(s)) on the median "
                                  f"instance {q.name} ({q.n_elems} elements), {k} LeapFrog "
                                  f"steps after 1, {t1:.1f} s"}
        res["speedup_vs_cpu"] = batched / res["cpu_baseline"]["value"]
    return res


def adjoint(args):
    """Adjoint design gradient of the 20^3 BCC impact with contact (see
    adjoint_measure), with a central finite-difference spot check; with
    --adjoint-sigma-y the elastoplastic variant."""
    print(json.dumps(adjoint_measure(args.adjoint_n, args.adjoint_steps, args.checkpoint_every,
                                     args.fd_struts, args.adjoint_sigma_y or None)), flush=True)


def adjoint_measure(n_cells=20, n=5000, every=100, fd_struts=0, sigma_y=None):
    """Adjoint design gradient of the n_cells^3 BCC impact with contact:
    gradient of J = absorbed energy + peak plate force with respect to all
    strut radii, by the checkpointed forward run and the reverse sweep;
    reported as (forward with checkpoints + reverse sweep) / plain forward
    time, with a central finite-difference spot check of the GPU forward run
    on fd_struts struts.  With sigma_y the elastoplastic variant, whose
    radius gradient includes the yield weights' dependence on the radius."""
    from dynlattice_amd import lattice as L
    from dynlattice_amd.adjoint import ImpactSimulation
    from dynlattice_amd.gpu import SECTION_FIELDS
    p = L.bcc_c5(n_cells, sigma_y)
    print(f"adjoint: {p.name}: {p.n_elems} elements, {p.n_struts} struts, {n} steps, "
          f"checkpoint every {every}", file=sys.stderr, flush=True)
    sim = ImpactSimulation(p, n, every)
    m = sim.model
    sec = {f: np.array(getattr(p.sections, f)) for f in SECTION_FIELDS}
    r0 = p.contact.radius.copy()
    v_plate = -20.0
    sgn = np.sign(v_plate)

    def objective(reaction):
        F = reaction[:, 0] * sgn
        return p.dt * abs(v_plate) * F.sum() + F.max(), F

    reaction = sim.run(sec, r0)  # warm-up: allocations, graph, caches
    J0, F = objective(reaction)
    rbar = np.full((n, 1), p.dt * abs(v_plate) * sgn)
    rbar[int(np.argmax(F)), 0] += sgn
    sim.gradients(rbar)
    # plain forward (no checkpoints)
    m.set_state(p.disp0, p.velo0)
    m.init()
    t0 = time.perf_counter()
    m.step(n)
    t_plain = time.perf_counter() - t0
    # checkpointed forward + reverse sweep
    m.set_state(p.disp0, p.velo0)
    m.init()
    t0 = time.perf_counter()
    m.adjoint_forward(n, p.dt, every)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    g = m.adjoint_backward(rbar)
    t_bwd = time.perf_counter() - t0
    pi = math.pi
    dJdr = (g["area"] * 2 * pi * r0 + (g["I1"] + g["I2"]) * pi * r0 ** 3 +
            g["J"] * 2 * pi * r0 ** 3 + g["contact_radius"])
    if p.plastic is not None:
        # weights 1/capacity: capacities go with r^2 (shear, axial), r^3 (moments)
        k = np.array([2.0, 2.0, 2.0, 3.0, 3.0, 3.0])
        dJdr += (g["yield_weight"] * (-k * p.plastic.weights / r0[:, None])).sum(1)
    active = float(m.plastic_active().mean()) if p.plastic is not None else 0.0
    # finite-difference spot check on the GPU forward run
    fd = []
    for s in np.random.default_rng(5).choice(p.n_struts, fd_struts, replace=False):
        h = 1e-6 * r0[s]
        Js = []
        for sg in (1, -1):
            r = r0.copy()
            r[s] += sg * h
            S = L.Sections.circle(r, p.sections.young, 0.3, p.sections.density)
            yw = None
            if p.plastic is not None:
                yw = L.Plastic.resultant_capacities(r, sigma_y, 0.01, 2.0).weights
            Js.append(objective(sim.run({f: getattr(S, f) for f in SECTION_FIELDS}, r, yw))[0])
        fdv = (Js[0] - Js[1]) / (2 * h)
        fd.append({"strut": int(s), "fd": fdv, "adjoint": float(dJdr[s]),
                   "rel_err": abs(fdv - dJdr[s]) / abs(fdv)})
    sim.close()
    res = {"metric": "adjoint/forward time ratio (full radius gradient)",
           "value": (t_fwd + t_bwd) / t_plain, "unit": "x", "higher_is_better": False,
           "n_gpus": 1, "dtype": "f64", "data": "synthetic (seeded BCC lattice generator)",
           "config": {"workload": p.name, "elements": p.n_elems, "struts": p.n_struts,
                      "steps": n, "checkpoint_every": every, "design_variables": p.n_struts},
           "t_forward_plain_s": t_plain, "t_forward_checkpointed_s": t_fwd,
           "t_reverse_s": t_bwd, "objective": J0, "fd_check": fd}
    if p.plastic is not None:
        res["config"]["yield"] = "quadratic resultant interaction (p=2)"
        res["config"]["sigma_y"] = sigma_y
        res["plastic_active_fraction_end"] = active
    return res


def milne(args):
    """MilneDevice attempts (predictor-corrector, adaptive step with the
    error estimate, accept/reject on the device) on the given lattice: one
    "step" is one attempt over the whole lattice, i.e. two internal-force
    evaluations, the predictor, corrector, step-size control and commit.
    value = elements x attempts / wall time."""
    from dynlattice_amd.gpu import GpuModel, milne_opts
    p = workload(args.workload)
    prec = args.milne_precision
    print(f"milne: {p.name}: {p.n_elems} elements, {p.n_nodes} nodes, first dt {p.dt:.3e}, "
          f"precision {prec:g}", file=sys.stderr, flush=True)
    g = GpuModel(p, device=0)
    if "bottom" in p.groups:
        g.set_reaction_group(p.groups["bottom"], 2)
    g.set_graph(args.milne_graph)
    o = milne_opts(p.dt, prec)
    q = g.milne(args.warmup, o)
    o = milne_opts(float(q["dt_next"][-1]), prec, dt_min=p.dt / 100, dt_max=p.dt * 100)
    g.sync()
    t0 = time.perf_counter()
    q = g.milne(args.steps, o)
    wall = time.perf_counter() - t0
    value = p.n_elems * args.steps / wall
    o = milne_opts(float(q["dt_next"][-1]), prec, dt_min=p.dt / 100, dt_max=p.dt * 100)
    g.set_timing(True)
    g.milne(args.steps, o)
    phases = {}
    for ph in ("step", "contact", "strut", "joint", "node"):
        ms, cnt = g.phase_time(ph)
        phases[ph] = ms / max(cnt, 1)
    g.set_timing(False)
    g.close()
    # force-mode strut kernel: element data + plastic state read, trial state
    # written (eps_p, alpha, strain, flow increment), private nodes: x, theta
    # read and the nodal force written
    npriv = private_node_count(p)
    b_elem = B_ELEM + (B_ELEM_PLASTIC + 8.0 * 7 if p.plastic is not None else 48.0)
    bytes_strut = b_elem * p.n_elems + 96.0 * npriv
    t_launch = phases["strut"] / 2 * 1e-3
    achieved = bytes_strut / t_launch / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "k_strut (force mode)", "algorithmic_bytes_per_launch": bytes_strut,
            "avg_launch_ms": round(phases["strut"] / 2, 5)}
    res = {"metric": "MilneDevice element-GP attempts/s (FP64, 2 force evaluations each)",
           "value": value, "unit": "element-GP attempts/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded BCC lattice generator, random per-strut radii)",
           "config": {"workload": p.name, "elements": p.n_elems, "nodes": p.n_nodes,
                      "precision": prec, "graph_attempts": args.milne_graph,
                      "material": "elastoplastic" if p.plastic is not None else "elastic",
                      "contact": p.contact is not None},
           "accepted_fraction": float(np.mean(q["accepted"])),
           "dt_range": [float(q["dt"].min()), float(q["dt"].max())],
           "roofline": roof, "phase_ms": {k: round(v, 5) for k, v in phases.items()}}
    if not args.no_cpu:
        from oracle.ref_cpu import RefModel
        from oracle.ref_cpu import milne_opts as ref_milne_opts
        for omp, key in ((True, "cpu_baseline"), (False, "cpu_baseline_serial")):
            ref = RefModel(p, grid=True, omp=omp)
            ro = ref_milne_opts(p.dt, prec)
            t0 = time.perf_counter()
            ref.milne(1, ro)
            t1 = time.perf_counter() - t0
            k = int(max(1, min(args.steps, args.cpu_budget / max(t1, 1e-9))))
            t0 = time.perf_counter()
            ref.milne(k, ro)
            dt = time.perf_counter() - t0
            cores = cpu_threads() if omp else 1
            res[key] = {"value": p.n_elems * k / dt, "unit": "element-GP attempts/s",
                        "cores": cores, "kind": "port",
                        "sample": f"oracle ref_milne (oracle/ref_cpu.c, {cores} thread(s)) on "
                                  f"the same {p.name} lattice, {k} attempts after 1, {dt:.1f} s"}
        res["speedup_vs_cpu"] = value / res["cpu_baseline"]["value"]
    print(json.dumps(res), flush=True)


def profiler_region(on):
    """Starts (on) or stops the collection of a rocprofv3 run with
    --selected-regions (roctxProfilerResume / roctxProfilerPause); a no-op
    without the profiler."""
    import ctypes as C
    try:
        lib = C.CDLL("librocprofiler-sdk-roctx.so")
    except OSError:
        return
    (lib.roctxProfilerResume if on else lib.roctxProfilerPause)(C.c_uint64(0))


def snapshot(g, p):
    """the model's state (displacements, velocities, plastic state) on the host"""
    d, v = g.get_state()
    snap = {"disp": d, "velo": v}
    if p.plastic is not None:
        snap["eps_p"], snap["alpha"] = g.plastic_state()
    return snap


def oracle_from(p, snap, omp):
    from oracle.ref_cpu import RefModel
    ref = RefModel(p, grid=True, omp=omp)
    if snap is not None:
        ref.disp[:] = snap["disp"].reshape(-1)
        ref.velo[:] = snap["velo"].reshape(-1)
        if p.plastic is not None:
            ref.set_plastic_state(snap["eps_p"], snap["alpha"])
    return ref


def window_cpu(p, w, snap, cpu_budget, window):
    """the all-cores oracle timed from a window's state (a sample of about
    cpu_budget seconds)"""
    t_cpu = time.perf_counter()
    ref = oracle_from(p, snap, True)
    t1 = time.perf_counter()
    ref.leapfrog(1)
    t1 = time.perf_counter() - t1
    k = int(max(1, min(window, cpu_budget / max(t1, 1e-9))))
    t1 = time.perf_counter()
    ref.leapfrog(k)
    t1 = time.perf_counter() - t1
    w["cpu_baseline"] = {"value": p.n_elems * k / t1, "unit": "element-GP updates/s",
                         "cores": cpu_threads(), "kind": "port",
                         "sample": f"oracle, all cores, {k} steps from this state"}
    w["speedup_vs_cpu"] = w["value"] / w["cpu_baseline"]["value"]
    w["cpu_s"] = time.perf_counter() - t_cpu


def crush_windows(g, p, strains, window, cpu_budget=0.0, log=sys.stderr, start=0, acc=None,
                  profile=False, deferred=None):
    """Runs the crush (top plate at its prescribed velocity) through the given
    nominal strains and times a window of `window` steps at each: wall time
    per step, kernel phases (HIP events), plastic-active fraction, active
    contact pairs per step, candidate-list rebuilds.  With cpu_budget > 0 the
    all-cores oracle is timed at the same state (state and plastic state
    copied over) on a sample of about cpu_budget seconds.  start: steps the
    model has taken; acc: {"steps", "wall_s"} accumulating the ordinary
    step calls (not the event-timed passes).  profile: the windows' timed
    steps are the regions a rocprofv3 --selected-regions run collects.
    deferred: a list that takes (window, state snapshot) for the oracle
    timing instead of running it between the GPU's timed work (a host
    pause of seconds lets the GPU clocks drop before the next timed steps)."""
    H = float(p.coords[:, 2].max() - p.coords[:, 2].min())
    v = float(np.abs(p.velo0[p.groups["top"], 2]).max())
    # k_strut's algorithmic bytes per launch (the headline roofline's figure)
    b_strut = (B_ELEM + (B_ELEM_PLASTIC if p.plastic is not None else 0.0)) * p.n_elems + \
        B_NODE * private_node_count(p)
    out = []
    done = start
    acc = {} if acc is None else acc

    def stepped(k, wall):
        acc["steps"] = acc.get("steps", 0) + k
        acc["wall_s"] = acc.get("wall_s", 0.0) + wall

    for eps in strains:
        target = int(round(eps * H / v / p.dt))
        while done < target:
            k = min(4000, target - done)
            t0 = time.perf_counter()
            g.step(k)  # with an output: device errors surface per chunk
            stepped(k, time.perf_counter() - t0)
            done += k
        g.sync()
        if profile:
            profiler_region(True)
        t0 = time.perf_counter()
        o = g.step(window)
        wall = time.perf_counter() - t0
        if profile:
            profiler_region(False)
        stepped(window, wall)
        h = g.history(window)
        act = g.plastic_active()
        g.set_timing(True)
        g.step(window, out=False)
        g.sync()
        ph = {}
        for name in ("step", "contact", "refresh", "react", "strut", "joint"):
            ms, cnt = g.phase_time(name)
            ph[name] = round(ms / max(cnt, 1), 5)
        g.set_timing(False)
        done += 2 * window
        w = {"strain": eps, "step": done, "ms_per_step": 1e3 * wall / window,
             "value": p.n_elems * window / wall, "phase_ms": ph,
             "plastic_active_fraction": float(act.mean()),
             "active_pairs_per_step": float(h["n_contacts"].mean()),
             "rebuilds_per_window": int(o.n_rebuilds),
             "candidates_last_step": int(o.n_candidates),
             "list_entries": int(g.contact_list_stats()[0]) if p.contact is not None else 0,
             "dissipation_J": float(h["e_diss"][-1]),
             "strut_roofline_frac": round(b_strut / (ph["strut"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        if cpu_budget > 0:
            if deferred is not None:
                deferred.append((w, snapshot(g, p)))
            else:
                window_cpu(p, w, snapshot(g, p), cpu_budget, window)
        print(f"bench: strain {eps:.2f} (step {done}): {w['ms_per_step']:.4f} ms/step, "
              f"plastic {w['plastic_active_fraction']:.3f}, pairs {w['active_pairs_per_step']:.0f}, "
              f"rebuilds {w['rebuilds_per_window']}", file=log, flush=True)
        out.append(w)
    return out


def crush(args):
    """The crush through 10 %, 20 % and 40 % nominal strain with a timed window at each"""
    from dynlattice_amd.gpu import DL_TRACK_DISSIPATION, GpuModel
    p = workload(args.workload)
    if args.max_iter and p.plastic is not None:
        p.plastic.max_iter = args.max_iter
    print(f"crush: {p.name}: {p.n_elems} elements, dt {p.dt:.3e}", file=sys.stderr, flush=True)
    g = GpuModel(p, device=0, flags=DL_TRACK_DISSIPATION)
    g.set_reaction_group(p.groups["bottom"], 2)
    # direct launches under a --selected-regions profile: that profiler mode
    # fails inside captured-graph launches
    g.set_graph(0 if args.profile_windows else 10)
    strains = [float(x) for x in args.strains.split(",")]
    if args.from_state:
        # a short run from a saved crush state (e.g. under rocprofv3)
        z = np.load(args.from_state)
        g.set_state(z["disp"], z["velo"])
        g.set_plastic_state(z["eps_p"], z["alpha"])
        g.step(args.warmup)
        t0 = time.perf_counter()
        o = g.step(args.steps)
        wall = time.perf_counter() - t0
        print(json.dumps({"from_state": args.from_state, "steps": args.steps,
                          "ms_per_step": 1e3 * wall / args.steps,
                          "active_pairs_last_step": int(o.n_contacts)}), flush=True)
        g.close()
        return
    res = {"metric": "element-GP updates/s per explicit step (FP64), crush windows",
           "config": {"workload": p.name, "elements": p.n_elems, "dt": p.dt,
                      "yield": yield_family(p), "window_steps": args.steps},
           "windows": crush_windows(g, p, strains, args.steps,
                                    0.0 if args.no_cpu else args.cpu_budget,
                                    profile=args.profile_windows)}
    if args.save_state:
        d, v = g.get_state()
        ep, al = g.plastic_state()
        np.savez(args.save_state, disp=d, velo=v, eps_p=ep, alpha=al)
    g.close()
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sweep", action="store_true",
                    help="batched design sweep vs sequential single-instance runs")
    ap.add_argument("--sweep-instances", type=int, default=256)
    ap.add_argument("--sweep-steps", type=int, default=10_000,
                    help="steps of the timed sweep (every instance; C4: 10,000)")
    ap.add_argument("--sweep-tune-steps", type=int, default=50)
    ap.add_argument("--sweep-seq-steps", type=int, default=200,
                    help="steps per instance of the one-by-one comparison")
    ap.add_argument("--adjoint", action="store_true",
                    help="adjoint gradient of the 20^3 impact: adjoint/forward time ratio")
    ap.add_argument("--adjoint-n", type=int, default=20)
    ap.add_argument("--adjoint-sigma-y", type=float, default=0.0,
                    help="with --adjoint: elastoplastic lattice of this yield stress")
    ap.add_argument("--adjoint-steps", type=int, default=5000)
    ap.add_argument("--checkpoint-every", type=int, default=100)
    ap.add_argument("--fd-struts", type=int, default=4)
    ap.add_argument("--milne", action="store_true",
                    help="MilneDevice adaptive predictor-corrector attempts on --workload")
    ap.add_argument("--milne-precision", type=float, default=1e-4)
    ap.add_argument("--milne-graph", type=int, default=10,
                    help="attempts per captured graph (0: direct launches)")
    ap.add_argument("--crush", action="store_true",
                    help="the crush through --strains with a timed window at each")
    ap.add_argument("--strains", default="0.1,0.2,0.4")
    ap.add_argument("--save-state", default="",
                    help="with --crush: after the last window, save the state (npz) here")
    ap.add_argument("--from-state", default="",
                    help="start from a state saved by --save-state (then time --steps steps)")
    ap.add_argument("--profile-windows", action="store_true",
                    help="with --crush: only the windows' timed steps are collected by a "
                         "rocprofv3 --selected-regions run")
    ap.add_argument("--max-iter", type=int, default=0,
                    help="return-map iteration cap (the material's maxIter; 0: the workload's)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="bcc40-elastoplastic-contact")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true",
                    help="skip the per-kernel event timing pass")
    ap.add_argument("--no-crush", action="store_true",
                    help="time the steps at the start of the run instead of in the crush")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the short batched-sweep comparison of the default line")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint/forward time ratio (20^3 lattice, 5000 steps)")
    ap.add_argument("--window", type=int, default=200, help="steps per crush window")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.sweep:
        sweep(args)
        return
    if args.adjoint:
        adjoint(args)
        return
    if args.milne:
        milne(args)
        return
    if args.crush:
        crush(args)
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus > 1 is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    from dynlattice_amd.gpu import DL_TRACK_DISSIPATION, GpuModel

    p = workload(args.workload)
    print(f"bench: {p.name}: {p.n_elems} elements, {p.n_nodes} nodes, {p.n_struts} struts, "
          f"dt {p.dt:.3e}", file=sys.stderr, flush=True)
    flags = DL_TRACK_DISSIPATION if p.plastic is not None else 0
    g = GpuModel(p, device=0, flags=flags)
    if "bottom" in p.groups:
        g.set_reaction_group(p.groups["bottom"], 2)
    from dynlattice_amd.batch import autotune_graph
    graph_steps, tune = autotune_graph(g, candidates=(0, 10), steps=10)
    crushing = p.contact is not None and p.plastic is not None and not args.no_crush
    windows = []
    acc = {}
    # the oracle timings wait for the end: all GPU timing runs back to back
    # (a host pause of seconds lets the GPU clocks drop before the next
    # timed steps)
    deferred = []
    if crushing:
        # the timed steps sit in the crush where it works: 20 % nominal strain
        windows += crush_windows(g, p, [0.1, 0.2], args.window,
                                 0.0 if args.no_cpu else args.cpu_budget / 2, acc=acc,
                                 deferred=deferred)
    t0 = time.perf_counter()
    g.step(args.warmup, out=False)
    g.sync()
    t_w = time.perf_counter() - t0
    t0 = time.perf_counter()
    o = g.step(args.steps, out=False)
    g.sync()
    wall = time.perf_counter() - t0
    o = g.step(0)  # error check only
    timed_state = snapshot(g, p) if not args.no_cpu else None
    acc["steps"] = acc.get("steps", 0) + args.warmup + args.steps
    acc["wall_s"] = acc.get("wall_s", 0.0) + t_w + wall
    value = p.n_elems * args.steps / wall
    ms_per_step = 1e3 * wall / args.steps
    h_act = g.plastic_active().mean() if p.plastic is not None else 0.0

    roof = None
    step_roof = None
    phases = {}
    if not args.no_timing:
        g.set_timing(True)
        g.step(args.steps, out=False)
        g.sync()
        for ph in ("step", "contact", "refresh", "react", "strut", "joint"):
            ms, cnt = g.phase_time(ph)
            phases[ph] = ms / max(cnt, 1)
        g.set_timing(False)
        npriv = private_node_count(p)
        b_elem = B_ELEM + (B_ELEM_PLASTIC if p.plastic is not None else 0.0)
        bytes_strut = b_elem * p.n_elems + B_NODE * npriv
        t_strut = phases["strut"] * 1e-3
        achieved = bytes_strut / t_strut / 1e9
        prof = profiled_counters(p.name)
        traffic = prof.get("traffic")
        copy_gbs = copy_bandwidth()
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_strut", "algorithmic_bytes_per_launch": bytes_strut,
                "avg_launch_ms": round(phases["strut"], 5),
                # practical ceiling on this device: a 1 GiB device copy
                "device_copy_GBs": round(copy_gbs, 1)}
        if traffic:
            actual = traffic / t_strut / 1e9
            roof["actual_GBs"] = round(actual, 1)
            roof["actual_frac_of_device_copy"] = round(actual / copy_gbs, 4)
        if prof.get("fp64_flops"):
            tf = prof["fp64_flops"] / t_strut / 1e12
            roof["fp64_TFLOPs"] = round(tf, 3)
            roof["fp64_frac"] = round(tf / FP64_PEAK_TFLOPS, 4)
        # the whole step against the same bound: every element's and node's
        # algorithmic bytes (the per-element and per-node figures above) over
        # the measured step time
        b_step = b_elem * p.n_elems + B_NODE * p.n_nodes
        a_step = b_step / (ms_per_step * 1e-3) / 1e9
        step_roof = {"bound": "hbm", "achieved": round(a_step, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(a_step / HBM_PEAK_GBS, 4),
                     "algorithmic_bytes_per_step": b_step}

    res = {
        "metric": "element-GP updates/s per explicit step (FP64)",
        "value": value,
        "unit": "element-GP updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded BCC lattice generator, random per-strut radii)",
        "config": {"workload": p.name, "elements": p.n_elems, "nodes": p.n_nodes,
                   "struts": p.n_struts, "dt": p.dt,
                   "material": "elastoplastic" if p.plastic is not None else "elastic",
                   "contact": p.contact is not None, "so3": bool(p.so3),
                   "yield": yield_family(p),
                   "state": ("crush at 20 % nominal strain" if crushing else "start of the run")},
        "roofline": roof,
        "step_roofline": step_roof,
        "graph_steps": graph_steps,
        "phase_ms": {k: round(v, 5) for k, v in phases.items()},
        "plastic_active_fraction": float(h_act),
    }
    if crushing:
        try:
            start = windows[-1]["step"] + args.warmup + (1 if args.no_timing else 2) * args.steps
            windows += crush_windows(g, p, [0.4], args.window,
                                     0.0 if args.no_cpu else args.cpu_budget / 2, start=start,
                                     acc=acc, deferred=deferred)
            # the crush from 0 to 40 % strain: the ordinary step calls (list
            # rebuilds and windows included, event-timed passes excluded)
            res["whole_crush"] = {"steps": acc["steps"], "wall_s": round(acc["wall_s"], 3),
                                  "value": p.n_elems * acc["steps"] / acc["wall_s"],
                                  "last_step": windows[-1]["step"]}
        except Exception as e:  # report, keep the line
            res["crush_error"] = str(e)
        res["crush_windows"] = windows
    g.close()
    if not args.no_adjoint:
        a = adjoint_measure(20, 5000, 100, 0)
        res["adjoint_forward_ratio"] = round(a["value"], 3)
        res["adjoint"] = {k: a[k] for k in ("config", "t_forward_plain_s",
                                            "t_forward_checkpointed_s", "t_reverse_s")}
        # the same gradient through the elastoplastic material's return map
        a = adjoint_measure(20, 5000, 100, 0, sigma_y=250e6)
        res["adjoint_plastic_forward_ratio"] = round(a["value"], 3)
        res["adjoint_plastic"] = {k: a[k] for k in ("config", "t_forward_plain_s",
                                                    "t_forward_checkpointed_s", "t_reverse_s",
                                                    "plastic_active_fraction_end")}
    if not args.no_sweep:
        # C4: the batched sweep against the same instances one by one (its
        # first 1000 steps; bench.py --sweep times all 10,000)
        w = sweep_measure(256, 1000, 10, 100, args.warmup, 0.0)
        res["sweep_batched_over_sequential"] = round(w["batched_over_sequential"], 3)
        res["sweep"] = {k: w[k] for k in ("value", "ms_per_step", "steps", "sequential_value",
                                          "sequential_steps_per_instance", "roofline")}
        res["sweep"]["config"] = {k: w["config"][k] for k in ("workload", "instances", "elements")}
        res["sweep"]["schedule"] = {k: w["schedule"][k]
                                    for k in ("instance_order", "graph_steps", "layout")}
    # the oracle on the host cores, from the states the GPU was timed at
    if not args.no_cpu:
        cb = cpu_baseline(p, args.cpu_budget, args.steps, omp=True, snap=timed_state)
        c1 = cpu_baseline(p, args.cpu_budget, args.steps, omp=False, snap=timed_state)
        res["cpu_baseline"] = cb
        res["speedup_vs_cpu"] = value / cb["value"]
        res["cpu_baseline_serial"] = c1
        res["speedup_vs_cpu_serial"] = value / c1["value"]
        for w, snap in deferred:
            window_cpu(p, w, snap, args.cpu_budget / 2, args.window)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
