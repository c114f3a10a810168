"""ctypes binding of libdlgpu.so (the C ABI in include/dl_gpu.h).

The HIP library is the only compute path: if ``libdlgpu.so`` is missing or
cannot be loaded this module raises ImportError -- there is no CPU fallback.
Build it with ``python -m dynlattice_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdlgpu.so")

DL_OK = 0
DL_ERR_INPUT = -1
DL_ERR_ZERO_MASS = -2
DL_ERR_PLASTIC_ITER = -3
DL_ERR_HIP = -4
DL_ERR_CAPACITY = -5
DL_ERR_STATE = -6

DL_INIT, DL_GET_INT_VECTOR, DL_GET_MATRIX2, DL_COMMIT, DL_CANCEL = range(5)
DL_LOAD_CASE_DEFAULT, DL_LOAD_CASE_OUTPUT = 0, 1
DL_TRACK_DISSIPATION = 1
DL_FRAMES_PER_ELEMENT = 2
DL_JOINTS_IN_MEMORY = 4

#: every symbol declared in include/dl_gpu.h
EXPORTS = (
    "dl_create", "dl_destroy", "dl_take_action", "dl_set_state", "dl_get_state",
    "dl_set_bcs", "dl_set_reaction_group", "dl_explicit_step", "dl_get_history",
    "dl_contact_pairs", "dl_contact_candidates", "dl_plastic_active", "dl_plastic_state",
    "dl_phase_time", "dl_set_timing", "dl_sync", "dl_last_error",
    "dl_set_reaction_groups", "dl_set_node_dt", "dl_set_graph", "dl_get_group_history",
    "dl_set_sections", "dl_adjoint_forward", "dl_adjoint_backward", "dl_milne_step",
    "dl_set_programs", "dl_set_time", "dl_get_time", "dl_program_eval", "dl_set_point_masses",
    "dl_get_table", "dl_get_energies", "dl_set_impact_groups", "dl_explicit_step_noalloc",
    "dl_contact_list_stats", "dl_set_plastic_state", "dl_set_yield_weights",
    "dl_set_mass_update",
)
#: DL_NOALLOC_SCRATCH_DOUBLES of include/dl_gpu.h
NOALLOC_SCRATCH_DOUBLES = 1 << 18
ELEM_TABLES = ("strain", "mat_strain", "stress", "mat_stress")
NODE_TABLES = ("potentialEnergy", "kineticEnergy", "mass")
DL_PROG_CONST, DL_PROG_STEP, DL_PROG_EXPR = 0, 1, 2

#: section fields of dl_sections / dl_section_grads, in ABI order
SECTION_FIELDS = ("area", "I1", "I2", "J", "shear_factor", "young", "shear_modulus", "density")

dp = C.POINTER(C.c_double)
ip32 = C.POINTER(C.c_int32)
ip64 = C.POINTER(C.c_int64)
u8p = C.POINTER(C.c_uint8)


class DLError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dl error {code}: {msg}")
        self.code = code


class dl_mesh(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("n_elems", C.c_int64), ("n_struts", C.c_int32),
                ("coords", dp), ("elem_nodes", ip32), ("elem_strut", ip32)]


class dl_sections(C.Structure):
    _fields_ = [("area", dp), ("I1", dp), ("I2", dp), ("J", dp), ("shear_factor", dp),
                ("young", dp), ("shear_modulus", dp), ("density", dp),
                ("inertia_correct", C.c_double)]


class dl_material(C.Structure):
    _fields_ = [("model", C.c_int32), ("weights", dp), ("yield_level", C.c_double),
                ("kin_coupling", C.c_double * 6), ("iso_coupling", C.c_double),
                ("has_iso", C.c_int32), ("iso_modulus", C.c_double), ("has_kin", C.c_int32),
                ("kin_const", C.c_double * 6), ("kin_ratio", C.c_double),
                ("precision", C.c_double), ("max_iter", C.c_int32),
                ("exponent", C.c_double * 6)]


class dl_contact(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("penalty_sts", C.c_double),
                ("penalty_nts", C.c_double), ("radius", dp), ("max_pairs", C.c_int64),
                ("skin", C.c_double)]


class dl_options(C.Structure):
    _fields_ = [("so3", C.c_int32), ("device", C.c_int32), ("stream", C.c_void_p),
                ("flags", C.c_uint32)]


class dl_params(C.Structure):
    _fields_ = [("int_vector", dp), ("matrix2", dp), ("load_case", C.c_int32),
                ("with_gyro", C.c_int32), ("lumped", C.c_int32), ("e_pot", C.c_double),
                ("e_diss", C.c_double)]


class dl_section_grads(C.Structure):
    _fields_ = [(f, dp) for f in SECTION_FIELDS] + [("contact_radius", dp), ("yield_weight", dp)]


class dl_step_out(C.Structure):
    _fields_ = [("reaction", C.c_double), ("e_diss", C.c_double), ("n_contacts", C.c_int64),
                ("n_candidates", C.c_int64), ("steps_done", C.c_int64), ("n_rebuilds", C.c_int64)]


class dl_milne_opts(C.Structure):
    """MilneDevice settings (defaults of the reference's solver module:
    dt_min = dt / 100, dt_max = 100 dt, safety 0.9, incr 1.2, decr 0.8,
    length scale 1e-3)"""
    _fields_ = [("dt", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double),
                ("precision", C.c_double), ("safety", C.c_double), ("incr", C.c_double),
                ("decr", C.c_double), ("len_scale", C.c_double)]


def milne_opts(dt, precision, dt_min=None, dt_max=None, safety=0.9, incr=1.2, decr=0.8,
               len_scale=1e-3):
    return dl_milne_opts(dt, dt / 100.0 if dt_min is None else dt_min,
                         dt * 100.0 if dt_max is None else dt_max, precision, safety, incr,
                         decr, len_scale)


class dl_program(C.Structure):
    _fields_ = [("kind", C.c_int32), ("a", C.c_double), ("b", C.c_double), ("c", C.c_double),
                ("expr", C.c_char_p)]


def program_eval(expr, t=0.0, i=0.0):
    """value of a program expression through the device's compiler and
    evaluator, run on the host"""
    L = load()
    v = C.c_double()
    rc = L.dl_program_eval(expr.encode(), float(t), float(i), C.byref(v))
    if rc != DL_OK:
        raise DLError(rc, L.dl_last_error().decode())
    return v.value


class dl_milne_out(C.Structure):
    _fields_ = [("attempts", C.c_int64), ("accepted", C.c_int64), ("time", C.c_double),
                ("dt_next", C.c_double)]


_lib = None


def load():
    """Load libdlgpu.so (raises ImportError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not built: run `python -m dynlattice_amd.build`")
    L = C.CDLL(LIB_PATH)
    L.dl_create.argtypes = [C.POINTER(dl_mesh), C.POINTER(dl_sections), C.POINTER(dl_material),
                            C.POINTER(dl_contact), C.POINTER(dl_options), C.POINTER(C.c_void_p)]
    L.dl_destroy.argtypes = [C.c_void_p]
    L.dl_take_action.argtypes = [C.c_void_p, C.c_int32, C.POINTER(dl_params)]
    L.dl_set_state.argtypes = [C.c_void_p, dp, dp, C.c_int64]
    L.dl_get_state.argtypes = [C.c_void_p, dp, dp, C.c_int64]
    L.dl_set_bcs.argtypes = [C.c_void_p, u8p, dp, dp, C.c_int64]
    L.dl_set_reaction_group.argtypes = [C.c_void_p, ip32, C.c_int64, C.c_int32]
    L.dl_explicit_step.argtypes = [C.c_void_p, C.c_int64, C.c_double, C.POINTER(dl_step_out)]
    L.dl_explicit_step_noalloc.argtypes = [C.c_void_p, C.c_int64, C.c_double, dp, dp, ip64]
    L.dl_contact_list_stats.argtypes = [C.c_void_p, ip64, ip64]
    L.dl_set_plastic_state.argtypes = [C.c_void_p, dp, dp, C.c_int64]
    L.dl_set_yield_weights.argtypes = [C.c_void_p, dp]
    L.dl_set_mass_update.argtypes = [C.c_void_p, C.c_char_p]
    L.dl_get_history.argtypes = [C.c_void_p, dp, dp, ip64, C.c_int64, ip64]
    L.dl_contact_pairs.argtypes = [C.c_void_p, ip32, ip32, C.c_int64, ip64]
    L.dl_contact_candidates.argtypes = [C.c_void_p, ip32, ip32, C.c_int64, ip64]
    L.dl_plastic_active.argtypes = [C.c_void_p, u8p, C.c_int64]
    L.dl_plastic_state.argtypes = [C.c_void_p, dp, dp, C.c_int64]
    L.dl_phase_time.argtypes = [C.c_void_p, C.c_char_p, dp, ip64]
    L.dl_set_timing.argtypes = [C.c_void_p, C.c_int32]
    L.dl_sync.argtypes = [C.c_void_p]
    L.dl_last_error.restype = C.c_char_p
    L.dl_set_reaction_groups.argtypes = [C.c_void_p, ip32, ip32, C.c_int32, C.c_int32]
    L.dl_set_node_dt.argtypes = [C.c_void_p, dp, C.c_int64]
    L.dl_set_graph.argtypes = [C.c_void_p, C.c_int32]
    L.dl_get_group_history.argtypes = [C.c_void_p, dp, C.c_int64, ip64]
    L.dl_set_sections.argtypes = [C.c_void_p, C.POINTER(dl_sections), dp]
    L.dl_adjoint_forward.argtypes = [C.c_void_p, C.c_int64, C.c_double, C.c_int32, dp]
    L.dl_adjoint_backward.argtypes = [C.c_void_p, dp, C.POINTER(dl_section_grads)]
    L.dl_milne_step.argtypes = [C.c_void_p, C.c_int64, C.POINTER(dl_milne_opts), dp,
                                C.POINTER(dl_milne_out)]
    L.dl_set_programs.argtypes = [C.c_void_p, C.POINTER(dl_program), C.c_int32, ip32, C.c_int64]
    L.dl_set_time.argtypes = [C.c_void_p, C.c_double, C.c_int64]
    L.dl_get_time.argtypes = [C.c_void_p, dp, ip64]
    L.dl_program_eval.argtypes = [C.c_char_p, C.c_double, C.c_double, dp]
    L.dl_set_point_masses.argtypes = [C.c_void_p, dp, dp, C.c_int64]
    L.dl_get_table.argtypes = [C.c_void_p, C.c_char_p, dp, C.c_int64, C.c_int32]
    L.dl_get_energies.argtypes = [C.c_void_p, dp, dp, dp, dp]
    L.dl_set_impact_groups.argtypes = [C.c_void_p, ip32, ip32, C.c_int32, ip32, dp]
    _lib = L
    return L


def _p(a, t=dp):
    return a.ctypes.data_as(t) if a is not None else None


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


class GpuModel:
    """Device-resident rod lattice behind the C ABI (one dl_ctx)."""

    def __init__(self, p, device=0, stream=None, flags=0, init=True, max_pairs=0):
        self.L = load()
        self.p = p
        k = []
        self._keep = k
        co = _c(p.coords)
        en = _c(p.elem_nodes, np.int32)
        es = _c(p.elem_strut, np.int32)
        k += [co, en, es]
        mesh = dl_mesh(p.n_nodes, p.n_elems, p.n_struts, _p(co), _p(en, ip32), _p(es, ip32))
        s = p.sections
        arrs = [_c(getattr(s, f)) for f in ("area", "I1", "I2", "J", "shear_factor", "young",
                                            "shear_modulus", "density")]
        k += arrs
        sec = dl_sections(*[_p(a) for a in arrs], s.inertia_correct)
        mat = dl_material()
        mat.model = 0
        if p.plastic is not None:
            q = p.plastic
            w = _c(q.weights)
            k.append(w)
            mat.model = 1
            mat.weights = _p(w)
            mat.yield_level = q.yield_level
            mat.kin_coupling[:] = list(q.kin_coupling)
            mat.iso_coupling = q.iso_coupling
            mat.has_iso = int(q.has_iso)
            mat.iso_modulus = q.iso_modulus
            mat.has_kin = int(q.has_kin)
            mat.kin_const[:] = list(q.kin_const)
            mat.kin_ratio = q.kin_ratio
            mat.precision = q.precision
            mat.max_iter = q.max_iter
            mat.exponent[:] = list(q.exponent)
        ct = dl_contact()
        if p.contact is not None:
            r = _c(p.contact.radius)
            k.append(r)
            ct.enabled = 1
            ct.penalty_sts = p.contact.penalty_sts
            ct.penalty_nts = p.contact.penalty_nts
            ct.radius = _p(r)
            ct.max_pairs = int(max_pairs)
            ct.skin = float(getattr(p.contact, "skin", 0.0))
        opt = dl_options(int(p.so3), int(device), stream, int(flags))
        h = C.c_void_p()
        self._check(self.L.dl_create(C.byref(mesh), C.byref(sec), C.byref(mat), C.byref(ct),
                                     C.byref(opt), C.byref(h)))
        self.h = h
        self.set_bcs(p.cmask, p.cval, p.fext)
        if getattr(p, "point_mass", None) is not None or getattr(p, "point_inertia", None) is not None:
            self.set_point_masses(p.point_mass, p.point_inertia)
        if getattr(p, "programs", None):
            self.set_programs(p.programs, p.dof_prog)
            self.set_time(p.t0, 0)
        if getattr(p, "impact", None):
            self.set_impact_groups(p.impact)
        self.set_state(p.disp0, p.velo0)
        if init:
            self.init()

    def _check(self, rc):
        if rc != DL_OK:
            raise DLError(rc, self.L.dl_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.dl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- actions -------------------------------------------------------
    def action(self, act, params=None):
        prm = params if params is not None else dl_params()
        self._check(self.L.dl_take_action(self.h, act, C.byref(prm)))
        return prm

    def init(self):
        self.action(DL_INIT)

    def int_vector(self, inelastic=True, gyro=False):
        f = np.zeros(6 * self.p.n_nodes)
        prm = dl_params()
        prm.int_vector = _p(f)
        prm.load_case = DL_LOAD_CASE_DEFAULT if inelastic else DL_LOAD_CASE_OUTPUT
        prm.with_gyro = int(gyro)
        self.action(DL_GET_INT_VECTOR, prm)
        return f.reshape(-1, 6)

    def mass_blocks(self, lumped=False):
        b = np.zeros(36 * self.p.n_nodes)
        prm = dl_params()
        prm.matrix2 = _p(b)
        prm.lumped = int(lumped)
        self.action(DL_GET_MATRIX2, prm)
        return b.reshape(-1, 6, 6)

    def commit(self):
        prm = self.action(DL_COMMIT)
        return prm.e_pot, prm.e_diss

    def cancel(self):
        self.action(DL_CANCEL)

    # --- state ---------------------------------------------------------
    def set_state(self, disp=None, velo=None):
        n = 6 * self.p.n_nodes
        d = _c(disp).reshape(-1) if disp is not None else None
        v = _c(velo).reshape(-1) if velo is not None else None
        self._check(self.L.dl_set_state(self.h, _p(d), _p(v), n))

    def get_state(self):
        n = 6 * self.p.n_nodes
        d = np.zeros(n)
        v = np.zeros(n)
        self._check(self.L.dl_get_state(self.h, _p(d), _p(v), n))
        return d.reshape(-1, 6), v.reshape(-1, 6)

    def set_bcs(self, cmask, cval, fext):
        n = 6 * self.p.n_nodes
        self._cm = _c(cmask, np.uint8).reshape(-1)
        self._cv = _c(cval).reshape(-1)
        self._fe = _c(fext).reshape(-1)
        self._check(self.L.dl_set_bcs(self.h, _p(self._cm, u8p), _p(self._cv), _p(self._fe), n))

    def set_reaction_group(self, nodes, dof=2):
        g = _c(nodes, np.int32)
        self._check(self.L.dl_set_reaction_group(self.h, _p(g, ip32), g.shape[0], int(dof)))

    def set_reaction_groups(self, nodes, group_ptr, dof=2):
        n = _c(nodes, np.int32)
        g = _c(group_ptr, np.int32)
        self._ngroups = g.shape[0] - 1
        self._check(self.L.dl_set_reaction_groups(self.h, _p(n, ip32), _p(g, ip32),
                                                  self._ngroups, int(dof)))

    def set_node_dt(self, dt):
        if dt is None:
            self._check(self.L.dl_set_node_dt(self.h, None, self.p.n_nodes))
        else:
            a = _c(dt)
            self._check(self.L.dl_set_node_dt(self.h, _p(a), a.shape[0]))

    def set_graph(self, steps_per_graph):
        self._check(self.L.dl_set_graph(self.h, int(steps_per_graph)))

    def group_history(self, nsteps):
        ng = max(getattr(self, "_ngroups", 1), 1)
        r = np.zeros(nsteps * ng)
        n = C.c_int64()
        self._check(self.L.dl_get_group_history(self.h, _p(r), nsteps, C.byref(n)))
        return r[: n.value * ng].reshape(-1, ng)

    # --- stepping ------------------------------------------------------
    def step(self, nsteps, dt=None, out=True):
        dt = self.p.dt if dt is None else dt
        o = dl_step_out()
        self._check(self.L.dl_explicit_step(self.h, int(nsteps), float(dt),
                                            C.byref(o) if out else None))
        return o

    def step_noalloc(self, nsteps, dt=None):
        """nsteps steps through the allocation-free entry point; returns its
        histories (reaction [nsteps, n_groups], e_diss, n_contacts)"""
        dt = self.p.dt if dt is None else dt
        ng = max(getattr(self, "_ngroups", 1), 1)
        r = np.zeros(max(nsteps, 1) * ng)
        e = np.zeros(max(nsteps, 1))
        c = np.zeros(max(nsteps, 1), np.int64)
        self._check(self.L.dl_explicit_step_noalloc(self.h, int(nsteps), float(dt), _p(r), _p(e),
                                                    _p(c, ip64)))
        return {"reaction": r[: nsteps * ng].reshape(-1, ng), "e_diss": e[:nsteps],
                "n_contacts": c[:nsteps]}

    def history(self, nsteps):
        r = np.zeros(nsteps)
        e = np.zeros(nsteps)
        c = np.zeros(nsteps, np.int64)
        n = C.c_int64()
        self._check(self.L.dl_get_history(self.h, _p(r), _p(e), _p(c, ip64), nsteps, C.byref(n)))
        return {"reaction": r[:n.value], "e_diss": e[:n.value], "n_contacts": c[:n.value]}

    def set_programs(self, programs, dof_prog):
        """programs: list of lattice.Program; dof_prog [N, 6] (-1: none)"""
        n = len(programs)
        arr = (dl_program * max(n, 1))()
        enc = [pr.expr.encode() for pr in programs]
        for j, pr in enumerate(programs):
            arr[j] = dl_program(pr.KINDS[pr.kind], pr.a, pr.b, pr.c, enc[j])
        dpg = _c(dof_prog if dof_prog is not None else np.full(6 * self.p.n_nodes, -1), np.int32)
        self._prog_keep = (arr, enc, dpg)
        self._check(self.L.dl_set_programs(self.h, arr, n, _p(dpg, ip32), dpg.size))

    def set_time(self, t, k=0):
        self._check(self.L.dl_set_time(self.h, float(t), int(k)))

    def get_time(self):
        t = C.c_double()
        k = C.c_int64()
        self._check(self.L.dl_get_time(self.h, C.byref(t), C.byref(k)))
        return t.value, k.value

    def set_impact_groups(self, groups):
        """groups: list of (nodes, dof, weight) (ImpactBC)"""
        nodes = _c(np.concatenate([np.asarray(g[0]) for g in groups]) if groups else
                   np.zeros(0), np.int32)
        ptr = _c(np.cumsum([0] + [len(g[0]) for g in groups]), np.int32)
        dofs = _c([g[1] for g in groups], np.int32)
        w = _c([g[2] for g in groups])
        self._imp_keep = (nodes, ptr, dofs, w)
        self._check(self.L.dl_set_impact_groups(self.h, _p(nodes, ip32), _p(ptr, ip32), len(groups),
                                                _p(dofs, ip32), _p(w)))

    def set_point_masses(self, mass=None, inertia=None):
        m = None if mass is None else _c(mass)
        i = None if inertia is None else _c(inertia)
        self._check(self.L.dl_set_point_masses(self.h, _p(m) if m is not None else None,
                                               _p(i) if i is not None else None,
                                               self.p.n_nodes))

    def table(self, name):
        """GET_TABLE at the current state: [n_elems, 6] for strain, mat_strain,
        stress, mat_stress; [n_nodes] for potentialEnergy, kineticEnergy, mass"""
        if name in ELEM_TABLES:
            rows, cols = self.p.n_elems, 6
        else:
            rows, cols = self.p.n_nodes, 1
        out = np.zeros(rows * cols)
        self._check(self.L.dl_get_table(self.h, name.encode(), _p(out), rows, cols))
        return out.reshape(rows, cols) if cols > 1 else out

    def energies(self):
        """{kinetic, potential, dissipated, mass} totals at the current state"""
        v = [C.c_double() for _ in range(4)]
        self._check(self.L.dl_get_energies(self.h, *[C.byref(x) for x in v]))
        return dict(zip(("kinetic", "potential", "dissipated", "mass"), [x.value for x in v]))

    def milne(self, n_attempts, opts):
        """n_attempts MilneDevice attempts from the current state (opts:
        :func:`milne_opts`); returns per-attempt arrays dt, err, accepted,
        reaction, n_contacts, dt_next and the totals (accepted count, time
        advanced)"""
        n = int(n_attempts)
        log = np.zeros(4 * max(n, 1))
        o = dl_milne_out()
        self._check(self.L.dl_milne_step(self.h, n, C.byref(opts), _p(log), C.byref(o)))
        h = self.history(n)
        lg = log[: 4 * n].reshape(-1, 4)
        return {"dt": lg[:, 0], "err": lg[:, 1], "accepted": lg[:, 2] > 0,
                "reaction": h["reaction"], "n_contacts": h["n_contacts"], "dt_next": lg[:, 3],
                "n_accepted": int(o.accepted), "time": float(o.time)}

    def sync(self):
        self._check(self.L.dl_sync(self.h))

    def set_timing(self, on=True):
        self._check(self.L.dl_set_timing(self.h, int(on)))

    def phase_time(self, phase):
        ms = C.c_double()
        cnt = C.c_int64()
        self._check(self.L.dl_phase_time(self.h, phase.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    # --- tables --------------------------------------------------------
    def plastic_active(self):
        m = np.zeros(self.p.n_elems, np.uint8)
        self._check(self.L.dl_plastic_active(self.h, _p(m, u8p), self.p.n_elems))
        return m

    def plastic_state(self):
        E = self.p.n_elems
        ep = np.zeros(6 * E)
        al = np.zeros(7 * E)
        self._check(self.L.dl_plastic_state(self.h, _p(ep), _p(al), E))
        return ep.reshape(E, 6), al.reshape(E, 7)

    def set_plastic_state(self, eps_p, alpha):
        """committed plastic strain [E, 6] and hardening parameters [E, 7]"""
        ep = _c(eps_p).reshape(-1)
        al = _c(alpha).reshape(-1)
        self._check(self.L.dl_set_plastic_state(self.h, _p(ep), _p(al), self.p.n_elems))

    def _pairs(self, fn):
        n = C.c_int64()
        self._check(fn(self.h, None, None, 0, C.byref(n)))
        k = n.value
        a = np.zeros(max(k, 1), np.int32)
        b = np.zeros(max(k, 1), np.int32)
        self._check(fn(self.h, _p(a, ip32), _p(b, ip32), k, C.byref(n)))
        return a[:k], b[:k]

    def contact_list_stats(self):
        """(partners listed over all elements, list builds so far)"""
        a, b = C.c_int64(), C.c_int64()
        self._check(self.L.dl_contact_list_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def contact_pairs(self):
        return self._pairs(self.L.dl_contact_pairs)

    # --- design sensitivities -------------------------------------------
    def set_sections(self, sections=None, contact_radius=None):
        """Replace the per-strut section data (a lattice.Sections) and/or the
        contact radii; call INIT afterwards to refresh the contact blacklist."""
        sec = None
        keep = []
        if sections is not None:
            arrs = [_c(getattr(sections, f)) for f in SECTION_FIELDS]
            keep += arrs
            sec = dl_sections(*[_p(a) for a in arrs], sections.inertia_correct)
        cr = _c(contact_radius) if contact_radius is not None else None
        self._check(self.L.dl_set_sections(self.h, C.byref(sec) if sec is not None else None,
                                           _p(cr)))

    def set_mass_update(self, cond):
        """updateWhen of the explicit modules: the mass is re-assembled at the
        start of the steps (attempts) where cond(t, i) holds, and after INIT
        or set_state (None: every step)"""
        self._check(self.L.dl_set_mass_update(self.h, None if cond is None else cond.encode()))

    def set_yield_weights(self, weights):
        """Replace the elastoplastic material's yield weights ([n_struts, 6])."""
        w = _c(weights).reshape(-1)
        if w.size != 6 * self.p.n_struts:
            raise ValueError(f"yield weights must have {self.p.n_struts} x 6 entries")
        self._check(self.L.dl_set_yield_weights(self.h, _p(w)))

    def adjoint_forward(self, nsteps, dt=None, checkpoint_every=100):
        """nsteps explicit steps with checkpoints; returns the reaction
        history [nsteps, n_groups]."""
        dt = self.p.dt if dt is None else dt
        ng = max(getattr(self, "_ngroups", 1), 1)
        r = np.zeros(nsteps * ng)
        self._check(self.L.dl_adjoint_forward(self.h, int(nsteps), float(dt),
                                              int(checkpoint_every), _p(r)))
        self._adj_steps = int(nsteps)
        return r.reshape(nsteps, ng)

    def adjoint_backward(self, reaction_bar):
        """Gradients of the section data (dict of [n_struts] arrays, plus
        contact_radius and, [n_struts, 6], yield_weight) for the objective
        adjoint reaction_bar [nsteps, n_groups]."""
        ng = max(getattr(self, "_ngroups", 1), 1)
        rb = _c(reaction_bar).reshape(-1)
        if rb.size != self._adj_steps * ng:
            raise ValueError(f"reaction_bar must have {self._adj_steps} x {ng} entries")
        R = self.p.n_struts
        out = {f: np.zeros(R) for f in SECTION_FIELDS + ("contact_radius",)}
        out["yield_weight"] = np.zeros((R, 6))
        g = dl_section_grads(*[_p(out[f]) for f in SECTION_FIELDS + ("contact_radius",)],
                             _p(out["yield_weight"]))
        self._check(self.L.dl_adjoint_backward(self.h, _p(rb), C.byref(g)))
        return out

    def contact_candidates(self):
        return self._pairs(self.L.dl_contact_candidates)
