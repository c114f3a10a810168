"""ctypes front end of the oracle (oracle/ref_cpu.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py -- never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

dp = C.POINTER(C.c_double)
ip32 = C.POINTER(C.c_int32)
ip64 = C.POINTER(C.c_int64)
u8p = C.POINTER(C.c_uint8)


class MilneOpts(C.Structure):
    """ref_milne_opts: first step, bounds, precision and the step-size
    factors of ExplicitModule (defaults: safety 0.9, incr 1.2, decr 0.8,
    length scale 1e-3)"""
    _fields_ = [("dt", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double),
                ("precision", C.c_double), ("safety", C.c_double), ("incr", C.c_double),
                ("decr", C.c_double), ("len_scale", C.c_double)]


def milne_opts(dt, precision, dt_min=None, dt_max=None, safety=0.9, incr=1.2, decr=0.8,
               len_scale=1e-3):
    return MilneOpts(dt, dt / 100.0 if dt_min is None else dt_min,
                     dt * 100.0 if dt_max is None else dt_max, precision, safety, incr, decr,
                     len_scale)


class _Mesh(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("n_elems", C.c_int64), ("n_struts", C.c_int32),
                ("coords", dp), ("elem_nodes", ip32), ("elem_strut", ip32)]


class _Sections(C.Structure):
    _fields_ = [("area", dp), ("I1", dp), ("I2", dp), ("J", dp), ("shear_factor", dp),
                ("young", dp), ("shear_modulus", dp), ("density", dp),
                ("inertia_correct", C.c_double)]


class _Plastic(C.Structure):
    _fields_ = [("model", C.c_int32), ("weights", dp), ("yield_level", C.c_double),
                ("kin_coupling", C.c_double * 6), ("iso_coupling", C.c_double),
                ("has_iso", C.c_int32), ("iso_modulus", C.c_double), ("has_kin", C.c_int32),
                ("kin_const", C.c_double * 6), ("kin_ratio", C.c_double),
                ("precision", C.c_double), ("max_iter", C.c_int32),
                ("exponent", C.c_double * 6)]


class _Contact(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("penalty_sts", C.c_double),
                ("penalty_nts", C.c_double), ("radius", dp)]


class _Options(C.Structure):
    _fields_ = [("so3", C.c_int32)]


def build(omp=False):
    subprocess.check_call(["make", "-s", "-C", _HERE], stdout=subprocess.DEVNULL)


def build_planted(k, outdir):
    """the oracle built with planted mistake k (REF_PLANT in ref_cpu.c) as
    outdir/libref_cpu_plant<k>.so, for the tests that show a pin fails on a
    wrong oracle"""
    out = os.path.join(outdir, f"libref_cpu_plant{k}.so")
    subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-std=c99", "-shared",
                           f"-DREF_PLANT={int(k)}", "-o", out, os.path.join(_HERE, "ref_cpu.c"),
                           "-lm"])
    return out


def lib(omp=False, path=None):
    name = path or ("libref_cpu_omp.so" if omp else "libref_cpu.so")
    if name in _LIBS:
        return _LIBS[name]
    if path is None:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            build()
    L = C.CDLL(path)
    L.ref_create.restype = C.c_void_p
    L.ref_create.argtypes = [C.POINTER(_Mesh), C.POINTER(_Sections), C.POINTER(_Plastic),
                             C.POINTER(_Contact), C.POINTER(_Options), dp]
    L.ref_destroy.argtypes = [C.c_void_p]
    L.ref_last_error.argtypes = [C.c_void_p]
    L.ref_init.argtypes = [C.c_void_p, dp]
    L.ref_get_int_vector.argtypes = [C.c_void_p, dp, dp, dp, C.c_int, C.c_int, C.c_int]
    L.ref_elem_fint.argtypes = [C.c_void_p, C.c_int64, dp, dp, C.c_int]
    L.ref_mass_blocks.argtypes = [C.c_void_p, dp, dp, C.c_int]
    L.ref_set_bcs.argtypes = [C.c_void_p, u8p, dp, dp]
    L.ref_leapfrog.argtypes = [C.c_void_p, dp, dp, C.c_double, C.c_int64, ip32, C.c_int64,
                               C.c_int32, dp, C.c_int, dp, dp, ip64]
    L.ref_commit.argtypes = [C.c_void_p, dp, dp, dp]
    L.ref_output_reaction.argtypes = [C.c_void_p, dp, ip32, C.c_int64, C.c_int32, dp]
    L.ref_milne.argtypes = [C.c_void_p, dp, dp, C.POINTER(MilneOpts), C.c_int64, ip32,
                            C.c_int64, C.c_int32, dp]
    L.ref_set_programs.argtypes = [C.c_void_p, C.c_int, ip32, dp, C.POINTER(C.c_char_p), ip32]
    L.ref_set_time.argtypes = [C.c_void_p, C.c_double, C.c_int64]
    L.ref_set_mass_update.argtypes = [C.c_void_p, C.c_char_p]
    L.ref_mass_in_use.argtypes = [C.c_void_p, dp]
    L.ref_get_time.argtypes = [C.c_void_p, dp, ip64]
    L.ref_expr_eval.argtypes = [C.c_char_p, C.c_double, C.c_double, dp]
    L.ref_set_point_masses.argtypes = [C.c_void_p, dp, dp]
    L.ref_elem_table.argtypes = [C.c_void_p, dp, C.c_int, dp, dp]
    L.ref_set_impact_groups.argtypes = [C.c_void_p, C.c_int, ip32, ip32, ip32, dp]
    L.ref_kinetic_energy.argtypes = [C.c_void_p, dp, dp, dp]
    L.ref_contact_candidates.restype = C.c_int64
    L.ref_contact_candidates.argtypes = [C.c_void_p, dp, ip32, ip32, C.c_int64]
    L.ref_contact_active.restype = C.c_int64
    L.ref_contact_active.argtypes = [C.c_void_p, ip32, ip32, C.c_int64]
    L.ref_contact_force.restype = C.c_int64
    L.ref_contact_force.argtypes = [C.c_void_p, dp, dp]
    L.ref_blacklist.restype = C.c_int64
    L.ref_blacklist.argtypes = [C.c_void_p, ip32, ip32, C.c_int64]
    L.ref_plastic_state.argtypes = [C.c_void_p, dp, dp, dp, u8p]
    L.ref_set_plastic_state.argtypes = [C.c_void_p, dp, dp]
    L.ref_curr_strains.argtypes = [C.c_void_p, dp]
    L.ref_frames.argtypes = [C.c_void_p, dp, dp]
    L.ref_set_grid.argtypes = [C.c_void_p, C.c_int]
    L.ref_expvec.argtypes = [dp, dp]
    L.ref_logmat.argtypes = [dp, dp]
    L.ref_expvecp.argtypes = [dp, dp, dp]
    _LIBS[name] = L
    return L


def _p(a, t=dp):
    return a.ctypes.data_as(t) if a is not None else None


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def expvec(psi):
    R = np.zeros(9)
    lib().ref_expvec(_p(_c(psi)), _p(R))
    return R.reshape(3, 3)


def logmat(R):
    rv = np.zeros(3)
    lib().ref_logmat(_p(_c(R).reshape(-1)), _p(rv))
    return rv


def expvecp(psi, psiP):
    out = np.zeros(9)
    lib().ref_expvecp(_p(_c(psi)), _p(_c(psiP)), _p(out))
    return out.reshape(3, 3)


def expr_eval(expr, t=0.0, i=0.0, omp=False):
    """the oracle's evaluation of a program expression (raises ValueError
    on a syntax error)"""
    v = C.c_double()
    if lib(omp).ref_expr_eval(expr.encode(), float(t), float(i), C.byref(v)):
        raise ValueError(f"bad expression: {expr!r}")
    return v.value


class RefModel:
    """Oracle instance for one Problem (see dynlattice_amd.lattice.Problem)."""

    def __init__(self, p, omp=False, grid=False, init=True, libpath=None):
        self.L = lib(omp, libpath)
        self.p = p
        self._keep = []
        k = self._keep.append
        co = _c(p.coords)
        en = _c(p.elem_nodes, np.int32)
        es = _c(p.elem_strut, np.int32)
        k(co); k(en); k(es)
        mesh = _Mesh(p.n_nodes, p.n_elems, p.n_struts, _p(co), _p(en, ip32), _p(es, ip32))
        s = p.sections
        arrs = [_c(getattr(s, f)) for f in ("area", "I1", "I2", "J", "shear_factor", "young",
                                            "shear_modulus", "density")]
        for a in arrs:
            k(a)
        sec = _Sections(*[_p(a) for a in arrs], s.inertia_correct)
        pl = _Plastic()
        pl.model = 0
        if p.plastic is not None:
            q = p.plastic
            w = _c(q.weights)
            k(w)
            pl.model = 1
            pl.weights = _p(w)
            pl.yield_level = q.yield_level
            pl.kin_coupling[:] = list(q.kin_coupling)
            pl.iso_coupling = q.iso_coupling
            pl.has_iso = int(q.has_iso)
            pl.iso_modulus = q.iso_modulus
            pl.has_kin = int(q.has_kin)
            pl.kin_const[:] = list(q.kin_const)
            pl.kin_ratio = q.kin_ratio
            pl.precision = q.precision
            pl.max_iter = q.max_iter
            pl.exponent[:] = list(q.exponent)
        ct = _Contact()
        if p.contact is not None:
            r = _c(p.contact.radius)
            k(r)
            ct.enabled = 1
            ct.penalty_sts = p.contact.penalty_sts
            ct.penalty_nts = p.contact.penalty_nts
            ct.radius = _p(r)
        opt = _Options(int(p.so3))
        self.h = self.L.ref_create(C.byref(mesh), C.byref(sec), C.byref(pl), C.byref(ct),
                                   C.byref(opt), None)
        self.L.ref_set_grid(self.h, int(grid))
        self._check()
        self.set_bcs(p.cmask, p.cval, p.fext)
        if getattr(p, "point_mass", None) is not None or getattr(p, "point_inertia", None) is not None:
            self.set_point_masses(p.point_mass, p.point_inertia)
        if getattr(p, "programs", None):
            self.set_programs(p.programs, p.dof_prog)
            self.set_time(p.t0, 0)
        if getattr(p, "impact", None):
            self.set_impact_groups(p.impact)
        self.disp = _c(p.disp0).reshape(-1).copy()
        self.velo = _c(p.velo0).reshape(-1).copy()
        if init:
            self.init()

    def _check(self, rc=0):
        e = self.L.ref_last_error(self.h)
        if rc or e:
            raise RuntimeError(f"oracle error {rc or e}")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_destroy(self.h)
            self.h = None

    def init(self):
        self._check(self.L.ref_init(self.h, _p(self.disp)))

    def set_bcs(self, cmask, cval, fext):
        self._cm = _c(cmask, np.uint8).reshape(-1)
        self._cv = _c(cval).reshape(-1)
        self._fe = _c(fext).reshape(-1)
        self.L.ref_set_bcs(self.h, _p(self._cm, u8p), _p(self._cv), _p(self._fe))

    def int_vector(self, disp=None, velo=None, inelastic=True, gyro=False, contact=True):
        d = self.disp if disp is None else _c(disp).reshape(-1)
        v = self.velo if velo is None else _c(velo).reshape(-1)
        f = np.zeros(6 * self.p.n_nodes)
        self._check(self.L.ref_get_int_vector(self.h, _p(d), _p(v), _p(f), int(inelastic),
                                              int(gyro), int(contact and self.p.contact is not None)))
        return f.reshape(-1, 6)

    def output_reaction(self, group, dof=2, disp=None):
        """GroupOutput's reaction at disp (default: the current state): the
        group dof's sum of the output-load-case internal force"""
        d = self.disp if disp is None else _c(disp).reshape(-1)
        g = _c(group, np.int32)
        r = C.c_double()
        self._check(self.L.ref_output_reaction(self.h, _p(d), _p(g, ip32), int(g.shape[0]),
                                               int(dof), C.byref(r)))
        return r.value

    def elem_fint(self, e, disp=None, inelastic=True):
        d = self.disp if disp is None else _c(disp).reshape(-1)
        f = np.zeros(12)
        self._check(self.L.ref_elem_fint(self.h, int(e), _p(d), _p(f), int(inelastic)))
        return f.reshape(2, 6)

    def mass_blocks(self, disp=None, lumped=False):
        d = self.disp if disp is None else _c(disp).reshape(-1)
        b = np.zeros(36 * self.p.n_nodes)
        self.L.ref_mass_blocks(self.h, _p(d), _p(b), int(lumped))
        return b.reshape(-1, 6, 6)

    def leapfrog(self, nsteps, dt=None, group=None, group_dof=2, energy=False):
        dt = self.p.dt if dt is None else dt
        react = np.zeros(nsteps)
        en = np.zeros(3 * nsteps) if energy else None
        nc = np.zeros(nsteps, np.int64)
        g = _c(group if group is not None else np.zeros(0), np.int32)
        rc = self.L.ref_leapfrog(self.h, _p(self.disp), _p(self.velo), float(dt), int(nsteps),
                                 _p(g, ip32), int(g.shape[0]), int(group_dof), _p(react),
                                 int(energy), _p(en), None, _p(nc, ip64))
        self._check(rc)
        out = {"reaction": react, "n_contacts": nc}
        if energy:
            out["energies"] = en.reshape(-1, 3)
        return out

    def set_programs(self, programs, dof_prog):
        """programs: list of lattice.Program; dof_prog [N, 6] (-1: none)"""
        n = len(programs)
        kind = np.array([pr.KINDS[pr.kind] for pr in programs], np.int32)
        abc = np.array([[pr.a, pr.b, pr.c] for pr in programs], np.float64).reshape(-1)
        ex = (C.c_char_p * max(n, 1))(*[pr.expr.encode() for pr in programs])
        dpg = _c(dof_prog if dof_prog is not None else np.full(6 * self.p.n_nodes, -1), np.int32)
        self._prog_keep = (kind, abc, ex, dpg)
        self._check(self.L.ref_set_programs(self.h, n, _p(kind, ip32), _p(abc), ex,
                                            _p(dpg, ip32)))

    def set_time(self, t, k=0):
        self._check(self.L.ref_set_time(self.h, float(t), int(k)))

    def set_mass_update(self, cond):
        """updateWhen of the explicit modules: the mass is re-assembled at the
        start of the steps where cond(t, i) holds (None: every step)"""
        self._check(self.L.ref_set_mass_update(self.h, None if cond is None else cond.encode()))

    def mass_in_use(self):
        """[N, 6, 6] blocks of the last mass update"""
        b = np.zeros(36 * self.p.n_nodes)
        self._check(self.L.ref_mass_in_use(self.h, _p(b)))
        return b.reshape(-1, 6, 6)

    def get_time(self):
        t = C.c_double()
        k = C.c_int64()
        self.L.ref_get_time(self.h, C.byref(t), C.byref(k))
        return t.value, k.value

    def set_impact_groups(self, groups):
        """groups: list of (nodes, dof, weight)"""
        nodes = _c(np.concatenate([np.asarray(g[0]) for g in groups]) if groups else
                   np.zeros(0), np.int32)
        ptr = _c(np.cumsum([0] + [len(g[0]) for g in groups]), np.int32)
        dofs = _c([g[1] for g in groups], np.int32)
        w = _c([g[2] for g in groups])
        self._imp_keep = (nodes, ptr, dofs, w)
        self._check(self.L.ref_set_impact_groups(self.h, len(groups), _p(nodes, ip32), _p(ptr, ip32),
                                                 _p(dofs, ip32), _p(w)))

    def set_point_masses(self, mass=None, inertia=None):
        m = None if mass is None else _c(mass)
        i = None if inertia is None else _c(inertia)
        self._pm_keep = (m, i)
        self._check(self.L.ref_set_point_masses(self.h, _p(m) if m is not None else None,
                                                _p(i) if i is not None else None))

    TABLE_KINDS = {"strain": 0, "mat_strain": 1, "stress": 2, "mat_stress": 3}

    def elem_table(self, name, disp=None):
        """[E, 6] strain / stress table at disp (default: the current state)"""
        d = self.disp if disp is None else _c(disp).reshape(-1)
        out = np.zeros(6 * self.p.n_elems)
        self._check(self.L.ref_elem_table(self.h, _p(d), self.TABLE_KINDS[name], _p(out), None))
        return out.reshape(-1, 6)

    def elem_potential_energy(self, disp=None):
        d = self.disp if disp is None else _c(disp).reshape(-1)
        ep = np.zeros(self.p.n_elems)
        self._check(self.L.ref_elem_table(self.h, _p(d), 0, None, _p(ep)))
        return ep

    def kinetic_energy(self, disp=None, velo=None):
        d = self.disp if disp is None else _c(disp).reshape(-1)
        v = self.velo if velo is None else _c(velo).reshape(-1)
        ek = np.zeros(self.p.n_nodes)
        self._check(self.L.ref_kinetic_energy(self.h, _p(d), _p(v), _p(ek)))
        return ek

    def milne(self, n_attempts, opts, group=None, group_dof=2):
        """n_attempts MilneDevice attempts from the current state; returns a
        dict of per-attempt arrays: dt, err, accepted, reaction, n_contacts,
        dt_next"""
        out = np.zeros(6 * n_attempts)
        g = _c(group if group is not None else np.zeros(0), np.int32)
        rc = self.L.ref_milne(self.h, _p(self.disp), _p(self.velo), C.byref(opts),
                              int(n_attempts), _p(g, ip32), int(g.shape[0]), int(group_dof),
                              _p(out))
        self._check(rc)
        o = out.reshape(-1, 6)
        return {"dt": o[:, 0], "err": o[:, 1], "accepted": o[:, 2] > 0, "reaction": o[:, 3],
                "n_contacts": o[:, 4].astype(np.int64), "dt_next": o[:, 5]}

    def contact_candidates(self, disp=None):
        d = self.disp if disp is None else _c(disp).reshape(-1)
        n = self.L.ref_contact_candidates(self.h, _p(d), None, None, 0)
        a = np.zeros(max(n, 1), np.int32)
        b = np.zeros(max(n, 1), np.int32)
        self.L.ref_contact_candidates(self.h, _p(d), _p(a, ip32), _p(b, ip32), n)
        return a[:n], b[:n]

    def contact_force(self, disp=None):
        d = self.disp if disp is None else _c(disp).reshape(-1)
        f = np.zeros(6 * self.p.n_nodes)
        n = self.L.ref_contact_force(self.h, _p(d), _p(f))
        return f.reshape(-1, 6), n

    def contact_active(self):
        n = self.L.ref_contact_active(self.h, None, None, 0)
        a = np.zeros(max(n, 1), np.int32)
        b = np.zeros(max(n, 1), np.int32)
        self.L.ref_contact_active(self.h, _p(a, ip32), _p(b, ip32), n)
        return a[:n], b[:n]

    def blacklist(self):
        n = self.L.ref_blacklist(self.h, None, None, 0)
        a = np.zeros(max(n, 1), np.int32)
        b = np.zeros(max(n, 1), np.int32)
        self.L.ref_blacklist(self.h, _p(a, ip32), _p(b, ip32), n)
        return a[:n], b[:n]

    def set_plastic_state(self, eps_p, alpha):
        """committed plastic strain [E, 6] and hardening parameters [E, 7]"""
        ep = _c(eps_p).reshape(-1)
        al = _c(alpha).reshape(-1)
        self._check(self.L.ref_set_plastic_state(self.h, _p(ep), _p(al)))

    def plastic_state(self):
        E = self.p.n_elems
        ep = np.zeros(6 * E)
        al = np.zeros(7 * E)
        df = np.zeros(E)
        act = np.zeros(E, np.uint8)
        self.L.ref_plastic_state(self.h, _p(ep), _p(al), _p(df), _p(act, u8p))
        return ep.reshape(E, 6), al.reshape(E, 7), df, act

    def curr_strains(self):
        s = np.zeros(6 * self.p.n_elems)
        self.L.ref_curr_strains(self.h, _p(s))
        return s.reshape(-1, 6)

    def frames(self):
        E = self.p.n_elems
        ln = np.zeros(18 * E)
        s0 = np.zeros(6 * E)
        self.L.ref_frames(self.h, _p(ln), _p(s0))
        return ln.reshape(E, 2, 3, 3), s0.reshape(E, 6)

    def commit(self):
        ep = C.c_double()
        ed = C.c_double()
        self.L.ref_commit(self.h, _p(self.disp), C.byref(ep), C.byref(ed))
        return ep.value, ed.value
