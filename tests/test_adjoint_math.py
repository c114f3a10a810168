"""Hand-written vector-Jacobian products of the step, checked on the CPU.

The device math (dl_device.h, dl_contact_math.h, dl_adjoint.h) is compiled
for the host into libdlhost.so.  Every VJP is compared with central finite
differences of the forward function it differentiates, component by
component, in the regimes where the code branches: element relative rotation
tiny (series branch of d log) and large, node updates with every constraint
mask class, contact pairs in the segment-to-segment and the node-to-segment
branches, from either side of the pair.
"""
import ctypes as C
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HOSTLIB = os.path.join(os.path.dirname(HERE), "dynlattice_amd", "libdlhost.so")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(HOSTLIB):
        from dynlattice_amd.build import build
        build()
    L = C.CDLL(HOSTLIB)
    P = C.POINTER(C.c_double)
    L.hc_elem_force.argtypes = [P] * 11
    L.hc_elem_force_vjp.argtypes = [P] * 12
    L.hc_node_advance.argtypes = [P, P, P, P, C.c_double, P, C.c_double, C.c_uint, P, P,
                                  C.c_double, P]
    L.hc_node_advance_vjp.argtypes = [P, P, P, C.c_double, P, C.c_double, C.c_uint, P, P,
                                      C.c_double, P, P, P, P]
    L.hc_pair_forces.argtypes = [P, P, C.c_double, C.c_double, C.c_double, P]
    L.hc_pair_forces.restype = C.c_int
    L.hc_pair_vjp.argtypes = [P, P, C.c_int, C.c_double, C.c_double, C.c_double, P, P, P]
    L.hc_pair_vjp.restype = C.c_int
    return L


def ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def qexp(t):
    th = np.linalg.norm(t)
    if th == 0:
        return np.array([1.0, 0, 0, 0])
    return np.r_[np.cos(th / 2), np.sin(th / 2) / th * np.asarray(t)]


def qmul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def fd_check(fun, x0, grad, h, rtol, atol_frac=1e-9):
    """central differences of the scalar fun at x0 against grad, per component"""
    g = np.empty_like(x0)
    for j in range(x0.size):
        xp = x0.copy(); xp[j] += h[j]
        xm = x0.copy(); xm[j] -= h[j]
        g[j] = (fun(xp) - fun(xm)) / (2 * h[j])
    scale = max(np.abs(g).max(), 1e-300)
    err = np.abs(g - grad)
    bad = err > rtol * np.abs(g) + atol_frac * scale
    assert not bad.any(), (np.nonzero(bad)[0], g[bad], grad[bad])
    return g


# ------------------------------------------------------------------ element
def _element(rng, rel_angle):
    L0 = 0.01 * (1 + rng.random())
    d = unit(rng.normal(size=3))
    Xa = rng.normal(size=3) * 0.01
    Xb = Xa + L0 * d
    qNa = unit(rng.normal(size=4))
    # frame at b: the frame at a turned by a rotation of the requested angle
    ax = unit(rng.normal(size=3))
    qNb = qmul(qexp(rel_angle * ax), qNa)
    ua = rng.normal(size=3) * 1e-4
    ub = rng.normal(size=3) * 1e-4
    qna = unit(np.r_[1.0, rng.normal(size=3) * 0.05])
    qnb = unit(np.r_[1.0, rng.normal(size=3) * 0.05])
    eps0 = rng.normal(size=6) * 1e-3
    Cm = np.array([2e4, 2e4, 6e4, 0.3, 0.3, 0.25]) * (1 + rng.random(6))
    return Xa, ua, qna, qNa, Xb, ub, qnb, qNb, eps0, Cm


@pytest.mark.parametrize("seed,rel_angle", [(0, 0.3), (1, 1.2), (2, 2.8), (3, 1e-7), (4, 0.0)])
def test_element_force_vjp(lib, seed, rel_angle):
    rng = np.random.default_rng(seed)
    Xa, ua, qna, qNa, Xb, ub, qnb, qNb, eps0, Cm = _element(rng, rel_angle)
    if rel_angle == 0.0:
        qnb = qna.copy()  # exactly equal rotations at both ends: phi = 0
    fbar = rng.normal(size=12)

    def f_of(x):
        f = np.zeros(12)
        lib.hc_elem_force(ptr(Xa), ptr(x[0:3]), ptr(x[3:7]), ptr(qNa), ptr(Xb), ptr(x[7:10]),
                          ptr(x[10:14]), ptr(qNb), ptr(eps0), ptr(x[14:20]), ptr(f))
        return float(fbar @ f)

    x0 = np.r_[ua, qna, ub, qnb, Cm]
    out = np.zeros(20)
    lib.hc_elem_force_vjp(ptr(Xa), ptr(ua), ptr(qna), ptr(qNa), ptr(Xb), ptr(ub), ptr(qnb),
                          ptr(qNb), ptr(eps0), ptr(Cm), ptr(fbar), ptr(out))
    h = np.r_[np.full(3, 1e-7), np.full(4, 1e-6), np.full(3, 1e-7), np.full(4, 1e-6), Cm * 1e-6]
    fd_check(f_of, x0, out, h, rtol=2e-5)


def test_element_vjp_rotation_gradient_is_tangent_consistent(lib):
    """the quaternion adjoint must not depend on the tangent splitting: the
    derivative along q -> exp(t w) q equals <qbar, (0, w/2) * q>"""
    rng = np.random.default_rng(7)
    Xa, ua, qna, qNa, Xb, ub, qnb, qNb, eps0, Cm = _element(rng, 0.9)
    fbar = rng.normal(size=12)
    out = np.zeros(20)
    lib.hc_elem_force_vjp(ptr(Xa), ptr(ua), ptr(qna), ptr(qNa), ptr(Xb), ptr(ub), ptr(qnb),
                          ptr(qNb), ptr(eps0), ptr(Cm), ptr(fbar), ptr(out))
    w = unit(rng.normal(size=3))

    def J(t):
        q = qmul(qexp(t * w), qna)
        f = np.zeros(12)
        lib.hc_elem_force(ptr(Xa), ptr(ua), ptr(q), ptr(qNa), ptr(Xb), ptr(ub), ptr(qnb),
                          ptr(qNb), ptr(eps0), ptr(Cm), ptr(f))
        return float(fbar @ f)

    h = 1e-6
    fd = (J(h) - J(-h)) / (2 * h)
    dq = qmul(np.r_[0.0, w / 2], qna)
    assert abs(out[3:7] @ dq - fd) <= 1e-6 * abs(fd) + 1e-12


# ------------------------------------------------------------------ node
def _node(rng, mask):
    q = unit(rng.normal(size=4))
    A = rng.normal(size=(3, 3))
    J0 = A @ A.T + 3 * np.eye(3)
    J0p = np.array([J0[0, 0], J0[1, 1], J0[2, 2], J0[0, 1], J0[0, 2], J0[1, 2]]) * 1e-9
    return dict(f=rng.normal(size=6), q=q, u=rng.normal(size=3) * 1e-3, v=rng.normal(size=6) * 3,
                mt=1e-3 * (1 + rng.random()), J0=J0p, gm=float(rng.integers(1, 5)), mask=mask,
                ac=rng.normal(size=6) * 10, fe=rng.normal(size=6) * 0.1, dt=2e-6)


@pytest.mark.parametrize("mask", [0, 0b000111, 0b111000, 0b101010, 0b010100, 0b111111])
@pytest.mark.parametrize("seed", [0, 1])
def test_node_advance_vjp(lib, mask, seed):
    rng = np.random.default_rng(10 * seed + mask)
    n = _node(rng, mask)
    ub, qb, vb = rng.normal(size=3), rng.normal(size=4), rng.normal(size=6)

    def J_of(x):
        out = np.zeros(13)
        lib.hc_node_advance(ptr(x[0:6]), ptr(x[6:10]), ptr(x[10:13]), ptr(x[13:19]), float(x[19]),
                            ptr(x[20:26]), n["gm"], mask, ptr(n["ac"]), ptr(n["fe"]), n["dt"],
                            ptr(out))
        return float(ub @ out[0:3] + qb @ out[3:7] + vb @ out[7:13])

    x0 = np.r_[n["f"], n["q"], n["u"], n["v"], n["mt"], n["J0"]]
    out = np.zeros(29)
    lib.hc_node_advance_vjp(ptr(n["f"]), ptr(n["q"]), ptr(n["v"]), n["mt"], ptr(n["J0"]), n["gm"],
                            mask, ptr(n["ac"]), ptr(n["fe"]), n["dt"], ptr(ub), ptr(qb), ptr(vb),
                            ptr(out))
    Jb = out[20:29].reshape(3, 3)
    J0p_bar = np.array([Jb[0, 0], Jb[1, 1], Jb[2, 2], Jb[0, 1] + Jb[1, 0], Jb[0, 2] + Jb[2, 0],
                        Jb[1, 2] + Jb[2, 1]])
    grad = np.r_[out[13:19], out[3:7], out[0:3], out[7:13], out[19], J0p_bar]
    h = np.r_[np.full(6, 1e-5), np.full(4, 1e-6), np.full(3, 1e-6), np.full(6, 1e-5),
              n["mt"] * 1e-6, np.abs(n["J0"]) * 1e-6 + 1e-18]
    fd_check(J_of, x0, grad, h, rtol=1e-5)


# ------------------------------------------------------------------ contact
# closed pairs of each branch of the pair logic, found by a seeded random
# search and rounded: (pA, pB) with closest-point parameters >= 0.05 away
# from the branch boundaries |u| = 1
PAIRS = {
    # both closest points inside: segment-to-segment
    "sts": ([0.00635, 0.00729, 0.00251, 0.00047, 0.00363, -0.01182],
            [0.03066, -0.00077, -0.02017, -0.00649, 0.00678, -0.005]),
    # closest point outside A only (uA = -1.42, uB = 0.10)
    "nts_a": ([-0.00067, -0.00241, 0.01027, 0.00296, -0.0032, 0.01734],
              [-0.00057, -0.00302, 0.01375, -0.00363, -0.00324, 0.00523]),
    # closest point outside B only (uA = -0.75, uB = -1.12)
    "nts_b": ([0.00171, 0.00919, -0.00713, -0.00134, 0.00237, 0.01375],
              [0.00116, 0.00804, -0.00504, 0.00358, 0.00415, -0.0125]),
    # both outside: node-to-segment from the end nodes
    "ends": ([-0.01697, -0.01249, 0.00112, -0.00783, -0.00237, 0.00382],
             [-0.01071, -0.00706, -0.00027, -0.00544, -0.00221, 0.00361]),
}


def _pair(kind, rng):
    pA, pB = (np.array(x) + rng.normal(size=6) * 1e-7 for x in PAIRS[kind])
    return pA, pB, 2e-3


@pytest.mark.parametrize("kind", ["sts", "nts_a", "nts_b", "ends"])
@pytest.mark.parametrize("XisA", [1, 0])
def test_contact_pair_vjp(lib, kind, XisA):
    rng = np.random.default_rng(2 * list(PAIRS).index(kind) + XisA)
    pA, pB, R = _pair(kind, rng)
    kS, kN = 1e6, 2e6
    f = np.zeros(12)
    assert lib.hc_pair_forces(ptr(pA), ptr(pB), R, kS, kN, ptr(f)) == 1
    assert np.abs(f).max() > 0
    fbA, fbB = rng.normal(size=6), rng.normal(size=6)
    pX, pY = (pA, pB) if XisA else (pB, pA)
    fbX, fbY = (fbA, fbB) if XisA else (fbB, fbA)

    def J_of(x):
        px, Rx = x[:6], float(x[6])
        a, b = (px, pY) if XisA else (pY, px)
        ff = np.zeros(12)
        lib.hc_pair_forces(ptr(np.ascontiguousarray(a)), ptr(np.ascontiguousarray(b)), Rx, kS, kN,
                           ptr(ff))
        return float(fbA @ ff[:6] + fbB @ ff[6:])

    out = np.zeros(7)
    assert lib.hc_pair_vjp(ptr(pX), ptr(pY), XisA, R, kS, kN, ptr(fbX), ptr(fbY), ptr(out)) == 1
    fd_check(J_of, np.r_[pX, R], out, np.full(7, 1e-9), rtol=1e-5)


def test_contact_pair_vjp_open_pair(lib):
    pA = np.r_[-0.01, 0, 0, 0.01, 0, 0]
    pB = np.r_[0, -0.01, 0.01, 0, 0.01, 0.01]
    out = np.full(7, 3.0)
    z = np.zeros(6)
    assert lib.hc_pair_vjp(ptr(pA), ptr(pB), 1, 2e-3, 1e6, 1e6, ptr(z), ptr(z), ptr(out)) == 0
    assert (out == 3.0).all()


# ------------------------------------------------------------------ plastic
# The cutting plane (return_map) is differentiated as executed: through each
# correction it takes.  Central differences are valid where the number of
# corrections is the same on both sides of the difference; every check
# asserts that.
def _plib(lib):
    P = C.POINTER(C.c_double)
    lib.hc_return_map.argtypes = [C.c_int] + [P] * 8
    lib.hc_return_map.restype = C.c_int
    lib.hc_return_map_vjp.argtypes = [C.c_int] + [P] * 11
    lib.hc_return_map_vjp.restype = C.c_int
    lib.hc_elem_force_pl.argtypes = [C.c_int] + [P] * 16
    lib.hc_elem_force_pl.restype = C.c_int
    lib.hc_elem_force_vjp_pl.argtypes = [C.c_int] + [P] * 15 + [C.c_int] + [P] * 5
    lib.hc_elem_force_vjp_pl.restype = C.c_int
    lib.hc_elem_strain.argtypes = [P] * 10
    return lib


def _params(p, iso, kin, max_iter=20):
    """flat PlasticParams: yield_level, kin_coupling 6, iso_coupling,
    iso_modulus, precision, exponent 6, kin_const 6, kin_ratio, has_iso,
    has_kin, max_iter"""
    return np.r_[1.0, np.full(6, 1.0), 0.4 if iso else 0.0, 50.0 if iso else 0.0, 1e-5,
                 np.full(6, float(p)), np.zeros(6), 0.01, float(iso), float(kin), float(max_iter)]


RM_CASES = [  # (expk, p, iso, kin, overshoot, balanced stiffness)
    (2, 2.0, 0, 1, 1.2, False),   # the crush's quadratic interaction, kinematic hardening
    (2, 2.0, 1, 1, 1.1, False),   # + isotropic hardening
    (1, 1.0, 0, 1, 1.3, True),    # sum-abs
    (0, 1.5, 1, 0, 1.2, False),   # general exponent, isotropic only
    (2, 2.0, 0, 1, 0.9, False),   # inside the surface: elastic, no correction
]
# The reference's cutting plane converges only from near the surface (a
# quadratic trial state 1.5 times outside already cycles past the 20
# iterations with these stiffness ratios, and the sum-abs surface diverges
# once a component changes sign), so the trial states lie where it converges
# -- as in an explicit step, whose strain increments are small.


def _rm_inputs(rng, p, over, balanced):
    Cm = (np.ones(6) if balanced else np.array([2e4, 2e4, 6e4, 0.3, 0.3, 0.25])) * \
        (1 + rng.random(6))
    ey = 1e-3 * (1 + 0.3 * rng.random(6))
    w = 1.0 / (Cm * ey)
    Hk = 0.01 * Cm * (1 + rng.random(6))
    ep0 = rng.normal(size=6) * 0.1 * ey
    al0 = np.r_[abs(rng.normal()) * 1e-3, rng.normal(size=6) * 0.1 * ey]
    # w |C (strain - ep0)| = over |d_i| / ||d||_p, so that sum |.|^p = over^p
    d = (0.5 + 0.5 * rng.random(6)) * rng.choice([-1.0, 1.0], 6)
    strain = ep0 + over * ey * d / (np.abs(d) ** p).sum() ** (1.0 / p)
    return Cm, w, Hk, strain, ep0, al0


@pytest.mark.parametrize("case", RM_CASES)
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_return_map_vjp(lib, case, seed):
    lib = _plib(lib)
    expk, p, iso, kin, over, bal = case
    rng = np.random.default_rng(100 * seed + RM_CASES.index(case))
    pa = _params(p, iso, kin)
    Cm, w, Hk, strain, ep0, al0 = _rm_inputs(rng, p, over, bal)
    sb, epb, alb = rng.normal(size=6) * 1e-3, rng.normal(size=6) * 1e3, rng.normal(size=7) * 1e3

    def run(x):
        ep, al, st = x[6:12].copy(), x[12:19].copy(), np.zeros(6)
        n = lib.hc_return_map(expk, ptr(pa), ptr(x[19:25]), ptr(x[25:31]), ptr(x[31:37]),
                              ptr(x[0:6]), ptr(ep), ptr(al), ptr(st))
        return n, float(sb @ st + epb @ ep + alb @ al)

    x0 = np.r_[strain, ep0, al0, Cm, w, Hk]
    n0, _ = run(x0)
    out = np.zeros(37)
    n = lib.hc_return_map_vjp(expk, ptr(pa), ptr(Cm), ptr(w), ptr(Hk), ptr(strain), ptr(ep0),
                              ptr(al0), ptr(sb), ptr(epb), ptr(alb), ptr(out))
    assert n == n0 and n >= 0
    if over > 1.0:
        assert n >= 1
    else:
        assert n == 0
    h = np.abs(x0) * 1e-7 + 1e-14

    def J_of(x):
        nx, J = run(x)
        assert nx == n0, "correction count changed inside the difference"
        return J

    # out = strain 6, ep0 6, al0 7, C 6, w 6, Hk 6: the order of x0
    fd_check(J_of, x0, out, h, rtol=2e-5, atol_frac=1e-8)


def test_return_map_cases_take_several_corrections(lib):
    """the quadratic and general-exponent cases above reverse more than one
    correction (so the recomputation of the intermediate states is tested),
    the sum-abs case exactly one (its surface is linear between sign changes)"""
    lib = _plib(lib)
    for case in RM_CASES:
        expk, p, iso, kin, over, bal = case
        rng = np.random.default_rng(RM_CASES.index(case))
        Cm, w, Hk, strain, ep0, al0 = _rm_inputs(rng, p, over, bal)
        ep, al, st = ep0.copy(), al0.copy(), np.zeros(6)
        n = lib.hc_return_map(expk, ptr(_params(p, iso, kin)), ptr(Cm), ptr(w), ptr(Hk),
                              ptr(strain), ptr(ep), ptr(al), ptr(st))
        expect = 0 if over < 1 else (1 if expk == 1 else 2)
        assert n >= expect and (expk != 1 or n == 1), (case, n)


def _pl_element(lib, rng, over):
    Xa, ua, qna, qNa, Xb, ub, qnb, qNb, eps0, Cm = _element(rng, 0.7)
    ep0 = rng.normal(size=6) * 1e-4
    strain = np.zeros(6)
    lib.hc_elem_strain(ptr(Xa), ptr(ua), ptr(qna), ptr(qNa), ptr(Xb), ptr(ub), ptr(qnb), ptr(qNb),
                       ptr(eps0), ptr(strain))
    # weights that put the trial state `over` times outside the quadratic surface
    w = over / (np.sqrt(6.0) * np.abs(Cm * (strain - ep0)))
    Hk = 0.01 * Cm
    al0 = np.r_[0.0, rng.normal(size=6) * 1e-5]
    return Xa, ua, qna, qNa, Xb, ub, qnb, qNb, eps0, Cm, w, Hk, ep0, al0


@pytest.mark.parametrize("seed,over,has_r,seed_mode", [(0, 1.2, False, 0), (1, 1.1, True, 0),
                                                       (2, 0.6, True, 0), (3, 1.15, True, 1)])
def test_element_force_vjp_plastic(lib, seed, over, has_r, seed_mode):
    """the element with the return map (step force), the output load case
    (elastic predictor from the committed state) and the returned plastic
    state, all reversed together; seed_mode: the output load case alone"""
    lib = _plib(lib)
    rng = np.random.default_rng(50 + seed)
    pa = _params(2.0, 0, 1)
    Xa, ua, qna, qNa, Xb, ub, qnb, qNb, eps0, Cm, w, Hk, ep0, al0 = _pl_element(lib, rng, over)
    fbar = rng.normal(size=12) * (0.0 if seed_mode else 1.0)
    rbar = rng.normal(size=12) if has_r else np.zeros(12)
    epb = rng.normal(size=6) * (0.0 if seed_mode else 1e2)
    alb = rng.normal(size=7) * (0.0 if seed_mode else 1e2)

    def run(x):
        out = np.zeros(37)
        n = lib.hc_elem_force_pl(2, ptr(pa), ptr(Xa), ptr(x[0:3]), ptr(x[3:7]), ptr(qNa), ptr(Xb),
                                 ptr(x[7:10]), ptr(x[10:14]), ptr(qNb), ptr(eps0), ptr(x[27:33]),
                                 ptr(x[33:39]), ptr(x[39:45]), ptr(x[14:20]), ptr(x[20:27]),
                                 ptr(out))
        return n, float(fbar @ out[0:12] + rbar @ out[12:24] + epb @ out[24:30] + alb @ out[30:37])

    x0 = np.r_[ua, qna, ub, qnb, ep0, al0, Cm, w, Hk]
    n0, _ = run(x0)
    assert (n0 >= 1) == (over > 1.0)
    out = np.zeros(45)
    n = lib.hc_elem_force_vjp_pl(2, ptr(pa), ptr(Xa), ptr(ua), ptr(qna), ptr(qNa), ptr(Xb),
                                 ptr(ub), ptr(qnb), ptr(qNb), ptr(eps0), ptr(Cm), ptr(w), ptr(Hk),
                                 ptr(ep0), ptr(al0), seed_mode, ptr(fbar),
                                 ptr(rbar) if has_r else None, ptr(epb), ptr(alb), ptr(out))
    assert n == (0 if seed_mode else n0)

    def J_of(x):
        nx, J = run(x)
        assert nx == n0, "correction count changed inside the difference"
        return J

    h = np.r_[np.full(3, 1e-9), np.full(4, 1e-8), np.full(3, 1e-9), np.full(4, 1e-8),
              np.full(6, 1e-9), np.full(7, 1e-9), Cm * 1e-7, w * 1e-7,
              Hk * 1e-7]
    fd_check(J_of, x0, out, h, rtol=2e-5, atol_frac=1e-8)
