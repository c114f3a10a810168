"""Yield conditions of the ElastoPlasticRod material (the ``yieldCond``
property, ElastoPlasticRodMaterial.cpp:131-134) read into the device's yield
family.

The reference compiles the string with JEM's function parser over the
arguments ``dx, dy, dz, rx, ry, rz`` (the stress resultants, named by the
dof types), ``h_0`` (when ``isotropicCoefficient`` is given) and ``h_dx ..
h_rz`` (when ``kinematicTensor`` is given; the hardening forces
q = -H alpha), ElastoPlasticRodMaterial.cpp:69-90.  The device evaluates
the power-sum family

    f = sum_i |w_i (s_i + ck_i q_i)|^p_i - Y (1 + ci q_0)

with analytic gradients.  :func:`to_family` parses the string (no Python
eval: a recursive-descent parser of the JEM syntax subset below), reads the
family's parameters off it by probing, and then checks the fitted family
against the parsed expression at seeded random arguments; a condition that
is not of that form (e.g. hardening inside a denominator,
tests/plastic/test3.pro) is refused.

Syntax: numbers, the argument names, PI, + - * / % ^ (right-associative,
above unary minus as in -x^2 = -(x^2)), unary + - !, comparisons, && ||,
parentheses, if(c,a,b), abs sqrt exp log sin cos tan asin acos atan sinh
cosh tanh floor ceil, min max pow.
"""
from __future__ import annotations

import math
import re

import numpy as np

STRESS = ("dx", "dy", "dz", "rx", "ry", "rz")


class YieldError(ValueError):
    pass


_TOK = re.compile(r"\s*(?:(?P<num>(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?)|"
                  r"(?P<id>[A-Za-z_][A-Za-z0-9_]*)|"
                  r"(?P<op>&&|\|\||<=|>=|==|!=|[-+*/%^()<>!,]))")

_FUN1 = {"abs": abs, "sqrt": math.sqrt, "exp": math.exp, "log": math.log, "sin": math.sin,
         "cos": math.cos, "tan": math.tan, "asin": math.asin, "acos": math.acos,
         "atan": math.atan, "sinh": math.sinh, "cosh": math.cosh, "tanh": math.tanh,
         "floor": math.floor, "ceil": math.ceil}


def _pow(a, b):
    # C's pow: a negative base with a non-integer exponent has no real value
    return a ** b if (a >= 0.0 or b == int(b)) else float("nan")


_FUN2 = {"min": min, "max": max, "pow": _pow}


def _tokens(s):
    pos, out = 0, []
    s = s.rstrip()
    while pos < len(s):
        m = _TOK.match(s, pos)
        if not m or m.end() == pos:
            raise YieldError(f"yield condition: unexpected text at {s[pos:pos + 10]!r}")
        pos = m.end()
        k = m.lastgroup
        out.append((k, m.group(k)))
    return out


class _Parser:
    """recursive descent to a tree of tuples (op, args...)"""

    def __init__(self, text):
        self.t = _tokens(text)
        self.i = 0

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else (None, None)

    def take(self, v=None):
        k, x = self.peek()
        if k is None or (v is not None and x != v):
            raise YieldError(f"yield condition: expected {v!r}")
        self.i += 1
        return k, x

    def parse(self):
        e = self.orx()
        if self.i != len(self.t):
            raise YieldError("yield condition: trailing text")
        return e

    def binary(self, ops, sub):
        e = sub()
        while self.peek()[1] in ops:
            op = self.take()[1]
            e = (op, e, sub())
        return e

    def orx(self):
        return self.binary(("||",), self.andx)

    def andx(self):
        return self.binary(("&&",), self.cmp)

    def cmp(self):
        return self.binary(("<", "<=", ">", ">=", "==", "!="), self.add)

    def add(self):
        return self.binary(("+", "-"), self.mul)

    def mul(self):
        return self.binary(("*", "/", "%"), self.unary)

    def unary(self):
        op = self.peek()[1]
        if op in ("-", "+", "!"):
            self.take()
            return ("u" + op, self.unary())
        return self.power()

    def power(self):
        b = self.atom()
        if self.peek()[1] == "^":
            self.take()
            return ("^", b, self.unary())
        return b

    def atom(self):
        k, x = self.take()
        if k == "num":
            return ("num", float(x))
        if k == "op" and x == "(":
            e = self.orx()
            self.take(")")
            return e
        if k == "id":
            if self.peek()[1] == "(":
                self.take("(")
                args = [self.orx()]
                while self.peek()[1] == ",":
                    self.take(",")
                    args.append(self.orx())
                self.take(")")
                return ("call", x, args)
            return ("var", x)
        raise YieldError(f"yield condition: unexpected {x!r}")


def _ev(e, env):
    op = e[0]
    if op == "num":
        return e[1]
    if op == "var":
        if e[1] == "PI":
            return math.pi
        if e[1] not in env:
            raise YieldError(f"yield condition: unknown argument {e[1]!r}")
        return env[e[1]]
    if op == "call":
        name, args = e[1], [_ev(a, env) for a in e[2]]
        if name == "if" and len(args) == 3:
            return args[1] if args[0] != 0.0 else args[2]
        if name in _FUN1 and len(args) == 1:
            return float(_FUN1[name](args[0]))
        if name in _FUN2 and len(args) == 2:
            return float(_FUN2[name](*args))
        raise YieldError(f"yield condition: unknown function {name}/{len(args)}")
    if op[0] == "u":
        x = _ev(e[1], env)
        return -x if op == "u-" else (x if op == "u+" else float(x == 0.0))
    a, b = _ev(e[1], env), _ev(e[2], env)
    if op == "+":
        return a + b
    if op == "-":
        return a - b
    if op == "*":
        return a * b
    if op == "/":
        return a / b
    if op == "%":
        return math.fmod(a, b)
    if op == "^":
        return _pow(a, b)
    return float({"<": a < b, "<=": a <= b, ">": a > b, ">=": a >= b, "==": a == b,
                  "!=": a != b, "&&": a != 0 and b != 0, "||": a != 0 or b != 0}[op])


def compile_condition(text):
    """the parsed expression as a function of a dict of arguments (a list of
    strings is one string: JEM's += on a string property appends)"""
    if isinstance(text, (list, tuple)):
        text = "".join(text)
    try:
        tree = _Parser(text).parse()
    except RecursionError:
        raise YieldError("yield condition: nested too deeply") from None

    def f(env):
        try:
            return _ev(tree, env)
        except RecursionError:
            raise YieldError("yield condition: nested too deeply") from None
    return f


def to_family(text, has_iso: bool, has_kin: bool, tol: float = 1e-10):
    """Reads ``yieldCond`` into the power-sum family.  Returns dict(weights[6],
    exponent[6], yield_level, kin_coupling[6], iso_coupling); raises
    YieldError when the condition is not of that form."""
    if isinstance(text, (list, tuple)):
        text = "".join(text)
    f = compile_condition(text)
    names = list(STRESS) + (["h_0"] if has_iso else []) + \
        ([f"h_{n}" for n in STRESS] if has_kin else [])

    def F(s=None, q0=0.0, qk=None):
        env = {n: 0.0 for n in names}
        for i, n in enumerate(STRESS):
            env[n] = 0.0 if s is None else float(s[i])
            if has_kin:
                env[f"h_{n}"] = 0.0 if qk is None else float(qk[i])
        if has_iso:
            env["h_0"] = q0
        try:
            return f(env)
        except (ZeroDivisionError, OverflowError, ValueError) as e:
            raise YieldError(f"yield condition not evaluable: {e}") from None

    Y = -F()
    if not Y > 0.0:
        raise YieldError("yield condition: f(0) must be negative (a positive yield level)")
    ci = (-F(q0=1.0) / Y - 1.0) if has_iso else 0.0
    w, p, ck = np.zeros(6), np.ones(6), np.zeros(6)
    for i in range(6):
        e = np.zeros(6)
        e[i] = 1.0
        g1 = F(e) + Y
        if abs(g1) <= 1e-14 * Y:
            # the component does not enter (w = 0); probe a larger value
            e[i] = 1e6
            if abs(F(e) + Y) > 1e-12 * Y:
                raise YieldError(f"yield condition: component {STRESS[i]} is not a power term")
            continue
        x = 1.0
        for _ in range(2):
            # probe at x and 2x; the second pass at the component's capacity
            # (|w x| = 1), where f + Y does not cancel against Y
            e[i] = x
            g1 = F(e) + Y
            e[i] = 2.0 * x
            g2 = F(e) + Y
            if not (g1 > 0.0 and g2 > 0.0):
                raise YieldError(f"yield condition: component {STRESS[i]} is not |w s|^p")
            p[i] = math.log2(g2 / g1)
            w[i] = g1 ** (1.0 / p[i]) / x
            x = 1.0 / w[i]
        if abs(p[i] - round(p[i])) < 1e-9:
            p[i] = float(round(p[i]))  # integer exponents select the device's fast forms
        e[i] = x
        if has_kin:
            # |w ck q|^p at s = 0 gives |ck|; s = q its sign
            qk = np.zeros(6)
            qk[i] = x
            c = max(F(qk=qk) + Y, 0.0) ** (1.0 / p[i]) / (w[i] * x)
            g = max(F(e, qk=qk) + Y, 0.0) ** (1.0 / p[i]) / (w[i] * x)
            ck[i] = c if abs(g - (1.0 + c)) <= abs(g - abs(1.0 - c)) else -c
    fam = dict(weights=w, exponent=p, yield_level=Y, kin_coupling=ck, iso_coupling=ci)
    # check the family against the expression at seeded arguments
    rng = np.random.default_rng(12345)
    scale = np.where(w > 0.0, 1.0 / np.where(w > 0.0, w, 1.0), 1.0)
    for _ in range(64):
        s = rng.uniform(-2.0, 2.0, 6) * scale
        qk = rng.uniform(-0.5, 0.5, 6) * scale if has_kin else np.zeros(6)
        q0 = float(rng.uniform(-0.5, 0.5)) if has_iso else 0.0
        ref = F(s, q0, qk)
        a = s + ck * qk
        fit = float((np.abs(w * a) ** p).sum() - Y * (1.0 + ci * q0))
        if not abs(fit - ref) <= tol * (abs(ref) + Y):
            raise YieldError(
                "yield condition is not of the power-sum form sum |w_i (s_i + c_i h_i)|^p_i "
                f"- Y (1 + c h_0) that the device evaluates: {text.strip()!r}")
    return fam
