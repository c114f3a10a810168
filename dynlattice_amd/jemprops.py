"""Reader of the reference's property files (.pro): the JEM property syntax
subset its test cases use (tests/*/test*.pro, input.pro, model.pro,
output.pro).

    key.path = value;          numbers, "strings", [arrays], true/false
    key.path += value;         array append / property-set merge
    key.path = other.path;     copy of another property (value or set)
    include "file.pro";        relative to the including file
    "$(key.path)"              substitution of a property's value in strings
    // and /* */ comments

The result is a nested dict.  Values stay as read: strings that hold
expressions ("PI * (5e-3)^2 * ...", "sqrt(12/2e3)") are evaluated where a
number is needed (:func:`number`), with the library's program-expression
compiler, as JEM evaluates them through its property converter.
"""
from __future__ import annotations

import os
import re

_TOKEN = re.compile(r"""
    (?P<ws>\s+) |
    (?P<lcom>//[^\n]*) |
    (?P<bcom>/\*.*?\*/) |
    (?P<str>"(?:[^"\\]|\\.)*") |
    (?P<num>[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?) |
    (?P<id>[A-Za-z_][A-Za-z0-9_]*(?:\.[A-Za-z_][A-Za-z0-9_]*)*) |
    (?P<op>\+=|=|;|\[|\]|,)
""", re.VERBOSE | re.DOTALL)


class ProError(ValueError):
    pass


def _tokens(text, fname):
    pos = 0
    out = []
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise ProError(f"{fname}: unexpected character {text[pos]!r} at {pos}")
        pos = m.end()
        kind = m.lastgroup
        if kind in ("ws", "lcom", "bcom"):
            continue
        out.append((kind, m.group(kind)))
    return out


def get(props: dict, path: str, default=None):
    node = props
    for k in path.split("."):
        if not isinstance(node, dict) or k not in node:
            return default
        node = node[k]
    return node


def _set(props: dict, path: str, value):
    keys = path.split(".")
    node = props
    for k in keys[:-1]:
        if not isinstance(node.get(k), dict):
            node[k] = {}
        node = node[k]
    node[keys[-1]] = value


def _merge(dst: dict, src: dict):
    for k, v in src.items():
        if isinstance(v, dict) and isinstance(dst.get(k), dict):
            _merge(dst[k], v)
        else:
            dst[k] = _copy(v)


def _copy(v):
    if isinstance(v, dict):
        return {k: _copy(x) for k, x in v.items()}
    if isinstance(v, list):
        return [_copy(x) for x in v]
    return v


def _fmt(v):
    if isinstance(v, float):
        return repr(v)
    if isinstance(v, list):
        return "[" + ", ".join(_fmt(x) for x in v) + "]"
    return str(v)


def _subst(s: str, props: dict) -> str:
    def rep(m):
        v = get(props, m.group(1))
        if v is None:
            raise ProError(f"undefined property $({m.group(1)})")
        return _fmt(v)
    return re.sub(r"\$\(([A-Za-z0-9_.]+)\)", rep, s)


def _value(toks, i, props, fname):
    kind, text = toks[i]
    if kind == "str":
        return _subst(bytes(text[1:-1], "utf-8").decode("unicode_escape"), props), i + 1
    if kind == "num":
        v = float(text)
        return (int(v) if re.fullmatch(r"[-+]?\d+", text) else v), i + 1
    if kind == "id":
        if text == "true":
            return True, i + 1
        if text == "false":
            return False, i + 1
        v = get(props, text)
        if v is None:
            raise ProError(f"{fname}: undefined property {text}")
        return _copy(v), i + 1
    if text == "[":
        arr = []
        i += 1
        while toks[i][1] != "]":
            v, i = _value(toks, i, props, fname)
            arr.append(v)
            if toks[i][1] == ",":
                i += 1
        return arr, i + 1
    raise ProError(f"{fname}: value expected, got {text!r}")


def parse(path: str, props: dict | None = None, case_name: str | None = None) -> dict:
    """properties of a .pro file (and its includes); CASE_NAME is the file's
    stem unless given"""
    props = {} if props is None else props
    if case_name is None and "CASE_NAME" not in props:
        props["CASE_NAME"] = os.path.splitext(os.path.basename(path))[0]
    elif case_name is not None:
        props["CASE_NAME"] = case_name
    text = open(path).read()
    toks = _tokens(text, path)
    i = 0
    while i < len(toks):
        kind, text_ = toks[i]
        if kind == "id" and text_ == "include":
            if toks[i + 1][0] != "str":
                raise ProError(f"{path}: include needs a file name")
            inc = _subst(toks[i + 1][1][1:-1], props)
            parse(os.path.join(os.path.dirname(path), inc), props)
            i += 2
            if i < len(toks) and toks[i][1] == ";":
                i += 1
            continue
        if kind != "id":
            raise ProError(f"{path}: property name expected, got {text_!r}")
        key = text_
        op = toks[i + 1][1]
        if op not in ("=", "+="):
            raise ProError(f"{path}: '=' expected after {key}")
        v, i = _value(toks, i + 2, props, path)
        if i < len(toks) and toks[i][1] == ";":
            i += 1
        if op == "=":
            _set(props, key, v)
        else:
            old = get(props, key)
            if old is None:
                _set(props, key, v)
            elif isinstance(old, dict) and isinstance(v, dict):
                _merge(old, v)
            elif isinstance(old, list):
                old.extend(v if isinstance(v, list) else [v])
            else:
                _set(props, key, ([old] if not isinstance(old, list) else old) +
                     (v if isinstance(v, list) else [v]))
    return props


def number(v, **variables) -> float:
    """a numeric property: numbers as they are, strings as expressions
    (JEM function syntax; PI and the given variables, e.g. t) evaluated with
    the library's program-expression compiler"""
    if isinstance(v, bool):
        return float(v)
    if isinstance(v, (int, float)):
        return float(v)
    if isinstance(v, str):
        from .gpu import program_eval
        return program_eval(v, variables.get("t", 0.0), variables.get("i", 0.0))
    raise ProError(f"number expected, got {v!r}")
