"""Reader for the MAD-X dialect the reference understands (ufo/mad.py).

Statements are ``LABEL: KIND, KEY=value, ...;`` and
``LABEL: LINE = (A, B, ...);``.  Values are numbers (no exponent, as in the
reference's token rule, ufo/mad.py:17), ``{a, b, ...}`` vectors or quoted
strings.  Keys are the reference's uppercase names.  Unknown statement kinds
are skipped with a warning.
"""
from __future__ import annotations

import re
from collections import namedtuple

Token = namedtuple("Token", ["type", "value"])

_TOKENS = [
    ("PARAMETER", r'(-?\d+(\.\d*)?)|(\s*\{(-?\d+(\.\d*)?\s*,?\s*)*\})|(".*")'),
    ("OPEN", r"\("),
    ("CLOSE", r"\)"),
    ("COLON", r":"),
    ("COMMA", r","),
    ("ASSIGN", r"="),
    ("DEFERRED_ASSIGN", r":="),
    ("END", r";"),
    ("ID", r"[A-Za-z_\-][A-Za-z0-9_]*"),
    ("NEWLINE", r"\n"),
    ("SKIP", r"[ \t\r]+"),
    ("MISMATCH", r"."),
]
_REGEX = re.compile("|".join(f"(?P<{k}>{v})" for k, v in _TOKENS))


def tokenize(text):
    for m in _REGEX.finditer(text):
        kind = m.lastgroup
        value = m.group()
        if kind == "PARAMETER":
            if '"' in value:
                value = value.replace('"', "")
            elif "{" in value:
                body = value.replace("{", "").replace("}", "")
                value = [float(v) for v in body.split(",") if v.strip()]
            else:
                value = float(value)
        elif kind in ("NEWLINE", "SKIP"):
            continue
        elif kind == "MISMATCH":
            raise RuntimeError(f"{value!r} unexpected")
        yield Token(kind, value)


def _expect(tokens, kind):
    tok = next(tokens)
    if tok.type != kind:
        raise RuntimeError(f"{tok.value!r} unexpected")
    return tok


def _element_keys(tokens):
    keys = {}
    while True:
        tok = next(tokens)
        if tok.type == "END":
            return keys
        if tok.type != "COMMA":
            print("There shoud have been a comma, but I forgive you")
        name = next(tokens).value
        _expect(tokens, "ASSIGN")
        keys[name] = _expect(tokens, "PARAMETER").value


def _line_items(tokens, table, line):
    _expect(tokens, "ASSIGN")
    _expect(tokens, "OPEN")
    while True:
        line.append(table[_expect(tokens, "ID").value])
        tok = next(tokens)
        if tok.type == "CLOSE":
            break
        if tok.type != "COMMA":
            raise RuntimeError(f"{tok.value!r} unexpected")
    _expect(tokens, "END")
    return line


def _skip_statement(tokens):
    while next(tokens).type != "END":
        pass


def _constructors():
    from . import elements as el
    return {
        "MARKER": lambda label: el.Marker(label),
        "DRIFT": lambda label, L=0.0: el.Drift(label, length=L),
        "QUADRUPOLE": lambda label, L=0.0, K1=0.0: el.Quadrupole(label, length=L, k1=K1),
        "SBEND": lambda label, L=0.0, ANGLE=0.0, K1=0.0, E1=0.0, E2=0.0, HGAP=0.0, FINT=0.0:
            el.Sbend(label, length=L, angle=ANGLE, k1=K1, e1=E1, e2=E2, hgap=HGAP, fint=FINT),
        "RBEND": lambda label, L=0.0, ANGLE=0.0, K1=0.0, E1=0.0, E2=0.0:
            el.Rbend(label, length=L, angle=ANGLE, k1=K1, e1=E1, e2=E2),
        "SEXTUPOLE": lambda label, L=0.0, K2=0.0, K2S=0.0: el.Sextupole(label, length=L, k2=K2, k2s=K2S),
        "OCTUPOLE": lambda label, L=0.0, K3=0.0, K3S=0.0: el.Octupole(label, length=L, k3=K3, k3s=K3S),
        "MULTIPOLE": lambda label, KNL=(), KSL=(): el.Multipole(label, knl=KNL, ksl=KSL),
        "APERTURE": lambda label, WINDOW=0: el.Aperture(label, window=WINDOW),
    }


def loads(text):
    """Parse MAD-X-dialect text into ``{label: element or Line}``."""
    from .lattice import Line
    ctors = _constructors()
    table = {}
    tokens = tokenize(text)
    while True:
        try:
            tok = next(tokens)
        except StopIteration:
            break
        if tok.type != "ID":
            raise RuntimeError(f"{tok.value!r} unexpected")
        label = tok.value
        _expect(tokens, "COLON")
        kind = _expect(tokens, "ID").value.upper()
        if kind in ctors:
            table[label] = ctors[kind](label, **_element_keys(tokens))
        elif kind == "LINE":
            table[label] = _line_items(tokens, table, Line(label))
        else:
            print(f"Warning: unknown statement '{kind}'")
            _skip_statement(tokens)
    return table


def load(path, flatten=True):
    """Read a lattice file (ufo/mad.py:101-147)."""
    with open(path, "r") as f:
        return loads(f.read())
