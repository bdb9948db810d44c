"""Reader of the lattice-file dialect of the reference (ufo/mad.py:14-147).

The dialect is a small subset of MAD-X.  Grammar (blanks, tabs, carriage
returns and newlines may separate any two tokens):

    file       = { statement } ;
    statement  = name ":" kind { sep key "=" value } ";"        element
               | name ":" "LINE" "=" "(" name { "," name } ")" ";"
    sep        = "," | any single token, accepted with a warning
    value      = number | "{" [ number { "," number } ] "}" | '"' text '"'
    number     = [ "-" ] digit { digit } [ "." { digit } ]
    name, key  = ( letter | "_" | "-" ) { letter | digit | "_" }

* Numbers have no exponent and no leading ".", the reference's token rule
  (ufo/mad.py:17); a number run into letters (``1e-3``) is refused.
* ``kind`` is case-insensitive; labels and keys are kept as written, and keys
  are the reference's uppercase names (``L``, ``K1``, ``ANGLE``...).  A key an
  element kind does not take raises TypeError, as the reference's
  constructor call does.
* A statement of an unknown kind is skipped, with a warning, up to its ";".
* A line may name only labels defined before it (KeyError otherwise).
* Differences from the reference's tokenizer, none of which changes a file
  it reads: blanks inside vector braces (``{ 0, 0.5 }``, the form its own
  ``dump`` writes) and an empty vector ``{}`` are accepted, and carriage
  returns count as blanks; a string ends at the next quote.

Errors are RuntimeError with the offending text and its line, as the
reference raises RuntimeError for an unexpected token.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any, List

# Element kinds: constructor name in ufo_amd.elements, and for every key the
# dialect allows, the constructor keyword it sets (defaults are the
# constructors' own).
_KINDS = {
    "MARKER": ("Marker", {}),
    "DRIFT": ("Drift", {"L": "length"}),
    "QUADRUPOLE": ("Quadrupole", {"L": "length", "K1": "k1"}),
    "SBEND": ("Sbend", {"L": "length", "ANGLE": "angle", "K1": "k1", "E1": "e1", "E2": "e2", "HGAP": "hgap",
                        "FINT": "fint"}),
    "RBEND": ("Rbend", {"L": "length", "ANGLE": "angle", "K1": "k1", "E1": "e1", "E2": "e2"}),
    "SEXTUPOLE": ("Sextupole", {"L": "length", "K2": "k2", "K2S": "k2s"}),
    "OCTUPOLE": ("Octupole", {"L": "length", "K3": "k3", "K3S": "k3s"}),
    "MULTIPOLE": ("Multipole", {"KNL": "knl", "KSL": "ksl"}),
    "APERTURE": ("Aperture", {"WINDOW": "window"}),
}

_PUNCT = {":": "colon", ",": "comma", "=": "equals", ";": "end", "(": "open", ")": "close"}
_BLANK = " \t\r\n"


@dataclass
class _Tok:
    kind: str          # name, number, vector, string, or one of _PUNCT's values
    value: Any
    line: int

    def describe(self):
        return f"{self.value!r} unexpected (line {self.line})"


def _is_name_start(c):
    return c.isascii() and (c.isalpha() or c in "_-")


def _is_name_char(c):
    return c.isascii() and (c.isalnum() or c == "_")


class _Scanner:
    """Turns the text into tokens, one character class at a time."""

    def __init__(self, text: str):
        self.text = text
        self.pos = 0
        self.line = 1

    def _peek(self, k=0):
        i = self.pos + k
        return self.text[i] if i < len(self.text) else ""

    def _skip_blanks(self):
        while self._peek() and self._peek() in _BLANK:
            if self._peek() == "\n":
                self.line += 1
            self.pos += 1

    def _fail(self, what):
        raise RuntimeError(f"{what!r} unexpected (line {self.line})")

    def _number(self):
        start = self.pos
        if self._peek() == "-":
            self.pos += 1
        while self._peek().isdigit():
            self.pos += 1
        if self._peek() == ".":
            self.pos += 1
            while self._peek().isdigit():
                self.pos += 1
        text = self.text[start:self.pos]
        if _is_name_start(self._peek()) and self._peek() != "-":
            self._fail(text + self._peek())
        return float(text)

    def _vector(self):
        self.pos += 1                                    # "{"
        items: List[float] = []
        while True:
            self._skip_blanks()
            c = self._peek()
            if c == "}":
                self.pos += 1
                return items
            if items:
                if c != ",":
                    self._fail(c)
                self.pos += 1
                self._skip_blanks()
                c = self._peek()
            if not (c.isdigit() or (c == "-" and self._peek(1).isdigit())):
                self._fail(c or "end of file")
            items.append(self._number())

    def _string(self):
        end = self.text.find('"', self.pos + 1)
        if end < 0:
            self._fail(self.text[self.pos:self.pos + 20])
        value = self.text[self.pos + 1:end]
        self.line += value.count("\n")
        self.pos = end + 1
        return value

    def tokens(self):
        while True:
            self._skip_blanks()
            c = self._peek()
            if not c:
                return
            line = self.line
            if c.isdigit() or (c == "-" and self._peek(1).isdigit()):
                yield _Tok("number", self._number(), line)
            elif c == "{":
                yield _Tok("vector", self._vector(), line)
            elif c == '"':
                yield _Tok("string", self._string(), line)
            elif c in _PUNCT:
                self.pos += 1
                yield _Tok(_PUNCT[c], c, line)
            elif _is_name_start(c):
                start = self.pos
                self.pos += 1
                while _is_name_char(self._peek()):
                    self.pos += 1
                yield _Tok("name", self.text[start:self.pos], line)
            else:
                self._fail(c)


class _Reader:
    """Statements over the token stream, building elements and lines."""

    def __init__(self, text: str):
        self.toks = list(_Scanner(text).tokens())
        self.i = 0

    def _next(self):
        if self.i >= len(self.toks):
            raise RuntimeError("unexpected end of file")
        t = self.toks[self.i]
        self.i += 1
        return t

    def _take(self, kind):
        t = self._next()
        if t.kind != kind:
            raise RuntimeError(t.describe())
        return t

    def _keys(self):
        keys = {}
        while True:
            t = self._next()
            if t.kind == "end":
                return keys
            if t.kind != "comma":
                print(f"Warning: expected ',' before a key, got {t.value!r} (line {t.line}); read on")
            key = self._next().value
            self._take("equals")
            v = self._next()
            if v.kind not in ("number", "vector", "string"):
                raise RuntimeError(v.describe())
            keys[key] = v.value

    def _members(self, defined):
        self._take("equals")
        self._take("open")
        names = []
        while True:
            names.append(self._take("name").value)
            t = self._next()
            if t.kind == "close":
                break
            if t.kind != "comma":
                raise RuntimeError(t.describe())
        self._take("end")
        return [defined[n] for n in names]

    def _skip(self):
        while self._next().kind != "end":
            pass

    def read(self):
        from . import elements
        from .lattice import Line
        defined = {}
        while self.i < len(self.toks):
            head = self._next()
            if head.kind != "name":
                raise RuntimeError(head.describe())
            self._take("colon")
            kind = self._take("name").value.upper()
            if kind == "LINE":
                defined[head.value] = Line(head.value, self._members(defined))
            elif kind in _KINDS:
                cls, allowed = _KINDS[kind]
                kwargs = {}
                for key, value in self._keys().items():
                    if key not in allowed:
                        raise TypeError(f"{kind} {head.value}: unexpected key {key!r}")
                    kwargs[allowed[key]] = value
                defined[head.value] = getattr(elements, cls)(head.value, **kwargs)
            else:
                print(f"Warning: unknown statement '{kind}'")
                self._skip()
        return defined


def loads(text: str) -> dict:
    """Elements and lines of a lattice text, ``{label: element or Line}`` in
    the order they are defined."""
    return _Reader(text).read()


def load(path, flatten=True):
    """Read a lattice file (ufo/mad.py:101-147); ``flatten`` is accepted and
    unused, as in the reference."""
    with open(path, "r") as f:
        return loads(f.read())
