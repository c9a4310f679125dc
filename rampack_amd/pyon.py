"""PYON, the reference's input format (docs/input-file.md, "PYON format"): Python-like literals and object
constructor calls.  parse() returns plain Python values for literals (int, float, bool, None, str, list, dict with
string keys) and Call(name, args, kwargs) for objects; a bare identifier is a call without arguments.

Integers take decimal, 0x and 0b forms; strings use double quotes with the escapes \\" \\\\ \\n \\t; # starts a
comment outside strings; whitespace is free."""
from __future__ import annotations

from dataclasses import dataclass, field


class PyonError(ValueError):
    """A syntax error in a PYON text (the reference reports these as input errors)."""


@dataclass
class Call:
    name: str
    args: list = field(default_factory=list)
    kwargs: dict = field(default_factory=dict)

    def arg(self, position: int, name: str, default=None, required: bool = False):
        """The argument given by keyword `name` or at `position`; `default` when absent."""
        if name in self.kwargs:
            return self.kwargs[name]
        if position is not None and position < len(self.args):
            return self.args[position]
        if required:
            raise PyonError(f"{self.name}: missing argument '{name}'")
        return default


_ESCAPES = {'"': '"', "\\": "\\", "n": "\n", "t": "\t"}
_PUNCT = "()[]{},=:"


def _tokens(text: str):
    i, n = 0, len(text)
    while i < n:
        ch = text[i]
        if ch.isspace():
            i += 1
        elif ch == "#":
            while i < n and text[i] != "\n":
                i += 1
        elif ch in _PUNCT:
            yield ("p", ch)
            i += 1
        elif ch == '"':
            i += 1
            out = []
            while True:
                if i >= n:
                    raise PyonError("unterminated string")
                c = text[i]
                if c == "\\":
                    if i + 1 >= n or text[i + 1] not in _ESCAPES:
                        raise PyonError(f"bad escape at offset {i}")
                    out.append(_ESCAPES[text[i + 1]])
                    i += 2
                elif c == '"':
                    i += 1
                    break
                else:
                    out.append(c)
                    i += 1
            yield ("s", "".join(out))
        elif ch.isdigit() or ch in "+-." or (ch.isalpha() or ch == "_"):
            j = i
            if ch in "+-":
                j += 1
            if j < n and (text[j].isdigit() or text[j] == "."):
                while j < n and (text[j].isalnum() or text[j] in "._" or (text[j] in "+-" and text[j - 1] in "eE")):
                    j += 1
                yield ("n", text[i:j])
            elif ch.isalpha() or ch == "_":
                while j < n and (text[j].isalnum() or text[j] == "_"):
                    j += 1
                yield ("i", text[i:j])
            else:
                raise PyonError(f"unexpected '{ch}' at offset {i}")
            i = j
        else:
            raise PyonError(f"unexpected '{ch}' at offset {i}")


def _number(tok: str):
    t = tok.replace("_", "")
    sign = -1 if t.startswith("-") else 1
    body = t.lstrip("+-")
    try:
        if body.lower().startswith("0x"):
            return sign * int(body[2:], 16)
        if body.lower().startswith("0b"):
            return sign * int(body[2:], 2)
        if any(c in body for c in ".eE"):
            return float(t)
        return int(t)
    except ValueError:
        raise PyonError(f"bad number '{tok}'") from None


class _Parser:
    def __init__(self, text: str):
        self.toks = list(_tokens(text))
        self.pos = 0

    def peek(self):
        return self.toks[self.pos] if self.pos < len(self.toks) else (None, None)

    def take(self, kind=None, value=None):
        tok = self.peek()
        if tok[0] is None or (kind and tok[0] != kind) or (value is not None and tok[1] != value):
            raise PyonError(f"expected {value or kind}, found {tok[1]!r}")
        self.pos += 1
        return tok

    def value(self):
        kind, tok = self.peek()
        if kind == "n":
            self.pos += 1
            return _number(tok)
        if kind == "s":
            self.pos += 1
            return tok
        if kind == "i":
            self.pos += 1
            if tok in ("True", "False"):
                return tok == "True"
            if tok == "None":
                return None
            call = Call(tok)
            if self.peek() == ("p", "("):
                self.take("p", "(")
                while self.peek() != ("p", ")"):
                    k2, t2 = self.peek()
                    if k2 == "i" and self.pos + 1 < len(self.toks) and self.toks[self.pos + 1] == ("p", "="):
                        self.pos += 2
                        if t2 in call.kwargs:
                            raise PyonError(f"{tok}: argument '{t2}' given twice")
                        call.kwargs[t2] = self.value()
                    else:
                        if call.kwargs:
                            raise PyonError(f"{tok}: positional argument after keyword arguments")
                        call.args.append(self.value())
                    if self.peek() != ("p", ")"):
                        self.take("p", ",")
                self.take("p", ")")
            return call
        if (kind, tok) == ("p", "["):
            self.take("p", "[")
            out = []
            while self.peek() != ("p", "]"):
                out.append(self.value())
                if self.peek() != ("p", "]"):
                    self.take("p", ",")
            self.take("p", "]")
            return out
        if (kind, tok) == ("p", "{"):
            self.take("p", "{")
            out = {}
            while self.peek() != ("p", "}"):
                key = self.take("s")[1]
                self.take("p", ":")
                out[key] = self.value()
                if self.peek() != ("p", "}"):
                    self.take("p", ",")
            self.take("p", "}")
            return out
        raise PyonError(f"unexpected {tok!r}")


def parse(text: str):
    """The single PYON value in `text`."""
    p = _Parser(text)
    v = p.value()
    if p.peek()[0] is not None:
        raise PyonError(f"trailing input after the value: {p.peek()[1]!r}")
    return v
