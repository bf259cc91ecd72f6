"""Rate expressions: the reference's ``PExpr`` parser and display, and their
lowering to the postfix bytecode the device library compiles.

Grammar (src/expr.rs:140-273, a winnow parser), reproduced rule for rule:

    expr        = add_sub | term
    term        = mul_div | factor
    factor      = pow | atom
    atom        = exp | max | min | variable | constant | parentheses | neg
    add_sub     = term (space0 [+-] space0 term)*        left fold
    mul_div     = factor (space0 [*/] space0 factor)*     left fold
    pow         = atom space0 '^' space0 atom             (not chained)
    neg         = '-' space0 term
    max / min   = "max" / "min" '(' space0 expr space0 ',' space0 expr space0 ')'
    exp         = "exp" parentheses
    parentheses = '(' space0 expr space0 ')'
    variable    = [A-Za-z_][A-Za-z0-9_]*
    constant    = a float literal: [+-]? (digits ('.' digits?)? | '.' digits)
                  ([eE] [+-]? digits)?, or "nan", or [+-]? "inf" / "infinity"
                  (case-insensitive)
    space0      = [ \\t]*

The whole input must be consumed.  Alternatives backtrack; a repetition
whose element fails after a separator stops before that separator.  The
quirks follow from the grammar as in the reference: ``variable`` comes
before ``constant`` (so ``inf``, ``nan``, ``e`` are names), ``-2`` is a
constant while ``- 2`` is a negation, ``2^3^4`` does not parse, and
``-infect`` does not parse (``-inf`` is taken as a constant).

``str()`` of a node is the reference's ``Display`` (src/expr.rs:108-124).
``to_postfix`` resolves names as ``PExpr::to_expr`` does (src/expr.rs:57-105):
a species name becomes its count, any other name must be a parameter.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Mapping

from . import _native

_OPS = {
    "Neg": _native.REBOP_EXPR_NEG,
    "Add": _native.REBOP_EXPR_ADD,
    "Sub": _native.REBOP_EXPR_SUB,
    "Mul": _native.REBOP_EXPR_MUL,
    "Div": _native.REBOP_EXPR_DIV,
    "Pow": _native.REBOP_EXPR_POW,
    "Max": _native.REBOP_EXPR_MAX,
    "Min": _native.REBOP_EXPR_MIN,
    "Exp": _native.REBOP_EXPR_EXP,
}


class RateParseError(ValueError):
    """The string is not a rate expression (src/expr.rs:132-138)."""


def rust_f64(x: float) -> str:
    """``Display`` of an f64 in Rust: shortest round-trip digits, never an exponent."""
    if x != x:
        return "NaN"
    if x in (float("inf"), float("-inf")):
        return "inf" if x > 0 else "-inf"
    if x == 0:
        return "-0" if str(x)[0] == "-" else "0"
    r = repr(x)
    if "e" in r or "E" in r:
        from decimal import Decimal

        r = format(Decimal(r), "f")
    if r.endswith(".0"):
        r = r[:-2]
    return r


@dataclass(frozen=True)
class Node:
    kind: str  # Constant, Variable, Neg, Add, Sub, Mul, Div, Pow, Max, Min, Exp
    args: tuple = ()
    value: float = 0.0
    name: str = ""

    def __str__(self) -> str:
        k, a = self.kind, self.args
        if k == "Constant":
            return rust_f64(self.value)
        if k == "Variable":
            return self.name
        if k == "Neg":
            return f"(-{a[0]})"
        if k == "Exp":
            return f"exp({a[0]})"
        if k in ("Max", "Min"):
            return f"{k.lower()}({a[0]}, {a[1]})"
        sym = {"Add": "+", "Sub": "-", "Mul": "*", "Div": "/", "Pow": "^"}[k]
        return f"({a[0]} {sym} {a[1]})"


def Constant(v: float) -> Node:
    return Node("Constant", value=float(v))


def Variable(n: str) -> Node:
    return Node("Variable", name=n)


class _Backtrack(Exception):
    pass


class _Parser:
    def __init__(self, s: str):
        self.s = s
        self.n = len(s)

    # -- tokens ---------------------------------------------------------
    def space0(self, i: int) -> int:
        while i < self.n and self.s[i] in " \t":
            i += 1
        return i

    def lit(self, i: int, text: str) -> int:
        if self.s.startswith(text, i):
            return i + len(text)
        raise _Backtrack

    def digits(self, i: int) -> int:
        j = i
        while j < self.n and self.s[j] in "0123456789":
            j += 1
        if j == i:
            raise _Backtrack
        return j

    def constant(self, i: int):
        s, n = self.s, self.n
        # [+-]? (digits ('.' digits?)? | '.' digits) exponent?
        j = i
        if j < n and s[j] in "+-":
            j += 1
        k = None
        if j < n and s[j].isdigit() and s[j].isascii():
            k = self.digits(j)
            if k < n and s[k] == ".":
                k += 1
                if k < n and s[k].isdigit() and s[k].isascii():
                    k = self.digits(k)
        elif j + 1 < n and s[j] == "." and s[j + 1].isdigit() and s[j + 1].isascii():
            k = self.digits(j + 1)
        if k is not None:
            if k < n and s[k] in "eE":
                e = k + 1
                if e < n and s[e] in "+-":
                    e += 1
                # an exponent marker must be followed by digits (no backtracking)
                if not (e < n and s[e].isdigit() and s[e].isascii()):
                    raise RateParseError("exponent without digits")
                k = self.digits(e)
            return Constant(float(s[i:k])), k
        if s[i:i + 3].lower() == "nan":
            return Constant(float("nan")), i + 3
        j = i + 1 if i < n and s[i] in "+-" else i
        for word in ("infinity", "inf"):
            if s[j:j + len(word)].lower() == word:
                return Constant(float(s[i:j + len(word)])), j + len(word)
        raise _Backtrack

    def variable(self, i: int):
        s = self.s
        if i < self.n and (s[i] == "_" or ("a" <= s[i] <= "z") or ("A" <= s[i] <= "Z")):
            j = i + 1
            while j < self.n and (s[j] == "_" or s[j].isascii() and s[j].isalnum()):
                j += 1
            return Variable(s[i:j]), j
        raise _Backtrack

    # -- rules ----------------------------------------------------------
    def alt(self, i: int, *rules):
        for rule in rules:
            try:
                return rule(i)
            except _Backtrack:
                continue
        raise _Backtrack

    def expr(self, i):
        return self.alt(i, self.add_sub, self.term)

    def term(self, i):
        return self.alt(i, self.mul_div, self.factor)

    def factor(self, i):
        return self.alt(i, self.pow, self.atom)

    def atom(self, i):
        return self.alt(i, self.exp, self.max, self.min, self.variable, self.constant,
                        self.parentheses, self.neg)

    def fold(self, i, elem, ops, kinds):
        left, i = elem(i)
        while True:
            j = self.space0(i)
            if j < self.n and self.s[j] in ops:
                op = self.s[j]
                j = self.space0(j + 1)
                try:
                    right, j = elem(j)
                except _Backtrack:
                    return left, i
                left = Node(kinds[op], (left, right))
                i = j
            else:
                return left, i

    def add_sub(self, i):
        return self.fold(i, self.term, "+-", {"+": "Add", "-": "Sub"})

    def mul_div(self, i):
        return self.fold(i, self.factor, "*/", {"*": "Mul", "/": "Div"})

    def pow(self, i):
        a, i = self.atom(i)
        i = self.lit(self.space0(i), "^")
        b, i = self.atom(self.space0(i))
        return Node("Pow", (a, b)), i

    def neg(self, i):
        i = self.lit(i, "-")
        a, i = self.term(self.space0(i))
        return Node("Neg", (a,)), i

    def parentheses(self, i):
        i = self.lit(i, "(")
        e, i = self.expr(self.space0(i))
        i = self.lit(self.space0(i), ")")
        return e, i

    def _pair(self, i, word, kind):
        i = self.lit(i, word)
        i = self.lit(i, "(")
        a, i = self.expr(self.space0(i))
        i = self.lit(self.space0(i), ",")
        b, i = self.expr(self.space0(i))
        i = self.lit(self.space0(i), ")")
        return Node(kind, (a, b)), i

    def max(self, i):
        return self._pair(i, "max", "Max")

    def min(self, i):
        return self._pair(i, "min", "Min")

    def exp(self, i):
        i = self.lit(i, "exp")
        a, i = self.parentheses(i)
        return Node("Exp", (a,)), i


def parse(s: str) -> Node:
    """``str::parse::<PExpr>`` (src/expr.rs:134-138): the whole string or RateParseError."""
    p = _Parser(s)
    try:
        node, i = p.expr(0)
    except _Backtrack:
        raise RateParseError(s) from None
    if i != p.n:
        raise RateParseError(s)
    return node


def to_postfix(node: Node, species: Mapping[str, int], params: Mapping[str, float]):
    """``PExpr::to_expr`` (src/expr.rs:57-105) lowered to postfix bytecode.

    Returns (tokens, constants): tokens are ``op | arg << 8`` (REBOP_EXPR_*).
    Raises ValueError("Parameter k should have a value") for an unknown name.
    """
    code: list[int] = []
    consts: list[float] = []

    def emit(n: Node):
        if n.kind == "Constant":
            code.append(_native.REBOP_EXPR_CONST | (len(consts) << 8))
            consts.append(n.value)
        elif n.kind == "Variable":
            if n.name in species:
                code.append(_native.REBOP_EXPR_SPECIES | (int(species[n.name]) << 8))
            elif n.name in params:
                code.append(_native.REBOP_EXPR_CONST | (len(consts) << 8))
                consts.append(float(params[n.name]))
            else:
                raise ValueError(f"Parameter {n.name} should have a value")
        else:
            for a in n.args:
                emit(a)
            code.append(_OPS[n.kind])

    emit(node)
    return code, consts
