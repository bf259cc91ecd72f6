"""``define_system`` -- rebop's macro DSL (src/gillespie_macro.rs:1-171) as a
run-time Python analogue over the device library.

rebop's ``define_system!`` turns a text such as

    rtx rtl rdi rdm rdp;
    Dimers { gene, mRNA, protein, dimer }
    r_tx : gene      => gene + mRNA      @ rtx
    r_tl : mRNA      => mRNA + protein   @ rtl
    r_di : 2 protein => dimer            @ rdi
    r_dm : mRNA      =>                  @ rdm
    r_dp : protein   =>                  @ rdp

into a struct with one ``isize`` field per species, one ``f64`` field per
parameter, ``t`` and an RNG, and the methods ``new`` (species 0, parameters
NaN), ``with_parameters``, ``seed`` and ``advance_until``
(gillespie_macro.rs:59-125).  ``define_system(text)`` returns a Python class
with the same fields and methods; its networks run on network-specialised
kernels compiled at run time (hipRTC) with the macro's own arithmetic
(``rebop_system_new`` in include/rebop_gpu.h):

* the propensity of ``m1 A + m2 B => ... @ rate`` is
  ``rate * lma(m1, A) * lma(m2, B)`` left to right, where ``lma(m, x)`` is the
  integer product ``x (x-1) ... (x-m+1)`` converted to f64 once (_rate_lma!,
  :128-142; ``A + A`` gives A * A, ``2 A`` gives A (A-1));
* the reaction fired is the first whose cumulative sum exceeds
  ``total * U`` strictly (_choice!, :146-170).

Beyond the macro, ``run`` executes the pyo3 grid loop for many copies of
the current state at once and ``batch`` gives N resumable copies.

Rate expressions (``@ ...``) may use numbers, parameters, ``+ - * /``, unary
minus and parentheses; they are evaluated once per ``advance_until`` in f64
(IEEE, as Rust does).  In the macro, a rate expression may also read species,
and then sees their values at the start of the ``advance_until`` call
(gillespie_macro.rs:101-104); that form is refused here with a ValueError.
"""

from __future__ import annotations

import math
import os
import re
import struct
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from ._native import REBOP_SYSTEM_MAX_COUNT
from .engine import Batch, GillespieNet

__all__ = ("define_system", "SystemSpec", "SystemReaction", "parse_system")

# names the generated class uses itself (struct fields t and rng, the methods)
_RESERVED = frozenset(
    {"t", "rng", "new", "seed", "with_parameters", "advance_until", "run", "batch", "network", "rates"}
)

_TOKEN = re.compile(
    r"""
    (?P<ws>\s+|//[^\n]*)
  | (?P<num>\d[\d_]*(?:\.(?![A-Za-z_.])[\d_]*)?(?:[eE][+-]?\d[\d_]*)?(?:_?f64)?)
  | (?P<ident>[A-Za-z_][A-Za-z0-9_]*)
  | (?P<arrow>=>)
  | (?P<punct>[;{},:+\-*/()@])
    """,
    re.X,
)


@dataclass(frozen=True)
class SystemReaction:
    name: str
    reactants: tuple[tuple[str, int], ...]  # (species, coefficient), written order
    products: tuple[tuple[str, int], ...]
    rate: tuple  # expression tree: ("num", v) | ("param", n) | ("neg", e) | (op, a, b)
    rate_text: str


@dataclass(frozen=True)
class SystemSpec:
    name: str
    params: tuple[str, ...]
    species: tuple[str, ...]
    reactions: tuple[SystemReaction, ...]


def _tokens(text: str) -> list[tuple[str, str, int]]:
    out, pos = [], 0
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise ValueError(f"define_system: unexpected character {text[pos]!r} at offset {pos}")
        kind = m.lastgroup
        if kind != "ws":
            out.append((kind, m.group(), pos))
        pos = m.end()
    out.append(("end", "", len(text)))
    return out


def _number(lit: str) -> float:
    s = lit.replace("_", "")
    if s.endswith("f64"):
        s = s[:-3]
    return float(s)


class _Parser:
    def __init__(self, text: str):
        self.text = text
        self.toks = _tokens(text)
        self.i = 0

    def peek(self, k: int = 0):
        return self.toks[min(self.i + k, len(self.toks) - 1)]

    def take(self):
        t = self.toks[self.i]
        self.i += 1
        return t

    def expect(self, value: str):
        t = self.take()
        if t[1] != value:
            raise ValueError(f"define_system: expected {value!r} at offset {t[2]}, found {t[1] or 'end of text'!r}")
        return t

    def ident(self, what: str) -> str:
        t = self.take()
        if t[0] != "ident":
            raise ValueError(f"define_system: expected {what} at offset {t[2]}, found {t[1] or 'end of text'!r}")
        return t[1]

    # `$($param:ident)*; $name:ident { $($species:ident),* } reactions*`
    # (gillespie_macro.rs:51-58)
    def system(self) -> SystemSpec:
        params = []
        while self.peek()[0] == "ident":
            params.append(self.take()[1])
        self.expect(";")
        name = self.ident("the system's name")
        self.expect("{")
        species = []
        if self.peek()[1] != "}":
            species.append(self.ident("a species name"))
            while self.peek()[1] == ",":
                self.take()
                species.append(self.ident("a species name"))
        self.expect("}")
        reactions = []
        while self.peek()[0] != "end":
            reactions.append(self.reaction())
        return SystemSpec(name, tuple(params), tuple(species), tuple(reactions))

    # `$rname: side => side @ $rate:expr` (gillespie_macro.rs:55-58)
    def reaction(self) -> SystemReaction:
        name = self.ident("a reaction name")
        self.expect(":")
        reactants = self.side()
        self.expect("=>")
        products = self.side()
        self.expect("@")
        start = self.peek()[2]
        rate = self.expr()
        end = self.peek()[2]
        return SystemReaction(name, reactants, products, rate, self.text[start:end].strip())

    # `$($($n:literal)? $s:ident)? $(+ $($n:literal)? $s:ident)*`: an optional
    # first term, then `+ term` repeated
    def side(self) -> tuple[tuple[str, int], ...]:
        terms = []
        if self.peek()[0] in ("num", "ident"):
            terms.append(self.term())
        while self.peek()[1] == "+":
            self.take()
            terms.append(self.term())
        return tuple(terms)

    def term(self) -> tuple[str, int]:
        n = 1
        if self.peek()[0] == "num":
            lit = self.take()[1].replace("_", "")
            if not lit.isdigit():
                raise ValueError(f"define_system: stoichiometric coefficient {lit!r} is not an integer")
            n = int(lit)
        return (self.ident("a species name"), n)

    # rate expression: sums of products of (unary minus) atoms
    def expr(self):
        a = self.product()
        while self.peek()[1] in ("+", "-"):
            op = self.take()[1]
            a = (op, a, self.product())
        return a

    def product(self):
        a = self.unary()
        while self.peek()[1] in ("*", "/"):
            op = self.take()[1]
            a = (op, a, self.unary())
        return a

    def unary(self):
        if self.peek()[1] == "-":
            self.take()
            return ("neg", self.unary())
        t = self.take()
        if t[0] == "num":
            return ("num", _number(t[1]))
        if t[0] == "ident":
            return ("param", t[1])
        if t[1] == "(":
            e = self.expr()
            self.expect(")")
            return e
        raise ValueError(f"define_system: expected a rate expression at offset {t[2]}, found {t[1] or 'end of text'!r}")


def parse_system(text: str) -> SystemSpec:
    """Parses a ``define_system!`` body and checks its names."""
    spec = _Parser(text).system()
    seen: dict[str, str] = {}
    for kind, names in (("parameter", spec.params), ("species", spec.species),
                        ("reaction", [r.name for r in spec.reactions])):
        for n in names:
            if n in _RESERVED:
                raise ValueError(f"define_system: {kind} name {n!r} is reserved")
            if n in seen:
                raise ValueError(f"define_system: {kind} name {n!r} is already a {seen[n]} name")
            seen[n] = kind
    sp = set(spec.species)
    for r in spec.reactions:
        for s, m in r.reactants + r.products:
            if s not in sp:
                raise ValueError(f"define_system: reaction {r.name}: unknown species {s!r}")
            if m > REBOP_SYSTEM_MAX_COUNT:
                raise ValueError(f"define_system: reaction {r.name}: coefficient {m} is beyond "
                                 f"{REBOP_SYSTEM_MAX_COUNT}")
        for n in _names(r.rate):
            if n in sp:
                raise ValueError(
                    f"define_system: reaction {r.name}: the rate reads species {n!r}; in the macro it "
                    "would see the count at the start of advance_until -- not supported here")
            if n not in spec.params:
                raise ValueError(f"define_system: reaction {r.name}: unknown parameter {n!r}")
    return spec


def _names(e) -> list[str]:
    if e[0] == "param":
        return [e[1]]
    if e[0] == "num":
        return []
    return [n for sub in e[1:] for n in _names(sub)]


def _eval(e, env: dict[str, float]) -> np.float64:
    op = e[0]
    if op == "num":
        return np.float64(e[1])
    if op == "param":
        return np.float64(env[e[1]])
    if op == "neg":
        return -_eval(e[1], env)
    a, b = _eval(e[1], env), _eval(e[2], env)
    with np.errstate(all="ignore"):  # IEEE results (inf, NaN) as in Rust
        if op == "+":
            return a + b
        if op == "-":
            return a - b
        if op == "*":
            return a * b
        return a / b


def _bits(v: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", float(v)))[0]


class SystemBase:
    """Base of the classes ``define_system`` makes (see the module docstring)."""

    _spec: SystemSpec

    def __init__(self):
        # `new` (gillespie_macro.rs:70-79): species 0, parameters NaN, t 0,
        # an entropy-seeded stream
        for s in self._spec.species:
            setattr(self, s, 0)
        for p in self._spec.params:
            setattr(self, p, math.nan)
        self.t = 0.0
        self._seed = int.from_bytes(os.urandom(8), "little")
        self._draws = 0
        self._net: GillespieNet | None = None
        self._net_key = None

    @classmethod
    def new(cls):
        return cls()

    @classmethod
    def with_parameters(cls, *values: float):
        """``with_parameters`` (gillespie_macro.rs:87-95): every parameter, in declaration order."""
        if len(values) != len(cls._spec.params):
            raise TypeError(f"{cls.__name__}.with_parameters takes {len(cls._spec.params)} values, "
                            f"got {len(values)}")
        o = cls()
        for p, v in zip(cls._spec.params, values):
            setattr(o, p, float(v))
        return o

    def seed(self, seed: int) -> None:
        """``seed`` (gillespie_macro.rs:81-85): restarts the stream of trajectory 0 of ``seed``."""
        self._seed = int(seed) & (2**64 - 1)
        self._draws = 0

    def _counts(self) -> np.ndarray:
        return np.array([int(getattr(self, s)) for s in self._spec.species], dtype=np.int64)

    def rates(self) -> list[float]:
        """Each reaction's rate expression at the current parameter values."""
        env = {p: float(getattr(self, p)) for p in self._spec.params}
        return [float(_eval(r.rate, env)) for r in self._spec.reactions]

    def network(self) -> GillespieNet:
        """The device network for the current parameter values, counts and seed."""
        rates = self.rates()
        key = (tuple(_bits(k) for k in rates), self._seed)
        if self._net is None or self._net_key != key:
            idx = {s: i for i, s in enumerate(self._spec.species)}
            net = GillespieNet.system(self._counts(), self._seed)
            for r, k in zip(self._spec.reactions, rates):
                net.add_system_reaction(k, [(idx[s], m) for s, m in r.reactants],
                                        [(idx[s], m) for s, m in r.products])
            self._net, self._net_key = net, key
        else:
            self._net.set_species(self._counts())
        return self._net

    def advance_until(self, tmax: float) -> None:
        """``advance_until`` (gillespie_macro.rs:97-124) on this one trajectory
        (stream (seed, 0)), resuming from the current counts, time and stream position."""
        b = self.network().batch(1, 0)
        b.set_time(float(self.t))
        b.set_draws(self._draws)
        b.advance_until(float(tmax))
        x = b.get_species()[0]
        for s, v in zip(self._spec.species, x):
            setattr(self, s, int(v))
        self.t = float(b.get_time()[0])
        self._draws = int(b.get_draws()[0])

    def run(self, tmax: float, nb_steps: int, n_trajectories: int = 1, traj_begin: int = 0):
        """The pyo3 grid loop (src/pyo3_gillespie.rs:197-208) for ``n_trajectories``
        copies of the current counts, from t = 0 and the start of their streams
        (seed, traj_begin + j).  Returns ``(times[nb_steps+1], counts[nb_steps+1, S, n])``."""
        return self.network().run(float(tmax), int(nb_steps), int(n_trajectories), int(traj_begin))

    def batch(self, n_trajectories: int, traj_begin: int = 0) -> Batch:
        """``n_trajectories`` resumable copies of the current counts at time ``t``,
        each at the start of its stream (seed, traj_begin + j)."""
        b = self.network().batch(int(n_trajectories), int(traj_begin))
        b.set_time(float(self.t))
        return b

    def __repr__(self) -> str:
        fields = [f"{s}: {getattr(self, s)}" for s in self._spec.species]
        fields += [f"{p}: {getattr(self, p)!r}" for p in self._spec.params]
        fields.append(f"t: {self.t!r}")
        return f"{self._spec.name} {{ {', '.join(fields)} }}"


def define_system(text: str) -> type:
    """Parses a ``define_system!`` body and returns the class it defines."""
    spec = parse_system(text)
    return type(spec.name, (SystemBase,), {"_spec": spec, "__doc__": f"define_system! {spec.name}"})
