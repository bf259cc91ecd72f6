"""Python API mirroring ``rebop.Gillespie`` on the batched device engine.

Follows ``python/rebop/gillespie.py`` (class ``Gillespie``, :22-160) and the
PyO3 class it wraps (``src/pyo3_gillespie.rs:11-253``): species are named,
indexed in order of first appearance (reactants, then products, then species
only seen in ``init``), a numeric rate is a Law-of-Mass-Action constant and a
string rate is a rate expression (``Rate::Expr``, parsed as the reference's
``PExpr``, names resolved to species counts or ``params`` at ``run``), and
``run`` samples ``nb_steps + 1`` equally spaced times.  ``run`` additionally accepts
``n_trajectories`` and then simulates that many independent trajectories in
one launch; trajectory j draws from the Philox stream (seed, j), where the
seed is derived from ``rng`` exactly as the reference wrapper does
(``gillespie.py:139-140``).
"""

from __future__ import annotations

import warnings
from typing import Mapping, Sequence

import numpy as np

from . import expr as _expr
from .engine import GillespieNet, Rate, RateExpr
from .result import make_dataset

_INIT_WARNING = (
    "Some species are not involved in any reactions. You should probably instead use parameters."
)


class Gillespie:
    """Reaction system composed of species and reactions."""

    def __init__(self) -> None:
        self._species: dict[str, int] = {}
        self._init: dict[str, int] = {}
        # (rate, reactants, products), as pyo3_gillespie.rs:15; rate is a float
        # (PRate::Lma) or a parsed expression node (PRate::Expr)
        self._reactions: list[tuple[float | _expr.Node, list[str], list[str]]] = []

    # -- model building (pyo3_gillespie.rs:76-117) ------------------------
    def add_species(self, species: str) -> None:
        """Register a species (pyo3_gillespie.rs:76-80)."""
        if species not in self._species:
            self._species[species] = len(self._species)

    def nb_species(self) -> int:
        return len(self._species)

    def nb_reactions(self) -> int:
        return len(self._reactions)

    def add_reaction(
        self,
        rate: float | str,
        reactants: Sequence[str],
        products: Sequence[str],
        reverse_rate: float | str | None = None,
    ) -> None:
        """Add a reaction (pyo3_gillespie.rs:86-113).

        A number is the Law-of-Mass-Action constant (``PRate::Lma``); a
        string is parsed as a rate expression (``PRate::Expr``), whose value
        is the propensity.  An unparsable string raises
        ``ValueError("Rate expression not understood")``.
        """
        rate_ = _prate(rate)
        reactants = [str(r) for r in reactants]
        products = [str(p) for p in products]
        for r in reactants:
            self.add_species(r)
        for p in products:
            self.add_species(p)
        self._reactions.append((rate_, reactants, products))
        if reverse_rate is not None:
            self.add_reaction(reverse_rate, products, reactants, None)

    def set_init(self, init: Mapping[str, int]) -> None:
        """Set the initial counts (pyo3_gillespie.rs:119-134); warns on unknown species."""
        warning = False
        clean: dict[str, int] = {}
        for name, value in init.items():
            v = int(value)
            if v < 0:
                raise OverflowError("can't convert negative int to unsigned")
            clean[str(name)] = v
        for name in clean:
            if name not in self._species:
                warning = True
                self.add_species(name)
        self._init = clean
        if warning:
            raise UserWarning(_INIT_WARNING)

    def __str__(self) -> str:
        """Textual form (pyo3_gillespie.rs:239-252)."""
        s = f"{len(self._species)} species and {len(self._reactions)} reactions\n"
        for rate, reactants, products in self._reactions:
            shown = str(rate) if isinstance(rate, _expr.Node) else f"LMA({_expr.rust_f64(rate)})"
            s += " + ".join(reactants) + " --> " + " + ".join(products) + f" @ {shown}\n"
        return s

    # -- network assembly (pyo3_gillespie.rs:161-196) -----------------------
    def _network(self, seed: int, sparse: bool, params: Mapping[str, float]) -> GillespieNet:
        S = len(self._species)
        x0 = np.zeros(S, dtype=np.int64)
        for name, value in self._init.items():
            if name in self._species:
                x0[self._species[name]] = value
        net = GillespieNet(x0, sparse=sparse, seed=seed)
        for rate, reactants, products in self._reactions:
            orders = np.zeros(S, dtype=np.uint32)
            actions = np.zeros(S, dtype=np.int64)
            for r in reactants:
                orders[self._species[r]] += 1
                actions[self._species[r]] -= 1
            for p in products:
                actions[self._species[p]] += 1
            if isinstance(rate, _expr.Node):
                # PRate::to_gillespie_rate -> PExpr::to_expr (pyo3_gillespie.rs:46-60)
                code, consts = _expr.to_postfix(rate, self._species, params)
                net.add_reaction(RateExpr.from_postfix(code, consts), actions)
            else:
                net.add_reaction(Rate.lma(rate, orders), actions)
        return net

    def run(
        self,
        init: Mapping[str, int],
        tmax: float,
        nb_steps: int,
        *,
        params: Mapping[str, float] | None = None,
        rng=None,
        sparse: bool | None = None,
        var_names: Sequence[str] | None = None,
        n_trajectories: int | None = None,
    ):
        """Run until ``tmax`` with ``nb_steps`` steps (gillespie.py:96-160, pyo3 :143-238).

        Returns an ``xarray.Dataset`` when xarray is importable, else a
        ``rebop_amd.result.Dataset`` with the same access pattern.  With
        ``n_trajectories`` set, every variable gains a trailing
        ``trajectory`` dimension.  ``nb_steps == 0`` returns every reaction
        of one trajectory (pyo3_gillespie.rs:209-223).
        """
        params = {} if params is None else dict(params)
        rng_ = np.random.default_rng(rng)
        seed = int(rng_.integers(np.iinfo(np.uint64).max, dtype=np.uint64))
        try:
            self.set_init(init)
        except UserWarning as e:
            warnings.warn(e, stacklevel=2)
        for p in params:
            if p in self._species:
                raise ValueError(f"Species {p} cannot also be a parameter.")
        if var_names is not None:
            for v in var_names:
                if v not in self._species:
                    raise KeyError(v)
            save = [(v, self._species[v]) for v in var_names]
        else:
            save = list(self._species.items())
        net = self._network(seed, bool(sparse) if sparse is not None else False, params)
        nb = int(nb_steps)
        if nb < 0:
            raise OverflowError("can't convert negative int to unsigned")
        if nb == 0:
            if n_trajectories not in (None, 1):
                raise ValueError("nb_steps == 0 returns every reaction of a single trajectory")
            times, counts = net.run_all_events(float(tmax))
            data = {name: counts[:, idx] for name, idx in save}
            return make_dataset(times, data, batched=False)
        n = 1 if n_trajectories is None else int(n_trajectories)
        times, out = net.run(float(tmax), nb, n_traj=n)
        if n_trajectories is None:
            data = {name: out[:, idx, 0] for name, idx in save}
            return make_dataset(times, data, batched=False)
        data = {name: out[:, idx, :] for name, idx in save}
        return make_dataset(times, data, batched=True)


def _prate(rate):
    """``PRate::new`` (pyo3_gillespie.rs:39-44): a number is an LMA constant, a
    string a rate expression."""
    if isinstance(rate, str):
        try:
            return _expr.parse(rate)
        except _expr.RateParseError:
            raise ValueError("Rate expression not understood") from None
    if isinstance(rate, (bool, int, float, np.integer, np.floating)):
        return float(rate)
    raise TypeError(f"rate must be a number or a string, not {type(rate).__name__}")
