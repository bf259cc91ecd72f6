"""Reaction networks of the benchmark workloads.

Each builder returns (nb_species, [(rate, orders, diffs)], x0); ``build``
turns one into a ``GillespieNet``.
"""

from __future__ import annotations

import numpy as np


def sir(transmission: float = 1e-4, recovery: float = 0.01):
    """S + I -> 2 I @ 1e-4, I -> R @ 0.01; init S=999, I=1 (tests/test_rebop.py:8-12, :18)."""
    return 3, [
        (transmission, [1, 1, 0], [-1, 1, 0]),
        (recovery, [0, 1, 0], [0, -1, 1]),
    ], [999, 1, 0]


def dimers():
    """Gene expression with dimerisation (src/gillespie.rs:302-308)."""
    return 4, [
        (25.0, [1, 0, 0, 0], [0, 1, 0, 0]),
        (1000.0, [0, 1, 0, 0], [0, 0, 1, 0]),
        (0.001, [0, 0, 2, 0], [0, 0, -2, 1]),
        (0.1, [0, 1, 0, 0], [0, -1, 0, 0]),
        (1.0, [0, 0, 1, 0], [0, 0, -1, 0]),
    ], [1, 0, 0, 0]


def michaelis_menten(kf=0.01, kb=0.1, kcat=0.1, e0=50, s0=200):
    """E + S -> ES, ES -> E + S, ES -> E + P (species E, S, ES, P)."""
    return 4, [
        (kf, [1, 1, 0, 0], [-1, -1, 1, 0]),
        (kb, [0, 0, 1, 0], [1, 1, -1, 0]),
        (kcat, [0, 0, 1, 0], [1, 0, -1, 1]),
    ], [e0, s0, 0, 0]


def synthetic(n_species: int, n_reactions: int, seed: int, init_mean: float = 100.0,
              bounded: bool = False):
    """Random mass-action network: reactant order 0/1/2 with p = .1/.6/.3, reactants
    uniform over species (with replacement), 0-2 products uniform over species,
    rates log-uniform in [1e-3, 1e1]; init counts Poisson(init_mean).

    With ``bounded``, every species gets a mass in {1, 2, 3} and the products
    of a reaction with reactants are redrawn until they weigh no more than
    the reactants: the total mass then grows only through the order-0
    (inflow) reactions, linearly in time, so no trajectory blows up
    exponentially and the event count of a time span stays moderate.
    """
    rng = np.random.default_rng(seed)
    mass = rng.integers(1, 4, size=n_species) if bounded else None
    rx = []
    for _ in range(n_reactions):
        order = int(rng.choice(3, p=[0.1, 0.6, 0.3]))
        orders = np.zeros(n_species, dtype=np.int64)
        for s in rng.integers(0, n_species, size=order):
            orders[s] += 1
        while True:
            diffs = -orders.copy()
            prods = rng.integers(0, n_species, size=int(rng.integers(0, 3)))
            for s in prods:
                diffs[s] += 1
            if mass is None or order == 0 or mass[prods].sum() <= (mass * orders).sum():
                break
        k = float(10.0 ** rng.uniform(-3, 1))
        rx.append((k, orders.tolist(), diffs.tolist()))
    x0 = rng.poisson(init_mean, size=n_species).astype(np.int64).tolist()
    return n_species, rx, x0


def build(spec, seed, sparse=False):
    """A ``GillespieNet`` for a network spec."""
    from .engine import GillespieNet, Rate

    S, rx, x0 = spec
    g = GillespieNet(x0, sparse=sparse, seed=seed)
    for k, o, d in rx:
        g.add_reaction(Rate.lma(k, o), d)
    return g
