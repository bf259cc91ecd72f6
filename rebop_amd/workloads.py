"""The benchmark workloads, built on the device engine.

- ``synthetic_large``: one random mass-action network with 128 species and
  512 reactions (reactant order 0/1/2 with p = .1/.6/.3, reactants uniform
  over species, 0-2 products, rates log-uniform in [1e-3, 1e1], initial
  counts Poisson(100), generator seed 7), tmax = 10, 200 samples.
- ``heterogeneous``: many random networks of different shapes (species
  uniform in [3, 64], reactions uniform in [4, 256], same distributions,
  initial counts Poisson(50), growth-bounded: a positive species weighting
  that no reaction with reactants increases), each with its own tmax chosen so that the mean
  number of events per trajectory spans 10^2 .. 10^5 log-uniformly across the
  networks, all run in one fused launch (``engine.run_jobs``).
"""

from __future__ import annotations

import math

import numpy as np

from . import networks
from .engine import GillespieNet, Rate, run_jobs


def synthetic_large(seed: int = 7):
    """The 128-species, 512-reaction network and its initial counts."""
    return networks.synthetic(128, 512, seed, init_mean=100.0)


def expression_large(n_species: int = 48, n_reactions: int = 96, seed: int = 11):
    """A mixed network beyond the specialised kernels' 32 species / 64
    reactions: the growth-bounded random mass-action network of that size,
    with every reaction whose only reactant is one molecule of X_a given a
    Rate::Expr rate (src/expr.rs) in place of k * X_a, taking turns among a
    Hill repression by another species X_b, an exponential damping by X_b,
    min(X_a, X_b + 1) and a threshold max(X_a - 5, 0).  Each is 0 when X_a
    is 0 and never negative, so counts stay non-negative.

    Returns (spec, names, exprs): spec = (S, [(k, orders, diffs)], x0) is
    the mass-action network (its orders give the operation count per event),
    names[s] = "X<s>", exprs[r] the expression text or None for a reaction
    that keeps its Rate::lma."""
    S, rx, x0 = networks.synthetic(n_species, n_reactions, seed, init_mean=50.0, bounded=True)
    rng = np.random.default_rng(seed + 1)
    names = [f"X{s}" for s in range(S)]
    exprs = []
    turn = 0
    for k, orders, _ in rx:
        if sum(orders) != 1:
            exprs.append(None)
            continue
        a, b = int(np.flatnonzero(orders)[0]), int(rng.integers(0, S))
        c = repr(k)
        exprs.append([f"{c} * X{a} / (1 + (X{b} / 40)^2)",
                      f"{c} * X{a} * exp(-X{b} / 100)",
                      f"{c} * min(X{a}, X{b} + 1)",
                      f"{c} * max(X{a} - 5, 0)"][turn % 4])
        turn += 1
    return (S, rx, x0), names, exprs


def heterogeneous_specs(n_networks: int, seed: int, s_range=(3, 64), r_range=(4, 256),
                        init_mean: float = 50.0):
    """Network specs (S, reactions, x0) of a heterogeneous batch."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n_networks):
        S = int(rng.integers(s_range[0], s_range[1] + 1))
        R = int(rng.integers(r_range[0], r_range[1] + 1))
        specs.append(networks.synthetic(S, R, int(rng.integers(0, 2**31 - 1)), init_mean=init_mean,
                                        bounded=True))
    return specs


def target_events(n_networks: int, lo: float = 1e2, hi: float = 1e5) -> np.ndarray:
    """Mean events per trajectory aimed at by each network, log-uniform over [lo, hi]."""
    if n_networks == 1:
        return np.array([math.sqrt(lo * hi)])
    return lo * (hi / lo) ** (np.arange(n_networks) / (n_networks - 1))


def initial_total_propensity(spec) -> float:
    """make_cumrates' total at the initial counts (falling-factorial LMA)."""
    S, rx, x0 = spec
    tot = 0.0
    for k, orders, _ in rx:
        a = k
        for s in np.flatnonzero(orders):
            n = int(x0[s])
            for f in range(n + 1 - int(orders[s]), n + 1):
                a = a * float(f)
        tot = tot + a
    return tot


def calibrate_tmax(nets, specs, targets, pilot_traj: int = 32, rounds: int = 8, seed: int = 99,
                   log=None):
    """Per-network tmax whose mean event count is near the target.

    Starts from target / (initial total propensity) and rescales by
    target / measured events (clipped to [0.1, 10] per round) on the device,
    a few rounds, all networks in one fused launch per round.
    """
    import time

    tmax = []
    for spec, tgt in zip(specs, targets):
        a0 = initial_total_propensity(spec)
        tmax.append(tgt / a0 if a0 > 0 else 1.0)
    tmax = np.array(tmax)
    for rnd in range(rounds):
        t0 = time.perf_counter()
        for net in nets:
            net.seed(seed + rnd)
        jobs = [dict(net=n, tmax=float(t), nb_steps=1, n_traj=pilot_traj) for n, t in zip(nets, tmax)]
        _, ev = run_jobs(jobs, per_job_events=True)
        ev = ev.astype(np.float64) / pilot_traj
        ratio = np.where(ev > 0, targets / np.maximum(ev, 1e-9), 10.0)
        bad = int(np.sum((ratio <= 0.7) | (ratio >= 1.4)))
        if log:
            log(f"calibration round {rnd}: {ev.sum() * pilot_traj:.3e} events, "
                f"{bad} networks outside [0.7, 1.4] x target, {time.perf_counter() - t0:.2f} s")
        if bad == 0:
            break
        tmax = tmax * np.clip(ratio, 0.1, 10.0)
    return tmax


def build_nets(specs, seed: int):
    """GillespieNet per spec (dense LMA rates, shared seed)."""
    out = []
    for S, rx, x0 in specs:
        g = GillespieNet(x0, sparse=False, seed=seed)
        for k, o, d in rx:
            g.add_reaction(Rate.lma(k, o), d)
        out.append(g)
    return out
