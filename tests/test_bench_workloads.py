"""Host logic of the benchmark's workloads (no GPU): the define_system! terms
of a mass-action reaction, the Rate::Expr form of the dimer network, and the
oracle builders the CPU baselines use for those modes."""

from __future__ import annotations

import numpy as np

import bench
from rebop_amd import networks


def test_system_terms_of_mass_action_reactions():
    # 2 protein -> dimer: reactant 2 P, product 1 D; gene -> gene + mRNA keeps the gene
    assert bench.system_terms([0, 0, 2, 0], [0, 0, -2, 1]) == ([(2, 2)], [(3, 1)])
    assert bench.system_terms([1, 0, 0, 0], [0, 1, 0, 0]) == ([(0, 1)], [(0, 1), (1, 1)])
    assert bench.system_terms([0, 1, 0, 0], [0, -1, 0, 0]) == ([(1, 1)], [])


def test_system_oracle_net_matches_lma_on_the_dimer():
    """The dimer's integer mass-action products equal Rate::lma's falling
    factorials for these counts, so both oracle nets give the same rates."""
    wl = bench.make_workload("dimer-system", 0)
    sysnet = wl.cpu_make_net()(wl.spec)
    lma = bench.oracle_lma_net(wl.spec)
    for x in ([1, 0, 0, 0], [1, 30, 200, 5], [1, 7, 1, 9]):
        assert [sysnet.rate(r, x) for r in range(5)] == [lma.rate(r, x) for r in range(5)]


def test_expr_workload_rates():
    """The expressions of the Rate::Expr dimer evaluate (in the oracle's
    interpreter) to the mass-action values of the same reactions."""
    wl = bench.make_workload("dimer-expr", 0)
    net = wl.cpu_make_net()(wl.spec)
    for x in ([1, 0, 0, 0], [1, 30, 200, 5], [1, 3, 2, 0]):
        got = [net.rate(r, x) for r in range(5)]
        k = [r[0] for r in networks.dimers()[1]]
        want = [k[0] * x[0], k[1] * x[1], k[2] * x[2] * (x[2] - 1), k[3] * x[1], k[4] * x[2]]
        assert np.allclose(got, want, rtol=1e-15, atol=0), (x, got, want)


def test_events_cpu_baseline_counts_events(monkeypatch):
    monkeypatch.setattr(bench, "CPU_MIN_WALL_S", 0.05)
    wl = bench.make_workload("sir-events", 0)
    r = bench.cpu_baseline_events(wl, wl.cpu_jobs()[1])
    assert r["value"] > 0 and r["cores"] >= 1 and r["kind"] == "port"
    assert r["philox_ln64_events_per_s_trajectories"] > 0


def test_large_expression_workload():
    """The mixed 48x96 workload: its Rate::Expr rates (in the oracle's
    interpreter) are the mass-action rate times the stated factor, 0 when
    the reactant is absent; reactions without an expression keep Rate::lma;
    the network is beyond the specialised kernels' 32 / 64."""
    import math

    wl = bench.make_workload("big-expr", 0)
    S, rx, x0 = wl.spec
    assert S > 32 and len(rx) > 64
    n_expr = sum(e is not None for e in wl.exprs)
    assert n_expr == sum(sum(o) == 1 for _, o, _ in rx) and n_expr > 40
    net = wl.cpu_make_net()(wl.spec)
    lma = bench.oracle_lma_net(wl.spec)
    rng = np.random.default_rng(5)
    for x in (list(x0), [0] * S, rng.integers(0, 120, S).tolist()):
        for r, ((k, o, _), e) in enumerate(zip(rx, wl.exprs)):
            got = net.rate(r, x)
            if e is None:
                assert got == lma.rate(r, x)
                continue
            a = o.index(1)
            b = int(e.rsplit("X", 1)[1].split(" ")[0].split("/")[0].split(")")[0])
            xa, xb = float(x[a]), float(x[b])
            want = {"/ (1": k * xa / (1 + (xb / 40) ** 2), "exp(": k * xa * math.exp(-xb / 100),
                    "min(": k * min(xa, xb + 1), "max(": k * max(xa - 5, 0)}
            key = next(t for t in want if t in e)
            assert math.isclose(got, want[key], rel_tol=1e-14, abs_tol=0), (r, e, x[a], x[b])
            if x[a] == 0:
                assert got == 0.0
