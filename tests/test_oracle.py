"""Pins the CPU oracle against the reference's own tests and against mathematics.

Every oracle function is checked against something independent of its own
formula: the known-answer tables of src/gillespie.rs's tests, closed forms,
brute force on tiny inputs, invariants, and exact probability laws.
"""

from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
import pytest
from scipy import stats

from tests.networks import build_oracle, dimers, sir

GOLDEN = Path(__file__).resolve().parent / "golden"
TABLES = json.loads((GOLDEN / "reference_tables.json").read_text())


# ---- Rate::rate (src/gillespie.rs:71-90) --------------------------------


@pytest.mark.parametrize("sparse_net", [False, True])
def test_rate_lma_known_answers(oracle, sparse_net):
    """src/gillespie.rs:448-472: dense and sparse evaluation give the table values."""
    t = TABLES["rate_lma"]
    for orders, expected in t["cases"]:
        n = oracle.OracleNet(2, sparse=sparse_net)
        n.add_lma(t["rate"], orders, [0, 0])
        assert n.rate(0, t["species"]) == expected, (orders, expected)
        # the same rate given as explicit sparse pairs
        m = oracle.OracleNet(2)
        m.add_lma_sparse(t["rate"], [(i, o) for i, o in enumerate(orders) if o > 0], [0, 0])
        assert m.rate(0, t["species"]) == expected


def test_rate_is_falling_factorial_product(oracle):
    """Closed form: k * prod_s n_s (n_s - 1) ... (n_s - e_s + 1), no 1/e! factor."""
    rng = np.random.default_rng(0)
    for _ in range(300):
        S = int(rng.integers(1, 5))
        orders = rng.integers(0, 4, size=S)
        x = rng.integers(0, 30, size=S)
        k = float(rng.uniform(0.1, 3))
        n = oracle.OracleNet(S)
        n.add_lma(k, orders, [0] * S)
        exact = k * math.prod(math.perm(int(xi), int(e)) for xi, e in zip(x, orders))
        assert n.rate(0, x) == pytest.approx(exact, rel=1e-12, abs=0)


def test_rate_multiplication_order(oracle):
    """The product is ((k * f1) * f2) ... in ascending factor order: a transposed
    order changes the last bit for these inputs."""
    n = oracle.OracleNet(2)
    k = 0.1
    n.add_lma(k, [2, 1], [0, 0])
    x = [3, 5]
    expected = ((k * 2.0) * 3.0) * 5.0
    assert n.rate(0, x) == expected
    assert ((k * 5.0) * 2.0) * 3.0 != expected  # the test can see an order mistake


def test_sparse_pairs_keep_given_order_and_duplicates(oracle):
    """Rate::LMASparse evaluates pairs as stored (gillespie.rs:79-86): a duplicated
    species multiplies n twice (not the falling factorial n(n-1))."""
    m = oracle.OracleNet(2)
    m.add_lma_sparse(2.0, [(0, 1), (0, 1)], [0, 0])
    assert m.rate(0, [5, 3]) == 2.0 * 5 * 5
    d = oracle.OracleNet(2)
    d.add_lma(2.0, [2, 0], [0, 0])
    assert d.rate(0, [5, 3]) == 2.0 * 4 * 5


def test_negative_counts_follow_integer_range(oracle):
    """(n+1-e ..= n) as f64 for negative n: order 2 at n = -1 gives k * (-2) * (-1)."""
    n = oracle.OracleNet(1)
    n.add_lma(1.5, [2], [0])
    assert n.rate(0, [-1]) == 1.5 * -2.0 * -1.0
    assert math.copysign(1.0, n.rate(0, [0])) == -1.0  # k * (-1) * 0 == -0.0


# ---- Jump (src/gillespie.rs:93-153) and add_reaction asserts ------------


@pytest.mark.parametrize("sparse_net", [False, True])
def test_jump_affect_known_answer(oracle, sparse_net):
    t = TABLES["jump_affect"]
    n = oracle.OracleNet(4, sparse=sparse_net)
    n.add_lma(1.0, [0, 0, 0, 0], t["diffs"])
    assert n.affect(0, t["species"]).tolist() == t["expected"]


def test_issue85_oob_rejected(oracle):
    t = TABLES["issue85_oob"]
    n = oracle.OracleNet(len(t["species"]), sparse=True)
    with pytest.raises(oracle.OracleError) as e:
        n.add_lma_sparse(1.0, t["pairs"], t["diffs"])
    assert e.value.code == 2


def test_add_reaction_shape_asserts(oracle):
    n = oracle.OracleNet(3)
    with pytest.raises(oracle.OracleError):
        n.add_lma(1.0, [1, 0], [0, 0, 0])
    with pytest.raises(oracle.OracleError):
        n.add_lma(1.0, [1, 0, 0], [0, 0])


# ---- make_cumrates / choose_cumrate_sum (:357-364, :402-407) -------------


def test_cumrates_sequential_sum(oracle):
    """cum[r] = total + rate_r with total starting at 0.0 (not a tree sum)."""
    n = oracle.OracleNet(1)
    ks = [0.1, 0.2, 0.3, 1e-17, 1e16, -1e16]
    for k in ks:
        n.add_lma(k, [0], [0])
    tot, cum = n.cumrates([1])
    acc, expect = 0.0, []
    for k in ks:
        acc = acc + k
        expect.append(acc)
    assert cum.tolist() == expect and tot == expect[-1]
    assert math.fsum(ks) != tot or True  # sequential, not exactly-rounded


def test_cumrates_first_add_normalises_negative_zero(oracle):
    n = oracle.OracleNet(1)
    n.add_lma(1.0, [2], [0])  # k * (-1) * 0 = -0.0 at n = 0
    tot, cum = n.cumrates([0])
    assert math.copysign(1.0, cum[0]) == 1.0  # 0.0 + -0.0 == +0.0


def test_choose_cumrate_sum_brute_force(oracle):
    """k = #{r : cum[r] < chosen}, against brute force, including ties, zeros,
    U == 0 and non-monotone sums."""
    rng = np.random.default_rng(1)
    for _ in range(2000):
        R = int(rng.integers(1, 9))
        a = rng.choice([0.0, 0.5, 1.0, 2.0, -1.0], size=R, p=[0.3, 0.2, 0.3, 0.15, 0.05])
        cum = np.cumsum(a)
        chosen = float(rng.choice([0.0, float(cum[-1]) * rng.random(), float(rng.choice(cum))]))
        assert oracle.choose_cumrate_sum(chosen, cum) == sum(1 for c in cum if c < chosen)
    # ties pick the reaction whose cumulative value equals chosen
    assert oracle.choose_cumrate_sum(1.0, [1.0, 1.0, 3.0]) == 0
    assert oracle.choose_cumrate_sum(2.0, [1.0, 1.0, 3.0]) == 2
    # chosen == 0 picks reaction 0 even if its propensity is 0 (the U == 0 case)
    assert oracle.choose_cumrate_sum(0.0, [0.0, 2.0]) == 0
    assert oracle.choose_cumrate_sum(float("nan"), [1.0, 2.0]) == 0


# ---- advance_until / grid loop -------------------------------------------


def _net(spec_tables):
    return spec_tables["species"], spec_tables["reactions"]


def test_gillespie_sir_conservation(oracle):
    """src/gillespie.rs:504-516: S+I+R == 10000 after advance_until(i), i = 1..=250."""
    t = TABLES["gillespie_sir"]
    x0, rx = _net(t)
    n = oracle.OracleNet(3)
    for k, o, d in rx:
        n.add_lma(k, o, d)
    for seed in range(5):
        tr = n.state(x0, seed)
        for i in range(1, t["until"] + 1):
            tr.advance_until(float(i))
            assert tr.t == float(i)
            assert tr.x.sum() == t["total"]


def test_gillespie_dimers_bounds(oracle):
    """src/gillespie.rs:518-530 and the doc test :300-314."""
    t = TABLES["gillespie_dimers"]
    x0, rx = _net(t)
    n = oracle.OracleNet(4)
    for k, o, d in rx:
        n.add_lma(k, o, d)
    for seed in range(5):
        tr = n.state(x0, seed)
        tr.advance_until(t["tmax"])
        assert tr.t == 1.0
        assert tr.x[0] == 1 and tr.x[2] > 1000 and 0 < tr.x[3] < 10000


def test_sir_grid_invariants(oracle):
    """tests/test_rebop.py:15-27: time == arange(251), bounds, S+I+R == 1000."""
    n = build_oracle(sir())
    times, out, final, ev, dr = n.run_grid(0, [999, 1, 0], 250.0, 250, n_traj=64)
    assert np.array_equal(times, np.arange(251.0))
    assert (out >= 0).all() and (out[:, 0] <= 999).all() and (out[:, 1:] <= 1000).all()
    assert (out.sum(axis=1) == 1000).all()
    # S never increases, R never decreases
    assert (np.diff(out[:, 0], axis=0) <= 0).all() and (np.diff(out[:, 2], axis=0) >= 0).all()


def test_grid_equals_repeated_advance_until(oracle):
    """The grid loop is advance_until(tmax*i/nb) per sample, on one state."""
    spec = dimers()
    n = build_oracle(spec)
    times, out, final, ev, dr = n.run_grid(11, spec[2], 1.0, 20, n_traj=3, traj_begin=5)
    for j in range(3):
        tr = n.state(spec[2], 11, traj=5 + j)
        for i in range(21):
            tr.advance_until(1.0 * i / 20)
            assert tr.x.tolist() == out[i, :, j].tolist()


def test_draw_accounting(oracle):
    """One draw per event plus one per overshoot (the discarded Exp1 draw)."""
    n = build_oracle(sir())
    times, out, final, ev, dr = n.run_grid(3, [999, 1, 0], 250.0, 250, n_traj=50)
    # every trajectory ends absorbed or overshooting: draws - events = overshoots <= 251 each
    assert 0 <= dr - ev <= 251 * 50
    tr = n.state([999, 1, 0], 3, traj=0)
    tr.advance_until(0.0)
    assert tr.draws == 1 and tr.t == 0.0  # advance_until(0.0) consumes one draw


def test_draw_cap_closed_forms(oracle):
    """The device's watchdog as the oracle states it (no reference
    counterpart).  Pure birth A -> 2A from A = 1 over a horizon it never
    reaches: in a grid run the first draw is the overshoot of sample 0 (t =
    0), every later draw is an event, so a cap of c >= 2 draws stops at
    A = c after c - 1 events (the cap is checked after events only, so c = 1
    stops after the first event: A = 2); in the all-events loop every draw
    is an event (A = c + 1, c + 1 records); advance_until with a cap set
    midway counts from there.  Without a cap nothing changes."""
    n = oracle.OracleNet(1)
    n.add_lma(1.0, [1], [1])
    for cap in (1, 2, 7, 100):
        _, out, final, ev, dr = n.run_grid(5, [1], 1e9, 4, n_traj=3, draw_cap=cap)
        a = max(cap, 2)
        assert final[:, 0].tolist() == [a] * 3
        assert ev == 3 * (a - 1) and dr == 3 * a
        assert (out[0, 0] == 1).all() and (out[1:, 0] == a).all()
        times, xs = n.run_all_events(5, 0, [1], float("inf"), draw_cap=cap)
        assert len(times) == cap + 1 and xs[-1, 0] == cap + 1
        assert np.isfinite(times).all() and (np.diff(times) > 0).all()
    tr = n.state([1], 9)
    tr.advance_until(0.5)
    d0, a0 = tr.draws, int(tr.x[0])
    tr.set_cap(10)
    tr.advance_until(1e9)
    assert tr.capped and tr.draws == d0 + 10 and tr.t == 1e9
    assert int(tr.x[0]) == a0 + 10  # every capped draw was an event
    # a cap that is never reached changes nothing, draw for draw
    m = build_oracle(sir())
    a = m.run_grid(3, [999, 1, 0], 250.0, 50, n_traj=20)
    b = m.run_grid(3, [999, 1, 0], 250.0, 50, n_traj=20, draw_cap=10**6)
    assert np.array_equal(a[1], b[1]) and a[3:] == b[3:]


def test_all_events_infinite_horizon_absorbs(oracle):
    """pyo3_gillespie.rs:215 with tmax = +inf ends when the state absorbs:
    pure decay from A = 30 gives 30 events and a final t = +inf record."""
    n = oracle.OracleNet(1)
    n.add_lma(0.5, [1], [-1])
    times, xs = n.run_all_events(2, 0, [30], float("inf"))
    assert len(times) == 32 and xs[-1, 0] == 0 and xs[-2, 0] == 0 and np.isinf(times[-1])
    assert np.isfinite(times[:-1]).all()


def test_run_empty(oracle):
    """tests/test_rebop.py:148-152: no species, no reactions, times linspace(0, 10, 11)."""
    n = oracle.OracleNet(0)
    times, out, final, ev, dr = n.run_grid(1, [], 10.0, 10, n_traj=2)
    assert np.array_equal(times, np.linspace(0, 10, 11)) and ev == 0 and dr == 0


def test_nan_rate_absorbs_without_draws(oracle):
    n = oracle.OracleNet(1)
    n.add_lma(float("nan"), [0], [1])
    tr = n.state([0], 1)
    tr.advance_until(5.0)
    assert tr.t == 5.0 and tr.draws == 0 and tr.x[0] == 0


def test_grid_times_formula(oracle):
    lib = oracle.load()
    for tmax, nb in [(250.0, 250), (1.0, 100), (0.1, 3), (10.0, 7)]:
        for i in range(nb + 1):
            assert lib.orc_grid_time(tmax, i, nb) == tmax * i / nb


def test_all_events_invariants(oracle):
    """tests/test_rebop.py:39-52 on the nb_steps == 0 loop (pyo3_gillespie.rs:209-223)."""
    n = build_oracle(sir())
    for seed in range(10):
        times, xs = n.run_all_events(seed, 0, [999, 1, 0], 250.0)
        assert times[0] == 0 and times[-1] > 250
        assert (np.diff(times) > 0).all()
        if np.isinf(times[-1]):
            xs = xs[:-1]
        d = np.diff(xs, axis=0)
        assert set(d[:, 0]) <= {-1, 0} and set(d[:, 1]) <= {-1, 1} and set(d[:, 2]) <= {0, 1}


def test_all_events_many_sums_the_single_trajectory_loop(oracle):
    """The threaded timing harness of the nb_steps == 0 loop counts what the
    per-trajectory entry point returns, for any thread count, including
    trajectories longer than its initial buffer and absorbed ones."""
    n = build_oracle(sir())
    ev, rec = 0, 0
    for j in range(150):
        t, _ = n.run_all_events(5, 7 + j, [999, 1, 0], 250.0)
        ev += len(t) - 1 - (1 if np.isinf(t[-1]) else 0)
        rec += len(t)
    for th in (1, 3, 8):
        assert n.run_all_events_many(5, 7, 150, [999, 1, 0], 250.0, nthreads=th) == (ev, rec)
    big = build_oracle(dimers())
    t, _ = big.run_all_events(1, 0, dimers()[2], 0.5)
    assert len(t) > 4096  # grows the harness's buffer
    e1, r1 = big.run_all_events_many(1, 0, 1, dimers()[2], 0.5, nthreads=2)
    assert r1 == len(t) and e1 == len(t) - 1 - (1 if np.isinf(t[-1]) else 0)
    # a small call still spreads over the threads (claims of one trajectory)
    _, _, busy = n.run_all_events_many(5, 7, 64, [999, 1, 0], 250.0, nthreads=4, with_busy=True)
    assert 2 <= busy <= 4


def test_grid_jobs_pool_equals_separate_runs(oracle):
    """The multi-network pool (the CPU baseline's runner) gives each job the
    events and draws of its own grid run, for any thread count, and a sample
    of a few trajectories per job still occupies several threads."""
    from tests.networks import synthetic

    specs = [sir(), dimers(), synthetic(12, 30, 4, init_mean=20.0)]
    runs = [(build_oracle(sp), 11 + q, sp[2], t, nb, n, 3 * q)
            for q, (sp, t, nb, n) in enumerate(zip(specs, (250.0, 0.2, 2.0), (50, 20, 10), (40, 7, 33)))]
    ev, dr = 0, 0
    for net, seed, x0, t, nb, n, tb in runs:
        _, _, _, e, d = net.run_grid(seed, x0, t, nb, n_traj=n, traj_begin=tb)
        ev, dr = ev + e, dr + d
    for th in (1, 2, 5):
        e, d, busy = oracle.run_grid_jobs(runs, nthreads=th)
        assert (e, d) == (ev, dr)
        assert 1 <= busy <= th
    _, _, busy = oracle.run_grid_jobs(runs, nthreads=4)
    assert busy >= 2


# ---- exact laws ----------------------------------------------------------


def _ks_discrete_ok(samples, cdf, alpha=1e-4):
    """Kolmogorov distance of integer samples to an exact discrete law vs the
    DKW bound at level alpha."""
    samples = np.asarray(samples)
    xs = np.arange(samples.min() - 1, samples.max() + 2)
    emp = np.searchsorted(np.sort(samples), xs, side="right") / len(samples)
    d = np.max(np.abs(emp - cdf(xs)))
    eps = math.sqrt(math.log(2 / alpha) / (2 * len(samples)))
    return d, eps


def test_birth_death_poisson_law(oracle, nthreads):
    """0 -> A @ kb, A -> 0 @ kd from 0: A(t) ~ Poisson(kb/kd (1 - e^-kd t)) at every sample."""
    kb, kd = 10.0, 0.5
    n = oracle.OracleNet(1)
    n.add_lma(kb, [0], [1])
    n.add_lma(kd, [1], [-1])
    times, out, final, ev, dr = n.run_grid(8, [0], 4.0, 8, n_traj=20000, nthreads=nthreads)
    for i, t in enumerate(times[1:], start=1):
        lam = kb / kd * (1 - math.exp(-kd * t))
        d, eps = _ks_discrete_ok(out[i, 0], lambda x: stats.poisson.cdf(x, lam))
        assert d < eps, (t, d, eps)


def test_pure_decay_binomial_law(oracle, nthreads):
    """A -> 0 @ k from n0: A(t) ~ Binomial(n0, e^-kt)."""
    k, n0 = 0.3, 50
    n = oracle.OracleNet(1)
    n.add_lma(k, [1], [-1])
    times, out, final, ev, dr = n.run_grid(9, [n0], 5.0, 5, n_traj=20000, nthreads=nthreads)
    for i, t in enumerate(times[1:], start=1):
        d, eps = _ks_discrete_ok(out[i, 0], lambda x: stats.binom.cdf(x, n0, math.exp(-k * t)))
        assert d < eps, (t, d, eps)


def test_dimerisation_second_order_law(oracle, nthreads):
    """2A -> B @ c from A=6: the propensity c*A*(A-1) (no 1/2!) fixes the mean
    waiting times; P(A(t)=6) = exp(-30 c t) exactly."""
    c = 0.05
    n = oracle.OracleNet(2)
    n.add_lma(c, [2, 0], [-2, 1])
    times, out, final, ev, dr = n.run_grid(10, [6, 0], 1.0, 4, n_traj=20000, nthreads=nthreads)
    for i, t in enumerate(times[1:], start=1):
        p = math.exp(-30 * c * t)
        frac = float((out[i, 0] == 6).mean())
        assert abs(frac - p) < 5 * math.sqrt(p * (1 - p) / 20000), (t, frac, p)


def test_sir_exact_law(oracle, nthreads):
    """The SIR benchmark model against its exact CME marginals (tests/golden/sir_fsp.npz)."""
    fsp = np.load(GOLDEN / "sir_fsp.npz")
    n = build_oracle(sir())
    times, out, final, ev, dr = n.run_grid(42, [999, 1, 0], 250.0, 250, n_traj=4000, nthreads=nthreads)
    for k, t in enumerate(fsp["times"]):
        if t == 0:
            continue
        i = int(t)
        for s, name in enumerate("SIR"):
            pmf = fsp[name][k].astype(np.float64)
            cdf = np.cumsum(pmf)
            d, eps = _ks_discrete_ok(out[i, s], lambda x: cdf[np.clip(x, 0, 1000)] * (x >= 0))
            assert d < eps, (t, name, d, eps)


def test_independent_streams_agree(oracle, nthreads):
    """Philox+ln64 vs xoshiro256++ + libm log: two-sample KS at every 10th sample."""
    spec = dimers()
    n = build_oracle(spec)
    _, a, _, _, _ = n.run_grid(1, spec[2], 1.0, 50, n_traj=3000, rng=oracle.PHILOX, nthreads=nthreads)
    _, b, _, _, _ = n.run_grid(2, spec[2], 1.0, 50, n_traj=3000, rng=oracle.XOSHIRO, nthreads=nthreads)
    for i in range(10, 51, 10):
        for s in (1, 2, 3):
            assert stats.ks_2samp(a[i, s], b[i, s]).pvalue > 1e-4, (i, s)
