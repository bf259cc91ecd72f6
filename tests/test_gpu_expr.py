"""Rate expressions (Rate::Expr) on the device: network-specialised kernels
with compiled expressions and the LDS kernel's interpreter (networks beyond
the specialised kernels' size, fused batches), bit for bit against the
oracle's interpreter, the reference's arbitrary-rate tests through the Python
API, and the route that refuses expression networks (rate tables)."""

from __future__ import annotations

import numpy as np
import pytest

from rebop_amd import Gillespie, _native
from rebop_amd import expr as E
from rebop_amd.engine import GillespieNet, Rate, RateExpr, run_jobs

pytestmark = pytest.mark.gpu
K = _native

# (S, x0, names, [(rate, orders or None, diffs)]): rate is a float (LMA) or a string
HILL = (
    2, [0, 0], {"M": 0, "P": 1},
    [
        ("20 / (1 + (P / 40)^2.5)", None, [1, 0]),      # repressed transcription
        (5.0, [1, 0], [0, 1]),                           # translation (LMA)
        (0.5, [1, 0], [-1, 0]),
        (0.1, [0, 1], [0, -1]),
        ("max(P - 100, 0) * 0.01 * exp(-M / 50)", None, [0, -1]),
        ("min(M, 3) * 0.2 + -0.05 * M ^ 0", None, [0, 1]),
    ],
)
REF2 = (  # tests/test_rebop.py:137-145
    4, [0, 1000, 0, 0], {"A": 0, "B": 1, "C": 2, "D": 3},
    [
        (14.0, [0, 0, 0, 0], [1, 0, 0, 0]),
        (0.1, [1, 1, 0, 0], [-1, -1, 1, 0]),
        (0.01, [0, 0, 1, 0], [1, 1, -1, 0]),
        ("0.2 * B * C / (5 + C)", None, [0, -1, 0, 1]),
    ],
)


def _build(net, seed):
    from oracle.pyoracle import OracleNet

    S, x0, names, rx = net
    g = GillespieNet(np.array(x0), seed=seed)
    o = OracleNet(S)
    for rate, orders, diffs in rx:
        if isinstance(rate, str):
            code, consts = E.to_postfix(E.parse(rate), names, {})
            g.add_reaction(RateExpr.from_postfix(code, consts), diffs)
            o.add_expr(code, consts, diffs)
        else:
            g.add_reaction(Rate.lma(rate, orders), diffs)
            o.add_lma(rate, orders, diffs)
    return g, o


def _device_run(g, tmax, nb, n, kernel, traj_begin=0):
    import torch

    out = torch.zeros((nb + 1, g.nb_species(), n), dtype=torch.int64, device="cuda")
    st = g.run_device(tmax, nb, n, out.data_ptr(), traj_begin=traj_begin, kernel=kernel)
    torch.cuda.synchronize()
    return out.cpu().numpy(), st


@pytest.mark.parametrize("kernel", [K.REBOP_KERNEL_AUTO, K.REBOP_KERNEL_PLAIN, K.REBOP_KERNEL_PERSISTENT])
@pytest.mark.parametrize("name,net,tmax,nb,n", [
    ("hill", HILL, 20.0, 40, 512),
    ("ref2", REF2, 100.0, 100, 256),
])
def test_expression_networks_bit_exact(name, net, tmax, nb, n, kernel, nthreads):
    g, o = _build(net, 99)
    out, st = _device_run(g, tmax, nb, n, kernel, traj_begin=5)
    assert st.kernel & K.REBOP_KERNEL_SPECIALIZED
    _, oout, _, oev, odr = o.run_grid(99, net[1], tmax, nb, n_traj=n, traj_begin=5, nthreads=nthreads)
    assert np.array_equal(out, oout), np.argwhere(out != oout)[:5].tolist()
    assert st.events == oev and st.draws == odr
    assert oev > 100 * n  # the expressions were exercised


def test_expression_all_events_bit_exact():
    g, o = _build(HILL, 3)
    ts, xs = g.run_all_events(5.0, n_traj=24, traj_begin=2)
    for j in range(24):
        ot, ox = o.run_all_events(3, 2 + j, HILL[1], 5.0)
        assert np.array_equal(ts[j].view(np.uint64), ot.view(np.uint64)), j
        assert np.array_equal(xs[j], ox), j


def test_expression_equal_to_lma_is_bit_identical():
    """'0.5 * A' is the same f64 product as LMA(0.5, A) (device side)."""
    a = GillespieNet(np.array([40, 0]), seed=8)
    a.add_reaction(Rate.lma(0.5, [1, 0]), [-1, 1])
    a.add_reaction(Rate.lma(2.0, [0, 0]), [1, 0])
    b = GillespieNet(np.array([40, 0]), seed=8)
    for text, d in [("0.5 * A", [-1, 1]), ("2", [1, 0])]:
        code, consts = E.to_postfix(E.parse(text), {"A": 0, "B": 1}, {})
        b.add_reaction(RateExpr.from_postfix(code, consts), d)
    ta, oa = a.run(10.0, 20, n_traj=1000)
    tb, ob = b.run(10.0, 20, n_traj=1000)
    assert np.array_equal(oa, ob)


@pytest.mark.parametrize("name,net,tmax,nb,n", [
    ("hill", HILL, 20.0, 40, 256),
    ("ref2", REF2, 100.0, 100, 128),
])
def test_expression_networks_on_the_lds_kernel_bit_exact(name, net, tmax, nb, n, nthreads):
    """The LDS kernel's interpreter (asked for explicitly on networks the
    specialised kernel also takes) equals the oracle bit for bit."""
    g, o = _build(net, 41)
    out, st = _device_run(g, tmax, nb, n, K.REBOP_KERNEL_LDS, traj_begin=3)
    assert st.kernel & K.REBOP_KERNEL_LDS
    _, oout, _, oev, odr = o.run_grid(41, net[1], tmax, nb, n_traj=n, traj_begin=3, nthreads=nthreads)
    assert np.array_equal(out, oout), np.argwhere(out != oout)[:5].tolist()
    assert st.events == oev and st.draws == odr


def _large_expression_network(S=40, R=80, seed=12):
    """S species, R reactions (beyond the specialised kernels' 32 / 64): a
    ring of conversions and decays, every fifth reaction with a Hill- or
    exp-type rate expression."""
    rng = np.random.default_rng(seed)
    names = {f"X{i}": i for i in range(S)}
    x0 = [int(v) for v in rng.integers(10, 60, S)]
    rx = []
    for r in range(R):
        a, b = int(rng.integers(0, S)), int(rng.integers(0, S))
        d = [0] * S
        if a == b:
            d[a] = -1
        else:
            d[a], d[b] = -1, 1
        if r % 5 == 0:
            text = [f"2 * X{a} / (1 + (X{b} / 30)^2)", f"0.05 * X{a} * exp(-X{b} / 80)",
                    f"max(X{a} - 20, 0) * 0.1", f"min(X{a}, X{b}) * 0.03 + 0.01"][(r // 5) % 4]
            rx.append((text, None, d))
        else:
            orders = [0] * S
            orders[a] = 1
            if r % 3 == 0 and a != b:
                orders[b] = 1
            rx.append((float(rng.uniform(0.01, 0.5)) / (10.0 if r % 3 == 0 else 1.0), orders, d))
    # inflow so that the network does not run dry
    for i in range(0, S, 8):
        d = [0] * S
        d[i] = 1
        rx.append((5.0, [0] * S, d))
    return (S, x0, names, rx)


def test_large_expression_network_runs_on_the_lds_kernel(nthreads):
    """A network with rate expressions beyond the specialised kernels' size
    runs (AUTO) on the LDS kernel, bit for bit against the oracle."""
    net = _large_expression_network()
    g, o = _build(net, 77)
    assert g.nb_species() == 40
    n, tmax, nb = 64, 2.0, 20
    out, st = _device_run(g, tmax, nb, n, K.REBOP_KERNEL_AUTO, traj_begin=11)
    assert st.kernel & K.REBOP_KERNEL_LDS
    _, oout, _, oev, odr = o.run_grid(77, net[1], tmax, nb, n_traj=n, traj_begin=11, nthreads=nthreads)
    assert np.array_equal(out, oout), np.argwhere(out != oout)[:5].tolist()
    assert st.events == oev and st.draws == odr
    assert oev > 50 * n
    # asking for the specialised kernel is refused: it does not fit one
    import torch
    buf = torch.zeros((nb + 1, 40, n), dtype=torch.int64, device="cuda")
    with pytest.raises(_native.RebopError) as e:
        g.run_device(tmax, nb, n, buf.data_ptr(), kernel=K.REBOP_KERNEL_SPECIALIZED)
    assert e.value.code == K.REBOP_ERR_UNSUPPORTED


def test_large_mixed_expression_workload_bit_exact(nthreads):
    """The benchmark's 48x96 mixed network (Rate::Expr on 53 reactions,
    Rate::lma on the rest) at a few hundred trajectories of its full horizon:
    AUTO runs it on the LDS kernel, bit for bit against the oracle."""
    from rebop_amd import workloads

    (S, rx, x0), names, exprs = workloads.expression_large()
    species = {n: i for i, n in enumerate(names)}
    net = (S, x0, species, [(e, None, d) if e is not None else (k, o, d)
                            for e, (k, o, d) in zip(exprs, rx)])
    g, o = _build(net, 42)
    n, tmax, nb = 384, 5.0, 50
    out, st = _device_run(g, tmax, nb, n, K.REBOP_KERNEL_AUTO, traj_begin=5)
    assert st.kernel & K.REBOP_KERNEL_LDS
    _, oout, _, oev, odr = o.run_grid(42, x0, tmax, nb, n_traj=n, traj_begin=5, nthreads=nthreads)
    assert np.array_equal(out, oout), np.argwhere(out != oout)[:5].tolist()
    assert st.events == oev and st.draws == odr
    assert oev > 1000 * n


def test_expression_networks_in_a_fused_batch(nthreads):
    """Expression networks in one fused launch with a mass-action network:
    each job equals its own oracle run."""
    import torch

    hill, _ = _build(HILL, 5)
    big, obig = _build(_large_expression_network(seed=4), 6)
    _, ohill = _build(HILL, 5)
    lma = GillespieNet(np.array([300, 0]), seed=9)
    lma.add_reaction(Rate.lma(0.1, [1, 0]), [-1, 1])
    from oracle.pyoracle import OracleNet
    olma = OracleNet(2)
    olma.add_lma(0.1, [1, 0], [-1, 1])
    specs = [(hill, ohill, HILL[1], 5, 10.0, 20, 96), (big, obig, _large_expression_network(seed=4)[1], 6, 1.0, 10, 64),
             (lma, olma, [300, 0], 9, 5.0, 10, 64)]
    outs = [torch.zeros((nb + 1, g.nb_species(), n), dtype=torch.int64, device="cuda")
            for g, _, _, _, _, nb, n in specs]
    st = run_jobs([dict(net=g, tmax=tm, nb_steps=nb, n_traj=n, out_ptr=o.data_ptr())
                   for (g, _, _, _, tm, nb, n), o in zip(specs, outs)])
    torch.cuda.synchronize()
    tot_ev = 0
    for (g, og, x0, seed, tm, nb, n), o in zip(specs, outs):
        _, oout, _, oev, _ = og.run_grid(seed, x0, tm, nb, n_traj=n, nthreads=nthreads)
        assert np.array_equal(o.cpu().numpy(), oout)
        tot_ev += oev
    assert st.events == tot_ev


def test_expression_networks_refuse_rate_tables():
    import torch

    g, _ = _build(REF2, 1)
    out = torch.zeros((11, 4, 64), dtype=torch.int64, device="cuda")
    rates = torch.ones((2, 4), dtype=torch.float64, device="cuda")
    with pytest.raises(_native.RebopError) as e:
        g.run_sweep_device(2, 32, 1.0, 10, out.data_ptr(), rates_ptr=rates.data_ptr())
    assert e.value.code == K.REBOP_ERR_UNSUPPORTED
    # an x0-only sweep is fine: expressions need no rate table
    x0 = torch.as_tensor(np.array([[0, 500, 0, 0], [0, 800, 0, 0]]), device="cuda")
    g.run_sweep_device(2, 32, 1.0, 10, out.data_ptr(), x0_ptr=x0.data_ptr())
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    assert (o[0, 1, :32] == 500).all() and (o[0, 1, 32:] == 800).all()


def test_infinite_expression_propensity_is_an_error():
    """exp(1000) is +inf: time would not advance; the run reports a range error
    (the kernel treats that trajectory as absorbed, so every lane finishes)."""
    g = GillespieNet(np.array([0]), seed=1)
    code, consts = E.to_postfix(E.parse("exp(1000)"), {"A": 0}, {})
    g.add_reaction(RateExpr.from_postfix(code, consts), [1])
    with pytest.raises(_native.RebopError) as e:
        g.run(1.0, 10, n_traj=64)
    assert e.value.code == K.REBOP_ERR_RANGE
    with pytest.raises(_native.RebopError) as e:
        g.run_all_events(1.0, n_traj=4)
    assert e.value.code == K.REBOP_ERR_RANGE


# ---- tests/test_rebop.py:92-166 through the Python API --------------------


def test_arbitrary_rates():
    s = Gillespie()
    with pytest.raises(ValueError, match="Rate expression not understood"):
        s.add_reaction("+", [], ["A"])
    with pytest.raises(ValueError, match="Rate expression not understood"):
        s.add_reaction("1+", [], ["A"])

    s = Gillespie()
    s.add_reaction("B", [], ["A"])
    with pytest.warns(UserWarning, match="species are not involved in any reaction"):
        ds = s.run({"B": 1}, tmax=10, nb_steps=100)
    assert ds.A[-1] >= 1
    np.testing.assert_equal(np.asarray(ds.B), [1] * 101)

    s = Gillespie()
    s.add_reaction("B", [], ["A"])
    s.add_reaction(1, [], ["B"])
    ds_with_b = s.run({}, tmax=10, nb_steps=100)
    assert ds_with_b.A[-1] >= 1


def test_arbitrary_rates_crossed():
    s = Gillespie()
    s.add_reaction("B", [], ["A"])
    s.add_reaction("A", [], ["B"])
    ds = s.run({}, tmax=10, nb_steps=10)
    np.testing.assert_equal(np.asarray(ds.A), [0] * 11)
    np.testing.assert_equal(np.asarray(ds.B), [0] * 11)
    np.testing.assert_equal(np.asarray(ds.time), np.linspace(0, 10, 11))

    ds = s.run({"A": 1}, tmax=10, nb_steps=10)
    assert ds.A[-1] > 1
    assert ds.B[-1] > 0

    ds = s.run({"B": 1}, tmax=10, nb_steps=10)
    assert ds.A[-1] > 0
    assert ds.B[-1] > 1


def test_arbitrary_rates_2():
    s = Gillespie()
    s.add_reaction(14, [], ["A"])
    s.add_reaction(0.1, ["A", "B"], ["C"], 0.01)
    s.add_reaction("0.2 * B * C / (5 + C)", ["B"], ["D"])
    ds = s.run({"B": 1000}, tmax=100, nb_steps=100)
    assert (np.asarray(ds.B) + np.asarray(ds.C) + np.asarray(ds.D) == 1000).all()
    assert ds.D[-1] >= 1


def test_parameters():
    """tests/test_rebop.py:155-166.  The last run is seeded here: unseeded, as
    there, a Poisson(4) count is 0 with probability e^-4 (about 1.8 %)."""
    s = Gillespie()
    s.add_reaction(4, ["A"], ["B"])
    with pytest.raises(ValueError, match="Species B cannot also be a parameter"):
        s.run({}, 10, 10, params={"B": 4.2})

    s = Gillespie()
    s.add_reaction("k", [], ["A"])
    with pytest.raises(ValueError, match="Parameter k should have a value"):
        s.run({}, 10, 10)
    ds = s.run({}, 10, 10, params={"k": 0.4}, rng=7)
    assert ds.A[-1] > 0


def test_python_expression_run_matches_oracle(nthreads):
    """The Python API's seed derivation + expression lowering, batched: every
    trajectory equals the oracle's."""
    s = Gillespie()
    s.add_reaction("kin * (1 + G)", [], ["P"])
    s.add_reaction(0.3, ["P"], [])
    s.add_reaction("0.5 * G", [], ["G"])
    s.add_reaction("0.1 * G^2 / (10 + G)", ["G"], [])
    ds = s.run({"G": 3}, tmax=8, nb_steps=16, rng=7, params={"kin": 2.0}, n_trajectories=200)
    seed = int(np.random.default_rng(7).integers(np.iinfo(np.uint64).max, dtype=np.uint64))
    from oracle.pyoracle import OracleNet

    names = {"P": 0, "G": 1}
    o = OracleNet(2)
    code, consts = E.to_postfix(E.parse("kin * (1 + G)"), names, {"kin": 2.0})
    o.add_expr(code, consts, [1, 0])
    o.add_lma(0.3, [1, 0], [-1, 0])
    code, consts = E.to_postfix(E.parse("0.5 * G"), names, {})
    o.add_expr(code, consts, [0, 1])
    code, consts = E.to_postfix(E.parse("0.1 * G^2 / (10 + G)"), names, {})
    o.add_expr(code, consts, [0, -1])
    _, oout, _, _, _ = o.run_grid(seed, [0, 3], 8.0, 16, n_traj=200, nthreads=nthreads)
    assert np.array_equal(np.asarray(ds.P), oout[:, 0, :])
    assert np.array_equal(np.asarray(ds.G), oout[:, 1, :])
