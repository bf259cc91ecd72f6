"""define_system networks on the device (src/gillespie_macro.rs semantics:
integer mass-action products, first-match choice): bit for bit against the
oracle's restatement over every specialised kernel variant, all-events runs,
resumable single trajectories, sweeps, the LDS kernel (networks beyond the
specialised kernels' size, fused batches); the macro's own unit tests and
the reference's Rust examples through the Python class."""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle.pyoracle import OracleNet
from rebop_amd import _native, define_system
from rebop_amd.engine import run_jobs

pytestmark = pytest.mark.gpu
K = _native
SPEC = K.REBOP_KERNEL_SPECIALIZED

SIR = define_system("r1 r2; SIR { S, I, R } r_infection: S + I => I + I @ r1 r_remission: I => R @ r2")
DIMERS = define_system("""
    rtx rtl rdi rdm rdp;
    Dimers { gene, mRNA, protein, dimer }
    r_tx : gene      => gene + mRNA      @ rtx
    r_tl : mRNA      => mRNA + protein   @ rtl
    r_di : 2 protein => dimer            @ rdi
    r_dm : mRNA      =>                  @ rdm
    r_dp : protein   =>                  @ rdp
""")
# A + A (A^2), a 0 coefficient (factor A, A unchanged), third order, a source
QUIRKS = define_system("""
    k1 k2 k3 k4;
    Quirks { A, B, C }
    r1 : A + A     => B          @ k1
    r2 : 0 A + B   => C          @ k2 / 2.
    r3 : 3 C       => A + 2 B    @ k3
    r4 :           => A          @ k4
""")


def _sir():
    s = SIR.with_parameters(1e-4, 0.01)
    s.S, s.I = 999, 1
    return s


def _dimers():
    d = DIMERS.with_parameters(25.0, 1000.0, 0.001, 0.1, 1.0)
    d.gene = 1
    return d


def _quirks():
    q = QUIRKS.with_parameters(0.002, 0.05, 0.0004, 3.0)
    q.A, q.B, q.C = 30, 10, 5
    return q


CASES = {"sir": (_sir, 250.0, 50, 512), "dimers": (_dimers, 1.0, 20, 512), "quirks": (_quirks, 20.0, 25, 512)}


def _oracle_of(obj, rates=None):
    spec = obj._spec
    idx = {s: i for i, s in enumerate(spec.species)}
    o = OracleNet.system(len(spec.species))
    for r, k in zip(spec.reactions, obj.rates() if rates is None else rates):
        o.add_system(float(k), [(idx[s], m) for s, m in r.reactants], [(idx[s], m) for s, m in r.products])
    return o


def _x0(obj):
    return [int(getattr(obj, s)) for s in obj._spec.species]


@pytest.mark.parametrize("kernel", [
    K.REBOP_KERNEL_AUTO,
    K.REBOP_KERNEL_PLAIN | SPEC,
    K.REBOP_KERNEL_PERSISTENT | SPEC,
    K.REBOP_KERNEL_PLAIN | SPEC | K.REBOP_KERNEL_GROUP2,
    K.REBOP_KERNEL_PLAIN | SPEC | K.REBOP_KERNEL_GROUP4,
    K.REBOP_KERNEL_PERSISTENT,  # data-driven request: the network forces the specialised kernel
])
@pytest.mark.parametrize("name", list(CASES))
def test_system_grid_bit_exact(name, kernel, nthreads):
    import torch

    make, tmax, nb, n = CASES[name]
    obj = make()
    obj.seed(77)
    g = obj.network()
    out = torch.zeros((nb + 1, len(obj._spec.species), n), dtype=torch.int64, device="cuda")
    st = g.run_device(tmax, nb, n, out.data_ptr(), traj_begin=3, kernel=kernel)
    torch.cuda.synchronize()
    assert st.kernel & SPEC
    o = _oracle_of(obj)
    _, oout, _, oev, odr = o.run_grid(77, _x0(obj), tmax, nb, n_traj=n, traj_begin=3, nthreads=nthreads)
    got = out.cpu().numpy()
    assert np.array_equal(got, oout), np.argwhere(got != oout)[:5].tolist()
    assert st.events == oev and st.draws == odr and oev > 20 * n


def test_system_run_host_and_per_trajectory_x0(nthreads):
    q = _quirks()
    q.seed(5)
    times, out = q.run(5.0, 10, n_trajectories=300, traj_begin=7)
    assert np.array_equal(times, np.linspace(0, 5.0, 11))
    _, oout, _, _, _ = _oracle_of(q).run_grid(5, _x0(q), 5.0, 10, n_traj=300, traj_begin=7, nthreads=nthreads)
    assert np.array_equal(out, oout)
    rng = np.random.default_rng(2)
    x0 = rng.integers(0, 40, size=(200, 3))
    _, out2 = q.network().run(5.0, 10, n_traj=200, x0=x0)
    _, oout2, _, _, _ = _oracle_of(q).run_grid(5, x0, 5.0, 10, n_traj=200, x0_per_traj=True, nthreads=nthreads)
    assert np.array_equal(out2, oout2)


def test_system_all_events_bit_exact():
    d = _dimers()
    d.seed(4)
    ts, xs = d.network().run_all_events(0.2, n_traj=12, traj_begin=1)
    o = _oracle_of(d)
    for j in range(12):
        ot, ox = o.run_all_events(4, 1 + j, _x0(d), 0.2)
        assert np.array_equal(ts[j].view(np.uint64), ot.view(np.uint64)), j
        assert np.array_equal(xs[j], ox), j


def test_advance_until_resumes_across_parameter_changes():
    """One trajectory advanced in steps, with a parameter changed between two
    calls: the stream continues where it stopped (oracle: the same state
    advanced over the two networks)."""
    s = _sir()
    s.seed(31)
    o1 = _oracle_of(s)
    tr = o1.state(_x0(s), 31)
    for t in (10.0, 25.0, 60.0):
        s.advance_until(t)
        tr.advance_until(t)
        assert _x0(s) == tr.x.tolist() and s.t == tr.t and s._draws == tr.draws
    s.r2 = 0.05  # a new network (new kernel); counts, t and stream carry over
    tr.net = _oracle_of(s)
    tr.cum = np.zeros(2)
    for t in (100.0, 250.0):
        s.advance_until(t)
        tr.advance_until(t)
        assert _x0(s) == tr.x.tolist() and s.t == tr.t and s._draws == tr.draws
    # counts set by hand between calls are used, as in the macro's struct
    s.S += 500
    tr.x[0] += 500
    s.advance_until(400.0)
    tr.advance_until(400.0)
    assert _x0(s) == tr.x.tolist()


def test_system_sweep_rates_and_x0(nthreads):
    import torch

    d = _dimers()
    d.seed(9)
    g = d.network()
    P, tpp, nb, tmax = 12, 48, 10, 0.5
    rng = np.random.default_rng(1)
    rates = np.array(d.rates()) * rng.uniform(0.5, 2.0, size=(P, 5))
    x0 = np.zeros((P, 4), dtype=np.int64)
    x0[:, 0] = rng.integers(1, 4, P)
    out = torch.zeros((nb + 1, 4, P * tpp), dtype=torch.int64, device="cuda")
    d_rates = torch.as_tensor(rates, device="cuda")
    d_x0 = torch.as_tensor(x0, device="cuda")
    st = g.run_sweep_device(P, tpp, tmax, nb, out.data_ptr(), rates_ptr=d_rates.data_ptr(),
                            x0_ptr=d_x0.data_ptr(), traj_begin=100)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ev = 0
    for p in range(P):
        o = _oracle_of(d, rates[p])
        _, oout, _, oev, _ = o.run_grid(9, x0[p], tmax, nb, n_traj=tpp, traj_begin=100 + p * tpp, nthreads=nthreads)
        assert np.array_equal(got[:, :, p * tpp:(p + 1) * tpp], oout), p
        ev += oev
    assert st.events == ev


# ---- gillespie_macro.rs:173-260 through the Python class ----------------------


def test_macro_unit_tests():
    s = SIR.new()
    s.r1, s.r2 = 0.1 / 10000.0, 0.01
    s.S, s.I = 9999, 1
    s.advance_until(1000.0)
    assert s.S + s.I + s.R == 10000

    d = DIMERS.with_parameters(25.0, 1000.0, 0.001, 0.1, 1.0)
    d.gene = 1
    d.advance_until(1.0)
    assert d.gene == 1 and 1000 < d.dimer < 10000

    BD = define_system("r_birth r_death; BirthDeath { A } birth: => A @ r_birth death: A => @ r_death")
    b = BD.new()
    b.r_birth, b.r_death = 10.0, 0.1
    b.advance_until(100.0)
    assert 50 < b.A < 200

    f = BD.new()
    f.r_birth = 10.0  # r_death stays NaN: no reaction happens
    f.advance_until(100.0)
    assert abs(f.t - 100.0) < np.finfo(float).eps and f.A == 0

    F = define_system(";\n FooBarBuz { Foo, Bar, Buz }")
    z = F.new()
    z.Foo, z.Bar = 42, 1337
    z.advance_until(1e20)
    assert abs(z.t - 1e20) < np.finfo(float).eps and (z.Foo, z.Bar, z.Buz) == (42, 1337, 0)


def test_reference_rust_examples():
    """examples/sir.rs (10^5 systems advanced to t = 250, mean R) and
    examples/dimers.rs (10^4 systems to t = 1, mean dimer) as resumable
    batches; a sample of trajectories equals the oracle bit for bit."""
    s = _sir()
    s.seed(2024)
    b = s.batch(100_000)
    b.advance_until(250.0)
    x = b.get_species()
    assert (x.sum(axis=1) == 1000).all()
    o = _oracle_of(s)
    for j in range(0, 100_000, 997):
        tr = o.state(_x0(s), 2024, traj=j)
        tr.advance_until(250.0)
        assert x[j].tolist() == tr.x.tolist(), j
    # an outbreak infects most of the population; some die out at once
    assert 500 < x[:, 2].mean() < 1000

    d = _dimers()
    d.seed(7)
    b = d.batch(10_000)
    b.advance_until(1.0)
    x = b.get_species()
    o = _oracle_of(d)
    for j in range(0, 10_000, 499):
        tr = o.state(_x0(d), 7, traj=j)
        tr.advance_until(1.0)
        assert x[j].tolist() == tr.x.tolist(), j
    assert (x[:, 0] == 1).all() and 1000 < x[:, 3].mean() < 10000


def test_same_law_as_run_time_semantics_at_scale():
    """10^5 trajectories each: the dimer network with define_system
    arithmetic and with the run-time arithmetic have the same law at t = 1
    (two-sample Kolmogorov-Smirnov on every species)."""
    from scipy.stats import ks_2samp

    from tests.networks import build_device, dimers

    d = _dimers()
    d.seed(1)
    _, a = d.run(1.0, 1, n_trajectories=100_000)
    g = build_device(dimers(), 2)
    _, b = g.run(1.0, 1, n_traj=100_000)
    for s in range(1, 4):
        assert ks_2samp(a[-1, s], b[-1, s]).pvalue > 1e-3, s


@pytest.mark.parametrize("name", list(CASES))
def test_system_on_the_lds_kernel_bit_exact(name, nthreads):
    """The LDS kernel's define_system arithmetic (term products as wrapping
    integers, first-match choice, `0 A`) equals the oracle bit for bit."""
    import torch

    make, tmax, nb, n = CASES[name]
    obj = make()
    obj.seed(31)
    g = obj.network()
    out = torch.zeros((nb + 1, len(obj._spec.species), n), dtype=torch.int64, device="cuda")
    st = g.run_device(tmax, nb, n, out.data_ptr(), traj_begin=2, kernel=K.REBOP_KERNEL_LDS)
    torch.cuda.synchronize()
    assert st.kernel & K.REBOP_KERNEL_LDS
    _, oout, _, oev, odr = _oracle_of(obj).run_grid(31, _x0(obj), tmax, nb, n_traj=n, traj_begin=2,
                                                    nthreads=nthreads)
    got = out.cpu().numpy()
    assert np.array_equal(got, oout), np.argwhere(got != oout)[:5].tolist()
    assert st.events == oev and st.draws == odr


def _large_system(S=36, R=72, seed=3):
    """A define_system-form network beyond the specialised kernels' 32 / 64:
    conversions, dimerisations (A + A), third-order terms, `0 A` catalysts
    and sources, as (GillespieNet, OracleNet, x0)."""
    rng = np.random.default_rng(seed)
    x0 = [int(v) for v in rng.integers(5, 40, S)]
    g = GillespieNet_system(x0, 13)
    o = OracleNet.system(S)
    for r in range(R):
        a, b = int(rng.integers(0, S)), int(rng.integers(0, S))
        kind = r % 6
        if kind == 0:
            reac, prod, k = [(a, 1)], [(b, 1)], 0.3
        elif kind == 1:
            reac, prod, k = [(a, 1), (a, 1)], [(b, 1)], 0.002
        elif kind == 2:
            reac, prod, k = [(a, 2)], [(b, 1)], 0.004
        elif kind == 3:
            reac, prod, k = [(a, 0), (b, 1)], [(a, 1)], 0.05
        elif kind == 4:
            reac, prod, k = [(a, 3)], [(b, 2)], 1e-4
        else:
            reac, prod, k = [], [(a, 1)], 2.0
        g.add_system_reaction(k, reac, prod)
        o.add_system(k, reac, prod)
    return g, o, x0


def GillespieNet_system(x0, seed):
    from rebop_amd.engine import GillespieNet

    return GillespieNet.system(x0, seed=seed)


def test_large_system_network_runs_on_the_lds_kernel(nthreads):
    import torch

    g, o, x0 = _large_system()
    n, tmax, nb = 64, 2.0, 20
    out = torch.zeros((nb + 1, len(x0), n), dtype=torch.int64, device="cuda")
    st = g.run_device(tmax, nb, n, out.data_ptr(), traj_begin=9)
    torch.cuda.synchronize()
    assert st.kernel & K.REBOP_KERNEL_LDS
    _, oout, _, oev, odr = o.run_grid(13, x0, tmax, nb, n_traj=n, traj_begin=9, nthreads=nthreads)
    got = out.cpu().numpy()
    assert np.array_equal(got, oout), np.argwhere(got != oout)[:5].tolist()
    assert st.events == oev and st.draws == odr and oev > 50 * n
    with pytest.raises(_native.RebopError) as e:
        g.run_device(tmax, nb, n, out.data_ptr(), kernel=K.REBOP_KERNEL_PLAIN | SPEC)
    assert e.value.code == K.REBOP_ERR_UNSUPPORTED


def test_system_networks_in_a_fused_batch(nthreads):
    """define_system networks in one fused launch beside a run-time network:
    each job equals its own oracle run."""
    import torch
    from rebop_amd.engine import GillespieNet, Rate

    d = _dimers()
    d.seed(8)
    big, obig, bx0 = _large_system(seed=5)
    lma = GillespieNet(np.array([300, 0]), seed=9)
    lma.add_reaction(Rate.lma(0.1, [1, 0]), [-1, 1])
    olma = OracleNet(2)
    olma.add_lma(0.1, [1, 0], [-1, 1])
    specs = [(d.network(), _oracle_of(d), _x0(d), 8, 0.5, 10, 96), (big, obig, bx0, 13, 1.0, 10, 64),
             (lma, olma, [300, 0], 9, 5.0, 10, 64)]
    outs = [torch.zeros((nb + 1, len(x0), n), dtype=torch.int64, device="cuda")
            for _, _, x0, _, _, nb, n in specs]
    st = run_jobs([dict(net=g, tmax=tm, nb_steps=nb, n_traj=n, out_ptr=o.data_ptr())
                   for (g, _, _, _, tm, nb, n), o in zip(specs, outs)])
    torch.cuda.synchronize()
    tot = 0
    for (_, og, x0, seed, tm, nb, n), o in zip(specs, outs):
        _, oout, _, oev, _ = og.run_grid(seed, x0, tm, nb, n_traj=n, nthreads=nthreads)
        assert np.array_equal(o.cpu().numpy(), oout)
        tot += oev
    assert st.events == tot


def test_system_nan_rate_absorbs():
    # NaN rate: every trajectory absorbs at once, no draw
    f = DIMERS.new()
    f.gene = 1
    times, x = f.run(1.0, 4, n_trajectories=64)
    assert (x[:, 0, :] == 1).all() and (x[:, 1:, :] == 0).all()
    assert math.isnan(f.rtx)
