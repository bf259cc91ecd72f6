"""Parameter sweeps (per-point rate constants and initial counts) on the
device, bit for bit against per-point oracle runs, and their capture in a
HIP graph replayed several times."""

from __future__ import annotations

import numpy as np
import pytest

from rebop_amd import _native
from tests.networks import build_device, build_oracle, michaelis_menten

pytestmark = pytest.mark.gpu


def _points(P, seed=0):
    rng = np.random.default_rng(seed)
    rates = 10.0 ** rng.uniform(-3, 1, size=(P, 3))
    x0 = np.zeros((P, 4), dtype=np.int64)
    x0[:, 0] = rng.integers(10, 201, P)
    x0[:, 1] = rng.integers(50, 1001, P)
    return rates, x0


def _oracle_point(spec, rates_p, x0_p, seed, tmax, nb, tpp, traj_begin, nthreads):
    from oracle.pyoracle import OracleNet

    S, rx, _ = spec
    o = OracleNet(S)
    for k, (_, orders, d) in zip(rates_p, rx):
        o.add_lma(float(k), orders, d)
    _, out, _, ev, _ = o.run_grid(seed, x0_p, tmax, nb, n_traj=tpp, traj_begin=traj_begin, nthreads=nthreads)
    return out, ev


@pytest.mark.parametrize("kernel", [
    _native.REBOP_KERNEL_PERSISTENT,
    _native.REBOP_KERNEL_PLAIN | _native.REBOP_KERNEL_SPECIALIZED,
    _native.REBOP_KERNEL_PERSISTENT | _native.REBOP_KERNEL_SPECIALIZED,
])
def test_sweep_bit_exact(kernel, nthreads):
    import torch

    spec = michaelis_menten()
    P, tpp, nb, tmax = 24, 40, 32, 50.0
    rates, x0 = _points(P)
    g = build_device(spec, 555)
    d_rates = torch.as_tensor(rates, device="cuda")
    d_x0 = torch.as_tensor(x0, device="cuda")
    out = torch.zeros((nb + 1, 4, P * tpp), dtype=torch.int64, device="cuda")
    st = g.run_sweep_device(P, tpp, tmax, nb, out.data_ptr(), rates_ptr=d_rates.data_ptr(),
                            x0_ptr=d_x0.data_ptr(), traj_begin=1000, kernel=kernel)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ev_total = 0
    for p in range(P):
        oout, ev = _oracle_point(spec, rates[p], x0[p], 555, tmax, nb, tpp, 1000 + p * tpp, nthreads)
        assert np.array_equal(got[:, :, p * tpp:(p + 1) * tpp], oout), p
        ev_total += ev
    assert st.events == ev_total


def test_sweep_graph_capture_replays():
    """Capture the sweep in a HIP graph; every replay (with eager work and
    synchronisations between replays) equals the eager result; rewriting the
    rate table in place changes the next replay exactly as an eager run."""
    import torch

    spec = michaelis_menten()
    P, tpp, nb, tmax = 64, 64, 16, 20.0
    rates, x0 = _points(P, seed=3)
    g = build_device(spec, 9)
    g.prepare()
    d_rates = torch.as_tensor(rates, device="cuda")
    d_x0 = torch.as_tensor(x0, device="cuda")
    out = torch.zeros((nb + 1, 4, P * tpp), dtype=torch.int64, device="cuda")
    kern = _native.REBOP_KERNEL_PERSISTENT | _native.REBOP_KERNEL_SPECIALIZED

    def eager():
        o = torch.zeros_like(out)
        g.run_sweep_device(P, tpp, tmax, nb, o.data_ptr(), rates_ptr=d_rates.data_ptr(),
                           x0_ptr=d_x0.data_ptr(), kernel=kern,
                           stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        return o

    ref = eager()
    s = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(graph, stream=s):
        g.run_sweep_device(P, tpp, tmax, nb, out.data_ptr(), rates_ptr=d_rates.data_ptr(),
                           x0_ptr=d_x0.data_ptr(), kernel=kern, stream=s.cuda_stream,
                           with_stats=False)
    for rep in range(4):
        out.fill_(-1)
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), rep
        del junk
    # a new parameter table, written in place
    d_rates.mul_(2.0)
    ref2 = eager()
    assert not torch.equal(ref2, ref)
    out.fill_(-1)
    graph.replay()
    torch.cuda.synchronize()
    assert torch.equal(out, ref2)
