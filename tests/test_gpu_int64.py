"""Counts over the reference's whole isize range (src/gillespie.rs:157-163
keeps them as `Vec<isize>`): the register and specialised kernels hold
counts as f64 (exact for |x| <= 2^52); runs that start or end up beyond that
run on the LDS kernel's int64 counts, bit-exact against the oracle (int64
counts, `x as f64` conversions as at src/gillespie.rs:77)."""

from __future__ import annotations

import numpy as np
import pytest

from rebop_amd import _native
from tests.networks import build_device, build_oracle

pytestmark = pytest.mark.gpu
LDS_WIDE = _native.REBOP_KERNEL_LDS | _native.REBOP_KERNEL_WIDE


def _big(a0, b0=3, inflow=0.5, d0=-(2**61) - 5):
    """A, B, C, D with A near a0: A -> A + B and 2A -> A + C with rates scaled
    so that their propensities are O(1) at A ~ 2^60 (products of two counts
    beyond 2^53), B and C decaying back, an inflow of A in jumps of 2^20, and
    a passive negative count D that A -> A + B decrements (negative counts
    take the LDS kernel's linear search)."""
    S = 4
    rx = [(2.0**-60, [1, 0, 0, 0], [0, 1, 0, -1]),
          (2.0**-121, [2, 0, 0, 0], [-1, 0, 1, 0]),
          (0.3, [0, 1, 0, 0], [0, -1, 0, 0]),
          (0.2, [0, 0, 1, 0], [1, 0, -1, 0]),
          (inflow, [0, 0, 0, 0], [2**20, 0, 0, 0])]
    return S, rx, [a0, b0, 0, d0]


def _dev(g, tmax, nb, n, kernel=_native.REBOP_KERNEL_AUTO, x0=None):
    import torch

    S = g.nb_species()
    out = torch.zeros((nb + 1, S, n), dtype=torch.int64, device="cuda")
    fin = torch.zeros((S, n), dtype=torch.int64, device="cuda")  # [S][n_traj]
    xp = None
    if x0 is not None:
        xt = torch.as_tensor(np.ascontiguousarray(x0, dtype=np.int64), device="cuda")
        xp = xt.data_ptr()
    st = g.run_device(tmax, nb, n, out.data_ptr(), kernel=kernel, x0_ptr=xp, final_ptr=fin.data_ptr())
    torch.cuda.synchronize()
    return out.cpu().numpy(), fin.cpu().numpy().T, st


@pytest.mark.parametrize("a0", [2**60 + 7, 2**53 + 1, 2**62 + 2**40 + 1])
def test_shared_counts_beyond_2p52_run_on_int64(a0, nthreads):
    """Shared initial counts beyond 2^52 go to the LDS kernel's int64 counts
    up front (also for a 4-species network the specialised kernel would
    take); samples and final states equal the oracle's."""
    spec = _big(a0)
    g = build_device(spec, 21)
    out, fin, st = _dev(g, 4.0, 16, 300)
    assert st.kernel == LDS_WIDE
    o = build_oracle(spec)
    _, oout, ofin, oev, _ = o.run_grid(21, spec[2], 4.0, 16, n_traj=300, nthreads=nthreads)
    assert np.array_equal(out, oout)
    assert np.array_equal(fin, ofin)
    assert st.events == oev
    # the low bits that f64 counts would have lost are there
    assert np.any(oout[:, 0, :] % 2 != 0) and np.any(oout[:, 3, :] % 2 != 0)


def test_host_run_and_events_beyond_2p52(nthreads):
    """The host-buffer grid loop and the every-event loop, same network."""
    spec = _big(2**60 + 7)
    g = build_device(spec, 5)
    times, out = g.run(3.0, 6, n_traj=40)
    o = build_oracle(spec)
    otimes, oout, _, _, _ = o.run_grid(5, spec[2], 3.0, 6, n_traj=40, nthreads=nthreads)
    assert np.array_equal(times, otimes)
    assert np.array_equal(out, oout)
    ts, xs = g.run_all_events(3.0, n_traj=8)
    for j in range(8):
        ot, ox = o.run_all_events(5, j, spec[2], 3.0)
        assert np.array_equal(ts[j], ot)
        assert np.array_equal(xs[j], ox)


@pytest.mark.parametrize("kernel", [_native.REBOP_KERNEL_AUTO, _native.REBOP_KERNEL_PLAIN,
                                    _native.REBOP_KERNEL_PERSISTENT | _native.REBOP_KERNEL_SPECIALIZED])
def test_per_trajectory_counts_beyond_2p52_are_rerun(kernel, nthreads):
    """Device-side initial counts are not visible to the host: the register
    or specialised kernel flags them at the first sample and the run is
    rerun on the int64 LDS kernel, every output rewritten."""
    spec = _big(5, d0=0)
    n = 256
    rng = np.random.default_rng(4)
    x0 = np.zeros((n, 4), dtype=np.int64)
    x0[:, 0] = 2**60 + rng.integers(0, 2**20, size=n)
    x0[:, 1] = rng.integers(0, 10, size=n)
    x0[: n // 2, 0] = rng.integers(0, 100, size=n // 2)  # half of them small
    g = build_device(spec, 8)
    out, fin, st = _dev(g, 3.0, 10, n, kernel=kernel, x0=x0)
    assert st.kernel == LDS_WIDE
    o = build_oracle(spec)
    _, oout, ofin, oev, _ = o.run_grid(8, x0, 3.0, 10, n_traj=n, x0_per_traj=True, nthreads=nthreads)
    assert np.array_equal(out, oout)
    assert np.array_equal(fin, ofin)
    assert st.events == oev


@pytest.mark.parametrize("kernel", [_native.REBOP_KERNEL_AUTO, _native.REBOP_KERNEL_PLAIN])
def test_counts_growing_past_2p52_are_rerun(kernel, nthreads):
    """A starts just below 2^52 and its inflow (jumps of 2^20) carries it
    past: the f64 kernels flag the first sample beyond 2^52 and the run is
    rerun on int64 counts.  A run that stays below keeps its kernel."""
    spec = _big(2**52 - 2**22, inflow=3.0, d0=0)
    g = build_device(spec, 13)
    out, fin, st = _dev(g, 4.0, 8, 200, kernel=kernel)
    assert st.kernel == LDS_WIDE
    o = build_oracle(spec)
    _, oout, ofin, oev, _ = o.run_grid(13, spec[2], 4.0, 8, n_traj=200, nthreads=nthreads)
    assert np.array_equal(out, oout) and np.array_equal(fin, ofin) and st.events == oev
    assert oout[-1, 0, :].max() > 2**52
    small = _big(2**40, inflow=3.0, d0=0)
    g2 = build_device(small, 13)
    out2, _, st2 = _dev(g2, 4.0, 8, 200, kernel=kernel)
    assert (st2.kernel & 0xF) != _native.REBOP_KERNEL_LDS
    _, oout2, _, _, _ = build_oracle(small).run_grid(13, small[2], 4.0, 8, n_traj=200, nthreads=nthreads)
    assert np.array_equal(out2, oout2)


def test_events_growing_past_2p52_are_rerun():
    """The every-event loop: the specialised kernel flags the first event
    beyond 2^52 and both passes (counting, recording) are rerun on int64
    counts; the records equal the oracle's."""
    spec = _big(2**52 - 2**22, inflow=3.0, d0=0)
    g = build_device(spec, 17)
    ts, xs = g.run_all_events(2.0, n_traj=6)
    o = build_oracle(spec)
    grew = False
    for j in range(6):
        ot, ox = o.run_all_events(17, j, spec[2], 2.0)
        assert np.array_equal(ts[j], ot)
        assert np.array_equal(xs[j], ox)
        grew = grew or int(ox[:, 0].max()) > 2**52
    assert grew


def test_fused_jobs_beyond_2p52(nthreads):
    """A fused batch holding a job whose counts are beyond 2^52 runs with
    int64 counts; every job equals the oracle."""
    import torch

    from rebop_amd.engine import run_jobs

    specs = [_big(2**60 + 7), _big(5, d0=0)]
    nets = [build_device(sp, 30 + i) for i, sp in enumerate(specs)]
    bufs = [torch.zeros((9, sp[0], 96), dtype=torch.int64, device="cuda") for sp in specs]
    st, ev = run_jobs([dict(net=g, tmax=2.0, nb_steps=8, n_traj=96, out_ptr=b.data_ptr())
                       for g, b in zip(nets, bufs)], per_job_events=True)
    torch.cuda.synchronize()
    assert st.kernel == LDS_WIDE
    for i, sp in enumerate(specs):
        _, oout, _, oev, _ = build_oracle(sp).run_grid(30 + i, sp[2], 2.0, 8, n_traj=96, nthreads=nthreads)
        assert np.array_equal(bufs[i].cpu().numpy(), oout) and int(ev[i]) == oev


def test_python_api_counts_beyond_2p52():
    """The reference's Python surface (Gillespie.run) with an initial count
    f64 cannot hold: A -> B conserves A + B = 2^60 + 7 exactly."""
    from rebop_amd import Gillespie

    s = Gillespie()
    s.add_reaction(2.0**-60, ["A"], ["B"])
    ds = s.run({"A": 2**60 + 7}, tmax=5.0, nb_steps=10, rng=3)
    A, B = np.asarray(ds.A, dtype=np.int64), np.asarray(ds.B, dtype=np.int64)
    assert ((A + B) == 2**60 + 7).all()
    assert B[-1] > 0 and A[0] == 2**60 + 7


def test_int64_overflow_and_batch_limits():
    """An int64 overflow is an error (the reference's isize addition would
    wrap in a release build, panic in a debug one); resumable batches keep
    their state in f64 arrays and stay within |x| <= 2^52."""
    spec = _big(2**63 - 2**21, inflow=5.0, d0=0)
    g = build_device(spec, 3)
    with pytest.raises(_native.RebopValueError):
        _dev(g, 2.0, 4, 64)
    with pytest.raises(_native.RebopValueError):
        build_device(_big(2**53, d0=0), 3).batch(8)
    b = build_device(_big(2**52 - 2**21, inflow=5.0, d0=0), 3).batch(16)
    with pytest.raises(_native.RebopValueError):
        b.advance_until(4.0)
