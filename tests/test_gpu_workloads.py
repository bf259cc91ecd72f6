"""The large benchmark workloads at their stated sizes and horizons, against
the oracle bit for bit.

- The 128-species x 512-reaction network (seed 7) at tmax = 10 with 200
  samples: 64 trajectories compared whole with the oracle, and the full
  2^18-trajectory launch, whose first and last trajectories must equal the
  oracle's (a trajectory does not depend on the launch it is part of) and
  whose event/draw accounting must close.
- The heterogeneous batch (4096 random networks, seed 2026, tmax per network
  calibrated so that mean events span 1e2 .. 1e5, 256 trajectories each,
  nb_steps = 50) run as the benchmark runs it, in one fused launch: every
  64th network (64 networks across the whole event range) compared with the
  oracle on all of its 256 trajectories, and each of those networks' event
  count.

Reference loop: src/gillespie.rs:315-364 driven by src/pyo3_gillespie.rs:197-208.
"""

from __future__ import annotations

import numpy as np
import pytest

from rebop_amd import _native, workloads
from rebop_amd.engine import run_jobs
from tests.networks import build_device, build_oracle

pytestmark = pytest.mark.gpu

SYN_TMAX, SYN_NB = 10.0, 200
HET_N, HET_TRAJ, HET_NB, HET_SEED = 4096, 256, 50, 2026


def _free():
    import torch

    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def test_synthetic_128x512_full_horizon(nthreads):
    """64 trajectories of the 128 x 512 network over the whole horizon
    (tmax = 10, 200 samples: ~27k events each, every event a full
    checkpointed make_cumrates over 512 reactions)."""
    import torch

    spec = workloads.synthetic_large()
    n = 64
    g = build_device(spec, 7)
    out = torch.zeros((SYN_NB + 1, spec[0], n), dtype=torch.int64, device="cuda")
    st = g.run_device(SYN_TMAX, SYN_NB, n, out.data_ptr(), kernel=_native.REBOP_KERNEL_AUTO)
    torch.cuda.synchronize()
    assert st.kernel == _native.REBOP_KERNEL_LDS | _native.REBOP_KERNEL_COMPACT  # counts < 2^15
    o = build_oracle(spec, sparse=True)
    _, oout, _, oev, odr = o.run_grid(7, spec[2], SYN_TMAX, SYN_NB, n_traj=n, nthreads=nthreads)
    got = out.cpu().numpy()
    assert np.array_equal(got, oout), np.argwhere(got != oout)[:5].tolist()
    assert st.events == oev and st.draws == odr and st.trajectories == n
    assert oev > n * 20000  # the long-horizon regime is exercised


def test_synthetic_128x512_full_size(nthreads):
    """The full 2^18-trajectory launch (54 GB of samples): its first 64 and
    last 16 trajectories equal the oracle's, every trajectory starts from
    x0 at sample 0, counts stay non-negative (every reaction consumes what
    its rate's order reads), and draws = events + discarded overshoots, at
    most one per sample."""
    import torch

    spec = workloads.synthetic_large()
    S, n = spec[0], 2**18
    g = build_device(spec, 7)
    out = torch.empty((SYN_NB + 1, S, n), dtype=torch.int64, device="cuda")
    try:
        st = g.run_device(SYN_TMAX, SYN_NB, n, out.data_ptr(), kernel=_native.REBOP_KERNEL_AUTO)
        torch.cuda.synchronize()
        assert st.kernel == _native.REBOP_KERNEL_LDS | _native.REBOP_KERNEL_COMPACT
        assert st.trajectories == n
        assert st.events <= st.draws <= st.events + n * (SYN_NB + 1)
        x0 = torch.as_tensor(np.asarray(spec[2], dtype=np.int64), device="cuda")
        assert bool((out[0] == x0[:, None]).all())
        assert int(out.min()) >= 0
        head = out[:, :, :64].cpu().numpy()
        tail = out[:, :, n - 16:].cpu().numpy()
    finally:
        del out
        _free()
    o = build_oracle(spec, sparse=True)
    _, ohead, _, _, _ = o.run_grid(7, spec[2], SYN_TMAX, SYN_NB, n_traj=64, nthreads=nthreads)
    _, otail, _, _, _ = o.run_grid(7, spec[2], SYN_TMAX, SYN_NB, n_traj=16, traj_begin=n - 16,
                                   nthreads=nthreads)
    assert np.array_equal(head, ohead)
    assert np.array_equal(tail, otail)


def test_heterogeneous_4096_fused_launch(nthreads):
    """The heterogeneous batch as the benchmark runs it: 4096 networks,
    calibrated horizons, 256 trajectories each, one fused launch.  Every
    64th network is compared with the oracle on all its trajectories."""
    import torch

    specs = workloads.heterogeneous_specs(HET_N, HET_SEED)
    nets = workloads.build_nets(specs, seed=0)
    targets = workloads.target_events(HET_N)
    tmax = workloads.calibrate_tmax(nets, specs, targets)
    R = np.array([len(sp[1]) for sp in specs], dtype=np.float64)
    cost = targets * (R + 16.0)
    seed = 4242
    for g in nets:
        g.seed(seed)
    idx = list(range(0, HET_N, 64))
    outs = {}
    try:
        bufs = [torch.empty((HET_NB + 1, sp[0], HET_TRAJ), dtype=torch.int64, device="cuda")
                for sp in specs]
        jobs = [dict(net=g, tmax=float(t), nb_steps=HET_NB, n_traj=HET_TRAJ, out_ptr=o.data_ptr(),
                     cost_hint=float(c)) for g, t, o, c in zip(nets, tmax, bufs, cost)]
        st, ev = run_jobs(jobs, per_job_events=True)
        torch.cuda.synchronize()
        assert st.kernel & ~_native.REBOP_KERNEL_COMPACT == _native.REBOP_KERNEL_LDS
        assert st.trajectories == HET_N * HET_TRAJ
        assert int(ev.sum()) == st.events
        for i in idx:
            outs[i] = bufs[i].cpu().numpy()
    finally:
        bufs = None
        _free()
    # the sample spans the whole event range of the batch
    per_traj = ev[idx].astype(np.float64) / HET_TRAJ
    assert per_traj.min() < 300 and per_traj.max() > 3e4, (per_traj.min(), per_traj.max())
    for i in idx:
        sp = specs[i]
        o = build_oracle(sp, sparse=True)
        _, oout, _, oev, _ = o.run_grid(seed, sp[2], float(tmax[i]), HET_NB, n_traj=HET_TRAJ,
                                        nthreads=nthreads)
        assert np.array_equal(outs[i], oout), (i, np.argwhere(outs[i] != oout)[:3].tolist())
        assert int(ev[i]) == oev, (i, int(ev[i]), oev)
