"""The tail hand-off of persistent specialised launches: once the trajectory
queue is empty, thin waves write their trajectories' state out and a grouped
launch (2 or 4 lanes per trajectory) finishes them.  Results, event and draw
counts must not change: against REBOP_KERNEL_WHOLE (no hand-off) and the
oracle, for grid and target (batch) runs and several thresholds."""

from __future__ import annotations

import os

import numpy as np
import pytest

from rebop_amd import _native
from tests.networks import build_device, build_oracle, dimers, sir

pytestmark = pytest.mark.gpu
K = _native
PS = K.REBOP_KERNEL_PERSISTENT | K.REBOP_KERNEL_SPECIALIZED


def _run(g, tmax, nb, n, kernel, traj_begin=0):
    import torch

    out = torch.zeros((nb + 1, g.nb_species(), n), dtype=torch.int64, device="cuda")
    st = g.run_device(tmax, nb, n, out.data_ptr(), traj_begin=traj_begin, kernel=kernel)
    torch.cuda.synchronize()
    return out.cpu().numpy(), st


@pytest.fixture
def tail_env():
    saved = {k: os.environ.get(k) for k in ("REBOP_TAIL_LANES", "REBOP_TAIL_GROUP")}
    yield os.environ
    for k, v in saved.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


def test_tail_equals_whole_and_oracle(nthreads):
    spec = dimers()
    g = build_device(spec, 61)
    n, tmax, nb = 20_000, 0.5, 25
    a, sa = _run(g, tmax, nb, n, PS, traj_begin=11)
    b, sb = _run(g, tmax, nb, n, PS | K.REBOP_KERNEL_WHOLE, traj_begin=11)
    assert sa.tail > 0 and sb.tail == 0
    assert np.array_equal(a, b)
    assert (sa.events, sa.draws, sa.trajectories) == (sb.events, sb.draws, sb.trajectories) == (sa.events, sa.draws, n)
    o = build_oracle(spec, sparse=True)
    _, oout, _, oev, odr = o.run_grid(61, spec[2], tmax, nb, n_traj=n, traj_begin=11, nthreads=nthreads)
    assert np.array_equal(a, oout) and sa.events == oev and sa.draws == odr


@pytest.mark.parametrize("lanes,group", [("1", "4"), ("40", "4"), ("63", "2"), ("16", "2")])
def test_tail_thresholds_and_groups(lanes, group, tail_env, nthreads):
    tail_env["REBOP_TAIL_LANES"] = lanes
    tail_env["REBOP_TAIL_GROUP"] = group
    spec = sir()
    g = build_device(spec, 5)
    n = 12_000
    a, sa = _run(g, 250.0, 50, n, PS)
    o = build_oracle(spec)
    _, oout, _, oev, odr = o.run_grid(5, spec[2], 250.0, 50, n_traj=n, nthreads=nthreads)
    assert np.array_equal(a, oout)
    assert sa.events == oev and sa.draws == odr and sa.trajectories == n
    assert sa.tail > 0


def test_tail_off_by_environment(tail_env):
    tail_env["REBOP_TAIL_LANES"] = "0"
    g = build_device(dimers(), 2)
    _, st = _run(g, 0.2, 10, 5000, PS)
    assert st.tail == 0


def test_tail_in_resumable_batches():
    """Target-mode launches (Batch.advance_until) with enough trajectories for
    a persistent launch hand off their tail too; states equal the oracle's."""
    spec = dimers()
    g = build_device(spec, 8)
    o = build_oracle(spec, sparse=True)
    n = 120_000
    b = g.batch(n, traj_begin=4)
    tails = 0
    for t in (0.1, 0.25):
        st = b.advance_until(t)
        assert st.trajectories == n
        tails += st.tail
    assert tails > 0
    x, tt, dr = b.get_species(), b.get_time(), b.get_draws()
    for j in range(0, n, 4999):
        tr = o.state(spec[2], 8, traj=4 + j)
        tr.advance_until(0.1)
        tr.advance_until(0.25)
        assert x[j].tolist() == tr.x.tolist() and tt[j] == tr.t and int(dr[j]) == tr.draws, j
