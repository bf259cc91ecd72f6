"""Bit-exact parity of the device SSA against the CPU oracle (same Philox stream)."""

from __future__ import annotations

import numpy as np
import pytest

from tests.networks import build_device, build_oracle, dimers, michaelis_menten, sir

pytestmark = pytest.mark.gpu


def _compare(spec, seed, tmax, nb, n, nthreads, traj_begin=0):
    g = build_device(spec, seed)
    o = build_oracle(spec)
    times, out, st = g.run(tmax, nb, n_traj=n, traj_begin=traj_begin, with_stats=True)
    otimes, oout, ofinal, oev, odr = o.run_grid(seed, spec[2], tmax, nb, n_traj=n, traj_begin=traj_begin, nthreads=nthreads)
    assert np.array_equal(times.view(np.uint64), otimes.view(np.uint64))
    mism = np.argwhere(out != oout)
    assert mism.size == 0, f"first mismatches (step, species, traj): {mism[:5].tolist()}"
    assert st.events == oev
    assert st.draws == odr
    assert st.trajectories == n
    return out


def test_sir_bit_exact_1e4(nthreads):
    """SIR, 10^4 trajectories, seed 42, tmax=250, nb_steps=250: every sample bit-exact."""
    out = _compare(sir(), 42, 250.0, 250, 10_000, nthreads)
    assert (out.sum(axis=1) == 1000).all()


def test_dimers_bit_exact(nthreads):
    """Dimer network, tmax=1, nb_steps=100, 4096 trajectories: every sample bit-exact."""
    out = _compare(dimers(), 42, 1.0, 100, 4096, nthreads)
    assert (out[:, 0, :] == 1).all()


def test_mm_bit_exact_offset(nthreads):
    """Michaelis-Menten with a trajectory-id offset (sharding by id range)."""
    _compare(michaelis_menten(), 7, 50.0, 32, 3000, nthreads, traj_begin=123456789)
