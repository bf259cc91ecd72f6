"""Pins the random stream shared by the device kernels and the oracle (rebop_rng.h).

Philox4x32-10 is checked against the published Random123 known-answer vectors
and against ATen's independent Philox engine (tests/golden/philox_aten.json);
ln64 against mpmath; the uniform mappings against their closed forms.
"""

from __future__ import annotations

import json
from pathlib import Path

import mpmath as mp
import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"

# Random123 kat_vectors, philox4x32 with 10 rounds: (counter, key) -> output
RANDOM123_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    (
        [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
        [0xA4093822, 0x299F31D0],
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1],
    ),
]


@pytest.mark.parametrize("ctr,key,expected", RANDOM123_KAT)
def test_philox_random123_kat(oracle, ctr, key, expected):
    assert oracle.philox(ctr, key) == expected


def test_philox_matches_aten_engine(oracle):
    """Our (seed, traj, draw) layout == ATen's (seed, subsequence, offset) layout."""
    data = json.loads((GOLDEN / "philox_aten.json").read_text())
    assert len(data["cases"]) > 200
    for c in data["cases"]:
        seed, traj, draw = int(c["seed"]), int(c["traj"]), int(c["draw"])
        ctr = [draw & 0xFFFFFFFF, draw >> 32, traj & 0xFFFFFFFF, traj >> 32]
        key = [seed & 0xFFFFFFFF, seed >> 32]
        assert oracle.philox(ctr, key) == c["words"], c


def test_stream_block_equals_philox(oracle):
    """The kernels' per-trajectory form (round 0's trajectory product taken
    once, keys recomputed per block) is the same function as the plain
    Philox4x32-10 block, on the ATen vectors and on random counters with
    high words set."""
    data = json.loads((GOLDEN / "philox_aten.json").read_text())
    for c in data["cases"][:64]:
        assert oracle.stream_block(int(c["seed"]), int(c["traj"]), int(c["draw"])) == c["words"], c
    rng = np.random.default_rng(11)
    for _ in range(2000):
        seed, traj, draw = (int(v) for v in rng.integers(0, 2**64, size=3, dtype=np.uint64))
        if rng.random() < 0.5:
            draw &= 0xFFFFFFFF
        ctr = [draw & 0xFFFFFFFF, draw >> 32, traj & 0xFFFFFFFF, traj >> 32]
        key = [seed & 0xFFFFFFFF, seed >> 32]
        assert oracle.stream_block(seed, traj, draw) == oracle.philox(ctr, key)


def test_philox_single_bit_avalanche(oracle):
    """A mistake in a round (dropped xor, swapped lanes) shows up as poor diffusion."""
    base = oracle.philox([1, 2, 3, 4], [5, 6])
    flips = []
    for word in range(4):
        for bit in (0, 13, 31):
            ctr = [1, 2, 3, 4]
            ctr[word] ^= 1 << bit
            out = oracle.philox(ctr, [5, 6])
            flips.append(sum(bin(a ^ b).count("1") for a, b in zip(base, out)))
    assert 40 < np.mean(flips) < 88


def _ulp_err(y: float, x: float) -> float:
    ex = mp.log(mp.mpf(x))
    if ex == 0:
        return 0.0 if y == 0 else float("inf")
    ulp = float(np.spacing(np.float64(abs(float(ex)))))
    return float(abs(mp.mpf(y) - ex) / ulp)


def test_ln64_faithful_on_the_draw_domain(oracle):
    """ln64 is used on U1 = j * 2^-53, j in [1, 2^53]: error < 1 ulp everywhere sampled."""
    mp.mp.dps = 40
    rng = np.random.default_rng(3)
    j = rng.integers(1, 2**53, size=3000, dtype=np.uint64)
    xs = list(j.astype(np.float64) * 2.0**-53)
    # both sides of the sqrt(2) reduction boundary, the ends of the domain, powers of two
    s2 = float(np.sqrt(2.0))
    xs += [2.0**-53, 1.0 - 2.0**-53, 1.0, 0.5, s2 / 2, np.nextafter(s2 / 2, 0), np.nextafter(s2 / 2, 1)]
    xs += [2.0**-k for k in range(1, 54)]
    xs += list(rng.uniform(0.5, 2.0, 2000)) + list(np.exp(rng.uniform(-700, 700, 500)))
    worst = max(_ulp_err(oracle.ln64(x), float(x)) for x in xs)
    assert worst < 1.0, worst


def test_ln64_exact_points(oracle):
    assert oracle.ln64(1.0) == 0.0
    # ln(2^k) = k ln 2 to within one rounding
    for k in (-53, -20, -1, 1, 7):
        assert _ulp_err(oracle.ln64(2.0**k), 2.0**k) < 1.0


def test_uniform_mappings(oracle):
    lib = oracle.load()
    # U[0,1) = (hi:lo >> 11) * 2^-53
    assert lib.orc_u01_closed_open(0, 0) == 0.0
    assert lib.orc_u01_closed_open(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0**-53
    assert lib.orc_u01_closed_open(0x800, 0) == 2.0**-53
    assert lib.orc_u01_closed_open(0x7FF, 0) == 0.0
    # U(0,1] = ((hi:lo >> 11) + 1) * 2^-53
    assert lib.orc_u01_open_closed(0, 0) == 2.0**-53
    assert lib.orc_u01_open_closed(0xFFFFFFFF, 0xFFFFFFFF) == 1.0
    rng = np.random.default_rng(5)
    for _ in range(200):
        lo, hi = (int(v) for v in rng.integers(0, 2**32, size=2))
        j = ((hi << 32) | lo) >> 11
        assert lib.orc_u01_closed_open(lo, hi) == j / 2.0**53
        assert lib.orc_u01_open_closed(lo, hi) == (j + 1) / 2.0**53


def test_exponential_stream_moments(oracle):
    """E = -ln U1 over a Philox stream is Exp(1): mean, variance and a KS test."""
    from scipy import stats

    lib = oracle.load()
    es, us = [], []
    for d in range(20000):
        w = oracle.philox([d, 0, 9, 0], [42, 0])
        es.append(-oracle.ln64(lib.orc_u01_open_closed(w[0], w[1])))
        us.append(lib.orc_u01_closed_open(w[2], w[3]))
    es = np.array(es)
    assert abs(es.mean() - 1) < 0.03 and abs(es.var() - 1) < 0.06
    assert stats.kstest(es, "expon").pvalue > 1e-4
    assert stats.kstest(np.array(us), "uniform").pvalue > 1e-4
