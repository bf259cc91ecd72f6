"""Test helpers: the benchmark networks in the oracle."""

from __future__ import annotations

from rebop_amd.networks import build as build_device  # noqa: F401
from rebop_amd.networks import dimers, michaelis_menten, sir, synthetic  # noqa: F401


def build_oracle(spec, sparse=False):
    from oracle.pyoracle import OracleNet

    S, rx, x0 = spec
    n = OracleNet(S, sparse=sparse)
    for k, o, d in rx:
        n.add_lma(k, o, d)
    return n
