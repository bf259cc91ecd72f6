"""Pins the FP64 CPU reference of the surrogate (oracle/surrogate_ref.py)
against independent implementations: torch's bfloat16 conversion and
autograd (CPU, float64), scipy's negative-binomial log-pmf, and finite
differences."""

from __future__ import annotations

import numpy as np
import pytest
import torch
from scipy import stats

from oracle import surrogate_ref as ref

K = 8


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(5000) * 10.0 ** rng.integers(-6, 6, 5000),
                        [1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-9, -2.5, 0.0, 65504.0, 1e-30]])
    mine = ref.bf16(x)
    theirs = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mine, theirs)
    assert ref.bf16(np.array([1.0 + 2**-8]))[0] == 1.0  # tie to even


def _random_head(B, S, rng):
    head = np.zeros((B, 128))
    for s in range(S):
        head[:, s * 3 * K:s * 3 * K + K] = rng.normal(size=(B, K))
        head[:, s * 3 * K + K:s * 3 * K + 2 * K] = rng.normal(np.log(40.0), 1.0, size=(B, K))
        head[:, s * 3 * K + 2 * K:(s + 1) * 3 * K] = rng.normal(1.5, 1.0, size=(B, K))
    return head


def test_nll_value_matches_scipy_nbinom():
    rng = np.random.default_rng(1)
    B, S, J = 3, 4, 16
    head = _random_head(B, S, rng)
    samples = rng.poisson(40, size=(B, S, J))
    total, _ = ref.nll(head, samples, K)
    expect = 0.0
    for b in range(B):
        for s in range(S):
            h = head[b, s * 3 * K:(s + 1) * 3 * K]
            w = np.exp(h[:K] - np.logaddexp.reduce(h[:K]))
            mu, r = np.exp(h[K:2 * K]), np.exp(h[2 * K:])
            for x in samples[b, s]:
                expect -= np.log(sum(w[c] * stats.nbinom.pmf(x, r[c], r[c] / (r[c] + mu[c])) for c in range(K)))
    assert total == pytest.approx(expect, rel=1e-10)


def test_nll_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    B, S, J = 2, 4, 8
    head = _random_head(B, S, rng)
    samples = rng.poisson(30, size=(B, S, J))
    _, g = ref.nll(head, samples, K)
    for idx in [(0, 0), (0, 3), (1, K + 2), (0, 2 * K + 5), (1, 3 * K + 1), (1, 3 * 3 * K + 2 * K + 7)]:
        h1, h2 = head.copy(), head.copy()
        eps = 1e-6
        h1[idx] += eps
        h2[idx] -= eps
        fd = (ref.nll(h1, samples, K)[0] - ref.nll(h2, samples, K)[0]) / (2 * eps)
        assert g[idx] == pytest.approx(fd, rel=1e-5, abs=1e-7), idx


def test_mlp_backward_matches_torch_autograd():
    rng = np.random.default_rng(3)
    dims = [64, 32, 32, 128]
    Ws = [rng.normal(0, 0.3, size=(dims[i], dims[i + 1])) for i in range(3)]
    bs = [rng.normal(0, 0.1, size=dims[i + 1]) for i in range(3)]
    X = rng.normal(size=(5, 64))
    G = rng.normal(size=(5, 128))
    head, H = ref.forward(X, Ws, bs, emulate_bf16=False)
    dW, db = ref.backward(G, H, Ws, emulate_bf16=False)
    tW = [torch.tensor(w, requires_grad=True) for w in Ws]
    tb = [torch.tensor(b, requires_grad=True) for b in bs]
    h = torch.tensor(X)
    for l in range(3):
        h = h @ tW[l] + tb[l]
        if l < 2:
            h = torch.relu(h)
    assert np.allclose(h.detach().numpy(), head, rtol=1e-12, atol=1e-12)
    (h * torch.tensor(G)).sum().backward()
    for l in range(3):
        assert np.allclose(tW[l].grad.numpy(), dW[l], rtol=1e-10, atol=1e-12)
        assert np.allclose(tb[l].grad.numpy(), db[l], rtol=1e-10, atol=1e-12)
