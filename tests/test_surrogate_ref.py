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


def test_nll_with_bounds_matches_scipy():
    """With bounds: negative-binomial log means relative to ln bound,
    components K/2.. binomial(bound, sigmoid(slot)) (scipy's pmfs, written out)."""
    rng = np.random.default_rng(4)
    B, S, J = 3, 4, 12
    head = _random_head(B, S, rng)
    for s in range(S):  # log-fractions of the bound
        head[:, s * 3 * K + K:s * 3 * K + 2 * K] = rng.normal(-1.0, 0.7, size=(B, K))
    bound = rng.integers(80, 200, size=(B, S)).astype(float)
    samples = np.minimum(rng.poisson(40, size=(B, S, J)), bound[:, :, None].astype(int))
    total, _ = ref.nll(head, samples, K, bound=bound)
    expect = 0.0
    for b in range(B):
        for s in range(S):
            h = head[b, s * 3 * K:(s + 1) * 3 * K]
            w = np.exp(h[:K] - np.logaddexp.reduce(h[:K]))
            mu, r = bound[b, s] * np.exp(h[K:2 * K]), np.exp(h[2 * K:])
            pbin = 1.0 / (1.0 + np.exp(-h[K:2 * K]))
            for x in samples[b, s]:
                p = 0.0
                for c in range(K):
                    if c < K // 2:
                        p += w[c] * stats.nbinom.pmf(x, r[c], r[c] / (r[c] + mu[c]))
                    else:
                        p += w[c] * stats.binom.pmf(x, int(bound[b, s]), pbin[c])
                expect -= np.log(p)
    assert total == pytest.approx(expect, rel=1e-10)


def test_nll_with_bounds_gradient_matches_finite_differences():
    rng = np.random.default_rng(6)
    B, S, J = 2, 4, 8
    head = _random_head(B, S, rng)
    for s in range(S):
        head[:, s * 3 * K + K:s * 3 * K + 2 * K] = rng.normal(-1.0, 0.5, size=(B, K))
    bound = rng.integers(60, 120, size=(B, S)).astype(float)
    samples = np.minimum(rng.poisson(30, size=(B, S, J)), bound[:, :, None].astype(int))
    _, g = ref.nll(head, samples, K, bound=bound)
    for idx in [(0, K + 1), (1, K + 6), (0, 2 * K + 2), (1, 3 * K + 2 * K + 1), (0, 3 * 3 * K + 7),
                (1, 3 * K + K + 5), (0, 3)]:
        h1, h2 = head.copy(), head.copy()
        eps = 1e-6
        h1[idx] += eps
        h2[idx] -= eps
        fd = (ref.nll(h1, samples, K, bound=bound)[0] - ref.nll(h2, samples, K, bound=bound)[0]) / (2 * eps)
        assert g[idx] == pytest.approx(fd, rel=1e-5, abs=1e-7), idx


def test_features_are_the_stated_groups():
    """Each feature against its definition computed from the raw parameters
    (catches swapped rate constants or species)."""
    rates = np.array([[0.01, 2.0, 0.3], [5.0, 0.002, 1.5]])
    x0 = np.array([[20, 800, 0, 0], [150, 60, 0, 0]])
    sample = np.array([4.0, 32.0])
    X, bnd = ref.features(rates, x0, sample, 50.0, 32, 64, with_offsets=True)
    t = 50.0 * sample / 32
    kf, kb, kc = rates.T
    e0, s0 = x0[:, 0].astype(float), x0[:, 1].astype(float)
    assert np.allclose(10 ** (2 * X[:, 0] - 1), kf) and np.allclose(10 ** (2 * X[:, 2] - 1), kc)
    assert np.allclose(X[:, 3] * 95 + 105, e0) and np.allclose(X[:, 4] * 475 + 525, s0)
    assert np.allclose((X[:, 5] + 1) / 2 * 50, t)
    assert np.allclose(10 ** (3 * X[:, 9]), kc * t)
    assert np.allclose(10 ** (3.4 * X[:, 10] + 1.6), kf * e0 * t)
    assert np.allclose(10 ** (3.4 * X[:, 11] + 2.3), kf * s0 * t)
    assert np.allclose(10 ** (4.65 * X[:, 12] - 2.05), (kb + kc) / kf / s0)
    assert np.allclose(10 ** (4 * X[:, 13]), kb / kc)
    assert np.allclose(10 ** (1.3 * X[:, 14] - 0.7), e0 / s0)
    assert np.allclose(10 ** (4.05 * X[:, 15] - 0.75), kc * e0 * t / s0)
    assert np.all(X[:, 16:] == 0)
    assert np.array_equal(bnd, np.stack([e0, s0, e0, s0], axis=1))


def test_w1_batch_equals_single_query_w1():
    """The gate's vectorised W1 equals the one-query definition (mixture CDF
    against the empirical CDF of integer samples), bounded and unbounded."""
    from rebop_amd import surrogate as sg

    rng = np.random.default_rng(5)
    P = 40
    params = np.zeros((P, 3, K))
    params[:, 0] = rng.normal(size=(P, K))
    params[:, 1] = rng.normal(np.log(30.0), 0.8, size=(P, K))
    params[:, 2] = rng.normal(1.0, 1.0, size=(P, K))
    samples = rng.poisson(rng.uniform(5, 60, size=(P, 1)), size=(P, 500))
    # one common largest sample, so that both forms sum the CDFs over the same range
    samples = np.minimum(samples, 80)
    samples[:, 0] = 80
    bounds = rng.integers(60, 120, size=P).astype(np.float64)
    for b in (None, bounds):
        got = sg.w1_batch(params, samples, b)
        want = [sg.w1_to_samples(params[q], samples[q], None if b is None else b[q]) for q in range(P)]
        assert np.allclose(got, want, rtol=1e-9, atol=1e-9)
    # a mixture at 0 (means clipped to e^-8) against samples at 0: W1 = its
    # mean; against samples at 1: W1 = 1 - its mean
    pm = np.zeros((1, 3, K))
    pm[0, 1] = -30.0
    pm[0, 2] = 12.0  # dispersion large: Poisson-like
    m = np.exp(-8.0)
    assert abs(sg.w1_batch(pm, np.zeros((1, 10), dtype=np.int64), None)[0] - m) < 1e-6
    assert abs(sg.w1_batch(pm, np.ones((1, 10), dtype=np.int64), None)[0] - (1.0 - m)) < 1e-6


def test_trust_map_admits_only_validated_passing_cells():
    """A (cell, time, species) is trusted only if the cell held min_points
    validated points and all passed there; unvalidated sample times are never
    trusted (the map does not extrapolate in time)."""
    from rebop_amd import surrogate as sg

    rates = np.array([[1e-2, 1e-2, 1e-2]] * 3 + [[5.0, 5.0, 5.0]] * 2)
    x0 = np.array([[50, 300, 0, 0]] * 5)
    times = (4, 8)
    rel = np.full((5, 2, sg.S), 0.01)
    rel[1, 1, 2] = 0.03  # one failing point in the first cell at t-index 8, species 2
    tm = sg.TrustMap.build(rel, rates, x0, times, bins=2, tol=0.02, min_points=3)
    q = np.array([[1e-2, 1e-2, 1e-2], [5.0, 5.0, 5.0]])
    qx = np.array([[50, 300, 0, 0]] * 2)
    assert tm.trusted(q, qx, 4, 0).tolist() == [True, False]  # the second cell has 2 < 3 points
    assert tm.trusted(q, qx, 8, 2).tolist() == [False, False]
    assert tm.trusted(q, qx, 8, 1).tolist() == [True, False]
    assert tm.trusted(q, qx, 5, 0).tolist() == [False, False]  # time 5 was not validated
    assert sg.route(tm, q, qx, np.array([4, 4])).tolist() == [True, False]
    assert sg.route(tm, q, qx, np.array([8, 8])).tolist() == [False, False]
