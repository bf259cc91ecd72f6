"""FP64 CPU reference of the surrogate's computations.

TEST INFRASTRUCTURE ONLY: the checker of the device kernels in
rebop_amd/csrc/surrogate.hip (loaded by tests/ only).

- ``bf16``: round-to-nearest-even to bfloat16, as the device casts do.
- ``forward`` / ``backward``: the ReLU MLP of rebop_amd.surrogate in FP64;
  with ``emulate_bf16`` the operands the device stores in BF16 (inputs,
  activations, weights, back-propagated gradients) are rounded the same way,
  so the only remaining difference is FP32 vs FP64 accumulation.
- ``nll``: the negative-binomial-mixture negative log-likelihood and its
  gradient w.r.t. the head, with scipy's gammaln / digamma.
- ``features``: the input features of features_kernel (and the bound of
  each species: E0 for E and ES, S0 for S and P).
"""

from __future__ import annotations

import numpy as np
from scipy.special import digamma, gammaln


def bf16(x: np.ndarray) -> np.ndarray:
    """Round (via float32) to the nearest bfloat16, ties to even; returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = ((u + 0x7FFF + lsb) >> 16) << 16
    out = (u.astype(np.uint32)).view(np.float32).astype(np.float64)
    nan = np.isnan(f)
    out[nan] = np.nan
    return out


def features(rates: np.ndarray, x0: np.ndarray, sample: np.ndarray, tmax: float, nb: int,
             width: int, with_offsets: bool = False):
    """The 16 input features of features_kernel (and optionally the log-mean offsets)."""
    B = len(sample)
    X = np.zeros((B, width))
    kf, kb, kc = rates[:, 0], rates[:, 1], rates[:, 2]
    e0, s0 = x0[:, 0].astype(np.float64), x0[:, 1].astype(np.float64)
    t = tmax * sample / nb
    lf, lb, lc = np.log10(kf), np.log10(kb), np.log10(kc)
    le, ls, lt = np.log10(e0), np.log10(s0), np.log10(t)
    lkm = np.log10(kb + kc) - lf
    X[:, 0] = (lf + 1.0) * 0.5
    X[:, 1] = (lb + 1.0) * 0.5
    X[:, 2] = (lc + 1.0) * 0.5
    X[:, 3] = (e0 - 105.0) / 95.0
    X[:, 4] = (s0 - 525.0) / 475.0
    X[:, 5] = 2.0 * t / tmax - 1.0
    X[:, 6] = (le - 1.65) / 0.65
    X[:, 7] = (ls - 2.35) / 0.65
    X[:, 8] = (lt - 0.95) / 0.75
    X[:, 9] = (lc + lt) / 3.0
    X[:, 10] = (lf + le + lt - 1.6) / 3.4
    X[:, 11] = (lf + ls + lt - 2.3) / 3.4
    X[:, 12] = (lkm - ls + 2.05) / 4.65
    X[:, 13] = (lb - lc) / 4.0
    X[:, 14] = (le - ls + 0.7) / 1.3
    X[:, 15] = (lc + le + lt - ls + 0.75) / 4.05
    if not with_offsets:
        return X
    bound = np.stack([e0, s0, e0, s0], axis=1)
    return X, bound


def forward(X, Ws, bs, emulate_bf16=True):
    """Returns (head, activations) in FP64."""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    H = [r(X)]
    for l, (W, b) in enumerate(zip(Ws, bs)):
        Z = H[-1] @ r(W) + b
        if l < len(Ws) - 1:
            H.append(r(np.maximum(Z, 0.0)))
        else:
            head = Z
    return head, H


def backward(dhead, H, Ws, emulate_bf16=True):
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    L = len(Ws)
    dW, db = [None] * L, [None] * L
    dH = np.asarray(dhead, dtype=np.float64)
    for l in reversed(range(L)):
        if l < L - 1:
            dH = dH * (H[l + 1] > 0)
        db[l] = dH.sum(axis=0)
        dZ = r(dH)
        dW[l] = H[l].T @ dZ
        if l > 0:
            dH = dZ @ r(Ws[l]).T
    return dW, db


def _softplus(z):
    return np.where(z > 0, z + np.log1p(np.exp(-np.abs(z))), np.log1p(np.exp(np.minimum(z, 0.0))))


def nll(head: np.ndarray, samples: np.ndarray, K: int, bound: np.ndarray | None = None):
    """head [B][>= S*3K], samples [B][S][J] -> (total NLL, d NLL / d head).

    Without bound every component is a negative binomial.  With bound [B][S]:
    components < K/2 are negative binomials with log means relative to
    ln bound, components >= K/2 are binomial(bound, sigmoid(head log-mean
    slot)) and their dispersion slot gets no gradient."""
    B, S, J = samples.shape
    grad = np.zeros_like(head, dtype=np.float64)
    total = 0.0
    for s in range(S):
        h = head[:, s * 3 * K:(s + 1) * 3 * K].astype(np.float64)
        logits, lmu_raw, lr_raw = h[:, :K], h[:, K:2 * K].copy(), h[:, 2 * K:]
        x = samples[:, s, :].astype(np.float64)[:, :, None]  # [B][J][1]
        nbk = K // 2 if bound is not None else K
        if bound is not None:
            bs = np.asarray(bound, dtype=np.float64)[:, s]
            lmu_raw[:, :nbk] += np.log(bs)[:, None]
        lz = np.logaddexp.reduce(logits, axis=1, keepdims=True)
        lpi = logits - lz  # [B][K]
        ll = np.zeros((B, J, K))
        # negative-binomial components
        lmu = np.clip(lmu_raw[:, :nbk], -8.0, 9.0)
        lr = np.clip(lr_raw[:, :nbk], -6.0, 12.0)
        mu3, r3 = np.exp(lmu)[:, None, :], np.exp(lr)[:, None, :]
        lrm = np.log(r3 + mu3)
        ll[:, :, :nbk] = (gammaln(x + r3) - gammaln(r3) - gammaln(x + 1.0) + r3 * (lr[:, None, :] - lrm)
                          + x * (lmu[:, None, :] - lrm))
        if bound is not None:
            th = np.clip(lmu_raw[:, nbk:], -30.0, 30.0)[:, None, :]
            bb = bs[:, None, None]
            lcb = gammaln(bb + 1.0) - gammaln(x + 1.0) - gammaln(np.maximum(bb - x, 0.0) + 1.0)
            ll[:, :, nbk:] = lcb - x * _softplus(-th) - (bb - x) * _softplus(th)
        lt = lpi[:, None, :] + ll  # [B][J][K]
        lp = np.logaddexp.reduce(lt, axis=2, keepdims=True)
        total -= lp.sum()
        gam = np.exp(lt - lp)
        pi = np.exp(lpi)[:, None, :]
        gl = (pi - gam).sum(axis=1)
        gn = gam[:, :, :nbk]
        gm = np.zeros((B, K))
        gr = np.zeros((B, K))
        gm[:, :nbk] = (-gn * r3 * (x - mu3) / (r3 + mu3)).sum(axis=1)
        gr[:, :nbk] = (-gn * r3 * (digamma(x + r3) - digamma(r3) + (lr[:, None, :] - lrm)
                                   + (mu3 - x) / (r3 + mu3))).sum(axis=1)
        clamp_m = np.zeros((B, K), dtype=bool)
        clamp_m[:, :nbk] = lmu_raw[:, :nbk] != lmu
        if bound is not None:
            p = 1.0 / (1.0 + np.exp(-th))
            gm[:, nbk:] = (-gam[:, :, nbk:] * (x - bb * p)).sum(axis=1)
            clamp_m[:, nbk:] = lmu_raw[:, nbk:] != th[:, 0, :]
        gm[clamp_m] = 0.0
        gr[:, :nbk][lr_raw[:, :nbk] != lr] = 0.0
        grad[:, s * 3 * K:s * 3 * K + K] = gl
        grad[:, s * 3 * K + K:s * 3 * K + 2 * K] = gm
        grad[:, s * 3 * K + 2 * K:(s + 1) * 3 * K] = gr
    return total, grad
