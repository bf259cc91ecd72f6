"""Learned surrogate of the SSA for the Michaelis-Menten network.

E + S -> ES @ kf, ES -> E + S @ kb, ES -> E + P @ kcat (species E, S, ES, P).
A multilayer perceptron maps (log10 kf, log10 kb, log10 kcat, E0, S0, t) to
a K-component negative-binomial mixture for each species.  It is trained on
trajectories this engine generates (``rebop_run_sweep_device``: one
launch over every parameter point) by the project's own device kernels
(include/rebop_surrogate.h): BF16 matrix-core GEMMs with FP32 accumulation
for the dense layers, an FP64 mixture-likelihood head, Adam.

The SSA stays the reference and the fallback: ``validate`` compares the
surrogate with fresh SSA histograms (10^4 trajectories per point) at every
sample time; ``TrustMap`` admits a (parameter cell, sample time, species)
only where the cell held enough validated points and every one had a
Wasserstein-1 distance <= 2 % of the species mean; ``route`` sends each
query to the surrogate only where the map admits all its species, and
``answer`` runs the SSA for the rest.
"""

from __future__ import annotations

import ctypes
import math
import os
from ctypes import c_double, c_float, c_int, c_size_t, c_uint64, c_void_p
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _native
from .engine import GillespieNet, Rate

LIB_PATH = Path(__file__).resolve().parent / "librebop_surrogate.so"
K = 8  # mixture components per species (REBOP_SG_COMPONENTS)
S = 4  # species E, S, ES, P
IN = 64  # input width: NF features, zero-padded for the matrix cores
NF = 16  # features per item (REBOP_SG_FEATURES)
OUT = 128  # head width, S * 3 * K = 96 used
WIDTH = 512
TMAX = 50.0
NB = 32  # samples at t = 50 i / 32, i = 1..32
J = 64  # SSA trajectories per parameter point in the training data
MLP_DIMS = [IN] + [WIDTH] * 4 + [OUT]  # five dense layers

_lib = None

_i, _v, _d, _f, _z = c_int, c_void_p, c_double, c_float, c_size_t
# every entry point of include/rebop_surrogate.h: name -> (restype, argtypes)
SIGNATURES: dict[str, tuple] = {
    "rebop_sg_last_error": (ctypes.c_char_p, []),
    "rebop_sg_gemm_bf16": (_i, [_i, _i, _i, _i, _i, _v, _i, _v, _i, _v, _i, _v]),
    "rebop_sg_gemm_bias_act_bf16": (_i, [_i, _i, _i, _i, _v, _i, _v, _i, _v, _i, _v, _i, _v, _i, _v]),
    "rebop_sg_transpose_bf16": (_i, [_i, _i, _v, _v, _v]),
    "rebop_sg_bias_act": (_i, [_i, _i, _v, _i, _v, _i, _v, _i, _v, _i, _v]),
    "rebop_sg_relu_grad": (_i, [_i, _i, _v, _i, _v, _i, _v, _i, _v]),
    "rebop_sg_cast_bf16": (_i, [_z, _v, _v, _v]),
    "rebop_sg_colsum": (_i, [_i, _i, _v, _i, _v, _v]),
    "rebop_sg_adam": (_i, [_z, _v, _v, _v, _v, _v, _f, _f, _f, _f, _f, _f, _v]),
    "rebop_sg_features": (_i, [_i, _v, _v, _v, _v, _d, _i, _v, _i, _v, _v]),
    "rebop_sg_head_offsets": (_i, [_i, _i, _i, _v, _i, _v, _v]),
    "rebop_sg_pack_kslices_bf16": (_i, [_i, _i, _v, _v, _v]),
    "rebop_sg_mlp_forward_bf16": (_i, [_i, _v, _i, _i, _v, _v, _v, _v, _i, _v]),
    "rebop_sg_infer": (_i, [_i, _v, _v, _v, _v, _d, _i, _i, _v, _v, _v, _v, _i, _v]),
    "rebop_sg_nb_mixture_nll": (_i, [_i, _i, _i, _i, _v, _i, _v, _v, _v, c_uint64, _v, _v, _d, _v, _v]),
}


def load():
    global _lib
    if _lib is not None:
        return _lib
    _native.load()  # one HIP runtime with torch (see _native.load)
    if not LIB_PATH.exists():
        raise _native.RebopError(4, f"{LIB_PATH} is missing: build with `python -m rebop_amd.build`")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ck(rc: int) -> None:
    if rc != 0:
        raise _native.RebopError(rc, load().rebop_sg_last_error().decode())


def _p(t) -> int:
    return t.data_ptr()


# ---------------------------------------------------------------------------
# thin wrappers of the device kernels (torch tensors as device buffers)


def gemm(A, B, C, ta=False, tb=False, stream=None):
    """C[M][N] = op(A) op(B); A, B bfloat16, C float32, all row-major."""
    M, N = C.shape
    Kd = A.shape[0] if ta else A.shape[1]
    _ck(load().rebop_sg_gemm_bf16(int(ta), int(tb), M, N, Kd, _p(A), A.shape[1], _p(B), B.shape[1],
                                  _p(C), C.shape[1], stream))


def gemm_bias_act(A, B, bias, relu, H=None, F=None, tb=False, stream=None):
    """H = bf16(act(A op(B) + bias)) and/or F = act(A op(B) + bias) in one
    launch (the values of gemm followed by bias_act, without the FP32
    product); tb: B is stored [N][K]."""
    out = H if H is not None else F
    M, N = out.shape
    _ck(load().rebop_sg_gemm_bias_act_bf16(int(tb), M, N, A.shape[1], _p(A), A.shape[1], _p(B), B.shape[1],
                                           _p(bias), int(relu),
                                           _p(H) if H is not None else None, H.shape[1] if H is not None else 0,
                                           _p(F) if F is not None else None, F.shape[1] if F is not None else 0,
                                           stream))


def bias_act(C, bias, relu, H=None, F=None, stream=None):
    M, N = C.shape
    _ck(load().rebop_sg_bias_act(M, N, _p(C), C.shape[1], _p(bias), int(relu),
                                 _p(H) if H is not None else None, H.shape[1] if H is not None else 0,
                                 _p(F) if F is not None else None, F.shape[1] if F is not None else 0,
                                 stream))


def relu_grad(dH, H, dZ, stream=None):
    M, N = dH.shape
    _ck(load().rebop_sg_relu_grad(M, N, _p(dH), dH.shape[1], _p(H), H.shape[1], _p(dZ), dZ.shape[1], stream))


def cast_bf16(x, y, stream=None):
    _ck(load().rebop_sg_cast_bf16(x.numel(), _p(x), _p(y), stream))


def colsum(X, out, stream=None):
    M, N = X.shape
    _ck(load().rebop_sg_colsum(M, N, _p(X), X.shape[1], _p(out), stream))


def transpose_bf16(x, y, stream=None):
    """y = bf16(x) transposed (x float32 [R][C], y bfloat16 [C][R])."""
    R, C = x.shape
    _ck(load().rebop_sg_transpose_bf16(R, C, _p(x), _p(y), stream))


def adam(w, g, m, v, wb, lr, b1, b2, eps, t, stream=None):
    bc1 = 1.0 - b1**t
    bc2 = 1.0 - b2**t
    _ck(load().rebop_sg_adam(w.numel(), _p(w), _p(g), _p(m), _p(v), _p(wb) if wb is not None else None,
                             lr, b1, b2, eps, bc1, bc2, stream))


def features(item_point, item_sample, rates, x0, X, bound=None, stream=None):
    """X [B][IN] bf16 features; bound [B][S] float32 (optional) the species' bounds."""
    B = item_point.numel()
    _ck(load().rebop_sg_features(B, _p(item_point), _p(item_sample), _p(rates), _p(x0), TMAX, NB,
                                 _p(X), X.shape[1], _p(bound) if bound is not None else None,
                                 stream))


def head_offsets(head, bound, stream=None):
    """Adds ln bound to the head's log means in place (head -> mixture parameters)."""
    B = head.shape[0]
    _ck(load().rebop_sg_head_offsets(B, S, K, _p(head), head.shape[1], _p(bound), stream))


def pack_kslices(W, Wp, stream=None):
    """Wp [K/32][N][32] bfloat16 = the k-slice-major copy of W [K][N] float32
    (the layout the fused forward pass reads)."""
    Kd, N = W.shape
    _ck(load().rebop_sg_pack_kslices_bf16(Kd, N, _p(W), _p(Wp), stream))


def packed(W):
    import torch

    Wp = torch.empty((W.shape[0] // 32, W.shape[1], 32), dtype=torch.bfloat16, device="cuda")
    pack_kslices(W, Wp)
    return Wp


def _layer_table(Wp, bias):
    """Host arrays (dims, weight pointers, bias pointers) of a packed layer stack."""
    depth = len(Wp)
    dims = (ctypes.c_int * (depth + 1))(Wp[0].shape[0] * 32, *[w.shape[1] for w in Wp])
    wp = (ctypes.c_void_p * depth)(*[_p(w) for w in Wp])
    bp = (ctypes.c_void_p * depth)(*[_p(b) for b in bias])
    return depth, dims, wp, bp


def mlp_forward(X, Wp, bias, out, stream=None):
    """out = the forward pass of X through the layers (Wp[l] packed by
    pack_kslices, bias[l] float32; ReLU and BF16 between layers) in one
    launch, with the activations kept on chip: the bits of gemm_bias_act
    layer by layer."""
    depth, dims, wp, bp = _layer_table(Wp, bias)
    _ck(load().rebop_sg_mlp_forward_bf16(X.shape[0], _p(X), X.shape[1], depth, ctypes.addressof(dims),
                                         ctypes.addressof(wp), ctypes.addressof(bp), _p(out),
                                         out.shape[1], stream))


def infer(Wp, bias, item_point, item_sample, rates, x0, head, stream=None):
    """head[b] = the mixture parameters of item b: features, the forward pass
    and head_offsets in one launch (the same bits as the three calls)."""
    depth, dims, wp, bp = _layer_table(Wp, bias)
    _ck(load().rebop_sg_infer(item_point.numel(), _p(item_point), _p(item_sample), _p(rates), _p(x0), TMAX,
                              NB, depth, ctypes.addressof(dims), ctypes.addressof(wp),
                              ctypes.addressof(bp), _p(head), head.shape[1], stream))


def nb_mixture_nll(head, item_point, item_sample, samples, grad, loss_sum, scale, bound=None,
                   stream=None):
    B = item_point.numel()
    _ck(load().rebop_sg_nb_mixture_nll(B, S, K, J, _p(head), head.shape[1], _p(item_point),
                                       _p(item_sample), _p(samples), samples.shape[2], _p(grad),
                                       _p(loss_sum), float(scale),
                                       _p(bound) if bound is not None else None, stream))


# ---------------------------------------------------------------------------
# training data


def mm_network(seed: int = 0) -> GillespieNet:
    """The Michaelis-Menten network; its constants are replaced per point."""
    g = GillespieNet([100, 500, 0, 0], seed=seed)
    g.add_reaction(Rate.lma(1.0, [1, 1, 0, 0]), [-1, -1, 1, 0])
    g.add_reaction(Rate.lma(1.0, [0, 0, 1, 0]), [1, 1, -1, 0])
    g.add_reaction(Rate.lma(1.0, [0, 0, 1, 0]), [1, 0, -1, 1])
    return g


def sample_points(n: int, seed: int):
    """Rate constants log-uniform in [1e-3, 1e1]; E0 ~ U{10..200}, S0 ~ U{50..1000}."""
    rng = np.random.default_rng(seed)
    rates = 10.0 ** rng.uniform(-3.0, 1.0, size=(n, 3))
    x0 = np.zeros((n, 4), dtype=np.int64)
    x0[:, 0] = rng.integers(10, 201, n)
    x0[:, 1] = rng.integers(50, 1001, n)
    return rates, x0


@dataclass
class SweepData:
    rates: "torch.Tensor"  # [P][3] float64
    x0: "torch.Tensor"  # [P][4] int64
    samples: "torch.Tensor"  # [NB+1][S][P*n_traj] int64
    n_traj: int
    stats: _native.Stats


def generate(rates: np.ndarray, x0: np.ndarray, n_traj: int, seed: int, traj_begin: int = 0) -> SweepData:
    """n_traj SSA trajectories per point, sampled at t = 50 i / 32 (one launch)."""
    import torch

    P = len(rates)
    g = mm_network(seed)
    d_rates = torch.as_tensor(np.ascontiguousarray(rates, dtype=np.float64), device="cuda")
    d_x0 = torch.as_tensor(np.ascontiguousarray(x0, dtype=np.int64), device="cuda")
    out = torch.empty((NB + 1, S, P * n_traj), dtype=torch.int64, device="cuda")
    st = g.run_sweep_device(P, n_traj, TMAX, NB, out.data_ptr(), rates_ptr=d_rates.data_ptr(),
                            x0_ptr=d_x0.data_ptr(), traj_begin=traj_begin,
                            stream=torch.cuda.current_stream().cuda_stream)
    return SweepData(d_rates, d_x0, out, n_traj, st)


# ---------------------------------------------------------------------------
# the network


class MLP:
    """Dense ReLU network IN -> WIDTH (x4) -> OUT with FP32 master weights."""

    def __init__(self, seed: int = 0, width: int = WIDTH, depth: int = 5):
        import torch

        g = torch.Generator(device="cpu").manual_seed(seed)
        dims = [IN] + [width] * (depth - 1) + [OUT]
        self.dims = dims
        # FP32 master weights W [in][out], their BF16 copy Wb (the backward's
        # operand), its transpose Wt [out][in] (the training forward's: the
        # layout the GEMM loads fastest) and its k-slice-major copy Wp (the
        # fused inference pass's)
        self.W, self.b, self.Wb, self.Wt, self.Wp = [], [], [], [], []
        for l in range(depth):
            fan_in = NF if l == 0 else dims[l]
            w = torch.randn(dims[l], dims[l + 1], generator=g, dtype=torch.float64) * math.sqrt(2.0 / fan_in)
            if l == 0:
                w[NF:, :] = 0.0  # padded inputs
            if l == depth - 1:
                w *= 0.1
                w[:, S * 3 * K:] = 0.0  # padded head columns
            b = torch.zeros(dims[l + 1], dtype=torch.float64)
            if l == depth - 1:
                for s in range(S):
                    b[s * 3 * K + K: s * 3 * K + K + K // 2] = math.log(0.5)  # log mean / bound
                    b[s * 3 * K + K + K // 2: s * 3 * K + 2 * K] = 0.0  # binomial logit
                    b[s * 3 * K + 2 * K: s * 3 * K + 3 * K] = math.log(10.0)  # log dispersion
            self.W.append(w.float().cuda())
            self.b.append(b.float().cuda())
            self.Wb.append(self.W[-1].to(torch.bfloat16))
            self.Wt.append(self.W[-1].t().contiguous().to(torch.bfloat16))
            self.Wp.append(packed(self.W[-1]))
        self.mW = [torch.zeros_like(w) for w in self.W]
        self.vW = [torch.zeros_like(w) for w in self.W]
        self.mb = [torch.zeros_like(b) for b in self.b]
        self.vb = [torch.zeros_like(b) for b in self.b]
        self.t = 0

    @property
    def depth(self) -> int:
        return len(self.W)

    def forward(self, X, keep: bool = True):
        """X [B][IN] bf16 -> head [B][OUT] float32 (activations kept for backward)."""
        import torch

        B = X.shape[0]
        H = [X]
        for l in range(self.depth):
            # product, bias and activation in one launch (same values as
            # gemm + bias_act)
            if l < self.depth - 1:
                h = torch.empty((B, self.dims[l + 1]), dtype=torch.bfloat16, device="cuda")
                gemm_bias_act(H[-1], self.Wt[l], self.b[l], True, H=h, tb=True)
                H.append(h)
            else:
                head = torch.empty((B, self.dims[l + 1]), dtype=torch.float32, device="cuda")
                gemm_bias_act(H[-1], self.Wt[l], self.b[l], False, F=head, tb=True)
        if keep:
            self._H = H
        return head

    def backward(self, dhead):
        """dhead [B][OUT] float32 -> (dW, db) per layer."""
        import torch

        H = self._H
        B = dhead.shape[0]
        dW = [None] * self.depth
        db = [None] * self.depth
        dH = dhead.clone()
        for l in reversed(range(self.depth)):
            dZ = torch.empty((B, self.dims[l + 1]), dtype=torch.bfloat16, device="cuda")
            if l == self.depth - 1:
                cast_bf16(dH, dZ)
            else:
                relu_grad(dH, H[l + 1], dZ)
            db[l] = torch.empty(self.dims[l + 1], dtype=torch.float32, device="cuda")
            colsum(dH, db[l])
            dW[l] = torch.empty((self.dims[l], self.dims[l + 1]), dtype=torch.float32, device="cuda")
            gemm(H[l], dZ, dW[l], ta=True)
            if l > 0:
                dH = torch.empty((B, self.dims[l]), dtype=torch.float32, device="cuda")
                gemm(dZ, self.Wb[l], dH, tb=True)
        return dW, db

    def step(self, dW, db, lr: float, b1=0.9, b2=0.999, eps=1e-8):
        self.t += 1
        for l in range(self.depth):
            adam(self.W[l], dW[l], self.mW[l], self.vW[l], self.Wb[l], lr, b1, b2, eps, self.t)
            transpose_bf16(self.W[l], self.Wt[l])
            pack_kslices(self.W[l], self.Wp[l])
            adam(self.b[l], db[l], self.mb[l], self.vb[l], None, lr, b1, b2, eps, self.t)


def train(mlp: MLP, data: SweepData, steps: int, batch: int = 4096, lr: float = 1e-3,
          seed: int = 0, log_every: int = 0, log=None, warmup: int = 200, lr_min: float = 1e-5):
    """Adam on the mean mixture NLL of the SSA samples; returns the loss history.

    The learning rate rises linearly over ``warmup`` steps, then follows a
    cosine from ``lr`` down to ``lr_min``.
    """
    import torch

    P = data.rates.shape[0]
    n_items = P * NB
    gen = torch.Generator(device="cuda").manual_seed(seed)
    X = torch.empty((batch, IN), dtype=torch.bfloat16, device="cuda")
    off = torch.empty((batch, S), dtype=torch.float32, device="cuda")
    grad = torch.zeros((batch, OUT), dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    hist = []
    for it in range(steps):
        q = torch.randint(0, n_items, (batch,), generator=gen, device="cuda")
        ip = (q // NB).to(torch.int32)  # read as uint32 by the kernels
        isamp = (q % NB + 1).to(torch.int32)
        features(ip, isamp, data.rates, data.x0, X, off)
        head = mlp.forward(X)
        grad.zero_()
        loss.zero_()
        nb_mixture_nll(head, ip, isamp, data.samples, grad, loss, 1.0 / (batch * J), off)
        dW, db = mlp.backward(grad)
        if it < warmup:
            lr_t = lr * (it + 1) / warmup
        else:
            x = (it - warmup) / max(steps - warmup, 1)
            lr_t = lr_min + (lr - lr_min) * 0.5 * (1.0 + math.cos(math.pi * x))
        mlp.step(dW, db, lr_t)
        if log_every and (it % log_every == 0 or it == steps - 1):
            hist.append((it, float(loss.item()) / (batch * J)))
            if log:
                log(f"train step {it}: NLL/sample {hist[-1][1]:.4f}")
    return hist


def predict(mlp: MLP, rates: np.ndarray, x0: np.ndarray, sample_idx: np.ndarray, chunk: int = 1 << 16):
    """Mixture parameters for queries (point rows, sample index 1..32): [Q][S][3][K],
    and the species' bounds [Q][S] (components K/2.. are binomials on the bound)."""
    import torch

    rates_t = torch.as_tensor(np.ascontiguousarray(rates, dtype=np.float64), device="cuda")
    x0_t = torch.as_tensor(np.ascontiguousarray(x0, dtype=np.int64), device="cuda")
    Q = len(sample_idx)
    out = np.zeros((Q, S, 3, K))
    bounds = np.zeros((Q, S))
    for a in range(0, Q, chunk):
        b = min(Q, a + chunk)
        B = ((b - a + 63) // 64) * 64
        ip = torch.zeros(B, dtype=torch.int32, device="cuda")
        isamp = torch.ones(B, dtype=torch.int32, device="cuda")
        ip[: b - a] = torch.arange(a, b, dtype=torch.int32, device="cuda")
        isamp[: b - a] = torch.as_tensor(sample_idx[a:b].astype(np.int32), device="cuda")
        X = torch.empty((B, IN), dtype=torch.bfloat16, device="cuda")
        off = torch.empty((B, S), dtype=torch.float32, device="cuda")
        features(ip, isamp, rates_t, x0_t, X, off)
        head = mlp.forward(X, keep=False)
        head_offsets(head, off)
        out[a:b] = head[: b - a, : S * 3 * K].double().cpu().numpy().reshape(b - a, S, 3, K)
        bounds[a:b] = off[: b - a].double().cpu().numpy()
    return out, bounds


def infer_device(mlp: MLP, d_rates, d_x0, d_point, d_sample, B: int):
    """Device-only inference: the mixture parameters [B][OUT] of B queries,
    in one launch (features, forward pass and head offsets fused)."""
    import torch

    head = torch.empty((B, OUT), dtype=torch.float32, device="cuda")
    infer(mlp.Wp, mlp.b, d_point, d_sample, d_rates, d_x0, head)
    return head


def infer_device_layers(mlp: MLP, d_rates, d_x0, d_point, d_sample, B: int):
    """infer_device as three steps (features, layer-by-layer forward pass,
    head offsets): the same values, with the activations through memory."""
    import torch

    X = torch.empty((B, IN), dtype=torch.bfloat16, device="cuda")
    off = torch.empty((B, S), dtype=torch.float32, device="cuda")
    features(d_point, d_sample, d_rates, d_x0, X, off)
    head = mlp.forward(X, keep=False)
    head_offsets(head, off)
    return head


# ---------------------------------------------------------------------------
# distributions and the Wasserstein-1 gate


def mixture_cdf(params: np.ndarray, xmax: int, bound: float | None = None) -> np.ndarray:
    """CDF on 0..xmax of one species' mixture (params [3][K]); with a bound,
    components K/2.. are binomial(bound, sigmoid(params[1][c]))."""
    from scipy import stats

    logits, lmu, lr = params
    w = np.exp(logits - logits.max())
    w /= w.sum()
    xs = np.arange(xmax + 1)
    pmf = np.zeros(xmax + 1)
    for c in range(K):
        if bound is not None and c >= K // 2:
            p = 1.0 / (1.0 + np.exp(-np.clip(lmu[c], -30, 30)))
            pc = stats.binom.pmf(xs, int(round(bound)), p)
        else:
            mu = np.exp(np.clip(lmu[c], -8, 9))
            r = np.exp(np.clip(lr[c], -6, 12))
            pc = stats.nbinom.pmf(xs, r, r / (r + mu))
        pmf += w[c] * pc
    return np.minimum(np.cumsum(pmf), 1.0)


def w1_to_samples(params: np.ndarray, samples: np.ndarray, bound: float | None = None) -> float:
    """W1 between the mixture and an empirical distribution of integer samples."""
    xmax = int(max(samples.max(), 1) * 2 + 50)
    if bound is not None:
        xmax = max(xmax, int(bound) + 1)
    F = mixture_cdf(params, xmax, bound)
    emp = np.searchsorted(np.sort(samples), np.arange(xmax + 1), side="right") / len(samples)
    return float(np.abs(F - emp).sum())


def mixture_cdf_batch(params: np.ndarray, xmax: int, bounds: np.ndarray | None) -> np.ndarray:
    """mixture_cdf for many queries at once: params [P][3][K], bounds [P]
    (or None) -> CDF [P][xmax + 1]."""
    from scipy import stats

    logits, lmu, lr = params[:, 0], params[:, 1], params[:, 2]
    w = np.exp(logits - logits.max(axis=1, keepdims=True))
    w /= w.sum(axis=1, keepdims=True)
    xs = np.arange(xmax + 1)[None, None, :]
    nbk = K // 2 if bounds is not None else K
    mu = np.exp(np.clip(lmu[:, :nbk], -8, 9))[..., None]
    r = np.exp(np.clip(lr[:, :nbk], -6, 12))[..., None]
    pmf = (w[:, :nbk, None] * stats.nbinom.pmf(xs, r, r / (r + mu))).sum(axis=1)
    if bounds is not None:
        pb = 1.0 / (1.0 + np.exp(-np.clip(lmu[:, nbk:], -30, 30)))[..., None]
        n = np.round(bounds).astype(np.int64)[:, None, None]
        pmf += (w[:, nbk:, None] * stats.binom.pmf(xs, n, pb)).sum(axis=1)
    return np.minimum(np.cumsum(pmf, axis=1), 1.0)


def w1_batch(params: np.ndarray, samples: np.ndarray, bounds: np.ndarray | None) -> np.ndarray:
    """w1_to_samples for P queries: params [P][3][K], integer samples [P][n],
    bounds [P] (or None) -> W1 [P]."""
    P, n = samples.shape
    xmax = int(max(samples.max(), 1) * 2 + 50)
    if bounds is not None:
        xmax = max(xmax, int(bounds.max()) + 1)
    out = np.zeros(P)
    for a in range(0, P, 128):  # bounded memory: 128 x K x (xmax + 1) doubles
        b = min(P, a + 128)
        F = mixture_cdf_batch(params[a:b], xmax, None if bounds is None else bounds[a:b])
        for q in range(b - a):
            cnt = np.bincount(np.clip(samples[a + q], 0, xmax), minlength=xmax + 1)
            out[a + q] = np.abs(F[q] - np.cumsum(cnt) / n).sum()
    return out


SAMPLE_TIMES = tuple(range(1, NB + 1))  # every sample index a query can ask for


@dataclass
class TrustMap:
    """Where the surrogate may answer: per cell of (log kf, log kb, log kcat,
    E0, S0), sample time and species, whether the cell held at least
    ``min_points`` validated points and every one of them passed there (W1 <=
    ``tol`` of the species mean).  A sample time that was not validated is
    never trusted; the SSA answers everything that is not."""

    bins: int
    times: tuple  # the validated sample indices
    passed: np.ndarray  # bool [bins]^3 x 2 x 2 x len(times) x S
    seen: np.ndarray  # int [bins]^3 x 2 x 2: validated points per cell

    @staticmethod
    def cell(rates: np.ndarray, x0: np.ndarray, bins: int) -> tuple:
        lk = np.clip(((np.log10(rates) + 3.0) / 4.0 * bins).astype(int), 0, bins - 1)
        e = (x0[:, 0] > 105).astype(int)
        s = (x0[:, 1] > 525).astype(int)
        return lk[:, 0], lk[:, 1], lk[:, 2], e, s

    @staticmethod
    def build(rel: np.ndarray, rates: np.ndarray, x0: np.ndarray, times, bins: int, tol: float,
              min_points: int) -> "TrustMap":
        """From validated relative W1 [P][len(times)][S] of points (rates, x0)."""
        ok = rel <= tol
        passed = np.ones((bins,) * 3 + (2, 2, len(times), S), dtype=bool)
        seen = np.zeros((bins,) * 3 + (2, 2), dtype=np.int64)
        cells = TrustMap.cell(rates, x0, bins)
        for p in range(len(rates)):
            c = tuple(ci[p] for ci in cells)
            seen[c] += 1
            passed[c] &= ok[p]
        passed &= (seen >= min_points)[..., None, None]
        return TrustMap(bins, tuple(int(t) for t in times), passed, seen)

    def trusted(self, rates: np.ndarray, x0: np.ndarray, sample, species) -> np.ndarray:
        """Per query (rows of rates/x0; sample indices 1..32 and species,
        scalars or per-query arrays): may the surrogate answer?"""
        sample = np.broadcast_to(np.asarray(sample, dtype=np.int64), (len(rates),))
        species = np.broadcast_to(np.asarray(species, dtype=np.int64), (len(rates),))
        lut = np.full(NB + 2, -1, dtype=np.int64)  # sample index -> validated position
        lut[np.asarray(self.times, dtype=np.int64)] = np.arange(len(self.times))
        ti = lut[np.clip(sample, 0, NB + 1)]
        out = np.zeros(len(rates), dtype=bool)
        v = ti >= 0
        if v.any():
            c = tuple(ci[v] for ci in self.cell(rates[v], x0[v], self.bins))
            out[v] = self.passed[c + (ti[v], species[v])]
        return out


def validate(mlp: MLP, rates: np.ndarray, x0: np.ndarray, n_ref: int, seed: int, tol: float = 0.02,
             bins: int = 2, times=SAMPLE_TIMES, min_points: int = 1, trust_tol: float | None = None,
             traj_begin: int = 1 << 40):
    """W1 of the surrogate against fresh n_ref-trajectory SSA histograms at
    every sample time in ``times``.

    Returns (relative W1 [P][len(times)][S], TrustMap, SSA stats): (cell,
    time, species) is trusted if the cell holds at least ``min_points``
    validated points and each has W1 <= trust_tol (default tol) * mean there.
    """
    data = generate(rates, x0, n_ref, seed=seed, traj_begin=traj_begin)
    samp = data.samples.cpu().numpy()
    P = len(rates)
    rel = np.zeros((P, len(times), S))
    for ti, i in enumerate(times):
        params, bounds = predict(mlp, rates, x0, np.full(P, i))
        for s in range(S):
            xs = samp[i, s].reshape(P, n_ref)
            w1 = w1_batch(params[:, s], xs, bounds[:, s])
            mean = xs.mean(axis=1)
            with np.errstate(divide="ignore", invalid="ignore"):
                rel[:, ti, s] = np.where(mean > 0, w1 / np.where(mean > 0, mean, 1.0),
                                         np.where(w1 <= tol, 0.0, np.inf))
    trust = TrustMap.build(rel, rates, x0, times, bins, tol if trust_tol is None else trust_tol,
                           min_points)
    return rel, trust, data.stats


def ssa_noise_floor(rates: np.ndarray, x0: np.ndarray, n_ref: int, seed: int,
                    times=SAMPLE_TIMES, traj_begin: int = 1 << 40) -> np.ndarray:
    """Relative W1 between validate's reference histograms (same seed and
    traj_begin) and a second, independent n_ref-trajectory SSA run of the
    same points [P][times][S]: the spread the 2 % gate has to sit above.
    Where a species' mean is near zero, 2 % of the mean is a small fraction
    of one molecule and two SSA runs already differ by more."""
    a = generate(rates, x0, n_ref, seed=seed, traj_begin=traj_begin).samples.cpu().numpy()
    b = generate(rates, x0, n_ref, seed=seed + 7, traj_begin=traj_begin + (1 << 39)).samples.cpu().numpy()
    P = len(rates)
    out = np.zeros((P, len(times), S))
    for ti, i in enumerate(times):
        for s in range(S):
            xa = np.sort(a[i, s].reshape(P, n_ref), axis=1)
            xb = np.sort(b[i, s].reshape(P, n_ref), axis=1)
            m = xa.mean(axis=1)
            w1 = np.abs(xa - xb).mean(axis=1)  # equal sizes: W1 of the two empirical laws
            with np.errstate(divide="ignore", invalid="ignore"):
                out[:, ti, s] = np.where(m > 0, w1 / np.where(m > 0, m, 1.0),
                                         np.where(w1 == 0, 0.0, np.inf))
    return out


def ssa_ms_per_point(stats: _native.Stats, n_points: int) -> float:
    """Device time of one parameter point's SSA (all its sample times) in a
    sweep launch."""
    return stats.kernel_ms / max(n_points, 1)


def route(trust: TrustMap, rates: np.ndarray, x0: np.ndarray, sample_idx: np.ndarray) -> np.ndarray:
    """Which (point, sample time) queries the surrogate answers: those whose
    every species is trusted there; the SSA answers the others.  bool [Q]."""
    ok = np.ones(len(rates), dtype=bool)
    for s in range(S):
        ok &= trust.trusted(rates, x0, sample_idx, s)
    return ok


def answer(mlp: MLP, trust: TrustMap, rates: np.ndarray, x0: np.ndarray, sample_idx: np.ndarray,
           n_ref: int = 10_000, seed: int = 0, traj_begin: int = 1 << 42):
    """A sweep query answered the gated way: mixture parameters from the
    surrogate where the trust map admits the query, and for every other query
    the n_ref-trajectory SSA histogram of its point (one sweep launch over
    those points).  Returns (surrogate mask [Q], params [Q][S][3][K] (valid
    where the mask is set), bounds [Q][S], {query index: samples [S][n_ref]})."""
    use = route(trust, rates, x0, sample_idx)
    params, bounds = predict(mlp, rates, x0, sample_idx)
    ssa = {}
    rest = np.flatnonzero(~use)
    if len(rest):
        data = generate(rates[rest], x0[rest], n_ref, seed=seed, traj_begin=traj_begin)
        samp = data.samples.cpu().numpy()
        for k, q in enumerate(rest):
            ssa[int(q)] = samp[int(sample_idx[q]), :, k * n_ref:(k + 1) * n_ref]
    return use, params, bounds, ssa


def benchmark(n_train: int = 100_000, steps: int = 12_000, batch: int = 8192, n_queries: int = 1_000_000,
              n_valid: int = 1024, n_holdout: int = 512, n_ref: int = 10_000, seed: int = 2026,
              lr: float = 1e-3, bins: int = 3, min_points: int = 4, trust_tol: float = 0.015,
              log=print) -> dict:
    """End-to-end: SSA training data -> training -> W1 validation at every
    sample time -> trust map -> check of the map on held-out points -> query
    timing and the gated sweep's effective speedup.

    A query is the distribution of every species at one (parameter point,
    sample time).  The plain GPU SSA's cost per query is that of n_ref
    trajectories of the point shared by its 32 sample times; the gated
    surrogate answers the queries the trust map admits and sends the rest
    to the SSA at that same cost per query, so its effective speedup is
    SSA time of all queries / (surrogate time + SSA time of the rest).  The
    same is reported for queries of one species' distribution (the map
    admits far more single species than whole queries).  The map admits a
    (cell, time, species) at trust_tol, below the 2 % criterion, as a margin
    for the points it has not seen.
    """
    import time

    import torch

    t0 = time.perf_counter()
    rates, x0 = sample_points(n_train, seed)
    data = generate(rates, x0, J, seed=seed)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    log(f"training data: {n_train} points x {J} trajectories, {data.stats.events} events, {gen_s:.2f} s")
    mlp = MLP(seed=seed)
    t0 = time.perf_counter()
    hist = train(mlp, data, steps=steps, batch=batch, lr=lr, seed=seed, log_every=max(1, steps // 10),
                 log=log)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    log(f"training: {steps} steps of {batch}, {train_s:.2f} s, NLL/sample {hist[0][1]:.3f} -> {hist[-1][1]:.3f}")
    del data
    torch.cuda.empty_cache()
    # validation against fresh n_ref-trajectory histograms at all 32 times
    vr, vx = sample_points(n_valid, seed + 1)
    t0 = time.perf_counter()
    rel, trust, vstats = validate(mlp, vr, vx, n_ref=n_ref, seed=seed + 2, bins=bins,
                                  min_points=min_points, trust_tol=trust_tol)
    val_s = time.perf_counter() - t0
    ssa_point_ms = ssa_ms_per_point(vstats, n_valid)
    ssa_qps = NB / (ssa_point_ms / 1e3)
    pass_frac = float((rel <= 0.02).mean())
    point_pass = float((rel <= 0.02).all(axis=(1, 2)).mean())
    floor = ssa_noise_floor(vr, vx, n_ref, seed + 2)
    repro = floor <= 0.02  # entries where two SSA runs agree to 2 % of the mean
    pass_repro = float((rel <= 0.02)[repro].mean()) if repro.any() else float("nan")
    log(f"validation: {n_valid} points x {n_ref} trajectories x {NB} times; W1 <= 2% of mean for "
        f"{pass_frac:.3f} of (point, time, species), {point_pass:.3f} of points; two SSA runs differ "
        f"by more than 2% in {1 - repro.mean():.3f} of the entries, the surrogate passes "
        f"{pass_repro:.3f} of the others; {val_s:.1f} s")
    # the trust map on held-out points and a fresh SSA
    hr, hx = sample_points(n_holdout, seed + 5)
    rel_h, _, _ = validate(mlp, hr, hx, n_ref=n_ref, seed=seed + 6, bins=bins, traj_begin=1 << 41)
    adm = np.stack([np.stack([trust.trusted(hr, hx, t, s_) for s_ in range(S)], axis=1)
                    for t in SAMPLE_TIMES], axis=1)  # [P][T][S]
    admitted = float(adm.mean())
    admitted_pass = float((rel_h[adm] <= 0.02).mean()) if adm.any() else float("nan")
    q_adm = adm.all(axis=2)
    log(f"held-out: {n_holdout} points; the trust map admits {admitted:.3f} of (point, time, "
        f"species) and {q_adm.mean():.3f} of (point, time) queries; W1 <= 2% holds for "
        f"{admitted_pass:.3f} of the admitted entries")
    # query timing (device-resident inputs, like the SSA's)
    qr, qx = sample_points(n_queries, seed + 3)
    d_r = torch.as_tensor(qr, device="cuda")
    d_x = torch.as_tensor(qx, device="cuda")
    qs = np.random.default_rng(seed + 4).integers(1, NB + 1, n_queries).astype(np.int32)
    chunk = 1 << 20  # the 10^6 queries in one launch of the fused kernel
    pts = [torch.arange(a, min(n_queries, a + chunk), dtype=torch.int32, device="cuda") for a in range(0, n_queries, chunk)]
    sps = [torch.as_tensor(qs[a:a + chunk], device="cuda") for a in range(0, n_queries, chunk)]
    for p_, s_ in zip(pts, sps):  # warm
        infer_device(mlp, d_r, d_x, p_, s_, p_.numel())
    torch.cuda.synchronize()
    times = []
    for _ in range(5):  # the median of five timed passes
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        for p_, s_ in zip(pts, sps):
            infer_device(mlp, d_r, d_x, p_, s_, p_.numel())
        ev1.record()
        torch.cuda.synchronize()
        times.append(ev0.elapsed_time(ev1))
    sg_ms = float(np.median(times))
    sg_qps = n_queries / (sg_ms / 1e3)
    flops = 2.0 * n_queries * sum(mlp.dims[l] * mlp.dims[l + 1] for l in range(mlp.depth))
    # the gated sweep: routing of the same queries, SSA for the rest
    t0 = time.perf_counter()
    use = route(trust, qr, qx, qs)
    route_s = time.perf_counter() - t0
    q_trusted = float(use.mean())
    ssa_all_ms = n_queries * ssa_point_ms / NB
    ssa_rest_ms = float((~use).sum()) * ssa_point_ms / NB
    effective = ssa_all_ms / (sg_ms + ssa_rest_ms)
    # one species per query: the same queries, one SSA quarter-cost each
    sp = np.random.default_rng(seed + 8).integers(0, S, n_queries)
    use_s = trust.trusted(qr, qx, qs, sp)
    e_trusted = float(use_s.mean())
    effective_s = ssa_all_ms / (sg_ms + float((~use_s).sum()) * ssa_point_ms / NB)
    log(f"surrogate: {n_queries} queries in {sg_ms:.2f} ms ({sg_qps:.3e} q/s, {flops / sg_ms / 1e9:.1f} TFLOP/s); "
        f"SSA: {ssa_qps:.3e} q/s; raw speedup {sg_qps / ssa_qps:.1f}x; the trust map admits "
        f"{q_trusted:.3f} of the queries, effective speedup of the gated sweep {effective:.2f}x; "
        f"one species per query: {e_trusted:.3f} admitted, {effective_s:.2f}x")
    return dict(surrogate_qps=sg_qps, ssa_qps=ssa_qps, speedup=sg_qps / ssa_qps, surrogate_ms=sg_ms,
                ssa_ms=vstats.kernel_ms, ssa_point_ms=ssa_point_ms, w1_pass_fraction=pass_frac,
                point_pass_fraction=point_pass, w1_pass_fraction_reproducible=pass_repro,
                ssa_self_w1_over_2pct=float(1 - repro.mean()),
                trusted_cells=float(trust.passed.mean()),
                trusted_validated_cells=float(trust.passed[trust.seen >= min_points].mean())
                if (trust.seen >= min_points).any() else 0.0,
                holdout_admitted_entries=admitted, holdout_admitted_queries=float(q_adm.mean()),
                holdout_admitted_pass_fraction=admitted_pass, trusted_query_fraction=q_trusted,
                effective_speedup=effective, trusted_species_query_fraction=e_trusted,
                effective_speedup_species=effective_s, trust_bins=bins, trust_tol=trust_tol,
                trust_min_points=min_points, n_valid=n_valid, route_s=route_s, train_s=train_s, gen_s=gen_s,
                val_s=val_s, nll_first=hist[0][1], nll_last=hist[-1][1],
                gemm_tflops=flops / sg_ms / 1e9, train_events=int(vstats.events),
                rel_w1_median=float(np.median(rel)))
