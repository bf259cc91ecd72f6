"""The surrogate's device kernels against the FP64 CPU reference
(oracle/surrogate_ref.py), a training run that must learn, and the W1 gate."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sg():
    from rebop_amd import surrogate

    surrogate.load()
    return surrogate


def _bf(x):
    import torch

    return torch.as_tensor(x, dtype=torch.float32, device="cuda").to(torch.bfloat16)


@pytest.mark.parametrize("M,N,Kd", [(128, 192, 96), (64, 128, 2048), (4096, 2048, 64)])
@pytest.mark.parametrize("ta,tb", [(False, False), (False, True), (True, False), (True, True)])
def test_gemm_bf16_matches_fp64(sg, ta, tb, M, N, Kd):
    """op(A) op(B) on the matrix cores equals the FP64 product of the same
    BF16 values to FP32-accumulation accuracy (the lane layout of
    v_mfma_f32_16x16x32_bf16 is checked here; the second shape has too few
    output tiles and runs split over K, the third runs on 128 x 128 tiles)."""
    import torch

    from oracle.surrogate_ref import bf16

    rng = np.random.default_rng(0)
    A = bf16(rng.normal(size=(M, Kd)))
    B = bf16(rng.normal(size=(Kd, N)))
    dA = _bf(A.T.copy() if ta else A)
    dB = _bf(B.T.copy() if tb else B)
    C = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
    sg.gemm(dA, dB, C, ta=ta, tb=tb)
    torch.cuda.synchronize()
    ref = A @ B
    got = C.double().cpu().numpy()
    assert np.isfinite(got).all()
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))


@pytest.mark.parametrize("M,N,Kd", [(128, 192, 64), (4096, 512, 512), (8192, 128, 512)])
@pytest.mark.parametrize("relu", [True, False])
def test_fused_gemm_epilogue_equals_gemm_then_bias_act(sg, M, N, Kd, relu):
    """The fused launch writes the same bits as gemm followed by bias_act
    (BF16 and FP32 outputs; 64 x 64 and 128 x 128 tiles) where gemm does not
    split K; the third shape has too few output tiles, gemm splits K there
    and adds partial sums atomically, so only FP32-rounding agreement holds."""
    import torch

    rng = np.random.default_rng(5)
    A = _bf(rng.normal(size=(M, Kd)))
    B = _bf(rng.normal(size=(Kd, N)) / np.sqrt(Kd))
    bias = torch.as_tensor(rng.normal(size=N), dtype=torch.float32, device="cuda")
    C = torch.empty((M, N), dtype=torch.float32, device="cuda")
    H0 = torch.empty((M, N), dtype=torch.bfloat16, device="cuda")
    F0 = torch.empty((M, N), dtype=torch.float32, device="cuda")
    sg.gemm(A, B, C)
    sg.bias_act(C, bias, relu, H=H0, F=F0)
    H1 = torch.full((M, N), float("nan"), dtype=torch.bfloat16, device="cuda")
    F1 = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
    sg.gemm_bias_act(A, B, bias, relu, H=H1, F=F1)
    torch.cuda.synchronize()
    F2 = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
    sg.gemm_bias_act(A, B, bias, relu, F=F2)  # one output only
    H3 = torch.full((M, N), float("nan"), dtype=torch.bfloat16, device="cuda")
    F3 = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
    sg.gemm_bias_act(A, B.t().contiguous(), bias, relu, H=H3, F=F3, tb=True)  # B stored [N][K]
    torch.cuda.synchronize()
    assert torch.equal(F1.view(torch.int32), F2.view(torch.int32))
    assert torch.equal(F1.view(torch.int32), F3.view(torch.int32))
    assert torch.equal(H1.view(torch.int16), H3.view(torch.int16))
    split = (M // 64) * (N // 64) < 512 and not (M % 128 == 0 and N % 128 == 0 and (M // 128) * (N // 128) >= 512)
    if not split:
        assert torch.equal(H0.view(torch.int16), H1.view(torch.int16))
        assert torch.equal(F0.view(torch.int32), F1.view(torch.int32))
    else:
        scale = float(F0.abs().max())
        assert float((F0 - F1).abs().max()) <= 1e-5 * scale
        assert float((H0.float() - H1.float()).abs().max()) <= 2.0 ** -7 * scale


def test_transposed_bf16_copy(sg):
    import torch

    x = torch.randn(192, 320, dtype=torch.float32, device="cuda")
    y = torch.empty(320, 192, dtype=torch.bfloat16, device="cuda")
    sg.transpose_bf16(x, y)
    torch.cuda.synchronize()
    assert torch.equal(y.view(torch.int16), x.t().contiguous().to(torch.bfloat16).view(torch.int16))


def _stack(sg, rng, dims):
    """Random layers: (Wt [N][K] bf16 for gemm_bias_act, the packed copy for
    the fused pass, bias)."""
    import torch

    W = [torch.as_tensor(rng.normal(size=(dims[l], dims[l + 1])) * np.sqrt(2.0 / dims[l]), dtype=torch.float32,
                         device="cuda") for l in range(len(dims) - 1)]
    Wt = [w.t().contiguous().to(torch.bfloat16) for w in W]
    Wp = [sg.packed(w) for w in W]
    b = [torch.as_tensor(rng.normal(size=dims[l + 1]) * 0.1, dtype=torch.float32, device="cuda")
         for l in range(len(dims) - 1)]
    return Wt, Wp, b


def _layer_by_layer(sg, X, Wt, b):
    import torch

    H = X
    for l in range(len(Wt)):
        N = Wt[l].shape[0]
        if l < len(Wt) - 1:
            h = torch.empty((X.shape[0], N), dtype=torch.bfloat16, device="cuda")
            sg.gemm_bias_act(H, Wt[l], b[l], True, H=h, tb=True)
            H = h
        else:
            F = torch.empty((X.shape[0], N), dtype=torch.float32, device="cuda")
            sg.gemm_bias_act(H, Wt[l], b[l], False, F=F, tb=True)
    return F


def test_packed_weight_layout(sg):
    import torch

    W = torch.randn(96, 256, dtype=torch.float32, device="cuda")
    Wp = sg.packed(W)
    torch.cuda.synchronize()
    ref = W.to(torch.bfloat16).reshape(3, 32, 256).permute(0, 2, 1).contiguous()
    assert torch.equal(Wp.view(torch.int16), ref.view(torch.int16))


@pytest.mark.parametrize("dims", [[64, 512, 512, 512, 512, 128], [128, 256, 128, 512, 128, 256], [64, 128]])
@pytest.mark.parametrize("M", [1088, 4096])
def test_fused_forward_equals_layer_by_layer(sg, dims, M):
    """One launch through the whole stack, activations kept in LDS, writes
    the bits of gemm_bias_act applied layer by layer (512 x 512 layers run
    the fully unrolled k loop, the others the generic one; 1088 rows end in a
    partial block of 128)."""
    import torch

    rng = np.random.default_rng(len(dims) * 7 + M)
    Wt, Wp, b = _stack(sg, rng, dims)
    X = _bf(rng.normal(size=(M, dims[0])))
    ref = _layer_by_layer(sg, X, Wt, b)
    out = torch.full((M, dims[-1]), float("nan"), dtype=torch.float32, device="cuda")
    sg.mlp_forward(X, Wp, b, out)
    torch.cuda.synchronize()
    assert torch.isfinite(out).all()
    assert torch.equal(out.view(torch.int32), ref.view(torch.int32))


def test_fused_forward_limits(sg):
    """The largest stack the kernel takes: eight layers, a 512-wide input
    (every LDS fragment block of a row in use) and all three output widths."""
    import torch

    dims = [512, 512, 256, 512, 128, 512, 512, 256, 128]
    rng = np.random.default_rng(29)
    Wt, Wp, b = _stack(sg, rng, dims)
    X = _bf(rng.normal(size=(640, dims[0])))
    ref = _layer_by_layer(sg, X, Wt, b)
    out = torch.full((640, dims[-1]), float("nan"), dtype=torch.float32, device="cuda")
    sg.mlp_forward(X, Wp, b, out)
    torch.cuda.synchronize()
    assert torch.equal(out.view(torch.int32), ref.view(torch.int32))


def test_fused_inference_wider_input(sg):
    """rebop_sg_infer with a 128-wide first layer: the features fill columns
    0..15 and the kernel zeroes 16..127, as features_kernel does."""
    import torch

    rates, x0 = sg.sample_points(40, 13)
    d_r, d_x = torch.as_tensor(rates, device="cuda"), torch.as_tensor(x0, device="cuda")
    rng = np.random.default_rng(14)
    B = 704
    pts = torch.as_tensor(rng.integers(0, 40, B).astype(np.int32), device="cuda")
    smp = torch.as_tensor(rng.integers(1, sg.NB + 1, B).astype(np.int32), device="cuda")
    dims = [128, 256, sg.OUT]
    Wt, Wp, b = _stack(sg, rng, dims)
    head = torch.full((B, sg.OUT), float("nan"), dtype=torch.float32, device="cuda")
    sg.infer(Wp, b, pts, smp, d_r, d_x, head)
    X = torch.empty((B, dims[0]), dtype=torch.bfloat16, device="cuda")
    off = torch.empty((B, sg.S), dtype=torch.float32, device="cuda")
    sg.features(pts, smp, d_r, d_x, X, off)
    ref = _layer_by_layer(sg, X, Wt, b)
    sg.head_offsets(ref, off)
    torch.cuda.synchronize()
    assert torch.equal(head.view(torch.int32), ref.view(torch.int32))


def test_fused_forward_ragged_rows(sg):
    """A row count that is not a multiple of 64: rows past M are neither read
    nor written."""
    import torch

    dims = [64, 512, 256, 128]
    rng = np.random.default_rng(17)
    Wt, Wp, b = _stack(sg, rng, dims)
    X = _bf(rng.normal(size=(1024, dims[0])))
    ref = _layer_by_layer(sg, X, Wt, b)
    out = torch.full((1024, dims[-1]), float("nan"), dtype=torch.float32, device="cuda")
    sg.mlp_forward(X[:1000], Wp, b, out[:1000])
    torch.cuda.synchronize()
    assert torch.equal(out[:1000].view(torch.int32), ref[:1000].view(torch.int32))
    assert torch.isnan(out[1000:]).all()


def test_fused_forward_rejects_unsupported_shapes(sg):
    import torch

    from rebop_amd._native import RebopError

    rng = np.random.default_rng(2)
    X = _bf(rng.normal(size=(128, 64)))
    for dims in ([64, 96], [64, 1024], [32, 128], [64, 64]):
        _, Wp, b = _stack(sg, rng, dims)
        out = torch.empty((128, dims[-1]), dtype=torch.float32, device="cuda")
        with pytest.raises(RebopError):
            sg.mlp_forward(X[:, : dims[0]].contiguous(), Wp, b, out)


def test_fused_inference_equals_three_steps(sg):
    """rebop_sg_infer (features, forward pass and head offsets in one launch)
    against the three separate launches: the same bits, for queries over many
    parameter points and sample times (5056 items: a partial last block)."""
    import torch

    rates, x0 = sg.sample_points(300, 11)
    d_r = torch.as_tensor(rates, device="cuda")
    d_x = torch.as_tensor(x0, device="cuda")
    rng = np.random.default_rng(12)
    B = 5056
    pts = torch.as_tensor(rng.integers(0, 300, B).astype(np.int32), device="cuda")
    smp = torch.as_tensor(rng.integers(1, sg.NB + 1, B).astype(np.int32), device="cuda")
    mlp = sg.MLP(seed=3)
    fused = sg.infer_device(mlp, d_r, d_x, pts, smp, B)
    steps = sg.infer_device_layers(mlp, d_r, d_x, pts, smp, B)
    torch.cuda.synchronize()
    assert torch.isfinite(fused).all()
    assert torch.equal(fused.view(torch.int32), steps.view(torch.int32))


def test_nll_kernel_matches_fp64_reference(sg):
    import torch

    from oracle import surrogate_ref as ref

    rng = np.random.default_rng(1)
    B, S, K, J, P = 64, 4, 8, 64, 64
    head = np.zeros((B, 128))
    for s in range(S):
        head[:, s * 3 * K:s * 3 * K + K] = rng.normal(size=(B, K))
        head[:, s * 3 * K + K:s * 3 * K + 2 * K] = rng.normal(np.log(60.0), 1.0, size=(B, K))
        head[:, s * 3 * K + 2 * K:(s + 1) * 3 * K] = rng.normal(1.0, 1.5, size=(B, K))
    head = head.astype(np.float32).astype(np.float64)
    nb = 32
    samples = rng.poisson(60, size=(nb + 1, S, P * J)).astype(np.int64)
    ip = rng.integers(0, P, B).astype(np.int32)
    isamp = rng.integers(1, nb + 1, B).astype(np.int32)
    bound = rng.integers(100, 300, size=(B, S)).astype(np.float32)  # > every sample (Poisson(60))
    head[:, [s * 3 * K + K + c for s in range(S) for c in range(K)]] -= np.repeat(np.log(bound), K, axis=1)
    head = head.astype(np.float32).astype(np.float64)
    d_head = torch.as_tensor(head, dtype=torch.float32, device="cuda")
    grad = torch.zeros((B, 128), dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    sg.nb_mixture_nll(d_head, torch.as_tensor(ip, device="cuda"), torch.as_tensor(isamp, device="cuda"),
                      torch.as_tensor(samples, device="cuda"), grad, loss, 1.0,
                      torch.as_tensor(bound, device="cuda"))
    torch.cuda.synchronize()
    obs = np.stack([samples[isamp[b], :, ip[b] * J:(ip[b] + 1) * J] for b in range(B)])
    total, g = ref.nll(head, obs, K, bound=bound.astype(np.float64))
    assert loss.item() == pytest.approx(total, rel=1e-10)
    got = grad.double().cpu().numpy()
    assert np.allclose(got, g, rtol=1e-5, atol=1e-5 * np.abs(g).max())


def test_training_step_gradients_match_reference(sg):
    """One forward + NLL + backward of the device MLP against the FP64
    reference with the same BF16 roundings (what remains is FP32 vs FP64
    accumulation), and the FP32-vs-FP64 error of the head stated."""
    import torch

    from oracle import surrogate_ref as ref

    rng = np.random.default_rng(2)
    mlp = sg.MLP(seed=3, width=128, depth=3)
    B, P, J, nb = 128, 32, 64, 32
    rates, x0 = sg.sample_points(P, seed=4)
    samples = rng.poisson(50, size=(nb + 1, 4, P * J)).astype(np.int64)
    # counts never exceed their species' bound (E0 for E and ES, S0 for S and P)
    bnd = np.repeat(x0[:, [0, 1, 0, 1]].T, J, axis=1)  # [4][P*J]
    samples = np.minimum(samples, bnd[None])
    ip = rng.integers(0, P, B).astype(np.int32)
    isamp = rng.integers(1, nb + 1, B).astype(np.int32)
    d_ip, d_is = torch.as_tensor(ip, device="cuda"), torch.as_tensor(isamp, device="cuda")
    X = torch.empty((B, sg.IN), dtype=torch.bfloat16, device="cuda")
    off = torch.empty((B, 4), dtype=torch.float32, device="cuda")
    sg.features(d_ip, d_is, torch.as_tensor(rates, device="cuda"), torch.as_tensor(x0, device="cuda"), X, off)
    head = mlp.forward(X)
    grad = torch.zeros((B, sg.OUT), dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    scale = 1.0 / (B * J)
    sg.nb_mixture_nll(head, d_ip, d_is, torch.as_tensor(samples, device="cuda"), grad, loss, scale, off)
    dW, db = mlp.backward(grad)
    torch.cuda.synchronize()

    Xr, offr = ref.features(rates[ip], x0[ip], isamp.astype(np.float64), sg.TMAX, sg.NB, sg.IN,
                            with_offsets=True)
    Xd = X.double().cpu().numpy()
    # device log10 and numpy's may differ in the last bit of a double, which
    # can move a BF16 rounding: allow one BF16 step on a few entries
    diff = np.abs(ref.bf16(Xr) - Xd)
    assert (diff <= np.abs(Xd) * 2.0**-7 + 1e-30).all() and (diff > 0).mean() < 1e-3
    assert np.array_equal(off.double().cpu().numpy(), offr)
    Xr = Xd  # the reference continues from the device's BF16 inputs
    Ws = [w.double().cpu().numpy() for w in mlp.W]
    bs = [b.double().cpu().numpy() for b in mlp.b]
    rhead, H = ref.forward(Xr, Ws, bs, emulate_bf16=True)
    dhead = head.double().cpu().numpy()
    err = np.abs(dhead - rhead).max() / np.abs(rhead).max()
    assert err < 1e-5, err
    # FP32 head vs an all-FP64 network without BF16 roundings: the stated error
    rhead64, _ = ref.forward(Xr, Ws, bs, emulate_bf16=False)
    err64 = np.abs(dhead - rhead64).max() / np.abs(rhead64).max()
    print(f"head error vs all-FP64 network: {err64:.3e} (vs BF16-emulating FP64: {err:.3e})")
    assert err64 < 5e-2
    obs = np.stack([samples[isamp[b], :, ip[b] * J:(ip[b] + 1) * J] for b in range(B)])
    total, g = ref.nll(dhead, obs, 8, bound=off.double().cpu().numpy())
    assert loss.item() == pytest.approx(total, rel=1e-9)
    rdW, rdb = ref.backward(g * scale, H, Ws, emulate_bf16=True)
    for l in range(mlp.depth):
        a = dW[l].double().cpu().numpy()
        assert np.abs(a - rdW[l]).max() <= 2e-3 * np.abs(rdW[l]).max() + 1e-12, l
        b = db[l].double().cpu().numpy()
        assert np.abs(b - rdb[l]).max() <= 2e-3 * np.abs(rdb[l]).max() + 1e-12, l


def test_adam_matches_reference(sg):
    import torch

    rng = np.random.default_rng(5)
    n = 1000
    w = rng.normal(size=n).astype(np.float32)
    g = rng.normal(size=n).astype(np.float32)
    m = rng.normal(size=n).astype(np.float32) * 0.1
    v = np.abs(rng.normal(size=n).astype(np.float32)) * 0.01
    dw, dg, dm, dv = (torch.as_tensor(a, device="cuda") for a in (w, g, m, v))
    wb = torch.empty(n, dtype=torch.bfloat16, device="cuda")
    sg.adam(dw, dg, dm, dv, wb, 1e-3, 0.9, 0.999, 1e-8, 3)
    torch.cuda.synchronize()
    m2 = 0.9 * m + 0.1 * g
    v2 = 0.999 * v + 0.001 * g * g
    w2 = w - 1e-3 * (m2 / (1 - 0.9**3)) / (np.sqrt(v2 / (1 - 0.999**3)) + 1e-8)
    assert np.allclose(dw.cpu().numpy(), w2, rtol=1e-5, atol=1e-7)
    assert np.array_equal(wb.float().cpu().numpy(),
                          torch.as_tensor(dw.cpu()).to(torch.bfloat16).float().numpy())


def test_training_learns_and_beats_untrained(sg):
    """A short training run lowers the NLL, and the trained network's W1 to
    fresh SSA histograms is far below the untrained network's: what the
    trained surrogate gains over the method without it."""
    P = 2048
    rates, x0 = sg.sample_points(P, seed=11)
    data = sg.generate(rates, x0, sg.J, seed=12)
    untrained = sg.MLP(seed=1)
    mlp = sg.MLP(seed=1)
    hist = sg.train(mlp, data, steps=300, batch=4096, lr=2e-3, log_every=50)
    assert hist[-1][1] < 0.7 * hist[0][1], hist
    vr, vx = sg.sample_points(12, seed=13)
    rel_t, _, _ = sg.validate(mlp, vr, vx, n_ref=4000, seed=14, times=(8, 32))
    rel_u, _, _ = sg.validate(untrained, vr, vx, n_ref=4000, seed=14, times=(8, 32))
    assert np.median(rel_t) < 0.5 * np.median(rel_u), (np.median(rel_t), np.median(rel_u))


def test_trust_map_holds_on_held_out_points(sg):
    """The gate as the sweep uses it: validate a trained surrogate against
    fresh 10^4-trajectory SSA histograms at every sample time, build the trust
    map, then on held-out points and a fresh SSA require W1 <= 2 % of the
    species mean for >= 95 % of the (point, time, species) entries the map
    admits.  The map must admit a useful share; route() and answer() send the
    rest to the SSA."""
    P = 20000
    rates, x0 = sg.sample_points(P, seed=21)
    data = sg.generate(rates, x0, sg.J, seed=22)
    mlp = sg.MLP(seed=3)
    sg.train(mlp, data, steps=3000, batch=8192, lr=1e-3)
    del data
    vr, vx = sg.sample_points(1024, seed=23)
    rel, trust, _ = sg.validate(mlp, vr, vx, n_ref=10_000, seed=24, bins=3, min_points=4,
                                trust_tol=0.015)
    hr, hx = sg.sample_points(256, seed=25)
    rel_h, _, _ = sg.validate(mlp, hr, hx, n_ref=10_000, seed=26, traj_begin=1 << 41)
    adm = np.stack([np.stack([trust.trusted(hr, hx, t, s) for s in range(sg.S)], axis=1)
                    for t in sg.SAMPLE_TIMES], axis=1)
    share = float(adm.mean())
    ok = float((rel_h[adm] <= 0.02).mean()) if adm.any() else float("nan")
    print(f"validated pass {float((rel <= 0.02).mean()):.3f}; held-out: admitted {share:.3f} of "
          f"entries, W1 <= 2% on {ok:.4f} of them (vs {float((rel_h <= 0.02).mean()):.3f} overall)")
    assert share >= 0.05, share
    assert ok >= 0.95, ok
    # routing: the admitted queries go to the surrogate, the rest to the SSA
    samp = np.random.default_rng(27).integers(1, sg.NB + 1, 12)
    use, params, bounds, ssa = sg.answer(mlp, trust, hr[:12], hx[:12], samp, n_ref=2000, seed=28)
    assert use.tolist() == sg.route(trust, hr[:12], hx[:12], samp).tolist()
    assert sorted(ssa) == np.flatnonzero(~use).tolist()
    for q, xs in ssa.items():
        assert xs.shape == (sg.S, 2000)
        assert (xs[0] + xs[2] == hx[q, 0]).all()  # E + ES = E0 in every SSA sample
