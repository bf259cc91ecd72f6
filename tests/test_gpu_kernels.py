"""GPU parity across kernel variants and the reference's edge cases.

Every variant (data-driven plain/persistent, network-specialised
plain/persistent) must give the oracle's counts bit for bit; the reference's
test invariants are checked through the Python API; full-size runs are
checked through size-independent properties.
"""

from __future__ import annotations

import numpy as np
import pytest

from rebop_amd import Gillespie, GillespieNet, Rate, _native
from tests.networks import build_device, build_oracle, dimers, michaelis_menten, sir, synthetic

pytestmark = pytest.mark.gpu

K = _native
VARIANTS = [
    K.REBOP_KERNEL_PLAIN,
    K.REBOP_KERNEL_PERSISTENT,
    K.REBOP_KERNEL_PLAIN | K.REBOP_KERNEL_SPECIALIZED,
    K.REBOP_KERNEL_PERSISTENT | K.REBOP_KERNEL_SPECIALIZED,
    K.REBOP_KERNEL_PLAIN | K.REBOP_KERNEL_SPECIALIZED | K.REBOP_KERNEL_GROUP2,
    K.REBOP_KERNEL_PLAIN | K.REBOP_KERNEL_SPECIALIZED | K.REBOP_KERNEL_GROUP4,
    K.REBOP_KERNEL_LDS,
]
# the data-driven register kernels hold at most 16 species and 32 reactions
REG_MAX_S, REG_MAX_R = 16, 32


def _device_run(g, tmax, nb, n, kernel, traj_begin=0, x0=None):
    import torch

    S = g.nb_species()
    out = torch.zeros((nb + 1, S, n), dtype=torch.int64, device="cuda")
    x0_ptr = None
    if x0 is not None:
        x0_t = torch.as_tensor(np.ascontiguousarray(x0, dtype=np.int64), device="cuda")
        x0_ptr = x0_t.data_ptr()
    st = g.run_device(tmax, nb, n, out.data_ptr(), traj_begin=traj_begin, x0_ptr=x0_ptr,
                      kernel=kernel)
    torch.cuda.synchronize()
    return out.cpu().numpy(), st


@pytest.mark.parametrize("kernel", VARIANTS)
@pytest.mark.parametrize(
    "name,spec,tmax,nb,n",
    [
        ("sir", sir(), 250.0, 250, 2048),
        ("dimers", dimers(), 1.0, 100, 1024),
        ("mm", michaelis_menten(), 50.0, 32, 2048),
        ("syn16x40", synthetic(16, 40, 3, init_mean=20.0), 5.0, 20, 512),
    ],
)
def test_variants_bit_exact(name, spec, tmax, nb, n, kernel, nthreads):
    if kernel in (K.REBOP_KERNEL_PLAIN, K.REBOP_KERNEL_PERSISTENT) and (
        spec[0] > REG_MAX_S or len(spec[1]) > REG_MAX_R
    ):
        pytest.skip("network beyond the data-driven register kernels (runs on LDS)")
    seed = 1234
    g = build_device(spec, seed)
    out, st = _device_run(g, tmax, nb, n, kernel, traj_begin=77)
    assert st.kernel & ~K.REBOP_KERNEL_COMPACT == kernel, (st.kernel, kernel)
    o = build_oracle(spec)
    _, oout, _, oev, odr = o.run_grid(seed, spec[2], tmax, nb, n_traj=n, traj_begin=77, nthreads=nthreads)
    assert np.array_equal(out, oout), np.argwhere(out != oout)[:5].tolist()
    assert st.events == oev and st.draws == odr and st.trajectories == n


def test_sparse_forms_bit_exact(nthreads):
    """A sparse Gillespie and explicit LMASparse rates (given order, duplicates)
    evaluate exactly as the oracle's Rate::rate does."""
    x0 = [30, 20, 5]
    rx = [
        ("dense", 0.01, [1, 1, 0], [-1, -1, 1]),
        ("pairs", 0.002, [(1, 1), (0, 1)], [1, 1, -1]),  # species 1 then 0 (not ascending)
        ("pairs", 0.0005, [(2, 1), (2, 1)], [0, 0, -1]),  # duplicate: k * n * n
        ("dense", 0.3, [0, 0, 1], [1, 0, -1]),
        ("dense", 1.5, [0, 0, 0], [0, 1, 0]),
    ]
    for sparse_net in (False, True):
        g = GillespieNet(x0, sparse=sparse_net, seed=5)
        o = build_oracle((3, [], x0), sparse=sparse_net)
        for kind, k, orders, d in rx:
            if kind == "dense":
                g.add_reaction(Rate.lma(k, orders), d)
                o.add_lma(k, orders, d)
            else:
                g.add_reaction(Rate.lma_sparse(k, orders), d)
                o.add_lma_sparse(k, orders, d)
        for kernel in VARIANTS:
            out, st = _device_run(g, 5.0, 25, 1024, kernel)
            _, oout, _, oev, _ = o.run_grid(5, x0, 5.0, 25, n_traj=1024, nthreads=nthreads)
            assert np.array_equal(out, oout), (sparse_net, kernel)
            assert st.events == oev


def test_dense_equals_sparse_net():
    """tests/test_rebop.py:55-65: the sparse flag does not change results."""
    a = build_device(sir(), 42, sparse=False)
    b = build_device(sir(), 42, sparse=True)
    _, oa = a.run(250.0, 250, n_traj=512)
    _, ob = b.run(250.0, 250, n_traj=512)
    assert np.array_equal(oa, ob)


def test_batch_advance_until_matches_grid():
    """advance_until(t_i) on a device batch, one call per sample, reproduces the grid run."""
    spec = dimers()
    g = build_device(spec, 9)
    times, out = g.run(1.0, 20, n_traj=300, traj_begin=11)
    b = g.batch(300, traj_begin=11)
    for i, t in enumerate(times):
        b.advance_until(float(t))
        assert np.array_equal(b.get_species().T, out[i]), i
        assert (b.get_time() == t).all()


def test_batch_state_roundtrip_and_backwards_time(nthreads):
    """set_species/set_time, and advance_until to an earlier time (the reference
    draws once and sets t = tmax)."""
    spec = sir()
    g = build_device(spec, 3)
    o = build_oracle(spec)
    b = g.batch(64)
    b.advance_until(40.0)
    b.advance_until(10.0)
    assert (b.get_time() == 10.0).all()
    for j in range(0, 64, 7):
        tr = o.state(spec[2], 3, traj=j)
        tr.advance_until(40.0)
        tr.advance_until(10.0)
        assert b.get_species()[j].tolist() == tr.x.tolist()
    b.set_species(np.tile([500, 10, 490], (64, 1)))
    b.set_time(0.0)
    b.advance_until(5.0)
    x = b.get_species()
    assert (x.sum(axis=1) == 1000).all()


def test_per_trajectory_x0(nthreads):
    spec = michaelis_menten()
    rng = np.random.default_rng(0)
    n = 700
    x0 = np.stack([rng.integers(10, 200, n), rng.integers(50, 1000, n), np.zeros(n), np.zeros(n)], axis=1)
    g = build_device(spec, 21)
    o = build_oracle(spec)
    for kernel in (K.REBOP_KERNEL_PERSISTENT, K.REBOP_KERNEL_PERSISTENT | K.REBOP_KERNEL_SPECIALIZED):
        out, st = _device_run(g, 50.0, 32, n, kernel, x0=x0)
        _, oout, _, oev, _ = o.run_grid(21, x0, 50.0, 32, n_traj=n, x0_per_traj=True, nthreads=nthreads)
        assert np.array_equal(out, oout)
    # the host-buffer path too
    _, hout = g.run(50.0, 32, n_traj=n, x0=x0)
    assert np.array_equal(hout, oout)


def test_absorbing_and_empty_networks():
    # NaN rate: absorbing from the start, no draws (gillespie_macro.rs:224-238)
    g = GillespieNet([3, 0], seed=1)
    g.add_reaction(Rate.lma(float("nan"), [1, 0]), [-1, 1])
    times, out, st = g.run(10.0, 10, n_traj=100, with_stats=True)
    assert (out[:, 0] == 3).all() and (out[:, 1] == 0).all() and st.draws == 0
    # zero propensity (reactant absent)
    g = GillespieNet([0, 5], seed=1)
    g.add_reaction(Rate.lma(2.0, [1, 0]), [-1, 1])
    times, out, st = g.run(10.0, 10, n_traj=100, with_stats=True)
    assert (out[:, 1] == 5).all() and st.events == 0
    # no reactions at all (tests/test_rebop.py:148-152)
    g = GillespieNet([4, 2], seed=1)
    times, out, st = g.run(10.0, 10, n_traj=10, with_stats=True)
    assert np.array_equal(times, np.linspace(0, 10, 11)) and (out[:, 0] == 4).all()
    # pure decay reaches 0 and absorbs mid-run
    g = GillespieNet([3], seed=1)
    g.add_reaction(Rate.lma(5.0, [1]), [-1])
    times, out, st = g.run(100.0, 10, n_traj=100, with_stats=True)
    assert (out[-1] == 0).all() and st.events == 300


def test_huge_counts_run_on_int64_counts():
    """2^53 + 1 is beyond the f64 counts of the register and specialised
    kernels: the run takes the LDS kernel's int64 counts (about 9e15 events
    per unit time, so a horizon of 2e-15)."""
    g = GillespieNet([2**53 + 1, 0], seed=1)
    g.add_reaction(Rate.lma(1.0, [1, 0]), [-1, 1])
    _, out, st = g.run(2e-15, 2, n_traj=4, with_stats=True)
    assert st.kernel == _native.REBOP_KERNEL_LDS | _native.REBOP_KERNEL_WIDE
    o = build_oracle((2, [(1.0, [1, 0], [-1, 1])], [2**53 + 1, 0]))
    _, oout, _, _, _ = o.run_grid(1, [2**53 + 1, 0], 2e-15, 2, n_traj=4)
    assert np.array_equal(out, oout) and out[-1, 1].min() > 0


def test_python_api_sir_invariants():
    """tests/test_rebop.py:15-27 through the Python API, on the device."""
    for seed in [None, *range(10)]:
        s = Gillespie()
        s.add_reaction(1e-4, ["S", "I"], ["I", "I"])
        s.add_reaction(0.01, ["I"], ["R"])
        ds = s.run({"S": 999, "I": 1}, tmax=250, nb_steps=250, rng=seed)
        assert np.array_equal(ds.time, np.arange(251))
        S, I, R = np.asarray(ds.S), np.asarray(ds.I), np.asarray(ds.R)
        assert (S >= 0).all() and (I >= 0).all() and (R >= 0).all()
        assert (S <= 999).all() and (I <= 1000).all() and (R <= 1000).all()
        assert ((S + I + R) == 1000).all()


def test_python_api_dense_vs_sparse_and_var_names():
    """tests/test_rebop.py:55-89 through the Python API."""
    s = Gillespie()
    s.add_reaction(1e-4, ["S", "I"], ["I", "I"])
    s.add_reaction(0.01, ["I"], ["R"])
    init = {"S": 999, "I": 1}
    a = s.run(init, tmax=250, nb_steps=250, rng=42, sparse=None)
    b = s.run(init, tmax=250, nb_steps=250, rng=42, sparse=False)
    c = s.run(init, tmax=250, nb_steps=250, rng=42, sparse=True)
    for v in "SIR":
        assert np.array_equal(np.asarray(a[v]), np.asarray(b[v]))
        assert np.array_equal(np.asarray(a[v]), np.asarray(c[v]))
    sub = s.run(init, tmax=250, nb_steps=250, rng=0, var_names=["S", "I"])
    full = s.run(init, tmax=250, nb_steps=250, rng=0)
    assert "S" in sub and "I" in sub and "R" not in sub
    assert np.array_equal(np.asarray(sub["S"]), np.asarray(full["S"]))


def test_python_api_batched_trajectories(nthreads):
    from oracle.pyoracle import OracleNet

    s = Gillespie()
    s.add_reaction(1e-4, ["S", "I"], ["I", "I"])
    s.add_reaction(0.01, ["I"], ["R"])
    ds = s.run({"S": 999, "I": 1}, tmax=250, nb_steps=250, rng=7, n_trajectories=3000)
    S = np.asarray(ds.S)
    assert S.shape == (251, 3000)
    seed = int(np.random.default_rng(7).integers(np.iinfo(np.uint64).max, dtype=np.uint64))
    o = OracleNet(3)
    o.add_lma(1e-4, [1, 1, 0], [-1, 1, 0])
    o.add_lma(0.01, [0, 1, 0], [0, -1, 1])
    _, oout, _, _, _ = o.run_grid(seed, [999, 1, 0], 250.0, 250, n_traj=3000, nthreads=nthreads)
    assert np.array_equal(S, oout[:, 0, :])


def test_dimers_full_size_invariants():
    """10^6 dimer trajectories (the benchmark workload): the gene count is 1 at
    every sample, counts are non-negative, the event/draw accounting closes,
    and the trajectory-id split is consistent (first 4096 == a separate run)."""
    import torch

    spec = dimers()
    g = build_device(spec, 42)
    n = 1_000_000
    out = torch.empty((101, 4, n), dtype=torch.int64, device="cuda")
    st = g.run_device(1.0, 100, n, out.data_ptr())
    assert st.trajectories == n
    assert bool((out[:, 0, :] == 1).all())
    assert bool((out >= 0).all())
    # every trajectory ends with an overshoot at t = 1 (never absorbed: gene == 1)
    assert st.draws - st.events == n * 101
    head, _ = _device_run(g, 1.0, 100, 4096, K.REBOP_KERNEL_PLAIN)
    assert np.array_equal(head, out[:, :, :4096].cpu().numpy())
    # the mean dimer count at t=1 sits in the range the reference's test brackets
    d = out[-1, 3, :].double().mean().item()
    assert 0 < d < 10000


def test_auto_groups_lanes_only_for_small_runs():
    """AUTO gives 4 (2) lanes per trajectory while that leaves at most one wave
    per SIMD, else one lane per trajectory with refill."""
    _, cus = K.device_info()
    lanes = cus * 4 * 64
    g = build_device(sir(), 5)
    for n, want in [(1000, K.REBOP_KERNEL_GROUP4), (lanes // 4, K.REBOP_KERNEL_GROUP4),
                    (lanes // 4 + 64, K.REBOP_KERNEL_GROUP2), (lanes // 2 + 64, 0)]:
        out, st = _device_run(g, 10.0, 4, n, K.REBOP_KERNEL_AUTO)
        assert st.kernel & (K.REBOP_KERNEL_GROUP2 | K.REBOP_KERNEL_GROUP4) == want, (n, st.kernel)
        if want == 0:
            assert st.kernel == K.REBOP_KERNEL_PERSISTENT | K.REBOP_KERNEL_SPECIALIZED


def test_grouped_lanes_refuse_persistent_and_generic():
    import torch

    g = build_device(sir(), 5)
    out = torch.zeros((5, 3, 64), dtype=torch.int64, device="cuda")
    for kern in (K.REBOP_KERNEL_PERSISTENT | K.REBOP_KERNEL_SPECIALIZED | K.REBOP_KERNEL_GROUP4,
                 K.REBOP_KERNEL_PLAIN | K.REBOP_KERNEL_GROUP2):
        with pytest.raises(K.RebopError) as e:
            g.run_device(10.0, 4, 64, out.data_ptr(), kernel=kern)
        assert e.value.code == K.REBOP_ERR_UNSUPPORTED


def test_output_only_8_byte_aligned(nthreads):
    """The sample conversion pass reads pairs of int64 when the output is
    16-byte aligned; an output at an odd int64 offset takes the one-by-one
    path and gives the same samples."""
    import torch

    spec = dimers()
    g = build_device(spec, 14)
    n, nb, tmax = 999, 20, 0.3
    total = (nb + 1) * 4 * n
    base = torch.zeros(total + 1, dtype=torch.int64, device="cuda")
    assert base.data_ptr() % 16 == 0
    st = g.run_device(tmax, nb, n, base.data_ptr() + 8)  # 8-byte aligned only
    torch.cuda.synchronize()
    got = base[1:].cpu().numpy().reshape(nb + 1, 4, n)
    assert int(base[0]) == 0  # nothing written before the output
    o = build_oracle(spec)
    _, oout, _, oev, _ = o.run_grid(14, spec[2], tmax, nb, n_traj=n, nthreads=nthreads)
    assert np.array_equal(got, oout) and st.events == oev
