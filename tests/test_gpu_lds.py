"""GPU parity of the LDS-resident kernel (large networks, heterogeneous fused
batches) against the CPU oracle, bit for bit."""

from __future__ import annotations

import numpy as np
import pytest

from rebop_amd import _native, workloads
from rebop_amd.engine import run_jobs
from tests.networks import build_device, build_oracle, dimers, michaelis_menten, sir, synthetic

pytestmark = pytest.mark.gpu
LDS = _native.REBOP_KERNEL_LDS
COMPACT = _native.REBOP_KERNEL_COMPACT


def _dev(g, tmax, nb, n, kernel, x0=None, traj_begin=0):
    import torch

    S = g.nb_species()
    out = torch.zeros((nb + 1, S, n), dtype=torch.int64, device="cuda")
    x0_ptr = None
    if x0 is not None:
        xt = torch.as_tensor(np.ascontiguousarray(x0, dtype=np.int64), device="cuda")
        x0_ptr = xt.data_ptr()
    st = g.run_device(tmax, nb, n, out.data_ptr(), kernel=kernel, x0_ptr=x0_ptr, traj_begin=traj_begin)
    torch.cuda.synchronize()
    return out.cpu().numpy(), st


@pytest.mark.parametrize(
    "name,spec,tmax,nb,n",
    [
        ("sir", sir(), 250.0, 250, 700),
        ("dimers", dimers(), 1.0, 100, 300),
        ("mm", michaelis_menten(), 50.0, 32, 1000),
        ("syn16x40", synthetic(16, 40, 3, init_mean=20.0), 5.0, 20, 300),
        ("syn64x256", synthetic(64, 256, 11, init_mean=30.0), 0.05, 10, 200),
    ],
)
@pytest.mark.parametrize("width", ["int16", "int32"])
def test_lds_kernel_bit_exact(name, spec, tmax, nb, n, width, nthreads, monkeypatch):
    """int16 counts (REBOP_LDS_COMPACT=1: whenever the counts fit, also
    where they would not run more waves) and int32 counts
    (REBOP_LDS_COMPACT=0) both equal the oracle."""
    monkeypatch.setenv("REBOP_LDS_COMPACT", "1" if width == "int16" else "0")
    g = build_device(spec, 77)
    out, st = _dev(g, tmax, nb, n, LDS, traj_begin=5)
    assert st.kernel == (LDS | COMPACT if width == "int16" else LDS)
    o = build_oracle(spec, sparse=True)  # same results as dense, faster
    _, oout, _, oev, odr = o.run_grid(77, spec[2], tmax, nb, n_traj=n, traj_begin=5, nthreads=nthreads)
    assert np.array_equal(out, oout), np.argwhere(out != oout)[:5].tolist()
    assert st.events == oev and st.draws == odr


def test_synthetic_128x512_auto_uses_lds(nthreads):
    """The 128-species, 512-reaction network: AUTO selects the LDS kernel; a
    short horizon is compared with the oracle bit for bit."""
    spec = workloads.synthetic_large()
    g = build_device(spec, 7)
    out, st = _dev(g, 0.05, 20, 64, _native.REBOP_KERNEL_AUTO)
    assert st.kernel == LDS | COMPACT
    o = build_oracle(spec, sparse=True)
    _, oout, _, oev, _ = o.run_grid(7, spec[2], 0.05, 20, n_traj=64, nthreads=nthreads)
    assert np.array_equal(out, oout)
    assert st.events == oev and st.events > 64 * 1000


def test_lds_linear_search_paths(nthreads):
    """Non-monotone cumulative sums: a negative rate constant, and a jump that
    consumes more than the rate's order (counts can go negative); both use the
    full linear count, which must match the oracle exactly."""
    x0 = [1, 3, 40]
    rx = [
        (5.0, [1, 0, 0], [-2, 1, 0]),   # consumes 2 A at order 1: A goes negative
        (-0.01, [0, 1, 0], [0, 0, 1]),  # negative constant: negative propensity
        (0.001, [0, 0, 1], [1, 0, -1]),
        (0.3, [0, 1, 1], [1, -1, 0]),
    ]
    spec = (3, rx, x0)
    g = build_device(spec, 3)
    for kernel in (LDS, _native.REBOP_KERNEL_PERSISTENT, _native.REBOP_KERNEL_PERSISTENT | _native.REBOP_KERNEL_SPECIALIZED):
        out, st = _dev(g, 4.0, 16, 400, kernel)
        o = build_oracle(spec)
        _, oout, _, oev, _ = o.run_grid(3, x0, 4.0, 16, n_traj=400, nthreads=nthreads)
        assert np.array_equal(out, oout), kernel
        assert st.events == oev
    assert (oout[:, 0, :] < 0).any()  # the negative-count path was exercised


def test_lds_batch_advance_until(nthreads):
    """The resumable batch state on a network beyond the register kernels."""
    spec = synthetic(40, 80, 5, init_mean=25.0)
    g = build_device(spec, 13)
    o = build_oracle(spec, sparse=True)
    b = g.batch(96, traj_begin=3)
    for t in (0.01, 0.02, 0.005):
        b.advance_until(t)
    x = b.get_species()
    for j in range(0, 96, 11):
        tr = o.state(spec[2], 13, traj=3 + j)
        for t in (0.01, 0.02, 0.005):
            tr.advance_until(t)
        assert x[j].tolist() == tr.x.tolist(), j


def test_lds_per_trajectory_x0_and_range_guard(nthreads):
    spec = synthetic(24, 48, 9, init_mean=10.0)
    rng = np.random.default_rng(1)
    x0 = rng.poisson(10, size=(200, 24))
    g = build_device(spec, 4)
    out, st = _dev(g, 0.5, 8, 200, LDS, x0=x0)
    o = build_oracle(spec, sparse=True)
    _, oout, _, _, _ = o.run_grid(4, x0, 0.5, 8, n_traj=200, x0_per_traj=True, nthreads=nthreads)
    assert np.array_equal(out, oout)
    # counts beyond the f64 kernels' exact range (2^52) run on int64 counts
    big = (40, [(1.0, [1] + [0] * 39, [-1, 1] + [0] * 38)], [2**53 + 1, 0] + [0] * 38)
    times, bout, bst = build_device(big, 1).run(2e-15, 2, n_traj=4, with_stats=True)
    assert bst.kernel == LDS | _native.REBOP_KERNEL_WIDE
    _, boout, _, _, _ = build_oracle(big, sparse=True).run_grid(1, big[2], 2e-15, 2, n_traj=4)
    assert np.array_equal(bout, boout) and bout[-1, 1].min() > 0


def _wide_spec(x0a):
    """40 species (beyond the register kernels): a pool A with an inflow in
    jumps of 2^20 that feeds B in jumps of 2^20, B decaying, plus a small
    network on the other species."""
    S = 40
    rx = [(2.0, [0] * 40, [2**20] + [0] * 39),
          (2e-9, [1] + [0] * 39, [-(2**20), 2**20] + [0] * 38),
          (0.05, [0, 1] + [0] * 38, [0, -(2**10)] + [0] * 38)]
    for i in range(2, S - 1):
        o = [0] * S
        o[i] = 1
        d = [0] * S
        d[i] = -1
        d[i + 1] = 1
        rx.append((0.3, o, d))
    return S, rx, [x0a, 0] + [20] * 38


def _int16_spec(a0):
    """100 species (int32 counts take 25.6 KB of LDS per wave, which bounds
    the waves per CU; int16 counts lift that bound): A with an inflow in
    jumps of 700 (so it passes 2^15 within the horizon when it starts near
    30000), feeding B, which decays, and a chain on the others."""
    S = 100
    rx = [(3.0, [0] * S, [700] + [0] * (S - 1)),
          (0.01, [1] + [0] * (S - 1), [-1, 1] + [0] * (S - 2)),
          (0.5, [0, 1] + [0] * (S - 2), [0, -1] + [0] * (S - 2))]
    for i in range(2, S - 1):
        o = [0] * S
        o[i] = 1
        d = [0] * S
        d[i] = -1
        d[i + 1] = 1
        rx.append((0.3, o, d))
    return S, rx, [a0, 0] + [20] * (S - 2)


@pytest.mark.parametrize("a0,kind", [(30000, LDS), (40000, LDS), (2**15 - 1, LDS), (100, LDS | COMPACT)])
def test_lds_int16_counts_and_rerun(a0, kind, nthreads):
    """int16 counts while every count stays within 2^15 - 1; a launch that
    meets a larger count stops early and is rerun with int32 counts (A starts
    near 30000 and its inflow carries it past); initial counts beyond int16
    start at int32.  All equal the oracle."""
    spec = _int16_spec(a0)
    g = build_device(spec, 12)
    out, st = _dev(g, 3.0, 12, 256, LDS)
    assert st.kernel == kind
    o = build_oracle(spec, sparse=True)
    _, oout, _, oev, _ = o.run_grid(12, spec[2], 3.0, 12, n_traj=256, nthreads=nthreads)
    assert np.array_equal(out, oout)
    assert st.events == oev
    assert (np.abs(oout).max() > 2**15 - 1) == (kind == LDS)


def test_lds_int16_to_int64_in_two_reruns(nthreads, monkeypatch):
    """Counts that fit int16 at the start and pass 2^30 during the run: the
    compact launch and then the narrow one stop early, the wide one runs."""
    monkeypatch.setenv("REBOP_LDS_COMPACT", "1")
    S, rx, x0 = _wide_spec(20)
    rx[0] = (600.0, rx[0][1], rx[0][2])  # inflow: ~3600 jumps of 2^20 by t = 6
    spec = (S, rx, x0)
    g = build_device(spec, 8)
    out, st = _dev(g, 6.0, 12, 128, LDS)
    assert st.kernel == LDS | _native.REBOP_KERNEL_WIDE
    o = build_oracle(spec, sparse=True)
    _, oout, _, oev, _ = o.run_grid(8, spec[2], 6.0, 12, n_traj=128, nthreads=nthreads)
    assert np.array_equal(out, oout) and st.events == oev
    assert np.abs(oout).max() > 2**30


def test_lds_jobs_int16_rerun(nthreads):
    """A fused batch starts compact when every job's counts fit int16; one
    job passing 2^15 - 1 reruns the whole launch with int32 counts."""
    import torch

    specs = [_int16_spec(100), _int16_spec(30000)]
    nets = [build_device(sp, 20 + i) for i, sp in enumerate(specs)]
    bufs = [torch.zeros((13, sp[0], 128), dtype=torch.int64, device="cuda") for sp in specs]
    st, ev = run_jobs([dict(net=g, tmax=3.0, nb_steps=12, n_traj=128, out_ptr=b.data_ptr())
                       for g, b in zip(nets, bufs)], per_job_events=True)
    torch.cuda.synchronize()
    assert st.kernel == LDS
    for i, sp in enumerate(specs):
        _, oout, _, oev, _ = build_oracle(sp, sparse=True).run_grid(20 + i, sp[2], 3.0, 12, n_traj=128,
                                                                    nthreads=nthreads)
        assert np.array_equal(bufs[i].cpu().numpy(), oout) and int(ev[i]) == oev
    st1, _ = run_jobs([dict(net=nets[0], tmax=3.0, nb_steps=12, n_traj=128, out_ptr=bufs[0].data_ptr())],
                      per_job_events=True)
    assert st1.kernel == LDS | COMPACT


@pytest.mark.parametrize("x0a", [2**31 + 5, 2**30 - 3 * 2**20])
def test_lds_counts_beyond_int32(x0a, nthreads):
    """Populations beyond int32 run on the LDS kernel with int64 counts: chosen
    up front when the initial counts are beyond int32 (2^31 + 5), or by a rerun
    when a narrow run's count leaves int32 during the run (A starts at
    2^30 - 3 * 2^20 and its inflow carries it past 2^30).  Both equal the
    oracle (int64)."""
    spec = _wide_spec(x0a)
    g = build_device(spec, 8)
    out, st = _dev(g, 6.0, 12, 256, LDS)
    assert st.kernel == LDS | _native.REBOP_KERNEL_WIDE
    o = build_oracle(spec, sparse=True)
    _, oout, _, oev, _ = o.run_grid(8, spec[2], 6.0, 12, n_traj=256, nthreads=nthreads)
    assert np.array_equal(out, oout)
    assert st.events == oev
    assert np.abs(oout).max() > 2**30


def test_lds_async_narrow_run_keeps_the_trajectories_in_range(nthreads):
    """Without counters (stats NULL) nothing reruns a launch, so its lanes
    do not drop their work when another trajectory leaves the count range:
    the trajectories that stay in range equal the oracle."""
    import torch

    spec = _wide_spec(100)
    n = 128
    x0 = np.tile(np.asarray(spec[2], dtype=np.int64), (n, 1))
    x0[64:, 0] = 2**30 - 3 * 2**20  # these pass 2^30: left wrong, as documented
    g = build_device(spec, 14)
    out = torch.zeros((13, spec[0], n), dtype=torch.int64, device="cuda")
    xt = torch.as_tensor(x0, device="cuda")
    st = g.run_device(6.0, 12, n, out.data_ptr(), kernel=LDS, x0_ptr=xt.data_ptr(), with_stats=False)
    torch.cuda.synchronize()
    assert st is None
    o = build_oracle(spec, sparse=True)
    _, oout, _, _, _ = o.run_grid(14, x0[:64], 6.0, 12, n_traj=64, x0_per_traj=True, nthreads=nthreads)
    assert np.array_equal(out[:, :, :64].cpu().numpy(), oout)


def test_lds_wide_per_trajectory_x0_and_jobs(nthreads):
    """Device-side initial counts (not visible to the host) start narrow and
    are rerun wide; so are fused job batches; the batch state runs wide."""
    import torch

    spec = _wide_spec(2**30 - 3 * 2**20)
    x0 = np.tile(np.asarray(spec[2], dtype=np.int64), (128, 1))
    g = build_device(spec, 9)
    out, st = _dev(g, 6.0, 12, 128, LDS, x0=x0)
    assert st.kernel == LDS | _native.REBOP_KERNEL_WIDE
    o = build_oracle(spec, sparse=True)
    _, oout, _, oev, _ = o.run_grid(9, x0, 6.0, 12, n_traj=128, x0_per_traj=True, nthreads=nthreads)
    assert np.array_equal(out, oout) and st.events == oev
    buf = torch.zeros((13, spec[0], 128), dtype=torch.int64, device="cuda")
    st2, ev = run_jobs([dict(net=g, tmax=6.0, nb_steps=12, n_traj=128, out_ptr=buf.data_ptr())],
                       per_job_events=True)
    torch.cuda.synchronize()
    assert st2.kernel == LDS | _native.REBOP_KERNEL_WIDE
    assert np.array_equal(buf.cpu().numpy(), oout) and int(ev[0]) == oev
    b = g.batch(32)
    b.advance_until(3.0)
    b.advance_until(6.0)
    xs = b.get_species()
    for j in (0, 31):
        tr = o.state(spec[2], 9, traj=j)
        tr.advance_until(3.0)
        tr.advance_until(6.0)
        assert xs[j].tolist() == tr.x.tolist()


def test_fused_heterogeneous_jobs_bit_exact(nthreads):
    """A fused launch over many different networks equals each network's own
    oracle run; per-job event counts are exact."""
    import torch

    specs = workloads.heterogeneous_specs(24, seed=5, s_range=(3, 64), r_range=(4, 256))
    nets = workloads.build_nets(specs, seed=31)
    targets = workloads.target_events(len(specs), lo=50, hi=400)
    tmax = [t / max(workloads.initial_total_propensity(sp), 1e-9) for sp, t in zip(specs, targets)]
    n = 96
    outs = [torch.zeros((6, sp[0], n), dtype=torch.int64, device="cuda") for sp in specs]
    jobs = [dict(net=g, tmax=float(t), nb_steps=5, n_traj=n, traj_begin=10 * i, out_ptr=o.data_ptr(),
                 cost_hint=float(i)) for i, (g, t, o) in enumerate(zip(nets, tmax, outs))]
    st, ev = run_jobs(jobs, per_job_events=True)
    torch.cuda.synchronize()
    assert st.trajectories == n * len(specs)
    for i, (sp, t) in enumerate(zip(specs, tmax)):
        o = build_oracle(sp, sparse=True)
        _, oout, _, oev, _ = o.run_grid(31, sp[2], float(t), 5, n_traj=n, traj_begin=10 * i, nthreads=nthreads)
        assert np.array_equal(outs[i].cpu().numpy(), oout), i
        assert int(ev[i]) == oev, i
    assert int(ev.sum()) == st.events


def test_lds_records_that_do_not_fit(nthreads):
    """Reactions whose factors (third order: three factors) or jump (five
    species) do not fit the LDS record take the global lists; the
    result stays bit-exact, in a network that mixes both kinds."""
    S = 6
    rx = [
        (0.002, [3, 0, 0, 0, 0, 0], [-3, 1, 0, 0, 0, 0]),  # 3A -> B: three factors
        (0.5, [0, 1, 0, 0, 0, 0], [1, -1, 1, 1, 1, 1]),    # B -> A + C + D + E + F: six jump entries
        (0.3, [0, 0, 1, 0, 0, 0], [0, 0, -1, 0, 0, 0]),
        (0.3, [0, 0, 0, 1, 0, 0], [0, 0, 0, -1, 0, 0]),
        (0.2, [0, 0, 0, 0, 1, 1], [2, 0, 0, 0, -1, -1]),  # E + F -> 2A
        (5.0, [0, 0, 0, 0, 0, 0], [1, 0, 0, 0, 0, 0]),
    ]
    spec = (S, rx, [40, 5, 0, 0, 0, 0])
    g = build_device(spec, 21)
    out, st = _dev(g, 4.0, 20, 300, LDS)
    o = build_oracle(spec)
    _, oout, _, oev, _ = o.run_grid(21, spec[2], 4.0, 20, n_traj=300, nthreads=nthreads)
    assert np.array_equal(out, oout)
    assert st.events == oev and oev > 300 * 50


def test_lds_jumps_on_species_beyond_255(nthreads):
    """A jump entry keeps 8 bits of species index in its record; jumps that
    touch species 256 and above take the global lists and stay bit-exact
    (300 species: molecules hop i -> i + 7 mod 300, with sources at 0 and 280)."""
    S = 300
    rx = []
    for i in range(S):
        o = [0] * S
        o[i] = 1
        d = [0] * S
        d[i] = -1
        d[(i + 7) % S] += 1
        rx.append((0.5, o, d))
    for s in (0, 280):
        d = [0] * S
        d[s] = 1
        rx.append((3.0, [0] * S, d))
    x0 = [0] * S
    for i in range(250, S):
        x0[i] = 2
    spec = (S, rx, x0)
    g = build_device(spec, 17)
    out, st = _dev(g, 2.0, 10, 128, LDS)
    o = build_oracle(spec)
    _, oout, _, oev, _ = o.run_grid(17, x0, 2.0, 10, n_traj=128, nthreads=nthreads)
    assert np.array_equal(out, oout)
    assert st.events == oev and oev > 128 * 50


def test_autotune_picks_a_variant_that_stays_bit_exact(nthreads, tmp_path, monkeypatch):
    """rebop_autotune times the applicable variants; the AUTO runs that follow
    use the winner and still equal the oracle bit for bit."""
    monkeypatch.setenv("REBOP_TUNE_CACHE", str(tmp_path / "tune.txt"))  # time, do not reuse
    spec = dimers()
    g = build_device(spec, 31)
    res = g.autotune(0.5, 50, n_pilot=4096)
    assert res.n_tried >= 8 and res.ms > 0
    assert res.threads in (64, 128, 256) and 0 < res.occupancy <= 1.0
    out, st = _dev(g, 0.5, 50, 4096, _native.REBOP_KERNEL_AUTO)
    assert st.kernel & ~COMPACT == res.kernel
    o = build_oracle(spec)
    _, oout, _, oev, _ = o.run_grid(31, spec[2], 0.5, 50, n_traj=4096, nthreads=nthreads)
    assert np.array_equal(out, oout) and st.events == oev

