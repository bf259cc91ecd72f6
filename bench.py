"""Benchmark: SSA reaction events/s of the batched direct method on one MI355X.

Workload (one "step"): the dimer gene-expression network (gene -> gene+mRNA
@25, mRNA -> mRNA+protein @1000, 2 protein -> dimer @0.001, mRNA -> @0.1,
protein -> @1), init gene=1, 10^6 independent trajectories, tmax=1.0,
nb_steps=100 samples (101 grid points), FP64 times/propensities, int64
output [101][4][10^6] written to HBM.  The network and the output buffer are
resident on the device before timing starts; each step uses a new seed.

Prints ONE JSON line (see the keys below).  ``roofline`` uses the device
time of the SSA launch sequence (the main kernel, the grouped launch that
finishes its last trajectories, and the pass that converts the f64 samples to
int64; HIP events on their launch stream) and the algorithmic FP64 operation
count per event; ``cpu_baseline`` times the CPU
oracle (the C restatement of the reference algorithm) on a bounded sample of
the same workload over the host's cores.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# FP64 VALU issue rate of MI355X for non-fused ops: 256 CUs x 64 lanes/clk x
# 2.4 GHz = 39.3e12 ops/s (the published 78.6 TFLOP/s counts an FMA as 2;
# this path forbids contraction, so every add/mul/compare is one issue).
PEAK_FP64_OPS = 256 * 64 * 2.4e9
N_TRAJ = 1_000_000
TMAX = 1.0
NB_STEPS = 100
CPU_SAMPLE = 32_768


def algorithmic_ops_per_event(rx) -> int:
    """F = sum_r order_r (propensity multiplies) + R (cumsum adds) + R (search
    compares) + 6 (ln counted once, the division, t add, t > T, tot*U, 0 < tot)."""
    R = len(rx)
    return int(sum(sum(o) for _, o, _ in rx)) + 2 * R + 6


def parse_traffic(path: Path):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    if not path.exists():
        return None
    vals = {}
    for line in path.read_text().splitlines():
        if ":" in line and not line.lstrip().startswith("#"):
            k, v = line.split(":", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def parse_issue_counters(path: Path):
    """The measured VALU issue figures of the committed rocprofv3 --pmc
    summary (instructions per event and the issue rate the kernel reaches),
    if any: the FP64-op fraction above counts only the reference's own
    arithmetic, these count every instruction the kernel issues."""
    if not path.exists():
        return None
    out = {}
    keys = {
        "VALU wave-instructions per event": "valu_wave_instructions_per_event",
        "active-lane fraction of VALU issue": "active_lane_fraction",
        "VALU issue per CU-cycle": "valu_issue_per_cu_cycle",
    }
    for line in path.read_text().splitlines():
        for k, name in keys.items():
            if line.startswith("- " + k) and ":" in line:
                try:
                    out[name] = float(line.rsplit(":", 1)[1].split()[0])
                except ValueError:
                    pass
    if not out:
        return None
    out["source"] = (f"{path.relative_to(ROOT)} (rocprofv3 --pmc of the same workload, a committed "
                     "measurement, not this run's)")
    return out


# The GPU host allots 16 CPU cores to each GPU (its process and thread
# limits are set for that share, whatever os.cpu_count() reports), so the CPU
# baseline uses at most 16 threads; the counts the host reports are printed
# beside it.
HOST_CORE_SHARE = 16


def host_cpu_facts() -> dict:
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = None
    return {"sched_getaffinity": aff, "os_cpu_count": os.cpu_count(),
            "core_share_per_gpu": HOST_CORE_SHARE}


def host_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(HOST_CORE_SHARE, n))


CPU_MIN_WALL_S = 1.0  # per random-stream variant: >= 16 s of CPU work on 16 threads


def oracle_lma_net(spec):
    from oracle import pyoracle

    net = pyoracle.OracleNet(spec[0], sparse=True)
    for k, o, d in spec[1]:
        net.add_lma(k, o, d)
    return net


def _time_oracle(jobs, threads: int, rng: int, make_net=oracle_lma_net):
    """Runs every job of the list in one thread pool (the threads take
    trajectories of any job), repeated with further seeds until it has taken
    CPU_MIN_WALL_S of wall time; returns (events, seconds, repetitions, the
    fewest threads that ran trajectories in a repetition)."""
    from oracle import pyoracle

    nets = [make_net(spec) for spec, *_ in jobs]
    events, reps, busy = 0, 0, threads
    t0 = time.perf_counter()
    while reps == 0 or time.perf_counter() - t0 < CPU_MIN_WALL_S:
        batch = [(net, seed + 1000 * reps, spec[2], tmax, nb, n, tb)
                 for net, (spec, seed, tmax, nb, n, tb) in zip(nets, jobs)]
        ev, _, b = pyoracle.run_grid_jobs(batch, rng=rng, nthreads=threads)
        events += ev
        busy = min(busy, b)
        reps += 1
    return events, time.perf_counter() - t0, reps, busy


def cpu_baseline(jobs, note: str, make_net=oracle_lma_net) -> dict:
    """Times the CPU oracle over `jobs` = [(spec, seed, tmax, nb_steps, n_traj, traj_begin)]
    twice: with the device's Philox stream and shared ln (identical arithmetic),
    and with xoshiro256++ + libm log (closer to the reference's own per-event
    cost).  The faster of the two is the baseline value."""
    from oracle import pyoracle

    threads = host_threads()
    ev_p, dt_p, rep_p, busy_p = _time_oracle(jobs, threads, pyoracle.PHILOX, make_net)
    ev_x, dt_x, rep_x, busy_x = _time_oracle(jobs, threads, pyoracle.XOSHIRO, make_net)
    rate_p, rate_x = ev_p / dt_p, ev_x / dt_x
    return {
        "value": max(rate_p, rate_x),
        "unit": "events/s",
        "cores": threads,
        "kind": "port",
        "threads_busy": min(busy_p, busy_x),
        "host": host_cpu_facts(),
        "philox_ln64_events_per_s": rate_p,
        "xoshiro_libm_events_per_s": rate_x,
        "sample": f"{note}, repeated with new seeds to >= {CPU_MIN_WALL_S:.0f} s wall per "
                  f"stream; C restatement of advance_until over {threads} threads, sparse rate "
                  f"form (bit-identical to dense, faster), timed with the device's Philox + ln64 "
                  f"stream ({rep_p} x, {ev_p} events, {dt_p:.2f} s) and with xoshiro256++ + libm "
                  f"log ({rep_x} x, {ev_x} events, {dt_x:.2f} s); value is the faster; the Rust "
                  "reference cannot be built here",
    }


class Workload:
    """One benchmark configuration: set up device buffers, run one step."""

    name = ""
    description = ""

    def setup(self):
        raise NotImplementedError

    def step(self, seed: int):
        """Runs one step; returns a Stats."""
        raise NotImplementedError

    def ops_per_event(self) -> float:
        raise NotImplementedError

    def cpu_jobs(self):
        raise NotImplementedError

    def output_bytes(self) -> int:
        """int64 samples written per step ((nb_steps + 1) * S * 8 per trajectory)."""
        raise NotImplementedError


class SingleNetwork(Workload):
    def __init__(self, name, spec, tmax, nb, n, kernel, description, cpu_n):
        self.name, self.spec, self.tmax, self.nb, self.n = name, spec, tmax, nb, n
        self.kernel, self.description, self.cpu_n = kernel, description, cpu_n

    def setup(self):
        import torch

        from rebop_amd import networks

        self.out = torch.empty((self.nb + 1, self.spec[0], self.n), dtype=torch.int64, device="cuda")
        self.stream = torch.cuda.current_stream().cuda_stream
        self.g = networks.build(self.spec, 0)

    def step(self, seed):
        self.g.seed(seed)
        return self.g.run_device(self.tmax, self.nb, self.n, self.out.data_ptr(), kernel=self.kernel,
                                 stream=self.stream, with_stats=True)

    def ops_per_event(self):
        return algorithmic_ops_per_event(self.spec[1])

    def output_bytes(self):
        return (self.nb + 1) * self.spec[0] * 8 * self.n

    def cpu_jobs(self):
        return [(self.spec, 42, self.tmax, self.nb, self.cpu_n, 0)], \
            f"first {self.cpu_n} trajectories of the same workload"


class Heterogeneous(Workload):
    name = "het"

    def __init__(self, n_networks, n_traj, nb, seed=2026):
        self.n_networks, self.n_traj, self.nb, self.seed = n_networks, n_traj, nb, seed
        self.description = (f"heterogeneous batch: {n_networks} random mass-action networks "
                            f"(species U{{3..64}}, reactions U{{4..256}}, init Poisson(50)), "
                            f"tmax per network for 1e2..1e5 mean events, {n_traj} trajectories "
                            f"each, nb_steps={nb}, one fused launch")

    def setup(self):
        import torch

        from rebop_amd import workloads

        log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
        t0 = time.perf_counter()
        self.specs = workloads.heterogeneous_specs(self.n_networks, self.seed)
        self.nets = workloads.build_nets(self.specs, seed=0)
        log(f"het: {self.n_networks} networks built in {time.perf_counter() - t0:.1f} s")
        self.targets = workloads.target_events(self.n_networks)
        self.tmax = workloads.calibrate_tmax(self.nets, self.specs, self.targets, log=log)
        self.outs = [torch.empty((self.nb + 1, sp[0], self.n_traj), dtype=torch.int64, device="cuda")
                     for sp in self.specs]
        self.stream = torch.cuda.current_stream().cuda_stream
        R = np.array([len(sp[1]) for sp in self.specs], dtype=np.float64)
        self.cost = self.targets * (R + 16.0)

    def step(self, seed):
        from rebop_amd.engine import run_jobs

        for g in self.nets:
            g.seed(seed)
        jobs = [dict(net=g, tmax=float(t), nb_steps=self.nb, n_traj=self.n_traj,
                     out_ptr=o.data_ptr(), cost_hint=float(c))
                for g, t, o, c in zip(self.nets, self.tmax, self.outs, self.cost)]
        st, ev = run_jobs(jobs, stream=self.stream, per_job_events=True)
        self.last_events = ev
        return st

    def ops_per_event(self):
        # events-weighted mean of each network's F
        F = np.array([algorithmic_ops_per_event(sp[1]) for sp in self.specs], dtype=np.float64)
        ev = getattr(self, "last_events", None)
        if ev is None or ev.sum() == 0:
            return float(F.mean())
        return float((F * ev).sum() / ev.sum())

    def output_bytes(self):
        return sum((self.nb + 1) * sp[0] * 8 * self.n_traj for sp in self.specs)

    def cpu_jobs(self):
        # every 32nd network, 16 trajectories each, all in one thread pool
        idx = list(range(0, self.n_networks, 32))
        jobs = [(self.specs[i], 42, float(self.tmax[i]), self.nb, 16, 0) for i in idx]
        return jobs, (f"{len(idx)} networks (every 32nd) x 16 trajectories of the same batch, "
                      "run concurrently in one pool")


def system_terms(orders, diffs):
    """define_system! reactant and product terms of a mass-action reaction."""
    reac = [(s, int(o)) for s, o in enumerate(orders) if o > 0]
    prod = [(s, int(o) + int(d)) for s, (o, d) in enumerate(zip(orders, diffs)) if int(o) + int(d) > 0]
    return reac, prod


class SystemNetwork(SingleNetwork):
    """The network compiled with define_system!'s arithmetic
    (src/gillespie_macro.rs:49-170): integer mass-action products and the
    first-match choice (the system mode of the engine)."""

    def setup(self):
        import torch

        from rebop_amd.engine import GillespieNet

        self.out = torch.empty((self.nb + 1, self.spec[0], self.n), dtype=torch.int64, device="cuda")
        self.stream = torch.cuda.current_stream().cuda_stream
        self.g = GillespieNet.system(self.spec[2], 0)
        for k, o, d in self.spec[1]:
            self.g.add_system_reaction(k, *system_terms(o, d))

    def cpu_make_net(self):
        from oracle import pyoracle

        def make(spec):
            net = pyoracle.OracleNet.system(spec[0])
            for k, o, d in spec[1]:
                net.add_system(k, *system_terms(o, d))
            return net

        return make


class ExprNetwork(SingleNetwork):
    """The network with Rate::Expr rates (src/expr.rs): exprs[r] is reaction
    r's expression over the species names, or None where the reaction keeps
    its mass-action rate from the spec."""

    def __init__(self, *a, names, exprs, **kw):
        super().__init__(*a, **kw)
        self.names, self.exprs = names, exprs

    def _postfix(self):
        from rebop_amd import expr as _expr

        species = {n: i for i, n in enumerate(self.names)}
        return [None if e is None else _expr.to_postfix(_expr.parse(e), species, {}) for e in self.exprs]

    def setup(self):
        import torch

        from rebop_amd.engine import GillespieNet, Rate, RateExpr

        self.out = torch.empty((self.nb + 1, self.spec[0], self.n), dtype=torch.int64, device="cuda")
        self.stream = torch.cuda.current_stream().cuda_stream
        self.g = GillespieNet(self.spec[2], seed=0)
        for pf, (k, o, d) in zip(self._postfix(), self.spec[1]):
            self.g.add_reaction(Rate.lma(k, o) if pf is None else RateExpr.from_postfix(*pf), d)

    def cpu_make_net(self):
        from oracle import pyoracle

        pf = self._postfix()

        def make(spec):
            net = pyoracle.OracleNet(spec[0])
            for p, (k, o, d) in zip(pf, spec[1]):
                if p is None:
                    net.add_lma(k, o, d)
                else:
                    net.add_expr(*p, d)
            return net

        return make


class EventsNetwork(Workload):
    """The nb_steps == 0 mode (src/pyo3_gillespie.rs:209-223): every event of
    every trajectory recorded, two passes (record counts, then the records)
    on device buffers (rebop_run_all_events_device), the offsets an exclusive
    scan of the counts on the device."""

    def __init__(self, name, spec, tmax, n, description, cpu_n):
        self.name, self.spec, self.tmax, self.n = name, spec, tmax, n
        self.description, self.cpu_n = description, cpu_n
        self.records = 0

    def setup(self):
        import torch

        from rebop_amd import networks

        self.g = networks.build(self.spec, 0)
        self.stream = torch.cuda.current_stream().cuda_stream
        self.counts = torch.zeros(self.n, dtype=torch.int64, device="cuda")  # uint64 bits
        self.cap = 0

    def step(self, seed):
        import torch

        self.g.seed(seed)
        st1 = self.g.run_all_events_device(self.tmax, self.n, self.counts.data_ptr(), stream=self.stream)
        offs = torch.cumsum(self.counts, 0) - self.counts
        total = int(self.counts.sum().item())
        if total > self.cap:
            self.times = torch.empty(total, dtype=torch.float64, device="cuda")
            self.states = torch.empty((total, self.spec[0]), dtype=torch.int64, device="cuda")
            self.cap = total
        st2 = self.g.run_all_events_device(self.tmax, self.n, self.counts.data_ptr(), offs.data_ptr(),
                                           self.times.data_ptr(), self.states.data_ptr(), stream=self.stream)
        self.records = total
        st2.kernel_ms += st1.kernel_ms  # both passes
        return st2

    def host_buffer_rate(self, seed: int = 7) -> float:
        """events/s of the same two passes through rebop_run_all_events (host
        buffers: the records cross PCIe), wall time."""
        t0 = time.perf_counter()
        ts, _ = self.g.run_all_events(self.tmax, n_traj=self.n)
        dt = time.perf_counter() - t0
        ev = sum(len(t) - 1 - (1 if np.isinf(t[-1]) else 0) for t in ts)
        return ev / dt

    def ops_per_event(self):
        return algorithmic_ops_per_event(self.spec[1])

    def output_bytes(self):
        return self.records * (8 + 8 * self.spec[0])

    def cpu_jobs(self):
        return None, (f"the oracle's all-events loop over trajectories 0.. of the same workload, "
                      f"blocks of 64 over {host_threads()} threads")


def cpu_baseline_events(wl: EventsNetwork, note: str) -> dict:
    """The oracle's nb_steps == 0 loop over trajectories 0, 1, ... of the
    workload on the host's threads (blocks of 64), repeated over further
    trajectories until CPU_MIN_WALL_S has passed; events exclude the initial
    and absorbing records."""
    from oracle import pyoracle

    threads = host_threads()
    net = oracle_lma_net(wl.spec)
    out = {}
    busy = threads
    for name, rng in (("philox_ln64_events_per_s", pyoracle.PHILOX),
                      ("xoshiro_libm_events_per_s", pyoracle.XOSHIRO)):
        ev, j, t0 = 0, 0, time.perf_counter()
        while time.perf_counter() - t0 < CPU_MIN_WALL_S:
            e, _, b = net.run_all_events_many(42, j, 64 * 16 * threads, wl.spec[2], wl.tmax, rng=rng,
                                              nthreads=threads, with_busy=True)
            ev += e
            busy = min(busy, b)
            j += 64 * 16 * threads
        out[name] = ev / (time.perf_counter() - t0)
        out[name + "_trajectories"] = j
    return {"value": max(out["philox_ln64_events_per_s"], out["xoshiro_libm_events_per_s"]),
            "unit": "events/s", "cores": threads, "kind": "port", "threads_busy": busy,
            "host": host_cpu_facts(), **out,
            "sample": note + f", >= {CPU_MIN_WALL_S:.0f} s wall per random stream; the faster is the value"}


def make_workload(name: str, kernel: int) -> Workload:
    from rebop_amd import networks, workloads

    if name == "dimer":
        return SingleNetwork("dimer", networks.dimers(), TMAX, NB_STEPS, N_TRAJ, kernel,
                             "dimer gene-expression network (4 species, 5 reactions), 10^6 "
                             "trajectories, tmax=1.0, nb_steps=100, int64 counts out [101][4][10^6]",
                             CPU_SAMPLE)
    if name == "sir":
        return SingleNetwork("sir", networks.sir(), 250.0, 250, 10_000, kernel,
                             "SIR (3 species, 2 reactions), 10^4 trajectories, tmax=250, "
                             "nb_steps=250", 10_000)
    if name == "syn":
        return SingleNetwork("syn", workloads.synthetic_large(), 10.0, 200, 2**18, kernel,
                             "synthetic network, 128 species x 512 reactions (seed 7), 2^18 "
                             "trajectories, tmax=10, nb_steps=200", 256)
    if name == "het":
        return Heterogeneous(4096, 256, 50)
    if name == "dimer-system":
        return SystemNetwork("dimer-system", networks.dimers(), TMAX, NB_STEPS, N_TRAJ, kernel,
                             "dimer network in define_system! form (integer mass-action products, "
                             "first-match choice), 10^6 trajectories, tmax=1.0, nb_steps=100",
                             CPU_SAMPLE)
    if name == "dimer-expr":
        return ExprNetwork("dimer-expr", networks.dimers(), TMAX, NB_STEPS, N_TRAJ, kernel,
                           "dimer network with Rate::Expr rates (25*gene, 1000*mRNA, "
                           "0.001*protein*(protein-1), 0.1*mRNA, protein), 10^6 trajectories, "
                           "tmax=1.0, nb_steps=100", CPU_SAMPLE,
                           names=["gene", "mRNA", "protein", "dimer"],
                           exprs=["25*gene", "1000*mRNA", "0.001*protein*(protein-1)", "0.1*mRNA", "protein"])
    if name == "big-expr":
        spec, names, exprs = workloads.expression_large()
        return ExprNetwork("big-expr", spec, 5.0, 50, 2**17, kernel,
                           "mixed network beyond the specialised kernels' size: 48 species, 96 "
                           "reactions (growth-bounded random mass action, seed 11), the 53 with "
                           "one reactant molecule given Hill / exp / min / threshold Rate::Expr "
                           "rates, 2^17 trajectories, tmax=5, nb_steps=50", 1024,
                           names=names, exprs=exprs)
    if name == "sir-events":
        return EventsNetwork("sir-events", networks.sir(), 250.0, 10_000,
                             "SIR (3 species, 2 reactions), 10^4 trajectories, tmax=250, nb_steps=0 "
                             "(every event recorded; two passes into device buffers)", 0)
    raise SystemExit(f"unknown workload {name}")


PEAK_BF16_DENSE = 2.5e15


def surrogate_summary(args) -> dict:
    """The metric's second half beside the SSA line: the Michaelis-Menten
    surrogate trained on this engine's SSA (10^5 points x 64 trajectories),
    gated at W1 <= 2 % of the mean against 10^4-trajectory SSA histograms
    (validation 512 points, held-out check 256), timed on 10^6 queries."""
    from rebop_amd import surrogate

    res = surrogate.benchmark(n_train=args.surrogate_points, steps=args.surrogate_train_steps,
                              n_valid=512, n_holdout=256,
                              log=lambda m: print(m, file=sys.stderr, flush=True))
    return {
        "metric": "surrogate vs GPU SSA speedup at W1 <= 2% of the mean",
        "effective_speedup_gated_sweep": res["effective_speedup"],
        "effective_speedup_species_queries": res["effective_speedup_species"],
        "raw_speedup_all_queries": res["speedup"],
        "surrogate_queries_per_s": res["surrogate_qps"],
        "ssa_queries_per_s": res["ssa_qps"],
        "trusted_query_fraction": res["trusted_query_fraction"],
        "trusted_species_query_fraction": res["trusted_species_query_fraction"],
        "holdout_admitted_pass_fraction": res["holdout_admitted_pass_fraction"],
        "w1_pass_fraction": res["w1_pass_fraction"],
        "ssa_self_w1_over_2pct": res["ssa_self_w1_over_2pct"],
        "mfma_frac": res["gemm_tflops"] * 1e12 / PEAK_BF16_DENSE,
        "train_s": res["train_s"],
        "note": ("a query is every species' distribution at one (parameter point, sample time); "
                 "the trust map sends the queries it does not admit to the SSA, whose cost per "
                 "query is a 10^4-trajectory run of the point shared by its 32 sample times"),
    }


def surrogate_main(args) -> int:
    """Surrogate workload: 10^5 log-uniform Michaelis-Menten parameter points x
    64 SSA trajectories at 32 times in [0, 50] (one sweep launch), training of
    the 5-layer width-512 BF16 MLP with an 8-component negative-binomial
    mixture head, W1 validation against 10^4-trajectory SSA histograms, and
    10^6 queries timed on the device.  ``value`` is surrogate queries/s; the
    GPU SSA's queries/s at the same statistical quality (10^4 trajectories
    per point, shared by its 32 times) is reported beside it."""
    import torch

    from rebop_amd import _native, surrogate

    dev_name, cus = _native.device_info()
    res = surrogate.benchmark(n_train=args.surrogate_points, steps=args.surrogate_train_steps,
                              log=lambda m: print(m, file=sys.stderr, flush=True))
    torch.cuda.synchronize()
    flops_per_query = 2.0 * sum(a * b for a, b in zip(surrogate.MLP_DIMS[:-1], surrogate.MLP_DIMS[1:]))
    achieved = res["gemm_tflops"]
    line = {
        "metric": "surrogate distribution queries/s on 1 MI355X (vs GPU SSA at 10^4 trajectories)",
        "value": res["surrogate_qps"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": res["surrogate_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 (fp32 accumulate)",
        "data": "synthetic",
        "config": {
            "workload": "Michaelis-Menten surrogate: %d training points x 64 trajectories x 32 "
                        "times, 5-layer width-512 MLP, 8-component NB mixture per species, "
                        "10^6 queries" % args.surrogate_points,
            "device": dev_name,
            "compute_units": cus,
            "ssa_queries_per_s": res["ssa_qps"],
            # at the W1 <= 2 % gate: surrogate where the trust map admits a
            # query, the SSA for the rest
            "speedup_vs_gpu_ssa": res["effective_speedup"],
            # every query answered by the surrogate, gate ignored
            "raw_speedup_all_queries": res["speedup"],
            "w1_pass_fraction": res["w1_pass_fraction"],
            "point_pass_fraction": res["point_pass_fraction"],
            "w1_pass_fraction_reproducible": res["w1_pass_fraction_reproducible"],
            "ssa_self_w1_over_2pct": res["ssa_self_w1_over_2pct"],
            "trusted_cells": res["trusted_cells"],
            "trusted_validated_cells": res["trusted_validated_cells"],
            "trusted_query_fraction": res["trusted_query_fraction"],
            "effective_speedup_gated_sweep": res["effective_speedup"],
            "trusted_species_query_fraction": res["trusted_species_query_fraction"],
            "effective_speedup_species_queries": res["effective_speedup_species"],
            "trust_map": {"bins": res["trust_bins"], "tol": res["trust_tol"],
                          "min_points": res["trust_min_points"], "validation_points": res["n_valid"]},
            "holdout_admitted_entries": res["holdout_admitted_entries"],
            "holdout_admitted_queries": res["holdout_admitted_queries"],
            "holdout_admitted_pass_fraction": res["holdout_admitted_pass_fraction"],
            "ssa_point_ms": res["ssa_point_ms"],
            "rel_w1_median": res["rel_w1_median"],
            "train_s": res["train_s"],
            "nll_first": res["nll_first"],
            "nll_last": res["nll_last"],
        },
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": PEAK_BF16_DENSE / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved * 1e12 / PEAK_BF16_DENSE,
            "traffic": parse_traffic(ROOT / "profiles" / "surrogate_mm_traffic.txt"),
            "flops_per_query": flops_per_query,
            "timed": "fused inference kernel (features, 5 layers, head offsets), one launch of "
                     "10^6 queries (HIP events)",
        },
    }
    print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--kernel", type=int, default=0, help="REBOP_KERNEL_* (0 = auto)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="before timing, time every kernel variant and launch shape at the "
                         "workload's full size (rebop_autotune); AUTO steps then use the fastest")
    ap.add_argument("--workload", default="dimer", choices=["dimer", "sir", "syn", "het", "surrogate", "dimer-system",
                                                              "dimer-expr", "big-expr", "sir-events"],
                    help="dimer (the headline configuration), sir, syn (128x512), het (4096 "
                         "networks), surrogate (Michaelis-Menten sweep -> MLP -> W1 gate), "
                         "dimer-system (define_system! arithmetic), dimer-expr (Rate::Expr rates), "
                         "big-expr (48x96 with Rate::Expr rates, LDS kernel), sir-events "
                         "(nb_steps = 0)")
    ap.add_argument("--surrogate-points", type=int, default=100_000)
    ap.add_argument("--no-surrogate", action="store_true",
                    help="leave the surrogate half of the metric out of the default (dimer) line")
    ap.add_argument("--surrogate-train-steps", type=int, default=12_000)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one device copy of 1 GiB (read 1 GiB, write 1 GiB) and exit; "
                         "scales rocprofv3 FETCH_SIZE/WRITE_SIZE to bytes")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 8, "copy_bytes_written": b.numel() * 8}))
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the SSA shards by trajectory id with no "
              f"communication, and this engine runs on a single MI355X (--gpus {args.gpus} "
              "is not supported)", file=sys.stderr)
        return 2

    if args.workload == "surrogate":
        return surrogate_main(args)

    import torch

    from rebop_amd import _native

    wl = make_workload(args.workload, args.kernel)
    dev_name, cus = _native.device_info()
    wl.setup()
    tuned = None
    if args.autotune and isinstance(wl, SingleNetwork) and args.kernel == 0:
        wl.g.seed(999)
        res = wl.g.autotune(wl.tmax, wl.nb, n_pilot=wl.n)
        tuned = {"kernel": int(res.kernel), "threads": int(res.threads),
                 "occupancy": float(res.occupancy), "pilot_ms": float(res.ms), "tried": int(res.n_tried)}
        print(f"autotune: {tuned}", file=sys.stderr, flush=True)

    for w in range(args.warmup):
        wl.step(1000 + w)
        torch.cuda.synchronize()
        print(f"warmup step {w} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    events = 0
    trajs = 0
    kernel_ms = []
    kinds = set()
    t0 = time.perf_counter()
    for k in range(args.steps):
        st = wl.step(42 + k)
        events += st.events
        trajs += st.trajectories
        kernel_ms.append(st.kernel_ms)
        kinds.add(int(st.kernel))
        if st.kernel_ms > 2000.0:
            print(f"step {k}: {st.events} events, {st.kernel_ms:.0f} ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    value = events / wall
    out_bytes = wl.output_bytes()
    avg_kernel_s = sum(kernel_ms) / len(kernel_ms) / 1e3
    ev_per_launch = events / args.steps
    F = wl.ops_per_event()
    achieved = ev_per_launch * F / avg_kernel_s / 1e12
    peak = PEAK_FP64_OPS / 1e12
    stem = {"dimer": "ssa_dimer_1e6", "sir": "ssa_sir_1e4", "syn": "ssa_syn", "het": "ssa_het"}.get(
        wl.name, "ssa_" + wl.name.replace("-", "_"))
    traffic = parse_traffic(ROOT / "profiles" / f"{stem}_traffic.txt")
    line = {
        "metric": "SSA reaction events/sec on 1 MI355X",
        "value": value,
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": wl.description,
            "events_per_step": ev_per_launch,
            "trajectories_per_s": trajs / wall,
            "output_gb_per_s": out_bytes * args.steps / wall / 1e9,
            "kernel": sorted(kinds),
            "autotuned": tuned,
            "device": dev_name,
            "compute_units": cus,
        },
        "roofline": {
            "bound": "fp64-valu-issue",
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            # the same against the published 78.6 TFLOP/s, which counts an
            # FMA as two operations (the SSA path has no contraction)
            "frac_of_published_fp64_peak": achieved / (2.0 * peak),
            "published_fp64_peak": 2.0 * peak,
            "traffic": traffic,
            "traffic_source": (f"profiles/{stem}_traffic.txt (rocprofv3 FETCH_SIZE + WRITE_SIZE of "
                               "the same workload, a committed measurement, not this run's)"
                               if traffic is not None else None),
            "ops_per_event": F,
            "kernel_ms": avg_kernel_s * 1e3,
            "timed": ("all-events mode on device buffers: the counting and the recording pass "
                      "(HIP events)"
                      if isinstance(wl, EventsNetwork) else
                      "SSA launch sequence: main kernel + tail launch + int64 sample conversion (HIP events)"),
            "measured_issue": parse_issue_counters(ROOT / "profiles" / f"{stem}_counters.md"),
        },
    }
    if isinstance(wl, EventsNetwork):
        line["config"]["host_buffer_events_per_s"] = wl.host_buffer_rate()
    if args.workload == "dimer" and not args.no_surrogate:
        line["surrogate"] = surrogate_summary(args)
    if not args.no_cpu_baseline:
        jobs, note = wl.cpu_jobs()
        if isinstance(wl, EventsNetwork):
            line["cpu_baseline"] = cpu_baseline_events(wl, note)
        else:
            make_net = wl.cpu_make_net() if hasattr(wl, "cpu_make_net") else oracle_lma_net
            line["cpu_baseline"] = cpu_baseline(jobs, note, make_net)
        line["speedup_vs_cpu_baseline"] = value / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
