"""Benchmark: SSA reaction events/s of the batched direct method on one MI355X.

Workload (one "step"): the dimer gene-expression network (gene -> gene+mRNA
@25, mRNA -> mRNA+protein @1000, 2 protein -> dimer @0.001, mRNA -> @0.1,
protein -> @1), init gene=1, 10^6 independent trajectories, tmax=1.0,
nb_steps=100 samples (101 grid points), FP64 times/propensities, int64
output [101][4][10^6] written to HBM.  The network and the output buffer are
resident on the device before timing starts; each step uses a new seed.

Prints ONE JSON line (see the keys below).  ``roofline`` uses the SSA
kernel's own device time (HIP events on its launch stream) and the
algorithmic FP64 operation count per event; ``cpu_baseline`` times the CPU
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


def cpu_baseline(spec, seed: int) -> dict:
    from oracle import pyoracle

    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        threads = os.cpu_count() or 1
    threads = max(1, min(16, threads))
    net = pyoracle.OracleNet(spec[0])
    for k, o, d in spec[1]:
        net.add_lma(k, o, d)
    t0 = time.perf_counter()
    _, _, _, ev, _ = net.run_grid(seed, spec[2], TMAX, NB_STEPS, n_traj=CPU_SAMPLE,
                                  nthreads=threads, want_out=False)
    dt = time.perf_counter() - t0
    return {
        "value": ev / dt,
        "unit": "events/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {CPU_SAMPLE} trajectories of the same workload ({ev} events, "
                  f"{dt:.2f} s wall on {threads} threads); C restatement of advance_until "
                  "with the same Philox stream (the Rust reference cannot be built here)",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--kernel", type=int, default=0, help="REBOP_KERNEL_* (0 = auto)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
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
              f"communication, and this build runs on a single MI355X (--gpus {args.gpus} "
              "is not supported)", file=sys.stderr)
        return 2

    import torch

    from rebop_amd import _native, networks

    spec = networks.dimers()
    S, rx, x0 = spec
    F = algorithmic_ops_per_event(rx)
    dev_name, cus = _native.device_info()
    out = torch.empty((NB_STEPS + 1, S, N_TRAJ), dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream

    g = networks.build(spec, 0)

    def step(seed: int):
        g.seed(seed)
        return g.run_device(TMAX, NB_STEPS, N_TRAJ, out.data_ptr(), kernel=args.kernel,
                            stream=stream, with_stats=True)

    for w in range(args.warmup):
        step(1000 + w)
    torch.cuda.synchronize()
    events = 0
    kernel_ms = []
    kinds = set()
    t0 = time.perf_counter()
    for k in range(args.steps):
        st = step(42 + k)
        events += st.events
        kernel_ms.append(st.kernel_ms)
        kinds.add(int(st.kernel))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    value = events / wall
    avg_kernel_s = sum(kernel_ms) / len(kernel_ms) / 1e3
    ev_per_launch = events / args.steps
    achieved = ev_per_launch * F / avg_kernel_s / 1e12
    peak = PEAK_FP64_OPS / 1e12
    traffic = parse_traffic(ROOT / "profiles" / "ssa_dimer_1e6_traffic.txt")
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
            "workload": "dimer gene-expression network (4 species, 5 reactions), 10^6 "
                        "trajectories, tmax=1.0, nb_steps=100, int64 counts out [101][4][10^6]",
            "n_trajectories": N_TRAJ,
            "nb_steps": NB_STEPS,
            "tmax": TMAX,
            "events_per_step": ev_per_launch,
            "kernel": sorted(kinds),
            "device": dev_name,
            "compute_units": cus,
        },
        "roofline": {
            "bound": "fp64-valu-issue",
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": traffic,
            "ops_per_event": F,
            "kernel_ms": avg_kernel_s * 1e3,
        },
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(spec, 42)
        line["speedup_vs_cpu_baseline"] = value / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
