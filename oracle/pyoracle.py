"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""

from __future__ import annotations

import ctypes
import subprocess
from ctypes import POINTER, c_double, c_int, c_int64, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

PHILOX = 0
XOSHIRO = 1

_lib = None


def build() -> Path:
    """(Re)build liboracle.so with the committed Makefile."""
    res = subprocess.run(["make", "-s", "-B", "-C", str(HERE)], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    return LIB_PATH


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    lib = ctypes.CDLL(str(LIB_PATH))
    P = POINTER
    sig = {
        "orc_net_new": (c_void_p, [c_uint32, c_int]),
        "orc_net_new_system": (c_void_p, [c_uint32]),
        "orc_net_add_reaction_system": (
            c_int,
            [c_void_p, c_double, P(c_uint32), P(c_uint32), c_uint32, P(c_uint32), P(c_uint32), c_uint32],
        ),
        "orc_choose_first_above": (c_uint64, [c_double, P(c_double), c_uint64]),
        "orc_net_free": (None, [c_void_p]),
        "orc_net_nb_species": (c_uint32, [c_void_p]),
        "orc_net_nb_reactions": (c_uint32, [c_void_p]),
        "orc_net_add_reaction": (
            c_int,
            [c_void_p, c_int, c_double, P(c_uint32), P(c_uint32), c_uint32, P(c_int64), c_uint32],
        ),
        "orc_net_add_reaction_expr": (
            c_int,
            [c_void_p, P(c_uint32), c_uint32, P(c_double), c_uint32, P(c_int64), c_uint32],
        ),
        "orc_rate": (c_double, [c_void_p, c_uint32, P(c_int64)]),
        "orc_make_cumrates": (c_double, [c_void_p, P(c_int64), P(c_double)]),
        "orc_choose_cumrate_sum": (c_uint64, [c_double, P(c_double), c_uint64]),
        "orc_affect": (None, [c_void_p, c_uint32, P(c_int64)]),
        "orc_run_grid": (
            c_int,
            [c_void_p, c_uint64, c_uint64, c_uint64, P(c_int64), c_int, c_double, c_uint64,
             P(c_int64), P(c_int64), P(c_double), P(c_uint64), P(c_uint64), c_int, c_int, c_uint64],
        ),
        "orc_run_all_events": (
            c_int,
            [c_void_p, c_uint64, c_uint64, P(c_int64), c_double, c_uint64, P(c_double),
             P(c_int64), P(c_uint64), c_int, c_uint64],
        ),
        "orc_run_all_events_many": (
            c_int,
            [c_void_p, c_uint64, c_uint64, c_uint64, P(c_int64), c_double, P(c_uint64), P(c_uint64), c_int,
             c_int, P(c_int)],
        ),
        "orc_run_grid_jobs": (
            c_int,
            [c_void_p, c_uint32, c_int, c_int, P(c_uint64), P(c_uint64), P(c_int)],
        ),
        "orc_philox4x32_10": (None, [P(c_uint32), P(c_uint32), P(c_uint32)]),
        "orc_stream_block": (None, [c_uint64, c_uint64, c_uint64, P(c_uint32)]),
        "orc_ln64": (c_double, [c_double]),
        "orc_exp64": (c_double, [c_double]),
        "orc_pow64": (c_double, [c_double, c_double]),
        "orc_fmax64": (c_double, [c_double, c_double]),
        "orc_fmin64": (c_double, [c_double, c_double]),
        "orc_u01_open_closed": (c_double, [c_uint32, c_uint32]),
        "orc_u01_closed_open": (c_double, [c_uint32, c_uint32]),
        "orc_grid_time": (c_double, [c_double, c_uint64, c_uint64]),
        "orc_hw_threads": (c_int, []),
        "orc_state_init": (None, [c_void_p, P(c_int64), c_uint64, c_uint64, c_int]),
        "orc_state_set_cap": (None, [c_void_p, c_uint64]),
        "orc_advance_until": (None, [c_void_p, c_void_p, c_double, P(c_double)]),
        "orc_advance_one_reaction": (None, [c_void_p, c_void_p, P(c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a, t):
    return a.ctypes.data_as(POINTER(t))


class OrcState(ctypes.Structure):
    _fields_ = [
        ("x", POINTER(c_int64)),
        ("t", c_double),
        ("seed", c_uint64),
        ("traj", c_uint64),
        ("draw", c_uint64),
        ("xs", c_uint64 * 4),
        ("rng_kind", c_int),
        ("events", c_uint64),
        ("draw_lim", c_uint64),
        ("capped", c_int),
    ]


class OracleError(RuntimeError):
    def __init__(self, code):
        super().__init__(f"oracle error {code}")
        self.code = code


class OracleNet:
    """A reaction network in the oracle (Gillespie's reactions + sparse flag)."""

    def __init__(self, nb_species: int, sparse: bool = False, system: bool = False):
        self.lib = load()
        if system:
            self.h = self.lib.orc_net_new_system(int(nb_species))
        else:
            self.h = self.lib.orc_net_new(int(nb_species), int(bool(sparse)))
        self.S = int(nb_species)

    @classmethod
    def system(cls, nb_species: int) -> "OracleNet":
        """A define_system! network (integer mass-action products, first-match choice)."""
        return cls(nb_species, system=True)

    def add_system(self, k: float, reactants, products) -> None:
        """define_system! reaction: reactants/products as [(species, multiplicity)] in written order."""
        ri = np.ascontiguousarray([p[0] for p in reactants], dtype=np.uint32)
        rm = np.ascontiguousarray([p[1] for p in reactants], dtype=np.uint32)
        pi = np.ascontiguousarray([p[0] for p in products], dtype=np.uint32)
        pm = np.ascontiguousarray([p[1] for p in products], dtype=np.uint32)
        rc = self.lib.orc_net_add_reaction_system(self.h, float(k), _p(ri, c_uint32), _p(rm, c_uint32), len(ri),
                                                  _p(pi, c_uint32), _p(pm, c_uint32), len(pi))
        if rc:
            raise OracleError(rc)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_net_free(self.h)
            self.h = None

    @property
    def R(self) -> int:
        return int(self.lib.orc_net_nb_reactions(self.h))

    def add_lma(self, k: float, orders, diffs) -> None:
        o = np.ascontiguousarray(orders, dtype=np.uint32)
        d = np.ascontiguousarray(diffs, dtype=np.int64)
        rc = self.lib.orc_net_add_reaction(self.h, 0, float(k), None, _p(o, c_uint32), len(o), _p(d, c_int64), len(d))
        if rc:
            raise OracleError(rc)

    def add_lma_sparse(self, k: float, pairs, diffs) -> None:
        i = np.ascontiguousarray([p[0] for p in pairs], dtype=np.uint32)
        o = np.ascontiguousarray([p[1] for p in pairs], dtype=np.uint32)
        d = np.ascontiguousarray(diffs, dtype=np.int64)
        rc = self.lib.orc_net_add_reaction(self.h, 1, float(k), _p(i, c_uint32), _p(o, c_uint32), len(o), _p(d, c_int64), len(d))
        if rc:
            raise OracleError(rc)

    def add_expr(self, code, consts, diffs) -> None:
        """Rate::expr reaction: postfix tokens (op | arg << 8) and constants."""
        c = np.ascontiguousarray(code, dtype=np.uint32)
        k = np.ascontiguousarray(consts, dtype=np.float64)
        d = np.ascontiguousarray(diffs, dtype=np.int64)
        rc = self.lib.orc_net_add_reaction_expr(self.h, _p(c, c_uint32), len(c), _p(k, c_double), len(k),
                                                _p(d, c_int64), len(d))
        if rc:
            raise OracleError(rc)

    def rate(self, r: int, x) -> float:
        xa = np.ascontiguousarray(x, dtype=np.int64)
        return float(self.lib.orc_rate(self.h, int(r), _p(xa, c_int64)))

    def cumrates(self, x):
        xa = np.ascontiguousarray(x, dtype=np.int64)
        cum = np.zeros(max(self.R, 1), dtype=np.float64)
        tot = self.lib.orc_make_cumrates(self.h, _p(xa, c_int64), _p(cum, c_double))
        return float(tot), cum[: self.R]

    def affect(self, r: int, x) -> np.ndarray:
        xa = np.ascontiguousarray(np.array(x, dtype=np.int64))
        self.lib.orc_affect(self.h, int(r), _p(xa, c_int64))
        return xa

    def run_grid(self, seed, x0, tmax, nb_steps, n_traj=1, traj_begin=0, rng=PHILOX,
                 nthreads=1, x0_per_traj=False, want_out=True, draw_cap=0):
        """The grid loop over trajectories traj_begin ..; draw_cap (0: none) is
        the device's watchdog, which the reference does not have."""
        x0a = np.ascontiguousarray(np.asarray(x0, dtype=np.int64).reshape(-1))
        out = np.zeros((nb_steps + 1, self.S, n_traj), dtype=np.int64) if want_out else None
        final = np.zeros((n_traj, self.S), dtype=np.int64)
        times = np.zeros(nb_steps + 1, dtype=np.float64)
        ev = c_uint64(0)
        dr = c_uint64(0)
        rc = self.lib.orc_run_grid(
            self.h, int(seed) & (2**64 - 1), int(traj_begin), int(n_traj), _p(x0a, c_int64),
            int(bool(x0_per_traj)), float(tmax), int(nb_steps),
            _p(out, c_int64) if want_out else None, _p(final, c_int64), _p(times, c_double),
            ctypes.byref(ev), ctypes.byref(dr), int(rng), int(nthreads), int(draw_cap),
        )
        if rc:
            raise OracleError(rc)
        return times, out, final, int(ev.value), int(dr.value)

    def run_all_events(self, seed, traj, x0, tmax, cap=1 << 16, rng=PHILOX, draw_cap=0):
        x0a = np.ascontiguousarray(np.asarray(x0, dtype=np.int64).reshape(-1))
        while True:
            times = np.zeros(cap, dtype=np.float64)
            xs = np.zeros((cap, self.S), dtype=np.int64)
            n = c_uint64(0)
            rc = self.lib.orc_run_all_events(
                self.h, int(seed) & (2**64 - 1), int(traj), _p(x0a, c_int64), float(tmax), int(cap),
                _p(times, c_double), _p(xs, c_int64), ctypes.byref(n), int(rng), int(draw_cap),
            )
            if rc:
                raise OracleError(rc)
            if n.value <= cap:
                return times[: n.value], xs[: n.value]
            cap = int(n.value)

    def run_all_events_many(self, seed, traj_begin, n_traj, x0, tmax, rng=PHILOX, nthreads=1,
                            with_busy=False):
        """(events, records) of run_all_events over n_traj trajectories on
        nthreads threads (the records are written and dropped); with_busy adds
        the number of threads that ran a trajectory."""
        x0a = np.ascontiguousarray(np.asarray(x0, dtype=np.int64).reshape(-1))
        ev = c_uint64(0)
        rec = c_uint64(0)
        busy = c_int(0)
        rc = self.lib.orc_run_all_events_many(self.h, int(seed) & (2**64 - 1), int(traj_begin), int(n_traj),
                                              _p(x0a, c_int64), float(tmax), ctypes.byref(ev), ctypes.byref(rec),
                                              int(rng), int(nthreads), ctypes.byref(busy))
        if rc:
            raise OracleError(rc)
        if with_busy:
            return int(ev.value), int(rec.value), int(busy.value)
        return int(ev.value), int(rec.value)

    def state(self, x0, seed, traj=0, rng=PHILOX) -> "OracleTrajectory":
        return OracleTrajectory(self, x0, seed, traj, rng)


class OrcGridJob(ctypes.Structure):
    _fields_ = [
        ("net", c_void_p),
        ("seed", c_uint64),
        ("traj_begin", c_uint64),
        ("n_traj", c_uint64),
        ("x0", POINTER(c_int64)),
        ("x0_per_traj", c_int),
        ("tmax", c_double),
        ("nb_steps", c_uint64),
        ("out", POINTER(c_int64)),
        ("final_x", POINTER(c_int64)),
        ("draw_cap", c_uint64),
    ]


def run_grid_jobs(jobs, rng=PHILOX, nthreads=1):
    """Grid runs of several networks in one thread pool (no sample output).

    jobs: [(OracleNet, seed, x0, tmax, nb_steps, n_traj, traj_begin)].
    Returns (events, draws, threads that ran at least one trajectory)."""
    lib = load()
    arr = (OrcGridJob * max(len(jobs), 1))()
    keep = []
    for q, (net, seed, x0, tmax, nb, n, tb) in enumerate(jobs):
        x0a = np.ascontiguousarray(np.asarray(x0, dtype=np.int64).reshape(-1))
        keep.append(x0a)
        arr[q].net = net.h
        arr[q].seed = int(seed) & (2**64 - 1)
        arr[q].traj_begin = int(tb)
        arr[q].n_traj = int(n)
        arr[q].x0 = _p(x0a, c_int64)
        arr[q].x0_per_traj = 0
        arr[q].tmax = float(tmax)
        arr[q].nb_steps = int(nb)
    ev, dr, busy = c_uint64(0), c_uint64(0), c_int(0)
    rc = lib.orc_run_grid_jobs(ctypes.cast(arr, c_void_p), len(jobs), int(rng), int(nthreads),
                               ctypes.byref(ev), ctypes.byref(dr), ctypes.byref(busy))
    if rc:
        raise OracleError(rc)
    return int(ev.value), int(dr.value), int(busy.value)


class OracleTrajectory:
    """One resumable trajectory (Gillespie's species, t and rng) in the oracle."""

    def __init__(self, net: OracleNet, x0, seed, traj, rng):
        self.net = net
        self.x = np.ascontiguousarray(np.array(x0, dtype=np.int64))
        self.st = OrcState()
        net.lib.orc_state_init(ctypes.byref(self.st), _p(self.x, c_int64), int(seed) & (2**64 - 1), int(traj), int(rng))
        self.cum = np.zeros(max(net.R, 1), dtype=np.float64)

    def advance_until(self, tmax: float) -> None:
        self.net.lib.orc_advance_until(self.net.h, ctypes.byref(self.st), float(tmax), _p(self.cum, c_double))

    def advance_one_reaction(self) -> None:
        self.net.lib.orc_advance_one_reaction(self.net.h, ctypes.byref(self.st), _p(self.cum, c_double))

    @property
    def t(self) -> float:
        return float(self.st.t)

    @property
    def draws(self) -> int:
        return int(self.st.draw)

    def set_cap(self, cap: int) -> None:
        """The device's watchdog from here on: at most `cap` more draws (0: none)."""
        self.net.lib.orc_state_set_cap(ctypes.byref(self.st), int(cap))

    @property
    def capped(self) -> bool:
        return bool(self.st.capped)


def philox(ctr, key):
    lib = load()
    c = (c_uint32 * 4)(*ctr)
    k = (c_uint32 * 2)(*key)
    o = (c_uint32 * 4)()
    lib.orc_philox4x32_10(c, k, o)
    return list(o)


def stream_block(seed, traj, draw):
    """rebop_stream_block(rebop_stream_of(seed, traj), draw): the kernels' form."""
    lib = load()
    o = (c_uint32 * 4)()
    lib.orc_stream_block(int(seed), int(traj), int(draw), o)
    return list(o)


def ln64(x: float) -> float:
    return float(load().orc_ln64(float(x)))


def exp64(x: float) -> float:
    return float(load().orc_exp64(float(x)))


def pow64(x: float, y: float) -> float:
    return float(load().orc_pow64(float(x), float(y)))


def fmax64(a: float, b: float) -> float:
    return float(load().orc_fmax64(float(a), float(b)))


def fmin64(a: float, b: float) -> float:
    return float(load().orc_fmin64(float(a), float(b)))


def choose_cumrate_sum(chosen: float, cum) -> int:
    c = np.ascontiguousarray(cum, dtype=np.float64)
    return int(load().orc_choose_cumrate_sum(float(chosen), _p(c, c_double), len(c)))


def choose_first_above(chosen: float, cum) -> int:
    c = np.ascontiguousarray(cum, dtype=np.float64)
    return int(load().orc_choose_first_above(float(chosen), _p(c, c_double), len(c)))


def hw_threads() -> int:
    return int(load().orc_hw_threads())
