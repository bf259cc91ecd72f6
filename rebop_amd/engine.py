"""Low-level mirror of rebop's Rust run-time API over the device library.

``Rate.lma`` / ``Rate.lma_sparse`` and ``GillespieNet`` follow
``src/gillespie.rs`` (``Rate`` :10-69, ``Gillespie`` :157-344), batched over
N independent trajectories: trajectory j of a run uses the Philox stream
(seed, traj_begin + j).  ``Batch`` holds N resumable trajectory states and
offers ``advance_until`` on all of them at once.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int64, c_uint32, c_void_p
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from . import _native
from ._native import Stats, check, load


def _i64(a: np.ndarray):
    return a.ctypes.data_as(POINTER(c_int64))


def _u32(a: np.ndarray):
    return a.ctypes.data_as(POINTER(c_uint32))


def _f64(a: np.ndarray):
    return a.ctypes.data_as(POINTER(c_double))


@dataclass(frozen=True)
class Rate:
    """An LMA rate: ``Rate::lma`` (dense orders) or ``Rate::lma_sparse`` (pairs)."""

    rate: float
    orders: tuple
    sparse: bool = False

    @staticmethod
    def lma(rate: float, orders: Sequence[int]) -> "Rate":
        """``Rate::lma(rate, orders)`` -- src/gillespie.rs:24-26."""
        return Rate(float(rate), tuple(int(o) for o in orders), False)

    @staticmethod
    def lma_sparse(rate: float, pairs: Sequence[tuple[int, int]]) -> "Rate":
        """``Rate::lma_sparse(rate, [(species, order)])`` -- src/gillespie.rs:29-31."""
        return Rate(float(rate), tuple((int(i), int(o)) for i, o in pairs), True)

    def dense(self) -> "Rate":
        """``Rate::dense`` (src/gillespie.rs:37-53): length max index + 1, orders summed."""
        if not self.sparse:
            return self
        n = max((i + 1 for i, _ in self.orders), default=0)
        d = [0] * n
        for i, o in self.orders:
            d[i] += o
        return Rate(self.rate, tuple(d), False)

    def sparse_form(self) -> "Rate":
        """``Rate::sparse`` (src/gillespie.rs:55-69): (index, order) for order > 0, ascending."""
        if self.sparse:
            return self
        return Rate(self.rate, tuple((i, o) for i, o in enumerate(self.orders) if o > 0), True)


@dataclass(frozen=True)
class RateExpr:
    """``Rate::Expr`` (src/gillespie.rs:18, :33-35): a rate expression in
    postfix bytecode (tokens ``op | arg << 8``, REBOP_EXPR_*) over constants.
    ``rebop_amd.expr.to_postfix`` builds one from a parsed string."""

    code: tuple
    consts: tuple

    @staticmethod
    def from_postfix(code: Sequence[int], consts: Sequence[float]) -> "RateExpr":
        return RateExpr(tuple(int(c) for c in code), tuple(float(k) for k in consts))

    # Rate::dense / Rate::sparse leave an expression unchanged (gillespie.rs:51, :67)
    def dense(self) -> "RateExpr":
        return self

    def sparse_form(self) -> "RateExpr":
        return self


@dataclass(frozen=True)
class Jump:
    """Net stoichiometry of a reaction: ``Jump::Flat`` or ``Jump::Sparse`` (src/gillespie.rs:93-100)."""

    diffs: tuple
    sparse: bool = False

    @staticmethod
    def new(differences: Sequence[int]) -> "Jump":
        return Jump(tuple(int(d) for d in differences), False)

    @staticmethod
    def new_sparse(pairs: Sequence[tuple[int, int]]) -> "Jump":
        return Jump(tuple((int(i), int(d)) for i, d in pairs), True)

    def dense(self) -> "Jump":
        """``Jump::dense`` (src/gillespie.rs:112-124)."""
        if not self.sparse:
            return self
        n = max((i + 1 for i, _ in self.diffs), default=0)
        d = [0] * n
        for i, v in self.diffs:
            d[i] += v
        return Jump(tuple(d), False)

    def sparse_form(self) -> "Jump":
        """``Jump::sparse`` (src/gillespie.rs:126-140): (index, diff) for diff != 0."""
        if self.sparse:
            return self
        return Jump(tuple((i, d) for i, d in enumerate(self.diffs) if d != 0), True)


class GillespieNet:
    """``gillespie::Gillespie`` (src/gillespie.rs:157-163) on the device.

    ``species`` is the initial state every trajectory starts from unless a
    per-trajectory ``x0`` is passed to ``run``/``batch``.
    """

    def __init__(self, species: Sequence[int], sparse: bool = False, seed: int | None = None):
        lib = load()
        x = np.ascontiguousarray(np.asarray(species, dtype=np.int64).reshape(-1))
        h = c_void_p()
        if seed is None:
            check(lib.rebop_gillespie_new(_i64(x), len(x), int(bool(sparse)), ctypes.byref(h)))
        else:
            check(
                lib.rebop_gillespie_new_with_seed(
                    _i64(x), len(x), int(bool(sparse)), int(seed) & (2**64 - 1), ctypes.byref(h)
                )
            )
        self._h = h
        self._lib = lib

    @classmethod
    def new_with_seed(cls, species: Sequence[int], sparse: bool, seed: int) -> "GillespieNet":
        """``Gillespie::new_with_seed`` -- src/gillespie.rs:179-187."""
        return cls(species, sparse, seed)

    @classmethod
    def system(cls, species: Sequence[int], seed: int) -> "GillespieNet":
        """A network with ``define_system!``'s arithmetic (src/gillespie_macro.rs:49-170):
        integer mass-action products and the first-match choice.  Reactions
        are added with ``add_system_reaction``."""
        self = cls.__new__(cls)
        lib = load()
        x = np.ascontiguousarray(np.asarray(species, dtype=np.int64).reshape(-1))
        h = c_void_p()
        check(lib.rebop_system_new(_i64(x), len(x), int(seed) & (2**64 - 1), ctypes.byref(h)))
        self._h = h
        self._lib = lib
        return self

    def add_system_reaction(self, rate: float, reactants: Sequence[tuple[int, int]],
                            products: Sequence[tuple[int, int]]) -> None:
        """One ``define_system!`` reaction ``m1 r1 + ... => p1 q1 + ... @ rate``
        (src/gillespie_macro.rs:55-58): (species, coefficient) terms in written order."""
        ri = np.ascontiguousarray([t[0] for t in reactants], dtype=np.uint32)
        rm = np.ascontiguousarray([t[1] for t in reactants], dtype=np.uint32)
        pi = np.ascontiguousarray([t[0] for t in products], dtype=np.uint32)
        pm = np.ascontiguousarray([t[1] for t in products], dtype=np.uint32)
        check(self._lib.rebop_system_add_reaction(self._h, float(rate), _u32(ri), _u32(rm), len(ri),
                                                  _u32(pi), _u32(pm), len(pi)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.rebop_gillespie_free(h)
            self._h = None

    @property
    def handle(self) -> c_void_p:
        return self._h

    def seed(self, seed: int) -> None:
        """``Gillespie::seed`` -- src/gillespie.rs:189-191."""
        check(self._lib.rebop_gillespie_seed(self._h, int(seed) & (2**64 - 1)))

    def set_draw_cap(self, max_draws: int) -> None:
        """Watchdog (no counterpart in the reference): later runs stop a
        trajectory after ``max_draws`` random draws in one call and raise
        ``RebopCapError``; 0 removes the cap (runs to a horizon that is not
        finite then use 2^20)."""
        check(self._lib.rebop_gillespie_set_draw_cap(self._h, int(max_draws)))

    def nb_species(self) -> int:
        return int(self._lib.rebop_gillespie_nb_species(self._h))

    def nb_reactions(self) -> int:
        return int(self._lib.rebop_gillespie_nb_reactions(self._h))

    def add_reaction(self, rate: "Rate | RateExpr", differences: Sequence[int]) -> None:
        """``Gillespie::add_reaction`` -- src/gillespie.rs:225-244."""
        d = np.ascontiguousarray(np.asarray(differences, dtype=np.int64).reshape(-1))
        if isinstance(rate, RateExpr):
            code = np.ascontiguousarray(rate.code, dtype=np.uint32)
            consts = np.ascontiguousarray(rate.consts, dtype=np.float64)
            check(
                self._lib.rebop_gillespie_add_reaction_expr(
                    self._h, _u32(code), len(code), _f64(consts), len(consts), _i64(d), len(d)
                )
            )
        elif rate.sparse:
            idx = np.ascontiguousarray([p[0] for p in rate.orders], dtype=np.uint32)
            orde = np.ascontiguousarray([p[1] for p in rate.orders], dtype=np.uint32)
            check(
                self._lib.rebop_gillespie_add_reaction_lma_sparse(
                    self._h, rate.rate, _u32(idx), _u32(orde), len(idx), _i64(d), len(d)
                )
            )
        else:
            orde = np.ascontiguousarray(rate.orders, dtype=np.uint32)
            check(
                self._lib.rebop_gillespie_add_reaction_lma(
                    self._h, rate.rate, _u32(orde), len(orde), _i64(d), len(d)
                )
            )

    def prepare(self, kernel: int = _native.REBOP_KERNEL_PERSISTENT | _native.REBOP_KERNEL_SPECIALIZED) -> None:
        """Compile (and on a device, load) the kernel a run will use."""
        check(self._lib.rebop_gillespie_prepare(self._h, int(kernel)))

    def autotune(self, tmax: float, nb_steps: int, n_pilot: int = 65536) -> "_native.TuneResult":
        """``rebop_autotune``: AUTO runs of this network then use the fastest variant."""
        res = _native.TuneResult()
        check(self._lib.rebop_autotune(self._h, float(tmax), int(nb_steps), int(n_pilot), ctypes.byref(res)))
        return res

    def get_species(self) -> np.ndarray:
        out = np.zeros(self.nb_species(), dtype=np.int64)
        check(self._lib.rebop_gillespie_get_species(self._h, _i64(out), len(out)))
        return out

    def set_species(self, species: Sequence[int]) -> None:
        x = np.ascontiguousarray(np.asarray(species, dtype=np.int64).reshape(-1))
        check(self._lib.rebop_gillespie_set_species(self._h, _i64(x), len(x)))

    def run(
        self,
        tmax: float,
        nb_steps: int,
        n_traj: int = 1,
        traj_begin: int = 0,
        x0: np.ndarray | None = None,
        with_stats: bool = False,
    ):
        """The pyo3 grid loop (src/pyo3_gillespie.rs:197-208) for ``n_traj`` trajectories.

        Returns ``(times[nb_steps+1], counts[nb_steps+1, S, n_traj])`` (plus a
        ``Stats`` if ``with_stats``).  Host buffers.
        """
        S = self.nb_species()
        nb = int(nb_steps)
        out = np.zeros((nb + 1, S, int(n_traj)), dtype=np.int64)
        times = np.zeros(nb + 1, dtype=np.float64)
        st = Stats()
        xp = None
        if x0 is not None:
            x0 = np.ascontiguousarray(np.asarray(x0, dtype=np.int64).reshape(int(n_traj), S))
            xp = _i64(x0)
        check(
            self._lib.rebop_run(
                self._h, float(tmax), nb, int(traj_begin), int(n_traj), xp, _i64(out),
                _f64(times), ctypes.byref(st),
            )
        )
        if with_stats:
            return times, out, st
        return times, out

    def run_device(
        self,
        tmax: float,
        nb_steps: int,
        n_traj: int,
        out_ptr: int | None,
        traj_begin: int = 0,
        x0_ptr: int | None = None,
        final_ptr: int | None = None,
        kernel: int = _native.REBOP_KERNEL_AUTO,
        stream: int | None = None,
        with_stats: bool = True,
    ) -> Stats | None:
        """``rebop_run_device``: device pointers (e.g. ``torch.Tensor.data_ptr()``)."""
        st = Stats()
        check(
            self._lib.rebop_run_device(
                self._h, float(tmax), int(nb_steps), int(traj_begin), int(n_traj),
                c_void_p(x0_ptr) if x0_ptr else None,
                c_void_p(out_ptr) if out_ptr else None,
                c_void_p(final_ptr) if final_ptr else None,
                int(kernel),
                c_void_p(stream) if stream else None,
                ctypes.byref(st) if with_stats else None,
            )
        )
        return st if with_stats else None

    def run_all_events(self, tmax: float, n_traj: int = 1, traj_begin: int = 0):
        """The nb_steps == 0 loop (src/pyo3_gillespie.rs:209-223).

        Returns ``(times, states)`` with ``states[record, species]`` for one
        trajectory, or lists of them for ``n_traj > 1``.
        """
        S = self.nb_species()
        n = int(n_traj)
        counts = np.zeros(max(n, 1), dtype=np.uint64)
        cp = counts.ctypes.data_as(POINTER(ctypes.c_uint64))
        check(self._lib.rebop_run_all_events(self._h, float(tmax), int(traj_begin), n, cp, None, None, None))
        total = int(counts[:n].sum())
        times = np.zeros(max(total, 1), dtype=np.float64)
        states = np.zeros((max(total, 1), S), dtype=np.int64)
        check(self._lib.rebop_run_all_events(self._h, float(tmax), int(traj_begin), n, cp,
                                             _f64(times), _i64(states), None))
        offs = np.concatenate([[0], np.cumsum(counts[:n].astype(np.int64))])
        ts = [times[offs[j]:offs[j + 1]] for j in range(n)]
        xs = [states[offs[j]:offs[j + 1]] for j in range(n)]
        if n == 1:
            return ts[0], xs[0]
        return ts, xs

    def run_all_events_device(self, tmax: float, n_traj: int, d_counts: int, d_offsets: int | None = None,
                              d_times: int | None = None, d_states: int | None = None, traj_begin: int = 0,
                              stream=None) -> "_native.Stats":
        """The nb_steps == 0 loop on device buffers (rebop_run_all_events_device):
        pass 1 fills ``d_counts`` (device pointer, uint64 [n_traj]); pass 2, given
        ``d_offsets`` (its exclusive scan), writes the records to ``d_times`` and
        ``d_states`` [total][S]."""
        st = _native.Stats()
        check(self._lib.rebop_run_all_events_device(self._h, float(tmax), int(traj_begin), int(n_traj), d_counts,
                                                    d_offsets, d_times, d_states, stream, ctypes.byref(st)))
        return st

    def run_sweep_device(self, n_points: int, traj_per_point: int, tmax: float, nb_steps: int,
                         out_ptr: int | None, rates_ptr: int | None = None,
                         x0_ptr: int | None = None, final_ptr: int | None = None,
                         traj_begin: int = 0, kernel: int = _native.REBOP_KERNEL_AUTO,
                         stream: int | None = None, with_stats: bool = True) -> Stats | None:
        """``rebop_run_sweep_device``: per-point rate constants / initial counts."""
        st = Stats()
        check(
            self._lib.rebop_run_sweep_device(
                self._h, int(n_points), int(traj_per_point),
                c_void_p(rates_ptr) if rates_ptr else None,
                c_void_p(x0_ptr) if x0_ptr else None,
                float(tmax), int(nb_steps), int(traj_begin),
                c_void_p(out_ptr) if out_ptr else None,
                c_void_p(final_ptr) if final_ptr else None,
                int(kernel), c_void_p(stream) if stream else None,
                ctypes.byref(st) if with_stats else None,
            )
        )
        return st if with_stats else None

    def batch(self, n_traj: int, traj_begin: int = 0, x0: np.ndarray | None = None) -> "Batch":
        return Batch(self, n_traj, traj_begin, x0)


class Batch:
    """N resumable trajectories of one network (device-resident state)."""

    def __init__(self, net: GillespieNet, n_traj: int, traj_begin: int = 0, x0=None):
        self._net = net
        self._lib = net._lib
        self.n = int(n_traj)
        self.S = net.nb_species()
        h = c_void_p()
        xp = None
        if x0 is not None:
            x0 = np.ascontiguousarray(np.asarray(x0, dtype=np.int64).reshape(self.n, self.S))
            xp = _i64(x0)
        check(self._lib.rebop_batch_new(net.handle, int(traj_begin), self.n, xp, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.rebop_batch_free(h)
            self._h = None

    def advance_until(self, tmax: float) -> Stats:
        """``Gillespie::advance_until`` on every trajectory -- src/gillespie.rs:315-344."""
        st = Stats()
        check(self._lib.rebop_batch_advance_until(self._h, float(tmax), ctypes.byref(st)))
        return st

    def get_species(self) -> np.ndarray:
        out = np.zeros((self.n, self.S), dtype=np.int64)
        check(self._lib.rebop_batch_get_species(self._h, _i64(out)))
        return out

    def set_species(self, species) -> None:
        x = np.ascontiguousarray(np.asarray(species, dtype=np.int64).reshape(self.n, self.S))
        check(self._lib.rebop_batch_set_species(self._h, _i64(x)))

    def get_time(self) -> np.ndarray:
        out = np.zeros(self.n, dtype=np.float64)
        check(self._lib.rebop_batch_get_time(self._h, _f64(out)))
        return out

    def set_time(self, t) -> None:
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(t, dtype=np.float64), (self.n,)))
        check(self._lib.rebop_batch_set_time(self._h, _f64(a)))

    def get_draws(self) -> np.ndarray:
        """Draws made so far by every trajectory (its random stream's position)."""
        out = np.zeros(self.n, dtype=np.uint64)
        check(self._lib.rebop_batch_get_draws(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        return out

    def set_draws(self, draws) -> None:
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(draws, dtype=np.uint64), (self.n,)))
        check(self._lib.rebop_batch_set_draws(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))


def run_jobs(jobs, stream: int | None = None, per_job_events: bool = False):
    """``rebop_run_jobs_device``: many networks' grid runs in one launch.

    ``jobs`` is a sequence of dicts with keys ``net`` (GillespieNet), ``tmax``,
    ``nb_steps``, ``n_traj`` and optionally ``traj_begin``, ``out_ptr``,
    ``x0_ptr``, ``final_ptr`` (device pointers) and ``cost_hint``.  Returns
    the ``Stats``, plus each job's event count if ``per_job_events``.
    """
    lib = load()
    arr = (_native.Job * len(jobs))()
    keep = []
    for a, j in zip(arr, jobs):
        net = j["net"]
        keep.append(net)
        a.net = net.handle.value
        a.tmax = float(j["tmax"])
        a.nb_steps = int(j["nb_steps"])
        a.traj_begin = int(j.get("traj_begin", 0))
        a.n_traj = int(j["n_traj"])
        a.d_x0 = j.get("x0_ptr") or None
        a.d_out = j.get("out_ptr") or None
        a.d_final = j.get("final_ptr") or None
        a.cost_hint = float(j.get("cost_hint", 0.0))
    st = Stats()
    ev = np.zeros(len(jobs), dtype=np.uint64)
    check(lib.rebop_run_jobs_device(arr, len(jobs), c_void_p(stream) if stream else None,
                                    ctypes.byref(st), ev.ctypes.data_as(POINTER(ctypes.c_uint64))
                                    if per_job_events else None))
    return (st, ev) if per_job_events else st
