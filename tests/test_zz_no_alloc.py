"""rebop_run_device_noalloc (no counterpart in the reference; for callers that
cannot allocate device memory) against rebop_run_device, bit for bit, through
the C ABI: two calls in a row on each kernel family, the second smaller, so a
stale counter, sample-time table or work table left in the shared scratch
region by the first call would show."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from rebop_amd import _native
from tests.networks import build_device, dimers, sir, synthetic

pytestmark = pytest.mark.gpu
K = _native

CASES = [
    # (name, spec, tmax, nb_steps, kernel, (n1, begin1), (n2, begin2))
    ("dimers-auto", dimers(), 1.0, 100, K.REBOP_KERNEL_AUTO, (20000, 0), (3000, 777)),
    ("sir-grouped", sir(), 250.0, 250, K.REBOP_KERNEL_AUTO, (4096, 5), (1000, 90)),
    ("syn12x24-data-driven", synthetic(12, 24, 4, init_mean=20.0), 2.0, 40,
     K.REBOP_KERNEL_PERSISTENT, (5000, 0), (700, 11)),
    ("syn40x80-lds", synthetic(40, 80, 5, init_mean=25.0), 0.2, 30, K.REBOP_KERNEL_AUTO,
     (3000, 2), (500, 64)),
]


def _run(lib_fn, g, tmax, nb, n, begin, kernel):
    import torch

    S = g.nb_species()
    out = torch.full((nb + 1, S, n), -7, dtype=torch.int64, device="cuda")
    fin = torch.full((S, n), -7, dtype=torch.int64, device="cuda")
    st = _native.Stats()
    rc = lib_fn(g.handle, float(tmax), nb, begin, n, None, ctypes.c_void_p(out.data_ptr()),
                ctypes.c_void_p(fin.data_ptr()), kernel, None, ctypes.byref(st))
    torch.cuda.synchronize()
    return rc, out.cpu().numpy(), fin.cpu().numpy(), st


@pytest.mark.parametrize("name,spec,tmax,nb,kernel,first,second", CASES)
def test_noalloc_equals_run_device(name, spec, tmax, nb, kernel, first, second):
    lib = _native.load()
    g = build_device(spec, 2024)
    _native.check(lib.rebop_gillespie_prepare(g.handle, kernel))
    for n, begin in (first, second):
        rc, out, fin, st = _run(lib.rebop_run_device_noalloc, g, tmax, nb, n, begin, kernel)
        assert rc == 0, lib.rebop_last_error()
        rc2, out2, fin2, st2 = _run(lib.rebop_run_device, g, tmax, nb, n, begin, kernel)
        assert rc2 == 0
        assert np.array_equal(out, out2), name
        assert np.array_equal(fin, fin2), name
        assert (st.events, st.draws, st.trajectories) == (st2.events, st2.draws, st2.trajectories)
        assert st.trajectories == n and st.kernel_ms == 0.0


def test_noalloc_refuses_what_it_cannot_do_without_allocating():
    """An unprepared handle and a sample table beyond the scratch limit are
    errors reported before anything is launched."""
    lib = _native.load()
    g = build_device(dimers(), 1)
    rc, *_ = _run(lib.rebop_run_device_noalloc, g, 1.0, 10, 64, 0, K.REBOP_KERNEL_AUTO)
    assert rc == K.REBOP_ERR_VALUE and b"prepare" in lib.rebop_last_error()
    _native.check(lib.rebop_gillespie_prepare(g.handle, K.REBOP_KERNEL_AUTO))
    st = _native.Stats()
    rc = lib.rebop_run_device_noalloc(g.handle, 1.0, 70000, 0, 64, None, None, None,
                                      K.REBOP_KERNEL_AUTO, None, ctypes.byref(st))
    assert rc == K.REBOP_ERR_VALUE and b"65536" in lib.rebop_last_error()
