"""The caches that spare later calls and processes work, none of which may
change a result: the autotuner's on-disk choice and the specialised kernels'
on-disk code objects."""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from rebop_amd import _native
from tests.networks import build_device, build_oracle, dimers, sir

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent


def _dev(g, tmax, nb, n, kernel):
    import torch

    out = torch.zeros((nb + 1, g.nb_species(), n), dtype=torch.int64, device="cuda")
    st = g.run_device(tmax, nb, n, out.data_ptr(), kernel=kernel)
    torch.cuda.synchronize()
    return out.cpu().numpy(), st


def test_autotune_cache_reuses_the_choice(tmp_path, monkeypatch, nthreads):
    """The autotuner's on-disk cache: a second handle of the same network
    shape takes the cached choice without timing anything, and its AUTO run
    still equals the oracle."""
    monkeypatch.setenv("REBOP_TUNE_CACHE", str(tmp_path / "tune.txt"))
    monkeypatch.delenv("REBOP_TUNE_FORCE", raising=False)
    spec = sir()
    a = build_device(spec, 5)
    r1 = a.autotune(50.0, 20, n_pilot=3000)
    assert r1.n_tried >= 8
    b = build_device(spec, 6)
    r2 = b.autotune(50.0, 20, n_pilot=3000)
    assert r2.n_tried == 0
    assert (r2.kernel, r2.threads, r2.occupancy) == (r1.kernel, r1.threads, r1.occupancy)
    assert (tmp_path / "tune.txt").read_text().count("\n") == 1
    out, st = _dev(b, 50.0, 20, 3000, _native.REBOP_KERNEL_AUTO)
    assert st.kernel & ~_native.REBOP_KERNEL_COMPACT == r1.kernel
    o = build_oracle(spec)
    _, oout, _, oev, _ = o.run_grid(6, spec[2], 50.0, 20, n_traj=3000, nthreads=nthreads)
    assert np.array_equal(out, oout) and st.events == oev


CHILD = """
import sys
sys.path.insert(0, {root!r})
import numpy as np, torch
from rebop_amd import networks, _native
from oracle import pyoracle
spec = networks.dimers()
g = networks.build(spec, 77)
out = torch.zeros((11, 4, 2048), dtype=torch.int64, device="cuda")
st = g.run_device(0.2, 10, 2048, out.data_ptr(),
                  kernel=_native.REBOP_KERNEL_PERSISTENT | _native.REBOP_KERNEL_SPECIALIZED)
torch.cuda.synchronize()
o = pyoracle.OracleNet(spec[0])
for k, orders, d in spec[1]:
    o.add_lma(k, orders, d)
_, oout, _, oev, _ = o.run_grid(77, spec[2], 0.2, 10, n_traj=2048, nthreads=8)
assert np.array_equal(out.cpu().numpy(), oout) and st.events == oev
print("child ok")
"""


def test_specialised_kernel_disk_cache(tmp_path):
    """A specialised kernel compiled in one process is written to the disk
    cache (code object + its source identity); a second process loads it from
    there, and its runs still equal the oracle bit for bit."""
    env = dict(os.environ, REBOP_JIT_CACHE=str(tmp_path / "jit"))
    code = CHILD.format(root=str(ROOT))
    for round_ in range(2):
        r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                           timeout=240)
        assert r.returncode == 0 and "child ok" in r.stdout, (round_, r.stdout[-2000:], r.stderr[-2000:])
        cos = sorted((tmp_path / "jit").glob("*.co"))
        assert len(cos) >= 1  # written by the first process, reused by the second
        srcs = sorted((tmp_path / "jit").glob("*.src"))
        assert [c.stem for c in cos] == [s_.stem for s_ in srcs]
        assert all(s_.read_text().startswith("hiprtc ") for s_ in srcs)
