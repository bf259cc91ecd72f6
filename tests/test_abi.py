"""CPU checks of the boundary: the C-ABI library loads, exports every entry
point include/rebop_gpu.h declares, validates arguments as the reference's
asserts do, and refuses to run without a gfx950 device (no silent fallback).
Also the host-side logic of the Python API (names, ordering, messages)."""

from __future__ import annotations

import ctypes
import json
import re
import warnings
from pathlib import Path

import numpy as np
import pytest

import rebop_amd
from rebop_amd import Gillespie, GillespieNet, Rate, _native
from rebop_amd.engine import Jump

ROOT = Path(__file__).resolve().parent.parent
TABLES = json.loads((ROOT / "tests" / "golden" / "reference_tables.json").read_text())


def _declared_symbols() -> set[str]:
    text = (ROOT / "include" / "rebop_gpu.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(rebop_[a-z0-9_]+)\s*\(", text))


def test_library_exports_every_declared_symbol():
    lib = _native.load()
    declared = _declared_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    # and the Python binding knows the signature of each
    assert declared == set(_native.SIGNATURES)


def test_surrogate_library_exports_every_declared_symbol():
    """include/rebop_surrogate.h: every declared entry point is exported by
    librebop_surrogate.so and has a signature in the Python binding."""
    from rebop_amd import surrogate

    text = (ROOT / "include" / "rebop_surrogate.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    declared = set(re.findall(r"\b(rebop_sg_[a-z0-9_]+)\s*\(", text))
    assert len(declared) >= 10
    lib = surrogate.load()
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(surrogate.SIGNATURES)


def _has_gpu() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


def test_device_calls_fail_loudly_without_gpu():
    if _has_gpu():
        pytest.skip("a GPU is present")
    g = GillespieNet([1], seed=1)
    g.add_reaction(Rate.lma(1.0, [1]), [-1])
    with pytest.raises(_native.RebopError) as e:
        g.run(1.0, 10, n_traj=4)
    assert e.value.code == _native.REBOP_ERR_DEVICE


def test_issue85_oob_rejected_at_the_boundary():
    """src/gillespie.rs:496-502: an out-of-range sparse species index is an error."""
    t = TABLES["issue85_oob"]
    g = GillespieNet(t["species"], sparse=True, seed=0)
    with pytest.raises(rebop_amd.RebopValueError) as e:
        g.add_reaction(Rate.lma_sparse(1.0, t["pairs"]), t["diffs"])
    assert e.value.code == _native.REBOP_ERR_INDEX
    assert g.nb_reactions() == 0


def test_draw_cap_setter_at_the_boundary():
    """The watchdog setter validates its handle and needs no device."""
    lib = _native.load()
    assert lib.rebop_gillespie_set_draw_cap(None, 5) == _native.REBOP_ERR_VALUE
    g = GillespieNet([1], seed=0)
    g.set_draw_cap(1000)
    g.set_draw_cap(0)
    st = _native.Stats()
    assert hasattr(st, "capped") and ctypes.sizeof(st) == 7 * 8


def test_shape_asserts_at_the_boundary():
    g = GillespieNet([1, 2, 3], seed=0)
    with pytest.raises(rebop_amd.RebopValueError) as e:
        g.add_reaction(Rate.lma(1.0, [1, 0]), [0, 0, 0])
    assert e.value.code == _native.REBOP_ERR_SHAPE
    with pytest.raises(rebop_amd.RebopValueError):
        g.add_reaction(Rate.lma(1.0, [1, 0, 0]), [0, 0])
    with pytest.raises(rebop_amd.RebopValueError):
        g.set_species([1, 2])
    with pytest.raises(rebop_amd.RebopValueError):
        g.add_reaction(Rate.lma(float("inf"), [1, 0, 0]), [0, 0, 0])
    with pytest.raises(rebop_amd.RebopValueError):
        g.add_reaction(Rate.lma(1.0, [1, 0, 0]), [2**21, 0, 0])
    g.add_reaction(Rate.lma(float("nan"), [1, 0, 0]), [0, 0, 0])  # NaN is accepted
    assert g.nb_reactions() == 1 and g.nb_species() == 3
    assert g.get_species().tolist() == [1, 2, 3]


def test_error_message_is_thread_local_text():
    lib = _native.load()
    g = GillespieNet([1], seed=0)
    o = np.array([1, 1], dtype=np.uint32)
    d = np.array([0], dtype=np.int64)
    rc = lib.rebop_gillespie_add_reaction_lma(
        g.handle, 1.0, o.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), 2,
        d.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 1)
    assert rc == _native.REBOP_ERR_SHAPE
    assert b"assertion" in lib.rebop_last_error()


# ---- Rate / Jump conversions (src/gillespie.rs:37-69, :112-140) -----------


def test_rate_sparse_table():
    t = TABLES["rate_sparse"]
    for dense, pairs in t["cases"]:
        assert Rate.lma(t["rate"], dense).sparse_form() == Rate.lma_sparse(t["rate"], [tuple(p) for p in pairs])


def test_rate_conversion_round_trips():
    t = TABLES["rate_conversion"]
    for dense, pairs in t["dense_to_sparse"]:
        r = Rate.lma(2.0, dense)
        assert r.dense() == r
        assert r.sparse_form().dense() == r
        assert r.sparse_form() == Rate.lma_sparse(2.0, [tuple(p) for p in pairs])
    for pairs, dense in t["sparse_to_dense"]:
        r = Rate.lma_sparse(2.0, [tuple(p) for p in pairs])
        assert r.sparse_form() == r
        assert r.dense() == Rate.lma(2.0, dense)
        assert r.dense().sparse_form() == r


def test_jump_conversion_table():
    t = TABLES["jump_conversion"]
    dj = Jump.new(t["dense"])
    sj = Jump.new_sparse([tuple(p) for p in t["sparse"]])
    assert dj.dense() == dj and dj.sparse_form() == sj
    assert sj.sparse_form() == sj and sj.dense() == dj


# ---- Python API host logic (pyo3_gillespie.rs, python/rebop/gillespie.py) --


def test_species_order_and_str():
    s = Gillespie()
    s.add_reaction(14, [], ["A"])
    s.add_reaction(0.1, ["A", "B"], ["C"], 0.01)
    assert s.nb_species() == 3 and s.nb_reactions() == 3
    assert list(s._species) == ["A", "B", "C"]
    assert str(s) == (
        "3 species and 3 reactions\n"
        " --> A @ LMA(14)\n"
        "A + B --> C @ LMA(0.1)\n"
        "C --> A + B @ LMA(0.01)\n"
    )


def test_set_init_warning_and_unknown_species():
    s = Gillespie()
    s.add_reaction(1.0, ["A"], ["B"])
    with pytest.raises(UserWarning, match="species are not involved in any reaction"):
        s.set_init({"Z": 3})
    assert "Z" in s._species


def test_parameter_species_clash_message():
    s = Gillespie()
    s.add_reaction(4, ["A"], ["B"])
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        with pytest.raises(ValueError, match="Species B cannot also be a parameter"):
            s.run({}, 10, 10, params={"B": 4.2})


def test_seed_derivation_matches_reference_wrapper():
    """gillespie.py:139-140: seed = default_rng(rng).integers(2**64 - 1, dtype=uint64)."""
    assert TABLES["test_fixed_seed"]["numpy_seed"] == str(
        np.random.default_rng(42).integers(np.iinfo(np.uint64).max, dtype=np.uint64)
    )
