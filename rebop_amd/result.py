"""Result container of ``Gillespie.run``.

The reference returns an ``xarray.Dataset`` with a ``time`` coordinate and one
variable per species (python/rebop/gillespie.py:153-159).  When xarray is
importable the same object is built here; otherwise ``Dataset`` offers the
access pattern the reference's tests use (``ds.time``, ``ds.S``, ``ds["S"]``,
``name in ds``, iteration over variable names).
"""

from __future__ import annotations

import numpy as np

try:  # optional, as in the reference's runtime dependencies
    import xarray as _xr
except Exception:  # pragma: no cover - depends on the environment
    _xr = None


class Dataset:
    """Minimal stand-in for ``xarray.Dataset`` of sampled counts."""

    def __init__(self, times: np.ndarray, data: dict[str, np.ndarray], dims: tuple[str, ...]):
        self.time = np.asarray(times)
        self.data_vars = dict(data)
        self.dims = dims
        self.coords = {"time": self.time}

    def __getattr__(self, name: str):
        data_vars = self.__dict__.get("data_vars", {})
        if name in data_vars:
            return data_vars[name]
        raise AttributeError(name)

    def __getitem__(self, key):
        if isinstance(key, (list, tuple)):
            return Dataset(self.time, {k: self.data_vars[k] for k in key}, self.dims)
        if key == "time":
            return self.time
        return self.data_vars[key]

    def __contains__(self, key) -> bool:
        return key in self.data_vars

    def __iter__(self):
        return iter(self.data_vars)

    def __len__(self) -> int:
        return len(self.data_vars)

    def __eq__(self, other) -> bool:  # equality of all variables and times
        if not isinstance(other, Dataset):
            return NotImplemented
        return (
            set(self.data_vars) == set(other.data_vars)
            and np.array_equal(self.time, other.time)
            and all(np.array_equal(self.data_vars[k], other.data_vars[k]) for k in self.data_vars)
        )

    def __repr__(self) -> str:
        return f"<rebop_amd.Dataset dims={self.dims} time={len(self.time)} vars={list(self.data_vars)}>"


def make_dataset(times, data: dict[str, np.ndarray], batched: bool):
    dims = ("time", "trajectory") if batched else ("time",)
    if _xr is not None:
        return _xr.Dataset(
            data_vars={name: _xr.DataArray(v, dims=dims, coords={"time": times}) for name, v in data.items()},
            coords={"time": times},
        )
    return Dataset(times, data, dims)
