"""Fields, scalar constants and option handles of the drop-in API.

Public names and constructor signatures follow glide/field.py (Field,
Constant, SubgridField, GridEntity, LocalOption, BroadcastOption; lines
12-16, 41-75, 212-248, 250-286) because grid.py, multigrid.py and user code
construct and query them by those names.  The implementation is this
package's own: a field is a thin view over one float64 torch tensor on the
ROCm device whose gradient buffer is created on first access, and an option
handle is a (read, write) pair over solver configuration objects.
"""
from __future__ import annotations

from enum import Enum
from typing import Any, Callable, Sequence

import numpy as np
import torch


class GridEntity(str, Enum):
    """Where a field lives on the MAC grid (the reference's string values)."""
    CELL = "cell"
    VERTICAL_FACET = "vertical_facet"      # u: (ny, nx + 1)
    HORIZONTAL_FACET = "horizontal_facet"  # v: (ny + 1, nx)


# Cell average of a facet array: the two slices whose mean gives the cell value.
_CELL_AVERAGE = {
    GridEntity.VERTICAL_FACET: (lambda a: a[:, :-1], lambda a: a[:, 1:]),
    GridEntity.HORIZONTAL_FACET: (lambda a: a[:-1], lambda a: a[1:]),
}


def _as_tensor_like(value, ref: torch.Tensor) -> torch.Tensor:
    t = value if isinstance(value, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(value))
    return t.to(device=ref.device, dtype=ref.dtype)


class Field:
    """One gridded array plus its metadata and a lazily created gradient."""

    __slots__ = ("data", "grid_entity", "dx", "grid", "name", "units", "attrs", "_grad")

    def __init__(self, data: torch.Tensor, grid_entity: GridEntity, dx: float, grid: Any = None,
                 name: str | None = None, units: str | None = None, attrs: dict | None = None,
                 _grad: torch.Tensor | None = None):
        self.data = data
        self.grid_entity = GridEntity(grid_entity)
        self.dx = dx
        self.grid = grid
        self.name = name
        self.units = units
        self.attrs = {} if attrs is None else attrs
        self._grad = _grad

    # values -------------------------------------------------------------
    def set(self, value) -> None:
        """In-place assignment: arrays must match the field's shape, scalars fill."""
        if not hasattr(value, "shape"):
            self.data.fill_(float(value))
            return
        shape = tuple(value.shape)
        if shape != tuple(self.data.shape):
            raise ValueError(f"cannot set field {self.name!r} of shape {tuple(self.data.shape)} from shape {shape}")
        self.data.copy_(_as_tensor_like(value, self.data))

    def zero(self) -> None:
        self.data.zero_()

    # gradient -----------------------------------------------------------
    @property
    def grad(self) -> torch.Tensor:
        if self._grad is None:
            self._grad = torch.zeros_like(self.data)
        return self._grad

    def has_grad(self) -> bool:
        return self._grad is not None

    def zero_grad(self) -> None:
        if self._grad is not None:
            self._grad.zero_()

    # conversions --------------------------------------------------------
    def to_cell(self) -> "Field":
        """This field on cell centres (facet fields averaged over the two faces)."""
        pick = _CELL_AVERAGE.get(self.grid_entity)
        if pick is None:
            return self
        lo, hi = pick
        return Field(0.5 * (lo(self.data) + hi(self.data)), GridEntity.CELL, self.dx, grid=self.grid,
                     units=self.units, attrs=self.attrs)

    @property
    def compact_string(self) -> str:
        ny, nx = self.data.shape[:2]
        return f"{self.name} [{self.units or '-'}] {self.grid_entity.value} {ny}x{nx}"

    def __repr__(self) -> str:
        return (f"Field({self.name!r}, entity={self.grid_entity.value}, shape={tuple(self.data.shape)}, "
                f"dtype={self.data.dtype}, units={self.units!r}, device={self.data.device})")


class SubgridField(Field):
    """A cell field with optional sub-grid quantiles (the bed of glide/grid.py:31)."""

    __slots__ = ("quantiles",)

    def __init__(self, *args, quantiles: Any = None, **kwargs):
        super().__init__(*args, **kwargs)
        self.quantiles = quantiles


class Constant:
    """A scalar physical parameter with a scalar gradient slot."""

    __slots__ = ("value", "name", "units", "attrs", "_grad")

    def __init__(self, value: float, name: str | None = None, units: str | None = None, attrs: dict | None = None,
                 _grad: float | None = None):
        self.value = float(value)
        self.name = name
        self.units = units
        self.attrs = {} if attrs is None else attrs
        self._grad = _grad

    def set(self, value) -> None:
        self.value = float(value)

    @property
    def grad(self) -> float:
        if self._grad is None:
            self._grad = 0.0
        return self._grad

    def has_grad(self) -> bool:
        return self._grad is not None

    def zero_grad(self) -> None:
        if self._grad is not None:
            self._grad = 0.0

    def __repr__(self) -> str:
        return f"{self.name} = {self.value:g} {self.units or ''}".rstrip()


class _OptionHandle:
    """A named (read, write) pair; prints as ``name=value``."""

    def __init__(self, label: str, read: Callable[[], Any], write: Callable[[Any], None]):
        self._label, self._read, self._write = label, read, write

    def get(self):
        return self._read()

    def set(self, value) -> None:
        self._write(value)

    def __repr__(self) -> str:
        return f"{self._label}={self.get()!r}"


class LocalOption(_OptionHandle):
    """An option of one configuration object (glide/field.py:250-265 signature)."""

    def __init__(self, getter: Callable[[], Any], setter: Callable[[Any], None], name: str):
        super().__init__(name, getter, setter)


class BroadcastOption(_OptionHandle):
    """An attribute shared by the per-level configurations of a hierarchy:
    reads the finest level's value, writes every level's (glide/field.py:268-286
    signature)."""

    def __init__(self, levels: Sequence, getter: Callable[[Any], Any], attr_name: str):
        configs = lambda: [getter(lev) for lev in levels]  # noqa: E731 (re-resolved on every access)

        def write(value):
            for cfg in configs():
                setattr(cfg, attr_name, value)

        super().__init__(attr_name, lambda: getattr(configs()[0], attr_name), write)
