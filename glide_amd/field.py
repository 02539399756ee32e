"""Fields, constants and option wrappers of the drop-in API.

Mirrors glide/field.py: a :class:`Field` owns one array (here a torch tensor
on the ROCm device), knows which MAC-grid entity it lives on, sets values in
place and allocates its gradient lazily (glide/field.py:41-75); a
:class:`Constant` is one scalar parameter (glide/field.py:212-248);
:class:`LocalOption` / :class:`BroadcastOption` wrap solver configs
(glide/field.py:254-286).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from enum import Enum
from typing import Any

import numpy as np
import torch


class GridEntity(str, Enum):
    CELL = "cell"
    VERTICAL_FACET = "vertical_facet"
    HORIZONTAL_FACET = "horizontal_facet"


@dataclass
class Field:
    data: torch.Tensor
    grid_entity: GridEntity
    dx: float
    grid: Any = field(default=None, repr=False, compare=False)
    name: str | None = None
    units: str | None = None
    attrs: dict = field(default_factory=dict)
    _grad: torch.Tensor | None = None

    def set(self, value) -> None:
        """Copy `value` into the field in place (scalars fill)."""
        if hasattr(value, "shape"):
            t = torch.as_tensor(np.asarray(value) if not isinstance(value, torch.Tensor) else value)
            if tuple(t.shape) != tuple(self.data.shape):
                raise ValueError(f"{self.name}: shape {tuple(t.shape)} != {tuple(self.data.shape)}")
            self.data.copy_(t.to(device=self.data.device, dtype=self.data.dtype))
        else:
            self.data.fill_(float(value))

    def zero(self) -> None:
        self.data.zero_()

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

    def to_cell(self) -> "Field":
        if self.grid_entity is GridEntity.CELL:
            return self
        if self.grid_entity is GridEntity.VERTICAL_FACET:
            d = 0.5 * (self.data[:, 1:] + self.data[:, :-1])
        else:
            d = 0.5 * (self.data[1:] + self.data[:-1])
        return Field(data=d, grid_entity=GridEntity.CELL, dx=self.dx, grid=self.grid, units=self.units,
                     attrs=self.attrs)

    @property
    def compact_string(self) -> str:
        return f"Field: {self.name}, {self.units}, ({self.data.shape[0]}, {self.data.shape[1]})"

    def __repr__(self) -> str:
        return f"Field: {self.name}\n{self.data}\n{tuple(self.data.shape)}, {self.data.dtype}, {self.units}"


@dataclass
class SubgridField(Field):
    quantiles: Any | None = None


@dataclass
class Constant:
    value: float
    name: str | None = None
    units: str | None = None
    attrs: dict = field(default_factory=dict)
    _grad: Any | None = None

    def set(self, value) -> None:
        self.value = float(value)

    @property
    def grad(self):
        if self._grad is None:
            self._grad = 0.0
        return self._grad

    def has_grad(self) -> bool:
        return self._grad is not None

    def zero_grad(self) -> None:
        if self._grad is not None:
            self._grad = 0.0

    def __repr__(self) -> str:
        return f"Constant: {self.name}, {self.value:.3f}, {self.units}"


class LocalOption:
    def __init__(self, getter, setter, name: str):
        self._getter = getter
        self._setter = setter
        self._name = name

    def get(self):
        return self._getter()

    def set(self, value):
        self._setter(value)

    def __repr__(self):
        return f"{self._name}={self.get()!r}"


class BroadcastOption:
    def __init__(self, levels, getter, attr_name: str):
        self._levels = levels
        self._getter = getter
        self._attr_name = attr_name

    def get(self):
        return getattr(self._getter(self._levels[0]), self._attr_name)

    def set(self, value):
        for lev in self._levels:
            setattr(self._getter(lev), self._attr_name, value)

    def __repr__(self):
        return f"{self._attr_name}={self.get()!r}"
