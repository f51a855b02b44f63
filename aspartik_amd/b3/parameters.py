"""Parameters: ``Real``, ``RealVector``, ``IntVector`` and ``ClassVector``.

They mirror b3's Parameter protocol (crates/b3/src/parameters/mod.rs):
``is_changed`` since the last ``accept``/``reject``, ``accept`` keeps the
edits, ``reject`` restores the last accepted values.  ``ClassVector``
follows parameters/class_vector.rs:12-97 over an EpochBuf
(crates/sk/src/epoch.rs): every ``set`` records its index (also when the
value does not change, and once per call), and ``changed_indices`` lists
them in call order until the next accept or reject.
"""
from __future__ import annotations

import numpy as np


class Real:
    """One real value (parameters/real.rs)."""

    def __init__(self, value: float):
        self._value = float(value)
        self._accepted = self._value
        self._changed = False

    @property
    def value(self) -> float:
        return self._value

    def set(self, value: float) -> None:
        self._value = float(value)
        self._changed = True

    def __float__(self) -> float:
        return self._value

    def is_changed(self) -> bool:
        return self._changed

    def accept(self) -> None:
        self._accepted = self._value
        self._changed = False

    def reject(self) -> None:
        self._value = self._accepted
        self._changed = False

    def __repr__(self):
        return f"Real({self._value})"


class _Vector:
    _dtype = np.float64

    def __init__(self, *values):
        if len(values) == 1 and np.ndim(values[0]) == 1:
            values = tuple(values[0])
        if not values:
            raise ValueError("a vector parameter must be non-empty")
        self._values = np.array(values, self._dtype)
        self._accepted = self._values.copy()
        self._indices: list[int] = []

    @classmethod
    def repeat(cls, value, length: int):
        return cls(*([value] * int(length)))

    def __len__(self) -> int:
        return self._values.size

    def __getitem__(self, index: int):
        return self._values[index].item()

    def __setitem__(self, index: int, value):
        self._values[index] = value
        self._indices.append(int(index))

    @property
    def values(self) -> np.ndarray:
        return self._values.copy()

    def changed_indices(self) -> list[int]:
        return list(self._indices)

    def is_changed(self) -> bool:
        return bool(self._indices)

    def accept(self) -> None:
        self._accepted[:] = self._values
        self._indices.clear()

    def reject(self) -> None:
        self._values[:] = self._accepted
        self._indices.clear()


class RealVector(_Vector):
    """parameters/real_vector.rs"""
    _dtype = np.float64


class IntVector(_Vector):
    """parameters/int_vector.rs"""
    _dtype = np.int64


class ClassVector:
    """``len`` class indices in ``0..num_classes`` (class_vector.rs:17-52)."""

    def __init__(self, num_classes: int, length: int, values=None):
        if length <= 0:
            raise ValueError("`ClassVector` must be non-empty")
        if num_classes <= 1:
            raise ValueError("`ClassVector` must have at least 2 classes")
        self._k = int(num_classes)
        self._values = np.zeros(int(length), np.uint32)
        if values is not None:
            v = np.asarray(values, np.int64).reshape(-1)
            if v.size != length or v.min() < 0 or v.max() >= num_classes:
                raise ValueError("initial classes out of range")
            self._values[:] = v
        self._accepted = self._values.copy()
        self._indices: list[int] = []

    @property
    def num_classes(self) -> int:
        return self._k

    def __len__(self) -> int:
        return self._values.size

    def __getitem__(self, index: int) -> int:
        return int(self._values[index])

    def set(self, index: int, cls: int) -> None:
        """class_vector.rs:39-42 (asserts the class is in range)."""
        if not 0 <= int(cls) < self._k:
            raise ValueError(f"class {cls} out of range 0..{self._k}")
        self._values[index] = cls
        self._indices.append(int(index))

    @property
    def values(self) -> np.ndarray:
        """The current classes (a copy)."""
        return self._values.copy()

    def changed_indices(self) -> list[int]:
        return list(self._indices)

    def is_changed(self) -> bool:
        return bool(self._indices)

    def accept(self) -> None:
        self._accepted[self._indices] = self._values[self._indices]
        self._indices.clear()

    def reject(self) -> None:
        self._values[self._indices] = self._accepted[self._indices]
        self._indices.clear()

    def __repr__(self):
        return f"ClassVector(num_classes={self._k}, len={len(self)})"
