"""Packed-Hermitian storage of a density matrix.

A Hermitian ``n x n`` matrix is stored as ``n*n`` float64 values: the real
diagonal first, then the ``(re, im)`` pairs of the strict upper triangle in
row-major order.  This is the layout of
``centrex_tlf/lindblad/state_layout.py:11-79`` and
``rust/src/lindblad/layout.rs:22-85``; pack/unpack here are vectorised with
index arrays so a 4096-trajectory batch packs without a Python loop per
element.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import numpy.typing as npt

__all__ = ["PackedHermitianLayout", "upper_offsets"]


def upper_offsets(n: int) -> np.ndarray:
    """Start index of each row in the row-major upper triangle (diagonal included).

    ``upper_index(i, j) = offsets[i] + (j - i)`` for ``i <= j``
    (``centrex_tlf/lindblad/ir.py:455-467``).
    """
    rows = np.arange(n, dtype=np.int64)
    return rows * n - rows * (rows - 1) // 2


@dataclass(frozen=True)
class PackedHermitianLayout:
    n: int
    _iu: tuple = field(init=False, repr=False, compare=False)

    def __post_init__(self) -> None:
        if self.n <= 0:
            raise ValueError("layout dimension must be positive")
        iu = np.triu_indices(self.n, k=1)
        object.__setattr__(self, "_iu", (iu[0].astype(np.int64), iu[1].astype(np.int64)))

    @property
    def packed_len(self) -> int:
        return self.n * self.n

    @property
    def upper_len(self) -> int:
        return self.n * (self.n + 1) // 2

    def diagonal_index(self, i: int) -> int:
        self._check_index(i)
        return i

    def upper_real_index(self, i: int, j: int) -> int:
        self._check_pair(i, j)
        return self.n + 2 * self._upper_offset(i, j)

    def upper_imag_index(self, i: int, j: int) -> int:
        self._check_pair(i, j)
        return self.n + 2 * self._upper_offset(i, j) + 1

    def pack(self, matrix: npt.ArrayLike) -> npt.NDArray[np.float64]:
        arr = np.asarray(matrix, dtype=np.complex128)
        if arr.shape != (self.n, self.n):
            raise ValueError(f"expected matrix shape {(self.n, self.n)}, got {arr.shape}")
        return self.pack_batch(arr[None])[0]

    def pack_batch(self, matrices: npt.ArrayLike) -> npt.NDArray[np.float64]:
        arr = np.asarray(matrices, dtype=np.complex128)
        if arr.ndim != 3 or arr.shape[1:] != (self.n, self.n):
            raise ValueError(
                f"expected matrices of shape (B, {self.n}, {self.n}), got {arr.shape}"
            )
        out = np.empty((arr.shape[0], self.packed_len), dtype=np.float64)
        out[:, : self.n] = np.real(np.diagonal(arr, axis1=1, axis2=2))
        upper = arr[:, self._iu[0], self._iu[1]]
        out[:, self.n :: 2] = upper.real
        out[:, self.n + 1 :: 2] = upper.imag
        return out

    def unpack(self, packed: npt.ArrayLike) -> npt.NDArray[np.complex128]:
        arr = np.asarray(packed, dtype=np.float64)
        if arr.shape != (self.packed_len,):
            raise ValueError(f"expected packed vector length {self.packed_len}, got {arr.shape}")
        return self.unpack_batch(arr[None])[0]

    def unpack_batch(self, packed: npt.ArrayLike) -> npt.NDArray[np.complex128]:
        arr = np.asarray(packed, dtype=np.float64)
        if arr.ndim != 2 or arr.shape[1] != self.packed_len:
            raise ValueError(f"expected packed batch of shape (B, {self.packed_len}), got {arr.shape}")
        out = np.zeros((arr.shape[0], self.n, self.n), dtype=np.complex128)
        idx = np.arange(self.n)
        out[:, idx, idx] = arr[:, : self.n]
        upper = arr[:, self.n :: 2] + 1j * arr[:, self.n + 1 :: 2]
        out[:, self._iu[0], self._iu[1]] = upper
        out[:, self._iu[1], self._iu[0]] = np.conjugate(upper)
        return out

    def _upper_offset(self, i: int, j: int) -> int:
        # Offset of (i, j), i < j, among the strict upper triangle in row-major order.
        return i * (self.n - 1) - i * (i - 1) // 2 + (j - i - 1)

    def _check_index(self, i: int) -> None:
        if not (0 <= i < self.n):
            raise IndexError(f"index {i} out of bounds for size {self.n}")

    def _check_pair(self, i: int, j: int) -> None:
        self._check_index(i)
        self._check_index(j)
        if not i < j:
            raise ValueError(f"expected upper-triangle pair with i < j, got {(i, j)}")
