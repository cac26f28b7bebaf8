"""Column-major device matrix mirroring elem::Matrix<T>
(``include/elemental/core/matrix.hpp:16-183``): a buffer, a height, a width
and a leading dimension, with element (i, j) at ``i + j*ldim``
(``src/core/matrix.cpp:228-240``).  Views share the owner's buffer
(``matrix.cpp:101-125``).  Storage is a flat torch tensor on the GPU."""
from __future__ import annotations

import ctypes

import numpy as np


class Matrix:
    def __init__(self, height: int, width: int, ldim: int | None = None, dtype=None, device="cuda",
                 _storage=None, _offset: int = 0):
        import torch

        self._h, self._w = int(height), int(width)
        self._ld = int(ldim) if ldim is not None else max(self._h, 1)
        if self._ld < max(self._h, 1):
            raise ValueError("Leading dimension must be no less than height")
        if _storage is None:
            dtype = dtype or torch.float64
            # like Memory<T>::Require the buffer is not zeroed by contract;
            # zero it anyway so unused padding is deterministic
            _storage = torch.zeros(max(self._ld * max(self._w, 1), 1), dtype=dtype, device=device)
        self._st = _storage
        self._off = int(_offset)

    # -- basic information (matrix.hpp:40-55)
    def Height(self) -> int:
        return self._h

    def Width(self) -> int:
        return self._w

    def LDim(self) -> int:
        return self._ld

    def DiagonalLength(self, offset: int = 0) -> int:
        if offset > 0:
            return max(0, min(self._h, self._w - offset))
        return max(0, min(self._h + offset, self._w))

    @property
    def dtype(self):
        return self._st.dtype

    def ptr(self) -> ctypes.c_void_p:
        return ctypes.c_void_p(self._st.data_ptr() + self._off * self._st.element_size())

    def Buffer(self) -> ctypes.c_void_p:
        return self.ptr()

    def tensor(self):
        """The (height x width) column-major view as a torch tensor."""
        import torch

        return torch.as_strided(self._st, (self._h, self._w), (1, self._ld), self._off)

    def diagonal(self):
        return self.tensor().diagonal()

    # -- host transfer
    @classmethod
    def from_numpy(cls, a: np.ndarray, ldim: int | None = None, device="cuda") -> "Matrix":
        import torch

        a = np.asarray(a)
        if a.ndim == 1:
            a = a.reshape(-1, 1)
        h, w = a.shape
        M = cls(h, w, ldim, dtype=torch.from_numpy(np.zeros(1, dtype=a.dtype)).dtype, device=device)
        M.tensor().copy_(torch.from_numpy(np.ascontiguousarray(a)).to(device))
        return M

    def to_numpy(self) -> np.ndarray:
        return np.asfortranarray(self.tensor().cpu().numpy())

    def Get(self, i: int, j: int) -> float:
        return float(self._st[self._off + i + j * self._ld])

    def Set(self, i: int, j: int, v: float) -> None:
        self._st[self._off + i + j * self._ld] = v

    def __repr__(self):
        return f"Matrix({self._h}x{self._w}, ldim={self._ld}, {self.dtype})"


def View(A: Matrix, i: int, j: int, height: int, width: int) -> Matrix:
    """elem::View(A, B, i, j, height, width): a submatrix sharing A's buffer."""
    if i < 0 or j < 0 or i + height > A.Height() or j + width > A.Width():
        raise ValueError("Trying to view outside of a Matrix")
    return Matrix(height, width, A.LDim(), _storage=A._st, _offset=A._off + i + j * A.LDim())
