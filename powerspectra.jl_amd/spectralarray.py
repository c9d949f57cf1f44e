"""SpectralArray: an array indexed by multipole l (offset indexing).

Mirrors the reference's SpectralArray (src/spectralarray.jl:5-90): a thin
wrapper whose indices start at a chosen l (0 by default) over a plain
column-major parent array.  `parent(A)` is the array the C ABI consumes.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np

from ._abi import AxisError


def _as_range(r) -> range:
    if isinstance(r, range):
        return r
    if isinstance(r, (int, np.integer)):
        return range(0, int(r))
    raise TypeError(f"expected an int size or a range, got {r!r}")


class SpectralArray:
    """N-dimensional array whose axis k runs over l = first[k] .. first[k]+size-1.

    The parent is stored column-major (Fortran order), like Julia's Array,
    so parent(A) of a 2-D block is exactly the layout of the C ABI."""

    __slots__ = ("_parent", "_first")

    def __init__(self, A, offsets: Sequence[int] | None = None, *, firsts: Sequence[int] | None = None):
        A = np.asarray(A)
        if A.ndim >= 2 and not A.flags.f_contiguous:
            A = np.asfortranarray(A)
        self._parent = A
        if firsts is not None:
            self._first = tuple(int(f) for f in firsts)
        elif offsets is not None:
            # Julia OffsetArray offsets: index = offset + position (1-based)
            self._first = tuple(int(o) + 1 for o in offsets)
        else:
            self._first = (0,) * A.ndim
        if len(self._first) != A.ndim:
            raise AxisError("one offset per dimension is required")

    # -- Julia-style accessors ----------------------------------------------
    @property
    def parent(self) -> np.ndarray:
        return self._parent

    @property
    def offsets(self) -> tuple:
        return tuple(f - 1 for f in self._first)

    def axes(self, dim: int | None = None):
        ax = tuple(range(f, f + n) for f, n in zip(self._first, self._parent.shape))
        return ax if dim is None else ax[dim]

    @property
    def shape(self):
        return self._parent.shape

    @property
    def ndim(self):
        return self._parent.ndim

    @property
    def dtype(self):
        return self._parent.dtype

    def firstindex(self, dim: int = 0) -> int:
        return self._first[dim]

    def lastindex(self, dim: int = 0) -> int:
        return self._first[dim] + self._parent.shape[dim] - 1

    def _pos(self, key):
        if not isinstance(key, tuple):
            key = (key,)
        if len(key) != self.ndim:
            raise IndexError("one index per dimension is required")
        out = []
        for k, f, n in zip(key, self._first, self._parent.shape):
            if isinstance(k, (int, np.integer)):
                i = int(k) - f
                if not 0 <= i < n:
                    raise IndexError(f"index {k} outside {f}:{f + n - 1}")
                out.append(i)
            elif isinstance(k, range):
                if k.step != 1:
                    raise IndexError("only unit-step ranges")
                i0, i1 = k.start - f, k.stop - f
                if i0 < 0 or i1 > n:
                    raise IndexError(f"range {k} outside {f}:{f + n - 1}")
                out.append(slice(i0, i1))
            elif isinstance(k, slice) and k == slice(None):
                out.append(slice(None))
            else:
                raise IndexError(f"unsupported index {k!r}")
        return tuple(out)

    def __getitem__(self, key):
        """A[l] / A[l1, l2] index by multipole; ranges return the parent block
        (like Julia's UnitRange slicing, src/spectralarray.jl:20-22)."""
        return self._parent[self._pos(key)]

    def __setitem__(self, key, value):
        self._parent[self._pos(key)] = value

    def view(self, *ranges: range) -> "SpectralArray":
        """IdentityRange slicing: a SpectralArray that keeps the l labels."""
        pos = self._pos(tuple(ranges))
        return SpectralArray(self._parent[pos], firsts=[r.start for r in ranges])

    def __array__(self, dtype=None, copy=None):
        return np.asarray(self._parent, dtype=dtype)

    def copy(self) -> "SpectralArray":
        return SpectralArray(self._parent.copy(order="F"), firsts=self._first)

    def __repr__(self):
        ax = ", ".join(f"{r.start}:{r.stop - 1}" for r in self.axes())
        return f"SpectralArray(axes=({ax}), parent={self._parent!r})"

    # elementwise arithmetic keeps the axes
    def _binop(self, other, op):
        o = other._parent if isinstance(other, SpectralArray) else other
        if isinstance(other, SpectralArray) and other.axes() != self.axes():
            raise AxisError("axes differ")
        return SpectralArray(op(self._parent, o), firsts=self._first)

    def __add__(self, o): return self._binop(o, np.add)
    def __sub__(self, o): return self._binop(o, np.subtract)
    def __mul__(self, o): return self._binop(o, np.multiply)
    def __truediv__(self, o): return self._binop(o, np.true_divide)
    __radd__ = __add__
    __rmul__ = __mul__

    def __neg__(self):
        return SpectralArray(-self._parent, firsts=self._first)


def SpectralVector(A, offset: int | None = None) -> SpectralArray:
    """0-indexed vector (src/spectralarray.jl:17), or with a Julia offset."""
    A = np.asarray(A)
    if A.ndim != 1:
        raise AxisError("SpectralVector needs a 1-D array")
    return SpectralArray(A, offsets=None if offset is None else (offset,))


def parent(A):
    return A.parent if isinstance(A, SpectralArray) else np.asarray(A)


def spectralzeros(*r) -> SpectralArray:
    """spectralzeros(n...) or spectralzeros(range...) (src/spectralarray.jl:73-74)."""
    rs = [_as_range(x) for x in r]
    A = np.zeros([len(x) for x in rs], order="F")
    return SpectralArray(A, firsts=[x.start for x in rs])


def spectralones(*r) -> SpectralArray:
    """spectralones(n...) or spectralones(range...) (src/spectralarray.jl:83-84)."""
    rs = [_as_range(x) for x in r]
    A = np.ones([len(x) for x in rs], order="F")
    return SpectralArray(A, firsts=[x.start for x in rs])


def inv(A: SpectralArray) -> SpectralArray:
    """inv keeps the offsets (src/spectralarray.jl:88-90)."""
    return SpectralArray(np.linalg.inv(A.parent), firsts=[A.firstindex(0), A.firstindex(1)])
