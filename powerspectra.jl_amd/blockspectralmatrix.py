"""BlockSpectralMatrix: dense block matrices of spectra and the FP64 LU solve.

Mirrors src/blockspectralmatrix.jl: block assembly (hvcat/vcat/hcat, :17-45),
`\\` with LU and partial pivoting (:89-129, here the library's blocked getrf
and triangular substitution through psx_lu_solve), getblock and the @spectra unpack
(:144-184).
"""
from __future__ import annotations

from typing import Sequence

import numpy as np

from ._abi import AxisError, Context, default_context
from .spectralarray import SpectralArray, SpectralVector


class BlockSpectralMatrix:
    """A contiguous column-major matrix with the l ranges of its row blocks
    (m_ells) and column blocks (n_ells)."""

    __slots__ = ("parent", "m_ells", "n_ells")

    def __init__(self, parent: np.ndarray, m_ells: Sequence[range], n_ells: Sequence[range]):
        self.parent = np.asfortranarray(parent)
        self.m_ells = tuple(m_ells)
        self.n_ells = tuple(n_ells)

    @property
    def shape(self):
        return self.parent.shape

    def __repr__(self):
        return (f"BlockSpectralMatrix(column blocks={len(self.m_ells)}, row blocks={len(self.n_ells)}, "
                f"row l: {self.m_ells}, column l: {self.n_ells})")


def hvcat(rows: Sequence[Sequence[SpectralArray]]) -> BlockSpectralMatrix:
    """[A B; C D] of SpectralArrays (src/blockspectralmatrix.jl:17-25)."""
    P = np.block([[np.asarray(b.parent) for b in row] for row in rows])
    m_ells = tuple(row[0].axes(0) for row in rows)
    n_ells = tuple(b.axes(1) for b in rows[0])
    return BlockSpectralMatrix(P, m_ells, n_ells)


def vcat(*values: SpectralArray) -> BlockSpectralMatrix:
    """[a; b] stacking of SpectralVectors / matrices (src/blockspectralmatrix.jl:27-35)."""
    P = np.concatenate([np.asarray(v.parent) for v in values], axis=0)
    m_ells = tuple(v.axes(0) for v in values)
    first = values[0]
    n_ells = (first.axes(1),) if first.ndim == 2 else (range(0, 1),)
    return BlockSpectralMatrix(P, m_ells, n_ells)


def hcat(*values: SpectralArray) -> BlockSpectralMatrix:
    P = np.concatenate([np.atleast_2d(np.asarray(v.parent).T).T for v in values], axis=1)
    m_ells = (values[0].axes(0),)
    n_ells = tuple(v.axes(1) if v.ndim == 2 else range(0, 1) for v in values)
    return BlockSpectralMatrix(P, m_ells, n_ells)


def getblock(A: BlockSpectralMatrix, i: int) -> SpectralArray:
    """Sub-block i (1-based, like the reference, src/blockspectralmatrix.jl:144-160)
    of a block vector, labelled with its l range."""
    if not 1 <= i <= len(A.m_ells):
        raise AxisError(f"block {i} of {len(A.m_ells)}")
    off = sum(len(r) for r in A.m_ells[: i - 1])
    n = len(A.m_ells[i - 1])
    col = A.parent[off: off + n]
    if col.ndim == 2 and col.shape[1] == 1:
        col = col[:, 0]
    return SpectralArray(np.asfortranarray(col), firsts=[A.m_ells[i - 1].start] + ([A.n_ells[0].start] if col.ndim == 2 else []))


def spectra(A: BlockSpectralMatrix) -> tuple:
    """@spectra a, b = A: every sub-block of a block vector (src/blockspectralmatrix.jl:170-184)."""
    return tuple(getblock(A, i) for i in range(1, len(A.m_ells) + 1))


def solve(A, B, ctx: Context | None = None):
    """`A \\ B`: FP64 LU with partial pivoting on the GPU (psx_lu_solve).

    SpectralArray \\ SpectralVector asserts firstindex(A,2) == firstindex(B,1)
    and returns a SpectralVector with B's offsets (src/blockspectralmatrix.jl:124-129);
    BlockSpectralMatrix \\ BlockSpectralMatrix checks the block l ranges (:113-122)."""
    ctx = ctx or default_context()
    if isinstance(A, BlockSpectralMatrix):
        if not isinstance(B, BlockSpectralMatrix):
            raise AxisError("block matrix \\ needs a block right-hand side")
        for a, b in zip(A.n_ells, B.m_ells):
            if a != b:
                raise AxisError("block l ranges of A and B do not match")
        X = ctx.lu_solve(A.parent, B.parent)
        return BlockSpectralMatrix(X.reshape(B.parent.shape, order="F"), B.m_ells, B.n_ells)
    if isinstance(A, SpectralArray):
        if A.firstindex(1) != B.firstindex(0):
            raise AxisError("firstindex(A, 2) != firstindex(B, 1)")
        X = ctx.lu_solve(A.parent, B.parent)
        return SpectralArray(X, firsts=[B.firstindex(d) for d in range(B.ndim)])
    X = ctx.lu_solve(np.asarray(A), np.asarray(B))
    return X


def block_vector(*vs: SpectralArray) -> BlockSpectralMatrix:
    """[pCl_EE; pCl_BB] -- a stacked block vector."""
    return vcat(*vs)
