"""Coefficient index sets, storage indices and the PDE projection.

Restates the combinatorics of ``sumpy/expansion/__init__.py``:

* full Taylor identifiers ``|mi| <= p`` in degree-lexicographic order
  (``FullExpansionTermsWrangler``, ``__init__.py:255-300,401-414``);
* the PDE-compressed set of ``LinearPDEBasedExpansionTermsWrangler``
  (``__init__.py:563-886``): for Laplace/Helmholtz in 3D the PDE's leading
  multi-index is ``(2,0,0)``, so the stored set is ``mi_x <= 1`` and every
  other derivative follows from ``u_xx = -u_yy - u_zz (- k^2 u)``;
* the projection matrix ``P`` (full <- stored), with the rscale factor
  ``rscale**(|row|-|col|)`` (``__init__.py:837-886``).  The reference applies
  ``P`` as a recursive CSE operator; here it is materialised once as a sparse
  matrix, which is what the GPU kernels consume.

All tables are plain numpy arrays; the HIP kernels receive them as device
buffers.
"""

from __future__ import annotations

from functools import lru_cache

import numpy as np


def _gnitstam(n: int, length: int):
    """Non-negative integer tuples summing to at most *n*, first entry
    varying fastest (the generation order the reference sorts stably)."""
    if length == 0:
        yield ()
        return
    for i in range(n + 1):
        for rem in _gnitstam(n - i, length - 1):
            yield (*rem, i)


@lru_cache(maxsize=None)
def full_ids_taylor(order: int, dim: int) -> tuple[tuple[int, ...], ...]:
    """Full identifiers in the ``FullExpansionTermsWrangler`` order
    (stable sort by total degree, ``__init__.py:285-300``)."""
    return tuple(sorted(_gnitstam(order, dim), key=sum))


@lru_cache(maxsize=None)
def full_ids_pde(order: int, dim: int) -> tuple[tuple[int, ...], ...]:
    """Full identifiers in the PDE-conforming order: key
    ``(|mi|, mi_0, mi_1, ...)`` (``__init__.py:621-654``; the on-axis PDE
    terms put axis 0 first for Laplace/Helmholtz)."""
    return tuple(sorted(_gnitstam(order, dim),
                        key=lambda mi: (sum(mi), *mi)))


def full_storage_index(mi) -> int:
    """``FullExpansionTermsWrangler.get_storage_index`` (``__init__.py:401-414``)."""
    order = sum(mi)
    if len(mi) == 3:
        return ((order*(order + 1)*(order + 2))//6
                + (order + 2)*mi[2] - (mi[2]*(mi[2] + 1))//2 + mi[1])
    if len(mi) == 2:
        return (order*(order + 1))//2 + mi[1]
    raise NotImplementedError


def pde_storage_index(mi) -> int:
    """``LinearPDEBasedExpansionTermsWrangler.get_storage_index`` for a PDE
    whose leading multi-index is ``2 e_0`` (``__init__.py:697-750``)."""
    c = 2
    order = sum(mi)
    if len(mi) == 3:
        if order < c - 1:
            return ((order*(order + 1)*(order + 2))//6
                    + (order + 2)*mi[0] - (mi[0]*(mi[0] + 1))//2 + mi[1])
        return ((c*(c-1)*(c-2))//6
                + (c*order*(2 + order - c) + mi[0]*(3 - mi[0] + 2*order))//2
                + mi[1])
    if len(mi) == 2:
        if order < c - 1:
            return (order*(order + 1))//2 + mi[0]
        return (c*(c-1))//2 + c*(order - c + 1) + mi[0]
    raise NotImplementedError


class TermsTable:
    """Index tables of one (expansion kind, order, dim).

    Attributes:
      ``full``   -- (nfull, dim) int array of full identifiers, in order;
      ``stored`` -- (nstored, dim) int array of stored identifiers;
      ``stored_in_full`` -- position of each stored id within ``full``.
    """

    def __init__(self, order: int, dim: int, compressed: bool):
        self.order = order
        self.dim = dim
        self.compressed = compressed
        if compressed:
            full = full_ids_pde(order, dim)
            stored = [mi for mi in full if mi[0] < 2] if order >= 2 else list(full)
        else:
            full = full_ids_taylor(order, dim)
            stored = list(full)
        self.full_list = list(full)
        self.stored_list = stored
        self.full = np.array(full, dtype=np.int32).reshape(-1, dim)
        self.stored = np.array(stored, dtype=np.int32).reshape(-1, dim)
        self.full_index = {mi: i for i, mi in enumerate(full)}
        self.stored_index = {mi: i for i, mi in enumerate(stored)}
        self.stored_in_full = np.array(
            [self.full_index[mi] for mi in stored], dtype=np.int32)

    @property
    def nfull(self):
        return len(self.full_list)

    @property
    def nstored(self):
        return len(self.stored_list)

    def projection(self, k2rscale2: float | complex = 0.0) -> np.ndarray:
        """Dense ``P`` (nfull x nstored) with ``D_full = P D_stored``.

        *k2rscale2* is ``(k rscale)^2`` for Helmholtz and 0 for Laplace
        (``__init__.py:757-835`` with the rscale powers of ``:837-886``).
        """
        dt = np.complex128 if isinstance(k2rscale2, complex) else np.float64
        P = np.zeros((self.nfull, self.nstored), dtype=dt)
        for i, mi in enumerate(self.full_list):
            if mi in self.stored_index:
                P[i, self.stored_index[mi]] = 1
                continue
            assert self.compressed
            base = (mi[0] - 2, *mi[1:])
            for ax in range(1, self.dim):
                other = list(base)
                other[ax] += 2
                P[i] -= P[self.full_index[tuple(other)]]
            if k2rscale2 != 0:
                P[i] -= k2rscale2 * P[self.full_index[base]]
        return P


@lru_cache(maxsize=None)
def get_terms_table(order: int, dim: int, compressed: bool) -> TermsTable:
    return TermsTable(order, dim, compressed)


def projection_csr(table: TermsTable, k2rscale2=0.0):
    """``P`` as (row_starts, col, val) over full rows, dropping zeros."""
    P = table.projection(k2rscale2)
    starts = [0]
    cols, vals = [], []
    for i in range(P.shape[0]):
        nz = np.nonzero(P[i])[0]
        cols.extend(nz.tolist())
        vals.extend(P[i, nz].tolist())
        starts.append(len(cols))
    return (np.array(starts, np.int32), np.array(cols, np.int32),
            np.array(vals, dtype=P.dtype))


def m2l_circulant_shape(src: TermsTable, tgt: TermsTable) -> tuple[int, ...]:
    """``max_mi + 1`` of the circulant layout (``m2l.py:403-450``)."""
    return tuple(int(src.stored[:, i].max() + tgt.stored[:, i].max()) + 1
                 for i in range(src.dim))


def circulant_offsets(ids: np.ndarray, shape) -> np.ndarray:
    """Row-major linear offset of each multi-index in a box of *shape*
    (the ``generate_nonnegative_integer_tuples_below`` order,
    ``m2l.py:436-440``).  The offset is linear in the multi-index, so
    ``off(t) + off(s) == off(t + s)``."""
    strides = np.ones(len(shape), dtype=np.int64)
    for i in range(len(shape) - 2, -1, -1):
        strides[i] = strides[i + 1] * shape[i + 1]
    return (ids.astype(np.int64) @ strides).astype(np.int32)
