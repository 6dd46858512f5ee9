"""Oracle-side index tables (test infrastructure only).

Built independently of ``sumpy_amd.expansion.terms`` so that the two can be
checked against each other.  Follows the reference directly:

* identifier enumeration: pytools' ``generate_nonnegative_integer_tuples_
  summing_to_at_most`` order, stably sorted by degree
  (``sumpy/expansion/__init__.py:285-300``), re-sorted by the PDE ordering key
  (``__init__.py:621-654,671-675``);
* stored set and projection: ``get_stored_ids_and_unscaled_projection_matrix``
  (``__init__.py:757-835``) as a ``CSEMatVecOperator`` (rows from input or
  from earlier output rows, ``__init__.py:493-557``), materialised by running
  its ``matvec`` on unit vectors; rscale powers as in ``__init__.py:837-886``;
* the circulant M2L index sets of ``m2l.py:403-459``.
"""

from __future__ import annotations

import itertools
from functools import lru_cache

import numpy as np


def gnitstam(n, length):
    if length == 0:
        yield ()
        return
    for i in range(n + 1):
        for rem in gnitstam(n - i, length - 1):
            yield rem + (i,)


class Wrangler:
    """A restatement of ``LinearPDEBasedExpansionTermsWrangler`` (Laplace or
    Helmholtz PDE) or ``FullExpansionTermsWrangler`` (``compressed=False``)."""

    def __init__(self, order, dim=3, compressed=True, max_mi=None,
                 helmholtz=False):
        self.order = order
        self.dim = dim
        self.compressed = compressed
        self.helmholtz = helmholtz
        res = sorted(gnitstam(order, dim), key=sum)
        if max_mi is not None:
            res = [mi for mi in res if all(mi[i] <= max_mi[i] for i in range(dim))]
        if compressed:
            res = sorted(res, key=lambda mi: (sum(mi), *mi))
        self.full = res
        idx = {mi: i for i, mi in enumerate(res)}
        # PDE: sum_i u_{2 e_i} (+ k^2 u) = 0; leading identifier 2 e_0
        pde = [(tuple(2 if j == i else 0 for j in range(dim)), 1.0)
               for i in range(dim)]
        if helmholtz:
            pde.append(((0,)*dim, "k2"))
        max_mi_pde = pde[0][0]
        self.rows_in = []     # (input index, coeff)
        self.rows_out = []    # (output index, coeff, rscale power, uses k2)
        self.stored = []
        if not compressed or any(p[0] not in idx for p in pde):
            self.stored = list(res)
            for i in range(len(res)):
                self.rows_in.append([(i, 1.0)])
                self.rows_out.append([])
            return
        for i, mi in enumerate(res):
            if any(mi[d] < max_mi_pde[d] for d in range(dim)):
                self.rows_in.append([(len(self.stored), 1.0)])
                self.rows_out.append([])
                self.stored.append(mi)
                continue
            diff = [mi[d] - max_mi_pde[d] for d in range(dim)]
            assignment = []
            for other_mi, coeff in pde:
                j = idx[tuple(o + df for o, df in zip(other_mi, diff))]
                if i == j:
                    continue
                deg_diff = sum(mi) - sum(res[j])
                assignment.append((j, -1.0, deg_diff, coeff == "k2"))
            self.rows_in.append([])
            self.rows_out.append(assignment)

    def projection(self, rscale=1.0, k=0.0):
        """Dense matrix of the rscale-adjusted projection operator."""
        nf, ns = len(self.full), len(self.stored)
        dt = np.complex128 if isinstance(k, complex) else np.float64
        P = np.zeros((nf, ns), dtype=dt)
        for col in range(ns):
            inp = np.zeros(ns)
            inp[col] = 1
            out = []
            for i in range(nf):
                v = 0
                for j, c in self.rows_in[i]:
                    v += inp[j]*c
                for j, c, deg, isk in self.rows_out[i]:
                    cc = c*(k*k if isk else 1.0)*rscale**deg
                    v += out[j]*cc
                out.append(v)
            P[:, col] = out
        return P


@lru_cache(maxsize=None)
def wrangler(order, dim=3, compressed=True, helmholtz=False):
    return Wrangler(order, dim, compressed, helmholtz=helmholtz)


def m2l_tables(order_src, order_tgt, compressed=True):
    """Circulant-layout tables for the non-FFT M2L (``m2l.py:403-459``)."""
    ws = wrangler(order_src, 3, compressed)
    wt = wrangler(order_tgt, 3, compressed)
    max_mi = [max(mi[i] for mi in ws.stored) + max(mi[i] for mi in wt.stored)
              for i in range(3)]
    circ = list(itertools.product(*[range(m + 1) for m in max_mi]))
    cidx = {mi: i for i, mi in enumerate(circ)}
    needed = set()
    for t in wt.stored:
        for s in ws.stored:
            needed.add(tuple(a + b for a, b in zip(s, t)))
    needed = sorted(needed)
    w2 = Wrangler(order_src + order_tgt, 3, compressed, max_mi=max_mi)
    f2idx = {mi: i for i, mi in enumerate(w2.full)}
    return dict(
        max_mi=np.array(max_mi, np.int32),
        ncirc=len(circ),
        src_circ_idx=np.array([cidx[mi] for mi in ws.stored], np.int32),
        tgt_circ_idx=np.array([cidx[mi] for mi in wt.stored], np.int32),
        tgt_degree=np.array([sum(mi) for mi in wt.stored], np.int32),
        w2=w2,
        needed_full_idx=np.array([f2idx[mi] for mi in needed], np.int32),
        needed_circ_idx=np.array([cidx[mi] for mi in needed], np.int32),
    )
