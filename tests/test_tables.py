"""The product's coefficient tables against the oracle's independent
construction and the reference's storage-index formulas
(``sumpy/test/test_misc.py:783-805``)."""

import numpy as np
import pytest

from oracle import tables as otab
from sumpy_amd.expansion.terms import (
    circulant_offsets,
    full_storage_index,
    get_terms_table,
    m2l_circulant_shape,
    pde_storage_index,
)


@pytest.mark.parametrize("order", list(range(0, 13)))
def test_storage_index_equals_position(order):
    t = get_terms_table(order, 3, True)
    assert all(pde_storage_index(mi) == i for i, mi in enumerate(t.stored_list))
    assert t.nstored == (order + 1)**2
    f = get_terms_table(order, 3, False)
    assert all(full_storage_index(mi) == i for i, mi in enumerate(f.full_list))
    assert f.nfull == (order + 1)*(order + 2)*(order + 3)//6


@pytest.mark.parametrize("order", [1, 2, 5, 10, 12])
@pytest.mark.parametrize("compressed", [True, False])
def test_tables_match_oracle(order, compressed):
    t = get_terms_table(order, 3, compressed)
    w = otab.wrangler(order, 3, compressed)
    assert t.full_list == w.full
    assert t.stored_list == w.stored
    for rs in (1.0, 0.37):
        assert np.array_equal(t.projection(), w.projection(rs))


def test_helmholtz_projection_matches_oracle():
    t = get_terms_table(6, 3, True)
    w = otab.Wrangler(6, 3, True, helmholtz=True)
    k, rs = 2.5, 0.3
    assert np.allclose(t.projection((k*rs)**2), w.projection(rs, k), rtol=1e-14, atol=0)


def test_projection_reproduces_harmonic_derivatives():
    """D_full = P D_stored for the derivatives of 1/r (harmonic)."""
    import oracle
    order = 8
    t = get_terms_table(order, 3, True)
    D = oracle.lap3d_derivs(np.array([0.4, 0.9, -0.3]), 0.25, order, (order,)*3)
    stored = np.array([D[mi] for mi in t.stored_list])
    full = np.array([D[mi] for mi in t.full_list])
    assert np.max(np.abs(t.projection() @ stored - full)) < 1e-12*np.abs(full).max()


@pytest.mark.parametrize("order", [3, 10])
def test_circulant_layout_matches_oracle(order):
    t = get_terms_table(order, 3, True)
    shape = m2l_circulant_shape(t, t)
    T = otab.m2l_tables(order, order)
    assert tuple(T["max_mi"] + 1) == shape
    assert np.array_equal(circulant_offsets(t.stored, shape), T["src_circ_idx"])
    # linearity: off(t) + off(s) == off(t + s)
    off = circulant_offsets(t.stored, shape)
    sums = t.stored[:, None, :] + t.stored[None, :, :]
    assert np.array_equal(off[:, None] + off[None, :],
                          circulant_offsets(sums.reshape(-1, 3), shape).reshape(len(off), -1))
