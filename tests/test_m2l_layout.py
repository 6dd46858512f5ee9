"""The tuned LDS layouts of the split-BF16 M2L contraction are relabellings:
the packed class words at ``off_t[row] + off_s[k]`` are the class value of
the multi-index sum of the row's and the k's coefficients, every stored
coefficient appears once among the rows and once among the k, and the
tuned gathers are no slower than the plain layout's."""

import numpy as np
import pytest

from sumpy_amd import m2l_layout as ml
from sumpy_amd.expansion.terms import get_terms_table


def _pack(data, shape, strides, ncirc):
    """numpy restatement of the packing: word mx S0 + my S1 + mz <- data."""
    out = np.zeros((data.shape[0], ncirc))
    mx, my, mz = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    w = (mx*strides[0] + my*strides[1] + mz).ravel()
    out[:, w] = data.reshape(data.shape[0], -1)
    return out


@pytest.mark.parametrize("order", list(range(1, 11)))
def test_layout_is_a_relabelling(order):
    lay = ml.layout(order)
    ids = get_terms_table(order, 3, True).stored
    nl = len(ids)
    assert np.array_equal(np.sort(lay.rows[lay.rows >= 0]), np.arange(nl))
    assert np.array_equal(np.sort(lay.ks[lay.ks >= 0]), np.arange(nl))
    assert len(lay.rows) % 16 == 0 and len(lay.ks) % 32 == 0 and len(lay.ks) <= 128
    assert np.array_equal(lay.rows[lay.row_pos], np.arange(nl))
    shape = lay.shape
    s0, s1 = lay.strides
    # injective packing of the circulant box
    assert s1 >= shape[2] and s0 >= (shape[1] - 1)*s1 + shape[2]
    assert lay.ncirc >= (shape[0] - 1)*s0 + (shape[1] - 1)*s1 + shape[2]
    assert lay.ncirc <= 2560
    rng = np.random.default_rng(order)
    data = rng.standard_normal((2, *shape))
    words = _pack(data, shape, lay.strides, lay.ncirc)
    ot, os_ = lay.off_t, lay.off_s
    for r in range(len(lay.rows)):
        if lay.rows[r] < 0:
            continue
        t = ids[lay.rows[r]]
        for k in range(len(lay.ks)):
            if lay.ks[k] < 0:
                continue
            m = t + ids[lay.ks[k]]
            assert words[1, ot[r] + os_[k]] == data[1, m[0], m[1], m[2]]


@pytest.mark.parametrize("order", [3, 6, 10])
def test_tuned_gathers_not_slower(order):
    lay, base = ml.layout(order), ml.identity_layout(order)
    assert ml.gather_cost(lay.off_t, lay.off_s) <= ml.gather_cost(base.off_t, base.off_s)


def test_gather_cost_model():
    """16 distinct words in one bank cost 16 cycles per half wave; a
    broadcast costs one, and so do 16 consecutive rows when the two k groups
    of a half wave sit 16 words apart (mod 32)."""
    assert ml.gather_cost(np.arange(16)*128, np.zeros(32, np.int64)) == 16.0
    assert ml.gather_cost(np.zeros(16, np.int64), np.zeros(32, np.int64)) == 1.0
    assert ml.gather_cost(np.arange(16), (np.arange(32)//8)*16) == 1.0
    assert ml.gather_cost(np.arange(16), np.arange(32)*16) == 2.0
