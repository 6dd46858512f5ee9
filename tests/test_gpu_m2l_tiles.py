"""The M2L contraction at its two tile widths (64 and 32 target boxes per
tile; the autotuner picks one per order and tree): with one split per tile
both widths add every target's class slots in the same order, so the work
rows are bitwise equal, in FP64, FP32 and split-BF16, for the allocating and
the allocation-free entry points; with splits they agree to rounding.  The
64-column contraction itself is checked against the oracle's Toeplitz M2L
(``tests/test_gpu_operators.py``)."""

import pytest
import torch

from test_zz_no_alloc import _level

pytestmark = pytest.mark.gpu

ONE_SPLIT = (1, 1 << 30)   # policy: at least 2^30 slots per split, i.e. one split


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
def test_tile_widths_give_the_same_work_rows(actx, precision):
    from sumpy_amd.e2e import m2l_apply_levels
    entry, new_work, keep = _level(actx, 40000, 6, precision, 0)
    w64, w32 = new_work(), new_work()
    e32 = entry(w32, tile_width=32)
    assert e32["tw"] == 32 and e32["sched"]["ntiles"] > entry(new_work())["sched"]["ntiles"]
    m2l_apply_levels(actx, [entry(w64)], policy=ONE_SPLIT)
    m2l_apply_levels(actx, [e32], policy=ONE_SPLIT)
    torch.cuda.synchronize()
    assert torch.count_nonzero(w64).item() > 0
    assert torch.equal(w64, w32), precision
    # with splits (and through the fixed scratch): equal to rounding
    ws, wn = new_work(), new_work()
    m2l_apply_levels(actx, [entry(ws, tile_width=32)], policy=(8, 6))
    m2l_apply_levels(actx, [entry(wn, tile_width=32)], policy=(8, 6), noalloc=True)
    torch.cuda.synchronize()
    assert torch.equal(ws, wn)
    tol = 1e-13 if precision == "fp64" else 1e-5
    assert (torch.linalg.norm((ws - w64).double())/torch.linalg.norm(w64.double())).item() < tol


def test_mixed_tile_widths_in_one_launch_are_refused(actx):
    from sumpy_amd.e2e import m2l_apply_levels
    entry, new_work, keep = _level(actx, 6000, 4, "fp64", 1)
    with pytest.raises(ValueError, match="tile width"):
        m2l_apply_levels(actx, [entry(new_work()), entry(new_work(), tile_width=32)])
