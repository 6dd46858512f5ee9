"""The N = 2^20, p = 10, 64-per-box Laplace-3D FMM workload at a size the
oracle finishes in seconds: N = 2^15 sources = targets, order 10, at most 64
particles per box, self excluded, strengths N(0, 1), seed 0, for the uniform
cube and for the mixture of 8 Gaussians (sigma = 0.02, centres
``default_rng(0).uniform(0.1, 0.9, (8, 3))``, clipped to the unit cube).

* The plain FP64 form equals the oracle-driven FMM (same tree, same stages)
  to 1e-12 relative.
* The mixed forms (FP32 near field, P2M moments, list-3/4 terms and L2P; FP32 or
  split-BF16 M2L) stay within 2x of the FP64 form's relative L2 error against the FP64
  direct sum on 4096 sampled targets (reference
  ``sumpy/test/test_fmm.py:78-299`` checks FMM results against the direct
  sum the same way).
"""

import numpy as np
import pytest

import oracle
from oracle_fmm import oracle_fmm
from test_gpu_fmm import rel, run_fmm

pytestmark = pytest.mark.gpu

N, ORDER, MAX_PER_BOX, NSAMPLE = 2**15, 10, 64, 4096


def workload(distribution, n=N, seed=0):
    rng = np.random.default_rng(seed)
    if distribution == "uniform":
        pts = rng.random((3, n))
    else:
        c = np.random.default_rng(0).uniform(0.1, 0.9, (8, 3))
        pts = np.clip(c[rng.integers(0, 8, n)].T + 0.02*rng.standard_normal((3, n)), 0, 1)
    return pts, rng.standard_normal(n)


@pytest.fixture(scope="module", params=["uniform", "gaussian8"])
def c3(request, actx):
    from sumpy_amd.tree import build_tree_and_traversal
    pts, q = workload(request.param)
    print(f"\nC3 {request.param}: N={N} p={ORDER} max {MAX_PER_BOX}/box", flush=True)
    tree, trav = build_tree_and_traversal(actx, pts, None, MAX_PER_BOX)
    sample = np.sort(np.random.default_rng(0).choice(N, size=NSAMPLE, replace=False))
    direct, _ = oracle.p2p_dense(3, pts[:, sample], pts, q, t2s=sample.astype(np.int32))
    (p64,), _, _ = run_fmm(actx, pts, q, ORDER, tree_trav=(tree, trav))
    return dict(name=request.param, pts=pts, q=q, tree=tree, trav=trav, sample=sample,
                direct=direct, p64=p64)


def test_c3_fp64_matches_oracle_fmm(c3):
    if c3["name"] == "gaussian8":
        vh = c3["trav"].host
        assert sum(len(x[2]) for x in vh["sep_smaller"]) > 0      # list 3 populated
        assert len(vh["from_sep_bigger_lists"]) > 0               # list 4 populated
    ref = oracle_fmm(c3["tree"], c3["trav"], c3["q"], ORDER, exclude_self=True)
    err = rel(c3["p64"], ref)
    print(f"{c3['name']}: FP64 FMM vs oracle FMM {err:.2e}")
    assert err < 1e-12


@pytest.mark.parametrize("m2l_precision", ["fp32", "bf16x3"])
def test_c3_mixed_error_within_twice_fp64(actx, c3, m2l_precision):
    s = c3["sample"]
    e64 = rel(c3["p64"][s], c3["direct"])
    (pm,), _, _ = run_fmm(actx, c3["pts"], c3["q"], ORDER, m2l_precision=m2l_precision,
                          p2p_precision="fp32", list34_precision="fp32",
                          l2p_precision="fp32", p2m_precision="fp32",
                          tree_trav=(c3["tree"], c3["trav"]))
    em = rel(pm[s], c3["direct"])
    print(f"{c3['name']}: rel. L2 error vs direct, FP64 {e64:.3e}, "
          f"{m2l_precision} {em:.3e} (ratio {em/e64:.3f})")
    assert e64 < 2e-5
    assert em <= 2*e64, (e64, em)
