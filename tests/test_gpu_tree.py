"""The GPU tree and traversal builder against the host builder: every tree,
traversal and translation-class array equal (box numbering, particle order,
lists 1-4 per source level, class numbering and distance vectors, centres
bit for bit), the host builder's invariants on the GPU builder's output, and
the FMM over either tree giving the same potentials bit for bit.

The reference takes these arrays from boxtree (``TreeBuilder``,
``FMMTraversalBuilder``, ``TranslationClassesBuilder``:
``sumpy/test/test_fmm.py:182-196``), which is absent here: the host builder
is the checker, and its own invariants are in ``tests/test_tree.py``.
"""

import time

import numpy as np
import pytest

import test_tree
from sumpy_amd.tree import build_tree_and_traversal
from sumpy_amd.tree_gpu import build_tree_and_traversal_gpu

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(1)
    c = rng.uniform(0.1, 0.9, (8, 3))
    g = rng.integers(0, 8, 2**15)
    same = np.full((3, 200), 0.25)
    same[:, 100:] = rng.random((3, 100))
    return {
        "uniform": (rng.random((3, 20000)), None, 30, None),
        "gaussian8": (np.clip(c[g].T + 0.02*rng.standard_normal((3, 2**15)), 0, 1), None, 64,
                      None),
        "separate_targets": (rng.standard_normal((3, 3000)),
                             0.1 + rng.standard_normal((3, 2000)), 30, None),
        "2d": (rng.random((2, 5000)), None, 20, None),
        "2d_separate_targets": (rng.random((2, 4000)), 0.5*rng.random((2, 700)), 16, None),
        "single_box": (rng.random((3, 7)), None, 30, None),
        "capped_levels": (rng.random((3, 5000)), None, 4, 4),
        "coincident": (same, None, 30, None),
    }


CASES = _cases()


@pytest.fixture(scope="module", params=list(CASES))
def both(request, actx):
    src, tgt, mp, ml = CASES[request.param]
    print(f"\ntree case {request.param}: src {src.shape} tgt "
          f"{None if tgt is None else tgt.shape} max {mp}/box levels {ml}", flush=True)
    host = build_tree_and_traversal(actx, src, tgt, mp, ml)
    gpu = build_tree_and_traversal_gpu(actx, src, tgt, mp, ml)
    return request.param, src, tgt, mp, host, gpu


def test_tree_arrays_equal_host_builder(both):
    name, _, _, _, (ht, hv), (gt, gv) = both
    assert (gt.dimensions, gt.nlevels, gt.sources_are_targets) == \
        (ht.dimensions, ht.nlevels, ht.sources_are_targets)
    assert gt.root_extent == ht.root_extent
    assert np.array_equal(gt.bounding_box[0], ht.bounding_box[0])
    assert set(gt.host) == set(ht.host)
    for k, v in ht.host.items():
        g = gt.host[k]
        assert g.dtype == v.dtype and g.shape == v.shape, (name, k, g.dtype, v.dtype)
        assert np.array_equal(g, v), (name, k)
        dev = getattr(gt, k)
        assert dev.is_cuda and np.array_equal(dev.cpu().numpy(), v), (name, k)


def test_traversal_arrays_equal_host_builder(both):
    name, _, _, _, (ht, hv), (gt, gv) = both
    assert set(gv.host) == set(hv.host)
    for k, v in hv.host.items():
        if k == "sep_smaller":
            continue
        g = gv.host[k]
        assert g.dtype == v.dtype and g.shape == v.shape, (name, k, g.shape, v.shape)
        assert np.array_equal(g, v), (name, k)
    assert len(gv.host["sep_smaller"]) == len(hv.host["sep_smaller"])
    for lev, (g3, h3) in enumerate(zip(gv.host["sep_smaller"], hv.host["sep_smaller"])):
        for a, b in zip(g3, h3):
            assert a.dtype == b.dtype and np.array_equal(a, b), (name, lev)
    for k in ("neighbor_source_boxes_lists", "from_sep_siblings_lists",
              "from_sep_bigger_lists", "target_boxes"):
        assert np.array_equal(getattr(gv, k).cpu().numpy(), hv.host[k])
    for lev, csr in enumerate(gv.from_sep_smaller_by_level):
        assert np.array_equal(csr.lists.cpu().numpy(), hv.host["sep_smaller"][lev][2])
    gc, hc = gv.__dict__["_translation_classes"], hv.__dict__["_translation_classes"]
    for f in ("from_sep_siblings_translation_classes",
              "from_sep_siblings_translation_class_to_distance_vector",
              "from_sep_siblings_translation_classes_level_starts"):
        a, b = getattr(gc, f).cpu().numpy(), getattr(hc, f).cpu().numpy()
        assert a.shape == b.shape and np.array_equal(a, b), (name, f)


def test_gpu_tree_invariants(both):
    name, src, tgt, mp, _, (gt, gv) = both
    if CASES[name][3] is not None or name == "coincident":
        # depth reached (given cap, or coincident points at the deepest
        # level): leaves may exceed max_particles_in_box
        mp = src.shape[1]
    built = (src, tgt, mp, gt, gv)
    test_tree.test_particles_box_contiguous_and_permuted(built)
    test_tree.test_lists_cover_every_source_once(built)
    test_tree.test_translation_classes_are_center_differences(built)


def test_fmm_over_gpu_tree_is_bitwise_equal(actx):
    from test_gpu_fmm import run_fmm
    src, _, mp, _ = CASES["gaussian8"]
    q = np.random.default_rng(2).standard_normal(src.shape[1])
    trees = [build_tree_and_traversal(actx, src, None, mp),
             build_tree_and_traversal_gpu(actx, src, None, mp)]
    (a,), _, _ = run_fmm(actx, src, q, 6, tree_trav=trees[0])
    (b,), _, _ = run_fmm(actx, src, q, 6, tree_trav=trees[1])
    assert np.array_equal(a, b)


@pytest.mark.parametrize("distribution", ["uniform", "gaussian8"])
def test_gpu_tree_at_2_20_equals_host_and_reports_time(actx, distribution):
    import torch
    from test_gpu_c3 import workload
    pts, _ = workload(distribution, n=2**20)
    dpts = torch.as_tensor(pts, device=actx.device)
    build_tree_and_traversal_gpu(actx, dpts, None, 64)          # warm-up
    torch.cuda.synchronize()
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        gt, gv = build_tree_and_traversal_gpu(actx, dpts, None, 64)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    timing = {}
    build_tree_and_traversal_gpu(actx, dpts, None, 64, timing=timing)
    t0 = time.perf_counter()
    ht, hv = build_tree_and_traversal(actx, pts, None, 64)
    th = time.perf_counter() - t0
    print(f"{distribution} 2^20: GPU build {1e3*min(times):.1f} ms (steps "
          + ", ".join(f"{k} {1e3*v:.1f}" for k, v in timing.items())
          + f" ms), host build {1e3*th:.0f} ms")
    for k, v in ht.host.items():
        assert np.array_equal(gt.host[k], v), k
    for k, v in hv.host.items():
        if k != "sep_smaller":
            assert np.array_equal(gv.host[k], v), k
