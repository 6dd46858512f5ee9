"""Tree and traversal invariants (boxtree is not available, so parity with
its numbering is unpinned; these invariants carry correctness): every source
is counted exactly once per target leaf across lists 1-4 and the ancestors'
lists 2 and 4, particles are box-contiguous, translation classes reproduce
centre differences."""

import numpy as np
import pytest

from sumpy_amd.array_context import TorchArrayContext
from sumpy_amd.tree import build_tree_and_traversal


def _cases():
    rng = np.random.default_rng(0)
    g = rng.integers(0, 8, 8000)
    c = rng.uniform(0.1, 0.9, (8, 3))
    return {
        "uniform": (rng.random((3, 2**13)), None, 30),
        "separate_targets": (rng.standard_normal((3, 1000)),
                             0.1 + rng.standard_normal((3, 300)), 30),
        "gaussian8": (np.clip(c[g].T + 0.02*rng.standard_normal((3, 8000)), 0, 1), None, 64),
        "2d": (rng.random((2, 3000)), None, 20),
    }


@pytest.fixture(scope="module", params=list(_cases()))
def built(request):
    src, tgt, mp = _cases()[request.param]
    tree, trav = build_tree_and_traversal(TorchArrayContext("cpu"), src, tgt, mp)
    return src, tgt, mp, tree, trav


def test_particles_box_contiguous_and_permuted(built):
    src, tgt, mp, tree, trav = built
    th = tree.host
    assert np.array_equal(np.sort(th["user_source_ids"]), np.arange(src.shape[1]))
    assert np.array_equal(th["sources"], src[:, th["user_source_ids"]])
    leaves = np.nonzero(th["box_source_counts_nonchild"])[0]
    h = tree.root_extent*2.0**-th["box_levels"]
    for b in leaves[:200]:
        s0, n = th["box_source_starts"][b], th["box_source_counts_nonchild"][b]
        p = th["sources"][:, s0:s0 + n]
        c = th["box_centers"][:, b:b + 1]
        assert np.all(np.abs(p - c) <= h[b]/2*(1 + 1e-12))
    # leaves hold at most max_particles_in_box particles; particles sit in leaves only
    npart = th["box_source_counts_cumul"] + (
        0 if tree.sources_are_targets else th["box_target_counts_cumul"])
    leaf = th["box_child_ids"].max(axis=0) == 0
    assert np.all(npart[leaf] <= mp)
    assert np.all(th["box_source_counts_nonchild"][~leaf] == 0)


def test_lists_cover_every_source_once(built):
    src, tgt, mp, tree, trav = built
    th, vh = tree.host, trav.host
    nsrc = th["sources"].shape[1]
    par = th["box_parent_ids"]
    bss, bsc = th["box_source_starts"], th["box_source_counts_cumul"]
    totp_idx = {b: i for i, b in enumerate(vh["target_or_target_parent_boxes"])}
    l3 = {}
    for tb, st, li in vh["sep_smaller"]:
        for i, t in enumerate(tb):
            l3.setdefault(t, []).extend(li[st[i]:st[i + 1]])
    for it, t in enumerate(vh["target_boxes"]):
        cnt = np.zeros(nsrc, int)
        for b in vh["neighbor_source_boxes_lists"][
                vh["neighbor_source_boxes_starts"][it]:vh["neighbor_source_boxes_starts"][it + 1]]:
            cnt[bss[b]:bss[b] + bsc[b]] += 1
        for b in l3.get(t, []):
            cnt[bss[b]:bss[b] + bsc[b]] += 1
        a = t
        while True:
            i = totp_idx[a]
            for key in ("from_sep_siblings", "from_sep_bigger"):
                st, li = vh[key + "_starts"], vh[key + "_lists"]
                for b in li[st[i]:st[i + 1]]:
                    cnt[bss[b]:bss[b] + bsc[b]] += 1
            if a == 0:
                break
            a = par[a]
        assert (cnt == 1).all()


def test_translation_classes_are_center_differences(built):
    src, tgt, mp, tree, trav = built
    th, vh = tree.host, trav.host
    cls = vh["from_sep_siblings_translation_classes"]
    vec = vh["from_sep_siblings_translation_class_to_distance_vector"]
    st, li = vh["from_sep_siblings_starts"], vh["from_sep_siblings_lists"]
    ctr = th["box_centers"]
    for i, t in enumerate(vh["target_or_target_parent_boxes"]):
        for e in range(st[i], st[i + 1]):
            d = ctr[:, t] - ctr[:, li[e]]
            assert np.allclose(vec[:, cls[e]], d, rtol=0, atol=1e-12*tree.root_extent)
    ls = vh["from_sep_siblings_translation_classes_level_starts"]
    assert ls[-1] == vec.shape[1]
    assert all(ls[lev + 1] - ls[lev] <= (316 if tree.dimensions == 3 else 40)
               for lev in range(tree.nlevels))


def test_boxtree_interface_names(built):
    from sumpy_amd.tree import FMMTraversalBuilder, TranslationClassesBuilder, TreeBuilder
    src, tgt, mp, tree, trav = built
    for name in ("box_source_starts", "box_source_counts_nonchild", "box_target_starts",
                 "box_target_counts_nonchild", "box_centers", "box_child_ids",
                 "box_parent_ids", "level_start_box_nrs", "sources", "targets",
                 "user_source_ids", "sorted_target_ids"):
        assert getattr(tree, name) is not None
    for name in ("target_boxes", "source_boxes", "neighbor_source_boxes_starts",
                 "from_sep_siblings_lists", "from_sep_bigger_lists",
                 "target_boxes_sep_smaller_by_source_level", "from_sep_smaller_by_level",
                 "level_start_target_or_target_parent_box_nrs"):
        assert getattr(trav, name) is not None
    actx = TorchArrayContext("cpu")
    t2, _ = TreeBuilder(actx)(actx, src, targets=tgt, max_particles_in_box=mp)
    v2, _ = FMMTraversalBuilder(actx)(actx, t2)
    tc, _ = TranslationClassesBuilder(actx)(actx, v2, t2)
    assert tc.from_sep_siblings_translation_classes.shape[0] == len(v2.host["from_sep_siblings_lists"])


def _pairs(vh):
    """(target box, source box) pairs of lists 1, 2, 3 and 4."""
    totp = vh["target_or_target_parent_boxes"]

    def csr(boxes, st, li):
        return {(int(t), int(s)) for i, t in enumerate(boxes) for s in li[st[i]:st[i + 1]]}
    l1 = csr(vh["target_boxes"], vh["neighbor_source_boxes_starts"],
             vh["neighbor_source_boxes_lists"])
    l2 = csr(totp, vh["from_sep_siblings_starts"], vh["from_sep_siblings_lists"])
    l3 = set()
    for tb, st, li in vh["sep_smaller"]:
        l3 |= csr(tb, st, li)
    l4 = csr(totp, vh["from_sep_bigger_starts"], vh["from_sep_bigger_lists"])
    return l1, l2, l3, l4


def test_swapped_traversal_transposes_every_list():
    """Sources and targets exchanged on the same boxes: lists 1 and 2 are
    transposed, and lists 3 and 4 exchange (the dual lists of the adaptive
    FMM), which makes the FMM with swapped roles the exact transpose of the
    forward FMM."""
    from sumpy_amd.tree import swapped_traversal, user_order_points
    rng = np.random.default_rng(5)
    c = rng.uniform(0.2, 0.8, (4, 3))
    src = np.clip(c[rng.integers(0, 4, 3000)].T + 0.05*rng.standard_normal((3, 3000)), 0, 1)
    tgt = rng.random((3, 1200))
    actx = TorchArrayContext("cpu")
    tree, trav = build_tree_and_traversal(actx, src, tgt, 20)
    assert np.array_equal(user_order_points(tree, "sources").numpy(), src)
    assert np.array_equal(user_order_points(tree, "targets").numpy(), tgt)
    strav = swapped_traversal(actx, trav)
    assert swapped_traversal(actx, trav) is strav
    st = strav.tree
    assert np.array_equal(st.host["box_centers"], tree.host["box_centers"])
    assert np.array_equal(user_order_points(st, "sources").numpy(), tgt)
    l1, l2, l3, l4 = _pairs(trav.host)
    s1, s2, s3, s4 = _pairs(strav.host)
    assert len(l3) > 0 and len(l4) > 0

    def tr(p):
        return {(b, a) for a, b in p}
    assert s1 == tr(l1) and s2 == tr(l2)
    assert s4 == tr(l3) and s3 == tr(l4)
