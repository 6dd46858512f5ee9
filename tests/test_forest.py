"""Forests of independent problems (many FMMs in one pass): the merged
traversal reaches, for every target, each source of its own problem exactly
once and none of another problem; translation classes stay centre
differences; and the oracle FMM over the forest equals the oracle FMM of each
problem's own tree in the same root box (the forest only renumbers)."""

import numpy as np
import pytest

import oracle
from sumpy_amd.array_context import TorchArrayContext
from sumpy_amd.tree import _build_host, _wrap, build_forest


def _sphere(n, rng):
    x = rng.standard_normal((3, n))
    x /= np.linalg.norm(x, axis=0)
    return x*(1 + 1e-3*rng.standard_normal(n))


@pytest.fixture(scope="module")
def forest():
    rng = np.random.default_rng(4)
    sizes = [300, 40, 1200, 700, 1]
    pts = [_sphere(n, rng) for n in sizes]
    pts[1] = 0.3*rng.random((3, 40))  # a problem in one corner of the box
    tree, trav, offs = build_forest(TorchArrayContext("cpu"), pts, max_particles_in_box=20)
    return pts, tree, trav, offs


def test_lists_cover_own_problem_once(forest):
    pts, tree, trav, offs = forest
    th, vh = tree.host, trav.host
    nsrc = th["sources"].shape[1]
    assert nsrc == offs[-1]
    prob_of_sorted = np.searchsorted(offs, th["user_source_ids"], side="right") - 1
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
            if par[a] == a or th["box_levels"][a] == 0:
                break
            a = par[a]
        p = prob_of_sorted[th["box_target_starts"][t]]
        own = prob_of_sorted == p
        assert (cnt[own] == 1).all() and (cnt[~own] == 0).all()


def test_levels_and_classes(forest):
    pts, tree, trav, offs = forest
    th, vh = tree.host, trav.host
    lbs = th["level_start_box_nrs"]
    for lev in range(tree.nlevels):
        assert (th["box_levels"][lbs[lev]:lbs[lev + 1]] == lev).all()
    cls = vh["from_sep_siblings_translation_classes"]
    vec = vh["from_sep_siblings_translation_class_to_distance_vector"]
    st, li = vh["from_sep_siblings_starts"], vh["from_sep_siblings_lists"]
    ctr = th["box_centers"]
    for i, t in enumerate(vh["target_or_target_parent_boxes"]):
        for e in range(st[i], st[i + 1]):
            assert np.allclose(vec[:, cls[e]], ctr[:, t] - ctr[:, li[e]], rtol=0,
                               atol=1e-12*tree.root_extent)
    ls = vh["from_sep_siblings_translation_classes_level_starts"]
    for lev in range(tree.nlevels):
        c = cls[st[vh["level_start_target_or_target_parent_box_nrs"][lev]]:
                st[vh["level_start_target_or_target_parent_box_nrs"][lev + 1]]]
        assert ((c >= ls[lev]) & (c < ls[lev + 1])).all()
        assert ls[lev + 1] - ls[lev] <= 316


def test_oracle_fmm_forest_equals_separate(forest):
    from oracle_fmm import oracle_fmm
    pts, tree, trav, offs = forest
    rng = np.random.default_rng(9)
    q = rng.standard_normal(offs[-1]) + 1j*rng.standard_normal(offs[-1])
    k, order = 3.0, 5
    pot = oracle_fmm(tree, trav, q, order, exclude_self=True, k=k)
    lo = tree.bounding_box[0]
    for i, p in enumerate(pts):
        sc, th, vh = _build_host(p, None, 20, None, (lo, tree.root_extent))
        t1, v1 = _wrap(TorchArrayContext("cpu"), sc, th, vh)
        qi = q[offs[i]:offs[i + 1]]
        ref = oracle_fmm(t1, v1, qi, order, exclude_self=True, k=k)
        assert np.allclose(pot[offs[i]:offs[i + 1]], ref, rtol=1e-12, atol=1e-12*np.abs(ref).max())
        if len(p[0]) > 1:
            direct, _ = oracle.p2p_dense(13, p, p, qi, t2s=np.arange(len(p[0])), k=k)
            assert np.linalg.norm(ref - direct)/np.linalg.norm(direct) < 1e-3
