"""An FMM pass driven entirely by the CPU oracle (test helper).

Same stage order as ``sumpy_amd.drive.drive_fmm`` / boxtree's ``drive_fmm``
over the same tree and traversal, with every operator taken from
``oracle`` (the FP64 restatements of the reference's CPU algorithms), so that
the GPU FMM can be compared with it to rounding.
"""

import numpy as np

import oracle
from oracle import helmholtz as oh
from oracle import tables


def rscale(tree, lev, order):
    return tree.root_extent*2**-lev*2/order


def helm_m2l(order, target_boxes, starts, lists, class_lists, c0, data, mp, mbase, loc,
             lbase):
    """L[t] = sum_s M[s] A_c[off(t) + off(s)] with complex class data
    (nclasses, 2, ncirc): the non-FFT M2L (``sumpy/expansion/m2l.py:363-389``)
    written as the dense class matrix."""
    T = tables.m2l_tables(order, order)
    A = data[:, 0] + 1j*data[:, 1]
    idx = T["tgt_circ_idx"][:, None] + T["src_circ_idx"][None, :]
    for i, tb in enumerate(target_boxes):
        acc = np.zeros(idx.shape[0], np.complex128)
        for q in range(int(starts[i]), int(starts[i + 1])):
            sb = int(lists[q])
            acc += A[int(class_lists[q]) - c0][idx] @ mp[sb - mbase]
        loc[tb - lbase] = acc


def oracle_fmm(tree, trav, weights, order, exclude_self=False, k=0.0, compressed=True):
    """*k* != 0: Helmholtz-3D with complex weights (complex expansions; the
    table-driven stages run on the real and imaginary planes).
    *compressed* = False: full Taylor expansions (Laplace), every multi-index
    of degree <= order stored."""
    if k:
        assert compressed
        return _oracle_fmm_helm(tree, trav, weights, order, exclude_self, k)
    th, vh = tree.host, trav.host
    lbs = th["level_start_box_nrs"]
    nlev = tree.nlevels
    cmp = {"compressed": compressed}
    ncoef = len(tables.wrangler(order, 3, compressed).stored)
    q = np.asarray(weights)[th["user_source_ids"]]
    centers = th["box_centers"]
    nboxes = len(th["box_levels"])

    def view(a, lev):
        return a[lbs[lev]*ncoef:lbs[lev + 1]*ncoef].reshape(-1, ncoef)

    # P2M
    mp = np.zeros(nboxes*ncoef)
    lsb = vh["level_start_source_box_nrs"]
    for lev in range(nlev):
        sb = vh["source_boxes"][lsb[lev]:lsb[lev + 1]]
        if len(sb):
            oracle.p2m(order, sb, th, centers, th["sources"], q, None,
                       rscale(tree, lev, order), view(mp, lev), int(lbs[lev]), **cmp)
    # M2M
    lsp = vh["level_start_source_parent_box_nrs"]
    for sl in range(nlev - 1, 2, -1):
        tl = sl - 1
        pb = vh["source_parent_boxes"][lsp[tl]:lsp[tl + 1]]
        if len(pb):
            oracle.m2m(order, order, pb, th["box_child_ids"], centers,
                       rscale(tree, sl, order), rscale(tree, tl, order), view(mp, sl),
                       int(lbs[sl]), view(mp, tl), int(lbs[tl]), **cmp)
    # P2P (list 1)
    t2s = np.argsort(th["user_source_ids"]).astype(np.int32)  # unused unless exclude_self
    nt = th["targets"].shape[1]
    t2s = np.arange(nt, dtype=np.int32) if exclude_self else None
    pot, _ = oracle.p2p_csr(3, th, vh["target_boxes"], vh["neighbor_source_boxes_starts"],
                            vh["neighbor_source_boxes_lists"], th["targets"], th["sources"],
                            q, t2s=t2s)
    # M2L (list 2)
    loc = np.zeros(nboxes*ncoef)
    totp = vh["target_or_target_parent_boxes"]
    lst = vh["level_start_target_or_target_parent_box_nrs"]
    cls_ls = vh["from_sep_siblings_translation_classes_level_starts"]
    vecs = vh["from_sep_siblings_translation_class_to_distance_vector"]
    for lev in range(nlev):
        c0, c1 = int(cls_ls[lev]), int(cls_ls[lev + 1])
        s0, s1 = lst[lev], lst[lev + 1]
        if c0 == c1 or s0 == s1:
            continue
        rs = rscale(tree, lev, order)
        data = oracle.m2l_class_data(order, order, vecs[:, c0:c1], rs, **cmp)
        oracle.m2l(order, order, totp[s0:s1], vh["from_sep_siblings_starts"][s0:s1 + 1],
                   vh["from_sep_siblings_lists"], vh["from_sep_siblings_translation_classes"],
                   c0, data, rs, rs, view(mp, lev), int(lbs[lev]), view(loc, lev),
                   int(lbs[lev]), **cmp)
    # M2P (list 3)
    for slev, (tbs, st, li) in enumerate(vh["sep_smaller"]):
        if len(tbs):
            oracle.m2p(order, tbs, st, li, th, centers, th["targets"],
                       rscale(tree, slev, order), view(mp, slev), int(lbs[slev]), pot, **cmp)
    # P2L (list 4)
    loc4 = np.zeros(nboxes*ncoef)
    st4, li4 = vh["from_sep_bigger_starts"], vh["from_sep_bigger_lists"]
    for lev in range(nlev):
        s0, s1 = lst[lev], lst[lev + 1]
        if s0 == s1 or st4[s1] == st4[s0]:
            continue
        oracle.p2l(order, totp[s0:s1], st4[s0:s1 + 1], li4, th, centers, th["sources"], q,
                   None, rscale(tree, lev, order), view(loc4, lev), int(lbs[lev]), **cmp)
    loc += loc4
    # L2L
    for tl in range(1, nlev):
        s0, s1 = lst[tl], lst[tl + 1]
        if s0 == s1:
            continue
        sl = tl - 1
        oracle.l2l(order, order, totp[s0:s1], th["box_parent_ids"], centers,
                   rscale(tree, sl, order), rscale(tree, tl, order), view(loc, sl),
                   int(lbs[sl]), view(loc, tl), int(lbs[tl]), **cmp)
    # L2P
    ltb = vh["level_start_target_box_nrs"]
    lpot = np.zeros(nt)
    for lev in range(nlev):
        tb = vh["target_boxes"][ltb[lev]:ltb[lev + 1]]
        if len(tb):
            oracle.l2p(order, tb, th, centers, th["targets"], rscale(tree, lev, order),
                       view(loc, lev), int(lbs[lev]), lpot, **cmp)
    pot = pot + lpot
    return pot[th["sorted_target_ids"]]


def _oracle_fmm_helm(tree, trav, weights, order, exclude_self, k):
    th, vh = tree.host, trav.host
    lbs = th["level_start_box_nrs"]
    nlev = tree.nlevels
    ncoef = (order + 1)**2
    q = np.asarray(weights, np.complex128)[th["user_source_ids"]]
    centers = th["box_centers"]
    nboxes = len(th["box_levels"])

    def view(a, lev):
        return a[lbs[lev]*ncoef:lbs[lev + 1]*ncoef].reshape(-1, ncoef)

    def both(fn, src, dst):
        """fn(src plane, dst plane) on real and imaginary parts; dst complex."""
        re, im = np.ascontiguousarray(dst.real), np.ascontiguousarray(dst.imag)
        fn(None if src is None else np.ascontiguousarray(src.real), re, 0)
        fn(None if src is None else np.ascontiguousarray(src.imag), im, 1)
        dst[...] = re + 1j*im

    mp = np.zeros(nboxes*ncoef, np.complex128)
    lsb = vh["level_start_source_box_nrs"]
    for lev in range(nlev):
        sb = vh["source_boxes"][lsb[lev]:lsb[lev + 1]]
        if len(sb):
            rs = rscale(tree, lev, order)
            both(lambda _, out, part: oracle.p2m(
                order, sb, th, centers, th["sources"],
                np.ascontiguousarray(q.imag if part else q.real), None, rs, out,
                int(lbs[lev]), k=k), None, view(mp, lev))
    lsp = vh["level_start_source_parent_box_nrs"]
    for sl in range(nlev - 1, 2, -1):
        tl = sl - 1
        pb = vh["source_parent_boxes"][lsp[tl]:lsp[tl + 1]]
        if len(pb):
            both(lambda src, out, _: oracle.m2m(
                order, order, pb, th["box_child_ids"], centers, rscale(tree, sl, order),
                rscale(tree, tl, order), src, int(lbs[sl]), out, int(lbs[tl]), k=k),
                view(mp, sl), view(mp, tl))
    nt = th["targets"].shape[1]
    t2s = np.arange(nt, dtype=np.int32) if exclude_self else None
    pot, _ = oracle.p2p_csr(13, th, vh["target_boxes"], vh["neighbor_source_boxes_starts"],
                            vh["neighbor_source_boxes_lists"], th["targets"], th["sources"],
                            q, t2s=t2s, k=k)
    loc = np.zeros(nboxes*ncoef, np.complex128)
    totp = vh["target_or_target_parent_boxes"]
    lst = vh["level_start_target_or_target_parent_box_nrs"]
    cls_ls = vh["from_sep_siblings_translation_classes_level_starts"]
    vecs = vh["from_sep_siblings_translation_class_to_distance_vector"]
    shape = tuple(int(m) + 1 for m in tables.m2l_tables(order, order)["max_mi"])
    for lev in range(nlev):
        c0, c1 = int(cls_ls[lev]), int(cls_ls[lev + 1])
        s0, s1 = lst[lev], lst[lev + 1]
        if c0 == c1 or s0 == s1:
            continue
        rs = rscale(tree, lev, order)
        data = oh.class_data(order, order, vecs[:, c0:c1], rs, k, shape)
        helm_m2l(order, totp[s0:s1], vh["from_sep_siblings_starts"][s0:s1 + 1],
                 vh["from_sep_siblings_lists"], vh["from_sep_siblings_translation_classes"],
                 c0, data, view(mp, lev), int(lbs[lev]), view(loc, lev), int(lbs[lev]))
    for slev, (tbs, st, li) in enumerate(vh["sep_smaller"]):
        if len(tbs):
            oh.m2p(order, tbs, st, li, th, centers, th["targets"], rscale(tree, slev, order),
                   k, view(mp, slev), int(lbs[slev]), pot)
    loc4 = np.zeros(nboxes*ncoef, np.complex128)
    st4, li4 = vh["from_sep_bigger_starts"], vh["from_sep_bigger_lists"]
    for lev in range(nlev):
        s0, s1 = lst[lev], lst[lev + 1]
        if s0 == s1 or st4[s1] == st4[s0]:
            continue
        oh.p2l(order, totp[s0:s1], st4[s0:s1 + 1], li4, th, centers, th["sources"], q, None,
               rscale(tree, lev, order), k, view(loc4, lev), int(lbs[lev]))
    loc += loc4
    for tl in range(1, nlev):
        s0, s1 = lst[tl], lst[tl + 1]
        if s0 == s1:
            continue
        sl = tl - 1
        both(lambda src, out, _: oracle.l2l(
            order, order, totp[s0:s1], th["box_parent_ids"], centers,
            rscale(tree, sl, order), rscale(tree, tl, order), src, int(lbs[sl]), out,
            int(lbs[tl]), k=k), view(loc, sl), view(loc, tl))
    ltb = vh["level_start_target_box_nrs"]
    lre, lim = np.zeros(nt), np.zeros(nt)
    for lev in range(nlev):
        tb = vh["target_boxes"][ltb[lev]:ltb[lev + 1]]
        if len(tb):
            lv = view(loc, lev)
            rs = rscale(tree, lev, order)
            oracle.l2p(order, tb, th, centers, th["targets"], rs,
                       np.ascontiguousarray(lv.real), int(lbs[lev]), lre, k=k)
            oracle.l2p(order, tb, th, centers, th["targets"], rs,
                       np.ascontiguousarray(lv.imag), int(lbs[lev]), lim, k=k)
    pot = pot + lre + 1j*lim
    return pot[th["sorted_target_ids"]]


def oracle_fmm_h2d(tree, trav, weights, order, k, exclude_self=False):
    """The 2D Helmholtz FMM with the H2D (Hankel/Bessel) expansions, every
    operator from ``oracle.h2d`` (scipy's Bessel functions), per box, in
    the stage order of ``drive_fmm``; the M2L through the translation-class
    Hankel rows (``sumpy/expansion/m2l.py:1008-1115``)."""
    from oracle import h2d
    th, vh = tree.host, trav.host
    nlev = tree.nlevels
    q = np.asarray(weights, np.complex128)[th["user_source_ids"]]
    C = th["box_centers"]
    S, T = th["sources"], th["targets"]
    nboxes = len(th["box_levels"])
    lev_of = th["box_levels"]
    bss, bsc = th["box_source_starts"], th["box_source_counts_nonchild"]
    bts, btc = th["box_target_starts"], th["box_target_counts_nonchild"]
    ncoef = 2*order + 1

    def rs(box):
        return rscale(tree, int(lev_of[box]), order)

    def srcs(b):
        return slice(int(bss[b]), int(bss[b] + bsc[b]))

    def tgts(b):
        return slice(int(bts[b]), int(bts[b] + btc[b]))

    mp = np.zeros((nboxes, ncoef), np.complex128)
    for b in vh["source_boxes"]:
        mp[b] = h2d.p2m(order, k, C[:, b], S[:, srcs(b)], q[srcs(b)], rs(b))
    lsp = vh["level_start_source_parent_box_nrs"]
    for sl in range(nlev - 1, 2, -1):
        for pb in vh["source_parent_boxes"][lsp[sl - 1]:lsp[sl]]:
            for ch in th["box_child_ids"][:, pb]:
                if ch:
                    mp[pb] += h2d.m2m(order, order, k, mp[ch], C[:, ch], C[:, pb], rs(ch),
                                      rs(pb))
    nt = T.shape[1]
    pot = np.zeros(nt, np.complex128)
    st1, li1 = vh["neighbor_source_boxes_starts"], vh["neighbor_source_boxes_lists"]
    for i, tb in enumerate(vh["target_boxes"]):
        sl = np.concatenate([np.arange(bss[s], bss[s] + bsc[s]) for s in li1[st1[i]:st1[i + 1]]]
                            + [np.zeros(0, np.int64)]).astype(np.int64)
        tt = np.arange(bts[tb], bts[tb] + btc[tb])
        g = np.zeros((len(tt), len(sl)), np.complex128)
        if len(sl):
            r = np.hypot(T[0, tt][:, None] - S[0, sl][None, :], T[1, tt][:, None] - S[1, sl][None, :])
            nz = r > 0
            if exclude_self:
                nz &= tt[:, None] != sl[None, :]
            from scipy import special
            g[nz] = h2d.SCALE*special.hankel1(0, k*r[nz])
        pot[tt] = g @ q[sl]
    loc = np.zeros((nboxes, ncoef), np.complex128)
    totp = vh["target_or_target_parent_boxes"]
    st2, li2 = vh["from_sep_siblings_starts"], vh["from_sep_siblings_lists"]
    cls = vh["from_sep_siblings_translation_classes"]
    vecs = vh["from_sep_siblings_translation_class_to_distance_vector"]
    for i, tb in enumerate(totp):
        for e in range(st2[i], st2[i + 1]):
            sb = li2[e]
            data = h2d.m2l_class_data(order, order, k, vecs[:, cls[e]])
            loc[tb] += h2d.m2l(order, order, k, mp[sb], None, rs(sb), rs(tb), data=data)
    for tbs, st, li in vh["sep_smaller"]:
        for i, tb in enumerate(tbs):
            for sb in li[st[i]:st[i + 1]]:
                pot[tgts(tb)] += h2d.m2p(order, k, C[:, sb], mp[sb], T[:, tgts(tb)], rs(sb))
    st4, li4 = vh["from_sep_bigger_starts"], vh["from_sep_bigger_lists"]
    for i, tb in enumerate(totp):
        for sb in li4[st4[i]:st4[i + 1]]:
            loc[tb] += h2d.p2l(order, k, C[:, tb], S[:, srcs(sb)], q[srcs(sb)], rs(tb))
    lst = vh["level_start_target_or_target_parent_box_nrs"]
    for tl in range(1, nlev):
        for tb in totp[lst[tl]:lst[tl + 1]]:
            pb = th["box_parent_ids"][tb]
            loc[tb] += h2d.l2l(order, order, k, loc[pb], C[:, pb], C[:, tb], rs(pb), rs(tb))
    for tb in vh["target_boxes"]:
        pot[tgts(tb)] += h2d.l2p(order, k, C[:, tb], loc[tb], T[:, tgts(tb)], rs(tb))
    return pot[th["sorted_target_ids"]]
