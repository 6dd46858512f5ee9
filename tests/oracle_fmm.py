"""An FMM pass driven entirely by the CPU oracle (test helper).

Same stage order as ``sumpy_amd.drive.drive_fmm`` / boxtree's ``drive_fmm``
over the same tree and traversal, with every operator taken from
``oracle`` (the FP64 restatements of the reference's CPU algorithms), so that
the GPU FMM can be compared with it to rounding.
"""

import numpy as np

import oracle


def rscale(tree, lev, order):
    return tree.root_extent*2**-lev*2/order


def oracle_fmm(tree, trav, weights, order, exclude_self=False):
    th, vh = tree.host, trav.host
    lbs = th["level_start_box_nrs"]
    nlev = tree.nlevels
    ncoef = (order + 1)**2
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
                       rscale(tree, lev, order), view(mp, lev), int(lbs[lev]))
    # M2M
    lsp = vh["level_start_source_parent_box_nrs"]
    for sl in range(nlev - 1, 2, -1):
        tl = sl - 1
        pb = vh["source_parent_boxes"][lsp[tl]:lsp[tl + 1]]
        if len(pb):
            oracle.m2m(order, order, pb, th["box_child_ids"], centers,
                       rscale(tree, sl, order), rscale(tree, tl, order), view(mp, sl),
                       int(lbs[sl]), view(mp, tl), int(lbs[tl]))
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
        data = oracle.m2l_class_data(order, order, vecs[:, c0:c1], rs)
        oracle.m2l(order, order, totp[s0:s1], vh["from_sep_siblings_starts"][s0:s1 + 1],
                   vh["from_sep_siblings_lists"], vh["from_sep_siblings_translation_classes"],
                   c0, data, rs, rs, view(mp, lev), int(lbs[lev]), view(loc, lev),
                   int(lbs[lev]))
    # M2P (list 3)
    for slev, (tbs, st, li) in enumerate(vh["sep_smaller"]):
        if len(tbs):
            oracle.m2p(order, tbs, st, li, th, centers, th["targets"],
                       rscale(tree, slev, order), view(mp, slev), int(lbs[slev]), pot)
    # P2L (list 4)
    loc4 = np.zeros(nboxes*ncoef)
    st4, li4 = vh["from_sep_bigger_starts"], vh["from_sep_bigger_lists"]
    for lev in range(nlev):
        s0, s1 = lst[lev], lst[lev + 1]
        if s0 == s1 or st4[s1] == st4[s0]:
            continue
        oracle.p2l(order, totp[s0:s1], st4[s0:s1 + 1], li4, th, centers, th["sources"], q,
                   None, rscale(tree, lev, order), view(loc4, lev), int(lbs[lev]))
    loc += loc4
    # L2L
    for tl in range(1, nlev):
        s0, s1 = lst[tl], lst[tl + 1]
        if s0 == s1:
            continue
        sl = tl - 1
        oracle.l2l(order, order, totp[s0:s1], th["box_parent_ids"], centers,
                   rscale(tree, sl, order), rscale(tree, tl, order), view(loc, sl),
                   int(lbs[sl]), view(loc, tl), int(lbs[tl]))
    # L2P
    ltb = vh["level_start_target_box_nrs"]
    lpot = np.zeros(nt)
    for lev in range(nlev):
        tb = vh["target_boxes"][ltb[lev]:ltb[lev + 1]]
        if len(tb):
            oracle.l2p(order, tb, th, centers, th["targets"], rscale(tree, lev, order),
                       view(loc, lev), int(lbs[lev]), lpot)
    pot = pot + lpot
    return pot[th["sorted_target_ids"]]
