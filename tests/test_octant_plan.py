"""The octree level translation (M2M / L2L as eight fixed matrices applied to
gathered coefficient vectors) reproduces the per-box translation: the group
plan built by the operators, the fragment order of the matrices and the
kernel's loop are emulated on the CPU, with the matrices made from the
oracle's M2M / L2L applied to unit vectors, and compared with the oracle's
per-box M2M / L2L on a real tree."""

import numpy as np
import pytest
import torch

import oracle
from sumpy_amd.array_context import TorchArrayContext
from sumpy_amd.e2e import GROUP, E2EFromChildren, E2EFromParent, _fragment_order
from sumpy_amd.expansion.local import LinearPDEConformingVolumeTaylorLocalExpansion
from sumpy_amd.expansion.multipole import LinearPDEConformingVolumeTaylorMultipoleExpansion
from sumpy_amd.kernel import LaplaceKernel
from sumpy_amd.tree import build_tree_and_traversal


def rel(a, b):
    return np.linalg.norm(a - b)/np.linalg.norm(b)


@pytest.fixture(scope="module")
def tree():
    rng = np.random.default_rng(3)
    return build_tree_and_traversal(TorchArrayContext("cpu"), rng.random((3, 4000)), None, 20)


def _unit_matrices(kind, order, offsets, present, rs_src, rs_tgt):
    """T[o][:, k] = oracle translation of the unit vector e_k across offsets[o]."""
    n = (order + 1)**2
    T = np.zeros((8, n, n))
    nb = 1 + 2*n
    par = 1 + np.arange(n)
    chd = 1 + n + np.arange(n)
    for o in range(8):
        if not present[o]:
            continue
        centers = np.zeros((3, nb))
        src = np.eye(n)
        out = np.zeros((n, n))
        if kind == "m2m":
            centers[:, chd] = -offsets[o][:, None]
            child = np.zeros((8, nb), np.int32)
            child[0, par] = chd
            oracle.m2m(order, order, par, child, centers, rs_src, rs_tgt, src, int(chd[0]),
                       out, int(par[0]))
        else:
            centers[:, chd] = offsets[o][:, None]
            parent = np.zeros(nb, np.int32)
            parent[chd] = par
            oracle.l2l(order, order, chd, parent, centers, rs_src, rs_tgt, src, int(par[0]),
                       out, int(chd[0]))
        T[o] = out.T
    return T


def _decode(frag, n_out, n_in):
    nrb, ks = -(-n_out//16), -(-n_in//4)
    T = np.zeros((8, n_out, n_in))
    r, c = np.meshgrid(np.arange(n_out), np.arange(n_in), indexing="ij")
    for o in range(8):
        T[o] = frag[((o*nrb + r//16)*ks + c//4)*64 + (c % 4)*16 + r % 16]
    return T


def _emulate(plan, T, src, sbase, tgt, tbase):
    targets = plan["targets"].reshape(-1, GROUP)
    for g in range(plan["ngroups"]):
        for o in range(8):
            if not (plan["mask"][g] >> o) & 1:
                continue
            for j in range(GROUP):
                tb, ib = targets[g, j], plan["inputs"][g, o, j]
                if tb >= 0 and ib >= 0:
                    tgt[tb - tbase] += T[o] @ src[ib - sbase]


@pytest.mark.parametrize("kind", ["m2m", "l2l"])
def test_octant_plan_reproduces_per_box_translation(tree, kind):
    tree, trav = tree
    th, vh = tree.host, trav.host
    order = 4
    n = (order + 1)**2
    knl = LaplaceKernel(3)
    lbs = th["level_start_box_nrs"]
    rng = np.random.default_rng(1)
    centers_t = torch.from_numpy(th["box_centers"])
    if kind == "m2m":
        ex = LinearPDEConformingVolumeTaylorMultipoleExpansion(knl, order)
        op = E2EFromChildren(ex, ex)
        tl = tree.nlevels - 2
        sl = tl + 1
        lsp = vh["level_start_source_parent_box_nrs"]
        tb = vh["source_parent_boxes"][lsp[tl]:lsp[tl + 1]]
        plan = op._plan(torch.from_numpy(tb.astype(np.int32)),
                        torch.from_numpy(th["box_child_ids"].astype(np.int32)), centers_t)
    else:
        ex = LinearPDEConformingVolumeTaylorLocalExpansion(knl, order)
        op = E2EFromParent(ex, ex)
        sl = tree.nlevels - 2
        tl = sl + 1
        tb = np.arange(lbs[tl], lbs[tl + 1])
        plan = op._plan(torch.from_numpy(tb.astype(np.int32)),
                        torch.from_numpy(th["box_parent_ids"].astype(np.int32)), centers_t)
    assert plan is not None
    # every target in exactly one group
    tg = plan["targets"][plan["targets"] >= 0]
    assert np.array_equal(np.sort(tg), np.sort(tb))
    h_src = tree.root_extent*2.0**-sl
    h_tgt = tree.root_extent*2.0**-tl
    rs_src, rs_tgt = 2*h_src/order, 2*h_tgt/order
    T = _unit_matrices(kind, order, plan["offsets"], plan["present"], rs_src, rs_tgt)
    frag = _fragment_order(torch.from_numpy(T)).numpy()
    assert np.array_equal(_decode(frag, n, n), T)

    src = rng.standard_normal((lbs[sl + 1] - lbs[sl], n))
    got = np.zeros((lbs[tl + 1] - lbs[tl], n))
    _emulate(plan, T, src, int(lbs[sl]), got, int(lbs[tl]))
    want = np.zeros_like(got)
    if kind == "m2m":
        oracle.m2m(order, order, tb, th["box_child_ids"], th["box_centers"], rs_src, rs_tgt,
                   src, int(lbs[sl]), want, int(lbs[tl]))
    else:
        oracle.l2l(order, order, tb, th["box_parent_ids"], th["box_centers"], rs_src, rs_tgt,
                   src, int(lbs[sl]), want, int(lbs[tl]))
    assert rel(got, want) < 1e-13
