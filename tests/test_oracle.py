"""Pinning the CPU oracle (no GPU): golden vectors from mpmath/sympy
(tests/golden/make_golden.py), closed forms, and the mathematical invariants
that a dropped term, a wrong sign or index, or a transposed operand breaks."""

import os

import numpy as np
import pytest

import oracle
from oracle import tables

HERE = os.path.dirname(os.path.abspath(__file__))
G = os.path.join(HERE, "golden")


def rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b))/np.linalg.norm(np.ravel(b))


@pytest.fixture(scope="module")
def p2p_gold():
    return np.load(os.path.join(G, "p2p_direct.npz"))


# {{{ P2P against 40-digit direct sums

@pytest.mark.parametrize("dim", [2, 3])
def test_p2p_laplace_golden(p2p_gold, dim):
    g = p2p_gold
    src, tgt, q, wd = (g[f"lap{dim}_{n}"] for n in ("src", "tgt", "q", "wd"))
    pot, grad = oracle.p2p_dense(dim, tgt, src, q, scale=1.0, want_grad=True)
    assert rel(pot, g[f"lap{dim}_pot"]) < 1e-14
    assert rel(np.array(grad), g[f"lap{dim}_grad"]) < 1e-14
    potd, _ = oracle.p2p_dense(dim, tgt, src, np.zeros_like(q), wd=list(wd), scale=1.0)
    assert rel(potd, g[f"lap{dim}_potdip"]) < 1e-14


def test_p2p_helmholtz_golden(p2p_gold):
    g = p2p_gold
    pot, grad = oracle.p2p_dense(13, g["helm3_tgt"], g["helm3_src"],
                                 g["helm3_q"].astype(complex), k=float(g["helm3_k"]),
                                 scale=1.0, want_grad=True)
    assert rel(pot, g["helm3_pot"]) < 1e-14
    assert rel(np.array(grad), g["helm3_grad"]) < 1e-14


def test_p2p_dipole_hessian_is_gradient_of_dipole():
    """Gradient of the dipole potential against central differences."""
    rng = np.random.default_rng(0)
    for knl, dim in ((3, 3), (2, 2), (13, 3)):
        src, tgt = rng.random((dim, 5)), rng.random((dim, 3)) + 1.0
        dt = complex if knl == 13 else float
        wd = [rng.standard_normal(5).astype(dt) for _ in range(dim)]
        _, grad = oracle.p2p_dense(knl, tgt, src, np.zeros(5, dt), wd=wd, k=2.0,
                                   want_grad=True)
        h = 1e-6
        for a in range(dim):
            tp, tm = tgt.copy(), tgt.copy()
            tp[a] += h
            tm[a] -= h
            pp, _ = oracle.p2p_dense(knl, tp, src, np.zeros(5, dt), wd=wd, k=2.0)
            pm, _ = oracle.p2p_dense(knl, tm, src, np.zeros(5, dt), wd=wd, k=2.0)
            assert rel(grad[a], (pp - pm)/(2*h)) < 1e-7


def test_p2p_exclude_self_and_csr():
    from sumpy_amd.array_context import TorchArrayContext
    from sumpy_amd.tree import build_tree_and_traversal
    rng = np.random.default_rng(1)
    n = 1500
    src = rng.random((3, n))
    q = rng.standard_normal(n)
    t2s = np.arange(n, dtype=np.int32)
    pot, _ = oracle.p2p_dense(3, src, src, q, t2s=t2s)
    d = src[:, :, None] - src[:, None, :]
    r = np.sqrt((d**2).sum(0))
    np.fill_diagonal(r, np.inf)
    assert rel(pot, (q/r).sum(1)/(4*np.pi)) < 1e-13
    # CSR over all-boxes-in-list-1 of a one-level tree equals the dense sum
    tree, trav = build_tree_and_traversal(TorchArrayContext("cpu"), src, None, n + 1)
    th, vh = tree.host, trav.host
    pcsr, _ = oracle.p2p_csr(3, th, vh["target_boxes"], vh["neighbor_source_boxes_starts"],
                             vh["neighbor_source_boxes_lists"], th["targets"], th["sources"],
                             q[th["user_source_ids"]], t2s=t2s)
    assert rel(pcsr[th["sorted_target_ids"]], pot) < 1e-13

# }}}


# {{{ derivatives

def test_laplace_derivatives_golden():
    g = np.load(os.path.join(G, "laplace3d_derivatives.npz"))
    boxes = [oracle.lap3d_derivs(p, float(g["rscale"]), 12, (12, 12, 12)) for p in g["points"]]
    got = np.array([boxes[pi][mx, my, mz] for pi, mx, my, mz in g["mis"]])
    assert np.max(np.abs(got - g["values"])/np.abs(g["values"]).max()) < 1e-13
    assert rel(got, g["values"]) < 1e-13


def test_laplace_derivatives_harmonic():
    """Every derivative of 1/r is harmonic: D(m+2x) + D(m+2y) + D(m+2z) = 0."""
    D = oracle.lap3d_derivs(np.array([0.3, -0.8, 0.45]), 0.2, 14, (14, 14, 14))
    for m in [(0, 0, 0), (1, 2, 3), (4, 0, 5), (2, 2, 2), (0, 6, 5)]:
        s = (D[m[0] + 2, m[1], m[2]] + D[m[0], m[1] + 2, m[2]] + D[m[0], m[1], m[2] + 2])
        assert abs(s) < 1e-12*abs(D[m[0] + 2, m[1], m[2]])

# }}}


# {{{ expansions

def _tree(n=2500, seed=3, maxp=30):
    from sumpy_amd.array_context import TorchArrayContext
    from sumpy_amd.tree import build_tree_and_traversal
    rng = np.random.default_rng(seed)
    src = rng.random((3, n))
    tree, trav = build_tree_and_traversal(TorchArrayContext("cpu"), src, None, maxp)
    return tree, trav, rng


@pytest.mark.parametrize("compressed", [True, False])
def test_p2m_m2p_against_direct(compressed):
    """A multipole of a box evaluated far away converges to the direct field
    (order by order), and the compressed and full forms agree."""
    tree, trav, rng = _tree()
    th, vh = tree.host, trav.host
    lev = tree.nlevels - 1
    lsb, lbs = vh["level_start_source_box_nrs"], th["level_start_box_nrs"]
    b = vh["source_boxes"][lsb[lev]]
    q = rng.standard_normal(th["sources"].shape[1])
    s0, ns = th["box_source_starts"][b], th["box_source_counts_nonchild"][b]
    c = th["box_centers"][:, b]
    h = tree.root_extent*2.0**-lev
    far = c[:, None] + 3*h*np.array([[1.0, 0.3, -0.2], [0.5, 1.1, 0.7]]).T
    direct, _ = oracle.p2p_dense(3, far, th["sources"][:, s0:s0 + ns], q[s0:s0 + ns])
    errs = []
    for order in (2, 4, 6, 8):
        w = tables.wrangler(order, 3, compressed)
        nb = lbs[lev + 1] - lbs[lev]
        M = np.zeros((nb, len(w.stored)))
        rs = h*2/order
        oracle.p2m(order, [b], th, th["box_centers"], th["sources"], q, None, rs, M,
                   int(lbs[lev]), compressed=compressed)
        # evaluate with the full derivative set (P2M stored coefficients times
        # rho^|m| d^m G): a single-target "tree" with target box = b
        Dv = []
        for t in range(far.shape[1]):
            D = oracle.lap3d_derivs(far[:, t] - c, rs, order, (order,)*3)
            Dv.append(sum(M[b - lbs[lev], i]*D[mi] for i, mi in enumerate(w.stored))
                      / (4*np.pi))
        errs.append(rel(np.array(Dv), direct))
    assert all(e1 < e0 for e0, e1 in zip(errs, errs[1:])), errs
    assert errs[-1] < 1e-4


def test_compressed_and_full_multipoles_evaluate_equal():
    tree, trav, rng = _tree()
    th, vh = tree.host, trav.host
    lev = tree.nlevels - 1
    lsb, lbs = vh["level_start_source_box_nrs"], th["level_start_box_nrs"]
    b = vh["source_boxes"][lsb[lev]]
    q = rng.standard_normal(th["sources"].shape[1])
    c = th["box_centers"][:, b]
    order, rs = 6, 0.05
    x = c + np.array([0.31, -0.22, 0.4])
    D = oracle.lap3d_derivs(x - c, rs, order, (order,)*3)
    vals = []
    for compressed in (True, False):
        w = tables.wrangler(order, 3, compressed)
        M = np.zeros((lbs[lev + 1] - lbs[lev], len(w.stored)))
        oracle.p2m(order, [b], th, th["box_centers"], th["sources"], q, None, rs, M,
                   int(lbs[lev]), compressed=compressed)
        vals.append(sum(M[b - lbs[lev], i]*D[mi] for i, mi in enumerate(w.stored)))
    assert abs(vals[0] - vals[1]) < 1e-12*abs(vals[1])


def test_m2m_shift_preserves_far_field():
    """Translating a multipole to the parent changes the far field only by
    the truncation error, which shrinks with the order."""
    tree, trav, rng = _tree()
    th, vh = tree.host, trav.host
    lbs = th["level_start_box_nrs"]
    tl = tree.nlevels - 2
    sl = tl + 1
    lsp = vh["level_start_source_parent_box_nrs"]
    pb = vh["source_parent_boxes"][lsp[tl]]
    q = rng.standard_normal(th["sources"].shape[1])
    c = th["box_centers"][:, pb]
    h = tree.root_extent*2.0**-tl
    x = c + 4*h*np.array([0.8, -0.5, 0.6])
    s0, ns = th["box_source_starts"][pb], th["box_source_counts_cumul"][pb]
    direct, _ = oracle.p2p_dense(3, x[:, None], th["sources"][:, s0:s0 + ns], q[s0:s0 + ns],
                                 scale=1.0)
    errs = []
    for order in (3, 6, 9):
        w = tables.wrangler(order)
        ncf = len(w.stored)
        Ms = np.zeros((lbs[sl + 1] - lbs[sl], ncf))
        lsb = vh["level_start_source_box_nrs"]
        sboxes = vh["source_boxes"][lsb[sl]:lsb[sl + 1]]
        rs_s, rs_t = h/2*2/order, h*2/order
        oracle.p2m(order, sboxes, th, th["box_centers"], th["sources"], q, None, rs_s, Ms,
                   int(lbs[sl]))
        Mt = np.zeros((lbs[tl + 1] - lbs[tl], ncf))
        oracle.m2m(order, order, [pb], th["box_child_ids"], th["box_centers"], rs_s, rs_t, Ms,
                   int(lbs[sl]), Mt, int(lbs[tl]))
        D = oracle.lap3d_derivs(x - c, rs_t, order, (order,)*3)
        v = sum(Mt[pb - lbs[tl], i]*D[mi] for i, mi in enumerate(w.stored))
        errs.append(abs(v - direct[0])/abs(direct[0]))
    assert errs[0] > errs[1] > errs[2] and errs[2] < 1e-4, errs


def test_m2l_toeplitz_equals_naive_sum():
    """Reference test_kernels.py:692-760: the Toeplitz M2L equals
    L[t] = sum_s M[s] rho^|t+s| d^{t+s} G (times the rscale ratio)."""
    order = 5
    rng = np.random.default_rng(44)
    T = tables.m2l_tables(order, order)
    w = tables.wrangler(order)
    d = np.array([[2.0], [1.0], [-1.5]])
    rs = 0.4
    data = oracle.m2l_class_data(order, order, d, rs)
    M = rng.standard_normal((1, len(w.stored)))
    L = np.zeros((1, len(w.stored)))
    oracle.m2l(order, order, [0], [0, 1], [0], [0], 0, data, rs, rs, M, 0, L, 0)
    Dfull = oracle.lap3d_derivs(d[:, 0], rs, 2*order, (2*order,)*3)
    naive = [sum(M[0, j]*Dfull[tuple(np.add(t, s))] for j, s in enumerate(w.stored))
             for t in w.stored]
    assert rel(L[0], naive) < 1e-12
    assert T["ncirc"] == 3*(2*order + 1)**2


def _local_setup(nsrc=20, ntgt=12, seed=5):
    """Three boxes: 0 = parent target box at the origin (size 1), 1 = its child
    at (1/4, -1/4, 1/4), 2 = a far source box at (3, 0.5, -0.7)."""
    rng = np.random.default_rng(seed)
    centers = np.array([[0.0, 0.25, 3.0], [0.0, -0.25, 0.5], [0.0, 0.25, -0.7]])
    src = centers[:, 2:3] + 0.3*(rng.random((3, nsrc)) - 0.5)
    tgt = centers[:, 1:2] + 0.2*(rng.random((3, ntgt)) - 0.5)
    q = rng.standard_normal(nsrc)
    wd = [rng.standard_normal(nsrc) for _ in range(3)]

    def arrays(tgt_box):
        cnt = np.zeros(3, np.int32)
        cnt[tgt_box] = ntgt
        return {"box_source_starts": np.zeros(3, np.int32),
                "box_source_counts_nonchild": np.array([0, 0, nsrc], np.int32),
                "box_target_starts": np.zeros(3, np.int32),
                "box_target_counts_nonchild": cnt}
    return centers, src, tgt, q, wd, arrays


@pytest.mark.parametrize("compressed", [True, False])
@pytest.mark.parametrize("dipole", [False, True])
def test_p2l_l2p_converge_and_l2l_is_exact(compressed, dipole):
    """P2L + L2P of a far cluster converges to the direct field with the
    order; L2L re-expands the truncated (harmonic) Taylor polynomial exactly,
    so L2P after L2L equals L2P of the parent expansion at the same points."""
    centers, src, tgt, q, wd, arrays = _local_setup()
    wdd = np.array(wd) if dipole else None
    direct, _ = oracle.p2p_dense(3, tgt, src, q, wd=wdd)
    errs = []
    for order in (3, 6, 9):
        w = tables.wrangler(order, 3, compressed)
        rs0, rs1 = 2.0/order, 1.0/order
        L0 = np.zeros((3, len(w.stored)))
        oracle.p2l(order, [0], [0, 1], [2], arrays(0), centers, src, q,
                   wd if dipole else None, rs0, L0, 0, compressed=compressed)
        pot0 = np.zeros(tgt.shape[1])
        oracle.l2p(order, [0], arrays(0), centers, tgt, rs0, L0, 0, pot0,
                   compressed=compressed)
        L1 = np.zeros((3, len(w.stored)))
        oracle.l2l(order, order, [1], np.zeros(3, np.int32), centers, rs0, rs1, L0, 0,
                   L1, 0, compressed=compressed)
        pot1 = np.zeros(tgt.shape[1])
        oracle.l2p(order, [1], arrays(1), centers, tgt, rs1, L1, 0, pot1,
                   compressed=compressed)
        assert rel(pot1, pot0) < 1e-12
        errs.append(rel(pot0, direct))
    assert errs[0] > errs[1] > errs[2] and errs[2] < 1e-4, errs


def test_m2p_matches_derivative_sum_and_direct():
    """M2P of a box's multipole equals sum_i M_i rho^|m_i| d^m_i G (derivatives
    pinned by the sympy fixture) and converges to the direct field."""
    centers, src, tgt, q, wd, arrays = _local_setup()
    direct, _ = oracle.p2p_dense(3, tgt, src, q)
    errs = []
    for order in (3, 6, 9):
        w = tables.wrangler(order)
        rs = 1.0/order
        M = np.zeros((3, len(w.stored)))
        oracle.p2m(order, [2], arrays(1), centers, src, q, None, rs, M, 0)
        pot = np.zeros(tgt.shape[1])
        oracle.m2p(order, [1], [0, 1], [2], arrays(1), centers, tgt, rs, M, 0, pot)
        expl = []
        for t in range(tgt.shape[1]):
            D = oracle.lap3d_derivs(tgt[:, t] - centers[:, 2], rs, order, (order,)*3)
            expl.append(sum(M[2, i]*D[mi] for i, mi in enumerate(w.stored))/(4*np.pi))
        assert rel(pot, expl) < 1e-12
        errs.append(rel(pot, direct))
    assert errs[0] > errs[1] > errs[2] and errs[2] < 1e-4, errs


@pytest.mark.parametrize("order", [4, 8])
def test_oracle_fmm_converges_to_direct(order):
    import sys
    sys.path.insert(0, HERE)
    from oracle_fmm import oracle_fmm
    tree, trav, rng = _tree(n=3000, seed=0)
    src = tree.host["sources"][:, tree.host["sorted_target_ids"]]
    q = rng.standard_normal(3000)
    pot = oracle_fmm(tree, trav, q, order, exclude_self=True)
    direct, _ = oracle.p2p_dense(3, src, src, q, t2s=np.arange(3000, dtype=np.int32))
    assert rel(pot, direct) < {4: 2e-3, 8: 5e-5}[order]


def test_oracle_fmm_full_taylor_equals_compressed():
    """Full Taylor expansions (every multi-index of degree <= p stored, the
    ``VolumeTaylor*`` path) and the Laplace-compressed ones represent the same
    harmonic polynomials, so the two oracle FMMs agree to rounding; a
    full-Taylor chain with a dropped or mis-indexed coefficient would not."""
    import sys
    sys.path.insert(0, HERE)
    from oracle_fmm import oracle_fmm
    tree, trav, rng = _tree(n=3000, seed=0)
    src = tree.host["sources"][:, tree.host["sorted_target_ids"]]
    q = rng.standard_normal(3000)
    full = oracle_fmm(tree, trav, q, 4, exclude_self=True, compressed=False)
    comp = oracle_fmm(tree, trav, q, 4, exclude_self=True)
    assert rel(full, comp) < 1e-12
    direct, _ = oracle.p2p_dense(3, src, src, q, t2s=np.arange(3000, dtype=np.int32))
    assert rel(full, direct) < 2e-3

# }}}


# {{{ QBX

def test_qbx_line_taylor_golden():
    g = np.load(os.path.join(G, "qbx_line_taylor.npz"))
    a, b, n, order = g["a"], g["b"], g["n"], int(g["order"])
    # a = c - s, b = t - c: choose s = 0, c = a, t = a + b
    s = np.zeros_like(a)
    c = a
    t = a + b
    got = np.zeros((4, a.shape[1]))
    for i in range(a.shape[1]):
        sl = (slice(None), slice(i, i + 1))
        args = dict(targets=t[sl], centers=c[sl], radii=np.ones(1), sources=s[sl],
                    normals=n[sl], density=np.ones(1), scale=1.0)
        got[0, i] = oracle.qbx2d(order, 0, -1, **args)[0]
        got[1, i] = oracle.qbx2d(order, 1, -1, **args)[0]
        got[2, i] = oracle.qbx2d(order, 0, 0, **args)[0]
        got[3, i] = oracle.qbx2d(order, 1, 1, **args)[0]
    assert np.max(np.abs(got - g["values"])) < 1e-12*np.max(np.abs(g["values"]))


def test_qbx_circle_eigenvalue():
    """Reference test_qbx.py:50-118: the 2D Laplace single layer on the unit
    circle has eigenvalue 1/(2n) for the density cos(n t)."""
    mode = 5
    errs = []
    for nsrc in (200, 400, 800):
        t = np.linspace(0, 2*np.pi, nsrc, endpoint=False)
        pts = np.array([np.cos(t), np.sin(t)])
        h = 2*np.pi/nsrc
        ctr = pts*(1 - 2*h)
        dens = np.cos(mode*t)*h  # quadrature weight folded into the strength
        pot = oracle.qbx2d(10, 0, -1, pts, ctr, np.ones(nsrc), pts, None, dens)
        errs.append(np.max(np.abs(pot - np.cos(mode*t)/(2*mode))))
    assert errs[-1] < errs[0] and errs[-1] < 1e-3, errs

# }}}
