"""LDS layout of the split-BF16 M2L contraction, and the offline tuner that
chooses it per expansion order.

The contraction (``m2l_apply_bf16_kernel``) stages one translation class's
coefficients in LDS as 32-bit words and gathers its MFMA A fragments from
them: lane ``l`` of a wave reads the word at ``off_t[row] + off_s[k]`` for
row ``16 rb + l % 16`` and the eight ``k = k0 + 8 (l / 16) + e``.  The words
are laid out like the reference's circulant array (``sumpy/expansion/m2l.py``
lines 403-537: the multi-index ``m`` of a (target + source) coefficient pair
sits at a linear offset, so that ``off(t + s) = off(t) + off(s)``), with
strides ``(S0, S1, 1)`` for ``(m_x, m_y, m_z)``.

LDS has 32 banks of 4 bytes and serves 32 lanes per cycle: a gather costs
one cycle per half wave plus one per extra distinct address in the busiest
bank.  With the plain layout (strides of the circulant shape, rows and k in
stored order) the order-10 gathers cost 2.9 cycles per half wave.  The
strides, the order of the target rows (the MFMA rows of the work array) and
the order of the source coefficients (the k of the packed multipoles) are
free: any choice gives the same sums.  :func:`tune` picks them by a stride
scan followed by simulated annealing over the two orders (deterministic
seed); ``python -m sumpy_amd.m2l_layout`` regenerates the committed table
``data/m2l_bf16_layouts.json`` for orders 1-10 (the orders the split-BF16
kernel supports).  The LDS bank model was measured on gfx950: a stride-32
ds_read_b32 pattern costs as much as a stride-64 one (32 banks per half
wave).
"""

from __future__ import annotations

import json
import os
from dataclasses import dataclass
from functools import lru_cache

import numpy as np

NBANKS = 32
HALF = 32
_TABLE = os.path.join(os.path.dirname(__file__), "data", "m2l_bf16_layouts.json")


def _pad(n, m):
    return -(-n//m)*m


def _stored(order):
    from sumpy_amd.expansion.terms import get_terms_table
    return get_terms_table(order, 3, True).stored.astype(np.int64)


@dataclass(frozen=True)
class Layout:
    """Strides of the packed class words and the row / k orders.

    ``rows[p]`` (``ks[p]``) is the stored coefficient index at work-row
    (packed-multipole) position ``p``, or -1 for padding."""
    order: int
    strides: tuple[int, int]
    rows: np.ndarray
    ks: np.ndarray

    @property
    def shape(self):
        n = 2*self.order + 1
        return (3, n, n)

    @property
    def ncirc(self):
        """Words per class (the last word is that of m = (2, 2p, 2p))."""
        p2 = 2*self.order
        return 2*self.strides[0] + p2*self.strides[1] + p2 + 1

    def offsets(self, order_ids):
        ids = _stored(self.order)
        o = ids[:, 0]*self.strides[0] + ids[:, 1]*self.strides[1] + ids[:, 2]
        return np.where(order_ids >= 0, o[np.maximum(order_ids, 0)], 0).astype(np.int32)

    @property
    def off_t(self):
        return self.offsets(self.rows)

    @property
    def off_s(self):
        return self.offsets(self.ks)

    @property
    def row_pos(self):
        """Work-row position of each stored target coefficient."""
        pos = np.zeros(int((self.rows >= 0).sum()), np.int32)
        live = np.nonzero(self.rows >= 0)[0]
        pos[self.rows[live]] = live
        return pos


def gather_cost(off_t, off_s):
    """Mean LDS cycles per half-wave A-fragment gather of the contraction."""
    total, n = _gather_cycles(off_t, off_s)
    return total/n


def _gather_cycles(off_t, off_s, nbanks=NBANKS, group=HALF):
    """(total LDS cycles, number of gathers of *group* lanes) over one class
    slot, for *nbanks* banks serving *group* lanes per cycle."""
    npad, kpad = len(off_t), len(off_s)
    lanes = np.arange(64)
    rb = np.arange(npad//16)[:, None, None, None]
    k0 = np.arange(0, kpad, 32)[None, :, None, None]
    e = np.arange(8)[None, None, :, None]
    a = (np.asarray(off_t, np.int64)[16*rb + lanes % 16]
         + np.asarray(off_s, np.int64)[k0 + 8*(lanes//16) + e]).reshape(-1, 64)
    tot = 0
    for h in range(0, 64, group):
        x = np.sort(a[:, h:h + group], axis=1)
        first = np.ones_like(x, bool)
        first[:, 1:] = x[:, 1:] != x[:, :-1]
        cnt = np.zeros((x.shape[0], nbanks), np.int64)
        r = np.broadcast_to(np.arange(x.shape[0])[:, None], x.shape)
        np.add.at(cnt, (r[first], (x % nbanks)[first]), 1)
        tot += int(cnt.max(axis=1).sum())
    return tot, a.shape[0]*(64//group)


def _objective(lay):
    """Annealing objective: the 32-bank half-wave cycles plus those of a
    64-bank whole-wave model.  The second term is not a cost of the hardware;
    it penalises same-bank pairs across the two halves as well, which makes
    the annealing reach much cheaper 32-bank layouts (order 10: 1.08 cycles
    per gather against 1.5 with the 32-bank term alone)."""
    return (_gather_cycles(lay.off_t, lay.off_s)[0]
            + _gather_cycles(lay.off_t, lay.off_s, 64, 64)[0])


def identity_layout(order):
    n = 2*order + 1
    nl = (order + 1)**2
    rows = np.full(_pad(nl, 16), -1, np.int64)
    rows[:nl] = np.arange(nl)
    ks = np.full(_pad(nl, 32), -1, np.int64)
    ks[:nl] = np.arange(nl)
    return Layout(order, (n*n, n), rows, ks)


def tune(order, iters=100000, nstrides=3, seed=0, log=None):
    """Stride scan, then simulated annealing over the row and k orders of
    the best *nstrides* stride pairs; returns the cheapest :class:`Layout`."""
    base = identity_layout(order)
    p2 = 2*order
    cands = []
    for s1 in range(p2 + 1, p2 + 21):
        for s0 in range(p2*s1 + p2 + 1, p2*s1 + p2 + 41):
            lay = Layout(order, (s0, s1), base.rows, base.ks)
            cands.append((gather_cost(lay.off_t, lay.off_s), lay.ncirc, s0, s1))
    cands.sort()
    best = None
    for c0, _, s0, s1 in cands[:nstrides]:
        rng = np.random.default_rng(seed)
        rows, ks = base.rows.copy(), base.ks.copy()
        cur = _objective(Layout(order, (s0, s1), rows, ks))
        for it in range(iters):
            # temperature in cycles summed over the slot's gathers
            temp = max(0.05, 2.0*0.9997**it)
            r2, k2 = rows.copy(), ks.copy()
            if rng.random() < 0.5:
                i, j = rng.integers(0, len(rows), 2)
                r2[i], r2[j] = r2[j], r2[i]
            else:
                i, j = rng.integers(0, len(ks), 2)
                k2[i], k2[j] = k2[j], k2[i]
            c = _objective(Layout(order, (s0, s1), r2, k2))
            if c <= cur or rng.random() < np.exp((cur - c)/temp):
                rows, ks, cur = r2, k2, c
        lay = Layout(order, (s0, s1), rows, ks)
        cost = gather_cost(lay.off_t, lay.off_s)
        if log:
            log(f"order {order}: strides ({s0}, {s1}) {c0:.3f} -> {cost:.3f}")
        if best is None or cost < best[0]:
            best = (cost, lay)
    return best[1]


@lru_cache(maxsize=None)
def _table():
    if not os.path.exists(_TABLE):
        return {}
    with open(_TABLE) as f:
        return json.load(f)


def layout(order):
    """The tuned layout of *order* from the committed table, else the plain
    one."""
    ent = _table().get(str(order))
    if ent is None:
        return identity_layout(order)
    return Layout(order, tuple(ent["strides"]), np.array(ent["rows"], np.int64),
                  np.array(ent["ks"], np.int64))


def main(orders=range(1, 11), iters=100000):
    import sys
    out = {}
    for p in orders:
        lay = tune(p, iters=iters, log=lambda s: print(s, file=sys.stderr, flush=True))
        base = identity_layout(p)
        out[str(p)] = {"strides": list(lay.strides), "rows": lay.rows.tolist(),
                       "ks": lay.ks.tolist(),
                       "cost": round(gather_cost(lay.off_t, lay.off_s), 4),
                       "plain_cost": round(gather_cost(base.off_t, base.off_s), 4)}
    os.makedirs(os.path.dirname(_TABLE), exist_ok=True)
    with open(_TABLE, "w") as f:
        json.dump(out, f, indent=None, separators=(",", ":"))
        f.write("\n")


if __name__ == "__main__":
    main()
