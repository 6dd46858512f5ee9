"""Expansion-to-expansion operators (reference ``sumpy/e2e.py``).

* :class:`E2EFromChildren` -- M2M (``e2e.py:866-978``), accumulates;
* :class:`E2EFromParent` -- L2L (``e2e.py:985-1083``), accumulates;
* :class:`E2EFromCSR` -- M2L without translation classes (``e2e.py:164-304``),
  overwrites;
* the translation-class M2L pipeline: :class:`M2LGenerateTranslationClassesDependentData`
  (``e2e.py:565-669``), :class:`M2LPreprocessMultipole` (``:676-760``),
  :class:`M2LUsingTranslationClassesDependentData` (``:310-558``) and
  :class:`M2LPostprocessLocal` (``:767-859``).
"""

from __future__ import annotations

import ctypes
import weakref

import numpy as np
import torch

from sumpy_amd import _capi, h2d, taylor2d
from sumpy_amd.expansion import m2l_device_tables
from sumpy_amd.kernel import kernel_id, remove_source_transforms, remove_target_transforms
from sumpy_amd.profiling import kernel_span
from sumpy_amd.tools import as_i32, expansion_k, helmholtz_k, planes, set_planes

TILE = 64  # target boxes per M2L tile by default (the HIP kernel's widest tile)
TILE_WIDTHS = (64, 32)  # the tile widths the contraction kernels are built for
GROUP = 16  # target boxes per octant-translation group (matches the HIP kernel)
MIN_SLOTS_PER_SPLIT = 12


class E2EBase:
    def __init__(self, src_expansion, tgt_expansion, name=None):
        strip = lambda e: e.with_kernel(  # noqa: E731
            remove_source_transforms(remove_target_transforms(e.kernel)))
        self.src_expansion = strip(src_expansion)
        self.tgt_expansion = strip(tgt_expansion)
        if self.src_expansion.dim != self.tgt_expansion.dim:
            raise ValueError("source and target expansions must have same dimensionality")
        self.name = name or self.default_name

    @property
    def dim(self):
        return self.tgt_expansion.dim

    def get_cache_key(self):
        return (type(self).__name__, self.src_expansion, self.tgt_expansion)

    def _tables(self, actx, src_rscale=None, tgt_rscale=None, k=0.0):
        """Device tables; for Helmholtz (*k* nonzero) those of the two
        expansion scales, whose PDE projections hold ``(k rscale)^2``."""
        if not k:
            return (self.src_expansion.device_tables(actx.device),
                    self.tgt_expansion.device_tables(actx.device))
        return (self.src_expansion.tables_for(actx.device, k, src_rscale),
                self.tgt_expansion.tables_for(actx.device, k, tgt_rscale))


class _OctantTranslation(E2EBase):
    """Shared machinery of M2M and L2L: when every child of the level sits at
    one of eight fixed offsets from its parent (an octree), the level's
    translation is eight fixed matrices applied to gathered coefficient
    vectors (``sumpyhip_octant_translate``, MFMA).  The matrices are the
    per-box kernel applied to unit vectors, so both paths compute the same map;
    other geometries use the per-box kernel."""

    def __init__(self, src_expansion, tgt_expansion, name=None,
                 use_octant_matrices=True):
        super().__init__(src_expansion, tgt_expansion, name)
        self.use_octant_matrices = use_octant_matrices
        self._mats = {}
        #: whether the last call took the octant-matrix (MFMA) path
        self.last_call_used_octant_matrices = False

    @staticmethod
    def _host_copy(t):
        """Host copy of the tree/traversal tensor *t*, held while the tensor
        (or the tensor it is a view of) lives."""
        per = _tensor_cache(_owner(t))
        key = ("host", t.data_ptr(), t.numel(), t._version, tuple(t.shape), str(t.dtype))
        v = per.get(key)
        if v is None:
            v = t.cpu().numpy()
            per[key] = v
        return v

    def _plan_cache(self, tb_t, others):
        """Plans of a level, held with the target-box tensor's owner (a
        traversal array; the tree arrays in *others* live as long as the
        traversal that refers to the tree)."""
        per = _tensor_cache(_owner(tb_t))
        key = (self.span_name,) + tuple(
            (t.data_ptr(), t.numel(), t._version) for t in (tb_t, *others))
        return per, key

    def _matrices(self, actx, offsets, present, src_rscale, tgt_rscale, k=0.0):
        key = (offsets.tobytes(), present.tobytes(), float(src_rscale), float(tgt_rscale),
               float(k))
        frag = self._mats.get(key)
        if frag is None:
            T = self._unit_matrices(actx, offsets, present, src_rscale, tgt_rscale, k)
            frag = _fragment_order(T)
            self._mats[key] = frag
        return frag

    def __call__(self, actx, **kwargs):
        if h2d.is_h2d(self.tgt_expansion):
            return (h2d.l2l if self.span_name == "l2l" else h2d.m2m)(self, actx, kwargs)
        if taylor2d.is_taylor2d(self.tgt_expansion):
            return (taylor2d.l2l if self.span_name == "l2l" else taylor2d.m2m)(
                self, actx, kwargs)
        src, tgt = kwargs["src_expansions"], kwargs["tgt_expansions"]
        k = expansion_k(self.tgt_expansion.kernel, kwargs, self.tgt_expansion)
        if not tgt.is_complex():
            return self._apply(actx, kwargs, src, tgt, k)
        # complex (Helmholtz) expansions: the translation is real-linear with
        # real matrices, applied to the real and the imaginary plane
        sr, si = planes(src)
        tr, ti = planes(tgt)
        self._apply(actx, kwargs, sr, tr, k)
        self._apply(actx, kwargs, si, ti, k)
        set_planes(tgt, tr, ti)
        return tgt

    def _unit_matrices(self, actx, offsets, present, src_rscale, tgt_rscale, k=0.0):
        """T[o] (n_out x n_in) from the per-box kernel on a synthetic level:
        box 0 unused, boxes 1..8n are the 'big' boxes (parents), 8n+1..16n the
        'small' ones (children), one pair per (octant, unit vector)."""
        st, tt = self._tables(actx, src_rscale, tgt_rscale, k)
        n_in, n_out = st.table.nstored, tt.table.nstored
        no = 8*n_in
        nb = 1 + 2*no
        par = 1 + np.arange(no, dtype=np.int32)
        chd = 1 + no + np.arange(no, dtype=np.int32)
        centers = np.zeros((3, nb))
        for o in range(8):
            centers[:, chd[o*n_in:(o + 1)*n_in]] = self._child_offset(offsets[o])[:, None]
        eye = torch.eye(n_in, dtype=torch.float64, device=actx.device).repeat(8, 1)
        out = torch.zeros((no, n_out), dtype=torch.float64, device=actx.device)
        self._unit_call(actx, st, tt, par, chd, nb, actx.from_numpy(centers), src_rscale,
                        tgt_rscale, eye, out)
        T = out.view(8, n_in, n_out).transpose(1, 2)
        return T*torch.from_numpy(present.astype(np.float64)).to(actx.device)[:, None, None]

    def _launch_octant(self, actx, plan, frag, src, sbase, tgt, tbase):
        st, tt = self._tables(actx)
        flops = 2.0*st.table.nstored*tt.table.nstored*plan["npairs"]
        partial = None
        if self.octant_inputs_per_target > 1:
            partial = torch.empty(plan["ngroups"]*8*GROUP*tt.table.nstored,
                                  dtype=torch.float64, device=tgt.device)
        with kernel_span(self.span_name, flops):
            _capi.call("sumpyhip_octant_translate", plan["ngroups"],
                       _capi.ptr(plan["targets"]), _capi.ptr(plan["inputs"]),
                       _capi.ptr(plan["mask"]), int(plan["items"].shape[0]),
                       _capi.ptr(plan["items"]), tt.table.nstored, st.table.nstored,
                       _capi.ptr(frag), _capi.ptr(src), int(sbase), _capi.ptr(tgt), int(tbase),
                       _capi.ptr(partial), actx.stream_handle)
        return tgt


def _fragment_order(T):
    """(8, n_out, n_in) -> flat MFMA A-fragment order
    ``[((o*nrb + rb)*ksteps + ks)*64 + l] = T[o, 16 rb + l%16, 4 ks + l/16]``."""
    no, n_out, n_in = T.shape
    nrb, ks = -(-n_out//16), -(-n_in//4)
    P = torch.zeros((no, nrb*16, ks*4), dtype=torch.float64, device=T.device)
    P[:, :n_out, :n_in] = T
    return P.view(no, nrb, 16, ks, 4).permute(0, 1, 3, 4, 2).contiguous().view(-1)


def _common_offsets(d, labels, rel_tol=1e-10):
    """Per label 0..7 the common offset of its pairs (columns of *d*), or None
    when the offsets of some label differ by more than *rel_tol*."""
    offsets = np.zeros((8, 3))
    present = np.zeros(8, bool)
    for o in range(8):
        sel = labels == o
        if not sel.any():
            continue
        do = d[:, sel]
        d0 = do[:, 0]
        if np.abs(do - d0[:, None]).max() > rel_tol*max(np.abs(d0).max(), 1e-300):
            return None
        offsets[o] = d0
        present[o] = True
    return offsets, present


def _octant_items(mask):
    """g * 8 + o for every octant o present in group g, in group order."""
    g, o = np.nonzero((np.asarray(mask)[:, None] >> np.arange(8)[None, :]) & 1)
    return (g*8 + o).astype(np.int32)


def _groups(actx, targets, inputs, labels):
    """Groups of GROUP targets sharing one label each (inputs[i] feeds target
    i through octant labels[i]); M2M passes all eight octants per target."""
    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(actx.device)
    return dev(targets), dev(inputs)


class E2EFromChildren(_OctantTranslation):
    """M2M: ``tgt[parent] += sum_children T(c_parent - c_child) src[child]``."""

    default_name = "e2e_from_children"
    span_name = "m2m"
    octant_inputs_per_target = 8  # a parent sums its children's octants

    @staticmethod
    def _child_offset(d):
        return -d  # the parent sits at the origin, d = c_parent - c_child

    def _unit_call(self, actx, st, tt, par, chd, nb, centers, src_rscale, tgt_rscale,
                   src, out):
        n_in = st.table.nstored
        child = np.zeros((8, nb), np.int32)
        for o in range(8):
            child[o, par[o*n_in:(o + 1)*n_in]] = chd[o*n_in:(o + 1)*n_in]
        child = actx.from_numpy(child)
        par_d = actx.from_numpy(par)
        _capi.call("sumpyhip_m2m", kernel_id(self.tgt_expansion.kernel), st.desc, tt.desc,
                   len(par), _capi.ptr(par_d), _capi.ptr(child), _capi.ptr(centers), nb,
                   float(src_rscale), float(tgt_rscale), _capi.ptr(src), int(chd[0]),
                   _capi.ptr(out), int(par[0]), actx.stream_handle)

    def _plan(self, tb_t, child_t, centers_t):
        plans, key = self._plan_cache(tb_t, (child_t, centers_t))
        if key in plans:
            return plans[key]
        tb = self._host_copy(tb_t).astype(np.int64)
        child = self._host_copy(child_t)
        centers = self._host_copy(centers_t)
        ch = child[:, tb].astype(np.int64)  # (8, n); 0 = no child
        has = ch != 0
        d = np.concatenate([centers[:, tb[has[o]]] - centers[:, ch[o, has[o]]]
                            for o in range(8)], axis=1)
        labels = np.concatenate([np.full(int(has[o].sum()), o) for o in range(8)])
        common = _common_offsets(d, labels) if len(labels) else None
        plan = None
        if common is not None:
            n = len(tb)
            ng = -(-n//GROUP)
            targets = np.full(ng*GROUP, -1, np.int64)
            targets[:n] = tb
            inputs = np.full((8, ng*GROUP), -1, np.int64)
            inputs[:, :n] = np.where(has, ch, -1)
            inputs = inputs.reshape(8, ng, GROUP).transpose(1, 0, 2)
            mask = ((inputs >= 0).any(axis=2)*(1 << np.arange(8))[None, :]).sum(axis=1)
            plan = dict(ngroups=ng, targets=targets, inputs=inputs, mask=mask,
                        offsets=common[0], present=common[1], npairs=int(has.sum()))
        plans[key] = plan
        return plan

    def _apply(self, actx, kwargs, src, tgt, k):
        st, tt = self._tables(actx, kwargs["src_rscale"], kwargs["tgt_rscale"], k)
        centers = kwargs["centers"].contiguous()
        tb = as_i32(kwargs["target_boxes"])
        child = as_i32(kwargs["box_child_ids"])
        self.last_call_used_octant_matrices = False
        if self.use_octant_matrices and tb.numel():
            plan = self._plan(tb, child, centers)
            if plan is not None:
                self.last_call_used_octant_matrices = True
                if "dev" not in plan:
                    plan["dev"] = dict(zip(("targets", "inputs"), _groups(
                        actx, plan["targets"], plan["inputs"], None)),
                        mask=actx.from_numpy(plan["mask"].astype(np.int32)),
                        items=actx.from_numpy(_octant_items(plan["mask"])))
                frag = self._matrices(actx, plan["offsets"], plan["present"],
                                      kwargs["src_rscale"], kwargs["tgt_rscale"], k)
                return self._launch_octant(actx, dict(plan["dev"], ngroups=plan["ngroups"],
                                                      npairs=plan["npairs"]),
                                           frag, src, kwargs["src_base_ibox"], tgt,
                                           kwargs["tgt_base_ibox"])
        with kernel_span(self.span_name, 0.0):
            _capi.call("sumpyhip_m2m", kernel_id(self.tgt_expansion.kernel), st.desc,
                       tt.desc, int(tb.shape[0]), _capi.ptr(tb), _capi.ptr(child),
                       _capi.ptr(centers), int(centers.shape[1]), float(kwargs["src_rscale"]),
                       float(kwargs["tgt_rscale"]), _capi.ptr(src),
                       int(kwargs["src_base_ibox"]), _capi.ptr(tgt),
                       int(kwargs["tgt_base_ibox"]), actx.stream_handle)
        return tgt


class E2EFromParent(_OctantTranslation):
    """L2L: ``tgt[child] += T(c_child - c_parent) src[parent]``."""

    default_name = "e2e_from_parent"
    span_name = "l2l"
    octant_inputs_per_target = 1  # a child has its parent in one octant

    @staticmethod
    def _child_offset(d):
        return d  # the parent sits at the origin, d = c_child - c_parent

    def _unit_call(self, actx, st, tt, par, chd, nb, centers, src_rscale, tgt_rscale,
                   src, out):
        parent = np.zeros(nb, np.int32)
        parent[chd] = par
        parent = actx.from_numpy(parent)
        chd_d = actx.from_numpy(chd)
        _capi.call("sumpyhip_l2l", kernel_id(self.tgt_expansion.kernel), st.desc, tt.desc,
                   len(chd), _capi.ptr(chd_d), _capi.ptr(parent), _capi.ptr(centers), nb,
                   float(src_rscale), float(tgt_rscale), _capi.ptr(src), int(par[0]),
                   _capi.ptr(out), int(chd[0]), actx.stream_handle)

    def _plan(self, tb_t, parent_t, centers_t):
        plans, key = self._plan_cache(tb_t, (parent_t, centers_t))
        if key in plans:
            return plans[key]
        tb = self._host_copy(tb_t).astype(np.int64)
        parent = self._host_copy(parent_t).astype(np.int64)
        centers = self._host_copy(centers_t)
        pb = parent[tb]
        d = centers[:, tb] - centers[:, pb]
        labels = ((d[0] > 0)*4 + (d[1] > 0)*2 + (d[2] > 0)).astype(np.int64)
        common = _common_offsets(d, labels) if len(tb) and (pb != tb).all() else None
        plan = None
        if common is not None:
            order = np.argsort(labels, kind="stable")
            cnt = np.bincount(labels, minlength=8)
            first = np.concatenate([[0], np.cumsum(cnt)[:-1]])
            tg, ip, mk = [], [], []
            for o in range(8):
                sel = order[first[o]:first[o] + cnt[o]]
                for c0 in range(0, len(sel), GROUP):
                    part = sel[c0:c0 + GROUP]
                    t = np.full(GROUP, -1, np.int64)
                    t[:len(part)] = tb[part]
                    i = np.full((8, GROUP), -1, np.int64)
                    i[o, :len(part)] = pb[part]
                    tg.append(t)
                    ip.append(i)
                    mk.append(1 << o)
            plan = dict(ngroups=len(tg), targets=np.concatenate(tg),
                        inputs=np.stack(ip), mask=np.array(mk), offsets=common[0],
                        present=common[1], npairs=len(tb))
        plans[key] = plan
        return plan

    def _apply(self, actx, kwargs, src, tgt, k):
        st, tt = self._tables(actx, kwargs["src_rscale"], kwargs["tgt_rscale"], k)
        centers = kwargs["centers"].contiguous()
        tb = as_i32(kwargs["target_boxes"])
        parents = as_i32(kwargs["box_parent_ids"])
        self.last_call_used_octant_matrices = False
        if self.use_octant_matrices and tb.numel():
            plan = self._plan(tb, parents, centers)
            if plan is not None:
                self.last_call_used_octant_matrices = True
                if "dev" not in plan:
                    plan["dev"] = dict(zip(("targets", "inputs"), _groups(
                        actx, plan["targets"], plan["inputs"], None)),
                        mask=actx.from_numpy(plan["mask"].astype(np.int32)),
                        items=actx.from_numpy(_octant_items(plan["mask"])))
                frag = self._matrices(actx, plan["offsets"], plan["present"],
                                      kwargs["src_rscale"], kwargs["tgt_rscale"], k)
                return self._launch_octant(actx, dict(plan["dev"], ngroups=plan["ngroups"],
                                                      npairs=plan["npairs"]),
                                           frag, src, kwargs["src_base_ibox"], tgt,
                                           kwargs["tgt_base_ibox"])
        with kernel_span(self.span_name, 0.0):
            _capi.call("sumpyhip_l2l", kernel_id(self.tgt_expansion.kernel), st.desc,
                       tt.desc, int(tb.shape[0]), _capi.ptr(tb), _capi.ptr(parents),
                       _capi.ptr(centers), int(centers.shape[1]), float(kwargs["src_rscale"]),
                       float(kwargs["tgt_rscale"]), _capi.ptr(src),
                       int(kwargs["src_base_ibox"]), _capi.ptr(tgt),
                       int(kwargs["tgt_base_ibox"]), actx.stream_handle)
        return tgt


def _m2l_tabs(src_expansion, tgt_expansion, device):
    return m2l_device_tables(src_expansion.order, tgt_expansion.order,
                             src_expansion.compressed, str(device))


_BF16_TABS = {}
_PER_TENSOR = {}  # id(tensor) -> (weak ref, {key: data derived from the tensor})


def _owner(t):
    """The tensor whose storage *t* views (``t`` itself if it is no view)."""
    return t._base if t._base is not None else t


def _tensor_cache(t):
    """A dict of data derived from the tensor object *t* (M2L tile schedules
    of a list tensor, packed class data of a class-data tensor), dropped when
    *t* is freed, so a later tensor at the same address or id never sees it."""
    k = id(t)
    ent = _PER_TENSOR.get(k)
    if ent is None or ent[0]() is not t:
        ent = (weakref.ref(t), {})
        _PER_TENSOR[k] = ent
        weakref.finalize(t, _PER_TENSOR.pop, k, None)
    return ent[1]


def _bf16_tabs(src_expansion, tgt_expansion, device):
    """Device tables of the split-BF16 contraction's LDS layout: the tuned
    layout of :mod:`sumpy_amd.m2l_layout` for equal compressed orders, else
    the plain circulant layout (stored row and k order)."""
    key = (src_expansion.order, tgt_expansion.order, src_expansion.compressed, str(device))
    tabs = _BF16_TABS.get(key)
    if tabs is not None:
        return tabs
    from sumpy_amd import m2l_layout
    m = _m2l_tabs(src_expansion, tgt_expansion, device)
    nl, nm = m["nl"], m["nm"]
    npad, kpad = -(-nl//16)*16, -(-nm//32)*32
    shape = tuple(int(x) for x in m["shape"])
    if src_expansion.order == tgt_expansion.order and src_expansion.compressed:
        lay = m2l_layout.layout(src_expansion.order)
        off_t, off_s, ks = lay.off_t, lay.off_s, lay.ks
        row_pos, strides, ncirc = lay.row_pos, lay.strides, lay.ncirc
    else:
        off_t = np.zeros(npad, np.int32)
        off_t[:nl] = m["off_t"].cpu().numpy()
        off_s = np.zeros(kpad, np.int32)
        off_s[:nm] = m["off_s"].cpu().numpy()
        ks = np.where(np.arange(kpad) < nm, np.arange(kpad), -1)
        row_pos = np.arange(nl)
        strides, ncirc = (shape[1]*shape[2], shape[2]), m["ncirc"]
    if len(off_t) != npad or len(off_s) != kpad:
        raise ValueError("split-BF16 layout does not match the expansion sizes")

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(device)

    tabs = dict(off_t=dev(off_t), off_s=dev(off_s), k_src=dev(ks), row_pos=dev(row_pos),
                strides=tuple(int(x) for x in strides), ncirc=int(ncirc), npad=npad,
                kpad=kpad, shape=_capi.np_i32(shape))
    _BF16_TABS[key] = tabs
    return tabs


class E2EFromCSR(E2EBase):
    """M2L without translation classes: kernel derivatives per (target,
    source) pair, overwrites (reference ``e2e.py:164-304``)."""

    default_name = "e2e_from_csr"

    def __call__(self, actx, **kwargs):
        from sumpy_amd.expansion.local import LocalExpansionBase
        from sumpy_amd.expansion.multipole import MultipoleExpansionBase
        if not (isinstance(self.src_expansion, MultipoleExpansionBase)
                and isinstance(self.tgt_expansion, LocalExpansionBase)):
            raise NotImplementedError("E2EFromCSR implements multipole-to-local only")
        if h2d.is_h2d(self.tgt_expansion):
            return h2d.m2l(self, actx, kwargs, with_classes=False)
        if taylor2d.is_taylor2d(self.tgt_expansion):
            return taylor2d.m2l(self, actx, kwargs, with_classes=False)
        m = _m2l_tabs(self.src_expansion, self.tgt_expansion, actx.device)
        centers = kwargs["centers"].contiguous()
        tb = as_i32(kwargs["target_boxes"])
        starts = as_i32(kwargs["src_box_starts"])
        lists = as_i32(kwargs["src_box_lists"])
        src, tgt = kwargs["src_expansions"], kwargs["tgt_expansions"]
        shape = _capi.np_i32(m["shape"])
        if self.tgt_expansion.is_complex_valued:
            _capi.call("sumpyhip_m2l_csr_helmholtz", self.src_expansion.order,
                       self.tgt_expansion.order, shape.ctypes.data, m["nl"], m["nm"],
                       _capi.ptr(m["off_t"]), _capi.ptr(m["off_s"]), _capi.ptr(m["tgt_degree"]),
                       int(tb.shape[0]), _capi.ptr(tb), _capi.ptr(starts), _capi.ptr(lists),
                       _capi.ptr(centers), int(centers.shape[1]), float(kwargs["src_rscale"]),
                       float(kwargs["tgt_rscale"]), *helmholtz_k(self.tgt_expansion.kernel, kwargs),
                       _capi.ptr(src), int(kwargs["src_base_ibox"]), _capi.ptr(tgt),
                       int(kwargs["tgt_base_ibox"]), actx.stream_handle)
            return tgt
        _capi.call("sumpyhip_m2l_csr", kernel_id(self.tgt_expansion.kernel),
                   self.src_expansion.order, self.tgt_expansion.order,
                   shape.ctypes.data, m["nl"], m["nm"],
                   _capi.ptr(m["off_t"]), _capi.ptr(m["off_s"]), _capi.ptr(m["tgt_degree"]),
                   int(tb.shape[0]), _capi.ptr(tb), _capi.ptr(starts), _capi.ptr(lists),
                   _capi.ptr(centers), int(centers.shape[1]), float(kwargs["src_rscale"]),
                   float(kwargs["tgt_rscale"]), _capi.ptr(src), int(kwargs["src_base_ibox"]),
                   _capi.ptr(tgt), int(kwargs["tgt_base_ibox"]), actx.stream_handle)
        return tgt


class M2LGenerateTranslationClassesDependentData(E2EBase):
    """Per translation class ``c``: ``data[c, off(m)] = rho^|m| d^m G(v_c)`` in
    the circulant layout of ``m2l.py:403-537``."""

    default_name = "m2l_generate_translation_classes_dependent_data"

    def __call__(self, actx, **kwargs):
        if h2d.is_h2d(self.tgt_expansion):
            return h2d.m2l_data(self, actx, kwargs)
        if taylor2d.is_taylor2d(self.tgt_expansion):
            return taylor2d.m2l_data(self, actx, kwargs)
        vecs = kwargs["m2l_translation_vectors"].contiguous()
        out = kwargs["m2l_translation_classes_dependent_data"]
        m = _m2l_tabs(self.src_expansion, self.tgt_expansion, actx.device)
        ncls = int(kwargs["ntranslation_classes"])
        shape = _capi.np_i32(m["shape"])
        if self.tgt_expansion.is_complex_valued:
            # real and imaginary plane per class: classes 2c, 2c + 1 of the
            # real embedding the contraction uses
            if out.shape != (ncls, 2*m["ncirc"]):
                raise ValueError(f"class data must have shape ({ncls}, {2*m['ncirc']})")
            _capi.call("sumpyhip_m2l_class_data_helmholtz", self.src_expansion.order,
                       self.tgt_expansion.order, shape.ctypes.data, ncls, _capi.ptr(vecs),
                       int(vecs.shape[1]), int(kwargs["translation_classes_level_start"]),
                       float(kwargs["src_rscale"]),
                       *helmholtz_k(self.tgt_expansion.kernel, kwargs), _capi.ptr(out),
                       actx.stream_handle)
            return out
        if out.shape != (ncls, m["ncirc"]):
            raise ValueError(f"class data must have shape ({ncls}, {m['ncirc']})")
        _capi.call("sumpyhip_m2l_class_data", kernel_id(self.src_expansion.kernel),
                   self.src_expansion.order, self.tgt_expansion.order, shape.ctypes.data,
                   ncls, _capi.ptr(vecs), int(vecs.shape[1]),
                   int(kwargs["translation_classes_level_start"]),
                   float(kwargs["src_rscale"]), _capi.ptr(out), actx.stream_handle)
        return out


class M2LPreprocessMultipole(E2EBase):
    """Pack the stored multipoles of a level into the M2L compute precision
    (zero padded to the MFMA K granule)."""

    default_name = "m2l_preprocess_multipole"

    def __call__(self, actx, **kwargs):
        src = kwargs["src_expansions"]
        out = kwargs["preprocessed_src_expansions"]
        tr = self.tgt_expansion.m2l_translation
        nb = int(src.shape[0])
        if out.shape[0] != nb:
            raise ValueError("preprocessed array has the wrong number of boxes")
        if src.is_complex():
            # rows Mr, Mi, -Mi per box (the real embedding of the complex
            # contraction); the packing kernel then sees 3 nb real boxes
            re, im = planes(src)
            src = torch.stack((re, im, -im), dim=1).reshape(3*nb, -1)
            out = out.view(3*nb, -1)
            nb = 3*nb
        k_src = None
        if tr.precision == "bf16x3":
            k_src = _bf16_tabs(self.src_expansion, self.tgt_expansion, actx.device)["k_src"]
            if int(k_src.shape[0]) != int(out.shape[1]):
                raise ValueError("preprocessed rows do not match the split-BF16 layout")
        _capi.call("sumpyhip_m2l_preprocess", tr.precision_id, nb, int(src.shape[1]),
                   _capi.ptr(src), int(out.shape[1]), _capi.ptr(k_src), _capi.ptr(out),
                   actx.stream_handle)
        return kwargs["preprocessed_src_expansions"]


class M2LUsingTranslationClassesDependentData(E2EBase):
    """Class-batched dense M2L on MFMA; overwrites the work rows of the
    target boxes (reference ``e2e.py:310-558``)."""

    default_name = "m2l_using_translation_classes_dependent_data"

    def packed_class_data(self, actx, data):
        """The split-BF16 form's class data: one 32-bit word bf16(x) |
        bf16(x - bf16(x)) << 16 per value, at the strides of the LDS layout.
        Made once per class-data tensor object (the wrangler's memoized class
        data, written once) and held while that tensor lives."""
        key = ("bf16x3_packed", self.src_expansion.order, self.tgt_expansion.order,
               self.src_expansion.compressed, data.data_ptr(), tuple(data.shape))
        per = _tensor_cache(_owner(data))
        out = per.get(key)
        if out is None:
            tabs = _bf16_tabs(self.src_expansion, self.tgt_expansion, actx.device)
            # complex kernels: the real and imaginary planes of a class are
            # the classes 2c and 2c + 1 of the real embedding
            ncirc = _m2l_tabs(self.src_expansion, self.tgt_expansion, actx.device)["ncirc"]
            ncls = int(data.numel())//int(ncirc)
            out = torch.empty((ncls, tabs["ncirc"]), dtype=torch.int32, device=data.device)
            _capi.call("sumpyhip_m2l_pack_bf16_class_data", ncls, tabs["shape"].ctypes.data,
                       *tabs["strides"], tabs["ncirc"], _capi.ptr(data), _capi.ptr(out),
                       actx.stream_handle)
            per[key] = out
        return out

    def schedule(self, actx, target_boxes, src_box_starts, src_box_lists, class_lists,
                 class_level_start, nclasses, centers, tile_width=TILE):
        """Tile schedule of one level, built on the host and cached;
        *tile_width* columns per tile of the contraction (64 or 32; a complex
        target takes two)."""
        if tile_width not in TILE_WIDTHS:
            raise ValueError(f"tile width {tile_width}: one of {TILE_WIDTHS}")
        tile = tile_width//2 if self.tgt_expansion.is_complex_valued else tile_width
        key = tuple((t.data_ptr(), t.numel(), t._version) for t in
                    (target_boxes, src_box_starts, src_box_lists, class_lists, centers)) \
            + (int(class_level_start), int(nclasses), tile, str(actx.device))
        # shared by every operator instance (the schedule depends on the lists
        # only), held while the traversal's list tensor lives
        per_lists = _tensor_cache(_owner(src_box_lists))
        sched = per_lists.get(key)
        if sched is None:
            sched = build_m2l_schedule(
                actx, target_boxes.cpu().numpy(), src_box_starts.cpu().numpy(),
                src_box_lists.cpu().numpy(), class_lists.cpu().numpy(),
                int(class_level_start), int(nclasses), centers.cpu().numpy(), tile=tile)
            per_lists[key] = sched
        return sched

    def level_entry(self, actx, tile_width=TILE, **kwargs):
        """The launch description of one level (for :func:`m2l_apply_levels`),
        or None when the level has no target tiles; *tile_width*: columns
        per tile of the contraction (64 or 32)."""
        data = kwargs["m2l_translation_classes_dependent_data"]
        src, work = kwargs["src_expansions"], kwargs["tgt_expansions"]
        sched = self.schedule(
            actx, kwargs["target_boxes"], kwargs["src_box_starts"], kwargs["src_box_lists"],
            kwargs["m2l_translation_classes_lists"],
            kwargs["translation_classes_level_start"], data.shape[0], kwargs["centers"],
            tile_width=tile_width)
        if sched["ntiles"] == 0:
            return None
        sbase, tbase = int(kwargs["src_base_ibox"]), int(kwargs["tgt_base_ibox"])
        if self.tgt_expansion.is_complex_valued:
            sched = self._complex_schedule(actx, sched, sbase, tbase, tile_width)
            nb_s, nb_t = int(src.shape[0]), int(work.shape[0])
            return dict(op=self, sched=sched, data=data.contiguous().view(2*data.shape[0], -1),
                        src=src.view(3*nb_s, -1), sbase=0, work=work.view(2*nb_t, -1),
                        tbase=0, tw=tile_width)
        return dict(op=self, sched=sched, data=data.contiguous(), src=src,
                    sbase=sbase, work=work, tbase=tbase, tw=tile_width)

    def _complex_schedule(self, actx, sched, sbase, tbase, tile_width=TILE):
        """The schedule of the real embedding.  A tile of tile_width/2 complex
        targets becomes one tile of tile_width columns: column 2j is the row Lr
        (2 (t_j - tbase)), column 2j + 1 the row Li.  Class slot c becomes the
        slots of Ar (class 2c) and Ai (2c + 1), fed by the source rows
        3 (s - sbase) + (0: Mr, 1: Mi, 2: -Mi):
        Lr = Ar Mr + Ai (-Mi), Li = Ar Mi + Ai Mr."""
        # held in the real schedule itself, so it lives and dies with it
        key = ("complex", sbase, tbase)
        out = sched.get(key)
        if out is not None:
            return out
        h = sched["host"]
        half = tile_width//2
        tt = h["tile_targets"].astype(np.int64).reshape(-1, half)
        ss = h["slot_sources"].astype(np.int64).reshape(-1, half)
        sc = h["slot_class"].astype(np.int64)

        def rows(ids, base, mult, offs):
            r = np.where(ids[..., None] >= 0,
                         mult*(ids[..., None] - base) + np.array(offs), -1)
            return r.reshape(ids.shape[0], -1)

        vt = rows(tt, tbase, 2, (0, 1))                      # (ntiles, tile_width)
        vs = np.stack([rows(ss, sbase, 3, (0, 1)),           # Ar <- (Mr, Mi)
                       rows(ss, sbase, 3, (2, 0))], axis=1)  # Ai <- (-Mi, Mr)
        vc = np.stack([2*sc, 2*sc + 1], axis=1)

        def dev(a):
            return torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(actx.device)

        ntiles, nslots = sched["ntiles"], 2*len(sc)
        out = dict(ntiles=ntiles, tile_targets=dev(vt.reshape(-1)),
                   tile_slot_starts=dev(2*h["tile_slot_starts"].astype(np.int64)),
                   slot_class=dev(vc.reshape(-1)), slot_sources=dev(vs.reshape(-1)),
                   nslots=nslots, npairs=4*sched["npairs"],
                   nsplit=choose_nsplit(ntiles, nslots, actx.device))
        sched[key] = out
        return out

    def __call__(self, actx, **kwargs):
        if h2d.is_h2d(self.tgt_expansion):
            return h2d.m2l(self, actx, kwargs, with_classes=True)
        if taylor2d.is_taylor2d(self.tgt_expansion):
            return taylor2d.m2l(self, actx, kwargs, with_classes=True)
        entry = self.level_entry(actx, **kwargs)
        if entry is not None:
            m2l_apply_levels(actx, [entry])
        return kwargs["tgt_expansions"]


def m2l_apply_levels(actx, entries, policy=None, noalloc=False):
    """One launch of the MFMA contraction over several levels (same
    translation operator orders), largest level first.  *policy* = (items
    per CU, minimum slots per split) overrides the split rule of the
    schedules (:mod:`sumpy_amd.autotune`).  *noalloc*: the split partial sums
    go to the library's fixed device scratch (``sumpyhip_m2l_apply_noalloc``)
    instead of buffers from torch's allocator."""
    entries = [e for e in entries if e is not None]
    if not entries:
        return
    op = entries[0]["op"]
    tr = op.tgt_expansion.m2l_translation
    m = _m2l_tabs(op.src_expansion, op.tgt_expansion, actx.device)
    for e in entries[1:]:
        if (e["op"].src_expansion.order, e["op"].tgt_expansion.order,
                e["op"].tgt_expansion.m2l_translation.precision) != \
                (op.src_expansion.order, op.tgt_expansion.order, tr.precision):
            raise ValueError("m2l_apply_levels: levels must share orders and precision")
    entries = sorted(entries, key=lambda e: -e["sched"]["npairs"])
    npad, kpad = int(entries[0]["work"].shape[1]), int(entries[0]["src"].shape[1])
    tw = entries[0].get("tw", TILE)
    if any(e.get("tw", TILE) != tw for e in entries):
        raise ValueError("m2l_apply_levels: levels of one launch share the tile width")
    nsplits = [e["sched"]["nsplit"] if policy is None else
               choose_nsplit(e["sched"]["ntiles"], e["sched"]["nslots"], actx.device, *policy)
               for e in entries]
    levels, keep = [], []
    npairs = 0
    for e, nsplit in zip(entries, nsplits):
        sched, work = dict(e["sched"], nsplit=nsplit), e["work"]
        if int(work.shape[1]) != npad or int(e["src"].shape[1]) != kpad:
            raise ValueError("m2l_apply_levels: inconsistent row padding")
        partial = None
        if sched["nsplit"] > 1 and not noalloc:
            partial = torch.empty((sched["nsplit"] - 1)*sched["ntiles"]*tw*npad,
                                  dtype=work.dtype, device=work.device)
            keep.append(partial)
        npairs += sched["npairs"]
        data = e["data"] if tr.precision != "bf16x3" else \
            e["op"].packed_class_data(actx, e["data"])
        keep.append(data)
        levels.append(_capi.M2LLevel(
            sched["ntiles"], sched["nsplit"], _capi.ptr(sched["tile_targets"]),
            _capi.ptr(sched["tile_slot_starts"]), _capi.ptr(sched["slot_class"]),
            _capi.ptr(sched["slot_sources"]), _capi.ptr(data), _capi.ptr(e["src"]),
            e["sbase"], _capi.ptr(work), e["tbase"], _capi.ptr(partial), tw))
    arr = (_capi.M2LLevel*len(levels))(*levels)
    flops = 2.0*m["nl"]*m["nm"]*npairs
    if tr.precision == "bf16x3":
        tabs = _bf16_tabs(op.src_expansion, op.tgt_expansion, actx.device)
        if (tabs["npad"], tabs["kpad"]) != (npad, kpad):
            raise ValueError("m2l_apply_levels: padding differs from the split-BF16 layout")
        geom = (npad, kpad, npad, kpad, _capi.ptr(tabs["off_t"]), _capi.ptr(tabs["off_s"]),
                tabs["ncirc"])
    else:
        geom = (m["nl"], m["nm"], npad, kpad, _capi.ptr(m["off_t"]), _capi.ptr(m["off_s"]),
                m["ncirc"])
    entry_point = "sumpyhip_m2l_apply_noalloc" if noalloc else "sumpyhip_m2l_apply"
    with kernel_span(f"m2l_apply_{tr.precision}", flops):
        _capi.call(entry_point, tr.precision_id, len(levels), ctypes.addressof(arr),
                   *geom, actx.stream_handle)


def build_m2l_schedule(actx, target_boxes, starts, lists, class_lists, class_start,
                       nclasses, centers, tile=TILE):
    """Group a level's target boxes into tiles of *tile* targets sharing most
    of their translation classes, and list for each tile the classes present
    with one source box per tile column (-1 if none).

    Targets are taken octant by octant (their octant within the parent fixes
    which classes they can have).  Within an octant they stay in box order,
    unless grouping them by their sets of classes (:func:`_class_set_order`)
    cuts the class slots of the level by 15 % or more (sparse levels, e.g.
    points on a surface)."""
    nt = len(target_boxes)
    e0, e1 = int(starts[0]), int(starts[-1])
    counts = np.diff(starts).astype(np.int64)
    itgt = np.repeat(np.arange(nt), counts)
    e_src = lists[e0:e1].astype(np.int64)
    e_cls = class_lists[e0:e1].astype(np.int64) - class_start
    if nt == 0:
        return dict(ntiles=0)
    # octant parity from centres: the smallest separation of a list-2 pair
    # along an axis is one box size h
    tb = target_boxes.astype(np.int64)
    ct = centers[:, tb]
    if len(e_src):
        diff = np.abs(centers[:, tb[itgt]] - centers[:, e_src])
        h = diff[diff > 1e-12*max(1.0, float(np.abs(centers).max()))].min()
    else:
        h = 1.0
    rel = np.rint((ct - ct[:, :1])/h).astype(np.int64)
    octant = ((rel[0] & 1) << 2) | ((rel[1] & 1) << 1) | (rel[2] & 1)
    tab = np.full((nt + 1, max(nclasses, 1)), -1, np.int32)
    tab[itgt, e_cls] = e_src
    member = tab >= 0  # row nt is all False
    cnt = np.bincount(octant, minlength=8)
    first = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    nchunk = (cnt + tile - 1)//tile
    # tiles: chunks of targets of one octant; chunk j of every octant before
    # chunk j + 1, so that a contiguous range of tiles covers one region of
    # space (the apply kernel hands each XCD such a range)
    chunks = sorted((j, o) for o in range(8) for j in range(nchunk[o]))
    ntiles = len(chunks)

    def tiles_of(order):
        padded = np.full(ntiles*tile, -1, np.int64)
        for it, (j, o) in enumerate(chunks):
            sel = order[first[o] + j*tile:first[o] + min((j + 1)*tile, cnt[o])]
            padded[it*tile:it*tile + len(sel)] = sel
        rows = np.where(padded >= 0, padded, nt)
        return padded, rows, int(member[rows].reshape(ntiles, tile, -1).any(axis=1).sum())

    padded, rows, nslots = tiles_of(np.argsort(octant, kind="stable"))
    if len(e_src) < 0.5*nslots*tile:
        cand = tiles_of(_class_set_order(member[:nt], octant, tile))
        if cand[2] <= nslots/1.15:
            padded, rows, nslots = cand
    tabT = tab[rows].reshape(ntiles, tile, -1)
    used = (tabT >= 0).any(axis=1)
    ti, ci = np.nonzero(used)
    slot_sources = tabT[ti, :, ci].astype(np.int32)
    tile_slot_starts = np.zeros(ntiles + 1, np.int32)
    np.cumsum(used.sum(axis=1), out=tile_slot_starts[1:])
    tile_targets = np.where(padded >= 0, tb[np.maximum(padded, 0)], -1).astype(np.int32)

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(actx.device)

    return dict(ntiles=ntiles, tile_targets=dev(tile_targets),
                tile_slot_starts=dev(tile_slot_starts), slot_class=dev(ci),
                slot_sources=dev(slot_sources.reshape(-1)), nslots=len(ci),
                npairs=int(len(e_src)),
                nsplit=choose_nsplit(ntiles, len(ci), actx.device),
                host=dict(tile_targets=tile_targets, tile_slot_starts=tile_slot_starts,
                          slot_class=ci.astype(np.int32),
                          slot_sources=slot_sources.reshape(-1)))


def _class_set_order(member, octant, tile, window=1024):
    """Targets octant by octant, grouped into tiles of *tile* that share
    classes: sorted by class set, then each tile grown greedily from its first
    free target by the target (among the next *window* free ones) that adds
    the fewest classes to the tile's set."""
    packed = np.packbits(member, axis=1)
    base = np.lexsort([packed[:, i] for i in range(packed.shape[1] - 1, -1, -1)] + [octant])
    out = []
    for o in range(8):
        seq = base[octant[base] == o]
        M = member[seq]
        Mi = M.astype(np.int16)
        free = np.ones(len(seq), bool)
        ptr = 0
        while True:
            while ptr < len(seq) and not free[ptr]:
                ptr += 1
            if ptr >= len(seq):
                break
            grp = [ptr]
            free[ptr] = False
            U = M[ptr].copy()
            while len(grp) < tile:
                cand = np.nonzero(free[ptr:ptr + window])[0] + ptr
                if len(cand) == 0:
                    break
                j = cand[np.argmin(Mi[cand] @ ~U)]
                free[j] = False
                grp.append(j)
                U |= M[j]
            out.extend(seq[grp])
    return np.array(out, np.int64)


def choose_nsplit(ntiles, nslots, device, per_cu=4, min_slots=MIN_SLOTS_PER_SPLIT):
    """Number of blocks per tile: enough (tile, split) items for about
    *per_cu* blocks per CU, with at least *min_slots* class slots per split
    (each split costs one extra write and read of its partial sums).  The
    defaults are the fixed rule; :mod:`sumpy_amd.autotune` may pick others."""
    if ntiles == 0:
        return 1
    ncu = torch.cuda.get_device_properties(device).multi_processor_count \
        if torch.device(device).type == "cuda" else 256
    want = -(-per_cu*ncu//ntiles)
    most = max(1, nslots//(ntiles*min_slots))
    return int(max(1, min(want, most)))


class M2LPostprocessLocal(E2EBase):
    """Scale the translated coefficients into FP64 locals:
    ``L[b, i] = work[b, i] * (tgt_rscale/src_rscale)^|mi_i|``."""

    default_name = "m2l_postprocess_local"

    def __call__(self, actx, **kwargs):
        tr = self.tgt_expansion.m2l_translation
        m = _m2l_tabs(self.src_expansion, self.tgt_expansion, actx.device)
        tgt = kwargs["tgt_expansions"]
        work = kwargs["tgt_expansions_before_postprocessing"]
        ratio = float(kwargs["tgt_rscale"])/float(kwargs["src_rscale"])
        row_pos = None
        if tr.precision == "bf16x3":
            row_pos = _bf16_tabs(self.src_expansion, self.tgt_expansion, actx.device)["row_pos"]
        out = tgt
        nb = int(tgt.shape[0])
        if tgt.is_complex():
            # work rows Lr, Li per box (the real embedding)
            work = work.reshape(2*nb, -1)
            out = torch.empty((2*nb, int(tgt.shape[1])), dtype=torch.float64,
                              device=tgt.device)
            nb = 2*nb
        _capi.call("sumpyhip_m2l_postprocess", tr.precision_id, nb,
                   int(out.shape[1]), int(work.shape[1]), _capi.ptr(work),
                   _capi.ptr(m["tgt_degree"]), _capi.ptr(row_pos), ratio, _capi.ptr(out),
                   actx.stream_handle)
        if out is not tgt:
            v = out.view(nb//2, 2, -1)
            set_planes(tgt, v[:, 0], v[:, 1])
        return tgt


__all__ = [
    "E2EFromCSR",
    "E2EFromChildren",
    "E2EFromParent",
    "M2LGenerateTranslationClassesDependentData",
    "M2LPostprocessLocal",
    "M2LPreprocessMultipole",
    "M2LUsingTranslationClassesDependentData",
]
