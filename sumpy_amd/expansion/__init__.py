"""Expansion classes with the reference's names (``sumpy/expansion``).

Only the data an operator needs is kept here: identifier tables, storage
indices and the device copies of the tables the HIP kernels consume.  The
mathematics lives in the kernels (``sumpy_amd/csrc``).
"""

from __future__ import annotations

from dataclasses import dataclass, field, replace
from functools import lru_cache

import numpy as np
import torch

from sumpy_amd.expansion.terms import (
    TermsTable,
    circulant_offsets,
    get_terms_table,
    m2l_circulant_shape,
    projection_csr,
)
from sumpy_amd.kernel import HelmholtzKernel, LaplaceKernel, ScalarKernel


# {{{ device tables

class DeviceExpnTables:
    """Device copies of one :class:`TermsTable` plus the C descriptor."""

    def __init__(self, table: TermsTable, device, k2rscale2=0.0):
        from sumpy_amd._capi import ExpnDesc
        p = table.order
        dev = torch.device(device)

        def t(a, dt=torch.int32):
            return torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)

        self.table = table
        self.full_ids = t(table.full.T.copy())
        self.stored_ids = t(table.stored.T.copy())
        self.stored_in_full = t(table.stored_in_full)
        cube = -np.ones((p + 1,)*3, dtype=np.int32)
        for i, mi in enumerate(table.full_list):
            cube[mi] = i
        self.cube_index = t(cube.ravel())
        rs, cols, vals = projection_csr(table, k2rscale2)
        self.proj_row_starts, self.proj_cols = t(rs), t(cols)
        self.proj_vals = t(vals.astype(np.float64), torch.float64)
        P = table.projection(k2rscale2)
        PT = P.T
        cs, rows, tv = [0], [], []
        for s in range(PT.shape[0]):
            nz = np.nonzero(PT[s])[0]
            rows.extend(nz.tolist())
            tv.extend(PT[s, nz].tolist())
            cs.append(len(rows))
        self.projT_col_starts = t(np.array(cs, np.int32))
        self.projT_rows = t(np.array(rows, np.int32))
        self.projT_vals = t(np.array(tv, np.float64), torch.float64)
        horner = []
        for i in range(p, -1, -1):
            for j in range(p - i, -1, -1):
                for k in range(p - i - j, -1, -1):
                    horner.append(table.full_index[(i, j, k)])
        self.horner_order = t(np.array(horner, np.int32))
        self.desc = ExpnDesc(
            p, table.nfull, table.nstored, self.full_ids.data_ptr(),
            self.stored_ids.data_ptr(), self.stored_in_full.data_ptr(),
            self.cube_index.data_ptr(), self.proj_row_starts.data_ptr(),
            self.proj_cols.data_ptr(), self.proj_vals.data_ptr(),
            self.projT_col_starts.data_ptr(), self.projT_rows.data_ptr(),
            self.projT_vals.data_ptr(), self.horner_order.data_ptr(),
            int(self.proj_vals.numel()))


_DEVICE_TABLES: dict = {}


def device_tables(order: int, compressed: bool, device, k2rscale2=0.0) -> DeviceExpnTables:
    key = (order, compressed, str(torch.device(device)), k2rscale2)
    tab = _DEVICE_TABLES.get(key)
    if tab is None:
        tab = DeviceExpnTables(get_terms_table(order, 3, compressed), device, k2rscale2)
        _DEVICE_TABLES[key] = tab
    return tab

# }}}


@dataclass(frozen=True)
class ExpansionBase:
    """Common interface (reference ``sumpy/expansion/__init__.py:70-245``)."""

    kernel: ScalarKernel
    order: int
    use_rscale: bool = field(default=True, kw_only=True)

    @property
    def dim(self):
        return self.kernel.dim

    @property
    def is_complex_valued(self):
        return self.kernel.is_complex_valued

    def with_kernel(self, kernel):
        return replace(self, kernel=kernel)

    def copy(self, **kwargs):
        return replace(self, **kwargs)

    def get_global_scaling_const(self):
        return self.kernel.get_global_scaling_const()

    def __len__(self):
        return len(self.get_coefficient_identifiers())


class VolumeTaylorExpansionMixin:
    compressed: bool = False

    @property
    def table(self) -> TermsTable:
        if self.kernel.dim not in (2, 3):
            raise NotImplementedError(
                "volume Taylor expansions are implemented for 2D and 3D kernels")
        from sumpy_amd.kernel import YukawaKernel
        base = self.kernel.get_base_kernel()
        yuk3 = isinstance(base, YukawaKernel) and base.dim == 3 and not self.compressed
        if not (isinstance(base, (LaplaceKernel, HelmholtzKernel)) or yuk3):
            raise NotImplementedError(
                "volume Taylor expansions on the GPU are implemented for the "
                "Laplace and Helmholtz kernels (and the full-Taylor 3D Yukawa kernel)")
        return get_terms_table(self.order, self.kernel.dim, self.compressed)

    def get_coefficient_identifiers(self):
        return self.table.stored_list

    def get_full_coefficient_identifiers(self):
        return self.table.full_list

    def get_storage_index(self, mi):
        return self.table.stored_index[tuple(mi)]

    def device_tables(self, device, k2rscale2=0.0) -> DeviceExpnTables:
        """Device tables; *k2rscale2* = ``(k rscale)^2`` enters the PDE
        projection of compressed Helmholtz expansions (0 for Laplace)."""
        if not self.compressed:
            k2rscale2 = 0.0
        return device_tables(self.order, self.compressed, device, float(k2rscale2))

    def tables_for(self, device, k, rscale) -> DeviceExpnTables:
        """Device tables at expansion scale *rscale* for wavenumber *k*."""
        return self.device_tables(device, (k*rscale)**2 if k else 0.0)


__all__ = [
    "DefaultExpansionFactory",
    "ExpansionBase",
    "VolumeTaylorExpansionFactory",
]


class VolumeTaylorExpansionFactory:
    """Volume-Taylor expansions for every kernel (reference
    ``__init__.py:1000-1030``)."""

    def get_local_expansion_class(self, base_kernel):
        from sumpy_amd.expansion.local import VolumeTaylorLocalExpansion
        return VolumeTaylorLocalExpansion

    def get_multipole_expansion_class(self, base_kernel):
        from sumpy_amd.expansion.multipole import VolumeTaylorMultipoleExpansion
        return VolumeTaylorMultipoleExpansion


class DefaultExpansionFactory:
    """PDE-compressed Taylor expansions (reference ``__init__.py:1033-1085``)."""

    def get_local_expansion_class(self, base_kernel):
        from sumpy_amd.expansion.local import LinearPDEConformingVolumeTaylorLocalExpansion
        return LinearPDEConformingVolumeTaylorLocalExpansion

    def get_multipole_expansion_class(self, base_kernel):
        from sumpy_amd.expansion.multipole import (
            LinearPDEConformingVolumeTaylorMultipoleExpansion,
        )
        return LinearPDEConformingVolumeTaylorMultipoleExpansion


@lru_cache(maxsize=None)
def m2l_device_tables(src_order: int, tgt_order: int, compressed: bool, device: str):
    """Circulant-layout tables of the dense class M2L (``m2l.py:403-459``)."""
    src = get_terms_table(src_order, 3, compressed)
    tgt = get_terms_table(tgt_order, 3, compressed)
    shape = m2l_circulant_shape(src, tgt)
    dev = torch.device(device)
    off_t = circulant_offsets(tgt.stored, shape)
    off_s = circulant_offsets(src.stored, shape)
    deg = tgt.stored.sum(axis=1).astype(np.int32)

    def t(a):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(dev)

    return dict(shape=shape, ncirc=int(np.prod(shape)), off_t=t(off_t), off_s=t(off_s),
                tgt_degree=t(deg), shape_dev=t(np.array(shape, np.int32)),
                nl=tgt.nstored, nm=src.nstored)


