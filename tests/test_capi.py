"""The C-ABI library loads without a GPU and exports every symbol that
include/sumpyhip.h declares (sumpyhip_* and sumpytree_gpu_*), with matching
argument counts."""

import os
import re

from sumpy_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_prototypes():
    hdr = open(os.path.join(ROOT, "include", "sumpyhip.h")).read()
    return re.findall(r"^(?:const char \*|int )(sumpy(?:hip|tree_gpu)_\w+)\(([^;]*)\);", hdr,
                      re.M)


def test_library_exports_header_symbols():
    lib = _capi.load()
    protos = header_prototypes()
    assert len(protos) >= 18
    for name, args in protos:
        assert hasattr(lib, name), name
        nargs = 0 if args.strip() == "void" else len(args.split(","))
        if name != "sumpyhip_last_error":
            assert len(_capi._SIGNATURES[name]) == nargs, name


def test_errors_without_gpu_do_not_crash():
    lib = _capi.load()
    assert lib.sumpyhip_last_error() is not None
