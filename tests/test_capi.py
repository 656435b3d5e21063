"""C-ABI library: loads, exports every declared symbol, rejects bad input.

Runs without a GPU: nothing here launches a kernel.
"""
import ctypes as C
import os
import re

import pytest

import branson_amd
from branson_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "branson_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bh_[a-z0-9_]+)\s*\(", src)))


def test_header_declarations_match_binding_list():
    assert _declared() == sorted(capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = branson_amd.lib()
    for name in _declared():
        assert hasattr(L, name), name


def test_abi_version():
    assert branson_amd.lib().bh_abi_version() == 4


def test_null_arguments_are_errors_not_crashes():
    L = branson_amd.lib()
    assert L.bh_create(0, None, None) == -1
    assert L.bh_sync(None) == -1
    assert L.bh_transport(None, 0.0, None, None, None) == -1
    assert L.bh_census_size(None, None) == -1
    assert L.bh_last_error(None) == b"null handle"


def _has_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device error path")
def test_create_without_device_reports_error():
    L = branson_amd.lib()
    h = C.c_void_p()
    rc = L.bh_create(0, None, C.byref(h))
    assert rc == -2
    assert b"no HIP device" in L.bh_last_error(h)
    L.bh_destroy(h)


def test_photon_layout_is_reference_120_bytes():
    assert capi.PHOTON_DTYPE.itemsize == 120
    off = {n: capi.PHOTON_DTYPE.fields[n][1] for n in capi.PHOTON_DTYPE.names}
    # src/photon.h:172-181 member order
    assert off == {"cell": 0, "group": 4, "source_type": 8, "descriptor": 12, "pos": 16,
                   "angle": 40, "E": 64, "E0": 72, "life_dx": 80, "rng": 88}
