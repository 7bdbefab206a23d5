"""CPU-side checks of the C ABI: the in-tree libraries load and export every
entry point that include/*.h declares (no compute calls without a GPU)."""
import ctypes as ct
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z0-9_]+)\s*\(", src)))


@pytest.mark.parametrize("header,lib", [("lbm_hip.h", "lbm_amd/liblbm_hip.so"), ("lbm_batch.h", "lbm_amd/liblbm_hip.so"),
                                        ("lbm_surrogate.h", "lbm_amd/liblbm_hip.so"),
                                        ("lbm_poisson.h", "lbm_amd/liblbm_hip.so"),
                                        ("lbm_host.h", "lbm_amd/liblbm_host.so")])
def test_library_exports_header(header, lib):
    path = os.path.join(ROOT, lib)
    if not os.path.exists(path):
        pytest.fail(f"{path} not built")
    so = ct.CDLL(path)
    names = declared(header)
    assert len(names) >= 5
    for n in names:
        assert hasattr(so, n), n


def test_python_binding_lists_all_exports():
    from lbm_amd import capi

    assert sorted(capi.EXPORTS) == declared("lbm_hip.h")


def test_python_bindings_list_batch_and_surrogate_exports():
    from lbm_amd import capi, poisson, surrogate

    assert sorted(capi.BATCH_EXPORTS) == declared("lbm_batch.h")
    assert sorted(surrogate.EXPORTS) == declared("lbm_surrogate.h")
    assert sorted(poisson.EXPORTS) == declared("lbm_poisson.h")


def test_every_header_is_checked():
    assert sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")) == \
        ["lbm_batch.h", "lbm_hip.h", "lbm_host.h", "lbm_poisson.h", "lbm_surrogate.h"]
