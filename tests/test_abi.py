"""The C-ABI library loads and exports every symbol include/tm_abi.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from everest_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tm_[a-z_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(abi.LIB_PATH)
    names = declared("tm_abi.h")
    assert "tm_count" in names and "tm_graph_upload" in names
    for n in names:
        assert hasattr(L, n), n


def test_version_and_no_device_ok():
    L = abi.lib()
    text = open(os.path.join(ROOT, "include", "tm_abi.h")).read()
    assert L.tm_abi_version() == int(re.search(r"#define TM_ABI_VERSION (\d+)", text).group(1))
    assert L.tm_device_count() >= 0


def test_plugin_exports_getworker():
    L = ctypes.CDLL(os.path.join(abi.LIB_DIR, "libTM.so"), mode=os.RTLD_GLOBAL if hasattr(os, "RTLD_GLOBAL") else 0)
    assert hasattr(L, "getWorker")


def test_handwritten_plugin_exports_getworker():
    L = ctypes.CDLL(os.path.join(abi.LIB_DIR, "libCSR.so"), mode=os.RTLD_GLOBAL if hasattr(os, "RTLD_GLOBAL") else 0)
    assert hasattr(L, "getWorker")
