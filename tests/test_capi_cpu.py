"""CPU-side checks of the C ABI: the library loads and exports every symbol include/mgo.h declares, the
ctypes mirror matches the header, and host-side argument validation needs no device."""
import ctypes as C
import os
import re

from memilio_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "mgo.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(mgo_[a-z_]+)\s*\(", text))


def test_header_matches_ctypes_mirror():
    assert header_functions() == set(capi.SIGNATURES)


def test_library_exports_every_symbol():
    lib = capi.load()
    for name in header_functions():
        assert hasattr(lib, name), name


def test_version_string():
    lib = capi.load()
    assert b"gfx950" in lib.mgo_version()


def test_struct_sizes():
    # plain C layout of the descriptors (x86-64): pointers 8, int32 4
    # 2 int32 + 10 pointers + int32 (padded) + 3 pointers + int32 (padded)
    assert C.sizeof(capi.GraphDesc) == 8 + 10 * 8 + 8 + 3 * 8 + 8
    assert C.sizeof(capi.IntegratorCfg) == 40
    assert C.sizeof(capi.SimCfg) == 32 + 40


def test_context_create_without_device_fails_cleanly():
    """No GPU in this container: context creation must return an error, not crash or fall back."""
    import torch
    if torch.cuda.is_available():
        return
    lib = capi.load()
    p = C.c_void_p()
    cfg = capi.DeviceCfg(0, 0, 0)
    st = lib.mgo_context_create(C.byref(cfg), C.byref(p))
    assert st == capi.ERR_DEVICE
    assert not p.value
