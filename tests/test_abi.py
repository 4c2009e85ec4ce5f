"""The C-ABI library loads and exports every symbol the public headers declare (no device calls)."""
import os
import re

import ferrite_amd as fa
from ferrite_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        with open(os.path.join(ROOT, "include", h)) as fh:
            txt = fh.read()
        names |= set(re.findall(r"^\s*(?:int32_t|void|const char\*)\s+(fem_\w+)\s*\(", txt, re.M))
    return names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(fa.LIB_PATH), "build the library first (__graft_entry__.build())"
    decl = declared_symbols()
    assert len(decl) >= 25
    lib = _lib.lib()
    missing = [n for n in sorted(decl) if getattr(lib, n, None) is None]
    assert not missing, missing
    # and the ctypes table covers all of them
    assert decl <= set(_lib.SIGNATURES)


def test_ctx_create_without_gpu_fails_cleanly():
    # count devices without initialising torch's own HIP runtime: a second runtime initialised
    # in this process before the library's would leave the library without a device
    import torch
    if torch.cuda.device_count() > 0:
        return
    try:
        fa.Context(0)
    except fa.FemError as e:
        assert e.code in (_lib.FEM_ERR_HIP, _lib.FEM_ERR_ARG)
    else:
        raise AssertionError("context creation must fail without a device")
