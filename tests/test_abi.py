"""The C-ABI library loads and exports every symbol include/*.h declares."""
import ctypes
import os
import re

import scp3_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"\b(scp3g_[a-z0-9_]+)\s*\(", src))
    return names


def test_library_exports_declared_symbols():
    L = ctypes.CDLL(scp3_amd.LIB_PATH)
    names = declared()
    assert len(names) >= 25
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_python_bindings_cover_header():
    assert declared() == set(scp3_amd.exported_symbols())


def test_create_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        return
    h = ctypes.c_void_p()
    rc = scp3_amd.lib().scp3g_create(ctypes.byref(h), 0)
    assert rc != 0 and not h.value


def test_cpu_pes_matches_oracle(hexamer):
    from oracle import oracle
    crd = hexamer["x0"] * scp3_amd.BOHR
    assert abs(scp3_amd.pes_cpu("qtip4pf", crd) - oracle.qtip4pf(crd)) < 1e-12 * 50
