"""The C-ABI library loads and exports every function include/canoe_gpu.h
declares (CPU only: no compute call is made)."""
import ctypes
import os
import re

import canoe_amd as ca
from canoe_amd import mlp as cm

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(
    __file__))), "include", "canoe_gpu.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cg_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_function_is_exported():
    names = declared()
    assert len(names) >= 30
    L = ctypes.CDLL(ca.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    bound = set(ca._SIGS) | set(cm._SIGS)
    missing = [n for n in declared() if n not in bound]
    assert not missing, missing


def test_version_string_names_the_arch():
    assert b"gfx950" in ca.lib().cg_version()


def test_python_constants_match_the_header():
    src = open(HEADER).read()
    nk = int(re.search(r"#define CG_NUM_KERNEL_IDS (\d+)", src).group(1))
    npar = int(re.search(r"#define CG_MLP_NPARAM (\d+)", src).group(1))
    assert ca.NUM_KERNEL_IDS == nk
    assert cm.NPARAM == npar
    m = re.search(r"#define CG_NOALLOC_SCRATCH_BYTES \(\(size_t\)(\d+) << (\d+)\)", src)
    assert ca.NOALLOC_SCRATCH_BYTES == int(m.group(1)) << int(m.group(2))
    assert ctypes.sizeof(ca.Tune) == 6 * 4 + 8  # 6 ints, double
    codes = dict((m.group(1), int(m.group(2)))
                 for m in re.finditer(r"CG_(OK|ERR_[A-Z]+) = (\d+)", src))
    for name, v in codes.items():
        assert getattr(ca, name) == v, name

