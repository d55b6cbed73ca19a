"""The C ABI: libfeaxhip.so loads and exports exactly what include/feaxhip.h
declares; the ctypes binding covers every declaration.  No device calls."""

import os
import re
import subprocess

from conftest import ROOT


def header_symbols():
    h = open(os.path.join(ROOT, "include", "feaxhip.h")).read()
    return set(re.findall(r"^(?:int|size_t|int64_t|int32_t) (fx_\w+)\(", h, re.M))


def test_library_exports_every_declared_symbol():
    from feax_amd import _native as nat
    L = nat.lib()
    syms = header_symbols()
    assert len(syms) >= 25
    for s in sorted(syms):
        assert hasattr(L, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", nat.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (fx_\w+)$", out, re.M))
    assert syms == exported, (syms ^ exported)


def test_ctypes_binding_matches_header():
    from feax_amd._native import SIGNATURES
    assert set(SIGNATURES) == header_symbols()


def test_version_and_workspace_query():
    import ctypes
    from feax_amd import _native as nat
    buf = ctypes.create_string_buffer(64)
    nat.check(nat.lib().fx_version(buf, 64))
    assert b"gfx950" in buf.value
    w = nat.lib().fx_pcg_workspace_bytes(1000, 8)
    assert w >= 4 * 8000


def test_product_has_no_oracle_or_cpu_fallback_imports():
    """feax_amd never imports the oracle, scipy or a CPU solver."""
    pkg = os.path.join(ROOT, "feax_amd")
    for dp, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                src = open(os.path.join(dp, f)).read()
                mods = re.findall(r"^\s*(?:from|import)\s+([\w.]+)", src, re.M)
                # scipy.sparse only as an export container (CSRMatrix.to_scipy)
                assert not any(m.split(".")[0] == "oracle" or m.startswith("scipy.sparse.linalg")
                               or m.startswith("scipy.linalg") for m in mods), (f, mods)
