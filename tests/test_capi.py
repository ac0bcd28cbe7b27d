"""The C-ABI library loads and exports every symbol include/sgpp_hip.h
declares; the ctypes binding covers all of them. CPU only (no compute)."""
import os
import re
import subprocess

from sgpp_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "sgpp_hip.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(sgpp_hip_[a-z0-9_]+)\s*\(", text))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "sgpp_hip_meval_mult" in syms and "sgpp_hip_meval_mult_transpose" in syms
    assert len(syms) >= 30


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (sgpp_hip_\w+)", out))
    missing = declared_symbols() - exported
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert declared_symbols() <= set(capi.SIGNATURES)
    for name in declared_symbols():
        assert hasattr(capi.lib, name)


def test_version_and_error_string():
    assert capi.F["sgpp_hip_version"]() >= 10000
    assert isinstance(capi.last_error(), str)
