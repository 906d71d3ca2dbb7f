"""The C-ABI library loads and exports every entry point include/cxg.h declares.

CPU-only: no compute call is made (no GPU here); only symbol presence, the pure
host queries and argument checks that fail before touching a device.
"""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cxg.h")
LIB = os.path.join(ROOT, "calculixpp_amd", "libcxg.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cxg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_path():
    names = declared()
    for must in ["cxg_create", "cxg_mesh_create", "cxg_pattern_build", "cxg_assemble_linear",
                 "cxg_internal_force", "cxg_assemble_j2", "cxg_commit", "cxg_vjp"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build libcxg.so first (python -m calculixpp_amd.build)"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\s[TW]\s+(cxg_\w+)$", out, flags=re.M))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from calculixpp_amd import _lib
    assert sorted(_lib.exported_symbols()) == declared()


def test_library_loads_and_host_queries():
    from calculixpp_amd import _lib
    L = _lib.lib()
    assert L.cxg_version().decode().startswith("cxg")
    assert [L.cxg_nodes_per_element(t) for t in (_lib.C3D8, _lib.C3D20, _lib.C3D20R, _lib.C3D6)] \
        == [8, 20, 20, 6]
    assert [L.cxg_gauss_points(t) for t in (_lib.C3D8, _lib.C3D20, _lib.C3D20R, _lib.C3D6)] \
        == [8, 27, 8, 2]
    assert L.cxg_nodes_per_element(_lib.C3D10) == 0  # outside the GPU path


def test_null_arguments_are_rejected_without_a_device():
    from calculixpp_amd import _lib
    L = _lib.lib()
    assert L.cxg_create(0, None, None) == _lib.ERR_ARG
    assert L.cxg_mesh_info(None, None, None, None) == _lib.ERR_ARG
    assert L.cxg_pattern_info(None, None, None) == _lib.ERR_ARG
    assert L.cxg_last_error(None) == b"null context"
    assert L.cxg_profile_count(None) == 0


def test_product_does_not_import_the_oracle():
    """The HIP path never routes through oracle/: no import of it in the package."""
    pkg = os.path.join(ROOT, "calculixpp_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in re.sub(r"(#|//).*", "", text), f


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from calculixpp_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libcxg.so"))
    with pytest.raises(_lib.CxgLibraryError):
        _lib.lib()
