"""C-ABI library checks that need no GPU: it loads, and it exports exactly
the entry points include/bem_hip.h declares (and the ctypes table binds)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bem_hip.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bem_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = _declared()
    assert "bem_assemble" in names and "bem_gmres" in names and len(names) >= 15


def test_library_exports_every_declared_symbol():
    from neuralat_amd import _lib

    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    from neuralat_amd import _lib

    assert sorted(_lib.SIGNATURES) == _declared()
    lib = _lib.load()
    assert b"gfx950" in lib.bem_version()


def test_library_built_for_gfx950_only():
    from neuralat_amd import _lib

    data = open(_lib.LIB_PATH, "rb").read()
    # every offload bundle in the fat binary targets gfx950, and no NVIDIA
    # target (nvptx / sm_NN) is embedded anywhere in the file
    targets = set(re.findall(rb"amdgcn-amd-amdhsa-[-a-z0-9:+]*", data))
    assert targets == {b"amdgcn-amd-amdhsa--gfx950"}, targets
    assert re.search(rb"nvptx|sm_[0-9]{2}", data) is None


def test_missing_library_fails_loudly(monkeypatch):
    from neuralat_amd import _lib

    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libbem_hip.so")
    with pytest.raises(ImportError):
        _lib.load()


def test_status_mapping():
    from neuralat_amd import _lib

    _lib.load()
    with pytest.raises(ValueError):
        _lib.check(_lib.BEM_ERR_BAD_ARG)
    with pytest.raises(_lib.NotConvergedError):
        _lib.check(_lib.BEM_ERR_NOT_CONVERGED)
    with pytest.raises(_lib.BemError):
        _lib.check(_lib.BEM_ERR_HIP)


def test_assemble_parts_validated_before_device_use():
    """parts must be 1, 2 or 3, optionally plus 4 (P1 ordered-pair sweep) and
    8 (no scratch matrix); the check precedes any pointer or device use."""
    from neuralat_amd import _lib

    lib = _lib.load()
    for parts in (0, 4, 8, 12, 16, -1):
        rc = lib.bem_assemble_ex(None, 1.0, 0, 0, 0, 1, parts, None, 1, None)
        assert rc == _lib.BEM_ERR_BAD_ARG
        assert b"parts" in lib.bem_last_error()
    for parts in (1, 5, 7, 11, 15):  # accepted: the null mesh is the error reported
        assert lib.bem_assemble_ex(None, 1.0, 0, 0, 0, 1, parts, None, 1, None) == _lib.BEM_ERR_BAD_ARG
        assert b"null" in lib.bem_last_error()


def test_fem_rejects_bad_meshes():
    """bem_fem_create validates on the host before any device call."""
    import ctypes

    import numpy as np

    from neuralat_amd import _lib

    lib = _lib.load()
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0]], dtype=np.float32)
    h = ctypes.c_void_p()
    for tets in ([[0, 1, 2, 7]], [[0, 1, 1, 3]], [[0, 1, 2, 4]]):  # out of range, repeated, flat
        t = np.array(tets, dtype=np.int32)
        rc = lib.bem_fem_create(v.ctypes.data, len(v), t.ctypes.data, len(t), ctypes.byref(h))
        assert rc == _lib.BEM_ERR_BAD_MESH, tets
        assert lib.bem_fem_last_error()
