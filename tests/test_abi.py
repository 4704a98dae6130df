"""The C-ABI library loads and exports every entry point include/commit_hip.h
declares (no GPU needed: no compute call is made)."""
import ctypes
import os
import re
import subprocess

import pytest

from commit_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "commit_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(commit_hip_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("commit_hip_plan_create", "commit_hip_A", "commit_hip_At", "commit_hip_plan_destroy"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert sorted(_lib.SIGNATURES) == declared_symbols()


def test_exports_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (commit_hip_\w+)", out))
    assert set(declared_symbols()) <= exported


def test_version_and_error_strings_without_gpu():
    lib = _lib.load()
    assert b"gfx950" in lib.commit_hip_version()
    assert isinstance(lib.commit_hip_last_error(), bytes)


def test_plan_create_rejects_bad_arguments_before_touching_the_device():
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.commit_hip_plan_create(ctypes.byref(h), 0, 1, 1, 0, 0, 0, 1, None, None, None, None, None,
                                    None, None, None, None, None, None, 1, 0, 0, 0, 0, 0)
    assert rc == -1 and b"size" in lib.commit_hip_last_error()
    rc = lib.commit_hip_plan_create(ctypes.byref(h), 8, 1, 1, 0, 0, 0, 1, None, None, None, None, None,
                                    None, None, None, None, None, None, 1, 0, 0, 0, 7, 0)
    assert rc == -1 and b"precision" in lib.commit_hip_last_error()     # unknown precision
    assert lib.commit_hip_A(None, None, None, None) == -1


def test_oracle_is_not_imported_by_the_package():
    """The product path never routes through the oracle."""
    pkg = os.path.join(ROOT, "commit_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".c", ".h")):
                src = open(os.path.join(dirpath, f), encoding="utf-8", errors="replace").read()
                assert "oracle" not in re.sub(r"#.*|//.*", "", src).lower().replace("oracle_", "x"), f


def test_plan_create_rejects_more_than_twenty_response_functions():
    """The reference generates code for at most 20 RFs per compartment
    (setup_operator.py:70, 128, 183); 21 is refused before any device call."""
    lib = _lib.load()
    h = ctypes.c_void_p()
    for rf in ((21, 0, 0), (1, 21, 0), (1, 0, 21)):
        rc = lib.commit_hip_plan_create(ctypes.byref(h), 8, 1, 1, 0, 0, 0, 1, None, None, None, None, None,
                                        None, None, None, None, None, None, rf[0], rf[1], rf[2], 0, 0, 0)
        assert rc == -4 and b"response functions" in lib.commit_hip_last_error(), rf
