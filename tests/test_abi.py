"""libdrh.so loads and exports every entry point include/drh.h declares; host-side error
paths.  CPU only: no kernel is launched here."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

from diffrender_tutorials_amd import drh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "drh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(drh_[a-z_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared() == sorted(drh.EXPORTS)


def test_library_exports_every_declared_symbol():
    assert os.path.exists(drh.LIB_PATH), "build libdrh.so first (python -m diffrender_tutorials_amd.build)"
    lib = ctypes.CDLL(drh.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_invalid_arguments_are_rejected_before_any_device_work():
    L = drh.load()
    h = ctypes.c_void_p()
    assert L.drh_create(None, ctypes.byref(h)) == -1
    s = drh.Settings()
    s.mode, s.spp = 9, 1
    assert L.drh_create(ctypes.byref(s), ctypes.byref(h)) == -1
    s.mode, s.spp = 2, 0
    assert L.drh_create(ctypes.byref(s), ctypes.byref(h)) == -1
    assert L.drh_commit(None, None) == -1
    assert L.drh_render(None, None, 0, None) == -1
    assert L.drh_grad_size(None) == -1
    assert L.drh_last_error(None) == b"null handle"


def test_oracle_is_not_imported_by_the_product():
    pkg = os.path.join(ROOT, "diffrender_tutorials_amd")
    for dp, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                txt = open(os.path.join(dp, f)).read()
                assert "import oracle" not in txt and "liboracle" not in txt and "cpu_ref" not in txt, f
