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


C_CALLER = r"""
#include "drh.h"
#include <stdio.h>
#include <string.h>
int main(void) {
  drh_handle h = 0;
  drh_scene_desc d;
  drh_camera cam;
  memset(&d, 0, sizeof d);
  memset(&cam, 0, sizeof cam);
  drh_status s = drh_create(NULL, &h);
  printf("%d %d %d\n", (int)s, d.n_meshes, (int)sizeof cam);
  return (s == DRH_E_INVALID && drh_commit(NULL, &d) == DRH_E_INVALID && drh_grad_size(NULL) == -1) ? 0 : 1;
}
"""


@pytest.mark.parametrize("lang", ["c", "c++"])
def test_header_is_plain_c_and_links(tmp_path, lang):
    """include/drh.h compiles as C99 and as C++ with warnings as errors, and a program built
    against it links libdrh.so and gets the host-side error codes (no device work)."""
    import subprocess
    src = tmp_path / "caller.c"
    src.write_text(C_CALLER)
    exe = tmp_path / "caller"
    libdir = os.path.dirname(drh.LIB_PATH)
    std = ["-x", "c", "-std=c99"] if lang == "c" else ["-x", "c++", "-std=c++17"]
    cc = "gcc" if lang == "c" else "g++"
    subprocess.run([cc, *std, "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                    "-L", libdir, "-ldrh", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.split()[0] == "-1"
