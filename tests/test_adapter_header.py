"""include/drh_idiffrender.h compiles against the reference's own IDiffRender header
(src/virtual_drender.h) with every warning an error, in C++14 (the reference's standard) and
C++17, with the compiler the reference builds with here.  Skipped where the reference's
headers are absent (the GPU host)."""
from __future__ import annotations

import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
INC = os.path.join(HERE, "..", "include")
REF = "/root/reference"
CXX = "/opt/rocm/llvm/bin/clang++"

TU = r"""
#include "drh_idiffrender.h"
// the adapter is an IDiffRender, made by the factory for the four shading models
static_assert(std::is_base_of<diff_render::IDiffRender, diff_render::DiffRenderHIP>::value, "IDiffRender");
int use(const diff_render::Scene& s, const diff_render::CamInfo* c, diff_render::Img* im,
        diff_render::DScene& d) {
  diff_render::DiffRenderSettings st;
  st.mode = diff_render::SHADING_MODEL::LAMBERT;
  st.spp = 4;
  std::shared_ptr<diff_render::IDiffRender> r = diff_render::MakeDifferentialRendererHIP(st);
  r->commit(s);
  r->render(s, c, im, 1);
  r->d_render(s, c, im, 1, 1000, d);
  return (int)r->d_render_and_compare(s, c, im, 1, 1000, d, im);
}
"""


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "src", "virtual_drender.h")),
                    reason="the reference's headers are not on this host")
@pytest.mark.skipif(not os.path.exists(CXX), reason="clang++ not installed")
@pytest.mark.parametrize("std", ["c++14", "c++17"])
def test_adapter_compiles_against_reference_interface(tmp_path, std):
    src = tmp_path / "use_adapter.cpp"
    src.write_text(TU)
    cmd = [CXX, f"-std={std}", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-I", INC,
           "-isystem", os.path.join(REF, "src"), "-isystem", os.path.join(REF, "common_src"), str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
