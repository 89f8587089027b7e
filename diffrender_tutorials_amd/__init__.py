"""MI355X-native differentiable-rendering path (render / d_render behind IDiffRender).

The product is libdrh.so (HIP kernels for gfx950 + a C ABI, include/drh.h).
``drh`` is its Python binding, ``scenes`` generates the synthetic benchmark scenes.
"""
from . import scenes  # noqa: F401
from .drh import DiffRender, MakeDifferentialRenderer, DrhError  # noqa: F401
