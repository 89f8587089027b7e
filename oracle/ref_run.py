"""Run the reference, compiled from its own sources by ``make -C oracle ref``
(oracle/_ref/ref_harness), on a scene and read its outputs.

TEST INFRASTRUCTURE ONLY: used by tests/golden/make_golden.py (in the container that holds the
reference sources) to write golden vectors, and by bench.py's CPU-baseline leg to time the
reference's own optimisation loop (OptSimple::Run with its brute-force tracer).  The shipped
package never imports it."""
from __future__ import annotations

import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "_ref", "ref_harness")

OP_RENDER, OP_HITS, OP_DRENDER, OP_COMPARE, OP_SEQ, OP_EDGES, OP_OPTIMIZE, OP_OPT_ALL = 1, 2, 4, 8, 16, 32, 64, 128
OP_TRANSFORMR = 256
TAGS = {1: "images", 2: "hit_prim", 3: "hit_geom", 4: "hit_t", 5: "hit_u", 6: "hit_v", 7: "grad", 8: "loss",
        9: "qmc", 10: "prng", 11: "hamm", 12: "edge_v0", 13: "edge_v1", 14: "edge_pmf", 15: "edge_cdf",
        16: "compare_grad", 17: "compare_images", 18: "proj", 19: "opt_pos", 20: "opt_col", 21: "opt_err",
        22: "proj_inv", 23: "wv_inv", 24: "wvp", 25: "opt_tr", 26: "opt_tex", 27: "transformr",
        28: "opt_seconds", 29: "omp_threads"}
INT_TAGS = {"hit_prim", "hit_geom", "qmc", "edge_v0", "edge_v1"}


def available() -> bool:
    return os.path.exists(HARNESS)


def run(meshes, cams, mode, spp, ops, edge_samples=0, images=None, opt_iters=0, opt_lrs=None, transformr=None):
    """opt_lrs = (position_lr, textures_lr, transforms_lr) runs the optimiser with OPT_ALL;
    transformr = (n, 7) parameter sets returns TransformR::to_mat of each."""
    buf = bytearray()
    if opt_lrs is not None:
        ops |= OP_OPT_ALL
    buf += struct.pack("<8i", 0x44524831, mode, spp, len(meshes), len(cams), edge_samples, ops, opt_iters)
    if opt_lrs is not None:
        buf += struct.pack("<3f", *opt_lrs)
    if transformr is not None:
        # the harness reads this block right after the header (and the rates)
        buf[24:28] = struct.pack("<i", ops | OP_TRANSFORMR)
        x = np.asarray(transformr, np.float32).reshape(-1, 7)
        buf += struct.pack("<i", x.shape[0]) + x.tobytes()
    for m in meshes:
        tw = 0 if m.tex is None else m.tex.shape[1]
        th = 0 if m.tex is None else m.tex.shape[0]
        buf += struct.pack("<8i", m.V, m.F, m.col is not None, m.tc is not None, m.nrm is not None, tw, th,
                           1 if m.differentiable else 0)
        buf += np.asarray(m.transform, np.float32).tobytes()
        buf += np.asarray(m.pos, np.float32).tobytes()
        for a in (m.col, m.tc, m.nrm):
            if a is not None:
                buf += np.asarray(a, np.float32).tobytes()
        buf += np.asarray(m.idx, np.uint32).tobytes()
        if m.tex is not None:
            buf += np.asarray(m.tex, np.float32).tobytes()
    for c in cams:
        buf += np.asarray(c.world_view, np.float32).tobytes() + np.asarray(c.proj, np.float32).tobytes()
        buf += struct.pack("<2f", float(c.width), float(c.height))
    if images is not None:
        for im in images:
            buf += np.asarray(im, np.float32).tobytes()
    with tempfile.TemporaryDirectory() as td:
        fi, fo = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fi, "wb") as f:
            f.write(buf)
        subprocess.run([HARNESS, fi, fo], check=True)
        data = open(fo, "rb").read()
    out, off = {}, 0
    while off < len(data):
        tag, n = struct.unpack_from("<2i", data, off)
        off += 8
        name = TAGS[tag]
        dt = np.uint32 if name in INT_TAGS else np.float32
        if name in ("edge_v0", "edge_v1", "omp_threads"):
            dt = np.int32
        if name == "opt_seconds":
            dt = np.float64
        out[name] = np.frombuffer(data[off:off + n], dtype=dt).copy()
        off += n
    return out
