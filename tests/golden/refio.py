"""Run the reference (compiled from its own sources by ``make -C oracle ref``) on a scene
and read its outputs (oracle/ref_run.py), and its mesh / image file code (oracle/_ref/
ref_meshio).  Used only by ``make_golden.py`` in this container, to write the golden vectors
under tests/golden/; the GPU tests never run it."""
from __future__ import annotations

import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from oracle.ref_run import (HARNESS, INT_TAGS, OP_COMPARE, OP_DRENDER, OP_EDGES, OP_HITS, OP_OPT_ALL,  # noqa: E402,F401
                            OP_OPTIMIZE, OP_RENDER, OP_SEQ, OP_TRANSFORMR, TAGS, run)


MESHIO = os.path.join(HERE, "..", "..", "oracle", "_ref", "ref_meshio")


def load_vsgf(path):
    """The reference's LoadMeshFromVSGF on `path`: float4 positions, normals, tangents,
    float2 texture coordinates, uint32 indices and material indices."""
    with tempfile.TemporaryDirectory() as td:
        fo = os.path.join(td, "out.bin")
        subprocess.run([MESHIO, "load", path, fo], check=True)
        data = open(fo, "rb").read()
    V, I = struct.unpack_from("<2i", data, 0)
    off, out = 8, {}
    for name, dt, n, w in (("pos4", np.float32, V, 4), ("nrm4", np.float32, V, 4), ("tang4", np.float32, V, 4),
                           ("tc", np.float32, V, 2), ("idx", np.uint32, I, 1), ("mat", np.uint32, I // 3, 1)):
        out[name] = np.frombuffer(data, dt, count=n * w, offset=off).reshape(n, w).copy()
        off += n * w * 4
    out["idx"] = out["idx"].reshape(-1, 3)
    out["mat"] = out["mat"].reshape(-1)
    return out


def save_vsgf(path, pos4, nrm4, tc, idx):
    """The reference's SaveMeshToVSGF of a mesh built with SimpleMesh(V, I) (material 0)."""
    V, I = pos4.shape[0], idx.size
    buf = struct.pack("<2i", V, I) + np.asarray(pos4, np.float32).tobytes() + np.asarray(nrm4, np.float32).tobytes() \
        + np.asarray(tc, np.float32).tobytes() + np.asarray(idx, np.uint32).tobytes()
    with tempfile.TemporaryDirectory() as td:
        fi = os.path.join(td, "in.bin")
        with open(fi, "wb") as f:
            f.write(buf)
        subprocess.run([MESHIO, "save", fi, path], check=True)


def save_image(path, img, gamma=2.2):
    """The reference's SaveImage<float3> of an (H, W, 3) float image, row 0 at the top."""
    H, W, _ = img.shape
    with tempfile.TemporaryDirectory() as td:
        fi = os.path.join(td, "in.bin")
        with open(fi, "wb") as f:
            f.write(struct.pack("<2i", W, H) + np.ascontiguousarray(img, np.float32).tobytes())
        subprocess.run([MESHIO, "image", fi, path, repr(float(gamma))], check=True)
