"""Mesh and image files for scenes and visual checks (host side, numpy).

* VSGF meshes, the reference's binary mesh format (common_src/cmesh.cpp:81-205): a 24-byte
  header (uint64 file size; uint32 vertex count, index count, material count, flags), then
  float4 positions, float4 normals unless flag 8 (no normals), float4 tangents if flag 1,
  float2 texture coordinates, uint32 indices and one uint32 material index per triangle.
* 8-bit images with the reference's tone map (common_src/Image2d.cpp:580-665): each channel
  int(x^(1/gamma) * 255 + 0.5) clamped to [0, 255], gamma 2.2 by default; ``.bmp`` (24-bit,
  rows bottom-up), ``.ppm`` (ASCII P3) and ``.png`` (8-bit RGBA with alpha 255, rows top-down:
  the pixels the reference hands to stb's PNG writer, common_src/Image2d.cpp:629-646; the
  stream is written here with zlib, so the compressed bytes differ from stb's, the pixels
  do not).  The BMP rows are padded to 4 bytes and the file
  size field holds the whole file, as the format specifies (the reference writes neither).
* ``.image3f``: uint32 width, height, then the float RGB pixels (the reference's raw dump).

Images are (H, W, 3) float arrays with row 0 at the top, the layout ``DiffRender.render``
returns.
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

from .scenes import Mesh

_HAS_TANGENT, _HAS_NO_NORMALS = 1, 8
_HEADER = struct.Struct("<QIIII")


def load_vsgf(path: str) -> Mesh:
    """Read a VSGF mesh: positions, normals (when stored), texture coordinates, triangles."""
    with open(path, "rb") as f:
        data = f.read()
    if len(data) < _HEADER.size:
        raise ValueError(f"{path}: shorter than a VSGF header")
    _, nv, ni, _, flags = _HEADER.unpack_from(data, 0)
    if ni % 3:
        raise ValueError(f"{path}: {ni} indices is not a triangle list")
    off = _HEADER.size

    def take(dtype, count, width):
        nonlocal off
        n = count * width * np.dtype(dtype).itemsize
        if off + n > len(data):
            raise ValueError(f"{path}: truncated")
        a = np.frombuffer(data, dtype=dtype, count=count * width, offset=off).reshape(count, width)
        off += n
        return a

    pos4 = take(np.float32, nv, 4)
    nrm4 = None if flags & _HAS_NO_NORMALS else take(np.float32, nv, 4)
    if flags & _HAS_TANGENT:
        take(np.float32, nv, 4)
    tc = take(np.float32, nv, 2)
    idx = take(np.uint32, ni // 3, 3)
    if ni and int(idx.max()) >= nv:
        raise ValueError(f"{path}: index out of range")
    return Mesh(pos=np.ascontiguousarray(pos4[:, :3]), idx=np.ascontiguousarray(idx),
                tc=np.ascontiguousarray(tc),
                nrm=None if nrm4 is None else np.ascontiguousarray(nrm4[:, :3]))


def save_vsgf(path: str, mesh: Mesh) -> None:
    """Write a VSGF mesh (w = 1 for positions, 0 for normals; material 0 for every triangle)."""
    V, F = mesh.V, mesh.F
    pos4 = np.concatenate([np.asarray(mesh.pos, np.float32), np.ones((V, 1), np.float32)], axis=1)
    parts = [pos4.tobytes()]
    flags = 0
    if mesh.nrm is not None:
        parts.append(np.concatenate([np.asarray(mesh.nrm, np.float32), np.zeros((V, 1), np.float32)], axis=1).tobytes())
    else:
        flags |= _HAS_NO_NORMALS
    tc = mesh.tc if mesh.tc is not None else np.zeros((V, 2), np.float32)
    parts.append(np.asarray(tc, np.float32).tobytes())
    parts.append(np.asarray(mesh.idx, np.uint32).tobytes())
    parts.append(np.zeros(F, np.uint32).tobytes())
    body = b"".join(parts)
    with open(path, "wb") as f:
        f.write(_HEADER.pack(_HEADER.size + len(body), V, 3 * F, F, flags))
        f.write(body)


def tonemap(img: np.ndarray, gamma: float = 2.2) -> np.ndarray:
    """(H, W, 3) float -> uint8 with the reference's tone map."""
    ginv = np.float32(1.0) / np.float32(gamma)             # 1.0f / gamma, in float as the reference
    with np.errstate(invalid="ignore"):                   # negative inputs give NaN, mapped to 0
        x = np.power(np.asarray(img, np.float32), ginv) * np.float32(255.0) + np.float32(0.5)
    x = np.nan_to_num(x, nan=0.0, posinf=255.0, neginf=0.0)
    return np.clip(np.trunc(x), 0, 255).astype(np.uint8)


def save_image(path: str, img: np.ndarray, gamma: float = 2.2) -> None:
    """Write an (H, W, 3) float image as .bmp, .ppm or .image3f (by extension)."""
    img = np.asarray(img, np.float32)
    if img.ndim != 3 or img.shape[2] != 3:
        raise ValueError("expected an (H, W, 3) image")
    H, W, _ = img.shape
    ext = path[path.rfind("."):].lower()
    if ext == ".image3f":
        with open(path, "wb") as f:
            f.write(struct.pack("<II", W, H))
            f.write(np.ascontiguousarray(img).tobytes())
        return
    px = tonemap(img, gamma)
    if ext == ".ppm":
        with open(path, "w") as f:
            f.write(f"P3\n{W} {H} 255\n")
            f.write(" ".join(str(int(v)) for v in px.reshape(-1)))
            f.write(" ")
        return
    if ext == ".png":
        rgba = np.empty((H, W, 4), np.uint8)
        rgba[:, :, :3] = px
        rgba[:, :, 3] = 255
        with open(path, "wb") as f:
            f.write(png_bytes(rgba))
        return
    if ext == ".bmp":
        row = 3 * W
        pad = (-row) % 4
        bgr = px[::-1, :, ::-1]                       # bottom-up rows, B G R bytes
        rows = np.zeros((H, row + pad), np.uint8)
        rows[:, :row] = bgr.reshape(H, row)
        data = rows.tobytes()
        header = struct.pack("<2sIHHI", b"BM", 54 + len(data), 0, 0, 54)
        info = struct.pack("<IiiHHIIiiII", 40, W, H, 1, 24, 0, len(data), 2835, 2835, 0, 0)
        with open(path, "wb") as f:
            f.write(header + info + data)
        return
    raise ValueError(f"unsupported image extension {ext!r} (.bmp, .ppm, .png, .image3f)")


_PNG_SIG = b"\x89PNG\r\n\x1a\n"


def _png_chunk(kind: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def png_bytes(px: np.ndarray) -> bytes:
    """An (H, W, 3 or 4) uint8 array as a PNG file (8-bit RGB or RGBA, no interlace): every
    scanline with filter type 0, the whole image in one zlib-compressed IDAT chunk."""
    px = np.ascontiguousarray(px, np.uint8)
    if px.ndim != 3 or px.shape[2] not in (3, 4):
        raise ValueError("expected an (H, W, 3) or (H, W, 4) uint8 array")
    H, W, C = px.shape
    raw = np.zeros((H, 1 + W * C), np.uint8)           # filter byte 0, then the row
    raw[:, 1:] = px.reshape(H, W * C)
    ihdr = struct.pack(">IIBBBBB", W, H, 8, 6 if C == 4 else 2, 0, 0, 0)
    return (_PNG_SIG + _png_chunk(b"IHDR", ihdr) + _png_chunk(b"IDAT", zlib.compress(raw.tobytes(), 6))
            + _png_chunk(b"IEND", b""))


def load_png(path: str) -> np.ndarray:
    """Read an 8-bit, non-interlaced RGB / RGBA / grey PNG as an (H, W, C) uint8 array, undoing
    each scanline's filter (none, sub, up, average, Paeth; PNG specification section 9)."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != _PNG_SIG:
        raise ValueError(f"{path}: not a PNG")
    off, idat, hdr = 8, [], None
    while off + 8 <= len(data):
        n, kind = struct.unpack_from(">I4s", data, off)
        body = data[off + 8: off + 8 + n]
        if len(body) != n or struct.unpack_from(">I", data, off + 8 + n)[0] != zlib.crc32(kind + body) & 0xFFFFFFFF:
            raise ValueError(f"{path}: bad {kind!r} chunk")
        if kind == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat.append(body)
        elif kind == b"IEND":
            break
        off += 12 + n
    if hdr is None:
        raise ValueError(f"{path}: no IHDR")
    W, H, depth, ctype, _, _, interlace = hdr
    C = {0: 1, 2: 3, 4: 2, 6: 4}.get(ctype)
    if depth != 8 or C is None or interlace:
        raise ValueError(f"{path}: only 8-bit non-interlaced grey / RGB / RGBA PNG is read")
    raw = np.frombuffer(zlib.decompress(b"".join(idat)), np.uint8)
    stride = W * C
    if raw.size != H * (stride + 1):
        raise ValueError(f"{path}: image data size")
    rows = raw.reshape(H, stride + 1)
    out = np.zeros((H, stride), np.uint8)
    prev = np.zeros(stride, np.int32)
    for y in range(H):
        ft, line = int(rows[y, 0]), rows[y, 1:].astype(np.int32)
        cur = np.zeros(stride, np.int32)
        if ft == 0:
            cur = line
        elif ft == 2:
            cur = (line + prev) & 255
        else:
            for x in range(stride):
                a = cur[x - C] if x >= C else 0
                b = int(prev[x])
                c = int(prev[x - C]) if x >= C else 0
                if ft == 1:
                    p = a
                elif ft == 3:
                    p = (a + b) >> 1
                elif ft == 4:
                    pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                    p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
                else:
                    raise ValueError(f"{path}: filter type {ft}")
                cur[x] = (int(line[x]) + p) & 255
        out[y] = cur
        prev = cur
    return out.reshape(H, W, C)


def load_bmp(path: str) -> np.ndarray:
    """Read a 24-bit uncompressed BMP as an (H, W, 3) uint8 array, row 0 at the top."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:2] != b"BM":
        raise ValueError(f"{path}: not a BMP")
    off = struct.unpack_from("<I", data, 10)[0]
    W, H = struct.unpack_from("<ii", data, 18)
    bpp = struct.unpack_from("<H", data, 28)[0]
    if bpp != 24:
        raise ValueError(f"{path}: {bpp}-bit BMP (24-bit only)")
    row = 3 * W + ((-3 * W) % 4)
    rows = np.frombuffer(data, np.uint8, count=row * abs(H), offset=off).reshape(abs(H), row)[:, :3 * W]
    img = rows.reshape(abs(H), W, 3)[:, :, ::-1]
    return np.ascontiguousarray(img[::-1] if H > 0 else img)


def load_image3f(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        W, H = struct.unpack("<II", f.read(8))
        return np.frombuffer(f.read(), np.float32, count=W * H * 3).reshape(H, W, 3).copy()
