"""Host-side mesh and image files (diffrender_tutorials_amd.meshio) against what the
reference's own writers wrote and its reader read back (tests/golden/meshio_files.npz, made
from common_src/cmesh.cpp:139-205 and common_src/Image2d.cpp:486-735 compiled from source).

Pinned there: the VSGF reader on a file the reference wrote (tangent block present), the
reference's reader on this package's writer (with and without normals), the tone map and the
PPM text byte for byte, the raw float dump byte for byte, the BMP pixel bytes and the PNG
pixels (decoded from the file the reference's stb writer produced).  The BMP
header differs from the reference's in the fields the reference leaves wrong or empty (file
size, image size, resolution); rows of widths that are not a multiple of 4 are padded here
and not in the reference's writer."""
from __future__ import annotations

import hashlib
import os
import struct

import numpy as np
import pytest

from diffrender_tutorials_amd import meshio, scenes as S

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "meshio_files.npz"))


def _mesh():
    return S.Mesh(pos=GOLD["pos4"][:, :3].copy(), idx=GOLD["idx"].copy(), tc=GOLD["tc"].copy(),
                  nrm=GOLD["nrm4"][:, :3].copy())


def test_reads_a_file_the_reference_wrote(tmp_path):
    f = tmp_path / "ref.vsgf"
    f.write_bytes(GOLD["ref_vsgf"].tobytes())
    size, nv, ni, nm, flags = struct.unpack_from("<QIIII", f.read_bytes())
    assert flags & 1 and size == f.stat().st_size          # the reference writes zero tangents
    m = meshio.load_vsgf(str(f))
    want = _mesh()
    assert (m.V, m.F) == (nv, ni // 3) == (want.V, want.F)
    for a, b in ((m.pos, want.pos), (m.nrm, want.nrm), (m.tc, want.tc), (m.idx, want.idx)):
        assert a.dtype == b.dtype and np.array_equal(a, b)


@pytest.mark.parametrize("tag,normals", [("ours", True), ("ours_nonrm", False)])
def test_the_reference_reads_what_we_write(tmp_path, tag, normals):
    m = _mesh()
    if not normals:
        m.nrm = None
    f = tmp_path / "m.vsgf"
    meshio.save_vsgf(str(f), m)
    data = f.read_bytes()
    # the reference's reader was run on exactly these bytes
    assert hashlib.sha256(data).hexdigest() == str(GOLD[tag + "_vsgf_sha"])
    assert np.array_equal(GOLD[tag + "_ref_pos4"][:, :3], m.pos)
    assert np.array_equal(GOLD[tag + "_ref_pos4"][:, 3], np.ones(m.V, np.float32))
    assert np.array_equal(GOLD[tag + "_ref_tc"], m.tc)
    assert np.array_equal(GOLD[tag + "_ref_idx"], m.idx)
    assert not GOLD[tag + "_ref_mat"].any()
    if normals:
        assert np.array_equal(GOLD[tag + "_ref_nrm4"][:, :3], m.nrm)
    else:
        assert not GOLD[tag + "_ref_nrm4"].any()            # the reference zero-fills absent normals
    size, nv, ni, nm, flags = struct.unpack_from("<QIIII", data)
    assert size == len(data) and (nv, ni, nm) == (m.V, 3 * m.F, m.F)
    assert flags == (0 if normals else 8)
    back = meshio.load_vsgf(str(f))
    assert np.array_equal(back.pos, m.pos) and np.array_equal(back.idx, m.idx)
    assert (back.nrm is None) == (not normals)


def test_vsgf_layout_by_hand(tmp_path):
    """One triangle: header, float4 positions (w = 1), float4 normals (w = 0), float2 texture
    coordinates, uint32 indices, one uint32 material per triangle."""
    m = S.Mesh(pos=np.array([[1, 2, 3], [4, 5, 6], [7, 8, 9]], np.float32), idx=np.array([[0, 2, 1]], np.uint32),
               tc=np.array([[0.25, 0.5], [0.75, 1], [0, 0]], np.float32),
               nrm=np.array([[0, 0, 1], [0, 1, 0], [1, 0, 0]], np.float32))
    f = tmp_path / "t.vsgf"
    meshio.save_vsgf(str(f), m)
    d = f.read_bytes()
    assert len(d) == 24 + 3 * 16 + 3 * 16 + 3 * 8 + 3 * 4 + 4
    assert struct.unpack_from("<QIIII", d) == (len(d), 3, 3, 1, 0)
    assert struct.unpack_from("<4f", d, 24) == (1, 2, 3, 1)
    assert struct.unpack_from("<4f", d, 24 + 48) == (0, 0, 1, 0)
    assert struct.unpack_from("<2f", d, 24 + 96 + 8) == (0.75, 1)
    assert struct.unpack_from("<3I", d, 24 + 96 + 24) == (0, 2, 1)


def test_vsgf_reader_rejects_bad_files(tmp_path):
    m = _mesh()
    f = tmp_path / "m.vsgf"
    meshio.save_vsgf(str(f), m)
    good = f.read_bytes()
    for bad in (good[:20], good[:24 + 16 * m.V]):
        f.write_bytes(bad)
        with pytest.raises(ValueError):
            meshio.load_vsgf(str(f))
    idx = good[:]
    off = 24 + (16 + 16 + 8) * m.V
    f.write_bytes(idx[:off] + struct.pack("<I", m.V) + idx[off + 4:])
    with pytest.raises(ValueError, match="out of range"):
        meshio.load_vsgf(str(f))


def test_tonemap_closed_form():
    img = np.array([[[0.0, 1.0, 4.0], [0.5, 1e-7, -1.0]]], np.float32)
    px = meshio.tonemap(img)
    assert px.dtype == np.uint8
    assert px[0, 0].tolist() == [0, 255, 255]
    assert px[0, 1, 0] == int(0.5 ** (1 / 2.2) * 255 + 0.5)       # 186
    assert px[0, 1, 1] == int((1e-7) ** (1 / 2.2) * 255 + 0.5)      # 0
    assert px[0, 1, 2] == 0
    assert meshio.tonemap(np.full((1, 1, 3), 0.25, np.float32), gamma=1.0)[0, 0, 0] == 64


@pytest.mark.parametrize("tag", ["img8", "img6"])
def test_ppm_and_raw_match_the_reference_byte_for_byte(tmp_path, tag):
    img = GOLD[tag]
    for ext, gamma, key in ((".ppm", 2.2, "_ref_ppm"), (".ppm", 1.0, "_ref_ppm_g1"), (".image3f", 2.2, "_ref_image3f")):
        f = tmp_path / ("x" + ext)
        meshio.save_image(str(f), img, gamma)
        assert f.read_bytes() == GOLD[tag + key].tobytes(), (tag, ext, gamma)
    assert np.array_equal(meshio.load_image3f(str(tmp_path / "x.image3f")), img)


def test_bmp_pixels_match_the_reference(tmp_path):
    img = GOLD["img8"]                       # width 8: rows need no padding in either writer
    ref = GOLD["img8_ref_bmp"].tobytes()
    f = tmp_path / "x.bmp"
    meshio.save_image(str(f), img)
    ours = f.read_bytes()
    assert len(ours) == len(ref) == 54 + img.size
    assert ours[54:] == ref[54:]
    assert ours[:2] == ref[:2] == b"BM"
    assert ours[10:34] == ref[10:34]         # data offset, info size, width, height, planes, bpp, no compression
    assert struct.unpack_from("<I", ours, 2)[0] == len(ours)
    # the reference's own BMP reader returns 0xRRGGBB per pixel, rows bottom-up; ours reads the same file
    f.write_bytes(ref)
    assert np.array_equal(meshio.load_bmp(str(f)), meshio.tonemap(img))


def test_bmp_rows_are_padded_for_any_width(tmp_path):
    img = GOLD["img6"]                       # width 6: 18-byte rows padded to 20
    ref = GOLD["img6_ref_bmp"].tobytes()     # the reference writes the rows unpadded
    f = tmp_path / "x.bmp"
    meshio.save_image(str(f), img)
    ours = f.read_bytes()
    H, W, _ = img.shape
    assert len(ours) == 54 + H * 20
    rows = np.frombuffer(ours, np.uint8, offset=54).reshape(H, 20)
    assert not rows[:, 18:].any()
    assert rows[:, :18].tobytes() == ref[54:]
    assert np.array_equal(meshio.load_bmp(str(f)), meshio.tonemap(img))


@pytest.mark.parametrize("tag", ["img8", "img6"])
def test_png_pixels_match_the_reference(tmp_path, tag):
    """The reference writes .png through stb (common_src/Image2d.cpp:629-646): tone-mapped RGBA
    with alpha 255, rows top-down.  Its file, decoded, holds exactly the pixels of ours; the
    compressed streams differ (zlib here, stb's own deflate there)."""
    img = GOLD[tag]
    ref = tmp_path / "ref.png"
    ref.write_bytes(GOLD[tag + "_ref_png"].tobytes())
    f = tmp_path / "x.png"
    meshio.save_image(str(f), img)
    ours, theirs = meshio.load_png(str(f)), meshio.load_png(str(ref))
    H, W, _ = img.shape
    assert ours.shape == theirs.shape == (H, W, 4)
    assert np.array_equal(ours, theirs)
    assert np.array_equal(ours[:, :, :3], meshio.tonemap(img)) and (ours[:, :, 3] == 255).all()


def test_png_reader_undoes_every_filter_type(tmp_path):
    """Rows filtered by hand with each of the five filter types (PNG spec section 9) decode to
    the original pixels; a corrupted CRC is refused."""
    rng = np.random.default_rng(3)
    px = rng.integers(0, 256, (5, 7, 3), dtype=np.uint8)
    H, W, C = px.shape
    stride = W * C
    rows = px.reshape(H, stride).astype(np.int32)
    raw = bytearray()
    for y in range(H):
        prev = rows[y - 1] if y else np.zeros(stride, np.int32)
        cur = rows[y]
        left = np.concatenate([np.zeros(C, np.int32), cur[:-C]])
        ul = np.concatenate([np.zeros(C, np.int32), prev[:-C]])
        ft = y % 5
        if ft == 0:
            f = cur
        elif ft == 1:
            f = cur - left
        elif ft == 2:
            f = cur - prev
        elif ft == 3:
            f = cur - (left + prev) // 2
        else:
            pa, pb, pc = np.abs(prev - ul), np.abs(left - ul), np.abs(left + prev - 2 * ul)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, prev, ul))
            f = cur - pred
        raw += bytes([ft]) + (f & 255).astype(np.uint8).tobytes()
    import zlib
    body = (meshio._PNG_SIG + meshio._png_chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, 8, 2, 0, 0, 0))
            + meshio._png_chunk(b"IDAT", zlib.compress(bytes(raw))) + meshio._png_chunk(b"IEND", b""))
    f = tmp_path / "f.png"
    f.write_bytes(body)
    assert np.array_equal(meshio.load_png(str(f)), px)
    bad = bytearray(body)
    bad[29] ^= 1                                          # inside IHDR: its CRC no longer holds
    f.write_bytes(bytes(bad))
    with pytest.raises(ValueError):
        meshio.load_png(str(f))


def test_save_image_rejects_bad_input(tmp_path):
    with pytest.raises(ValueError):
        meshio.save_image(str(tmp_path / "x.jpg"), np.zeros((2, 2, 3), np.float32))
    with pytest.raises(ValueError):
        meshio.save_image(str(tmp_path / "x.bmp"), np.zeros((2, 2), np.float32))
