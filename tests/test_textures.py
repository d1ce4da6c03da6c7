"""Textures on the host: the image readers against images written here from known pixels (PNG through Python's zlib
at every colour type, bit depth, filter and compression mode, Adam7; TGA; PPM / PGM), the zlib decoder, the sRGB
table and the texture lookup against closed forms and the oracle's restatement, tangent frames against the oracle
and their closed forms, OBJ / MTL texture statements, and texture closed forms of the oracle's renderer. The file
formats follow their specifications (PNG: RFC 2083 / ISO 15948, zlib / deflate: RFC 1950 / 1951); the semantics
of the maps follow the reference (src/scene/texture.h:15-25, src/render/wavefront.cu:252-271)."""
import math
import struct
import zlib

import numpy as np
import pytest

import cwp
import pyoracle as po

C2W = "1,0,0,0,0,0.7071068,-0.7071068,-2,0,0.7071068,0.7071068,1,0,0,0,1"


# ------------------------------------------------------------------------------------------------ image writers
def png_chunk(t, data):
    return struct.pack(">I", len(data)) + t + data + struct.pack(">I", zlib.crc32(t + data) & 0xFFFFFFFF)


def filter_row(ftype, row, prev, bpp):
    out = bytearray(len(row))
    for i in range(len(row)):
        a = row[i - bpp] if i >= bpp else 0
        b = prev[i]
        c = prev[i - bpp] if i >= bpp else 0
        if ftype == 0:
            p = 0
        elif ftype == 1:
            p = a
        elif ftype == 2:
            p = b
        elif ftype == 3:
            p = (a + b) >> 1
        else:
            pp = a + b - c
            pa, pb, pc = abs(pp - a), abs(pp - b), abs(pp - c)
            p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
        out[i] = (row[i] - p) & 255
    return bytes([ftype]) + bytes(out)


def pack_row(samples, depth):
    """samples: flat list of integer samples of one row"""
    if depth == 8:
        return bytes(samples)
    if depth == 16:
        return b"".join(struct.pack(">H", s) for s in samples)
    bits, out, acc, n = depth, bytearray(), 0, 0
    for s in samples:
        acc = (acc << bits) | s
        n += bits
        if n == 8:
            out.append(acc); acc, n = 0, 0
    if n:
        out.append(acc << (8 - n))
    return bytes(out)


def write_png(path, samples, ctype, depth, palette=None, trns=None, interlace=False, level=6, strategy=None,
              split_idat=1, filters=(0, 1, 2, 3, 4)):
    """samples: (h, w, channels) integer array of raw sample values"""
    h, w, ch = samples.shape
    bpp = max(1, ch * depth // 8)
    raw = bytearray()
    passes = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)] \
        if interlace else [(0, 0, 1, 1)]
    k = 0
    for xs, ys, dx, dy in passes:
        sub = samples[ys::dy, xs::dx]
        if sub.size == 0:
            continue
        prev = bytes(len(pack_row(list(sub[0].reshape(-1)), depth)))
        for y in range(sub.shape[0]):
            row = pack_row([int(x) for x in sub[y].reshape(-1)], depth)
            raw += filter_row(filters[k % len(filters)], row, prev, bpp)
            prev = row
            k += 1
    if strategy is None:
        comp = zlib.compress(bytes(raw), level)
    else:
        c = zlib.compressobj(level, zlib.DEFLATED, 15, 9, strategy)
        comp = c.compress(bytes(raw)) + c.flush()
    out = b"\x89PNG\r\n\x1a\n" + png_chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 1 if interlace else 0))
    out += png_chunk(b"tEXt", b"Comment\x00ancillary chunk")
    if palette is not None:
        out += png_chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    if trns is not None:
        out += png_chunk(b"tRNS", trns)
    step = max(1, len(comp) // split_idat + 1)
    for i in range(0, len(comp), step):
        out += png_chunk(b"IDAT", comp[i:i + step])
    out += png_chunk(b"IEND", b"")
    path.write_bytes(out)


def rgba_words(r, g, b, a):
    return (np.asarray(r, np.uint32) | (np.asarray(g, np.uint32) << 8) | (np.asarray(b, np.uint32) << 16) |
            (np.asarray(a, np.uint32) << 24))


def scale8(v, depth):
    v = np.asarray(v, np.int64)
    return v >> 8 if depth == 16 else (v if depth == 8 else v * 255 // ((1 << depth) - 1))


# ------------------------------------------------------------------------------------------------ readers
@pytest.mark.parametrize("ctype,depth", [(0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (3, 1), (3, 2),
                                         (3, 4), (3, 8), (4, 8), (4, 16), (6, 8), (6, 16)])
@pytest.mark.parametrize("interlace", [False, True])
def test_png_colour_types_depths_filters(tmp_path, ctype, depth, interlace):
    rng = np.random.default_rng(ctype * 100 + depth + 7 * interlace)
    w, h = 13, 11   # odd sizes: partial bytes, empty Adam7 passes
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    maxv = (1 << depth) - 1
    pal = None
    if ctype == 3:
        npal = min(1 << depth, 40)
        pal = rng.integers(0, 256, (npal, 3))
        s = rng.integers(0, npal, (h, w, 1))
    else:
        s = rng.integers(0, maxv + 1, (h, w, ch))
    p = tmp_path / "a.png"
    write_png(p, s, ctype, depth, palette=pal, interlace=interlace)
    got = cwp.image_read(p)
    if ctype == 3:
        want = rgba_words(pal[s[..., 0], 0], pal[s[..., 0], 1], pal[s[..., 0], 2], 255)
    elif ctype in (0, 4):
        g = scale8(s[..., 0], depth)
        want = rgba_words(g, g, g, scale8(s[..., 1], depth) if ctype == 4 else 255)
    else:
        want = rgba_words(scale8(s[..., 0], depth), scale8(s[..., 1], depth), scale8(s[..., 2], depth),
                          scale8(s[..., 3], depth) if ctype == 6 else 255)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("level,strategy,split", [(0, None, 1), (1, None, 3), (9, None, 1), (6, zlib.Z_FIXED, 2),
                                                  (6, zlib.Z_HUFFMAN_ONLY, 1), (6, zlib.Z_RLE, 5)])
def test_png_compression_modes_and_idat_splits(tmp_path, level, strategy, split):
    y, x = np.mgrid[0:64, 0:96]
    s = np.stack([(x * 3) % 256, (y * 5) % 256, (x ^ y) % 256], -1)   # compressible, long matches
    p = tmp_path / "c.png"
    write_png(p, s, 2, 8, level=level, strategy=strategy, split_idat=split)
    np.testing.assert_array_equal(cwp.image_read(p), rgba_words(s[..., 0], s[..., 1], s[..., 2], 255))


def test_png_transparency_chunks(tmp_path):
    s = np.array([[[0], [1], [2], [3]]])
    write_png(tmp_path / "p.png", s, 3, 2, palette=[[10, 20, 30], [40, 50, 60], [70, 80, 90], [1, 2, 3]], trns=bytes([0, 128]))
    np.testing.assert_array_equal(cwp.image_read(tmp_path / "p.png")[0] >> 24, [0, 128, 255, 255])
    g = np.array([[[5], [7]]])
    write_png(tmp_path / "g.png", g, 0, 8, trns=struct.pack(">H", 7))
    np.testing.assert_array_equal(cwp.image_read(tmp_path / "g.png")[0] >> 24, [255, 0])
    c = np.array([[[1, 2, 3], [1, 2, 4]]])
    write_png(tmp_path / "r.png", c, 2, 16, trns=struct.pack(">HHH", 1, 2, 3))
    np.testing.assert_array_equal(cwp.image_read(tmp_path / "r.png")[0] >> 24, [0, 255])


def test_png_damaged_files_are_rejected(tmp_path):
    s = np.random.default_rng(3).integers(0, 256, (8, 8, 3))
    p = tmp_path / "d.png"
    write_png(p, s, 2, 8)
    good = p.read_bytes()
    i = good.index(b"IDAT") + 10
    bad = bytearray(good); bad[i] ^= 0x40
    p.write_bytes(bytes(bad))
    with pytest.raises(cwp.CwpError, match="CRC"):
        cwp.image_read(p)
    p.write_bytes(good[:len(good) // 2])
    with pytest.raises(cwp.CwpError):
        cwp.image_read(p)
    # a valid chunk CRC over a damaged zlib stream: the Adler-32 check catches it
    comp = bytearray(zlib.compress(bytes(filter_row(0, bytes(s[0].astype(np.uint8).reshape(-1)), bytes(24), 3)) * 8, 0))
    comp[-5] ^= 1
    body = b"\x89PNG\r\n\x1a\n" + png_chunk(b"IHDR", struct.pack(">IIBBBBB", 8, 8, 8, 2, 0, 0, 0)) + \
        png_chunk(b"IDAT", bytes(comp)) + png_chunk(b"IEND", b"")
    p.write_bytes(body)
    with pytest.raises(cwp.CwpError, match="checksum"):
        cwp.image_read(p)


def test_inflate_matches_zlib():
    rng = np.random.default_rng(5)
    blobs = [b"", b"a", bytes(rng.integers(0, 256, 70000, dtype=np.uint8)), b"abc" * 30000,
             bytes(rng.integers(0, 4, 100000, dtype=np.uint8))]
    for data in blobs:
        for level in (0, 1, 6, 9):
            for strategy in (zlib.Z_DEFAULT_STRATEGY, zlib.Z_FIXED, zlib.Z_RLE):
                c = zlib.compressobj(level, zlib.DEFLATED, 15, 9, strategy)
                comp = c.compress(data) + c.flush()
                assert cwp.inflate(comp, len(data)) == data
    comp = zlib.compress(b"x" * 1000)
    with pytest.raises(cwp.CwpError, match="more data"):
        cwp.inflate(comp, 999)
    with pytest.raises(cwp.CwpError):
        cwp.inflate(comp[:-6], 1000)
    with pytest.raises(cwp.CwpError, match="checksum"):
        cwp.inflate(comp[:-1] + bytes([comp[-1] ^ 1]), 1000)


def write_tga(path, rgba, bits, rle=False, top=False, grey=False):
    h, w = rgba.shape[:2]
    hdr = struct.pack("<BBBHHBHHHHBB", 0, 0, (11 if rle else 3) if grey else (10 if rle else 2), 0, 0, 0, 0, 0, w, h,
                      8 if grey else bits, (0x20 if top else 0) | (8 if bits == 32 else 0))
    rows = rgba if top else rgba[::-1]
    px = []
    for row in rows:
        for p in row:
            px.append(bytes([p[0]]) if grey else (bytes([p[2], p[1], p[0], p[3]]) if bits == 32 else bytes([p[2], p[1], p[0]])))
    if not rle:
        body = b"".join(px)
    else:   # runs of equal pixels, raw packets otherwise
        body, i = bytearray(), 0
        while i < len(px):
            j = i
            while j + 1 < len(px) and px[j + 1] == px[i] and j - i < 127:
                j += 1
            if j > i:
                body += bytes([0x80 | (j - i)]) + px[i]; i = j + 1
            else:
                k = i
                while k + 1 < len(px) and px[k + 1] != px[k] and k - i < 127:
                    k += 1
                body += bytes([k - i]) + b"".join(px[i:k + 1]); i = k + 1
    path.write_bytes(hdr + bytes(body))


@pytest.mark.parametrize("bits,rle,top,grey", [(24, False, False, False), (32, False, True, False), (24, True, True, False),
                                               (32, True, False, False), (8, False, False, True), (8, True, True, True)])
def test_tga(tmp_path, bits, rle, top, grey):
    rng = np.random.default_rng(bits + 2 * rle + 4 * top)
    rgba = rng.integers(0, 256, (9, 14, 4))
    rgba[2:5, 3:11] = rgba[2, 3]   # runs
    if grey:
        rgba[..., 1] = rgba[..., 2] = rgba[..., 0]
    if bits != 32:
        rgba[..., 3] = 255
    p = tmp_path / "t.tga"
    write_tga(p, rgba, bits, rle, top, grey)
    np.testing.assert_array_equal(cwp.image_read(p), rgba_words(rgba[..., 0], rgba[..., 1], rgba[..., 2], rgba[..., 3]))


def test_ppm_pgm(tmp_path):
    rng = np.random.default_rng(9)
    c = rng.integers(0, 256, (5, 7, 3))
    (tmp_path / "a.ppm").write_bytes(b"P6\n# comment\n7 5\n255\n" + bytes(c.astype(np.uint8).reshape(-1)))
    np.testing.assert_array_equal(cwp.image_read(tmp_path / "a.ppm"), rgba_words(c[..., 0], c[..., 1], c[..., 2], 255))
    (tmp_path / "b.ppm").write_text("P3 7 5 255\n" + " ".join(str(int(x)) for x in c.reshape(-1)) + "\n")
    np.testing.assert_array_equal(cwp.image_read(tmp_path / "b.ppm"), rgba_words(c[..., 0], c[..., 1], c[..., 2], 255))
    g = rng.integers(0, 65536, (4, 6))
    (tmp_path / "c.pgm").write_bytes(b"P5 6 4 65535\n" + b"".join(struct.pack(">H", int(x)) for x in g.reshape(-1)))
    gg = (g * 255 + 32767) // 65535
    np.testing.assert_array_equal(cwp.image_read(tmp_path / "c.pgm"), rgba_words(gg, gg, gg, 255))
    (tmp_path / "d.pgm").write_text("P2 2 1 15\n0 15\n")
    np.testing.assert_array_equal(cwp.image_read(tmp_path / "d.pgm"), rgba_words([[0, 255]], [[0, 255]], [[0, 255]], 255))
    with pytest.raises(cwp.CwpError, match="unsupported"):
        (tmp_path / "x.bmp").write_bytes(b"BM0000")
        cwp.image_read(tmp_path / "x.bmp")


# ------------------------------------------------------------------------------------------------ lookup
def test_srgb_table_closed_forms():
    t = po.srgb_table()
    assert t[0] == 0.0 and t[255] == 1.0
    assert t[10] == np.float32(10 / 255 / 12.92)            # linear segment
    assert abs(t[128] - ((128 / 255 + 0.055) / 1.055) ** 2.4) < 1e-7
    assert (np.diff(t) > 0).all()


def lookup_both(img, texels, u, v, nearest):
    a = cwp.texture_sample(img, texels, u, v, nearest)
    b = po.texture_sample(img, texels, u, v, nearest)
    np.testing.assert_array_equal(a, b)
    return a


def test_lookup_closed_forms():
    w, h = 4, 2
    vals = np.arange(w * h, dtype=np.uint32) * 30   # r = 0, 30, ..., 210: texel (x, y) = 30 (y w + x)
    tex = rgba_words(vals, vals, vals, 255)
    img = (0, w, h, 0)
    # nearest: texel floor(u w), row floor((1 - v) h) (rows top first)
    for x in range(w):
        for y in range(h):
            u, v = (x + 0.25) / w, 1 - (y + 0.75) / h
            assert lookup_both(img, tex, u, v, True)[0] == np.float32(vals[y * w + x]) * np.float32(1 / 255)
    # bilinear at a texel centre is the texel; halfway between two neighbours is their mean (to rounding)
    c = lookup_both(img, tex, 1.5 / w, 1 - 0.5 / h, False)
    assert c[0] == np.float32(30) * np.float32(1 / 255)
    m = lookup_both(img, tex, 2.0 / w, 1 - 0.5 / h, False)
    assert abs(m[0] - (30 + 60) / 2 / 255) < 1e-6
    # wrap: u + 1, v - 2 give the same value; the left border blends with the right column
    for u, v in ((0.3125, 0.1875), (-0.6875, 1.25), (0.015625, 0.984375)):   # dyadic: u + 1, v - 2 are exact
        np.testing.assert_array_equal(lookup_both(img, tex, u, v, False), lookup_both(img, tex, u + 1, v - 2, False))
    edge = lookup_both(img, tex, 0.0, 1 - 0.5 / h, False)
    assert abs(edge[0] - (0 + 90) / 2 / 255) < 1e-6
    # NaN / infinite coordinates read texel (0, 0) area, never outside
    for u in (np.nan, np.inf, -np.inf):
        assert np.isfinite(lookup_both(img, tex, u, 0.5, False)).all()


def test_lookup_srgb_and_random_parity():
    rng = np.random.default_rng(11)
    w, h = 7, 5
    tex = rng.integers(0, 2 ** 32, w * h * 2, dtype=np.uint64).astype(np.uint32)
    lin, srgb = (0, w, h, 0), (w * h, w, h, 1)
    t = po.srgb_table()
    c = lookup_both(srgb, tex, 0.5 / w, 1 - 0.5 / h, True)
    px = int(tex[w * h])
    np.testing.assert_array_equal(c, [t[px & 255], t[(px >> 8) & 255], t[(px >> 16) & 255], np.float32(px >> 24) * np.float32(1 / 255)])
    for k in range(2000):
        u, v = rng.uniform(-4, 4, 2).astype(np.float32)
        lookup_both(lin if k & 1 else srgb, tex, u, v, bool(k & 2))


# ------------------------------------------------------------------------------------------------ tangents
def test_tangents_match_oracle_and_closed_forms(tmp_path):
    s = cwp.Setup(["--gen", "terrain", "--gen_grid", "30", "--gen_boxes", "12", "--gen_textures", "16", "--width", "32",
                   "--height", "32"])
    order = s.array(cwp.ARR_ORDER)
    n = s.array(cwp.ARR_NORMALS) if s.params().has_normals else None
    ot = po.corner_tangents(s.array(cwp.ARR_VERTICES), n, s.array(cwp.ARR_TRIANGLES), s.array(cwp.ARR_MESH_UVS))
    pt = s.array(cwp.ARR_TANGENTS)
    np.testing.assert_array_equal(pt, ot[order])
    assert np.allclose(np.linalg.norm(pt[..., :3], axis=-1), 1, atol=1e-6)
    assert set(np.unique(pt[..., 3])) <= {-1.0, 1.0}
    # a quad in the z = 0 plane with u = x, v = y: tangent +x, handedness +1; with u mirrored: tangent -x, ... w = -1
    for mirror, tx, w in ((False, 1.0, 1.0), (True, -1.0, -1.0)):
        (tmp_path / "q.mtl").write_text("newmtl m\nKd 1 1 1\nmap_Kd t.ppm\nnorm t.ppm\n")
        (tmp_path / "t.ppm").write_bytes(b"P6 1 1 255\n" + bytes([128, 128, 255]))
        us = [1, 0, 0, 1] if mirror else [0, 1, 1, 0]
        (tmp_path / "q.obj").write_text("mtllib q.mtl\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\n" +
                                        "".join("vt %d %d\n" % (us[i], [0, 0, 1, 1][i]) for i in range(4)) +
                                        "usemtl m\nf 1/1 2/2 3/3\nf 1/1 3/3 4/4\n")
        q = cwp.Setup(["--mesh", str(tmp_path / "q.obj"), "--hfov", "40", "--c2w", C2W])
        tg = q.array(cwp.ARR_TANGENTS)
        np.testing.assert_allclose(tg[..., 0], tx, atol=1e-6)
        np.testing.assert_allclose(tg[..., 1:3], 0, atol=1e-6)
        np.testing.assert_array_equal(tg[..., 3], w)


# ------------------------------------------------------------------------------------------------ OBJ / MTL
def test_obj_mtl_texture_statements(tmp_path):
    (tmp_path / "img").mkdir()
    for name, px in (("base.ppm", [200, 100, 50]), ("rough.ppm", [0, 64, 0]), ("n.ppm", [128, 128, 255])):
        (tmp_path / "img" / name).write_bytes(b"P6 1 1 255\n" + bytes(px))
    (tmp_path / "m.mtl").write_text(
        "newmtl plain\nKd 0.5 0.5 0.5\n"
        "newmtl tex\nKd 1 1 1\nmap_Kd -s 1 1 1 -o 0 0 img/base.ppm\nmap_Pr -imfchan g img/rough.ppm\n"
        "map_Bump -bm 0.5 img/n.ppm\nmap_Ke img/missing.png\n")
    (tmp_path / "m.obj").write_text(
        "mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n"
        "usemtl plain\nf 1 2 3\nusemtl tex\nf 1/1 3/-2 4/-1\n")
    s = cwp.Setup(["--mesh", str(tmp_path / "m.obj"), "--hfov", "40", "--c2w", C2W])
    i = s.info
    assert i["textured"] == 1 and i["full"] == 1 and i["num_images"] == 3
    maps = s.array(cwp.ARR_MAPS)   # material 0: the OBJ default, 1: plain, 2: tex
    assert tuple(maps[1][:5]) == (-1, -1, -1, -1, -1)
    base, emis, nrm, metal, rough, mch, rch, scale = maps[2]
    assert emis == -1 and metal == -1 and min(base, nrm, rough) >= 0 and rch == 1
    assert np.float32(np.int32(scale).view(np.float32)) == np.float32(0.5)
    ims = s.array(cwp.ARR_TEX_IMAGES)
    assert ims[base][3] == 1 and ims[rough][3] == 0 and ims[nrm][3] == 0   # colour maps decode from sRGB
    np.testing.assert_array_equal(s.array(cwp.ARR_MESH_UVS), [[0, 0, 0, 0, 0, 0], [0, 0, 1, 1, 0, 1]])
    # texture kinds, --albedo and missing coordinates
    s2 = cwp.Setup(["--mesh", str(tmp_path / "m.obj"), "--hfov", "40", "--c2w", C2W, "--textures", "normal"])
    assert tuple(s2.array(cwp.ARR_MAPS)[2][[0, 2, 4]]) == (-1, nrm, -1)
    s3 = cwp.Setup(["--mesh", str(tmp_path / "m.obj"), "--hfov", "40", "--c2w", C2W, "--albedo", "0.5", "0.5", "0.5"])
    assert s3.array(cwp.ARR_MAPS)[2][0] == -1
    (tmp_path / "nouv.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 1 1 0\nusemtl tex\nf 1 2 3\n")
    s4 = cwp.Setup(["--mesh", str(tmp_path / "nouv.obj"), "--hfov", "40", "--c2w", C2W])
    assert s4.info["textured"] == 0 and len(s4.array(cwp.ARR_UVS)) == 0
    with pytest.raises(cwp.CwpError, match="--textures"):
        cwp.Setup(["--mesh", str(tmp_path / "m.obj"), "--hfov", "40", "--c2w", C2W, "--textures", "base,specular"])
    with pytest.raises(cwp.CwpError, match="--tex_filter"):
        cwp.Setup(["--mesh", str(tmp_path / "m.obj"), "--hfov", "40", "--c2w", C2W, "--tex_filter", "cubic"])


def test_uvs_follow_the_leaf_order():
    s = cwp.Setup(["--gen", "boxes", "--gen_boxes", "30", "--gen_textures", "8", "--width", "16", "--height", "16"])
    uv = s.array(cwp.ARR_UVS)
    muv = s.array(cwp.ARR_MESH_UVS)
    order = s.array(cwp.ARR_ORDER)
    np.testing.assert_array_equal(uv[:, 0, :], muv[order][:, :4])
    np.testing.assert_array_equal(uv[:, 1, :2], muv[order][:, 4:])
    assert (uv[:, 1, 2:] == 0).all()


# ------------------------------------------------------------------------------------------------ oracle closed forms
def plane_scene(tmp_path, kd_px, extra_mtl="", n_px=None):
    (tmp_path / "t.ppm").write_bytes(b"P6 1 1 255\n" + bytes(kd_px))
    if n_px is not None:
        (tmp_path / "n.ppm").write_bytes(b"P6 1 1 255\n" + bytes(n_px))
    (tmp_path / "p.mtl").write_text("newmtl m\nKd 0.6 0.6 0.6\nmap_Kd t.ppm\n" + extra_mtl)
    (tmp_path / "p.obj").write_text("mtllib p.mtl\nv -1000 -1000 0\nv 1000 -1000 0\nv 1000 1000 0\nv -1000 1000 0\n"
                                    "vt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\nusemtl m\nf 1/1 2/2 3/3\nf 1/1 3/3 4/4\n")
    return str(tmp_path / "p.obj")


def test_oracle_textured_plane_closed_form(tmp_path):
    """A Lambert plane whose base-colour map is one texel c under a constant sky L (no sun, no bounce beyond the
    first): a pixel is L (1 - f (1 - a c)), f = fraction of samples hitting the plane, with c the sRGB decode of the
    texel; the same scene without the map gives a = 0.6 alone."""
    sky = (0.4, 0.5, 0.6)
    common = ["--c2w", C2W, "--width", "24", "--height", "24", "--sky", *map(str, sky), "--sun_rgb", "0", "0", "0",
              "--filter", "box", "--batch", "8", "--spp", "8", "--integrator", "full", "--bounces", "0"]
    s = cwp.Setup(["--mesh", plane_scene(tmp_path, [200, 100, 50])] + common)
    assert s.info["textured"]
    sc, _ = po.scene_from_setup(s)
    acc, _ = sc.render(s.params(), 8, batch=8, bounces=0)
    img = acc[..., :3] / 8
    c = po.srgb_table()[[200, 100, 50]]
    a = 0.6 * c
    f = (1.0 - img / np.float32(sky)) / (1.0 - a)
    np.testing.assert_allclose(f * 8, np.round(f * 8), atol=2e-3)
    assert (np.abs(f - 1) < 1e-4).all(axis=-1).sum() > 0.3 * 24 * 24
    # a normal map of strength 0 leaves the shading normal as it is: the same image
    s0 = cwp.Setup(["--mesh", plane_scene(tmp_path, [200, 100, 50], "map_Bump -bm 0 n.ppm\n", [10, 200, 30])] + common)
    assert (s0.array(cwp.ARR_MAPS)[:, 2] >= 0).any()
    sc0, _ = po.scene_from_setup(s0)
    acc0, _ = sc0.render(s0.params(), 8, batch=8, bounces=0)
    np.testing.assert_allclose(acc0, acc, rtol=1e-5, atol=1e-6)
