"""Asset ingest: the OBJ / MTL reader (obj.cs), the RGBE .hdr reader (image.cs:307-397) and the TGA
reader (image.cs:217-266). The restatement in oracle/assets_ref.py is pinned by hand-built files
whose answers are written out here; the library (host code, no GPU) is then checked against it."""
import math
import struct

import numpy as np
import pytest

from oracle import assets_ref as R


QUAD = """# unit quad in z = 0, two uv sets
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vt 0 0
vt 1 0
vt 1 1
vt 0 1
f 1/1 2/2 3/3 4/4
"""


def test_oracle_fan_triangulation_vflip_and_defaults():
    objs = R.parse_obj(QUAD)
    assert len(objs) == 1
    o = objs[0]
    # Fan from the first vertex: (1, 2, 3), (1, 3, 4).
    assert o["pos"].tolist() == [[0, 0, 0, 1, 0, 0, 1, 1, 0], [0, 0, 0, 1, 1, 0, 0, 1, 0]]
    # V flipped: vt (u, v) -> (u, 1 - v).
    assert o["uv"].tolist() == [[0, 1, 1, 1, 1, 0], [0, 1, 1, 0, 0, 0]]
    assert not o["has_nrm"].any() and o["smooth"]
    # No vt: the defaults (0,0), (1,0), (0,1) for the fan's A, B, C.
    p = R.parse_obj("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\nf 1 2 3 4\n")[0]
    assert p["uv"].tolist() == [[0, 0, 1, 0, 0, 1]] * 2


def test_oracle_indices_normals_and_objects():
    text = """v 0 0 0
v 1 0 0
v 0 1 0
vn 0 0 1
vn 0 1 0
o first
f -3//-2 -2//-2 -1//-1
f 1 2 3
g second extra words
f 1 2 3
usemtl missing
f 3 2 1
"""
    objs = R.parse_obj(text)
    assert len(objs) == 3
    a = objs[0]
    # Negative indices count back from the end; normals only where all three corners have one.
    assert a["pos"][0].tolist() == [0, 0, 0, 1, 0, 0, 0, 1, 0]
    assert a["has_nrm"].tolist() == [True, False]
    assert a["nrm"][0].tolist() == [0, 0, 1, 0, 0, 1, 0, 1, 0]
    assert not a["smooth"]  # the file has vn
    assert objs[1]["pos"].shape == (1, 9) and objs[2]["pos"][0].tolist() == [0, 1, 0, 1, 0, 0, 0, 0, 0]
    # usemtl naming an unknown material falls back to White / roughness 1.
    assert objs[2]["material"]["albedo"] == [1.0, 1.0, 1.0] and objs[2]["material"]["roughness"] == 1.0


@pytest.mark.parametrize("text,msg", [
    ("v 0 0 0\nv 1 0 0\nf 1 2\n", "fewer than 3"),
    ("v 0 0 0\nf 1 2 3\n", "index out of bounds (1, count: 1)"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n", "index out of bounds (-1, count: 3)"),
    ("v 0 x 0\n", "invalid number 'x'"),
    ("v 0 0\n", "expected number, got Newline"),
    ("v 0 0 0\n", "no objects found"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3.5\n", "invalid int '3.5'"),
    # The lexer skips blanks, so after "1/" the next word is read as a texture index.
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1/ 2 3\n", "index out of bounds (1, count: 0)"),
])
def test_oracle_errors(text, msg):
    with pytest.raises(R.ObjError, match=msg.replace("(", r"\(").replace(")", r"\)")):
        R.parse_obj(text)


MTL = """newmtl shiny
Kd 0.5 0.25 0.125
Ns 98
Pm 1
Ni 1.33
d 0.75
newmtl rough # trailing comment
Pr 0.6
Ns 1000
Ke 2 3 4
Tr 0.5
"""


def test_oracle_mtl_fields():
    m = R.parse_mtl(MTL)
    s, r = m["shiny"], m["rough"]
    assert s["albedo"] == [0.5, 0.25, 0.125]
    assert s["roughness"] == pytest.approx(math.sqrt(2 / 100), rel=1e-7)  # sqrt(2 / (Ns + 2))
    assert s["metallic"] == 1.0 and s["ior"] == pytest.approx(1.33) and s["transparency"] == 0.25
    assert r["roughness"] == pytest.approx(0.6) and r["radiance"] == [2.0, 3.0, 4.0] and r["transparency"] == 0.5
    assert r["albedo"] == [1.0, 1.0, 1.0] and r["metallic"] == 0.0


def _tga(w, h, pixels_rgb, top_left, gray=False):
    hdr = bytes([0, 0, 3 if gray else 2]) + bytes(9) + struct.pack("<HH", w, h) + bytes([8 if gray else 24,
                                                                                           0x20 if top_left else 0])
    rows = pixels_rgb if top_left else pixels_rgb[::-1]
    body = rows[..., 0].tobytes() if gray else rows[..., ::-1].tobytes()
    return hdr + body


def test_tga_decode_oracle_and_library():
    import sharpray_amd as A
    px = np.array([[[255, 0, 0], [0, 255, 0], [0, 0, 255]], [[1, 2, 3], [4, 5, 6], [7, 8, 9]]], np.uint8)
    for top_left in (True, False):
        data = _tga(3, 2, px, top_left)
        assert np.array_equal(R.decode_tga(data), px)
        assert np.array_equal(A.decode_tga(data), px)
    g = np.repeat(np.array([[[10], [20]], [[30], [40]]], np.uint8), 3, 2)
    data = _tga(2, 2, g, False, gray=True)
    assert np.array_equal(R.decode_tga(data), g) and np.array_equal(A.decode_tga(data), g)
    with pytest.raises(A.SharpRayError, match="Only 24-bit"):
        A.decode_tga(bytes([0, 0, 2]) + bytes(9) + struct.pack("<HH", 1, 1) + bytes([32, 0]) + bytes(4))


def _rle_channel(vals):
    """Textbook Radiance RLE of one channel: runs of >= 4 equal bytes as (128 + n, v), else literals."""
    out, i, n = bytearray(), 0, len(vals)
    while i < n:
        j = i
        while j < n and j - i < 127 and vals[j] == vals[i]:
            j += 1
        if j - i >= 4:
            out += bytes([128 + j - i, vals[i]])
            i = j
            continue
        k = i
        while k < n and k - i < 128:
            if k + 3 < n and vals[k] == vals[k + 1] == vals[k + 2] == vals[k + 3]:
                break
            k += 1
        k = max(k, i + 1)
        out += bytes([k - i]) + bytes(vals[i:k])
        i = k
    return bytes(out)


def _hdr(rgbe, rle):
    h, w, _ = rgbe.shape
    data = bytearray(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0\n\n" + f"-Y {h} +X {w}\n".encode())
    for y in range(h):
        if rle:
            data += bytes([2, 2, w >> 8, w & 255])
            for c in range(4):
                data += _rle_channel(bytes(rgbe[y, :, c]))
        else:
            data += rgbe[y].tobytes()
    return bytes(data)


def test_hdr_known_values_and_rle():
    import sharpray_amd as A
    # (128, 64, 32, E = 129): scale 2^(129-128) / 256 -> (1, 0.5, 0.25); E = 0 is black.
    rgbe = np.array([[[128, 64, 32, 129], [255, 255, 255, 0]], [[1, 2, 3, 128], [200, 100, 50, 136]]], np.uint8)
    want = np.array([[[1.0, 0.5, 0.25], [0, 0, 0]], [[1 / 256, 2 / 256, 3 / 256], [200.0, 100.0, 50.0]]], np.float32)
    for rle in (False, True):
        if rle:  # RLE scanlines need w >= 8 in practice; any w decodes here
            pass
        data = _hdr(rgbe, rle)
        assert np.array_equal(R.decode_hdr(data), want)
        assert np.array_equal(A.decode_hdr(data), want)
    # Luminance clamp of Texture.FromHdr: (200, 100, 50) has luminance 0.2126*200 + 0.7152*100 + 0.0722*50.
    lum = 0.2126 * 200 + 0.7152 * 100 + 0.0722 * 50
    c = A.decode_hdr(_hdr(rgbe, False), max_luminance=10.0)
    assert c[1, 1] == pytest.approx(want[1, 1] * (10.0 / lum), rel=1e-6) and np.array_equal(c[0], want[0])


def test_hdr_rle_round_trip_wide():
    import sharpray_amd as A
    rng = np.random.default_rng(2)
    h, w = 5, 300
    rgbe = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    rgbe[:, 40:140] = rgbe[:, 40:41]  # long runs
    rgbe[..., 3] = rng.integers(120, 140, (h, w))
    data = _hdr(rgbe, True)
    ref = R.decode_hdr(data)
    assert np.array_equal(A.decode_hdr(data), ref)
    e = rgbe[..., 3].astype(np.float64)
    direct = rgbe[..., :3] * (np.ldexp(1.0, (e - 128).astype(int)) / 256.0)[..., None]
    assert np.array_equal(ref, direct.astype(np.float32))


@pytest.mark.parametrize("data,msg", [
    (b"#?NOPE\n\n-Y 1 +X 1\n\x00\x00\x00\x00", "Invalid magic"),
    (b"#?RGBE\n\n+Y 1 +X 1\n\x00\x00\x00\x00", "Unsupported size/orientation"),
    (b"#?RGBE\n\n-Y 1 +X 8\n\x02\x02\x00\x09", "Scanline width mismatch"),
    (b"#?RGBE\n\n-Y 1 +X 8\n\x01\x01\x01\x00", "Old RLE format"),
    (b"#?RGBE\n\n-Y 1 +X 8\n\x02\x02\x00\x08\x00", "length 0"),
    (b"#?RGBE\n\n-Y 2 +X 1\n\x01\x02\x03\x80", "unexpected end of file"),
    # A flat scanline's 4-byte header does not fit a zero-width scanline (image.cs:371 throws).
    (b"#?RADIANCE\n\n-Y 2 +X 0\n" + bytes(16), "width 0"),
])
def test_hdr_errors(data, msg):
    import sharpray_amd as A
    with pytest.raises(A.SharpRayError, match=msg):
        A.decode_hdr(data)


def test_hdr_zero_width_rle_scanline_is_an_empty_image():
    """An RLE scanline header of width 0 matches a width-0 image: the reference decodes nothing and
    returns an empty image (image.cs:336-365), while a flat scanline of width 0 is an error."""
    import sharpray_amd as A
    out = A.decode_hdr(b"#?RGBE\n\n-Y 2 +X 0\n\x02\x02\x00\x00\x02\x02\x00\x00")
    assert out.shape == (2, 0, 3)


def _load(tmp_path, files, config=None):
    import sharpray_amd as A
    R.write_files(str(tmp_path), files)
    scene = A.Scene(-1)
    n = A.ObjLoader.load(tmp_path / "model.obj", scene, config)
    return scene, n


def _check_against_oracle(scene, objs, first_object=0):
    import sharpray_amd as A
    for k, o in enumerate(objs):
        rec = scene.object_record(first_object + k)
        tris = scene.mesh_triangles(rec["mesh"])
        assert tris.shape[0] == o["pos"].shape[0]
        pa, ab, ac = tris[:, 0:3], tris[:, 3:6], tris[:, 6:9]
        pos = o["pos"].reshape(-1, 3, 3)
        assert np.array_equal(pa, pos[:, 0])
        assert np.array_equal(ab, (pos[:, 1] - pos[:, 0]).astype(np.float32))
        assert np.array_equal(ac, (pos[:, 2] - pos[:, 0]).astype(np.float32))
        assert np.array_equal(tris[:, 30:36], o["uv"])
        # Triangle ctors on the oracle's parse give the same records (smooth normals included).
        if o["has_nrm"].all() or not o["has_nrm"].any():
            want = A.triangles_build(o["pos"], o["nrm"] if o["has_nrm"].all() else None, o["uv"],
                                     smooth=o["smooth"] and not o["has_nrm"].any())
            assert np.array_equal(tris, want.reshape(-1, 36))
        else:
            nr = tris[:, 9:18]
            assert np.array_equal(nr[o["has_nrm"]], o["nrm"][o["has_nrm"]])
        m = o["material"]
        assert rec["material"]["albedo"] == pytest.approx(m["albedo"]) and \
            rec["material"]["roughness"] == pytest.approx(m["roughness"], rel=1e-7)


def test_library_obj_matches_oracle(tmp_path):
    text = """mtllib model.mtl
v -1 0 -1
v 1 0 -1
v 1 0 1
v -1 0 1
v 0 2 0
vt 0 0
vt 1 0
vt 0.5 1
o base
usemtl shiny
f 1/1 2/2 3/3 4/2
o apex
usemtl rough
f 1/1 2/2 5/3
f 2/2 3/2 5
f 3 4 5
f 4 1 5
"""
    scene, n = _load(tmp_path, {"model.obj": text, "model.mtl": MTL})
    objs = R.parse_obj(text, {"model.mtl": MTL})
    assert n == len(objs) == 2
    assert scene.counts() == (0, 2, 2)
    _check_against_oracle(scene, objs)
    rec = scene.object_record(1)
    assert rec["material"]["radiance"] == [2.0, 3.0, 4.0] and rec["material"]["transparency"] == 0.5


def test_library_obj_normals_transform_default_material(tmp_path):
    import sharpray_amd as A
    text = """v 0 0 0
v 1 0 0
v 0 1 0
v 1 1 0
vn 0 0 1
vn 0.6 0 0.8
f 1//1 2//2 3//1
f 2 4 3
"""
    mat = A.Material((0.2, 0.3, 0.4), 0.7)
    cfg = A.ObjConfig(A.Transform((1, 2, 3), A.Quat.angle_axis(0.5, (0, 1, 0)), (2, 2, 2)), mat, True)
    scene, n = _load(tmp_path, {"model.obj": text}, cfg)
    objs = R.parse_obj(text, default_material=dict(albedo=[0.2, 0.3, 0.4], roughness=0.7))
    assert n == 1
    _check_against_oracle(scene, objs)
    rec = scene.object_record(0)
    assert rec["pos"] == [1, 2, 3] and rec["scale"] == [2, 2, 2]
    assert rec["rot"] == pytest.approx(list(A.Quat.angle_axis(0.5, (0, 1, 0)).xyzw))


def test_library_obj_smooth_normals_and_textures(tmp_path):
    import sharpray_amd as A
    # An octahedron without vn: smooth normals make every vertex normal radial.
    text = "mtllib m.mtl\nusemtl tex\n" + "\n".join(
        f"v {x} {y} {z}" for x, y, z in [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]) + """
f 1 3 5
f 3 2 5
f 2 4 5
f 4 1 5
f 3 1 6
f 2 3 6
f 4 2 6
f 1 4 6
"""
    px = np.array([[[255, 128, 0], [0, 0, 0]], [[64, 64, 64], [255, 255, 255]]], np.uint8)
    files = {"model.obj": text, "m.mtl": "newmtl tex\nmap_Kd c.tga\nmap_Pr r.tga\n", "c.tga": _tga(2, 2, px, True),
             "r.tga": _tga(2, 2, px, False)}
    scene, n = _load(tmp_path, files)
    objs = R.parse_obj(text, {"m.mtl": files["m.mtl"]})
    _check_against_oracle(scene, objs)
    tris = scene.mesh_triangles(0)
    pos = objs[0]["pos"].reshape(-1, 3, 3)
    nr = tris[:, 9:18].reshape(-1, 3, 3)
    assert np.allclose(nr, pos / np.linalg.norm(pos, axis=-1, keepdims=True), atol=1e-6)
    rec = scene.object_record(0)
    m = rec["material"]
    assert m["color_texture"] == 0 and m["roughness_texture"] == 1 and m["metallic_texture"] == -1
    srgb = scene.texture_texels(0)
    assert srgb[0, 0, 0] == np.float32(1.0) and srgb[0, 0, 2] == 0.0
    assert srgb[0, 0, 1] == pytest.approx((128 / 255) ** 2.233333333, rel=1e-6)  # Color.FromPixel sRGB decode
    lin = scene.texture_texels(1)
    assert np.array_equal(lin, (px / np.float32(255)).astype(np.float32))  # FromLinear, rows top-down


def test_library_obj_errors(tmp_path):
    import sharpray_amd as A
    for text, msg in [("v 0 0 0\nv 1 0 0\nf 1 2\n", "fewer than 3"), ("v 0 0 0\n", "no objects found"),
                      ("v 0 0 0\nf 1 2 3\n", r"index out of bounds \(1, count: 1\)")]:
        with pytest.raises(A.SharpRayError, match=msg):
            _load(tmp_path, {"model.obj": text})
    with pytest.raises(A.SharpRayError, match="Unsupported file format"):
        _load(tmp_path, {"model.obj": QUAD.replace("v 0 0 0", "mtllib m.mtl\nusemtl a\nv 0 0 0"),
                         "m.mtl": "newmtl a\nmap_Kd tex.png\n"})


def test_icosphere_workload_through_obj_and_hdr_files(tmp_path):
    """The 256^2 icosphere workload's assets written as the reference would read them (an OBJ
    without vn, an RGBE .hdr sky) and loaded through ObjLoader / load_hdr give bit-identical
    triangle records and texels to the in-memory construction."""
    import sharpray_amd as A
    from sharpray_amd import scenes
    ico = scenes.icosphere(3, seed=scenes.SEED_BASE)
    scenes.write_obj(tmp_path / "ico.obj", ico)
    scene = A.Scene(-1)
    assert A.ObjLoader.load(tmp_path / "ico.obj", scene) == 1
    want = A.triangles_build(ico, None, None, smooth=True).reshape(-1, 36)
    assert np.array_equal(scene.mesh_triangles(0), want)
    sky = scenes.procedural_sky(64, 32, peak=50.0)
    rgbe = scenes.rgbe_encode(sky)
    scenes.write_hdr(tmp_path / "sky.hdr", rgbe)
    assert np.array_equal(A.load_hdr(tmp_path / "sky.hdr"), A.Texture.from_hdr(rgbe).rgb)
    # A ground quad with uvs round-trips its uvs through the V flip.
    gp, guv = scenes.quad(-1.0, 6.0)
    scenes.write_obj(tmp_path / "ground.obj", gp, guv)
    scene2 = A.Scene(-1)
    A.ObjLoader.load(tmp_path / "ground.obj", scene2)
    assert np.array_equal(scene2.mesh_triangles(0), A.triangles_build(gp, None, guv).reshape(-1, 36))
