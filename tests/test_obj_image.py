"""Scene ingestion on the host: image files (PNG / PNM / PFM / Radiance HDR) and
Wavefront OBJ + MTL, with the conventions of the reference's cpu::Image::open /
save (cpu/image.cpp:6-72) and TriangleMesh::from_obj (cpu/object.cpp:454-611).

The decoders are checked against an independent decoder (PIL) or against files
written here from the format definitions; the OBJ / MTL mapping against what
object.cpp:454-611 states.  OpenCV's Drago tonemap (used for float images) and
tinyobjloader (unrecorded commit) are not available, so those two stay unpinned:
only invariants are checked for the tonemap.
"""
import math
import struct

import numpy as np
import pytest
from PIL import Image as PILImage

import tinypt_amd as T

GAMMA = 2.2


def linear(a8):
    """8-bit file samples -> the texture values cpu::Image::open produces (row 0 = bottom)."""
    return np.power(np.asarray(a8, dtype=np.float64) / 255.0, GAMMA)[::-1]


def rgb_of(path, grayscale=False):
    rgb, alpha = T.Image.open(str(path), grayscale)
    return (None if rgb is None else rgb.raw()), (None if alpha is None else alpha.raw())


@pytest.mark.parametrize("mode", ["RGB", "RGBA", "L", "LA", "P", "1"])
def test_png_decodes_like_pil(tmp_path, mode):
    rng = np.random.default_rng(5)
    base = rng.integers(0, 256, (13, 17, 4), dtype=np.uint8)
    img = PILImage.fromarray(base, "RGBA").convert(mode)
    if mode == "P":
        img.info["transparency"] = 3  # one transparent palette entry
    p = tmp_path / f"x_{mode}.png"
    img.save(p)
    ref = np.asarray(PILImage.open(p).convert("RGBA" if mode in ("RGBA", "LA", "P") else "RGB"))
    rgb, alpha = rgb_of(p)
    assert rgb.shape == (13, 17, 3)
    np.testing.assert_allclose(rgb, linear(ref[..., :3]), rtol=1e-5, atol=1e-7)
    if mode in ("RGBA", "LA", "P"):
        assert alpha is not None
        np.testing.assert_allclose(alpha[..., 0], (ref[..., 3] / 255.0)[::-1], atol=1e-7)
    else:
        assert alpha is None


def test_png_16_bit_and_grayscale_mode(tmp_path):
    v = np.arange(0, 65536, 257, dtype=np.uint16).reshape(16, 16)
    p = tmp_path / "g16.png"
    PILImage.fromarray(v.astype(np.int32), "I").convert("I;16").save(p)
    _, alpha = rgb_of(p, grayscale=True)
    np.testing.assert_allclose(alpha[..., 0], (v / 65535.0)[::-1], atol=1e-6)
    # colour read as grayscale: BT.601 luma in OpenCV's fixed point, not gamma-corrected
    rng = np.random.default_rng(6)
    c = rng.integers(0, 256, (9, 11, 3), dtype=np.uint8)
    p2 = tmp_path / "c.png"
    PILImage.fromarray(c, "RGB").save(p2)
    _, a2 = rgb_of(p2, grayscale=True)
    c = c.astype(np.int64)
    g = (c[..., 0] * 4899 + c[..., 1] * 9617 + c[..., 2] * 1868 + 8192) >> 14
    np.testing.assert_array_equal(a2[..., 0], (g / np.float32(255.0)).astype(np.float32)[::-1])


@pytest.mark.parametrize("mode,ext", [("RGB", "ppm"), ("L", "pgm")])
def test_pnm_decodes_like_pil(tmp_path, mode, ext):
    rng = np.random.default_rng(7)
    a = rng.integers(0, 256, (10, 12, 3), dtype=np.uint8)
    img = PILImage.fromarray(a, "RGB").convert(mode)
    p = tmp_path / f"x.{ext}"
    img.save(p)
    ref = np.asarray(img.convert("RGB"))
    rgb, _ = rgb_of(p)
    np.testing.assert_allclose(rgb, linear(ref), rtol=1e-5, atol=1e-7)


def test_save_then_open_round_trips(tmp_path):
    rng = np.random.default_rng(8)
    lin = rng.uniform(0, 1, (6, 9, 3)).astype(np.float32)
    im = T.Image.from_linear(lin)  # row 0 = bottom, as stored
    for ext in ("png", "ppm"):
        p = tmp_path / f"r.{ext}"
        im.save(str(p))
        # the file holds round(255 * v^(1/2.2)), top row first
        f = np.asarray(PILImage.open(p).convert("RGB"))
        expect = np.clip(np.rint(255 * np.power(lin, 1 / GAMMA)), 0, 255)[::-1]
        assert np.max(np.abs(f.astype(np.int64) - expect)) <= 1
        back, _ = rgb_of(p)
        np.testing.assert_allclose(back, np.power(np.round(255 * np.power(lin, 1 / GAMMA)) / 255, GAMMA), atol=2e-3)


def write_pfm(path, a, little=True):
    h, w = a.shape[:2]
    c = 1 if a.ndim == 2 else a.shape[2]
    hdr = f"{'PF' if c == 3 else 'Pf'}\n{w} {h}\n{-1.0 if little else 1.0}\n".encode()
    rows = a[::-1].astype("<f4" if little else ">f4")  # PFM stores the bottom row first
    path.write_bytes(hdr + rows.tobytes())


def rgbe(v):
    """Ward's float -> RGBE: mantissas m with v = m * 2^(e - 136)."""
    m = max(v)
    if m < 1e-32:
        return bytes(4)
    e = math.frexp(m)[1]
    s = 256.0 / 2.0 ** e
    return bytes([int(x * s) for x in v] + [e + 128])


def test_float_images_are_tonemapped_into_unit_range(tmp_path):
    rng = np.random.default_rng(9)
    hdr = np.exp(rng.normal(0, 2, (8, 10, 3))).astype(np.float32)  # high dynamic range
    p = tmp_path / "x.pfm"
    write_pfm(p, hdr)
    rgb, _ = rgb_of(p)
    assert rgb.shape == (8, 10, 3) and np.all(np.isfinite(rgb))
    assert rgb.min() >= 0 and rgb.max() <= 1 and rgb.max() > 0.5
    # brighter pixels stay brighter (luminance order preserved by the tonemap)
    lum_in = (hdr @ [0.299, 0.587, 0.114])[::-1].ravel()
    lum_out = (rgb @ [0.299, 0.587, 0.114]).ravel()
    i, j = np.argmax(lum_in), np.argmin(lum_in)
    assert lum_out[i] > lum_out[j]


def test_radiance_hdr_decodes_flat_and_run_length_scanlines(tmp_path):
    w, h = 12, 3
    rng = np.random.default_rng(10)
    vals = rng.uniform(0.01, 50, (h, w, 3))
    enc = [[rgbe(vals[y, x]) for x in range(w)] for y in range(h)]
    body = b""
    for y in range(h):
        if y == 1:  # new-style run-length scanline: 4 components, each as one literal run
            body += bytes([2, 2, w >> 8, w & 255])
            for c in range(4):
                body += bytes([w]) + bytes(enc[y][x][c] for x in range(w))
        else:
            body += b"".join(enc[y])
    p = tmp_path / "x.hdr"
    p.write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n" % (h, w) + body)
    # decoded floats: m * 2^(e - 136), read through the PFM path of the same code for comparison
    dec = np.array([[[b[c] * 2.0 ** (b[3] - 136) for c in range(3)] for b in row] for row in enc], np.float32)
    q = tmp_path / "same.pfm"
    write_pfm(q, dec)
    a, _ = rgb_of(p)
    b, _ = rgb_of(q)
    np.testing.assert_array_equal(a, b)


def test_image_errors_name_the_file(tmp_path):
    p = tmp_path / "x.jpg"
    p.write_bytes(b"\xff\xd8\xff\xe0 not supported")
    with pytest.raises(RuntimeError, match="x.jpg"):
        T.Image.open(str(p))
    with pytest.raises(RuntimeError, match="cannot open image"):
        T.Image.open(str(tmp_path / "missing.png"))


# ---------------------------------------------------------------- OBJ / MTL

OBJ = """\
# two shapes; y-up coordinates
mtllib scene.mtl
o first
v 0 0 0
v 2 0 0
v 2 0 1
v 0 0 3
v 1 1 1
vn 0 1 0
vt 0 0
vt 1 0
vt 1 1
usemtl diffuse_tex
f 1/1/1 2/2/1 3/3/1 4/1/1
usemtl shiny
f -5//1 -3//1 -1//1
o second
usemtl nosuch
f 1 2 5
usemtl metal
f 1 2 3 4 5
usemtl glass
f 2/1 3/2 5/3
g third
usemtl veryshiny
f 1 3 5
usemtl dissolve
f 2 4 5
"""

MTL = """\
newmtl diffuse_tex
Kd 0.5 0.5 0.5
map_Kd -bm 1.0 Tex.PNG
bump -bm 0.5 height.png
newmtl shiny
Kd 0.2 0.3 0.4
Ks 0.5 0.5 0.5
Ns 64
Ke 1 2 3
newmtl metal
illum 3
Ks 0.9 0.8 0.7
newmtl glass
illum 7
Ni 1.45
newmtl veryshiny
Kd 1 1 1
Ks 0.4 0.4 0.4
Ns 2000
newmtl dissolve
Kd 0.7 0.7 0.7
d 0.25
"""


@pytest.fixture
def obj_dir(tmp_path):
    (tmp_path / "scene.obj").write_text(OBJ)
    (tmp_path / "scene.mtl").write_text(MTL)
    rng = np.random.default_rng(11)
    PILImage.fromarray(rng.integers(0, 256, (4, 4, 4), dtype=np.uint8), "RGBA").save(tmp_path / "tex.png")
    PILImage.fromarray(rng.integers(0, 256, (4, 4), dtype=np.uint8), "L").save(tmp_path / "height.png")
    return tmp_path


def test_obj_geometry(obj_dir):
    m = T.TriangleMesh.from_obj(str(obj_dir / "scene.obj"))
    v = m.vertices
    assert v.shape == (5, 3)
    np.testing.assert_array_equal(v[2], [2, -1, 0])  # yup2zup: (x, y, z) -> (x, -z, y)
    np.testing.assert_array_equal(m.normals, [[0, 0, 1]])
    f = m.faces
    # quad 1-2-3-4: diagonal 0-2 has length^2 4 + 1 = 5 < diagonal 1-3 (4 + 9 = 13) -> (0,1,2), (0,2,3)
    assert f[0, 0::3].tolist() == [0, 1, 2] and f[1, 0::3].tolist() == [0, 2, 3]
    assert f[0, 1::3].tolist() == [0, 0, 0] and f[0, 2::3].tolist() == [0, 1, 2]  # (v, n, t) per corner
    assert f[2, 0::3].tolist() == [0, 2, 4]  # negative indices count from the end
    assert f[2, 2::3].tolist() == [-1, -1, -1]  # v//n: no texcoord
    assert f[3, 1::3].tolist() == [-1, -1, -1]  # bare v: no normal
    # pentagon 1 2 3 4 5 fans from its first corner
    assert [f[i, 0::3].tolist() for i in (4, 5, 6)] == [[0, 1, 2], [0, 2, 3], [0, 3, 4]]
    assert m.num_faces == 10
    assert m.shape_offsets.tolist() == [0, 3, 8, 10]  # "o first", "o second", "g third"


def test_obj_materials(obj_dir):
    default = T.Material(T.Metal(T.RGBTexture([0.1, 0.2, 0.3])))
    m = T.TriangleMesh.from_obj(str(obj_dir / "scene.obj"), default)
    mats = m.materials
    names = ["diffuse_tex", "shiny", "metal", "glass", "veryshiny", "dissolve"]
    assert len(mats) == len(names) + 1
    fm = m.face_material.tolist()
    assert fm == [0, 0, 1, 6, 2, 2, 2, 3, 4, 5]  # unknown material -> the default, appended last
    d = mats[0]
    assert isinstance(d.surface, T.Lambertian) and d.surface.texture.map is not None  # map_Kd (case-insensitive)
    assert d.alpha_texture.map is not None  # the diffuse map's alpha channel
    assert d.bump_texture.map is not None
    s = mats[1].surface
    assert isinstance(s, T.AddBXDF) and isinstance(s.bxdf1, T.Lambertian) and isinstance(s.bxdf2, T.Glossy)
    assert s.bxdf2.shininess == 64
    np.testing.assert_allclose(mats[1].emission_texture.value, [1, 2, 3])
    assert isinstance(mats[2].surface, T.Metal)
    np.testing.assert_allclose(mats[2].surface.texture.value, [0.9, 0.8, 0.7], rtol=1e-6)
    assert isinstance(mats[3].surface, T.Dielectric) and abs(mats[3].surface.ior - 1.45) < 1e-6
    vs = mats[4].surface
    assert isinstance(vs, T.AddBXDF) and isinstance(vs.bxdf2, T.Metal)  # Ns >= 1024
    assert isinstance(mats[5].surface, T.Lambertian) and abs(mats[5].alpha_texture.value - 0.25) < 1e-6
    assert isinstance(mats[6].surface, T.Metal)  # the given default material


def test_obj_errors(tmp_path):
    with pytest.raises(RuntimeError, match="Failed to parse obj file"):
        T.TriangleMesh.from_obj(str(tmp_path / "none.obj"))
    (tmp_path / "bad.obj").write_text("v 0 0 0\nf 1 2 3\n")
    with pytest.raises(RuntimeError, match="index out of range"):
        T.TriangleMesh.from_obj(str(tmp_path / "bad.obj"))
    (tmp_path / "tex.obj").write_text("mtllib t.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl a\nf 1 2 3\n")
    (tmp_path / "t.mtl").write_text("newmtl a\nmap_Kd nothere.png\n")
    with pytest.raises(RuntimeError, match="Cannot find file nothere.png"):
        T.TriangleMesh.from_obj(str(tmp_path / "tex.obj"))


def test_obj_scene_flattens(obj_dir):
    # the loaded mesh goes through the same scene description as array meshes
    m = T.TriangleMesh.from_obj(str(obj_dir / "scene.obj"))
    cam = T.Camera()
    cam.location = [1, -5, 1]
    cam.width, cam.height = 8, 6
    scene = T.Scene(cam, [m], [], None)
    with pytest.raises(RuntimeError):
        scene.to("cuda")  # not built: raises like a reference build without CUDA
