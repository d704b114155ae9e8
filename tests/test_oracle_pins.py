"""Closed-form pins of the oracle parts that no golden of the reference fixes:
the Glossy lobe's microfacet models (Blinn-Phong and Trowbridge-Reitz / GGX,
cpu/material.cpp:35-161), the sphere light's cone sampling and pdf
(cpu/object.cpp:66-104), the camera and its thin lens (cpu/camera.cpp:6-34) and
BumpTexture::bump_normal (cpu/texture.cpp:52-63).  Each check is a property the
mathematics fixes (a normalisation integral, a sampling histogram, a symmetry,
an energy bound, an independent published formula, a geometric identity), chosen
so that a plausible slip (a dropped factor, a wrong exponent, a swapped vector,
a transposed axis) fails it.  CPU only."""
import ctypes
import math

import numpy as np
import pytest
from scipy import special, stats

import oracle_py as O
from test_oracle_goldens import _mat, _ray, _scene

KINDS = [(O.BLINN, "blinn-phong"), (O.GGX, "ggx")]
SHININESS = [1.0, 50.0, 1000.0]
N_Z = np.array([0, 0, 1], dtype=np.float32)


def _unit(v):
    v = np.asarray(v, dtype=np.float64)
    return (v / np.linalg.norm(v)).astype(np.float32)


def _h_at(cos_b, phi=0.3):
    s = math.sqrt(max(0.0, 1.0 - cos_b * cos_b))
    return np.array([s * math.cos(phi), s * math.sin(phi), cos_b], dtype=np.float32)


def _cos_grid(n=4000):
    """cos(beta) from 0 to 1, dense near 1 where the sharp lobes live."""
    t = np.concatenate([[0.0], np.geomspace(1e-9, 1.0, n)])
    return np.unique(np.clip(1.0 - t, 0.0, 1.0))


def _D_of_cos(kind, s, c):
    return np.array([O.micro_terms(kind, s, N_Z, N_Z, N_Z, _h_at(x))[0] for x in c], dtype=np.float64)


# ---------------------------------------------------------------- microfacet distributions
@pytest.mark.parametrize("kind,name", KINDS)
@pytest.mark.parametrize("s", SHININESS)
def test_microfacet_D_is_a_density_of_h(kind, name, s):
    """Both models sample h with pdf_h = D (material.cpp:68, 101), so D must
    integrate to 1 over the hemisphere of h: the reference's D carries the
    cos(beta) of the projected-area normalisation inside it
    (Blinn-Phong (s + 2) / 2pi cos^(s+1); GGX alpha^2 cos / (pi (alpha^2 cos^2 + sin^2)^2),
    Walter et al. 2007).  A wrong exponent or constant moves the integral."""
    c = _cos_grid()
    D = _D_of_cos(kind, s, c)
    total = np.trapezoid(D, c) * 2 * math.pi
    assert abs(total - 1.0) < 2e-3, (name, s, total)
    assert np.all(D >= 0)


def _cdf_closed_form(kind, s, c):
    """P(cos(beta) <= c) of the sampled microfacet normal."""
    c = np.asarray(c, dtype=np.float64)
    if kind == O.BLINN:
        return c ** (s + 2.0)
    a2 = 2.0 / (s + 2.0)
    return a2 * c * c / (1.0 - c * c * (1.0 - a2))


@pytest.mark.parametrize("kind,name", KINDS)
@pytest.mark.parametrize("s", SHININESS)
def test_microfacet_sampling_follows_D(kind, name, s):
    """chi^2 of the sampled cos(beta) histogram against the bin masses of D
    (integrated from the oracle's own D, so sampling and density must agree),
    the closed-form CDF of the same distribution, and a uniform azimuth about a
    tilted normal (Rotation::from_direction)."""
    n = _unit([0.3, -0.5, 0.8])
    N = 60000
    h = np.concatenate([O.micro_sample_h(kind, s, n, seed, N // 6) for seed in range(6)]).astype(np.float64)
    assert np.allclose(np.linalg.norm(h, axis=1), 1.0, atol=2e-6)
    cb = np.clip(h @ n.astype(np.float64), -1, 1)
    assert cb.min() > -1e-6
    K = 20
    # equal-mass bin edges from the closed form, masses from the oracle's D
    qs = np.linspace(0, 1, K + 1)
    grid = _cos_grid(6000)
    cdf_closed = _cdf_closed_form(kind, s, grid)
    edges = np.interp(qs, cdf_closed, grid)
    edges[0], edges[-1] = 0.0, 1.0
    D = _D_of_cos(kind, s, grid)
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (D[1:] + D[:-1]) * np.diff(grid))]) * 2 * math.pi
    assert np.max(np.abs(cum - cdf_closed)) < 3e-3  # the closed form is the integral of D
    mass = np.diff(np.interp(edges, grid, cum))
    mass /= mass.sum()
    counts, _ = np.histogram(cb, bins=edges)
    chi2 = float(np.sum((counts - N * mass) ** 2 / (N * mass)))
    assert chi2 < stats.chi2.ppf(0.999, K - 1), (name, s, chi2)
    # azimuth about n is uniform
    t = np.cross(n, [1.0, 0.0, 0.0])
    t /= np.linalg.norm(t)
    b = np.cross(n, t)
    ph = np.arctan2(h @ b, h @ t)
    keep = cb < 0.999999  # the azimuth of h == n is undefined
    pc, _ = np.histogram(ph[keep], bins=16, range=(-math.pi, math.pi))
    e = keep.sum() / 16
    assert float(np.sum((pc - e) ** 2 / e)) < stats.chi2.ppf(0.999, 15)


def _glossy_scene(kind, s, albedo=(1.0, 1.0, 1.0)):
    import tinypt_amd as T

    mt = T.Glossy.MICRO_BLINN_PHONG if kind == O.BLINN else T.Glossy.MICRO_TROWBRIDGE_REITZ
    sp = T.Sphere()
    sp.material = T.Material(T.Glossy(s, T.RGBTexture(list(albedo)), mt))
    return O.OracleScene(_scene([sp]))


def _f(o, l, v):
    """Glossy::f(ray = -l, out = v) and pdf on a hit with normal +z."""
    rd = np.ascontiguousarray(-np.asarray(l, np.float32))
    vv = np.ascontiguousarray(v, dtype=np.float32)
    f = np.zeros(3, dtype=np.float32)
    p = ctypes.c_float()
    O.lib().oracle_bxdf_eval(o.h, 0, rd.ctypes.data, N_Z.ctypes.data, 1, vv.ctypes.data, f.ctypes.data, ctypes.byref(p))
    return f.astype(np.float64), p.value


@pytest.mark.parametrize("kind,name", KINDS)
def test_glossy_f_is_reciprocal(kind, name):
    """f(l, v) = f(v, l): D and G are symmetric, 4 (n.l)(n.v) too, and Schlick's
    F uses v.h = l.h for the half vector (material.cpp:123-134, 157-161)."""
    rng = np.random.default_rng(5)
    for s in SHININESS:
        o = _glossy_scene(kind, s, albedo=(0.9, 0.5, 0.2))
        for _ in range(60):
            l = _unit(rng.normal(size=3) * [1, 1, 0.4] + [0, 0, 0.8])
            v = _unit(rng.normal(size=3) * [1, 1, 0.4] + [0, 0, 0.8])
            if l[2] < 0.05 or v[2] < 0.05:
                continue
            a, _ = _f(o, l, v)
            b, _ = _f(o, v, l)
            assert np.allclose(a, b, rtol=2e-5, atol=1e-7 * max(1.0, a.max())), (name, s, a, b)


def _albedo_integral(o, kind, s, l, nb=320, nphi=48):
    """int f(l, v) cos(theta_v) dv over the hemisphere, taken over the half
    vector h (dv = 4 (v.h) dh) on a grid dense about the lobe's peak."""
    tb = np.concatenate([[0.0], np.geomspace(1e-6, math.pi / 2, nb)])
    ph = (np.arange(nphi) + 0.5) * 2 * math.pi / nphi
    l64 = np.asarray(l, np.float64)
    rows = []
    for beta in tb:
        acc = 0.0
        for p in ph:
            h = np.array([math.sin(beta) * math.cos(p), math.sin(beta) * math.sin(p), math.cos(beta)])
            v = 2 * np.dot(l64, h) * h - l64
            if v[2] <= 0:
                continue
            f, _ = _f(o, l, _unit(v))
            acc += f[0] * v[2] * 4 * np.dot(v, h)
        rows.append(acc * (2 * math.pi / nphi) * math.sin(beta))
    return float(np.trapezoid(rows, tb))


@pytest.mark.parametrize("kind,name", KINDS)
def test_glossy_energy_bound(kind, name):
    """White albedo (F = 1): int f cos(theta_v) dv <= 1 (no energy created); the
    reference's f is the microfacet BRDF F D G / (4 n.l n.v) with D = D_std cos(beta),
    so the bound holds and, for a near-mirror lobe at normal incidence, the
    integral is close to 1 (a dropped 4 or doubled D is off by 4x / 2x)."""
    for s in (50.0, 1000.0):
        o = _glossy_scene(kind, s)
        for th in (0.0, 0.6, 1.2):
            l = _unit([math.sin(th), 0.0, math.cos(th)])
            e = _albedo_integral(o, kind, s, l)
            assert e <= 1.0 + 2e-3, (name, s, th, e)
            if s == 1000.0 and th == 0.0:
                assert e > 0.95, (name, e)


def _beckmann_g1(a):
    """Exact Smith G1 of the Beckmann distribution (Walter et al. 2007, eq. 25)."""
    return 2.0 / (1.0 + special.erf(a) + math.exp(-a * a) / (a * math.sqrt(math.pi)))


def test_blinn_phong_G1_is_the_beckmann_approximation():
    """The Blinn-Phong G1 is Walter et al.'s rational fit of the Beckmann Smith
    G1 at a = sqrt(s / 2 + 1) / tan(theta) (material.cpp:37-56); it must track the
    exact erf form to the fit's accuracy, be continuous at a = 1.6 and reach 1
    at normal incidence."""
    for s in (1.0, 20.0, 300.0):
        for th in np.linspace(0.05, 1.5, 40):
            w = np.array([math.sin(th), 0, math.cos(th)], dtype=np.float32)
            # G(l = n, v = w) = G1(n) G1(w) = G1(w)
            _, g = O.micro_terms(O.BLINN, s, N_Z, N_Z, w, N_Z)
            a = math.sqrt(0.5 * s + 1) * math.cos(th) / math.sin(th)
            assert abs(g - _beckmann_g1(a)) < 6e-3, (s, th, g, _beckmann_g1(a))
    _, g = O.micro_terms(O.BLINN, 10.0, N_Z, N_Z, N_Z, N_Z)
    assert g == 1.0
    a_lo, a_hi = 1.6 - 1e-4, 1.6 + 1e-4
    fit = lambda a: (3.535 * a + 2.181 * a * a) / (1 + 2.276 * a + 2.577 * a * a)  # noqa: E731
    assert abs(fit(a_lo) - 1.0) < 1e-3 and fit(a_hi) == fit(a_hi)


@pytest.mark.parametrize("s", [1.0, 20.0, 300.0])
def test_ggx_G1_satisfies_the_smith_projection_identity(s):
    """Heitz 2014 (eq. 2 / 17): with the Smith masking term, the visible projected
    microfacet area equals the macro surface's, int G1(v) max(0, v.h) D_std(h) dh =
    n.v, where D_std = D / cos(beta) is the standard GGX density.  Exact for the
    GGX G1 2 / (1 + sqrt(1 + alpha^2 tan^2)) (material.cpp:77-86)."""
    c = _cos_grid(3000)
    beta = np.arccos(c)
    nphi = 64
    ph = (np.arange(nphi) + 0.5) * 2 * math.pi / nphi
    for th in (0.2, 0.8, 1.3):
        v = np.array([math.sin(th), 0, math.cos(th)], dtype=np.float32)
        _, g1 = O.micro_terms(O.GGX, s, N_Z, N_Z, v, N_Z)
        rows = []
        for cb, b in zip(c, beta):
            D, _ = O.micro_terms(O.GGX, s, N_Z, N_Z, N_Z, _h_at(cb))
            if cb <= 0:
                rows.append(0.0)
                continue
            hx = np.sin(b) * np.cos(ph)
            vh = np.maximum(0.0, v[0] * hx + v[2] * cb)
            rows.append(D / cb * vh.mean() * 2 * math.pi)
        proj = g1 * float(np.trapezoid(rows, c))
        assert abs(proj - math.cos(th)) < 3e-3 * max(1.0, math.cos(th)), (s, th, proj, math.cos(th))


# ---------------------------------------------------------------- sphere light: cone sampling
def _sphere_light_scene(center, radius):
    import tinypt_amd as T

    s = T.Sphere()
    s.location = list(center)
    s.radius = radius
    s.material = _mat(T, emission=(1, 1, 1))
    return O.OracleScene(_scene([s]))


def _light_samples(o, pos, N):
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    dirs = np.empty((N, 3), dtype=np.float32)
    pdfs = np.empty(N, dtype=np.float64)
    d = np.zeros(3, dtype=np.float32)
    p = ctypes.c_float()
    for i in range(N):
        O.lib().oracle_light_sample(o.h, 0, pos.ctypes.data, 7919 * i + 1, d.ctypes.data, ctypes.byref(p))
        dirs[i] = d
        pdfs[i] = p.value
    return dirs.astype(np.float64), pdfs


def test_sphere_light_cone_sampling_and_pdf():
    """From outside: directions fill the cone of half-angle theta_max,
    cos(theta_max) = sqrt(1 - r^2 / d^2), with cos(theta) uniform on
    [cos(theta_max), 1] and uniform azimuth, and every sample's pdf is
    1 / (2 pi (1 - cos(theta_max))) (object.cpp:66-104).  From inside: the whole
    sphere of directions, pdf 1 / (4 pi)."""
    center, r = np.array([3.0, -2.0, 5.0]), 1.5
    o = _sphere_light_scene(center, r)
    pos = np.array([0.5, 0.25, -1.0])
    d = np.linalg.norm(center - pos)
    ctm = math.sqrt(1 - r * r / (d * d))
    N = 6000
    dirs, pdfs = _light_samples(o, pos, N)
    axis = (center - pos) / d
    ct = dirs @ axis
    assert ct.min() >= ctm - 2e-6
    assert stats.kstest((ct - ctm) / (1 - ctm), "uniform").pvalue > 1e-3
    t = np.cross(axis, [0.0, 0.0, 1.0])
    t /= np.linalg.norm(t)
    b = np.cross(axis, t)
    ph = np.arctan2(dirs @ b, dirs @ t)
    assert stats.kstest((ph + math.pi) / (2 * math.pi), "uniform").pvalue > 1e-3
    hit = pdfs > 0
    assert hit.mean() > 0.995  # only rays grazing the silhouette may miss in FP32
    want = 1 / (2 * math.pi * (1 - ctm))
    assert np.allclose(pdfs[hit], want, rtol=2e-4)
    # from inside the sphere
    dirs, pdfs = _light_samples(o, center + [0.2, 0.1, -0.3], 4000)
    assert np.allclose(pdfs, 1 / (4 * math.pi), rtol=1e-5)
    assert np.all(np.abs(dirs.mean(0)) < 0.05)
    assert stats.kstest((dirs[:, 2] + 1) / 2, "uniform").pvalue > 1e-3


# ---------------------------------------------------------------- camera and thin lens
def _camera_scene(w, h, fov_deg, forward, loc=(1.0, 2.0, 3.0), aperture=0.0, focus=1.0):
    import tinypt_amd as T
    from tinypt_amd import scenes

    cam = T.Camera()
    cam.location = list(loc)
    cam.rotation = scenes.lookat_rotation(forward)
    cam.width, cam.height = w, h
    cam.fov = math.radians(fov_deg)
    cam.aperture = aperture
    cam.focus_dist = focus
    return O.OracleScene(T.Scene(cam, [], [], None)), np.array(cam.rotation, dtype=np.float64)


def _angle(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return math.atan2(np.linalg.norm(np.cross(a, b)), np.dot(a, b))


@pytest.mark.parametrize("w,h", [(64, 48), (48, 64)])
def test_camera_rays_span_the_field_of_view(w, h):
    """The film centre (W/2, H/2) looks along -w; the field of view fov spans
    the longer axis (view = 2 tan(fov / 2) there, scaled by the aspect on the
    other, camera.cpp:6-21); x grows along u and y along v (row 0 is the bottom)."""
    fov = 50.0
    o, R = _camera_scene(w, h, fov, (0.3, 1.0, -0.2))
    u, v, wv = R[:, 0], R[:, 1], R[:, 2]
    org, d = O.camera_ray(o, w / 2, h / 2)
    assert np.allclose(org, [1, 2, 3])
    assert np.allclose(d, -wv, atol=2e-7)
    long_ = math.radians(fov)
    short = 2 * math.atan(math.tan(long_ / 2) * min(w, h) / max(w, h))
    hor = _angle(O.camera_ray(o, 0, h / 2)[1], O.camera_ray(o, w, h / 2)[1])
    ver = _angle(O.camera_ray(o, w / 2, 0)[1], O.camera_ray(o, w / 2, h)[1])
    assert abs(hor - (long_ if w > h else short)) < 2e-6
    assert abs(ver - (short if w > h else long_)) < 2e-6
    assert np.dot(O.camera_ray(o, w, h / 2)[1], u) > 0.3
    assert np.dot(O.camera_ray(o, w / 2, h)[1], v) > 0.3
    # a corner ray makes equal angles with its two edge neighbours' planes: it is the
    # normalised sum of the two half-extents and -w
    _, dc = O.camera_ray(o, w, h)
    half_w = math.tan(math.radians(fov) / 2) * (1 if w > h else w / h)
    half_h = math.tan(math.radians(fov) / 2) * (h / w if w > h else 1)
    want = _unit(half_w * u + half_h * v - wv)
    assert np.allclose(dc, want, atol=2e-6)


def test_thin_lens_rays_meet_on_the_focal_plane():
    """aperture A, focus distance F: the origins lie in the lens disk (radius A
    in the u-v plane, uniform: E|o|^2 = A^2 / 2), and all rays of one film point
    pass through one point at distance F along -w (camera.cpp:23-34)."""
    A, F = 0.5, 7.0
    o, R = _camera_scene(64, 48, 40.0, (1.0, 0.2, -0.1), aperture=A, focus=F)
    u, v, wv = R[:, 0], R[:, 1], R[:, 2]
    c = np.array([1.0, 2.0, 3.0])
    for x, y in ((32, 24), (5.5, 40.25)):
        pts, offs = [], []
        for seed in range(400):
            org, d = O.camera_ray(o, x, y, seed)
            off = org.astype(np.float64) - c
            offs.append(off)
            s = F / np.dot(d, -wv)
            pts.append(org + s * d)
        offs, pts = np.array(offs), np.array(pts)
        assert np.allclose(offs @ wv, 0, atol=1e-6)
        assert np.all(np.linalg.norm(offs, axis=1) <= A * (1 + 1e-6))
        assert abs(np.mean(np.sum(offs * offs, 1)) / (A * A / 2) - 1) < 0.12
        assert np.all(np.linalg.norm(pts - pts.mean(0), axis=1) < 2e-5 * F)
        # the common point is the pinhole ray's point at distance F
        _, d0 = _camera_scene(64, 48, 40.0, (1.0, 0.2, -0.1))[0], None
        assert np.dot(pts.mean(0) - c, -wv) == pytest.approx(F, rel=1e-5)


# ---------------------------------------------------------------- bump mapping
def test_bump_normal_on_a_planar_ramp():
    """Height map b(x, y) = kx x + ky y texels on a square of side L with uv
    spanning it: the Sobel gradient (scale / 8) is (kx s, ky s) per texel inside,
    dp/du = (L, 0, 0), dp/dv = (0, L, 0), so bump_normal tilts the normal to
    normalize(n - (L kx s, L ky s, 0) / (sqrt(L L) + EPS)) (texture.cpp:52-63);
    against the slope, with no swap of u and v."""
    import tinypt_amd as T

    L, kx, ky, scale = 4.0, 0.01, -0.004, 8.0
    W = 32
    yy, xx = np.mgrid[0:W, 0:W]
    hmap = (kx * xx + ky * yy).astype(np.float32)
    verts = np.array([[0, 0, 0], [L, 0, 0], [L, L, 0], [0, L, 0]], dtype=np.float32)
    faces = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32)
    uvs = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=np.float32)
    m = T.Material(T.Lambertian(T.RGBTexture([0.5, 0.5, 0.5])))
    m.bump_texture = T.BumpTexture(T.Image.from_linear(hmap), scale)
    faces9 = np.stack([faces[:, 0], faces[:, 0], faces[:, 0], faces[:, 1], faces[:, 1], faces[:, 1],
                       faces[:, 2], faces[:, 2], faces[:, 2]], 1).astype(np.int32)
    faces9[:, [1, 4, 7]] = -1
    mesh = T.TriangleMesh.from_arrays(verts, faces9, [m], np.zeros(2, np.int32), None, uvs)
    o = O.OracleScene(_scene([mesh]))
    EPS = 1e-3
    want = _unit([-L * kx * scale / (L + EPS), -L * ky * scale / (L + EPS), 1.0])
    for px, py in ((1.3, 1.1), (2.9, 0.7), (0.8, 3.1)):  # both triangles, away from the border texels
        prim, out = o.intersect_surface(_ray((px, py, 2.0), (0, 0, -1)))
        assert prim >= 0
        assert np.allclose(out[1:4], [0, 0, 1], atol=1e-6)  # geometric normal, facing the ray
        assert np.allclose(out[4:7], want, atol=2e-6), (out[4:7], want)
