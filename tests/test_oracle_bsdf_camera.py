"""Pins the oracle's BSDFs and cameras with closed forms, sampling-density
checks and physical laws (the reference's Sampling.cpp / material.h /
Camera.h do not compile standalone, so these are checked from the
mathematics).  A dropped term, a flipped sign or a swapped operand in the
restatement fails one of these."""
import math

import numpy as np
import pytest

import pyoracle as po
import scenes
import sppm_amd as S

PI = np.float32(3.14159)


def norm(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def sample(kind, p, seeds):
    out = np.zeros((len(seeds), 3), np.float32)
    pdf = np.zeros(len(seeds), np.float32)
    nd = po.C.c_int()
    for j, s in enumerate(seeds):
        po.orc().orc_bsdf_sample(kind, po.C.byref(p), int(s), po.fp(out[j]), po.fp(pdf[j:j + 1]), po.C.byref(nd))
    return out, pdf


def test_diffuse_brdf_and_pdf_closed_form():
    n = (0.0, 1.0, 0.0)
    rng = np.random.default_rng(0)
    for _ in range(200):
        i = norm(rng.normal(size=3))
        o = norm(rng.normal(size=3))
        p = po.fparam(i, o, n)
        b = po.orc().orc_bsdf_brdf(S.MAT_DIFFUSE, po.C.byref(p))
        ci, co = -np.float32(i[1]), np.float32(o[1])
        want = np.float32(1.0) / PI if (ci > 1e-4 and co > 1e-4) else 0.0
        assert b == np.float32(want)
        pdf = po.orc().orc_bsdf_pdf(S.MAT_DIFFUSE, po.C.byref(p))
        assert pdf == max(np.float32(1e-16), np.float32(np.float32(o[1]) / PI))


def test_cosine_sampling_density():
    # E[cos] = 2/3 and E[cos^2] = 1/2 under a cosine-weighted hemisphere
    n = norm((0.3, 0.8, -0.2))
    p = po.fparam(-n, n, n)
    d, pdf = sample(S.MAT_DIFFUSE, p, range(1, 40001))
    c = d.astype(np.float64) @ n
    assert (c > -1e-6).all()
    assert abs(c.mean() - 2 / 3) < 0.01
    assert abs((c * c).mean() - 0.5) < 0.01
    np.testing.assert_allclose(pdf, np.maximum(c / 3.14159, 1e-16), rtol=2e-5, atol=1e-7)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)


def test_mirror_is_exact_reflection():
    rng = np.random.default_rng(2)
    for s in range(50):
        nrm = norm(rng.normal(size=3))
        i = norm(rng.normal(size=3))
        if i @ nrm > 0:
            i = -i
        p = po.fparam(i, i, nrm)
        d, pdf = sample(S.MAT_METAL, p, [s + 1])
        want = i - 2 * (i @ nrm) * nrm
        np.testing.assert_allclose(d[0], want, atol=1e-6)
        assert pdf[0] == 1.0


def hemisphere_dirs(n, count, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(count, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    v[v @ n < 0] *= -1
    return v


@pytest.mark.parametrize("alpha", [0.2, 0.3, 0.5])
def test_ggx_conductor_pdf_normalisation(alpha):
    # At normal incidence o = reflect(i, h) stays above the horizon iff the
    # microfacet tilt is below 45 degrees; GGX gives P(tan^2 <= 1) =
    # 1 / (1 + alpha^2), which the pdf D(h)(h.n)/(4 o.h) must integrate to.
    n = np.array([0.0, 1.0, 0.0])
    i = np.array([0.0, -1.0, 0.0])
    dirs = hemisphere_dirs(n, 120000, 3)
    tot = 0.0
    for o in dirs:
        p = po.fparam(i, o, n, rough=alpha)
        tot += po.orc().orc_bsdf_pdf(S.MAT_ROUGH_CONDUCTOR, po.C.byref(p))
    integral = tot * 2 * math.pi / len(dirs)
    assert abs(integral - 1.0 / (1.0 + alpha * alpha)) < 0.015, integral


def test_ggx_conductor_sampling_matches_pdf():
    # sampled directions are distributed by the returned pdf: E_q[f/q] for
    # f = indicator(o.n > 0.7) equals the hemisphere integral of f
    n = np.array([0.0, 1.0, 0.0])
    i = norm((0.3, -1.0, 0.0))
    alpha = 0.3
    p = po.fparam(i, i, n, rough=alpha)
    d, pdf = sample(S.MAT_ROUGH_CONDUCTOR, p, range(1, 30001))
    ok = np.linalg.norm(d - 1e-3, axis=1) > 1e-3  # INVALID_DIR is (1e-3,1e-3,1e-3)
    d = d[ok]
    # pdf of the sample equals the PDF() evaluation at the sampled direction
    for k in range(0, len(d), 997):
        q = po.fparam(i, d[k], n, rough=alpha)
        v = po.orc().orc_bsdf_pdf(S.MAT_ROUGH_CONDUCTOR, po.C.byref(q))
        assert abs(v - pdf[ok][k]) <= 1e-4 * max(1.0, abs(v))
    frac = (d @ n > 0.9).mean() * ok.mean()
    # same probability by quadrature of the pdf
    dirs = hemisphere_dirs(n, 80000, 4)
    tot = 0.0
    for o in dirs[dirs @ n > 0.9]:
        q = po.fparam(i, o, n, rough=alpha)
        tot += po.orc().orc_bsdf_pdf(S.MAT_ROUGH_CONDUCTOR, po.C.byref(q))
    quad = tot * 2 * math.pi / len(dirs)
    assert abs(frac - quad) < 0.02, (frac, quad)


def test_dielectric_fresnel_normal_incidence_and_tir():
    # smooth glass: reflect probability = Fresnel reflectance; at normal
    # incidence ((n1-n2)/(n1+n2))^2 = 0.04 for 1 -> 1.5
    n = np.array([0.0, 1.0, 0.0])
    i = np.array([0.0, -1.0, 0.0])
    p = po.fparam(i, i, n, fi=1.0, ti=1.5)
    d, pdf = sample(S.MAT_GLASS, p, range(1, 20001))
    refl = (d @ n > 0)
    assert abs(refl.mean() - 0.04) < 0.006
    np.testing.assert_allclose(pdf[refl], 0.04, rtol=0.02)
    # grazing from inside glass beyond the critical angle: always reflects
    th = math.radians(60)
    i2 = np.array([math.sin(th), math.cos(th), 0.0])  # travelling up inside, normal facing up
    p2 = po.fparam(i2, i2, n, fi=1.5, ti=1.0)
    d2, _ = sample(S.MAT_GLASS, p2, range(1, 200))
    assert (d2 @ n < 0).all()


def test_glass_refraction_obeys_snell():
    n = np.array([0.0, 1.0, 0.0])
    for deg in (10, 30, 50, 70):
        th = math.radians(deg)
        i = np.array([math.sin(th), -math.cos(th), 0.0])
        p = po.fparam(i, i, n, fi=1.0, ti=1.5)
        d, _ = sample(S.MAT_GLASS, p, range(1, 400))
        t = d[d @ n < 0]
        assert len(t) > 0
        sin_t = np.sqrt(1 - (t @ n) ** 2)
        np.testing.assert_allclose(sin_t, math.sin(th) / 1.5, atol=2e-6)


def test_rough_glass_tends_to_snell_for_small_roughness():
    n = np.array([0.0, 1.0, 0.0])
    th = math.radians(40)
    i = np.array([math.sin(th), -math.cos(th), 0.0])
    p = po.fparam(i, i, n, fi=1.0, ti=1.5, rough=0.01)
    d, _ = sample(S.MAT_GGX_GLASS, p, range(1, 3000))
    ok = np.linalg.norm(d - 1e-3, axis=1) > 1e-3
    t = d[ok & (d @ n < 0)]
    sin_t = np.sqrt(1 - (t @ n) ** 2)
    assert abs(np.median(sin_t) - math.sin(th) / 1.5) < 0.01


def test_pinhole_camera_geometry():
    cam = scenes.pinhole_camera(64, 64)
    # centre pixel with zero jitter looks along the camera direction
    o, d, pdf = po.camera_ray(cam, 0.5, 0.5, [0.0, 0.0, 0.0])
    want = norm(np.subtract(cam["target"], cam["pos"]))
    np.testing.assert_allclose(d, want, atol=1e-6)
    assert pdf == 1.0
    # every ray passes through the pinhole (camera position)
    rng = np.random.default_rng(5)
    for _ in range(50):
        u, v = rng.random(2)
        o, d, _ = po.camera_ray(cam, u, v, rng.random(3))
        t = np.dot(np.subtract(cam["pos"], o), d)
        np.testing.assert_allclose(o + t * d, cam["pos"], atol=1e-5)
    # the half field of view is PI*Fov/360: the image corner ray makes that angle vertically
    o, d, _ = po.camera_ray(cam, 0.5, 0.0, [0.0, 0.0, 0.0])
    ang = math.degrees(math.acos(float(np.dot(d, want))))
    assert abs(ang - 22.5) < 1e-3


def lens_uv(cam, o):
    """Lens (u, v) of a thin-lens ray origin: o = pos + (u - 1/2) e1 LW + (v - 1/2) e2 LH
    (Camera.h:34-53 for the basis and the lens size)."""
    pos = np.asarray(cam["pos"], np.float64)
    d = norm(np.subtract(cam["target"], cam["pos"]))
    e1 = norm(np.cross([0.0, 1.0, 0.0], d))
    e2 = norm(np.cross(d, e1))
    half = np.float32(np.pi) * np.float32(cam["fov_deg"]) / np.float32(360.0)
    LH = 2.0 * cam["dist_to_lens"] * math.tan(float(half))
    LW = LH * cam["aspect"]
    r = np.asarray(o, np.float64) - pos
    return r @ e1 / LW + 0.5, r @ e2 / LH + 0.5


def test_diaphragm_single_open_pixel_fixes_the_lens_point():
    """ThinLensCamera with a diaphragm image (Camera.h:106-111, 136-158):
    with one open pixel every lens sample lands on that pixel's lens point
    ((i % W) / W, 1 - (i / W) / H) -- pins the pixel order and orientation."""
    W, H = 5, 3
    base = scenes.thinlens_camera(32, 32, aperture=0.05, focus=1.5)
    rng = np.random.default_rng(8)
    for i in (0, 4, 7, 10, 14):
        img = np.ones((H, W, 4), np.float32)
        img.reshape(-1, 4)[i, 0] = 0.0
        cam = dict(base, diaphragm=img)
        want = ((i % W) / W, 1.0 - (i // W) / H)
        for _ in range(8):
            o, d, pdf = po.camera_ray(cam, *rng.random(2), rng.random(3))
            u, v = lens_uv(cam, o)
            assert abs(u - want[0]) < 1e-4 and abs(v - want[1]) < 1e-4, (i, u, v, want)
            assert pdf > 0


def test_diaphragm_pick_frequencies_follow_one_minus_red():
    """Lens pixels are picked with probability (1 - red) / sum (1 - red): a
    deterministic grid of sampler pairs (x through every table bin, y over
    [0, 1)) against the weights (alpha, green and blue are ignored)."""
    W, H = 8, 4
    rng = np.random.default_rng(9)
    img = rng.random((H, W, 4)).astype(np.float32)
    img[0, :3, 0] = 0.0  # holes
    img[2, 5, 0] = 1.0   # closed
    n = W * H
    cam = dict(scenes.thinlens_camera(32, 32, aperture=0.05, focus=1.5), diaphragm=img)
    B = 400
    counts = np.zeros(n)
    for a in range(2 * n):
        for b in range(B):
            o, _, _ = po.camera_ray(cam, 0.5, 0.5, [(a + 0.5) / (2 * n), (b + 0.5) / B, 0.0])
            u, v = lens_uv(cam, o)
            x, y = int(round(u * W)), int(round((1.0 - v) * H))
            counts[y * W + x] += 1
    freq = counts / counts.sum()
    w = 1.0 - img[..., 0].reshape(-1).astype(np.float64)
    assert freq[2 * W + 5] == 0.0
    np.testing.assert_allclose(freq, w / w.sum(), atol=4.0 / (n * B))


def test_diaphragm_open_image_keeps_plain_lens_sampling():
    """m_HoleRatio > 0.999 (holes = pixels with red < 1e-6, ratio in float)
    keeps the plain lens sampling: all open, and 999 of 1000 open (0.999f is
    above 0.999); 998 of 1000 open uses the table."""
    base = scenes.thinlens_camera(32, 32, aperture=0.05, focus=1.5)
    rng = np.random.default_rng(10)
    smp = [rng.random(3) for _ in range(6)]
    plain = [po.camera_ray(base, 0.3, 0.6, s) for s in smp]
    for closed, same in ((0, True), (1, True), (2, False)):
        img = np.zeros((20, 50, 4), np.float32)
        img.reshape(-1, 4)[:closed, 0] = 1.0
        got = [po.camera_ray(dict(base, diaphragm=img), 0.3, 0.6, s) for s in smp]
        eq = all(np.array_equal(g[0], p_[0]) and np.array_equal(g[1], p_[1]) for g, p_ in zip(got, plain))
        assert eq == same, closed


def test_diaphragm_without_open_pixel_is_refused():
    img = np.ones((4, 4, 4), np.float32)
    with pytest.raises(ValueError):
        po.camera_ray(dict(scenes.thinlens_camera(32, 32), diaphragm=img), 0.5, 0.5, [0.5, 0.5, 0.5])


def test_thinlens_focus_and_pdf():
    cam = scenes.thinlens_camera(64, 64, aperture=0.02, focus=2.0)
    rng = np.random.default_rng(6)
    for _ in range(20):
        u, v = rng.random(2)
        pts = []
        for _ in range(4):
            o, d, pdf = po.camera_ray(cam, u, v, rng.random(3))
            # all lens samples of one image point meet on the focal plane (distance 2 along the axis)
            axis = norm(np.subtract(cam["target"], cam["pos"]))
            t = (2.0 - np.dot(o - np.asarray(cam["pos"]), axis)) / np.dot(d, axis)
            pts.append(o + t * d)
            assert pdf > 0
        pts = np.array(pts)
        assert np.abs(pts - pts.mean(0)).max() < 1e-4
