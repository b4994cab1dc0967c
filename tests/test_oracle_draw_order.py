"""Pins which RandomMT draw feeds which variable in the oracle, per routine and
along the photon and eye paths, with scripted draws (orc_script_set) and
closed-form expectations.  The HIP path is compared with the oracle bit for
bit on whole paths (tests/test_gpu_parity.py), so this pins it too.

Draw order of the reference (statement order; SPPM/SPPM/ of the reference):
  light pick           real2 -> bin, real1 -> coin      TableSampler.h:126-129
  PointOnTriangle      u, v                             Sampling.cpp:340-341
  Cos_Sampling         ph (azimuth), th                 Sampling.cpp:20-21
  photon RR            1 draw at a surface, before the
                       photon is stored                 SPPM.cpp:580-583, SPPM.h:33-52
  GGX_Sampling         phi, e                           Sampling.cpp:127-128
  GGXGlass_Sampling    phi, e, then the Fresnel draw    Sampling.cpp:191-192, 218
  GlassSurface         1 draw (reflect if F > draw)     material.h:486
  eye RR               1 draw only at depth 5 and 6     SPPM.cpp:113-122
"""
import math

import numpy as np
import pytest

import pyoracle as po
import scenes
import sppm_amd as S

PI = 3.14159  # PTUtility::PI (Vec.h:10)


def two_light_scene():
    """Box walls, two untextured ceiling lights of different emission: the
    light pick table has bins whose coin matters."""
    attrs = [
        dict(name="white", kind=S.MAT_DIFFUSE, albedo=(0.73, 0.73, 0.73), is_sppm=1, if_opaque=1),
        dict(name="red", kind=S.MAT_DIFFUSE, albedo=(0.63, 0.06, 0.05), is_sppm=1, if_opaque=1),
        dict(name="green", kind=S.MAT_DIFFUSE, albedo=(0.14, 0.45, 0.09), is_sppm=1, if_opaque=1),
        dict(name="lightA", kind=S.MAT_DIFFUSE, albedo=(0.78, 0.78, 0.78), radiance=(15.0, 15.0, 15.0), is_sppm=1,
             if_opaque=1),
        dict(name="lightB", kind=S.MAT_DIFFUSE, albedo=(0.78, 0.78, 0.78), radiance=(4.0, 4.0, 4.0), is_sppm=1,
             if_opaque=1),
    ]
    b = scenes._Builder()
    scenes.box_walls(b, 0, 1, 2)
    scenes.light_quad(b, 3, size=0.2, cx=0.3, cz=0.3)
    scenes.light_quad(b, 4, size=0.3, cx=0.7, cz=0.6)
    return b.scene(attrs)


def render_for(scene, N=16):
    os_ = po.OracleScene(scene)
    r = po.OracleRender(os_, scenes.pinhole_camera(8, 8), 0.9, 0.015, N, nthreads=1)
    return os_, r


def run_path(r, draws, step=1, path=0):
    with po.script(draws) as s:
        res = r.photon_path(step, path)
    return res, s.used


def in_triangle(p, tri, tol=1e-5):
    a, b, c = (np.asarray(tri[3 * k:3 * k + 3], np.float64) for k in range(3))
    v0, v1, v2 = b - a, c - a, np.asarray(p, np.float64) - a
    d00, d01, d11 = v0 @ v0, v0 @ v1, v1 @ v1
    d20, d21 = v2 @ v0, v2 @ v1
    den = d00 * d11 - d01 * d01
    v = (d11 * d20 - d01 * d21) / den
    w = (d00 * d21 - d01 * d20) / den
    off = abs(np.cross(v0, v1) @ v2) / np.linalg.norm(np.cross(v0, v1))
    return v >= -tol and w >= -tol and v + w <= 1 + tol and off < 2e-4


KILL = 0.99  # above every 0.98 * max(albedo): the path ends at its first surface


def test_light_pick_draws_bin_then_coin():
    sc = two_light_scene()
    os_, r = render_for(sc)
    L = os_.lights()
    n = len(L["facet"])
    assert n == 4
    prob, left, right = L["prob"], L["left"], L["right"]
    seen = set()
    for bin_ in range(n):
        for coin in (0.02, 0.98):
            d_bin = (bin_ + 0.5) / n
            want = left[bin_] if coin < prob[bin_] else right[bin_]
            if want < 0:
                want = left[bin_]
            res, used = run_path(r, [d_bin, coin, 0.3, 0.6, 0.1, 0.4, KILL])
            assert res["k"] == 0 and used == 7
            f = L["facet"][want]
            assert in_triangle(res["pos"][0], sc.pos[f]), (bin_, coin)
            seen.add(int(want))
            # the same two draws the other way round pick by the other rule
            swapped = left[int(coin * n)] if d_bin < prob[int(coin * n)] else right[int(coin * n)]
            swapped = left[int(coin * n)] if swapped < 0 else swapped
            if swapped != want:
                res2, _ = run_path(r, [coin, d_bin, 0.3, 0.6, 0.1, 0.4, KILL])
                assert not in_triangle(res2["pos"][0], sc.pos[f])
    assert len(seen) >= 3  # the coin changes the pick for some bins


def test_point_on_triangle_draws_u_then_v():
    sc = two_light_scene()
    os_, r = render_for(sc)
    L = os_.lights()
    # bin 0, coin 0.01: the light the alias table gives for these draws
    li = L["left"][0] if 0.01 < L["prob"][0] else L["right"][0]
    li = L["left"][0] if li < 0 else li
    f = L["facet"][li]
    P = sc.pos[f].astype(np.float64).reshape(3, 3)
    nrm = np.array([0.0, -1.0, 0.0])
    for u, v in ((0.8, 0.3), (0.3, 0.8), (0.55, 0.05)):
        res, used = run_path(r, [0.01, 0.01, u, v, 0.2, 0.3, KILL])
        assert used == 7
        uu, vv = u, v
        if vv > uu:
            uu *= 0.5
            vv -= uu
        else:
            vv *= 0.5
            uu -= vv
        want = uu * P[0] + vv * P[1] + (1 - uu - vv) * P[2] + 1e-4 * nrm
        np.testing.assert_allclose(res["pos"][0], want, atol=2e-6)


def test_cosine_emission_draws_azimuth_then_elevation():
    sc = scenes.box_scene(with_spheres=False)
    os_, r = render_for(sc)
    L = os_.lights()
    area = 0.5 * 0.25 * 0.25
    # frame of Cos_Sampling about n = (0,-1,0): e2 = (1,0,0) x n normalised, e3 = n x e2
    n = np.array([0.0, -1.0, 0.0])
    e2 = np.cross([1.0, 0.0, 0.0], n)
    e2 /= np.linalg.norm(e2)
    e3 = np.cross(n, e2)
    for a, b in ((0.25, 0.36), (0.36, 0.25), (0.9, 0.05)):
        res, used = run_path(r, [0.2, 0.5, 0.4, 0.4, a, b, KILL])
        assert used == 7 and res["k"] == 0
        ph = a * 2.0 * PI
        want = n * math.sqrt(1 - b) + e2 * math.cos(ph) * math.sqrt(b) + e3 * math.sin(ph) * math.sqrt(b)
        np.testing.assert_allclose(res["indir"][0], want, atol=2e-6)
        # PDF = (1 / area) * P(pick) * cos / PI  (SPPM.cpp:460-466)
        pick = np.linalg.norm(L["flux"][0]) / L["avg"] / len(L["facet"])
        np.testing.assert_allclose(res["pdf"][0], (1 / area) * pick * math.sqrt(1 - b) / PI, rtol=2e-5)


def test_photon_rr_draw_at_surface_before_store():
    sc = scenes.box_scene(with_spheres=False)
    os_, r = render_for(sc)
    emit = [0.2, 0.5, 0.4, 0.4, 0.3, 0.5]
    # RR fails at the first surface: nothing stored, exactly one draw after emission
    res, used = run_path(r, emit + [KILL, 0.5, 0.5, 0.5])
    assert res["k"] == 0 and used == 7
    # RR survives: the surface photon is stored with PDF x 0.98 max(albedo)
    # (SPPM.cpp:583, 599-603) and the BSDF draws follow
    res0, _ = run_path(r, emit + [KILL])
    res, used = run_path(r, emit + [0.0, 0.3, 0.5, KILL, KILL, KILL])
    assert res["k"] >= 1 and used >= 9
    np.testing.assert_array_equal(res["indir"][0], res0["indir"][0])
    p = res["pos"][0]
    # the first surface is a white wall or floor (albedo 0.73) or a coloured side wall
    if p[0] < 1e-3:
        alb = 0.63
    elif p[0] > 1 - 1e-3:
        alb = 0.45
    else:
        alb = 0.73
    np.testing.assert_allclose(res["pdf"][0], res0["pdf"][0] * np.float32(0.98 * alb), rtol=1e-6)


def mirror_corridor():
    """Two facing mirror planes (z = 0 and z = 1, MetalSurface, no hit points)
    and a far small light: an eye ray along z bounces between them forever."""
    attrs = [
        dict(name="mirror", kind=S.MAT_METAL, albedo=(0.9, 0.9, 0.9), if_delta=1, is_sppm=0, if_opaque=1),
        dict(name="light", kind=S.MAT_DIFFUSE, albedo=(0.78, 0.78, 0.78), radiance=(15.0, 15.0, 15.0), is_sppm=1,
             if_opaque=1),
    ]
    b = scenes._Builder()
    b.quad((-50, -50, 0), (-50, 50, 0), (50, 50, 0), (50, -50, 0), (0, 0, 1), 0)
    b.quad((-50, -50, 1), (50, -50, 1), (50, 50, 1), (-50, 50, 1), (0, 0, -1), 0)
    scenes.light_quad(b, 1, size=0.2, y=60.0)
    return b.scene(attrs)


def test_eye_rr_draws_only_at_depth_5_and_6():
    sc = mirror_corridor()
    os_ = po.OracleScene(sc)
    cam = dict(type=S.CAM_PINHOLE, width=8, height=8, pos=(0.5, 0.5, 0.5), target=(0.5, 0.5, 0.0), fov_deg=10.0,
               F=1.0, dist_to_lens=0.01, aspect=1.0, sampler=S.SAMPLER_UNIFORM, sampler_id=0, sampler_nsample=1)
    r = po.OracleRender(os_, cam, 0.9, 0.015, 16, nthreads=1)
    pix = [4 * 8 + 4]
    smp = np.zeros((1, 3), np.float32)
    rus5, rus6 = math.exp(-1.0), math.exp(-2.0)
    cases = [
        ([0.0, 0.0], 2),               # survives both rouletes, stops at depth 7
        ([0.99], 1),                    # stopped at depth 5
        ([rus5 * 0.99, 0.99], 2),       # survives depth 5, stopped at depth 6
        ([rus5 * 1.01], 1),             # just above exp(-1): stopped at depth 5
        ([rus5 * 0.99, rus6 * 1.01], 2),
    ]
    for draws, want in cases:
        with po.script(draws + [0.5] * 4) as s:
            h = r.eye_pixels(1, pix, smp)
        assert h["type"][0] == 0
        assert s.used == want, (draws, s.used)


def bsdf_scripted(kind, p, draws):
    out = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    nd = po.C.c_int()
    with po.script(draws) as s:
        po.orc().orc_bsdf_sample(kind, po.C.byref(p), 1, po.fp(out), po.fp(pdf), po.C.byref(nd))
    return out.astype(np.float64), float(pdf[0]), s.used


def microfacet_angles(h):
    """Tilt cosine and azimuth of h in Cos/GGX_Sampling's frame about n = +y:
    e2 = (0,0,1), e3 = (1,0,0); h = e3 sth cph + e2 sth sph + n cth."""
    h = h / np.linalg.norm(h)
    return h[1], math.atan2(h[2], h[0]) % (2 * math.pi)


@pytest.mark.parametrize("alpha", [0.3, 0.6])
def test_ggx_draws_azimuth_then_tilt(alpha):
    n = (0.0, 1.0, 0.0)
    i = np.array([0.0, -1.0, 0.0])
    p = po.fparam(i, i, n, rough=alpha)
    for a, b in ((0.1, 0.3), (0.3, 0.1), (0.7, 0.45)):
        o, _, used = bsdf_scripted(S.MAT_ROUGH_CONDUCTOR, p, [a, b])
        assert used == 2
        cth, az = microfacet_angles(o - i)
        want_c = math.sqrt((1.0 - b) / (b * (alpha * alpha - 1.0) + 1.0))
        assert abs(cth - want_c) < 2e-5, (a, b)
        assert abs(az - (a * 2 * PI) % (2 * math.pi)) < 2e-4


def test_ggx_glass_draws_azimuth_tilt_then_fresnel():
    n = (0.0, 1.0, 0.0)
    i = np.array([0.0, -1.0, 0.0])
    alpha = 0.3
    p = po.fparam(i, i, n, fi=1.0, ti=1.5, rough=alpha)
    for a, b in ((0.15, 0.2), (0.2, 0.15)):
        # Fresnel draw below F (~0.04): reflection, whose half vector is the microfacet normal
        o, _, used = bsdf_scripted(S.MAT_GGX_GLASS, p, [a, b, 0.001])
        assert used == 3 and o[1] > 0
        cth, az = microfacet_angles(o - i)
        want_c = math.sqrt((1.0000001 - b) / (b * (alpha * alpha - 0.99999) + 1.0))
        assert abs(cth - want_c) < 2e-5
        assert abs(az - (a * 2 * PI) % (2 * math.pi)) < 2e-4
        # the same two draws with a Fresnel draw above F: refraction into the glass
        o2, _, used2 = bsdf_scripted(S.MAT_GGX_GLASS, p, [a, b, 0.5])
        assert used2 == 3 and o2[1] < 0


def test_smooth_glass_one_draw_decides_reflection():
    n = (0.0, 1.0, 0.0)
    i = np.array([0.0, -1.0, 0.0])
    p = po.fparam(i, i, n, fi=1.0, ti=1.5)
    o, pdf, used = bsdf_scripted(S.MAT_GLASS, p, [0.01])
    assert used == 1 and o[1] > 0 and abs(pdf - 0.04) < 1e-3
    o, pdf, used = bsdf_scripted(S.MAT_GLASS, p, [0.5])
    assert used == 1 and o[1] < 0 and abs(pdf - 0.96) < 1e-3


def test_transmit_surfaces_choice_draw_first():
    n = (0.0, 1.0, 0.0)
    i = np.array([0.0, -1.0, 0.0])
    # DiffuseTransmit (material.h:651-665): choice draw, then Cos_Sampling (ph, th)
    p = po.fparam(i, i, n, refl=0.3)
    o, _, used = bsdf_scripted(S.MAT_DIFFUSE_TRANSMIT, p, [0.1, 0.25, 0.0])
    assert used == 3
    np.testing.assert_allclose(o, [0.0, 1.0, 0.0], atol=1e-6)  # th = 0: along the reflected side's normal
    o, _, used = bsdf_scripted(S.MAT_DIFFUSE_TRANSMIT, p, [0.9, 0.25, 0.0])
    assert used == 3
    np.testing.assert_allclose(o, [0.0, -1.0, 0.0], atol=1e-6)
    # GlossTransmit (material.h:698-717): choice draw, then GGX (phi, e)
    p = po.fparam(i, i, n, refl=0.5, rough=0.3)
    o, _, used = bsdf_scripted(S.MAT_GLOSS_TRANSMIT, p, [0.2, 0.1, 0.3])
    assert used == 3 and o[1] > 0
    cth, _ = microfacet_angles(o - i)
    assert abs(cth - math.sqrt((1.0 - 0.3) / (0.3 * (0.09 - 1.0) + 1.0))) < 2e-5
    o, _, used = bsdf_scripted(S.MAT_GLOSS_TRANSMIT, p, [0.8, 0.1, 0.3])
    assert used == 3 and o[1] < 0
    # mirrors draw nothing
    for kind in (S.MAT_METAL, S.MAT_BIMETAL):
        _, _, used = bsdf_scripted(kind, po.fparam(i, i, n), [0.5])
        assert used == 0
