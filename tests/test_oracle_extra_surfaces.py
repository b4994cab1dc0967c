"""Pins the oracle's BiMetalSurface, DiffuseTransmitSurface and
GlossTransmitSurface (material.h:427-463,646-766) with closed forms,
normalisation of their densities, symmetries between their reflected and
transmitted halves and agreement with the already pinned MetalSurface and
Roughconductor, so that a flipped normal, a swapped share or a wrong mirror
fails one of them."""
import math

import numpy as np
import pytest

import pyoracle as po
import sppm_amd as S

PI = np.float32(3.14159)
SMIN = np.float32(1e-16)


def norm(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def brdf(kind, p):
    return po.orc().orc_bsdf_brdf(kind, po.C.byref(p))


def pdf(kind, p):
    return po.orc().orc_bsdf_pdf(kind, po.C.byref(p))


def sample(kind, p, seeds):
    out = np.zeros((len(seeds), 3), np.float32)
    pd = np.zeros(len(seeds), np.float32)
    nd = po.C.c_int()
    draws = []
    for j, s in enumerate(seeds):
        po.orc().orc_bsdf_sample(kind, po.C.byref(p), int(s), po.fp(out[j]), po.fp(pd[j:j + 1]), po.C.byref(nd))
        draws.append(nd.value)
    return out, pd, np.array(draws)


def sphere_dirs(count, seed):
    v = np.random.default_rng(seed).normal(size=(count, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def valid(d):
    return (d.astype(np.float64) ** 2).sum(1) > 0.9999  # RUtil::isValidDirection (RayUtility.cpp:38-40)


# ---- BiMetalSurface ----------------------------------------------------------
def test_bimetal_is_metal_about_the_facing_normal():
    rng = np.random.default_rng(1)
    for _ in range(300):
        n = norm(rng.normal(size=3))
        i = norm(rng.normal(size=3))
        if abs(i @ n) < 1e-3:
            continue
        front = n if i @ n < 0 else -n
        d_b, p_b, k_b = sample(S.MAT_BIMETAL, po.fparam(i, i, n), [1])
        d_m, p_m, k_m = sample(S.MAT_METAL, po.fparam(i, i, front), [1])
        np.testing.assert_array_equal(d_b, d_m)
        assert p_b[0] == p_m[0] == np.float32(1.0) and k_b[0] == k_m[0] == 0   # DeltaPDF = 1, no draws
        # PDF about the facing normal whichever way the normal points
        for o in (d_m[0], norm(rng.normal(size=3))):
            assert pdf(S.MAT_BIMETAL, po.fparam(i, o, n)) == pdf(S.MAT_METAL, po.fparam(i, o, front))
        # BRDF: the opaque test uses the given normal, so only a front-facing one reflects
        b = brdf(S.MAT_BIMETAL, po.fparam(i, d_m[0], n))
        if i @ n < 0:
            assert b == brdf(S.MAT_METAL, po.fparam(i, d_m[0], n))
            if -(i @ n) > 1e-3:  # DeltaPDF / (o.n + 1e-5)
                assert abs(b - 1.0 / (np.float64(np.float32(d_m[0] @ n)) + 1e-5)) <= 1e-5 * b
        else:
            assert b == 0.0


# ---- DiffuseTransmitSurface ----------------------------------------------------
@pytest.mark.parametrize("R", [0.7, 0.88])
def test_diffuse_transmit_closed_form_and_normalisation(R):
    n = norm((0.2, 1.0, -0.1))
    i = norm((0.5, -1.0, 0.3))  # arriving from the +n side
    R32 = np.float32(R)
    for o in sphere_dirs(400, 2):
        p = po.fparam(i, o, n, refl=R)
        reflect = np.float32(n @ i) * np.float32(n @ o) < 0
        assert brdf(S.MAT_DIFFUSE_TRANSMIT, p) == (R32 / PI if reflect else (np.float32(1) - R32) / PI)
        c = np.float32(abs(np.float64(np.float32(o @ n)) + 1e-5))
        want = (R32 * c / PI) if reflect else ((np.float32(1) - R32) * c / PI)
        assert abs(pdf(S.MAT_DIFFUSE_TRANSMIT, p) - want) <= 2e-7
    # the density over the whole sphere integrates to R + (1 - R) = 1
    tot = sum(pdf(S.MAT_DIFFUSE_TRANSMIT, po.fparam(i, o, n, refl=R)) for o in sphere_dirs(100000, 3))
    assert abs(tot * 4 * math.pi / 100000 - 1.0) < 0.01


def test_diffuse_transmit_sampling():
    n = norm((0.0, 1.0, 0.0))
    i = norm((0.3, -1.0, 0.1))
    R = 0.7
    d, pd, draws = sample(S.MAT_DIFFUSE_TRANSMIT, po.fparam(i, i, n, refl=R), range(1, 30001))
    assert (draws == 3).all()  # the choice, then Cos_Sampling's two draws
    c = d.astype(np.float64) @ n
    up = c > 0  # back towards the arriving side
    assert abs(up.mean() - R) < 0.012
    assert abs(np.abs(c[up]).mean() - 2 / 3) < 0.01 and abs(np.abs(c[~up]).mean() - 2 / 3) < 0.01
    share = np.where(up, R, 1 - R)
    np.testing.assert_allclose(pd, np.maximum(np.abs(c) / 3.14159, 1e-16) * share, rtol=1e-4, atol=1e-6)
    for k in range(0, len(d), 1499):  # the returned pdf is PDF() at the sample
        assert abs(pdf(S.MAT_DIFFUSE_TRANSMIT, po.fparam(i, d[k], n, refl=R)) - pd[k]) <= 1e-4 * max(pd[k], 1e-2)


# ---- GlossTransmitSurface --------------------------------------------------------
def mirror(o, n):
    return o - 2.0 * (o @ n) * n


@pytest.mark.parametrize("alpha", [0.1, 0.25])
def test_gloss_transmit_halves_are_mirror_images(alpha):
    n = np.array([0.0, 1.0, 0.0])
    i = norm((0.4, -1.0, 0.2))
    R = 0.1
    for o in sphere_dirs(300, 5):
        if o @ n < 0.05:
            continue
        refl = brdf(S.MAT_GLOSS_TRANSMIT, po.fparam(i, o, n, rough=alpha, refl=R))
        tran = brdf(S.MAT_GLOSS_TRANSMIT, po.fparam(i, mirror(o, n), n, rough=alpha, refl=R))
        assert abs(tran * R - refl * (1 - R)) <= 1e-4 * max(refl * (1 - R), 1e-6)
        pr = pdf(S.MAT_GLOSS_TRANSMIT, po.fparam(i, o, n, rough=alpha, refl=R))
        pt = pdf(S.MAT_GLOSS_TRANSMIT, po.fparam(i, mirror(o, n), n, rough=alpha, refl=R))
        assert abs(pt * R - pr * (1 - R)) <= 1e-4 * max(pr * (1 - R), 1e-6)


def test_gloss_transmit_shares_the_conductor_microfacet_terms():
    # Roughconductor = F(n.-i) G D / fx and GlossTransmit (reflection side) =
    # R G D / fx with the same D, G and fx: their ratio is constant in o.
    n = np.array([0.0, 1.0, 0.0])
    i = norm((0.6, -1.0, -0.2))
    ratios = []
    for o in sphere_dirs(400, 6):
        if o @ n < 0.1:
            continue
        g = brdf(S.MAT_GLOSS_TRANSMIT, po.fparam(i, o, n, rough=0.3, refl=1.0))
        c = brdf(S.MAT_ROUGH_CONDUCTOR, po.fparam(i, o, n, fi=1.0, ti=1.1978, rough=0.3, absorb=7.0488))
        if g > 1e-3:
            ratios.append(c / g)
    ratios = np.array(ratios)
    assert len(ratios) > 100
    assert np.ptp(ratios) <= 2e-5 * ratios.mean(), (ratios.min(), ratios.max())
    assert 0.5 < ratios.mean() < 1.0  # a metal's Fresnel reflectance


@pytest.mark.parametrize("alpha", [0.2, 0.5])
def test_gloss_transmit_pdf_normalisation(alpha):
    # at normal incidence each half integrates to P(tan^2 <= 1) = 1 / (1 + alpha^2)
    n = np.array([0.0, 1.0, 0.0])
    i = np.array([0.0, -1.0, 0.0])
    dirs = sphere_dirs(200000, 7)
    tot = sum(pdf(S.MAT_GLOSS_TRANSMIT, po.fparam(i, o, n, rough=alpha, refl=0.3)) for o in dirs)
    assert abs(tot * 4 * math.pi / len(dirs) - 1.0 / (1.0 + alpha * alpha)) < 0.015


def test_gloss_transmit_sampling():
    n = np.array([0.0, 1.0, 0.0])
    i = norm((0.3, -1.0, 0.0))
    R, alpha = 0.25, 0.25
    d, pd, draws = sample(S.MAT_GLOSS_TRANSMIT, po.fparam(i, i, n, rough=alpha, refl=R), range(1, 30001))
    assert (draws == 3).all()  # the choice, then GGX_Sampling's phi and e
    ok = valid(d)
    d, pd = d[ok].astype(np.float64), pd[ok]
    up = d @ n > 0
    assert abs(up.mean() - R) < 0.015
    # transmitted samples are reflected ones mirrored through the surface plane
    m = d[~up] - 2.0 * (d[~up] @ n)[:, None] * n
    assert np.abs(m.mean(0) - d[up].mean(0)).max() < 0.02
    for k in range(0, len(d), 1499):
        v = pdf(S.MAT_GLOSS_TRANSMIT, po.fparam(i, d[k], n, rough=alpha, refl=R))
        assert abs(v - pd[k]) <= 2e-4 * max(1.0, abs(v))
