"""Pins the oracle's GGX dielectric (GGXGlass, material.h:280-394;
GGXGlass_Sampling / _PDF, Sampling.cpp:181-288; Fresnel_Reflectance,
Sampling.cpp:297-313) from the mathematics.

The microfacet normal h is drawn from D(h)(h.n); the ray goes to the
reflected direction with probability F and to the refracted one with 1 - F.
So the reflected part of PDF() integrates to E_h[F] over the h whose
reflection is valid (the Jacobian 1 / (4 o.h) is the textbook one).

The refracted part is where the reference departs from the textbook: its
Jacobian (Sampling.cpp:281-283, and the BTDF's at material.h:385) is
eta_o^2 |o.h| / (eta_i |i.h| + eta_o |o.h|)^2, where Walter et al. 2007
(eq. 17) have (eta_i (w_i.h) + eta_o (w_o.h))^2 with w_i.h and w_o.h of
opposite sign, i.e. a difference.  PDF() therefore integrates over the
refracted lobe to E_h[(1 - F) r(h)] with r = ((eta_i c_i - eta_o c_o) /
(eta_i c_i + eta_o c_o))^2, about 1/25 at normal incidence from air into
IOR 1.5, while the sampler's directions follow the true density.  The
restatement keeps the reference's form (parity), and these tests pin it:
the h-space side below is computed in float64 from the textbook GGX and
Fresnel formulas and the ratio r, so a wrong Jacobian (either one), a
swapped F or dropped D fails them.

* PDF() over the sphere: reflected mass E_h[F], refracted mass E_h[(1-F) r].
* Sampled directions: each sample's own pdf equals PDF() at it; the share
  of samples in a reflection region equals the region's PDF() integral, in
  a refraction region the integral of PDF() / r.
* BRDF closed forms at normal incidence, from the reference's expressions:
  reflection F D / 4; transmission eta_o^2 (1 - F) D / (eta_i + eta_o)^2
  (same sum in the denominator), with D(n) = 1 / (pi alpha^2) and
  F = ((eta_o - eta_i) / (eta_o + eta_i))^2; and the reflection lobe is
  reciprocal, f(i -> o) = f(-o -> -i).
"""
import math

import numpy as np
import pytest

import pyoracle as po
import sppm_amd as S

N_UP = np.array([0.0, 1.0, 0.0])


def sphere_dirs(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def fresnel(cos_i, n1, n2):
    """Unpolarised dielectric Fresnel reflectance (textbook)."""
    cos_i = np.clip(cos_i, 0.0, 1.0)
    sin_t = n1 / n2 * np.sqrt(1 - cos_i ** 2)
    tir = sin_t >= 1.0
    cos_t = np.sqrt(np.maximum(0.0, 1 - sin_t ** 2))
    with np.errstate(invalid="ignore", divide="ignore"):
        rs = (n1 * cos_i - n2 * cos_t) / (n1 * cos_i + n2 * cos_t)
        rp = (n2 * cos_i - n1 * cos_t) / (n2 * cos_i + n1 * cos_t)
    return np.where(tir, 1.0, 0.5 * (rs ** 2 + rp ** 2))


def ggx_h(alpha, n, seed):
    """Microfacet normals about +y drawn from D(h)(h.n) (exact inverse CDF)."""
    rng = np.random.default_rng(seed)
    u1, u2 = rng.random(n), rng.random(n)
    t2 = alpha * alpha * u1 / (1 - u1)
    c = 1 / np.sqrt(1 + t2)
    s = np.sqrt(1 - c * c)
    ph = 2 * math.pi * u2
    return np.stack([s * np.cos(ph), c, s * np.sin(ph)], 1)


def hspace_masses(i, alpha, n1, n2, n=400000, seed=3):
    """E_h[F valid_reflection], E_h[(1 - F) r valid_refraction] for incoming
    direction i (travelling toward the surface, i.n < 0), r the ratio of the
    reference's refraction Jacobian to the textbook one."""
    h = ggx_h(alpha, n, seed)
    ih = h @ i
    front = ih < 0  # microfacets facing the ray
    cos_i = -ih
    F = fresnel(cos_i, n1, n2)
    o_r = i - 2 * ih[:, None] * h
    ok_r = front & (o_r @ N_UP > 0)
    eta = n1 / n2
    k = 1 - eta * eta * (1 - cos_i ** 2)
    o_t = eta * i + (eta * cos_i - np.sqrt(np.maximum(k, 0)))[:, None] * h
    ok_t = front & (k > 0) & (o_t @ N_UP < 0)
    co = np.abs(np.einsum("ij,ij->i", o_t, h))
    r = ((n1 * cos_i - n2 * co) / (n1 * cos_i + n2 * co)) ** 2
    return float(np.mean(F * ok_r)), float(np.mean((1 - F) * r * ok_t))


def refraction_ratio(i, o, n1, n2):
    """r(h) at the refraction microfacet normal of (i -> o), per direction o."""
    h = -(n1 * -i[None, :] + n2 * o)
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    h *= np.sign(h @ N_UP)[:, None]
    ci, co = np.abs(h @ i), np.abs(np.einsum("ij,ij->i", o, h))
    return ((n1 * ci - n2 * co) / (n1 * ci + n2 * co)) ** 2


def pdf_masses(i, alpha, n1, n2, ndir=600000, seed=5):
    p = po.fparam(i, i, N_UP, fi=n1, ti=n2, rough=alpha)
    d = sphere_dirs(ndir, seed)
    pdf, _ = po.bsdf_eval(S.MAT_GGX_GLASS, p, d)
    pdf = pdf.astype(np.float64)
    up = d @ N_UP > 0
    return 4 * math.pi * np.mean(pdf * up), 4 * math.pi * np.mean(pdf * ~up)


@pytest.mark.parametrize("alpha", [0.3, 0.5])
def test_pdf_mass_normal_incidence(alpha):
    # at normal incidence the reflected lobe covers the h tilted below 45
    # degrees: its mass is int_{<45deg} F D cos dh (1-D quadrature)
    i = -N_UP
    th = np.linspace(0, math.pi / 2, 200001)
    D = alpha ** 2 / (math.pi * (1 - (1 - alpha ** 2) * np.cos(th) ** 2) ** 2)
    F = fresnel(np.cos(th), 1.0, 1.5)
    integrand = 2 * math.pi * D * np.cos(th) * np.sin(th)
    assert abs(np.trapezoid(integrand, th) - 1.0) < 1e-6
    refl = np.trapezoid((F * integrand)[th < math.pi / 4], th[th < math.pi / 4])
    r, t = pdf_masses(i, alpha, 1.0, 1.5)
    hr, ht = hspace_masses(i, alpha, 1.0, 1.5)
    assert abs(r - refl) < 0.004, (r, refl)
    assert abs(r - hr) < 0.004, (r, hr)
    assert abs(t - ht) < 0.004 + 0.03 * ht, (t, ht)
    assert t < 0.1  # the reference's refraction Jacobian (see the module docstring)


@pytest.mark.parametrize("deg,n1,n2", [(40, 1.0, 1.5), (25, 1.5, 1.0)])
def test_pdf_mass_oblique(deg, n1, n2):
    th = math.radians(deg)
    i = np.array([math.sin(th), -math.cos(th), 0.0])
    alpha = 0.3
    r, t = pdf_masses(i, alpha, n1, n2)
    hr, ht = hspace_masses(i, alpha, n1, n2)
    assert abs(r - hr) < 0.004 + 0.02 * hr, (r, hr)
    assert abs(t - ht) < 0.004 + 0.03 * ht, (t, ht)


def test_sampling_follows_pdf():
    th = math.radians(35)
    i = np.array([math.sin(th), -math.cos(th), 0.0])
    alpha = 0.3
    p = po.fparam(i, i, N_UP, fi=1.0, ti=1.5, rough=alpha)
    d, spdf = po.bsdf_sample(S.MAT_GGX_GLASS, p, 60000, seed0=11)
    d = d.astype(np.float64)
    valid = np.linalg.norm(d - 1e-3, axis=1) > 1e-3  # RUtil::INVALID_DIR
    # each sample's pdf is PDF() at the sample (the sampler's form carries
    # its own fudge constants: 1.0000001, 0.99999, 1.00000001)
    pdf, _ = po.bsdf_eval(S.MAT_GGX_GLASS, p, d[valid])
    np.testing.assert_allclose(spdf[valid], pdf, rtol=3e-3, atol=1e-6)
    # share of samples in two regions == the regions' PDF integral
    snell = np.array([math.sin(th) / 1.5, -math.sqrt(1 - (math.sin(th) / 1.5) ** 2), 0.0])
    refl = np.array([math.sin(th), math.cos(th), 0.0])
    dirs = sphere_dirs(600000, 12)
    qpdf, _ = po.bsdf_eval(S.MAT_GGX_GLASS, p, dirs)
    qpdf = qpdf.astype(np.float64)
    down = dirs @ N_UP < 0
    true = qpdf.copy()
    true[down] /= refraction_ratio(i, dirs[down], 1.0, 1.5)
    for axis, cosmax in ((snell, 0.95), (refl, 0.9)):
        frac = np.mean(valid & (d @ axis > cosmax))
        quad = 4 * math.pi * np.mean(true * (dirs @ axis > cosmax))
        assert abs(frac - quad) < 0.01 + 0.05 * quad, (frac, quad)


@pytest.mark.parametrize("alpha", [0.1, 0.3, 0.5])
def test_brdf_normal_incidence_closed_forms(alpha):
    i = -N_UP
    D = 1.0 / (math.pi * alpha * alpha)
    F = ((1.5 - 1.0) / (1.5 + 1.0)) ** 2
    p = po.fparam(i, N_UP, N_UP, fi=1.0, ti=1.5, rough=alpha)
    _, b = po.bsdf_eval(S.MAT_GGX_GLASS, p, N_UP[None, :])
    assert abs(b[0] - F * D / 4) <= 2e-4 * F * D / 4
    _, b = po.bsdf_eval(S.MAT_GGX_GLASS, p, -N_UP[None, :])
    want = 1.5 ** 2 * (1 - F) * D / (1.0 + 1.5) ** 2
    assert abs(b[0] - want) <= 2e-4 * want
    # from inside the glass the same directions: reflectance is the same at
    # normal incidence, the transmitted term has eta_o = 1
    p2 = po.fparam(i, N_UP, N_UP, fi=1.5, ti=1.0, rough=alpha)
    _, b = po.bsdf_eval(S.MAT_GGX_GLASS, p2, -N_UP[None, :])
    want = 1.0 ** 2 * (1 - F) * D / (1.5 + 1.0) ** 2
    assert abs(b[0] - want) <= 2e-4 * want


def test_reflection_lobe_is_reciprocal():
    rng = np.random.default_rng(9)
    for _ in range(300):
        a = sphere_dirs(1, int(rng.integers(1 << 30)))[0]
        b = sphere_dirs(1, int(rng.integers(1 << 30)))[0]
        i = -np.abs(a[1]) * N_UP + a * np.array([1, 0, 1])   # travelling down
        o = np.abs(b[1]) * N_UP + b * np.array([1, 0, 1])    # leaving up
        i /= np.linalg.norm(i)
        o /= np.linalg.norm(o)
        if -i[1] < 0.05 or o[1] < 0.05:
            continue
        p = po.fparam(i, o, N_UP, fi=1.0, ti=1.5, rough=0.4)
        q = po.fparam(-o, -i, N_UP, fi=1.0, ti=1.5, rough=0.4)
        _, f1 = po.bsdf_eval(S.MAT_GGX_GLASS, p, o[None, :])
        _, f2 = po.bsdf_eval(S.MAT_GGX_GLASS, q, -i[None, :])
        assert abs(f1[0] - f2[0]) <= 1e-4 * max(f1[0], 1e-6), (f1, f2)
