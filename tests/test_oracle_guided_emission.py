"""Pins the oracle's guided emission (the learned proposal's mixture draw and
density; no counterpart in the reference, it replaces generateLightRay's
sampling, SPPM.cpp:453-471, and the photon PDF, SPPM.cpp:527) from the
mathematics, with proposal tables of random logits:

* the light frame: P0, P0 + e1, P0 + e2, P0 + e1 + e2 are the rectangle's
  corners, e1 is perpendicular to e2, n = e1 x e2 / |.| faces like the
  facets' normal;
* the mixture density integrates to one over light area x hemisphere
  (untextured light), and over the 1000-point tables x hemisphere with the
  reference's table measure (textured light, LightSampler.cpp:68-76);
* the draws follow that density: the sample mean of f / q equals the exact
  measure of f's support (a rectangle of the light times a cone of
  directions), so the weight flux / PDF that SPPM accumulates is unbiased
  whichever component drew the photon.

A missing hemi-octahedral Jacobian (2|P|^3), a wrong cell area, a dropped
mixture weight or a density of the wrong cell fails one of these.
"""
import math

import numpy as np
import pytest

import pyoracle as po
import scenes


def random_tables(seed, sigma=1.0, mask=None):
    rng = np.random.default_rng(seed)
    prob = np.exp(sigma * rng.normal(size=po.GUIDE_CELLS)).astype(np.float32)
    if mask is not None:
        prob *= mask
    cdf = np.cumsum(prob, dtype=np.float64).astype(np.float32)
    return prob, cdf


def hemisphere(n, normal, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    v[v @ normal < 0] *= -1
    return v


@pytest.fixture(scope="module")
def box():
    sc = scenes.box_scene(with_spheres=False)
    return sc, po.OracleScene(sc)


@pytest.fixture(scope="module")
def textured():
    sc = scenes.full_scene(freq=8, tex_size=32)
    return sc, po.OracleScene(sc)


def test_light_frame(box):
    sc, os_ = box
    g = po.OracleGuide(os_, 0.5)
    f = {k: v.astype(np.float64) for k, v in g.frame().items()}
    L = os_.lights()
    verts = np.concatenate([sc.pos[k].reshape(3, 3) for k in L["facet"]]).astype(np.float64)
    for c in (f["P0"], f["P0"] + f["e1"], f["P0"] + f["e2"], f["P0"] + f["e1"] + f["e2"]):
        assert np.min(np.linalg.norm(verts - c, axis=1)) < 1e-6
    assert abs(f["e1"] @ f["e2"]) < 1e-7
    n = np.cross(f["e1"], f["e2"])
    np.testing.assert_allclose(f["n"], n / np.linalg.norm(n), atol=1e-7)
    assert f["n"] @ np.array([0.0, -1.0, 0.0]) > 0.999
    np.testing.assert_allclose(f["t1"], f["e1"] / np.linalg.norm(f["e1"]), atol=1e-7)
    np.testing.assert_allclose(f["t2"], np.cross(f["n"], f["t1"]), atol=1e-7)


@pytest.mark.parametrize("w", [0.25, 0.5, 1.0])
def test_density_integrates_to_one_untextured(box, w):
    sc, os_ = box
    g = po.OracleGuide(os_, w)
    g.set_tables(*random_tables(1))
    f = {k: v.astype(np.float64) for k, v in g.frame().items()}
    area = np.linalg.norm(np.cross(f["e1"], f["e2"]))
    rng = np.random.default_rng(2)
    M = 1_000_000
    uv = rng.random((M, 2))
    pos = f["P0"] + uv[:, :1] * f["e1"] + uv[:, 1:] * f["e2"]
    d = hemisphere(M, f["n"], 3)
    q = g.density_batch(pos, d).astype(np.float64)
    est = area * 2 * math.pi * q.mean()
    err = area * 2 * math.pi * q.std() / math.sqrt(M)
    assert abs(est - 1.0) < max(5 * err, 2e-3), (est, err)


def test_density_integrates_to_one_textured(textured):
    sc, os_ = textured
    L = os_.lights()
    assert L["textured"].all() and len(L["facet"]) == 2
    g = po.OracleGuide(os_, 0.5)
    mass, start, pt_cell = g.point_cells()
    # proposal cells whose position cell holds no table point are masked out
    pmask = np.repeat(mass > 0, 256).astype(np.float32)
    g.set_tables(*random_tables(4, mask=pmask))
    f = g.frame()
    area = float(0.5 * np.linalg.norm(np.cross(*(sc.pos[L["facet"][0]].reshape(3, 3)[1:] -
                                                 sc.pos[L["facet"][0]].reshape(3, 3)[0]))))
    # the mixture's probability of (table point k, direction) is pi_k q / D_k
    # with D_k = pi_k 1000 / area: sum_k (area / 1000) int q_k = 1
    K = 2000
    per = 400
    light = np.repeat(np.arange(2), 1000 * per).astype(np.int32)
    point = np.tile(np.repeat(np.arange(1000), per), 2).astype(np.int32)
    d = hemisphere(K * per, f["n"].astype(np.float64), 5)
    q = g.density_batch(np.zeros((K * per, 3), np.float32), d, light, point).astype(np.float64)
    est = (area / 1000.0) * K * 2 * math.pi * q.mean()
    err = (area / 1000.0) * K * 2 * math.pi * q.std() / math.sqrt(K * per)
    assert abs(est - 1.0) < max(5 * err, 3e-3), (est, err)


def cone_measure(half_deg):
    return 2 * math.pi * (1 - math.cos(math.radians(half_deg)))


@pytest.mark.parametrize("w", [0.25, 0.5])
def test_samples_follow_density_untextured(box, w):
    sc, os_ = box
    g = po.OracleGuide(os_, w)
    g.set_tables(*random_tables(6, sigma=1.5))
    r = po.OracleRender(os_, scenes.pinhole_camera(8, 8), 0.9, 0.015, 1 << 20, nthreads=1)
    r.set_guide(g)
    f = {k: v.astype(np.float64) for k, v in g.frame().items()}
    area = np.linalg.norm(np.cross(f["e1"], f["e2"]))
    M = 400_000
    pos, d, q, _ = r.guide_emit_batch(3, 0, M)
    lpos = pos.astype(np.float64) - 1e-4 * f["n"]
    rel = lpos - f["P0"]
    u = rel @ f["e1"] / (f["e1"] @ f["e1"])
    v = rel @ f["e2"] / (f["e2"] @ f["e2"])
    assert (u > -1e-5).all() and (u < 1 + 1e-5).all() and (v > -1e-5).all() and (v < 1 + 1e-5).all()
    axis = math.cos(math.radians(30)) * f["n"] + math.sin(math.radians(30)) * f["t1"]
    inside = (u > 0.2) & (u < 0.6) & (v > 0.1) & (v < 0.7) & (d.astype(np.float64) @ axis > math.cos(math.radians(40)))
    vals = inside / q.astype(np.float64)
    est, err = vals.mean(), vals.std() / math.sqrt(M)
    want = 0.4 * 0.6 * area * cone_measure(40)
    assert abs(est - want) < max(5 * err, 0.01 * want), (est, want, err)
    # the whole domain
    vals = 1.0 / q.astype(np.float64)
    assert abs(vals.mean() - area * 2 * math.pi) < 5 * vals.std() / math.sqrt(M)


def test_samples_follow_density_textured(textured):
    sc, os_ = textured
    g = po.OracleGuide(os_, 0.5)
    mass, start, _ = g.point_cells()
    pmask = np.repeat(mass > 0, 256).astype(np.float32)
    g.set_tables(*random_tables(7, sigma=1.5, mask=pmask))
    L = os_.lights()
    area = float(0.5 * np.linalg.norm(np.cross(*(sc.pos[L["facet"][0]].reshape(3, 3)[1:] -
                                                 sc.pos[L["facet"][0]].reshape(3, 3)[0]))))
    r = po.OracleRender(os_, scenes.pinhole_camera(8, 8), 0.9, 0.015, 1 << 20, nthreads=1)
    r.set_guide(g)
    f = {k: v.astype(np.float64) for k, v in g.frame().items()}
    M = 400_000
    _, d, q, _ = r.guide_emit_batch(2, 0, M)
    axis = math.cos(math.radians(25)) * f["n"] + math.sin(math.radians(25)) * f["t2"]
    inside = d.astype(np.float64) @ axis > math.cos(math.radians(45))
    # E[g (1000 / area) / q] = sum over the points that can be drawn of int g
    vals = inside * (1000.0 / area) / q.astype(np.float64)
    est, err = vals.mean(), vals.std() / math.sqrt(M)
    want = int(start[-1]) * cone_measure(45)
    assert abs(est - want) < max(5 * err, 0.01 * want), (est, want, err)
