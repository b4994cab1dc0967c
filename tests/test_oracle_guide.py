"""Pins the CPU oracle of the learned emission proposal: the analytic
gradient against central finite differences (float64), the bf16 evaluation
against the exact one (and the same for the wide fp16 network, 128 x 6),
Adam's first step in closed form, and the
hemi-octahedral direction map (round trip, Jacobian integrates to the
hemisphere's 2*pi, Jacobian against a finite-difference area ratio)."""
import numpy as np
import pytest

import mlp_oracle as mo


@pytest.fixture(scope="module")
def small_problem():
    rng = np.random.default_rng(5)
    cells = rng.choice(mo.NCELLS, 300, replace=False)
    X = mo.features(mo.cell_coords(cells))
    theta = mo.init_params(3).astype(np.float64)
    # non-zero output layer so every layer gets gradient
    p = mo.unpack(theta)
    p["w5"][:] = rng.normal(0, 0.3, mo.HID)
    p["b5"][:] = 0.1
    theta = mo.pack(p)
    W = rng.gamma(0.5, 1.0, len(cells)) * (rng.random(len(cells)) < 0.4)
    return X, theta, W


def test_gradient_matches_finite_differences(small_problem):
    X, theta, W = small_problem
    loss, g = mo.loss_and_grad(theta, X, W, "exact")
    rng = np.random.default_rng(9)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in mo.SHAPES])
    idx = np.concatenate([rng.choice(np.arange(offs[k], offs[k + 1]), min(12, offs[k + 1] - offs[k]), replace=False)
                          for k in range(len(mo.SHAPES))])
    eps = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += eps
        tm[i] -= eps
        num = (mo.loss_and_grad(tp, X, W, "exact")[0] - mo.loss_and_grad(tm, X, W, "exact")[0]) / (2 * eps)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), (i, num, g[i])


def test_loss_gradient_wrt_logits_is_softmax_minus_target():
    # one-layer sanity: sum of the logit gradient is zero (softmax shift invariance) -> db5 == 0
    rng = np.random.default_rng(2)
    X = mo.features(mo.cell_coords(rng.choice(mo.NCELLS, 200, replace=False)))
    theta = mo.init_params(1).astype(np.float64)
    W = rng.random(200)
    _, g = mo.loss_and_grad(theta, X, W, "exact")
    assert abs(g[-1]) < 1e-12
    # zero output layer: uniform proposal; loss = log(n) exactly
    loss, _ = mo.loss_and_grad(theta, X, W, "exact")
    assert abs(loss - np.log(200)) < 1e-12


def test_bf16_evaluation_tracks_exact(small_problem):
    X, theta, W = small_problem
    l0, g0 = mo.loss_and_grad(theta, X, W, "exact")
    l1, g1 = mo.loss_and_grad(theta.astype(np.float32), X, W, "bf16")
    assert abs(l0 - l1) < 2e-2 * abs(l0)
    cos = g0 @ g1 / np.linalg.norm(g0) / np.linalg.norm(g1)
    assert cos > 0.98  # bf16 rounding of 4 layers; a wrong term drops this far lower


@pytest.fixture(scope="module")
def wide_problem():
    net = mo.WIDE
    rng = np.random.default_rng(6)
    cells = rng.choice(mo.NCELLS, 200, replace=False)
    X = mo.features(mo.cell_coords(cells))
    theta = mo.init_params(4, net).astype(np.float64)
    p = mo.unpack(theta, net)
    p["w7"][:] = rng.normal(0, 0.3, net.hidden)
    p["b7"][:] = 0.1
    theta = mo.pack(p, net)
    W = rng.gamma(0.5, 1.0, len(cells)) * (rng.random(len(cells)) < 0.4)
    return X, theta, W


def test_wide_network_shapes():
    net = mo.WIDE
    assert [n for n, _ in mo.shapes(net)][-2:] == ["w7", "b7"]
    assert mo.nparams(net) == 64 * 128 + 128 + 5 * (128 * 128 + 128) + 128 + 1
    assert mo.nparams(mo.DEFAULT) == 16705 == mo.NPARAMS


def test_wide_gradient_matches_finite_differences(wide_problem):
    net = mo.WIDE
    X, theta, W = wide_problem
    loss, g = mo.loss_and_grad(theta, X, W, "exact", net)
    rng = np.random.default_rng(10)
    sh = mo.shapes(net)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in sh])
    idx = np.concatenate([rng.choice(np.arange(offs[k], offs[k + 1]), min(6, offs[k + 1] - offs[k]), replace=False)
                          for k in range(len(sh))])
    eps = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += eps
        tm[i] -= eps
        num = (mo.loss_and_grad(tp, X, W, "exact", net)[0] - mo.loss_and_grad(tm, X, W, "exact", net)[0]) / (2 * eps)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), (i, num, g[i])


def test_fp16_evaluation_tracks_exact(wide_problem):
    net = mo.WIDE
    X, theta, W = wide_problem
    l0, g0 = mo.loss_and_grad(theta, X, W, "exact", net)
    l1, g1 = mo.loss_and_grad(theta.astype(np.float32), X, W, "fp16", net)
    assert abs(l0 - l1) < 1e-2 * abs(l0)
    cos = g0 @ g1 / np.linalg.norm(g0) / np.linalg.norm(g1)
    assert cos > 0.99  # fp16 keeps 11 bits; the scaled gradients stay out of its subnormal range


def test_fp16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, 65504.0, 1e-8], np.float32)
    r = mo.to_fp16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0 + 2.0 ** -9)
    assert r[3] == 65504.0 and r[4] == 0.0


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.01171875, -2.5, 3.0e38, 1e-40], np.float32)
    r = mo.to_bf16(x)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1.0
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.015625)
    assert r[3] == -2.5
    assert (r.view(np.uint32) & 0xFFFF == 0).all()


def test_adam_first_step_closed_form():
    g = np.array([0.5, -2.0, 0.0, 1e-3], np.float32)
    th = np.ones(4, np.float32)
    t1, m, v = mo.adam(th, np.zeros(4, np.float32), np.zeros(4, np.float32), g, 1, lr=1e-3)
    # bias-corrected first step: theta - lr * g / (|g| + eps)
    want = th - 1e-3 * g / (np.abs(g) + 1e-8)
    np.testing.assert_allclose(t1, want, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(m, 0.1 * g, rtol=1e-6)
    np.testing.assert_allclose(v, 0.001 * g * g, rtol=1e-5)


def test_hemi_octahedral_map():
    rng = np.random.default_rng(4)
    d = rng.normal(size=(5000, 3))
    d[:, 2] = np.abs(d[:, 2])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    st = mo.dir_to_st(d)
    assert (np.abs(st) <= 1 + 1e-12).all()
    np.testing.assert_allclose(mo.st_to_dir(st), d, atol=1e-12)
    # the Jacobian integrates to the hemisphere's solid angle
    n = 800
    g = (np.arange(n) + 0.5) / n * 2 - 1
    S, T = np.meshgrid(g, g, indexing="ij")
    total = mo.st_jacobian(np.stack([S, T], -1)).sum() * (2.0 / n) ** 2
    assert abs(total - 2 * np.pi) < 1e-3
    # local area ratio: a small (s, t) square maps to solid angle ~ J * ds * dt
    for s0, t0 in ((0.1, 0.2), (-0.5, 0.3), (0.7, -0.1)):
        h = 1e-4
        c = mo.st_to_dir(np.array([[s0, t0], [s0 + h, t0], [s0, t0 + h]]))
        area = np.linalg.norm(np.cross(c[1] - c[0], c[2] - c[0]))
        assert abs(area / (h * h) - mo.st_jacobian(np.array([s0, t0]))) < 1e-3 * area / (h * h)


def test_cell_layout():
    co = mo.cell_coords(np.array([0, mo.NCELLS - 1, 1, mo.GT]))
    np.testing.assert_allclose(co[0], [-1 + 1 / 32, -1 + 1 / 32, -1 + 1 / 16, -1 + 1 / 16], atol=1e-7)
    np.testing.assert_allclose(co[1], [1 - 1 / 32, 1 - 1 / 32, 1 - 1 / 16, 1 - 1 / 16], atol=1e-7)
    assert co[2][3] > co[0][3] and co[3][2] > co[0][2]
