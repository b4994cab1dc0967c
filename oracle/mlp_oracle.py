"""CPU oracle of the learned emission proposal (test infrastructure: only
tests/, __graft_entry__.smoke() and bench.py's CPU leg use it).

The proposal has no counterpart in the reference.  It hooks the emission
step generateLightRay (SPPM/SPPM/SPPM.cpp:453-471) and the stored PDF
(SPPM.cpp:527); this file restates the network, its loss and its gradients
as the GPU kernels compute them.

Cells: 32 x 32 light-surface (u, v) positions x 16 x 16 hemi-octahedral
directions (s, t), 262,144 cells, index ((iu * 32 + iv) * 16 + is) * 16 + it.
Input of a cell: its centre (2u-1, 2v-1, s, t), each coordinate encoded as
sin / cos at 8 frequencies pi * 2^(k/2), k = 0..7 (64 features).
Network: 64 features -> `layers` hidden ReLU layers of `hidden` units -> 1
logit; parameters in the order W1 [in][out], b1, ..., W_L, b_L, w_out, b_out
(the default, hidden 64 x 4 layers, has 16,705).  Networks are `Net`
tuples; the module-level SHAPES / NPARAMS / HID / LAYERS describe the
default.
Proposal P(cell) = softmax over all cells of the logits.
Loss over a deposit histogram W (luminance deposited by the photons emitted
in each cell): L = -sum_c (W_c / sum W) log P_c, so dL/dlogit_c =
P_c - W_c / sum W.

Two evaluations: `exact` (float64, no rounding; pinned by finite
differences in the tests) and low precision, `bf16` or `fp16` (what the MFMA
kernels do: layer inputs and weights rounded, products accumulated in
float32, the output layer in float32 on the last hidden layer before
rounding; the backward carries dL/dlogit scaled by GRAD_SCALE = 1024 through
the rounded layers and divides the sums by it).
"""
from collections import namedtuple

import numpy as np

GU, GV, GS, GT = 32, 32, 16, 16
NCELLS = GU * GV * GS * GT
NFREQ = 8
FEAT = 4 * 2 * NFREQ
GRAD_SCALE = 1024.0
Net = namedtuple("Net", "hidden layers dtype")
DEFAULT = Net(64, 4, "bf16")
WIDE = Net(128, 6, "fp16")


def shapes(net=DEFAULT):
    """Parameter names and shapes; the output layer is named w{L+1}, b{L+1}."""
    out = []
    for l in range(1, net.layers + 1):
        out += [(f"W{l}", (FEAT if l == 1 else net.hidden, net.hidden)), (f"b{l}", (net.hidden,))]
    return out + [(f"w{net.layers + 1}", (net.hidden,)), (f"b{net.layers + 1}", (1,))]


def nparams(net=DEFAULT):
    return sum(int(np.prod(s)) for _, s in shapes(net))


HID = DEFAULT.hidden
LAYERS = DEFAULT.layers
SHAPES = shapes(DEFAULT)
NPARAMS = nparams(DEFAULT)


def unpack(theta, net=DEFAULT):
    out, o = {}, 0
    for name, shp in shapes(net):
        n = int(np.prod(shp))
        out[name] = theta[o:o + n].reshape(shp)
        o += n
    return out


def pack(d, net=DEFAULT):
    return np.concatenate([np.asarray(d[name], np.float64).reshape(-1) for name, _ in shapes(net)])


def freqs():
    return (np.pi * 2.0 ** (np.arange(NFREQ) / 2.0)).astype(np.float32)


def cell_coords(cells=None):
    c = np.arange(NCELLS) if cells is None else np.asarray(cells)
    it = c % GT
    is_ = (c // GT) % GS
    iv = (c // (GS * GT)) % GV
    iu = c // (GV * GS * GT)
    f = np.float32
    a = f(2.0) * ((iu.astype(f) + f(0.5)) / f(GU)) - f(1.0)
    b = f(2.0) * ((iv.astype(f) + f(0.5)) / f(GV)) - f(1.0)
    s = f(-1.0) + (is_.astype(f) + f(0.5)) * (f(2.0) / f(GS))
    t = f(-1.0) + (it.astype(f) + f(0.5)) * (f(2.0) / f(GT))
    return np.stack([a, b, s, t], 1)


def features(coords):
    """[n, 4] -> [n, 64]: feature 16 * j + 2 * k (+1) = sin (cos) of freq_k * coord_j."""
    fr = freqs()
    n = coords.shape[0]
    x = coords[:, :, None].astype(np.float64) * fr[None, None, :].astype(np.float64)
    out = np.empty((n, 4, NFREQ, 2), np.float64)
    out[..., 0] = np.sin(x)
    out[..., 1] = np.cos(x)
    return out.reshape(n, FEAT)


def to_bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(np.asarray(x, np.float32))
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def to_fp16(x):
    """Round float32 values to IEEE half (nearest even), returned as float32."""
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float32)


def rounder(mode, net):
    if mode == "bf16" or (mode == "lowp" and net.dtype == "bf16"):
        return to_bf16
    if mode == "fp16" or (mode == "lowp" and net.dtype == "fp16"):
        return to_fp16
    raise ValueError(mode)


def init_params(seed, net=DEFAULT):
    """He-uniform hidden layers, zero output layer (uniform proposal at start)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in shapes(net):
        if name.startswith("W"):
            lim = np.sqrt(6.0 / shp[0])
            d[name] = rng.uniform(-lim, lim, shp)
        else:
            d[name] = np.zeros(shp)
    return pack(d, net).astype(np.float32)


def forward(theta, X, mode="exact", net=DEFAULT):
    """Returns logits [n] and the list of post-ReLU hidden activations."""
    p = unpack(np.asarray(theta, np.float64), net)
    L = net.layers
    if mode == "exact":
        h = np.asarray(X, np.float64)
        hs = [h]
        for l in range(1, L + 1):
            h = np.maximum(h @ p[f"W{l}"] + p[f"b{l}"], 0.0)
            hs.append(h)
        return h @ p[f"w{L + 1}"] + p[f"b{L + 1}"][0], hs
    rnd = rounder(mode, net)
    h = rnd(X)
    hs = [h]
    for l in range(1, L + 1):
        z = (h.astype(np.float32) @ rnd(p[f"W{l}"].astype(np.float32))).astype(np.float32) + \
            p[f"b{l}"].astype(np.float32)
        hf = np.maximum(z, np.float32(0.0))
        hs.append(hf)  # float32 (the output layer uses it unrounded)
        h = rnd(hf)
    logit = (hs[-1] @ p[f"w{L + 1}"].astype(np.float32)).astype(np.float32) + np.float32(p[f"b{L + 1}"][0])
    return logit, hs


def softmax(logits):
    m = np.max(logits)
    e = np.exp(np.asarray(logits, np.float64) - m)
    return e / e.sum(), m + np.log(e.sum())


def loss_and_grad(theta, X, W, mode="exact", net=DEFAULT):
    """Loss and gradient (flat, like theta) for deposit histogram W."""
    W = np.asarray(W, np.float64)
    tot = W.sum()
    L = net.layers
    logit, hs = forward(theta, X, mode, net)
    P, lse = softmax(logit)
    loss = -np.sum(W * (np.asarray(logit, np.float64) - lse)) / tot
    g = P - W / tot  # dL/dlogit
    p = unpack(np.asarray(theta, np.float64), net)
    grads = {}
    wo, bo = f"w{L + 1}", f"b{L + 1}"
    if mode == "exact":
        grads[wo] = hs[-1].T @ g
        grads[bo] = np.array([g.sum()])
        dh = np.outer(g, p[wo]) * (hs[-1] > 0)
        for l in range(L, 0, -1):
            grads[f"W{l}"] = hs[l - 1].T @ dh
            grads[f"b{l}"] = dh.sum(0)
            if l > 1:
                dh = (dh @ p[f"W{l}"].T) * (hs[l - 1] > 0)
        return loss, pack(grads, net)
    rnd = rounder(mode, net)
    S = GRAD_SCALE
    g32 = (g * S).astype(np.float32)  # scaled dL/dlogit
    hL = rnd(hs[-1])
    grads[wo] = hL.astype(np.float64).T @ g32.astype(np.float64) / S
    grads[bo] = np.array([g32.astype(np.float64).sum() / S])
    dh = np.outer(g32, p[wo].astype(np.float32)).astype(np.float32) * (hL > 0)
    for l in range(L, 0, -1):
        hin = hs[l - 1] if l == 1 else rnd(hs[l - 1])  # hs[0] is already rounded
        grads[f"W{l}"] = rnd(hin).astype(np.float64).T @ rnd(dh).astype(np.float64) / S
        grads[f"b{l}"] = dh.astype(np.float64).sum(0) / S
        if l > 1:
            dh = (rnd(dh) @ rnd(p[f"W{l}"].astype(np.float32)).T).astype(np.float32) * (rnd(hs[l - 1]) > 0)
    return loss, pack(grads, net)


def adam(theta, m, v, grad, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (step counts from 1); returns new theta, m, v (float32)."""
    grad = np.asarray(grad, np.float32)
    m = np.float32(b1) * m + np.float32(1 - b1) * grad
    v = np.float32(b2) * v + np.float32(1 - b2) * grad * grad
    mh = m / np.float32(1 - b1 ** step)
    vh = v / np.float32(1 - b2 ** step)
    theta = theta - np.float32(lr) * mh / (np.sqrt(vh) + np.float32(eps))
    return theta.astype(np.float32), m.astype(np.float32), v.astype(np.float32)


# ---- hemi-octahedral direction map (local frame, z = light normal) --------
def dir_to_st(d):
    """Unit d with z >= 0 -> (s, t) in [-1, 1]^2."""
    d = np.asarray(d, np.float64)
    n = np.abs(d[..., 0]) + np.abs(d[..., 1]) + d[..., 2]
    px, py = d[..., 0] / n, d[..., 1] / n
    return np.stack([px + py, px - py], -1)


def st_to_dir(st):
    st = np.asarray(st, np.float64)
    px, py = 0.5 * (st[..., 0] + st[..., 1]), 0.5 * (st[..., 0] - st[..., 1])
    pz = 1.0 - np.abs(px) - np.abs(py)
    P = np.stack([px, py, pz], -1)
    return P / np.linalg.norm(P, axis=-1, keepdims=True)


def st_jacobian(st):
    """d(omega) / d(s, t) = 1 / (2 |P|^3), P the unnormalised octahedron point."""
    st = np.asarray(st, np.float64)
    px, py = 0.5 * (st[..., 0] + st[..., 1]), 0.5 * (st[..., 0] - st[..., 1])
    pz = 1.0 - np.abs(px) - np.abs(py)
    return 1.0 / (2.0 * np.sqrt(px * px + py * py + pz * pz) ** 3)
