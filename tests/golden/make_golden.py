"""Writes tests/golden/ref_vectors.json from the reference-built checker.

Runs in the build container where /root/reference exists: the checker
oracle/_ref/libsppm_ref.so is compiled from the unmodified reference sources
(MT.h, CameraSampler.h, QRandSequence.h, TableSampler.h, TextureSampler.h,
TextureUtility.cpp, Matrix.h/.cpp) by oracle/Makefile.ref.  The JSON holds
inputs and the reference's outputs only (data, no reference code).

    make -f oracle/Makefile.ref && python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402


def f32list(a):
    return [float(x) for x in np.asarray(a, np.float32).ravel()]


def main():
    R = po.ref()
    out = {"source": "FriendsRender-R reference sources compiled by oracle/Makefile.ref (g++ 11, strict FP)"}
    # RandomMT(seed).genrand64_real1() sequences (MT.h)
    mt = {}
    for seed in (0, 1, 135, 5489, 2147483647, -5, 65536 * 3 + 135):
        buf = np.zeros(64, np.float64)
        R.ref_mt_real1_seq(seed, 64, buf.ctypes.data_as(po.D))
        mt[str(seed)] = [float.hex(float(x)) for x in buf]
    out["mt_real1"] = mt
    # UniformSample(ID).Sample() (CameraSampler.h) -- pins the argument order
    us = {}
    for sid in (0, 7):
        buf = np.zeros((300, 3), np.float32)
        R.ref_uniform_sampler_seq(sid, 300, po.fp(buf))
        us[str(sid)] = f32list(buf)
    out["uniform_sampler"] = us
    # Halton / Hammersley (QRandSequence.h)
    out["halton"] = {str(b): [float(R.ref_halton(b, n)) for n in range(0, 1000)] for b in (2, 3, 5)}
    ham = np.zeros((1000, 3), np.float32)
    for i in range(1000):
        R.ref_hammersley(1000, i, po.fp(ham[i]))
    out["hammersley_1000"] = f32list(ham)
    # alias tables (TableSampler.h)
    rng = np.random.default_rng(2026)
    cases = [np.array([5.0, 5.0], np.float32), np.array([1.0], np.float32),
             np.array([0.0, 3.0, 0.0, 1.0], np.float32),
             rng.random(17).astype(np.float32), (rng.random(1000) ** 4).astype(np.float32),
             np.where(rng.random(64) < 0.3, 0.0, rng.random(64)).astype(np.float32)]
    tabs = []
    for v in cases:
        n = len(v)
        prob, left, right, avg = np.zeros(n, np.float32), np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(1, np.float32)
        R.ref_table_build(po.fp(v), n, po.fp(prob), po.ip(left), po.ip(right), po.fp(avg))
        pdf = np.zeros(n, np.float32)
        R.ref_table_pdf(po.fp(v), n, po.fp(pdf))
        draws = np.zeros(200, np.int32)
        R.ref_table_sample_seq(po.fp(v), n, 31337, 200, po.ip(draws))
        tabs.append(dict(values=f32list(v), prob=f32list(prob), left=left.tolist(), right=right.tolist(),
                         avg=float(avg[0]), pdf=f32list(pdf), draws_seed=31337, draws=draws.tolist()))
    out["tables"] = tabs
    # bilinear fetch (TextureSampler.h / TextureUtility.cpp)
    tex = (rng.random((5, 7, 4)) * 255).astype(np.uint8).astype(np.float32) / np.float32(255.0)
    tex = np.ascontiguousarray(tex, np.float32)
    uv = np.concatenate([rng.random((200, 2)) * 3 - 1, np.array([[0, 0], [1, 1], [0.5, 0.5], [1 / 7, 2 / 5],
                                                                    [0.9999999, 0.9999999], [-0.25, 1.75]])]).astype(np.float32)
    res = np.zeros((len(uv), 4), np.float32)
    R.ref_bilinear_batch(7, 5, po.fp(tex), len(uv), po.fp(uv), po.fp(res))
    nm = np.zeros((len(uv), 3), np.float32)
    for i in range(len(uv)):
        R.ref_normal_from_map(7, 5, po.fp(tex), float(uv[i, 0]), float(uv[i, 1]), po.fp(nm[i]))
    out["bilinear"] = dict(width=7, height=5, rgba=f32list(tex), uv=f32list(uv), out=f32list(res), normal=f32list(nm))
    # Matrix3::inv (Matrix.h)
    mats = (rng.random((20, 9)) * 2 - 1).astype(np.float32)
    inv = np.zeros_like(mats)
    for i in range(20):
        R.ref_matrix3_inv(po.fp(mats[i]), po.fp(inv[i]))
    out["matrix3_inv"] = dict(m=f32list(mats), inv=f32list(inv))
    path = os.path.join(HERE, "ref_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
