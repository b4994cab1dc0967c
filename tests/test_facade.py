"""The C++ host facade (reference class names over the C ABI):
Scene::CreateScene reads an OBJ/MTL scene written by scenes.write_obj into
exactly the facet records, attribute table and textures the Python scene
generators hand to the C ABI (CPU), and a render through
Render::StochasticProgressivePhotonMapping matches the same render through
the C ABI and the oracle (GPU)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import scenes
import sppm_amd

PKG = os.path.dirname(sppm_amd.LIB_PATH)
RENDER_OBJ = os.path.join(PKG, "render_obj")


def read_dump(path):
    b = open(path, "rb").read()
    nf, na, nt = np.frombuffer(b[:12], np.int32)
    off = 12
    def take(n, dt):
        nonlocal off
        a = np.frombuffer(b[off:off + n * np.dtype(dt).itemsize], dt)
        off += n * np.dtype(dt).itemsize
        return a
    pos = take(nf * 9, np.float32).reshape(nf, 9)
    nrm = take(nf * 9, np.float32).reshape(nf, 9)
    uv = take(nf * 9, np.float32).reshape(nf, 9)
    attr = take(nf, np.int32)
    recs = []
    for _ in range(na):
        raw = b[off:off + C.sizeof(sppm_amd.Attr)]
        recs.append(sppm_amd.Attr.from_buffer_copy(raw))
        off += C.sizeof(sppm_amd.Attr)
    tex = []
    for _ in range(nt):
        w, h = take(2, np.int32)
        tex.append(take(w * h * 4, np.float32).reshape(h, w, 4))
    return pos, nrm, uv, attr, recs, tex


@pytest.mark.parametrize("which", ["box", "full"])
def test_obj_scene_matches_generator(which, tmp_path, box_scene, full_scene_small):
    sc = box_scene if which == "box" else full_scene_small
    obj = scenes.write_obj(sc, str(tmp_path), which)
    dump = str(tmp_path / "dump.bin")
    p = subprocess.run([RENDER_OBJ, obj, str(tmp_path / "x.pfm"), "--dump-scene", dump], capture_output=True,
                       text=True)
    assert p.returncode == 0, p.stderr
    pos, nrm, uv, attr, recs, tex = read_dump(dump)
    np.testing.assert_array_equal(pos, sc.pos)
    np.testing.assert_array_equal(nrm, sc.nrm)
    np.testing.assert_array_equal(uv, sc.uv)
    np.testing.assert_array_equal(attr, sc.attr)
    want = sppm_amd.attr_records(sc.attrs)
    assert len(recs) == len(want)
    for r, w, a in zip(recs, want, sc.attrs):
        for f in ("kind", "index", "if_delta", "if_opaque", "roughness", "reflectance", "extinction", "is_sppm"):
            assert getattr(r, f) == getattr(w, f), (a.get("name"), f)
        assert list(r.albedo) == list(w.albedo) and list(r.radiance) == list(w.radiance)
        # texture ids are numbered in first-use order: compare the images they name
        for f in ("tex_albedo", "tex_roughness", "tex_bump"):
            assert (getattr(r, f) < 0) == (getattr(w, f) < 0), (a.get("name"), f)
            if getattr(w, f) >= 0:
                np.testing.assert_array_equal(tex[getattr(r, f)], sc.textures[getattr(w, f)])


def test_obj_errors_are_reported(tmp_path):
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    p = subprocess.run([RENDER_OBJ, str(bad), str(tmp_path / "x.pfm"), "--dump-scene", str(tmp_path / "d")],
                       capture_output=True, text=True)
    assert p.returncode == 1 and "face without material" in p.stderr
    bad.write_text("mtllib missing.mtl\n")
    p = subprocess.run([RENDER_OBJ, str(bad), str(tmp_path / "x.pfm"), "--dump-scene", str(tmp_path / "d")],
                       capture_output=True, text=True)
    assert p.returncode == 1 and "cannot open material library" in p.stderr


def read_pfm(path):
    b = open(path, "rb").read()
    parts = b.split(b"\n", 3)
    assert parts[0] == b"PF"
    w, h = map(int, parts[1].split())
    img = np.frombuffer(parts[3], "<f4").reshape(h, w, 3)
    return img[::-1]


def rel_rmse(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-30))


@pytest.mark.gpu
@pytest.mark.parametrize("env", [False, True], ids=["box", "box+environment"])
def test_facade_render_matches_capi_and_oracle(tmp_path, box_scene, env):
    import pyoracle
    W, H, N, P = 64, 48, 20000, 3
    if env:
        box_scene = scenes.box_scene_ibl()
    obj = scenes.write_obj(box_scene, str(tmp_path), "box")
    out = str(tmp_path / "img.pfm")
    cam = scenes.pinhole_camera(W, H)
    args = [RENDER_OBJ, obj, out, "--size", str(W), str(H), "--photons", str(N), "--passes", str(P), "--pinhole",
            "--pos", *map(str, cam["pos"]), "--target", *map(str, cam["target"]), "--fov", str(cam["fov_deg"]),
            "--alpha", "0.9", "--radius", "0.015", "--bvh", "4"]
    if env:  # write_obj wrote the environment map as box_tex<id>.ppm
        args += ["--ibl", str(tmp_path / f"box_tex{box_scene.ibl}.ppm")]
    print("facade render:", " ".join(args[2:]), flush=True)
    p = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    img = read_pfm(out)
    # the same render through the C ABI
    r = sppm_amd.Renderer(box_scene, cam, 0.9, 0.015, N, bvh_width=4)
    for _ in range(P):
        r.pass_(stats=False)
    ref_gpu = r.reconstruct(N * P, P).reshape(H, W, 4)[..., :3]
    r.close()
    assert rel_rmse(img, ref_gpu) <= 1e-6
    # and the oracle
    o = pyoracle.OracleRender(pyoracle.OracleScene(box_scene), cam, 0.9, 0.015, N)
    for _ in range(P):
        o.full_pass()
    ref = o.reconstruct(N * P, P).reshape(H, W, 4)[..., :3]
    assert rel_rmse(img, ref) <= 1e-4
    assert "passes 3" in p.stdout


@pytest.mark.gpu
def test_facade_thinlens_diaphragm_matches_capi_and_oracle(tmp_path, box_scene):
    """ThinLensCamera(..., DiaphragmImageFileName) through the facade: the
    image is read from a binary PPM and renders as the C ABI with the same
    image and as the oracle."""
    import pyoracle
    W, H, N, P = 48, 48, 15000, 2
    obj = scenes.write_obj(box_scene, str(tmp_path), "box")
    img = scenes.aperture_image(24, 18, "hexagon")
    dia = str(tmp_path / "aperture.ppm")
    scenes.write_ppm(img, dia)
    cam = dict(scenes.thinlens_camera(W, H, aperture=0.05, focus=1.5), diaphragm=img)
    out = str(tmp_path / "img.pfm")
    args = [RENDER_OBJ, obj, out, "--size", str(W), str(H), "--photons", str(N), "--passes", str(P),
            "--thinlens", repr(cam["dist_to_lens"]), repr(cam["F"]), "--pos", *map(str, cam["pos"]),
            "--target", *map(str, cam["target"]), "--fov", str(cam["fov_deg"]), "--alpha", "0.9", "--radius",
            "0.015", "--bvh", "4", "--diaphragm", dia]
    print("facade render:", " ".join(args[2:]), flush=True)
    p = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    got = read_pfm(out)
    r = sppm_amd.Renderer(box_scene, cam, 0.9, 0.015, N, bvh_width=4)
    for _ in range(P):
        r.pass_(stats=False)
    ref_gpu = r.reconstruct(N * P, P).reshape(H, W, 4)[..., :3]
    r.close()
    assert rel_rmse(got, ref_gpu) <= 1e-6
    o = pyoracle.OracleRender(pyoracle.OracleScene(box_scene), cam, 0.9, 0.015, N)
    for _ in range(P):
        o.full_pass()
    assert rel_rmse(got, o.reconstruct(N * P, P).reshape(H, W, 4)[..., :3]) <= 1e-4
