"""CPU-side checks of the product boundary and host logic (no GPU calls):
the C-ABI library loads and exports every function include/sppm_amd.h
declares; error paths return codes instead of crashing; the scene
generators produce valid facet records; bench.py refuses N > 1 GPUs."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import scenes
import sppm_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "sppm_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sppm_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_function():
    lib = C.CDLL(sppm_amd.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(sppm_amd.EXPORTED) == names


def test_abi_version_and_error_paths_without_gpu():
    lib = sppm_amd.load_library()
    assert lib.sppm_abi_version() == 6
    # NULL arguments are rejected before any device work
    assert lib.sppm_create(None, 0, None) == -1
    assert b"NULL" in lib.sppm_last_error()
    assert lib.sppm_destroy(None) == 0
    assert lib.sppm_sync(None) == -1
    assert lib.sppm_scene_upload(None, None, 8) == -1
    assert lib.sppm_camera_set(None, None) == -1


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK), reason="a GPU is visible")
def test_create_without_device_fails_cleanly():
    lib = sppm_amd.load_library()
    h = C.c_void_p()
    rc = lib.sppm_create(C.byref(h), 0, None)
    assert rc in (-1, -2)
    assert lib.sppm_last_error()


def test_box_scene_records():
    sc = scenes.box_scene()
    assert sc.pos.shape == (sc.nfacets, 9) and sc.nrm.shape == sc.pos.shape and sc.uv.shape == sc.pos.shape
    assert sc.nfacets == 12 + 2 * 1520
    # two emitting facets (the 0.25 x 0.25 ceiling quad), facing down
    emit = [i for i in range(sc.nfacets) if sum(sc.attrs[sc.attr[i]].get("radiance", (0, 0, 0))) > 1e-5]
    assert emit == [10, 11]
    n = sc.nrm[10].reshape(3, 3).sum(0)
    assert n[1] < 0
    area = 0.5 * np.linalg.norm(np.cross(sc.pos[10, 3:6] - sc.pos[10, :3], sc.pos[10, 6:] - sc.pos[10, :3]))
    assert abs(area * 2 - 0.0625) < 1e-6
    # no degenerate triangles
    e1 = sc.pos[:, 3:6] - sc.pos[:, :3]
    e2 = sc.pos[:, 6:] - sc.pos[:, :3]
    assert (np.linalg.norm(np.cross(e1, e2), axis=1) > 1e-9).all()
    assert abs(sc.diagonal() - np.sqrt(3)) < 1e-6


def test_full_scene_records():
    sc = scenes.full_scene(freq=10, tex_size=64)
    assert sc.nfacets == 12 + 2 * 1520 + 20 * 100
    assert len(sc.textures) == 4
    chk = sc.textures[3]
    assert chk.shape == (256, 256, 4)
    vals = np.unique(np.round(chk[..., 0] * 255).astype(int))
    np.testing.assert_array_equal(vals, [5, 20])
    assert (chk[..., 3] == 1).all()


def test_thinlens_parameters():
    cam = scenes.thinlens_camera(1024, 1024, aperture=0.02, focus=2.0)
    # lens width == image plane width == aperture; focus B = A F / (A - F) = 2
    half = np.float32(np.pi) * np.float32(45.0) / np.float32(360.0)
    plane = 2.0 * cam["dist_to_lens"] * np.tan(float(half)) * cam["aspect"]
    assert abs(plane - 0.02) < 1e-9
    A, F = cam["dist_to_lens"], cam["F"]
    assert abs(A * F / (A - F) - 2.0) < 1e-9


def test_bench_refuses_multi_gpu():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert p.returncode == 2
    assert "one GPU" in p.stderr


def test_float_thresholds_equal_double_compares():
    """The device replaces (double)x > c and (double)x < c by float compares
    with lo(c) (largest float <= c) and hi(c) (smallest float >= c); each
    constant must be exactly that neighbour of the double constant it names."""
    import re
    src = open(os.path.join(os.path.dirname(__file__), "..", "friendsrender-r-public_amd", "csrc",
                            "sppm_device.h")).read()
    found = re.findall(r"constexpr float (\w+) = (0x[0-9a-fp.+-]+)f;\s*// (lo|hi)\(([0-9.e-]+)\)", src)
    assert len(found) == 10
    for name, hexv, side, c in found:
        v = np.float32(float.fromhex(hexv))
        assert float(v) == float.fromhex(hexv), name  # representable as written
        c = float(c)
        up, down = np.nextafter(v, np.float32(np.inf)), np.nextafter(v, np.float32(-np.inf))
        if side == "lo":
            assert float(v) <= c < float(up), name
        else:
            assert float(down) < c <= float(v), name
        # and the comparisons agree on the floats around c
        x = down
        for _ in range(5):
            if side == "lo":
                assert (float(x) > c) == bool(x > v)
            else:
                assert (float(x) < c) == bool(x < v)
            x = np.nextafter(x, np.float32(np.inf))
