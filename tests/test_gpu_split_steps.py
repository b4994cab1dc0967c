"""GPU tests of the split pass steps (sppm_pass_photons / _build / _eye /
_gather, the reference's createPhotonMap(step), kd-tree build,
createHitpoints(step) and applyHitPointInfo called one by one, SPPM.cpp:770-843)
against whole passes (sppm_pass): the full pass runs the eye pass on a second
stream beside the photon pass and the build, the split steps run everything
in order on one stream, and both must give the same render -- photon lists
and hit points bit-exact, N / NumGather / radius exact, flux to 1e-5 (the sum
order inside a grid cell follows atomic ranks)."""
import numpy as np
import pytest

import scenes
import sppm_amd

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-30)


def same_render(a, b):
    pa, pb = a.read_photons(), b.read_photons()
    for k in ("pos", "flux", "indir", "index"):
        np.testing.assert_array_equal(pa[k], pb[k])
    ha, hb = a.read_hitpoints(), b.read_hitpoints()
    for k in ("type", "facet", "pos", "weight"):
        np.testing.assert_array_equal(ha[k], hb[k])
    sa, sb = a.read_shared(), b.read_shared()
    np.testing.assert_array_equal(sa["num_photon"], sb["num_photon"])
    np.testing.assert_array_equal(sa["num_gather"], sb["num_gather"])
    np.testing.assert_array_equal(sa["radius"], sb["radius"])
    assert rel(sa["radiance"], sb["radiance"]) <= 1e-5


@pytest.mark.parametrize("which", ["box", "full", "box_nee"])
def test_split_steps_match_whole_passes(which, box_scene, full_scene_small):
    nee = sppm_amd.NEE_KEEP if which == "box_nee" else sppm_amd.NEE_OFF
    if which == "full":
        scene, cam, n, w = full_scene_small, scenes.thinlens_camera(48, 48, aperture=0.02, focus=2.0), 15000, 8
    else:
        scene, cam, n, w = box_scene, scenes.pinhole_camera(48, 40), 20000, 4
    whole = sppm_amd.Renderer(scene, cam, 0.7, 0.03, n, bvh_width=w, nee=nee)
    split = sppm_amd.Renderer(scene, cam, 0.7, 0.03, n, bvh_width=w, nee=nee)
    passes = 4
    for _ in range(passes):
        whole.pass_(stats=False)
    for p in range(passes):
        assert split.photons(p + 1) == n
        stored = split.build()
        split.eye(p + 1)
        split.gather()
        print(f"{which}: split pass {p + 1}, {stored} photons stored", flush=True)
    whole.sync()
    split.sync()
    same_render(whole, split)
    np.testing.assert_allclose(split.reconstruct(total_photons=passes * n, passes=passes),
                               whole.reconstruct(total_photons=passes * n, passes=passes), rtol=1e-4, atol=1e-6)
    whole.close()
    split.close()


def test_calls_between_passes_keep_the_render(box_scene):
    """Whole passes with readbacks, ray queries and counter reads between
    them (each of which makes the next pass wait for the main stream instead
    of starting its photon pass beside the previous gather) == the same
    passes back to back."""
    cam = scenes.pinhole_camera(48, 40)
    plain = sppm_amd.Renderer(box_scene, cam, 0.7, 0.03, 20000, bvh_width=4)
    mixed = sppm_amd.Renderer(box_scene, cam, 0.7, 0.03, 20000, bvh_width=4)
    for _ in range(6):
        plain.pass_(stats=False)
    rng = np.random.default_rng(3)
    org = np.tile(np.array([[0.0, 0.5, 0.5]], np.float32), (256, 1))
    dirs = rng.normal(size=(256, 3)).astype(np.float32)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    mixed.pass_(stats=False)
    mixed.pass_(stats=False)
    n1 = len(mixed.read_photons()["pos"])
    mixed.pass_(stats=False)
    mixed.trace(org, dirs)
    mixed.pass_(stats=True)
    mixed.counters()
    mixed.pass_(stats=False)
    mixed.read_hitpoints()
    mixed.pass_(stats=False)
    print(f"interleaved: {n1} photons stored after pass 2", flush=True)
    plain.sync()
    mixed.sync()
    same_render(plain, mixed)
    plain.close()
    mixed.close()
