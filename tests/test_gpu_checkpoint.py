"""GPU tests of checkpoint / resume through the C ABI: a render saved after
some passes and restored into a fresh context continues exactly as the
uninterrupted render -- the photon lists and hit points of the following
passes bit-exact (camera stream, path seeds and stale photon slots carried
over), N / NumGather / radius exact, flux to 1e-5 (the sum order inside a
grid cell follows atomic ranks)."""
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
    for k in ("pos", "indir", "index"):
        np.testing.assert_array_equal(pa[k], pb[k])
    ha, hb = a.read_hitpoints(), b.read_hitpoints()
    np.testing.assert_array_equal(ha["type"], hb["type"])
    np.testing.assert_array_equal(ha["pos"], hb["pos"])
    sa, sb = a.read_shared(), b.read_shared()
    np.testing.assert_array_equal(sa["num_photon"], sb["num_photon"])
    np.testing.assert_array_equal(sa["num_gather"], sb["num_gather"])
    np.testing.assert_array_equal(sa["radius"], sb["radius"])
    assert rel(sa["radiance"], sb["radiance"]) <= 1e-5
    assert rel(a.reconstruct(), b.reconstruct()) <= 1e-5


@pytest.mark.parametrize("which", ["box", "full"])
def test_resume_continues_the_render(which, box_scene, full_scene_small):
    if which == "box":
        scene, cam, a, r0, n, w = box_scene, scenes.pinhole_camera(48, 40), 0.7, 0.03, 20000, 4
    else:
        scene, cam, a, r0, n, w = full_scene_small, scenes.thinlens_camera(48, 48, aperture=0.02, focus=2.0), 0.7, \
            0.03, 15000, 8
    straight = sppm_amd.Renderer(scene, cam, a, r0, n, bvh_width=w)
    first = sppm_amd.Renderer(scene, cam, a, r0, n, bvh_width=w)
    for _ in range(5):
        straight.pass_(stats=False)
    for _ in range(2):
        first.pass_(stats=False)
    ckpt = first.checkpoint()
    first.close()
    print(f"{which}: checkpoint {ckpt[0].size} bytes after {ckpt[1]} passes", flush=True)
    resumed = sppm_amd.Renderer(scene, cam, a, r0, n, bvh_width=w)
    resumed.restore(ckpt)
    for _ in range(3):
        resumed.pass_(stats=False)
    straight.sync()
    resumed.sync()
    assert resumed.passes == 5 and resumed.total_photons == 5 * n
    same_render(straight, resumed)
    straight.close()
    resumed.close()


def test_resume_restores_the_proposal(box_scene):
    cam = scenes.pinhole_camera(32, 32)
    r = sppm_amd.Renderer(box_scene, cam, 0.9, 0.015, 20000, bvh_width=4)
    r.guide_set(True, mix_plain=0.5, learning_rate=1e-2, train_steps=1, seed=4, history_decay=0.95)
    for _ in range(3):
        r.pass_(stats=False)
    r.sync()
    ckpt = r.checkpoint()
    w0 = r.guide_weights()
    p0, h0 = r.guide_read()
    r2 = sppm_amd.Renderer(box_scene, cam, 0.9, 0.015, 20000, bvh_width=4)
    r2.guide_set(True, mix_plain=0.5, learning_rate=1e-2, train_steps=1, seed=99, history_decay=0.95)
    r2.restore(ckpt)
    np.testing.assert_array_equal(r2.guide_weights(), w0)
    p1, h1 = r2.guide_read()
    np.testing.assert_array_equal(h1, h0)
    np.testing.assert_array_equal(p1, p0)  # the same forward of the same weights
    for _ in range(2):
        r2.pass_(stats=False)
    r2.sync()
    assert np.isfinite(r2.reconstruct()).all()
    r.close()
    r2.close()


def test_guided_resume_equals_uninterrupted(box_scene):
    """Guided renders are bit-reproducible (fixed-point training sums), so a
    render checkpointed after 3 passes and resumed for 4 more equals 7
    straight passes: image, proposal and histogram, with deposits pending
    between two trainings at the checkpoint (training every 2nd pass)."""
    cam = scenes.pinhole_camera(48, 40)
    guide = dict(mix_plain=0.5, learning_rate=1e-2, train_steps=1, seed=4, history_decay=0.95, train_interval=2,
                 target=sppm_amd.GUIDE_TARGET_RMS)

    def new():
        r = sppm_amd.Renderer(box_scene, cam, 0.7, 0.03, 20000, bvh_width=4)
        r.guide_set(True, **guide)
        return r
    straight, first = new(), new()
    for _ in range(7):
        straight.pass_(stats=False)
    for _ in range(3):
        first.pass_(stats=False)
    ckpt = first.checkpoint()
    first.close()
    resumed = new()
    resumed.restore(ckpt)
    for _ in range(4):
        resumed.pass_(stats=False)
    straight.sync()
    resumed.sync()
    same_render(straight, resumed)
    ps, hs = straight.guide_read()
    pr, hr = resumed.guide_read()
    np.testing.assert_array_equal(pr, ps)
    np.testing.assert_array_equal(hr, hs)
    np.testing.assert_array_equal(resumed.guide_weights(), straight.guide_weights())
    straight.close()
    resumed.close()


def test_checkpoint_of_another_setup_is_refused(box_scene):
    a = sppm_amd.Renderer(box_scene, scenes.pinhole_camera(32, 32), 0.9, 0.015, 5000, bvh_width=4)
    a.pass_(stats=False)
    ckpt = a.checkpoint()
    b = sppm_amd.Renderer(box_scene, scenes.pinhole_camera(32, 16), 0.9, 0.015, 5000, bvh_width=4)
    with pytest.raises(RuntimeError, match="another resolution"):
        b.restore(ckpt)
    a.close()
    b.close()
