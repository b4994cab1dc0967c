"""GPU tests of graph-replayed passes, multi-scene batches and the
autotuner, through the C ABI.  Each replay is checked against the same
render run eagerly (with eager work interleaved between replays): the whole
shared state bit for bit (photons sit in their grid cells in flattened
order, so the flux sums repeat exactly)."""
import numpy as np
import pytest

import scenes
import sppm_amd

pytestmark = pytest.mark.gpu


def assert_same_state(a, b, tol=0.0):
    sa, sb = a.read_shared(), b.read_shared()
    np.testing.assert_array_equal(sa["num_photon"], sb["num_photon"])
    np.testing.assert_array_equal(sa["num_gather"], sb["num_gather"])
    np.testing.assert_array_equal(sa["radius"], sb["radius"])
    ra, rb = sa["radiance"].astype(np.float64), sb["radiance"].astype(np.float64)
    err = np.sqrt(np.mean((ra - rb) ** 2)) / max(np.sqrt(np.mean(rb ** 2)), 1e-30)
    assert err <= tol, err
    np.testing.assert_array_equal(a.read_counts(), b.read_counts())


def test_graph_replay_matches_eager(box_scene):
    cam = scenes.pinhole_camera(48, 40)
    eager = sppm_amd.Renderer(box_scene, cam, 0.9, 0.015, 20000, bvh_width=4)
    graph = sppm_amd.Renderer(box_scene, cam, 0.9, 0.015, 20000, bvh_width=4)
    graph.graph_capture()
    for p in range(4):
        eager.pass_(stats=False)   # eager work between the replays
        graph.pass_(stats=False)
        eager.sync()
        graph.sync()
        print(f"pass {p + 1}", flush=True)
        assert_same_state(eager, graph)
    ie = eager.reconstruct()
    ig = graph.reconstruct()
    np.testing.assert_allclose(ig, ie, rtol=1e-4, atol=1e-6)
    eager.close()
    graph.close()


def test_graph_replay_back_to_back_matches_eager(full_scene_small):
    """Replays queued with no synchronisation between them: each photon pass
    runs beside the previous gather and each eye pass reads the camera words
    the previous eye graph made, alternating between the two word buffers.
    A step offset moves every stream's step counter alike."""
    cam = scenes.thinlens_camera(48, 32, aperture=0.02, focus=2.0)
    eager = sppm_amd.Renderer(full_scene_small, cam, 0.7, 0.03, 30000, bvh_width=8, step_offset=5)
    graph = sppm_amd.Renderer(full_scene_small, cam, 0.7, 0.03, 30000, bvh_width=8, step_offset=5)
    graph.graph_capture()
    for _ in range(7):
        eager.pass_(stats=False)
        graph.pass_(stats=False)
    eager.sync()
    graph.sync()
    assert_same_state(eager, graph)
    ce, cg = eager.counters(), graph.counters()
    for k in ("hit_points", "neighbours", "stored", "photon_rays", "eye_rays"):
        assert ce[k] == cg[k], k
    eager.close()
    graph.close()


@pytest.mark.parametrize("mode", [sppm_amd.BATCH_MULTI, sppm_amd.BATCH_BRANCHES], ids=["multi", "branches"])
def test_batch_matches_individual_renders(box_scene, full_scene_small, mode):
    occl = scenes.occluded_scene(clutter_freq=6)
    setups = [(box_scene, scenes.pinhole_camera(32, 32), 0.9, 0.015, 8000, 4),
              (full_scene_small, scenes.thinlens_camera(48, 32, aperture=0.02, focus=2.0), 0.7, 0.03, 12000, 8),
              (occl, scenes.occluded_camera(40, 40), 0.7, 0.05, 10000, 2)]
    solo = [sppm_amd.Renderer(s, c, a, r0, n, bvh_width=w) for s, c, a, r0, n, w in setups]
    inb = [sppm_amd.Renderer(s, c, a, r0, n, bvh_width=w) for s, c, a, r0, n, w in setups]
    batch = sppm_amd.Batch(inb, mode=mode)
    for p in range(3):
        for r in solo:
            r.pass_(stats=False)
        batch.pass_()
        batch.sync()
        for r in solo:
            r.sync()
        print(f"batch pass {p + 1}", flush=True)
        for a, b in zip(solo, inb):
            assert_same_state(a, b)
    batch.close()
    for r in solo + inb:
        r.close()


def test_multi_scene_batch_with_mixed_setups(box_scene, full_scene_small):
    """Twelve scenes in one multi-scene batch: BVH widths 2 / 4 / 8 (one
    photon and eye launch per width), eye-path NEE (its own eye launch), the
    wavefront photon kernel (its photon pass joins the shared megakernel
    launch), a diaphragm
    camera, work from 2k to 60k photons (waves move between scenes): every
    scene's state after each pass equals its own eager passes."""
    occl = scenes.occluded_scene(clutter_freq=6)
    box_ibl = scenes.box_scene_ibl()
    dia = dict(scenes.thinlens_camera(40, 40, aperture=0.05, focus=1.5), diaphragm=scenes.aperture_image(24, 18))
    setups = []
    for k in range(12):
        scene, cam = [(box_scene, scenes.pinhole_camera(24 + 4 * k, 24)),
                      (full_scene_small, scenes.thinlens_camera(32, 24 + 4 * k, aperture=0.02, focus=2.0)),
                      (occl, scenes.occluded_camera(32, 32)),
                      (box_ibl, dia)][k % 4]
        kw = dict(bvh_width=(2, 4, 8)[k % 3])
        if k == 5:
            kw["nee"] = sppm_amd.NEE_KEEP
        if k == 7:
            kw["photon_kernel"] = sppm_amd.PHOTON_WAVEFRONT
        setups.append((scene, cam, 0.8, 0.02 * scene.diagonal(), 2000 + 5000 * k, kw))
    solo = [sppm_amd.Renderer(s, c, a, r0, n, **kw) for s, c, a, r0, n, kw in setups]
    inb = [sppm_amd.Renderer(s, c, a, r0, n, **kw) for s, c, a, r0, n, kw in setups]
    batch = sppm_amd.Batch(inb)  # automatic mode: small passes -> multi-scene launches
    assert batch.mode == sppm_amd.BATCH_MULTI
    for p in range(3):
        for r in solo:
            r.pass_(stats=False)
        batch.pass_()
        batch.sync()
        for r in solo:
            r.sync()
        print(f"mixed batch pass {p + 1}", flush=True)
        for k, (a, b) in enumerate(zip(solo, inb)):
            assert_same_state(a, b)
            ca, cb = a.counters(), b.counters()
            for key in ("stored", "neighbours", "photon_rays", "eye_rays"):
                assert ca[key] == cb[key], (k, key, ca[key], cb[key])
    batch.close()
    for r in solo + inb:
        r.close()


def test_multi_scene_batch_guided_runs(box_scene):
    """Guided scenes in a multi-scene batch: the guided photon launch is its
    own group; the proposals train and the images are finite."""
    rs = []
    for k in range(3):
        r = sppm_amd.Renderer(box_scene, scenes.pinhole_camera(32, 32), 0.9, 0.015, 10000 + 5000 * k, bvh_width=4)
        if k != 1:
            r.guide_set(True, mix_plain=0.5, learning_rate=1e-2, train_steps=1, seed=2 + k, history_decay=0.95)
        rs.append(r)
    batch = sppm_amd.Batch(rs)
    for _ in range(3):
        batch.pass_()
    batch.sync()
    for k, r in enumerate(rs):
        if k != 1:
            prob, hist = r.guide_read()
            assert abs(prob.sum() - 1.0) < 1e-3 and hist.sum() > 0
        img = r.reconstruct()
        assert np.isfinite(img).all() and img[..., :3].mean() > 0
    batch.close()
    for r in rs:
        r.close()


def test_autotune_keeps_results(full_scene_small):
    cam = scenes.thinlens_camera(48, 48, aperture=0.02, focus=2.0)
    ref = sppm_amd.Renderer(full_scene_small, cam, 0.7, 0.03, 15000, bvh_width=8)
    tuned = sppm_amd.Renderer(full_scene_small, cam, 0.7, 0.03, 15000, bvh_width=8)
    res = tuned.autotune(full_scene_small, passes=2)
    print("autotune", res, flush=True)
    assert res["bvh_width"] in (2, 4, 8) and res["grid_blocks"] > 0
    assert res["bvh_builder"] in (0, sppm_amd.BVH_LBVH)
    assert res["photon_kernel"] in (sppm_amd.PHOTON_MEGAKERNEL, sppm_amd.PHOTON_WAVEFRONT)
    assert all(t > 0 for t in res["trial_ms"] + res["grid_ms"] + res["kernel_ms"] + res["photons_per_s"])
    # each stage keeps its fastest trial; the result is the last stage's choice
    k = 3 * (res["bvh_builder"] == sppm_amd.BVH_LBVH) + (2, 4, 8).index(res["bvh_width"])
    assert res["trial_ms"][k] == min(res["trial_ms"])
    assert res["ms_per_pass"] == min(res["kernel_ms"])
    assert res["photons_suggested"] in (7500, 15000, 30000)
    assert res["photons_per_s"][(7500, 15000, 30000).index(res["photons_suggested"])] == max(res["photons_per_s"])
    for _ in range(2):
        ref.pass_(stats=False)
        tuned.pass_(stats=False)
    ref.sync()
    tuned.sync()
    assert_same_state(ref, tuned)
    ref.close()
    tuned.close()


def test_guided_graph_replay_runs(box_scene):
    cam = scenes.pinhole_camera(32, 32)
    r = sppm_amd.Renderer(box_scene, cam, 0.9, 0.015, 20000, bvh_width=4)
    r.guide_set(True, mix_plain=0.5, learning_rate=1e-2, train_steps=1, seed=2, history_decay=0.95)
    r.graph_capture()
    for _ in range(3):
        r.pass_(stats=False)
    r.sync()
    prob, hist = r.guide_read()
    assert abs(prob.sum() - 1.0) < 1e-3 and hist.sum() > 0
    img = r.reconstruct()
    assert np.isfinite(img).all() and img[..., :3].mean() > 0
    r.close()
