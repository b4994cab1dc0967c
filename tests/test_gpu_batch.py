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


def test_batch_matches_individual_renders(box_scene, full_scene_small):
    occl = scenes.occluded_scene(clutter_freq=6)
    setups = [(box_scene, scenes.pinhole_camera(32, 32), 0.9, 0.015, 8000, 4),
              (full_scene_small, scenes.thinlens_camera(48, 32, aperture=0.02, focus=2.0), 0.7, 0.03, 12000, 8),
              (occl, scenes.occluded_camera(40, 40), 0.7, 0.05, 10000, 2)]
    solo = [sppm_amd.Renderer(s, c, a, r0, n, bvh_width=w) for s, c, a, r0, n, w in setups]
    inb = [sppm_amd.Renderer(s, c, a, r0, n, bvh_width=w) for s, c, a, r0, n, w in setups]
    batch = sppm_amd.Batch(inb)
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


def test_autotune_keeps_results(full_scene_small):
    cam = scenes.thinlens_camera(48, 48, aperture=0.02, focus=2.0)
    ref = sppm_amd.Renderer(full_scene_small, cam, 0.7, 0.03, 15000, bvh_width=8)
    tuned = sppm_amd.Renderer(full_scene_small, cam, 0.7, 0.03, 15000, bvh_width=8)
    res = tuned.autotune(full_scene_small, passes=2)
    print("autotune", res, flush=True)
    assert res["bvh_width"] in (2, 4, 8) and res["grid_blocks"] > 0
    assert all(t > 0 for t in res["trial_ms"]) and res["ms_per_pass"] == min(res["trial_ms"])
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
