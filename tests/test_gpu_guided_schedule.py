"""GPU tests of guided passes in the overlapped schedule.  The training
signal is summed in fixed point, so a guided render is bit-reproducible and
the overlapped passes (photon pass beside the previous gather, photons staged
out of the slots, proposal published before the build) must equal the same
passes run one step after another, and captured passes replayed."""
import numpy as np
import pytest

import scenes
import sppm_amd

pytestmark = pytest.mark.gpu

GUIDE = dict(mix_plain=0.5, learning_rate=1e-2, train_steps=1, seed=2, history_decay=0.95,
             target=sppm_amd.GUIDE_TARGET_RMS)


def renderer(box_scene, interval):
    r = sppm_amd.Renderer(box_scene, scenes.pinhole_camera(48, 40), 0.9, 0.015, 20000, bvh_width=4)
    r.guide_set(True, train_interval=interval, **GUIDE)
    return r


def assert_same(a, b):
    sa, sb = a.read_shared(), b.read_shared()
    for k in ("num_photon", "num_gather", "radius", "radiance"):
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)
    pa, ha = a.guide_read()
    pb, hb = b.guide_read()
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(ha, hb)
    assert ha.sum() > 0


def test_guided_render_is_reproducible(box_scene):
    a, b = renderer(box_scene, 2), renderer(box_scene, 2)
    for _ in range(6):
        a.pass_(stats=False)
        b.pass_(stats=False)
    a.sync()
    b.sync()
    assert_same(a, b)
    a.close()
    b.close()


@pytest.mark.parametrize("interval", [1, 3])
def test_guided_overlapped_passes_match_split_steps(box_scene, interval):
    whole, split = renderer(box_scene, interval), renderer(box_scene, interval)
    passes = 6
    for _ in range(passes):
        whole.pass_(stats=False)   # back to back: each photon pass leads
    for p in range(passes):
        split.photons(p + 1)
        split.build()
        split.eye(p + 1)
        split.gather()
    whole.sync()
    split.sync()
    print(f"guided, train every {interval}: {passes} passes", flush=True)
    assert_same(whole, split)
    whole.close()
    split.close()


def test_guided_graph_replay_matches_eager(box_scene):
    eager, graph = renderer(box_scene, 1), renderer(box_scene, 1)
    graph.graph_capture()
    for _ in range(5):
        eager.pass_(stats=False)
        graph.pass_(stats=False)
    eager.sync()
    graph.sync()
    assert_same(eager, graph)
    eager.close()
    graph.close()
