"""Setup changes against captured passes, batches and checkpoints, through the
C ABI:

* a scene upload drops a captured pass and un-begins the render: the next
  sppm_pass is refused (SPPM_E_STATE) until sppm_begin, and the render after
  begin equals a fresh one;
* changing guided emission drops a captured pass (the graph held the old
  proposal's buffers), and the passes after it run eagerly;
* a context in a batch refuses its own passes and setup changes, its reads
  wait for the batch's last launch, and destroying it makes the batch refuse
  to launch;
* a checkpoint taken with another scene, camera or alpha is refused.
"""
import numpy as np
import pytest

import scenes
import sppm_amd

pytestmark = pytest.mark.gpu


def test_scene_upload_drops_captured_pass(box_scene):
    cam = scenes.pinhole_camera(32, 32)
    r = sppm_amd.Renderer(box_scene, cam, 0.9, 0.015, 8000, bvh_width=4)
    r.graph_capture()
    r.pass_(stats=False)
    other = scenes.box_scene(with_spheres=False)
    r.upload_scene(other, 4)
    with pytest.raises(sppm_amd.SppmError) as e:
        r.pass_(stats=False)
    assert "sppm_begin" in str(e.value)
    r.begin()
    fresh = sppm_amd.Renderer(other, cam, 0.9, 0.015, 8000, bvh_width=4)
    for _ in range(2):
        r.pass_(stats=False)
        fresh.pass_(stats=False)
    a, b = r.read_shared(), fresh.read_shared()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])
    r.close()
    fresh.close()


def test_guide_change_drops_captured_pass(box_scene):
    cam = scenes.pinhole_camera(32, 32)
    r = sppm_amd.Renderer(box_scene, cam, 0.9, 0.015, 8000, bvh_width=4)
    r.graph_capture()
    r.pass_(stats=False)
    # turning guided emission on after the capture: the replay would ignore it
    r.guide_set(True, mix_plain=0.5, learning_rate=1e-2, train_steps=1, seed=2)
    r.pass_(stats=False)  # eager, guided
    prob, hist = r.guide_read()
    assert hist.sum() > 0  # the pass deposited training records: it ran guided
    # split steps are available again (they are refused while a graph is held)
    r.photons(3)
    r.close()


def test_batch_membership(box_scene):
    cam = scenes.pinhole_camera(32, 32)
    a = sppm_amd.Renderer(box_scene, cam, 0.9, 0.015, 8000, bvh_width=4)
    b = sppm_amd.Renderer(box_scene, cam, 0.9, 0.015, 8000, bvh_width=4)
    solo = sppm_amd.Renderer(box_scene, cam, 0.9, 0.015, 8000, bvh_width=4)
    batch = sppm_amd.Batch([a, b])
    for _ in range(3):
        batch.pass_()
        solo.pass_(stats=False)
    # reads of a member wait for the batch's last launch (no batch sync here)
    sa, ss = a.read_shared(), solo.read_shared()
    for k in sa:
        np.testing.assert_array_equal(sa[k], ss[k])
    for call in (lambda: a.pass_(stats=False), lambda: a.begin(), lambda: a.upload_scene(box_scene, 4),
                 lambda: a.photons(1), lambda: a.guide_set(True)):
        with pytest.raises(sppm_amd.SppmError) as e:
            call()
        assert "batch" in str(e.value)
    # a destroyed member: the batch can no longer launch
    b.close()
    with pytest.raises(sppm_amd.SppmError):
        batch.pass_()
    batch.close()
    # after the batch is gone its former member is free again
    a.begin()
    a.pass_(stats=False)
    a.close()
    solo.close()


def test_checkpoint_refuses_other_setups(box_scene):
    cam = scenes.pinhole_camera(32, 32)
    r = sppm_amd.Renderer(box_scene, cam, 0.9, 0.015, 8000, bvh_width=4)
    r.pass_(stats=False)
    ck = r.checkpoint()
    cam2 = dict(cam)
    cam2["pos"] = (0.5, 0.55, 2.2)
    others = [sppm_amd.Renderer(scenes.box_scene(with_spheres=False), cam, 0.9, 0.015, 8000, bvh_width=4),
              sppm_amd.Renderer(box_scene, cam2, 0.9, 0.015, 8000, bvh_width=4),
              sppm_amd.Renderer(box_scene, cam, 0.7, 0.015, 8000, bvh_width=4)]
    for o in others:
        with pytest.raises(sppm_amd.SppmError) as e:
            o.restore(ck)
        assert "another scene, camera" in str(e.value)
        o.close()
    # the same setup, other BVH width: accepted (closest hits do not depend on it)
    same = sppm_amd.Renderer(box_scene, cam, 0.9, 0.015, 8000, bvh_width=8)
    same.restore(ck)
    same.pass_(stats=False)
    r.pass_(stats=False)
    a, b = r.read_shared(), same.read_shared()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])
    r.close()
    same.close()
