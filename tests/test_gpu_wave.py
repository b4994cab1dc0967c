"""The wavefront photon pass against the one-lane-per-path megakernel (the
default, which the whole-pass oracle tests run), through the C ABI: both run
createPhotonMap (R/SPPM.cpp:495-680), so every stored photon (position, flux,
direction, IOR, PDF), every path's photon count and the closest-hit ray and
tie counts must be identical, for every BVH width, lanes per ray and number of
wavefront bounces before the one-lane-per-path tail."""
import numpy as np
import pytest

import scenes
import sppm_amd
from gpu_compare import say

pytestmark = pytest.mark.gpu


def same_photons(a, b, label):
    pa, pb = a.read_photons(), b.read_photons()
    for k in pa:
        np.testing.assert_array_equal(pa[k], pb[k], err_msg=f"{label}: {k}")
    np.testing.assert_array_equal(a.read_path_counts(), b.read_path_counts(), err_msg=f"{label}: path counts")
    ca, cb = a.counters(), b.counters()
    for k in ("photon_rays", "photon_tie_rays", "stored"):
        assert ca[k] == cb[k], f"{label}: {k} {ca[k]} != {cb[k]}"
    return pa["pos"].shape[0]


@pytest.mark.parametrize("width,lanes", [(2, 2), (4, 2), (4, 4), (8, 2), (8, 4), (8, 8)])
def test_wavefront_equals_megakernel(box_scene, full_scene_small, width, lanes):
    for name, scene, cam in (("box", box_scene, scenes.pinhole_camera(32, 32)),
                             ("full", full_scene_small, scenes.thinlens_camera(48, 32, aperture=0.02, focus=2.0))):
        mk = sppm_amd.Renderer(scene, cam, 0.8, 0.02, 30_000, bvh_width=width)
        wv = sppm_amd.Renderer(scene, cam, 0.8, 0.02, 30_000, bvh_width=width,
                               photon_kernel=sppm_amd.PHOTON_WAVEFRONT, wave_lanes=lanes)
        for p in range(3):
            mk.pass_(stats=False)
            wv.pass_(stats=False)
            n = same_photons(wv, mk, f"{name} W{width} L{lanes} pass {p + 1}")
        sa, sb = wv.read_shared(), mk.read_shared()
        for k in sa:
            np.testing.assert_array_equal(sa[k], sb[k], err_msg=f"{name} W{width} L{lanes}: shared {k}")
        say(f"{name} W{width} L{lanes}: wavefront == megakernel over 3 passes ({n} photons in the last)")
        mk.close()
        wv.close()


@pytest.mark.parametrize("bounces", [1, 3, 32])
def test_wavefront_bounce_counts(full_scene_small, bounces):
    cam = scenes.thinlens_camera(32, 32, aperture=0.02, focus=2.0)
    mk = sppm_amd.Renderer(full_scene_small, cam, 0.8, 0.02, 20_000, bvh_width=8)
    wv = sppm_amd.Renderer(full_scene_small, cam, 0.8, 0.02, 20_000, bvh_width=8,
                           photon_kernel=sppm_amd.PHOTON_WAVEFRONT, wave_bounces=bounces)
    for p in range(2):
        mk.pass_(stats=False)
        wv.pass_(stats=False)
        same_photons(wv, mk, f"{bounces} bounces, pass {p + 1}")
    mk.close()
    wv.close()


def test_wavefront_graph_replay(box_scene):
    cam = scenes.pinhole_camera(32, 32)
    wf = sppm_amd.PHOTON_WAVEFRONT
    eager = sppm_amd.Renderer(box_scene, cam, 0.9, 0.015, 20_000, bvh_width=8, photon_kernel=wf)
    graph = sppm_amd.Renderer(box_scene, cam, 0.9, 0.015, 20_000, bvh_width=8, photon_kernel=wf)
    graph.graph_capture()
    for p in range(3):
        eager.pass_(stats=False)
        graph.pass_(stats=False)
        eager.sync()
        graph.sync()
        same_photons(graph, eager, f"graph replay {p + 1}")
    eager.close()
    graph.close()
