"""sppm_trace_closest_static, the closest-hit query that allocates nothing
(rays and hits staged in one fixed device array per GPU), against
sppm_trace_closest, which allocates its staging per call: facet ids, t, u, v
bit-exact on two calls in a row, the second with fewer rays (the array keeps
the first call's contents), with and without a MaxT; and more rays than the
array holds are refused."""
import numpy as np
import pytest

import scenes
import sppm_amd

pytestmark = pytest.mark.gpu


def rays(scene, n, seed):
    rng = np.random.default_rng(seed)
    org = rng.uniform(0.02, 0.98, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    k = n // 8  # a share aimed at triangle vertices (ties)
    tri = scene.pos[rng.integers(0, scene.nfacets, size=k)].reshape(k, 3, 3)
    d[:k] = tri[np.arange(k), rng.integers(0, 3, size=k)] - org[:k]
    d[:k] /= np.linalg.norm(d[:k], axis=1, keepdims=True)
    return org, d.astype(np.float32)


def test_static_trace_matches_allocating_trace(full_scene_small):
    r = sppm_amd.Renderer(full_scene_small, scenes.pinhole_camera(8, 8), 0.9, 0.015, 64, bvh_width=8)
    for n, seed, tmax in ((sppm_amd.TRACE_STATIC_MAX_RAYS, 1, 3.4028234663852886e38), (5000, 2, 0.4), (777, 3, 1.0)):
        print(f"static trace: {n} rays, tmax {tmax}", flush=True)
        org, d = rays(full_scene_small, n, seed)
        fa, ta, ua, va = r.trace(org, d, tmax=tmax)
        fs, ts, us, vs = r.trace_static(org, d, tmax=tmax)
        assert (fa >= 0).sum() > n // 4
        np.testing.assert_array_equal(fs, fa)
        np.testing.assert_array_equal(ts, ta)
        np.testing.assert_array_equal(us, ua)
        np.testing.assert_array_equal(vs, va)
    org, d = rays(full_scene_small, sppm_amd.TRACE_STATIC_MAX_RAYS + 1, 4)
    with pytest.raises(sppm_amd.SppmError) as e:
        r.trace_static(org, d)
    assert "fixed buffer" in str(e.value)
    r.close()
