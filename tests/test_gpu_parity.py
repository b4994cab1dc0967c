"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bar: integer outputs bit-exact -- closest-hit facet ids (ties to
the lowest id), flattened photon lists (count and order), hit point types
and facets, per-hit-point neighbour index sets, M / N / NumGather; flux and
image within 1e-4 relative RMSE (FP32).  Geometry (photon positions,
directions, hit positions) is compared bit-exact as well, since both sides
execute the same float operations without contraction.
"""
import numpy as np
import pytest

import pyoracle
import scenes
import sppm_amd
from gpu_compare import REL_TOL, compare_pass, rel_rmse, say

pytestmark = pytest.mark.gpu


def random_rays(scene, n, seed):
    rng = np.random.default_rng(seed)
    org = rng.uniform(0.02, 0.98, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # a share of rays aimed at triangle vertices / edges (tie cases)
    k = n // 8
    tri = scene.pos[rng.integers(0, scene.nfacets, size=k)].reshape(k, 3, 3)
    w = rng.integers(0, 3, size=k)
    tgt = tri[np.arange(k), w]
    d[:k] = tgt - org[:k]
    d[:k] /= np.linalg.norm(d[:k], axis=1, keepdims=True)
    return org, d.astype(np.float32)


@pytest.mark.parametrize("builder", ["sah", "lbvh"])
@pytest.mark.parametrize("width", [2, 4, 8])
@pytest.mark.parametrize("which", ["box", "full"])
def test_closest_hit_bit_exact(width, which, builder, box_scene, full_scene_small):
    scene = box_scene if which == "box" else full_scene_small
    say(f"closest hit: scene={which} width={width} builder={builder}")
    org, d = random_rays(scene, 20000, seed=width * 7 + len(which))
    os_ = pyoracle.OracleScene(scene)
    fo, to, uo, vo = os_.closest_hit(org, d, mode=0)  # brute force
    flags = width | (sppm_amd.BVH_LBVH if builder == "lbvh" else 0)
    r = sppm_amd.Renderer(scene, scenes.pinhole_camera(8, 8), 0.9, 0.015, 64, bvh_width=flags)
    fg, tg, ug, vg = r.trace(org, d)
    r.close()
    np.testing.assert_array_equal(fg, fo)
    hit = fo >= 0
    np.testing.assert_array_equal(tg[hit], to[hit])
    np.testing.assert_array_equal(ug[hit], uo[hit])
    np.testing.assert_array_equal(vg[hit], vo[hit])


def small_scene(case):
    """The light quad plus a few triangles: random ones (2, 3, 5, 9, 33 in
    all), 64 coincident copies of one triangle (every centroid extent zero:
    the SAH build splits at the middle), or a 4 x 4 grid of coplanar quads."""
    attrs = [dict(name="white", kind=scenes.MAT_DIFFUSE, albedo=(0.7, 0.7, 0.7), is_sppm=1, if_opaque=1),
             dict(name="light", kind=scenes.MAT_DIFFUSE, albedo=(0.78, 0.78, 0.78), radiance=(15.0, 15.0, 15.0),
                  is_sppm=1, if_opaque=1)]
    b = scenes._Builder()
    scenes.light_quad(b, 1)
    rng = np.random.default_rng(11)
    z3 = [(0.0, 0.0, 1.0)] * 3
    t3 = [(0.0, 0.0, 0.0)] * 3
    if case == "coincident":
        tri = rng.uniform(0.2, 0.8, size=(3, 3))
        for _ in range(64):
            b.tri(tri, z3, t3, 0)
    elif case == "grid":
        for i in range(4):
            for j in range(4):
                x0, z0 = 0.1 + 0.2 * i, 0.1 + 0.2 * j
                b.quad((x0, 0.3, z0), (x0 + 0.2, 0.3, z0), (x0 + 0.2, 0.3, z0 + 0.2), (x0, 0.3, z0 + 0.2),
                       (0, 1, 0), 0)
    else:
        for _ in range(int(case) - 2):
            c = rng.uniform(0.1, 0.9, size=3)
            b.tri(c + rng.uniform(-0.15, 0.15, size=(3, 3)), z3, t3, 0)
    return b.scene(attrs)


@pytest.mark.parametrize("builder", ["sah", "lbvh"])
@pytest.mark.parametrize("width", [2, 4, 8])
@pytest.mark.parametrize("case", ["2", "3", "5", "9", "33", "coincident", "grid"])
def test_closest_hit_small_and_degenerate_scenes(case, width, builder):
    """The wide BVH (binary SAH or LBVH tree, collapsed by least SAH cost) on
    scenes of a few triangles and on degenerate ones: closest hits bit-exact
    against brute force."""
    scene = small_scene(case)
    say(f"closest hit, small scene {case} ({scene.nfacets} triangles), width={width} builder={builder}")
    org, d = random_rays(scene, 4000, seed=len(case) * 13 + width)
    os_ = pyoracle.OracleScene(scene)
    fo, to, uo, vo = os_.closest_hit(org, d, mode=0)
    flags = width | (sppm_amd.BVH_LBVH if builder == "lbvh" else 0)
    r = sppm_amd.Renderer(scene, scenes.pinhole_camera(8, 8), 0.9, 0.015, 64, bvh_width=flags)
    fg, tg, ug, vg = r.trace(org, d)
    r.close()
    np.testing.assert_array_equal(fg, fo)
    hit = fo >= 0
    assert hit.any()
    np.testing.assert_array_equal(tg[hit], to[hit])
    np.testing.assert_array_equal(ug[hit], uo[hit])
    np.testing.assert_array_equal(vg[hit], vo[hit])


def run_parity(scene, cam, alpha, r0, N, passes, width, label, record=True, nee=sppm_amd.NEE_OFF, guide=None,
               max_pairs=0):
    """record: True lists from the recording gather kernel, RECORD_PAIRS the
    production gather's own pairs, False none (counts only).  guide: kwargs of
    Renderer.guide_set; each pass then emits from the proposal the GPU holds
    before it, which the oracle is given too."""
    gpu = sppm_amd.Renderer(scene, cam, alpha, r0, N, bvh_width=width, record=record, nee=nee, max_pairs=max_pairs)
    oscene = pyoracle.OracleScene(scene)
    ref = pyoracle.OracleRender(oscene, cam, alpha, r0, N, nee=nee)
    og = None
    if guide is not None:
        gpu.guide_set(**guide)
        og = pyoracle.OracleGuide(oscene, guide.get("mix_plain", 0.5))
        ref.set_guide(og)
    for p in range(passes):
        say(f"{label}: pass {p}")
        if og is not None:
            og.set_tables(*gpu.guide_tables())
        st = gpu.pass_()
        assert st.flags == 0, f"device flags {st.flags:#x}"
        ref.full_pass(record=bool(record))
        stored_ref = pyoracle.orc().orc_flatten(ref.h)
        assert st.photons_stored == stored_ref
        compare_pass(gpu, ref, stored_ref, f"{label} pass {p}", record=record)
    img_g, img_r = gpu.reconstruct(), ref.reconstruct()
    assert np.isfinite(img_g).all()
    e = rel_rmse(img_g[..., :3], img_r[..., :3])
    gpu.close()
    assert e <= REL_TOL, f"{label}: image rel RMSE {e}"


def test_pass_parity_box_reference_constants(box_scene):
    # reference constants: alpha 0.9, FirstRadius 0.015 (SPPM.cpp:20-21)
    run_parity(box_scene, scenes.pinhole_camera(64, 64), 0.9, 0.015, 20000, 3, 4, "box a=0.9")


def test_pass_parity_box_config_constants(box_scene):
    # configured constants: alpha 0.7, r0 = 0.02 x scene diagonal
    r0 = 0.02 * box_scene.diagonal()
    run_parity(box_scene, scenes.pinhole_camera(64, 48), 0.7, r0, 15000, 3, 8, "box a=0.7")


def test_pass_parity_full_materials_thinlens(full_scene_small):
    r0 = 0.02 * full_scene_small.diagonal()
    run_parity(full_scene_small, scenes.thinlens_camera(48, 48), 0.7, r0, 15000, 2, 8, "full thinlens")


@pytest.mark.parametrize("kind,sampler", [("hexagon", sppm_amd.SAMPLER_UNIFORM), ("noise", sppm_amd.SAMPLER_UNIFORM),
                                          ("hexagon", sppm_amd.SAMPLER_HALTON)])
def test_pass_parity_thinlens_diaphragm(full_scene_small, kind, sampler):
    """ThinLensCamera with a diaphragm image (Camera.h:106-111, 136-158): lens
    points from the alias table over the aperture's pixels."""
    r0 = 0.02 * full_scene_small.diagonal()
    cam = dict(scenes.thinlens_camera(48, 48, aperture=0.05, focus=1.5), sampler=sampler,
               diaphragm=scenes.aperture_image(40, 30, kind, seed=3))
    run_parity(full_scene_small, cam, 0.7, r0, 15000, 2, 8, f"diaphragm {kind} sampler {sampler}")


def test_pass_parity_balanced_gather_box(box_scene):
    # production gather (chunked, load-balanced): counts and update exact, flux within tolerance
    r0 = 0.02 * box_scene.diagonal()
    run_parity(box_scene, scenes.pinhole_camera(96, 64), 0.7, r0, 30000, 4, 4, "box balanced", record=False)


def test_pass_parity_balanced_gather_full(full_scene_small):
    r0 = 0.02 * full_scene_small.diagonal()
    run_parity(full_scene_small, scenes.thinlens_camera(64, 64), 0.7, r0, 20000, 3, 8, "full balanced",
               record=False)


def test_pass_parity_environment_light():
    """Box lit also by an environment map through its open side: eye paths
    continue past their hit points and add the sky when they leave
    (SPPM.cpp:290-300); DirectLight, its counts and the image against the
    oracle."""
    sc = scenes.box_scene_ibl()
    cam = scenes.pinhole_camera(48, 40)
    cam["fov_deg"] = 70.0
    run_parity(sc, cam, 0.9, 0.015, 20000, 3, 4, "box + environment")


def test_halton_and_hammersley_samplers(box_scene):
    for smp, nsample in ((sppm_amd.SAMPLER_HALTON, 1), (sppm_amd.SAMPLER_HAMMERSLEY, 16)):
        cam = scenes.pinhole_camera(32, 32, sampler=smp)
        cam["sampler_nsample"] = nsample
        run_parity(box_scene, cam, 0.9, 0.015, 5000, 2, 2, f"sampler {smp}")


def test_full_size_box_properties(box_scene):
    """Box workload at full size (256^2, 100k photons, 16 passes, BVH4):
    properties that need no oracle run -- radius never grows, counters add
    up, the run is deterministic in its integer outputs, image finite."""
    cam = scenes.pinhole_camera(256, 256)
    r0 = 0.02 * box_scene.diagonal()
    runs = []
    for rep in range(2):
        say(f"full-size box run {rep}")
        g = sppm_amd.Renderer(box_scene, cam, 0.7, r0, 100000, bvh_width=4)
        prev_r = None
        msum = 0
        for p in range(16):
            st = g.pass_()
            assert st.flags == 0
            assert st.photons_emitted == 100000
            sh = g.read_shared()
            if prev_r is not None:
                assert (sh["radius"] <= prev_r).all()
            prev_r = sh["radius"]
            msum += int(g.read_counts().sum())
        c = g.counters()
        assert c["neighbours"] == msum and c["passes"] == 16
        img = g.reconstruct()
        assert np.isfinite(img).all()
        runs.append((sh["num_photon"].copy(), sh["num_gather"].copy(), c))
        g.close()
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
    assert runs[0][2] == runs[1][2]


@pytest.mark.parametrize("width", [2, 4, 8])
def test_occlusion_bit_exact(width):
    """Scene::IsOccluded through alpha-textured screens: GPU and oracle agree
    on every ray (random origins, directions, per-ray MaxDist, two
    thresholds)."""
    scene = scenes.screens_scene()
    say(f"occlusion: width={width}")
    org, d = random_rays(scene, 20000, seed=40 + width)
    rng = np.random.default_rng(width)
    tmax = rng.uniform(0.02, 1.8, size=len(org)).astype(np.float32)
    os_ = pyoracle.OracleScene(scene)
    r = sppm_amd.Renderer(scene, scenes.pinhole_camera(8, 8), 0.9, 0.015, 64, bvh_width=width)
    for tr in (0.5, 0.3):
        og = r.occluded(org, d, 1e-4, tmax, tr)
        oo = os_.occluded(org, d, 1e-4, tmax, tr)
        np.testing.assert_array_equal(og, oo)
        say(f"  threshold {tr}: {og.mean():.3f} occluded")
    # shadow rays as SPPM.cpp:205 casts them: to points on the light, MaxDist = distance - 2 RAY_OFFSET
    lp = np.stack([rng.uniform(0.375, 0.625, 4000), np.full(4000, 0.999), rng.uniform(0.375, 0.625, 4000)], 1)
    p = rng.uniform(0.02, 0.98, size=(4000, 3))
    v = lp - p
    dist = np.linalg.norm(v, axis=1)
    dirs = (v / dist[:, None]).astype(np.float32)
    mx = (dist.astype(np.float32) - np.float32(2e-4)).astype(np.float32)
    np.testing.assert_array_equal(r.occluded(p.astype(np.float32), dirs, 1e-4, mx, 0.5),
                                  os_.occluded(p.astype(np.float32), dirs, 1e-4, mx, 0.5))
    r.close()


@pytest.mark.parametrize("width", [4, 8])
def test_pass_parity_transmit_surfaces(width):
    """DiffuseTransmit, GlossTransmit and BiMetal surfaces (material.h:427-463,
    646-766): photons through and off them, hit points on them, against the
    oracle."""
    sc = scenes.transmit_scene()
    run_parity(sc, scenes.pinhole_camera(48, 48), 0.9, 0.015, 20000, 3, width, f"transmit w{width}")


@pytest.mark.parametrize("nee", [sppm_amd.NEE_AS_WRITTEN, sppm_amd.NEE_KEEP])
def test_pass_parity_eye_nee(nee):
    """Next-event estimation on eye paths (SPPM.cpp:144-168, 187-229): light
    samples with shadow rays at the conductor and glass spheres, MIS-weighted
    emitter hits, paths ending at the light."""
    run_parity(scenes.nee_scene(), scenes.pinhole_camera(48, 48), 0.9, 0.015, 20000, 3, 8, f"nee={nee}", nee=nee)


def test_pass_parity_eye_nee_textured_light(full_scene_small):
    """The same on the full-material scene: light samples from the textured
    light's 1000-point table, emitter hits priced by its table density."""
    r0 = 0.02 * full_scene_small.diagonal()
    run_parity(full_scene_small, scenes.thinlens_camera(48, 48), 0.7, r0, 15000, 2, 8, "full nee",
               nee=sppm_amd.NEE_KEEP)


def test_ray_counts_are_consistent(box_scene):
    """Every photon path traces at least its emission ray and at most the 33
    loop iterations of createPhotonMap (SPPM.cpp:549); every pixel traces at
    least its camera ray and at most 7 (eye depth <= 6, SPPM.cpp:99-100)."""
    N, W, H = 20000, 32, 32
    r = sppm_amd.Renderer(box_scene, scenes.pinhole_camera(W, H), 0.9, 0.015, N, bvh_width=4)
    for _ in range(3):
        r.pass_(stats=False)
    c = r.counters()
    r.close()
    say(f"rays: photon {c['photon_rays']} eye {c['eye_rays']}")
    assert 3 * N <= c["photon_rays"] <= 3 * N * 33
    assert 3 * W * H <= c["eye_rays"] <= 3 * W * H * 7


@pytest.mark.parametrize("which", ["box", "full"])
def test_production_gather_neighbour_sets(which, box_scene, full_scene_small):
    """The neighbour index sets of the production gather (k_gather_chunks,
    the kernel every measured pass runs) against the oracle's, bit-exact,
    with the update that follows."""
    if which == "box":
        r0 = 0.02 * box_scene.diagonal()
        run_parity(box_scene, scenes.pinhole_camera(96, 64), 0.7, r0, 30000, 3, 4, "pairs box",
                   record=sppm_amd.RECORD_PAIRS, max_pairs=40_000_000)
    else:
        r0 = 0.02 * full_scene_small.diagonal()
        run_parity(full_scene_small, scenes.thinlens_camera(64, 64), 0.7, r0, 20000, 3, 8, "pairs full",
                   record=sppm_amd.RECORD_PAIRS, max_pairs=40_000_000)


def test_pass_parity_guided_untextured(box_scene):
    """Guided emission (the proposal mixture, untextured light) against the
    oracle's restatement of it, each pass sampling the proposal trained by the
    GPU so far: photon lists bit-exact, PDFs and flux within 1e-4, counts and
    update exact."""
    r0 = 0.02 * box_scene.diagonal()
    run_parity(box_scene, scenes.pinhole_camera(64, 64), 0.7, r0, 20000, 4, 4, "guided box", record=False,
               guide=dict(mix_plain=0.5, learning_rate=1e-2, train_steps=4))


def test_pass_parity_guided_textured(full_scene_small):
    """The same with the textured light's 1000-point tables (table points
    picked inside the proposal's position cell)."""
    r0 = 0.02 * full_scene_small.diagonal()
    run_parity(full_scene_small, scenes.thinlens_camera(48, 48), 0.7, r0, 15000, 3, 8, "guided full",
               record=sppm_amd.RECORD_PAIRS, max_pairs=20_000_000,
               guide=dict(mix_plain=0.25, learning_rate=1e-2, train_steps=4))


def test_flux_is_reproducible(full_scene_small):
    """Photons sit in their grid cells in flattened order (a stable sort by
    cell), so two renders of the same setup give bit-identical flux."""
    r0 = 0.02 * full_scene_small.diagonal()
    out = []
    for rep in range(2):
        g = sppm_amd.Renderer(full_scene_small, scenes.thinlens_camera(96, 96), 0.7, r0, 40000, bvh_width=8)
        for _ in range(4):
            g.pass_(stats=False)
        out.append(g.read_shared())
        g.close()
    for k in ("radiance", "radius", "num_photon", "num_gather"):
        np.testing.assert_array_equal(out[0][k], out[1][k])
