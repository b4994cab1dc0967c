"""Every configured workload at its configured size on the GPU, against the
CPU oracle (the reference's pass loop, SPPM.cpp:770-854, restated), through
the C ABI.

  box          256^2, 100k photons/pass, 16 passes, BVH4, pinhole, alpha 0.7,
               r0 = 0.02 x diagonal: every pass whole against the oracle
               (photon lists, hit points, M / N / NumGather bit-exact, the
               production gather's neighbour index sets on three passes,
               radiance within 1e-4), and the final image.
  full-material 1024^2, 1M photons, 203k triangles, thin lens, BVH8: three
               whole passes against the oracle.
  64 scenes    the batched workload (31M triangles, 26M photons per pass):
               each scene's state after two batched graph passes equals the
               same render run eagerly, bit for bit; the two smallest scenes
               also whole-pass against the oracle.
  occluded     1024^2, 1M photons, 297,700 triangles: two plain passes whole
               against the oracle; two guided passes (64 x 4 bf16 network,
               mixture 0.5) against the oracle given the GPU's proposal; at a
               fixed radius guided and plain estimate the same image.
  caustic      2M triangles of GGX glass, 2048^2 textures, thin lens 2048^2,
               4M photons: spot parity of plain passes (3000 pixels, 1500
               photon paths per pass) and of guided ones (128 x 6 fp16).
"""
import numpy as np
import pytest

import pyoracle
import scenes
import sppm_amd
from gpu_compare import compare_pass, oracle_pass, rel_rmse, say, spot_check

pytestmark = pytest.mark.gpu


def test_box_full_size_every_pass(box_scene):
    cam = scenes.pinhole_camera(256, 256)
    r0 = 0.02 * box_scene.diagonal()
    N = 100_000
    gpu = sppm_amd.Renderer(box_scene, cam, 0.7, r0, N, bvh_width=4, record=sppm_amd.RECORD_PAIRS,
                            max_pairs=60_000_000)
    ref = pyoracle.OracleRender(pyoracle.OracleScene(box_scene), cam, 0.7, r0, N)
    for p in range(16):
        st = gpu.pass_()
        assert st.flags == 0 and st.photons_emitted == N
        lists = p in (0, 7, 15)
        stored = oracle_pass(ref, record=lists)
        assert st.photons_stored == stored
        compare_pass(gpu, ref, stored, f"box 256^2 pass {p}", record=lists)
    e = rel_rmse(gpu.reconstruct()[..., :3], ref.reconstruct()[..., :3])
    say(f"box 256^2: image rel RMSE {e:.2e}")
    assert e <= 1e-4
    gpu.close()


@pytest.fixture(scope="module")
def full_material():
    scene = scenes.full_scene(freq=100, tex_size=512)
    return scene, scenes.thinlens_camera(1024, 1024), pyoracle.OracleScene(scene)


def test_full_material_whole_passes(full_material):
    scene, cam, oscene = full_material
    r0 = 0.02 * scene.diagonal()
    N = 1_000_000
    gpu = sppm_amd.Renderer(scene, cam, 0.7, r0, N, bvh_width=8)
    ref = pyoracle.OracleRender(oscene, cam, 0.7, r0, N)
    for p in range(3):
        st = gpu.pass_()
        assert (st.flags & ~0x31) == 0, f"flags {st.flags:#x}"
        stored = oracle_pass(ref)
        assert st.photons_stored == stored
        compare_pass(gpu, ref, stored, f"full-material 1024^2 pass {p}", record=False)
    gpu.close()


@pytest.fixture(scope="module")
def batch64():
    say("generating the 64-scene workload")
    return scenes.batch_workload(64, progress=lambda s: s % 16 == 0 and say(f"  scene {s}"))


def test_batch_64_scenes_match_eager_and_oracle(batch64):
    work = batch64
    K = 2
    rs = []
    for k, w in enumerate(work):
        rs.append(sppm_amd.Renderer(w["scene"], w["cam"], w["alpha"], w["r0"], w["photons"], bvh_width=8))
        if k % 16 == 15:
            say(f"  {k + 1} scenes uploaded")
    b = sppm_amd.Batch(rs)
    for _ in range(K):
        b.pass_()
    b.sync()
    saved = [(r.read_shared(), r.read_counts()) for r in rs]  # readbacks wait for the batch
    b.close()
    say(f"batch: {K} graph passes of {len(rs)} scenes done")
    size = [w["scene"].nfacets * w["cam"]["width"] ** 2 for w in work]
    spot = set(int(j) for j in np.argsort(size)[:2])
    for j, (r, w) in enumerate(zip(rs, work)):
        r.begin()
        ref = pyoracle.OracleRender(pyoracle.OracleScene(w["scene"]), w["cam"], w["alpha"], w["r0"],
                                    w["photons"]) if j in spot else None
        for p in range(K):
            st = r.pass_()
            assert (st.flags & ~0x31) == 0
            if ref is not None:
                stored = oracle_pass(ref)
                compare_pass(r, ref, stored, f"scene {j} pass {p}", record=False)
        sh, m = r.read_shared(), r.read_counts()
        for key in ("radiance", "radius", "num_photon", "num_gather"):
            np.testing.assert_array_equal(sh[key], saved[j][0][key], err_msg=f"scene {j} {key}")
        np.testing.assert_array_equal(m, saved[j][1])
        r.close()
        if j % 16 == 15:
            say(f"  scenes 0..{j}: batch == eager")


@pytest.fixture(scope="module")
def occluded():
    scene = scenes.occluded_scene()
    return scene, scenes.occluded_camera(1024, 1024), pyoracle.OracleScene(scene)


def test_occluded_plain_whole_passes(occluded):
    scene, cam, oscene = occluded
    assert scene.nfacets > 290_000
    r0 = 0.02 * scene.diagonal()
    gpu = sppm_amd.Renderer(scene, cam, 0.7, r0, 1_000_000, bvh_width=8)
    ref = pyoracle.OracleRender(oscene, cam, 0.7, r0, 1_000_000)
    for p in range(2):
        st = gpu.pass_()
        assert st.flags == 0
        stored = oracle_pass(ref)
        compare_pass(gpu, ref, stored, f"occluded plain pass {p}", record=False)
    gpu.close()


def test_occluded_guided_whole_passes(occluded):
    scene, cam, oscene = occluded
    r0 = 0.02 * scene.diagonal()
    gpu = sppm_amd.Renderer(scene, cam, 0.7, r0, 1_000_000, bvh_width=8)
    gpu.guide_set(True, mix_plain=0.5, learning_rate=1e-3, train_steps=4, seed=3, history_decay=0.9)
    ref = pyoracle.OracleRender(oscene, cam, 0.7, r0, 1_000_000)
    og = pyoracle.OracleGuide(oscene, 0.5)
    ref.set_guide(og)
    for p in range(3):
        og.set_tables(*gpu.guide_tables())
        st = gpu.pass_()
        assert st.flags == 0
        stored = oracle_pass(ref)
        compare_pass(gpu, ref, stored, f"occluded guided pass {p}", record=False)
    gpu.close()


def fixed_radius_image(scene, cam, r0, photons, passes, guided):
    r = sppm_amd.Renderer(scene, cam, 1.0, r0, photons, bvh_width=8)
    if guided:
        r.guide_set(True, mix_plain=0.5, learning_rate=1e-3, train_steps=4, seed=5, history_decay=0.9)
    for _ in range(passes):
        st = r.pass_()
        assert st.flags == 0
    img = r.reconstruct()
    sh = r.read_shared()
    r.close()
    assert np.isfinite(img).all()
    assert np.all(sh["radius"] == np.float32(r0))
    return sh["radiance"].astype(np.float64).reshape(cam["height"], cam["width"], 3)


def test_occluded_guided_matches_plain_at_fixed_radius(occluded):
    scene, cam, _ = occluded
    r0 = 0.02 * scene.diagonal()
    plain = fixed_radius_image(scene, cam, r0, 1_000_000, 24, False)
    guided = fixed_radius_image(scene, cam, r0, 1_000_000, 24, True)
    mp, mg = plain.mean(), guided.mean()
    say(f"occluded 1024^2, fixed radius: mean flux plain {mp:.5g} guided {mg:.5g}")
    assert abs(mg - mp) < 0.03 * mp
    bp = plain.reshape(16, 64, 16, 64, 3).mean((1, 3))
    bg = guided.reshape(16, 64, 16, 64, 3).mean((1, 3))
    lit = bp > 0.05 * bp.max()
    assert np.median(np.abs(bg - bp)[lit] / bp[lit]) < 0.1


@pytest.fixture(scope="module")
def caustic():
    scene = scenes.caustic_scene()
    return scene, scenes.caustic_cameras(n=2), pyoracle.OracleScene(scene)


def test_caustic_plain_spot_parity(caustic):
    scene, cams, oscene = caustic
    assert scene.nfacets > 1_900_000
    cam = cams[0]
    r0 = 0.01 * scene.diagonal()
    N = 4_000_000
    gpu = sppm_amd.Renderer(scene, cam, 0.7, r0, N, bvh_width=8)
    orender = pyoracle.OracleRender(oscene, cam, 0.7, r0, N)
    for step in (1, 2):
        st = gpu.pass_()
        assert (st.flags & ~0x31) == 0
        say(f"caustic camera 0 pass {step}: {st.photons_stored} photons stored")
        spot_check(gpu, orender, step, 3000, 1500, seed=step, sampler_id=cam["sampler_id"])
    gpu.close()


def test_caustic_guided_spot_parity(caustic):
    scene, cams, oscene = caustic
    cam = cams[1]
    r0 = 0.01 * scene.diagonal()
    N = 4_000_000
    gpu = sppm_amd.Renderer(scene, cam, 0.7, r0, N, bvh_width=8)
    gpu.guide_set(True, mix_plain=0.5, learning_rate=1e-3, train_steps=4, seed=9, history_decay=0.9, hidden=128,
                  layers=6, fp16=True)
    orender = pyoracle.OracleRender(oscene, cam, 0.7, r0, N)
    og = pyoracle.OracleGuide(oscene, 0.5)
    orender.set_guide(og)
    for step in (1, 2, 3):
        og.set_tables(*gpu.guide_tables())
        st = gpu.pass_()
        assert (st.flags & ~0x31) == 0
        say(f"caustic camera 1 guided pass {step}: {st.photons_stored} photons stored")
        spot_check(gpu, orender, step, 2000, 1500, seed=10 + step, sampler_id=cam["sampler_id"])
    # the softmax over the cells whose position cell holds table points (the
    # others are masked to 0; sampling normalises by the CDF's total)
    prob, cdf = gpu.guide_tables()
    assert np.isfinite(prob).all() and (prob >= 0).all() and 0.5 < prob.sum() <= 1.0 + 1e-3
    assert cdf[-1] > 0
    assert np.isfinite(gpu.reconstruct()).all()
    gpu.close()
