"""Pins the oracle's closest hit, photon slot semantics, neighbour sets,
progressive update and reconstruction with structure-independent checks:
brute force vs the oracle's own BVH, analytic hits and ties, an independent
brute-force neighbour search, closed-form update cases, and the reference's
stale-slot rule (SPPM.cpp:530-540, 606, 672, 812-822)."""
import numpy as np
import pytest

import pyoracle as po
import scenes


def rays(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(0.02, 0.98, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


@pytest.mark.parametrize("which", ["box", "full"])
def test_bvh_equals_brute_force(which, box_scene, full_scene_small):
    sc = box_scene if which == "box" else full_scene_small
    os_ = po.OracleScene(sc)
    o, d = rays(20000, 11)
    a = os_.closest_hit(o, d, mode=0)
    b = os_.closest_hit(o, d, mode=1)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert (a[0] >= 0).mean() > 0.5


def test_analytic_hit_and_tie_goes_to_lowest_id(box_scene):
    os_ = po.OracleScene(box_scene)
    # straight down onto the floor (facets 0, 1 span y = 0): t = height
    o = np.array([[0.7, 0.5, 0.2], [0.2, 0.25, 0.8]], np.float32)
    d = np.array([[0, -1, 0], [0, -1, 0]], np.float32)
    f, t, u, v = os_.closest_hit(o, d, mode=0)
    assert set(f.tolist()) <= {0, 1}
    np.testing.assert_allclose(t, [0.5, 0.25], rtol=1e-6)
    # the floor's diagonal x == z is shared by facets 0 and 1 at equal t
    o = np.array([[0.15, 0.5, 0.15], [0.9, 0.9, 0.9], [0.5, 0.95, 0.5]], np.float32)
    f, t, _, _ = os_.closest_hit(o, np.array([[0, -1, 0]] * 3, np.float32), mode=0)
    np.testing.assert_array_equal(f, [0, 0, 0])
    # MinT is strict: an origin on the surface does not hit it at t = 0
    f, _, _, _ = os_.closest_hit(np.array([[0.5, 0.0, 0.5]], np.float32), np.array([[0, 1, 0]], np.float32))
    assert f[0] != 0 and f[0] != 1


@pytest.fixture(scope="module")
def small_run(box_scene):
    cam = scenes.pinhole_camera(48, 40)
    r = po.OracleRender(po.OracleScene(box_scene), cam, 0.9, 0.015, 12000)
    return r, cam


def brute_neighbours(pos, c, R):
    """Independent restatement of isIncludePoint + strict radius in float32."""
    c = c.astype(np.float32)
    R = np.float32(R)
    lo, hi = c - R, c + R
    inbox = ((lo <= pos) & (pos <= hi)).all(axis=1)
    dd = (pos - c).astype(np.float32)
    d2 = (dd[:, 0] * dd[:, 0] + dd[:, 1] * dd[:, 1]) + dd[:, 2] * dd[:, 2]
    return np.where(inbox & (d2 < R * R))[0]


def test_neighbour_sets_and_update_closed_form(small_run):
    r, cam = small_run
    r.full_pass(record=True)
    n = po.orc().orc_flatten(r.h)
    ph = r.read_photons(n)
    hp = r.read_hitpoints()
    off, idx = r.read_neighbours()
    m = r.read_counts()
    sh = r.read_shared()
    hits = np.where(hp["type"] == 1)[0]
    rng = np.random.default_rng(1)
    for v in rng.choice(hits, min(400, len(hits)), replace=False):
        want = brute_neighbours(ph["pos"], hp["pos"][v], np.float32(0.015))
        np.testing.assert_array_equal(idx[off[v]:off[v + 1]], want)
        assert m[v] == len(want)
    # first pass from N = 0: N = int(0.9 M) (int truncation: M = 1 keeps N = 0),
    # NumGather = 1, R = r0 * sqrt(0.9 M / M)
    g = m > 0
    np.testing.assert_array_equal(sh["num_photon"][g], (np.float32(0.9) * m[g].astype(np.float32)).astype(np.int32))
    assert (sh["num_photon"][m == 1] == 0).all()
    np.testing.assert_array_equal(sh["num_gather"][g], 1)
    np.testing.assert_array_equal(sh["num_gather"][~g], 0)
    np.testing.assert_allclose(sh["radius"][g], 0.015 * np.sqrt(0.9), rtol=1e-6)
    np.testing.assert_array_equal(sh["radius"][~g], np.float32(0.015))


def test_stale_slot_rule(box_scene):
    """Pass p >= 2: a path that stores k >= 1 photons keeps slot k from an
    earlier pass; it is flattened iff it holds a photon, i.e. iff the path
    stored more than k photons when it last wrote that slot.  With fresh
    slots (pass 1) no stale photon exists; k = 0 yields the emission photon."""
    cam = scenes.pinhole_camera(16, 16)
    N = 4000
    r = po.OracleRender(po.OracleScene(box_scene), cam, 0.9, 0.015, N)
    k_hist = []
    for step in (1, 2, 3):
        po.orc().orc_photon_pass(r.h, step)
        counts = r.read_path_counts()
        ks = np.array([r.photon_path(step, i)["k"] for i in range(N)])
        # slot k's validity before this pass: valid iff some earlier pass wrote a
        # photon there last, i.e. the latest earlier pass touching slot k had k' > k
        stale = np.zeros(N, bool)
        for i in range(N):
            k = ks[i]
            if k == 0:
                continue
            valid = False
            for kp in k_hist:  # oldest to newest
                if kp[i] > k:
                    valid = True
                elif kp[i] + 1 == k or (kp[i] == 0 and k == 1):
                    valid = False  # terminator written at slot kp+1 (slot 1 for an emission-only path)
            stale[i] = valid
        want = np.where(ks == 0, 1, ks + stale.astype(np.int32))
        np.testing.assert_array_equal(counts, want)
        if step == 1:
            assert not stale.any()
        else:
            assert stale.any()
        k_hist.append(ks)


def test_reconstruct_closed_form(small_run):
    r, cam = small_run
    sh = r.read_shared()
    total, passes = 12000 * 3, 3
    img = r.reconstruct(total, passes)
    f = np.float32
    den = (f(total) * (sh["num_gather"].astype(f) + f(1e-4)) / (f(1e-4) + f(passes))) * f(3.14159) * sh["radius"] * sh["radius"]
    L = (f(0.0) / f(1e-4)) + sh["radiance"] / den[:, None]
    want = np.power(L.astype(np.float64), 1 / 1.2)
    np.testing.assert_allclose(img.reshape(-1, 4)[:, :3], want, rtol=2e-6, atol=1e-30)
    assert (img.reshape(-1, 4)[:, 3] == 1.0).all()


# ---- environment light (IBL.h, SPPM.cpp:290-300, 404-418, 485) -----------------
def test_ibl_sphere_mapping_and_conversion():
    """GetUVFromDirection (SPHERE): u = azimuth from +x towards +z over 2 pi,
    v = polar angle from +y over pi (PI = 3.14159f, so values sit within
    1e-6 of the exact ones); straight up maps to (0, 0).  The colour is the
    bilinear fetch at (u, v) of the map stored as 0.4 * c^2."""
    for d, u, v in (((1, 0, 0), 0.0, 0.5), ((0, 0, 1), 0.25, 0.5), ((-1, 0, 0), 0.5, 0.5), ((0, 0, -1), 0.75, 0.5),
                    ((0.6, 0.8, 0), 0.0, np.arccos(0.8) / np.pi), ((0.0, -0.6, -0.8), 0.75, np.arccos(-0.6) / np.pi),
                    ((0, 1, 0), 0.0, 0.0)):
        uv = po.OracleScene.ibl_uv(np.array(d, np.float32))
        assert abs(uv[0] - u) < 2e-6 and abs(uv[1] - v) < 2e-6, (d, uv)
    sc = scenes.box_scene_ibl()
    o = po.OracleScene(sc)
    conv = (np.float32(0.4) * (sc.textures[sc.ibl] * sc.textures[sc.ibl])).astype(np.float32)
    conv[..., 3] = 0.0
    import sppm_amd
    td = sppm_amd.Texture()
    td.height, td.width = conv.shape[0], conv.shape[1]
    td.rgba = po.fp(conv)
    rng = np.random.default_rng(3)
    for _ in range(20):
        d = rng.normal(size=3).astype(np.float32)
        d /= np.linalg.norm(d)
        uv = po.OracleScene.ibl_uv(d)
        want = np.zeros(4, np.float32)
        po.orc().orc_bilinear(po.C.byref(td), float(uv[0]), float(uv[1]), po.fp(want))
        np.testing.assert_array_equal(o.ibl_color(d), want[:3])
    o2 = po.OracleScene(scenes.box_scene())
    assert (o2.ibl_color(np.array([1, 0, 0], np.float32)) == 0).all()


def test_ibl_direct_light_accumulates_and_reconstructs():
    """Pixels whose camera ray leaves the box see the sky: type
    ONLY_DIRECTLIGHT with DirectLight = sky colour / camera pdf (pinhole: 1,
    SPPM.cpp:346-347); the shared sum and count follow SPPM.cpp:415-418 and
    reconstruct adds DirectLight / (NumDirectLight + 1e-4) (SPPM.cpp:485)."""
    sc = scenes.box_scene_ibl()
    osc = po.OracleScene(sc)
    cam = scenes.pinhole_camera(24, 24, sampler=2)  # HammersleySample
    cam["sampler_nsample"] = 1
    cam["pos"] = (0.5, 0.5, 3.0)
    cam["fov_deg"] = 80.0
    o = po.OracleRender(osc, cam, 0.9, 0.015, 2000)
    o.full_pass(record=False)
    hp = o.read_hitpoints()
    dl = o.read_direct()
    only = hp["type"] == 2
    assert only.any() and (hp["type"] == 1).any()
    # an escaping camera ray: DirectLight is the sky in its direction (throughput 1, pdf 1)
    smp = np.zeros(3, np.float32)
    po.orc().orc_hammersley(1, 1, po.fp(smp))  # HammersleySample(NumSample 1) at step 1
    for p in np.nonzero(only)[0][:5]:
        _, d, pdf = po.camera_ray(cam, np.float32(p % 24) / np.float32(24), np.float32(p // 24) / np.float32(24), smp)
        np.testing.assert_array_equal(dl["hit"][p], (osc.ibl_color(d) / np.float32(pdf)).astype(np.float32))
    nz = np.sum(dl["hit"].astype(np.float64) ** 2, 1) > 1e-8
    np.testing.assert_array_equal(dl["num"], (nz & (hp["type"] != 0)).astype(np.int32))
    np.testing.assert_array_equal(dl["shared"], np.where((nz & (hp["type"] != 0))[:, None], dl["hit"], 0))
    img = o.reconstruct(2000, 1).reshape(-1, 4)
    want = np.power(dl["shared"][only] / np.float32(1.0001), np.float32(1 / 1.2))
    np.testing.assert_allclose(img[only][:, :3], want, rtol=1e-5)


def test_occlusion_through_alpha_screens():
    """Scene::IsOccluded (Scene.cpp:246-281) on rays along -z from z = 0.95
    whose answers follow from the screens' layout: an opaque half, a
    translucent half in front of an opaque screen or of nothing, a stack of
    clear layers, MinDist past a screen, and the translucency threshold."""
    os_ = po.OracleScene(scenes.screens_scene())
    d = np.array([[0, 0, -1]], np.float32)

    def occ(x, tmax, tmin=1e-4, tr=0.5):
        return int(os_.occluded(np.array([[x, 0.5, 0.95]], np.float32), d, tmin, tmax, tr)[0])

    # back wall at t = 0.95, A at 0.45, B at 0.65, C at 0.8, stack at 0.05 .. 0.4
    assert occ(0.7, 0.4) == 0 and occ(0.7, 0.5) == 1      # A's opaque half
    assert occ(0.4, 0.9) == 0 and occ(0.4, 1.0) == 1      # A's clear half, then the back wall
    assert occ(0.2, 0.6) == 0 and occ(0.2, 0.7) == 1      # A's clear half, then B
    assert occ(0.7, 0.6, tmin=0.5) == 0                    # MinDist beyond A
    assert occ(0.2, 0.7, tmin=0.3) == 1                    # past A, MinDist restarts at RAY_OFFSET: B
    assert occ(0.4, 0.9, tr=0.85) == 1                     # 1 - 0.2 = 0.8 no longer passes
    assert occ(0.95, 0.9) == 0 and occ(0.95, 1.0) == 1    # 8 clear layers, then the back wall
    assert occ(0.95, 0.9, tr=1.0) == 1                     # 1 - 0 = 1 is not > 1
    assert occ(0.05, 0.9) == 0                             # nothing in front of the back wall


# ---- eye-path next-event estimation (SPPM.cpp:144-168, 187-229, compiled out there) ----
def light_filled_camera(n=16):
    """Pinhole under the box light looking up at it with a narrow field:
    every camera ray hits the light first."""
    cam = scenes.pinhole_camera(n, n)
    cam.update(pos=(0.5, 0.6, 0.5), target=(0.5, 1.0, 0.51), fov_deg=6.0, dist_to_lens=0.05)
    return cam


def test_nee_modes_on_directly_seen_light():
    """Off: the light is a hit-point surface like any other.  Otherwise the
    path ends at the emitter with DirectLight = emitter albedo (1) x radiance
    x T / PDF / camera pdf = 15 exactly (pinhole pdf 1, depth 0, MIS weight
    1).  As written the pixel keeps type NONE, so applyHitPointInfo drops it;
    with KEEP it is ONLY_DIRECTLIGHT and reaches the shared sums."""
    sc = scenes.nee_scene()
    os_ = po.OracleScene(sc)
    cam = light_filled_camera()
    for nee, want_type in ((0, 1), (1, 0), (2, 2)):
        r = po.OracleRender(os_, cam, 0.9, 0.015, 1000, nee=nee)
        r.full_pass()
        hp, d = r.read_hitpoints(), r.read_direct()
        assert (hp["type"] == want_type).all(), (nee, np.bincount(hp["type"]))
        if nee == 0:
            assert (sc.attr[hp["facet"]] == 3).all()
        assert (d["hit"] == np.float32(0.0 if nee == 0 else 15.0)).all()
        assert (d["shared"] == np.float32(15.0 if nee == 2 else 0.0)).all()
        assert (d["num"] == (1 if nee == 2 else 0)).all()


def test_nee_mis_matches_plain_bsdf_sampling():
    """The MIS combination of light samples and emitter hits (KEEP) has the
    expectation of plain BSDF sampling with emitter hits at full weight (the
    oracle-only eye_nee = 3), and a lower variance.  Both renders share every
    camera ray, so the per-pass difference of their DirectLight totals
    isolates the light reaching the eye through the conductor and glass
    spheres (surfaces without hit points).  Dropping cosL from the light
    sample moves the mean by ~4 standard errors here."""
    sc = scenes.nee_scene()
    os_ = po.OracleScene(sc)
    cam = scenes.pinhole_camera(64, 64)
    tot = {}
    for nee in (2, 3):
        r = po.OracleRender(os_, cam, 0.9, 0.015, 500, nee=nee)
        t = []
        for _ in range(300):
            r.full_pass()
            t.append(r.read_direct()["hit"].astype(np.float64).sum(0))
        tot[nee] = np.array(t)
    diff = tot[2] - tot[3]
    se = diff.std(0) / np.sqrt(len(diff))
    assert (np.abs(diff.mean(0)) < 3.0 * se).all(), (diff.mean(0), se)
    assert (tot[2].var(0) < tot[3].var(0)).all()
