"""Benchmark: plain SPPM photons/s on one MI355X.

Workload (full-material SPPM): Cornell box + 200,000-triangle displaced
geodesic sphere, GGX conductor and GGX dielectric spheres, 256x256 RGB8
checker-textured area light, 512x512 albedo/roughness/normal textures,
1024x1024 thin-lens camera (aperture 0.02, focus 2.0), 1,000,000 photons per
pass, alpha 0.7, r0 = 0.02 x scene diagonal, BVH8.  Synthetic seeded scene.

A step is one SPPM pass (photon pass, gather-structure build, eye pass,
gather + progressive update) over the whole image; the timed region is K
passes bracketed by device synchronisation.  value = photons emitted / s.
The defaults (1 warm-up + 255 timed passes) time the configured 256-pass
render; the radius shrinks from pass to pass, so early passes alone would
not represent it.

Prints one JSON line.  Extra fields: per-phase kernel times, the roofline of
the dominant phase (algorithmic bytes per launch / measured launch time,
hipEvents on the launch stream), and a CPU baseline (the oracle restatement,
OpenMP, on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "friendsrender-r-public_amd"))

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes (lower bounds), per unit of work
STORE_BYTES_PER_PHOTON = 44           # pos 12 + flux 12 + in-dir 12 + pdf 4 + IOR 4
BUILD_BYTES_PER_PHOTON = 2 * 44 + 8   # one read, one reordered write, key/index
GATHER_BYTES_PER_HIT = 68 + 2 * 24    # hit point record + shared state read/write
GATHER_BYTES_PER_NEIGHBOUR = 44       # each in-radius photon record


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=255)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["full", "box"], default="full")
    ap.add_argument("--bvh", type=str, default=None,
                    help="closest-hit structure of the main workload: BVH width 2, 4 or 8, with an 'l' suffix "
                         "for the LBVH builder (e.g. 4l); default 8 (binned SAH)")
    ap.add_argument("--photon-kernel", choices=["megakernel", "wavefront"], default=None,
                    help="photon pass kernels (default: the library's default)")
    ap.add_argument("--wave-bounces", type=int, default=0, help="wavefront: bounces from ray queues (0 = default)")
    ap.add_argument("--wave-lanes", type=int, default=0, help="wavefront: lanes per ray (0 = default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--rmse", action="store_true",
                    help="time-to-RMSE mode on the occluded-light workload: plain vs guided emission against a "
                         "long plain reference render")
    ap.add_argument("--rmse-scene", choices=["occluded", "caustic"], default="occluded",
                    help="time-to-RMSE workload: the occluded-light scene (1024^2, 1M photons, 64x4 bf16 network) "
                         "or camera 0 of the caustic workload (2048^2, 4M photons, 128x6 fp16 network)")
    ap.add_argument("--ref-passes", type=int, default=10000)
    ap.add_argument("--ref-save", default=None, metavar="NPY",
                    help="time-to-RMSE mode: write the reference's linear image and stop")
    ap.add_argument("--ref-load", default=None, metavar="NPY",
                    help="time-to-RMSE mode: read the reference's linear image instead of rendering it")
    ap.add_argument("--no-rmse", action="store_true",
                    help="leave out the time-to-RMSE leg (guided vs plain on the occluded-light scene)")
    ap.add_argument("--mix", type=float, default=0.5, help="guided emission: probability of the plain emission")
    ap.add_argument("--lr", type=float, default=1e-3, help="guided emission: Adam step size")
    ap.add_argument("--train-steps", type=int, default=2, help="guided emission: Adam steps per training pass")
    ap.add_argument("--train-interval", type=int, default=4, help="guided emission: train after every k-th pass")
    ap.add_argument("--decay", type=float, default=0.95, help="guided emission: training histogram decay")
    ap.add_argument("--guide-target", choices=["mean", "rms"], default="rms",
                    help="guided emission: the proposal's training target (mean or root-mean-square deposit)")
    ap.add_argument("--max-passes", type=int, default=2000)
    ap.add_argument("--no-graph-leg", action="store_true",
                    help="leave out the graph-replayed timing of the main workload")
    ap.add_argument("--no-bound", action="store_true",
                    help="time-to-RMSE mode: leave out the emission-proposal bound")
    ap.add_argument("--check-every", type=int, default=16)
    ap.add_argument("--target-rmse", type=float, default=0.01)
    ap.add_argument("--resolution", type=int, default=1024)
    ap.add_argument("--alpha", type=float, default=1.0,
                    help="time-to-RMSE mode: alpha = 1 keeps the radius at r0, so guided and plain estimate the same "
                         "r0-smoothed image and the RMSE measures variance; with alpha < 1 the radius schedule "
                         "depends on the photons gathered and a finite plain reference is not bias-converged")
    ap.add_argument("--batch", action="store_true",
                    help="batched mode: many uneven seeded scenes, one graph launch per pass over all of them")
    ap.add_argument("--batch-mode", choices=["auto", "multi", "branches"], default="auto",
                    help="batch mode: one launch per phase for all scenes (multi), a graph branch per scene "
                         "(branches), or by the mean photons per pass (auto)")
    ap.add_argument("--scenes", type=int, default=64)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--caustic", action="store_true",
                    help="caustic workload: ~2M triangles of GGX glass on a 2048^2-textured floor, small textured "
                         "light, CAMERAS thin-lens cameras at RES^2 batched per graph launch; plain and guided "
                         "(128 x 6 fp16 network) photons/s")
    ap.add_argument("--cameras", type=int, default=8)
    ap.add_argument("--caustic-res", type=int, default=2048)
    ap.add_argument("--caustic-photons", type=int, default=4_000_000)
    ap.add_argument("--guide-train", action="store_true",
                    help="training rate of the learned proposal's networks (64x4 bf16, 128x6 fp16): Adam steps "
                         "over all 262,144 cells on a fixed histogram")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="before the workload, run a device elementwise copy of MIB MiB (float32, torch) three "
                         "times: a kernel of known traffic for scaling the HBM byte counters of a profiler run")
    return ap.parse_args()


def photon_kernel_args(args):
    """Renderer keyword arguments choosing the photon pass kernels."""
    import sppm_amd
    kw = {}
    if args.photon_kernel is not None:
        kw["photon_kernel"] = (sppm_amd.PHOTON_WAVEFRONT if args.photon_kernel == "wavefront"
                               else sppm_amd.PHOTON_MEGAKERNEL)
    if args.wave_bounces:
        kw["wave_bounces"] = args.wave_bounces
    if args.wave_lanes:
        kw["wave_lanes"] = args.wave_lanes
    return kw


def read_traffic(path):
    """HBM bytes per launch of a phase from its committed counter summary
    (line 'bytes_per_launch: N'), or None."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.strip().startswith("bytes_per_launch:"):
            return int(float(line.split(":", 1)[1]))
    return None


def read_valu(path):
    """Per-launch VALU counts of the pass kernels from their committed counter
    summary (lines 'lane_ops_per_launch.<kernel>: N' and
    'issued_slots_per_launch.<kernel>: N'): {kernel: (lane_ops, slots)}."""
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        t = line.strip()
        for key in ("lane_ops_per_launch.", "issued_slots_per_launch."):
            if t.startswith(key):
                k, v = t[len(key):].split(":", 1)
                out.setdefault(k, [0.0, 0.0])[0 if key.startswith("lane") else 1] = float(v)
    return out


VALU_PEAK_T = 39.3  # MI355X vector ALUs: 256 CUs x 64 lanes x 2.4 GHz, T lane operations / s


BATCH_MODES = {"multi": 0, "branches": 1, "auto": 2}  # sppm_batch_mode
BATCH_MODE_NAMES = {0: "multi", 1: "branches"}


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def workload(name):
    import scenes
    if name == "full":
        scene = scenes.full_scene(freq=100, tex_size=512)
        cam = scenes.thinlens_camera(1024, 1024, aperture=0.02, focus=2.0)
        return scene, cam, dict(name="full-material SPPM", photons=1_000_000, width=1024, height=1024, bvh=8,
                                alpha=0.7, r0=0.02 * scene.diagonal(), triangles=scene.nfacets, passes=256)
    scene = scenes.box_scene()
    cam = scenes.pinhole_camera(256, 256)
    return scene, cam, dict(name="box SPPM", photons=100_000, width=256, height=256, bvh=4, alpha=0.7,
                            r0=0.02 * scene.diagonal(), triangles=scene.nfacets, passes=16)


def host_cores():
    """The host cores this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS when the machine sets it (the GPU box gives each GPU a
    share of its cores and says so there)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_baseline(scene, cam, cfg, seconds):
    """Oracle restatement (OpenMP, on every host core this process may use) on the same
    workload: one untimed pass (allocates the photon slots), then whole
    passes of the full-size workload until the time budget is used (at
    least one)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    cores = host_cores()
    os_ = pyoracle.OracleScene(scene)
    r = pyoracle.OracleRender(os_, cam, cfg["alpha"], cfg["r0"], cfg["photons"], nthreads=cores)
    r.full_pass()
    t0 = time.perf_counter()
    passes = 0
    while True:
        r.full_pass()
        passes += 1
        el = time.perf_counter() - t0
        if el > seconds or passes >= 4:
            break
    return dict(value=round(cfg["photons"] * passes / el, 1), unit="photons/s", cores=cores, kind="port",
                sample=f"passes 2..{passes + 1} of the same workload at full size ({cam['width']}x{cam['height']} px, "
                       f"{cfg['photons']} photons/pass), oracle/sppm_oracle.c with OpenMP; {el:.1f} s timed")


def copy_calibration(mib):
    """Reads and writes exactly mib MiB per launch: out = in * 1 (elementwise)."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(f"copy calibration: 3 launches of {mib} MiB read + {mib} MiB write", file=sys.stderr, flush=True)
    del a, b


def rmse_workload(args):
    import scenes
    if args.rmse_scene == "caustic":
        scene = scenes.caustic_scene()
        res = args.caustic_res
        cam = scenes.caustic_cameras(1, res, res)[0]
        w = dict(scene=scene, cam=cam, res=res, N=args.caustic_photons, alpha=args.alpha, r0=0.01 * scene.diagonal(),
                 net=dict(hidden=128, layers=6, fp16=True), name="caustic, camera 0",
                 dtype="f32 (MLP fp16 MFMA, f32 accumulation)")
    else:
        scene = scenes.occluded_scene()
        res = args.resolution
        cam = scenes.occluded_camera(res, res)
        w = dict(scene=scene, cam=cam, res=res, N=1_000_000, alpha=args.alpha, r0=0.02 * scene.diagonal(), net={},
                 name="occluded light", dtype="f32 (MLP bf16 MFMA, f32 accumulation)")
    return w


def guided_settings(args):
    import sppm_amd
    return dict(mix_plain=args.mix, learning_rate=args.lr, train_steps=args.train_steps, seed=1,
                history_decay=args.decay, train_interval=args.train_interval,
                target=sppm_amd.GUIDE_TARGET_RMS if args.guide_target == "rms" else sppm_amd.GUIDE_TARGET_MEAN)


def linear_image(r, N, passes, res):
    """The linear image: reconstruct output ^ 1.2 (before the reference's
    1/1.2 gamma, SPPM.cpp:485-488)."""
    import numpy as np
    img = r.reconstruct(N * passes, passes).reshape(res, res, 4)[..., :3].astype(np.float64)
    return np.power(np.maximum(img, 0.0), 1.2)


def reference_image(args, w):
    """A long plain render with random streams independent of the timed
    renders (pass steps offset by 10^6, another camera sample stream)."""
    import numpy as np
    import sppm_amd
    if args.ref_load:
        log(f"reference read from {args.ref_load}")
        return np.load(args.ref_load).astype(np.float64)
    t0 = time.perf_counter()
    r = sppm_amd.Renderer(w["scene"], dict(w["cam"], sampler_id=7919), w["alpha"], w["r0"], w["N"], bvh_width=8,
                          step_offset=1_000_000)
    for p in range(args.ref_passes):
        r.pass_(stats=False)
        if (p + 1) % 1000 == 0:
            r.sync()
            log(f"reference: {p + 1}/{args.ref_passes} passes, {time.perf_counter() - t0:.1f} s")
    r.sync()
    ref = linear_image(r, w["N"], args.ref_passes, w["res"])
    r.close()
    log(f"reference: {args.ref_passes} passes in {time.perf_counter() - t0:.1f} s")
    if args.ref_save:
        np.save(args.ref_save, ref.astype(np.float32))
    return ref


def _fill_proposal(p, s, wmix):
    """q_x = max(w p_x, t s_x), t by bisection so that q sums to 1."""
    import numpy as np
    lo, hi = 0.0, 1.0
    while np.maximum(wmix * p, hi * s).sum() < 1.0:
        hi *= 2.0
    for _ in range(80):
        t = 0.5 * (lo + hi)
        if np.maximum(wmix * p, t * s).sum() > 1.0:
            hi = t
        else:
            lo = t
    q = np.maximum(wmix * p, hi * s)
    return q / q.sum()


def _pool_cells(x, blocks=(4, 4, 2, 2)):
    """Sums over blocks of (u, v, s, t) emission cells, spread back over the cells."""
    import numpy as np
    a, b, c, d = blocks
    g = x.reshape(32 // a, a, 32 // b, b, 16 // c, c, 16 // d, d).sum((1, 3, 5, 7))
    for ax, k in enumerate(blocks):
        g = np.repeat(g, k, ax)
    return g.ravel()


def emission_gain_bound(w, plain_ms, plain_gather_ms, mixes=(0.5, 0.25), passes=8):
    """What any emission proposal over the 262,144 cells, mixed with the
    plain emission at weight w, can gain on this workload: per-pass variance
    and time.

    A photon path's contribution c (the sum, over its stored photons in the
    camera's room, x < 1 in the occluded-light scene, of luminance(flux) /
    PDF) depends on its emission cell x only through the cell's plain
    probability p_x and its second moment m_x = E[c^2 | x].  Emitting with
    q = w p + (1 - w) g and weighting by p / q, the per-path second moment is
    sum_x p_x^2 m_x / q_x; it is smallest for q_x = max(w p_x, t p_x
    sqrt(m_x)).  The cost of a path is a + b r_x, r_x its mean number of
    photons in the camera's room (the gather's work follows them): a and b
    split the measured plain pass time (plain_ms) into its gather part
    (plain_gather_ms, per room photon) and the rest (per path).  Time per
    unit variance is smallest for q_x = max(w p_x, t p_x sqrt(m_x / (a + b
    r_x + k))) (k by search).

    p, m and r are measured on plain-emission passes (guided emission with
    mix_plain = 1 draws plain photons and records each path's cell).  Fitted
    and scored on the same passes the optimum overfits the noise of m (the
    in-sample figure); the cross-validated figures fit on the even passes,
    score on the odd ones and the other way round, with m and r pooled over
    blocks of 4 x 4 position x 2 x 2 direction cells (the pooling that scores
    best).  The mean share of paths whose first segment crosses the door
    (their first stored photon lies in the room) is reported beside it."""
    import numpy as np
    import sppm_amd
    r = sppm_amd.Renderer(w["scene"], w["cam"], 1.0, w["r0"], w["N"], bvh_width=8)
    r.guide_set(True, mix_plain=1.0, learning_rate=1e-3, train_steps=1, seed=1, history_decay=0.95, **w["net"])
    ncell = sppm_amd.GUIDE_CELLS
    half = np.zeros((2, 4, ncell))  # count, sum c, sum c^2, sum of room photons
    door = np.zeros(2)
    room_photons = 0
    lum = np.array([0.2126, 0.7152, 0.0722])
    for k in range(passes):
        r.pass_(stats=False)
        counts = r.read_path_counts()
        cells = r.guide_path_cells()
        ph = r.read_photons()
        path = np.repeat(np.arange(len(counts)), counts)
        first = np.concatenate([[0], np.cumsum(counts)[:-1]])
        inroom = ph["pos"][:, 0] < 1.0
        val = np.where(inroom, (ph["flux"].astype(np.float64) @ lum) / ph["pdf"].astype(np.float64), 0.0)
        c = np.bincount(path, weights=val, minlength=len(counts))
        nr = np.bincount(path, weights=inroom.astype(np.float64), minlength=len(counts))
        h = half[k % 2]
        h[0] += np.bincount(cells, minlength=ncell)
        h[1] += np.bincount(cells, weights=c, minlength=ncell)
        h[2] += np.bincount(cells, weights=c * c, minlength=ncell)
        h[3] += np.bincount(cells, weights=nr, minlength=ncell)
        through = ph["pos"][first, 0] < 1.0
        door += [c[through].sum(), c.sum()]
        room_photons += int(inroom.sum())
    r.close()
    tot = half.sum(0)
    n = tot[0].sum()
    p = tot[0] / n
    ok = p > 0
    a_cost = (plain_ms - plain_gather_ms) / w["N"]
    b_cost = plain_gather_ms / (room_photons / passes)

    def cell_mean(h, k, blocks=None):
        num, den = (h[k], h[0]) if blocks is None else (_pool_cells(h[k], blocks), _pool_cells(h[0], blocks))
        return np.where(den > 0, num / np.maximum(den, 1), 0.0)

    def score(q, h):
        """Per-path variance and cost of proposal q on the paths of h."""
        m, rr = cell_mean(h, 2), cell_mean(h, 3)
        mu = h[1].sum() / h[0].sum()
        return (p[ok] ** 2 * m[ok] / q[ok]).sum() - mu * mu, (q * (a_cost + b_cost * rr)).sum()

    out = dict(paths=int(n), cells_hit=int(ok.sum()), door_mean_share=round(float(door[0] / door[1]), 4),
               passes=passes, room_photons_per_pass=round(room_photons / passes),
               cost_model_ms=dict(per_million_paths=round(a_cost * 1e6, 4),
                                  per_thousand_room_photons=round(b_cost * 1e3, 5)))
    for wmix in mixes:
        m_all = cell_mean(tot, 2)
        v_plain, c_plain = score(p, tot)
        q = _fill_proposal(p, np.sqrt(p * p * m_all), wmix)
        out[f"in_sample_variance_gain_mix_{wmix}"] = round(float(v_plain / score(q, tot)[0]), 3)
        acc = []
        for fit, ev in ((0, 1), (1, 0)):
            m, rr = cell_mean(half[fit], 2, (4, 4, 2, 2)), cell_mean(half[fit], 3, (4, 4, 2, 2))
            vp, cp = score(p, half[ev])
            qv = _fill_proposal(p, np.sqrt(p * p * m), wmix)
            vv, cv = score(qv, half[ev])
            best = None
            for kap in np.concatenate([[0.0], np.geomspace(1e-4, 1e2, 25) * a_cost]):
                qc = _fill_proposal(p, np.sqrt(p * p * m / (a_cost + b_cost * rr + kap)), wmix)
                vc, cc = score(qc, half[ev])
                if best is None or vc * cc < best[0] * best[1]:
                    best = (vc, cc)
            acc.append((vp / vv, vp * cp / (vv * cv), vp / best[0], vp * cp / (best[0] * best[1])))
        g = np.mean(acc, 0)
        out[f"cv_variance_gain_mix_{wmix}"] = round(float(g[0]), 3)
        out[f"cv_time_gain_of_variance_optimum_mix_{wmix}"] = round(float(g[1]), 3)
        out[f"cv_cost_aware_variance_gain_mix_{wmix}"] = round(float(g[2]), 3)
        out[f"cv_cost_aware_time_gain_mix_{wmix}"] = round(float(g[3]), 3)
    out["note"] = ("gains of the best emission proposal over the cells against plain emission, per pass (variance) "
                   "and per unit of variance (time, from the cost model; guided sampling and training overhead not "
                   "counted): bounds for the learned proposal.  in_sample: fitted and scored on the same paths "
                   "(overfits); cv: fitted and scored on disjoint passes")
    return out


def time_to_rmse(args, w, ref):
    """Seconds of rendering until the linear image, normalised to the
    reference's mean, is within RMSE target of the reference, for plain and
    for guided emission.  The timed region is the passes only (synchronised
    chunks of check_every passes, guided training included); the RMSE checks
    between chunks are not timed."""
    import numpy as np
    import sppm_amd
    N, res = w["N"], w["res"]
    # pixels whose value is not finite (a NaN pixel weight of the reference's
    # own arithmetic, propagated as it does) are left out of the RMSE, in the
    # reference and in the renders alike; their number is reported
    ref_ok = np.isfinite(ref).all(axis=-1)
    scale = float(ref[ref_ok].mean())
    out = {}
    for mode in ("plain", "guided"):
        r = sppm_amd.Renderer(w["scene"], w["cam"], w["alpha"], w["r0"], N, bvh_width=8)
        if mode == "guided":
            r.guide_set(True, **guided_settings(args), **w["net"])
        el, passes, e = 0.0, 0, float("inf")
        curve = []
        r.kernel_ms("reset")
        while passes < args.max_passes:
            ts = time.perf_counter()
            for _ in range(args.check_every):
                r.pass_(stats=False)
            r.sync()
            el += time.perf_counter() - ts
            passes += args.check_every
            img = linear_image(r, N, passes, res)
            ok = ref_ok & np.isfinite(img).all(axis=-1)
            e = float(np.sqrt(np.mean((img[ok] - ref[ok]) ** 2))) / scale
            curve.append((passes, round(el, 4), round(e, 5)))
            if e <= args.target_rmse:
                break
        out[mode] = dict(seconds=round(el, 3), passes=passes, rmse=round(e, 5), reached=e <= args.target_rmse,
                         nonfinite_pixels=int((~ok).sum()), ms_per_pass=round(el / passes * 1e3, 3),
                         phases_ms={k: round(r.kernel_ms(k), 3) for k in ("photons", "build", "eye", "gather",
                                                                            "guide")},
                         curve=curve[:: max(1, len(curve) // 12)])
        r.close()
        log(f"{mode}: {out[mode]['seconds']} s, {passes} passes, RMSE {e:.4g}")
    both = out["plain"]["reached"] and out["guided"]["reached"]
    return dict(
        metric=f"seconds to RMSE<={args.target_rmse} against a {args.ref_passes}-pass plain render (image "
               f"normalised to its mean), guided vs plain",
        workload=w["name"], resolution=f"{res}x{res}", photons_per_pass=N, triangles=w["scene"].nfacets,
        alpha=w["alpha"], r0=round(w["r0"], 6), reference_passes=args.ref_passes,
        reference_mean_linear_radiance=scale,
        guided_settings=dict(guided_settings(args),
                             network=f"{w['net'].get('hidden', 64)}x{w['net'].get('layers', 4)} "
                                     f"{'fp16' if w['net'].get('fp16') else 'bf16'}"),
        plain_s=out["plain"]["seconds"], guided_s=out["guided"]["seconds"],
        speedup_guided_vs_plain=round(out["plain"]["seconds"] / out["guided"]["seconds"], 3) if both else None,
        plain=out["plain"], guided=out["guided"])


def rmse_mode(args):
    """The time-to-RMSE measurement alone (one JSON line)."""
    w = rmse_workload(args)
    log(f"{w['name']}: {w['scene'].nfacets} triangles, {w['res']}x{w['res']}, {w['N']} photons/pass")
    ref = reference_image(args, w)
    if args.ref_save:
        return
    res = time_to_rmse(args, w, ref)
    if args.rmse_scene == "occluded" and not args.no_bound:
        res["emission_bound"] = emission_gain_bound(w, res["plain"]["ms_per_pass"],
                                                    res["plain"]["phases_ms"]["gather"])
    line = {"metric": res["metric"], "value": res["guided_s"], "unit": "s", "higher_is_better": False,
            "n_gpus": 1, "dtype": w["dtype"], "data": "synthetic (seeded procedural scene)",
            "config": {"workload": w["name"]}, "time_to_rmse": res}
    print(json.dumps(line), flush=True)


def batch_mode(args):
    """Aggregate photons/s of K passes over all scenes of the batched workload,
    each pass one graph launch whose branches are the scenes' passes; beside
    it the same passes run scene after scene without graphs."""
    import scenes
    import sppm_amd
    t0 = time.perf_counter()
    work = scenes.batch_workload(n=args.scenes, progress=lambda s: s % 8 == 0 and log(f"generating scene {s}"))
    tris = [w["scene"].nfacets for w in work]
    log(f"{len(work)} scenes generated in {time.perf_counter() - t0:.1f} s: {sum(tris)} triangles, "
        f"{sum(w['photons'] for w in work)} photons per pass")
    rs, tuned = [], []
    for k, w in enumerate(work):
        r = sppm_amd.Renderer(w["scene"], w["cam"], w["alpha"], w["r0"], w["photons"], bvh_width=8)
        if not args.no_autotune:
            res = r.autotune(w["scene"], passes=2)
            tuned.append(res)
        rs.append(r)
        if (k + 1) % 8 == 0:
            log(f"set up {k + 1}/{len(work)} scenes")
    K, W = args.steps, args.warmup
    # sequential eager passes (reference point): scene after scene
    ts = time.perf_counter()
    for r in rs:
        for _ in range(W + K):
            r.pass_(stats=False)
    for r in rs:
        r.sync()
    seq = time.perf_counter() - ts
    log(f"sequential eager passes: {seq:.3f} s")
    # re-begin every scene (autotuned configuration kept) and capture the batch graph
    for r in rs:
        sppm_amd._check(r.lib.sppm_begin(r.h))
        r.passes = r.total_photons = 0
    batch = sppm_amd.Batch(rs, mode=BATCH_MODES[args.batch_mode])
    used_mode = BATCH_MODE_NAMES[batch.mode]
    for _ in range(W):
        batch.pass_()
    batch.sync()
    tb = time.perf_counter()
    for _ in range(K):
        batch.pass_()
    batch.sync()
    el = time.perf_counter() - tb
    photons = sum(w["photons"] for w in work) * K
    def tally(key):
        out = {}
        for t in tuned:
            out[str(key(t))] = out.get(str(key(t)), 0) + 1
        return out
    choices = dict(bvh=tally(lambda t: ("lbvh" if t["bvh_builder"] else "sah") + str(t["bvh_width"])),
                   grid=tally(lambda t: t["grid_blocks"]),
                   photon_kernel=tally(lambda t: "wavefront" if t["photon_kernel"] else "megakernel"),
                   photons_suggested_vs_set=tally(lambda t: "half" if t["photons_per_s"][0] == max(t["photons_per_s"])
                                                  else ("same" if t["photons_per_s"][1] == max(t["photons_per_s"])
                                                        else "twice")))
    line = {
        "metric": "photons/s (batched plain SPPM, aggregate over scenes, one graph launch per pass)",
        "batch_mode": f"{args.batch_mode} ({used_mode})" if args.batch_mode == "auto" else used_mode,
        "value": round(photons / el, 1), "unit": "photons/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(el / K * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded procedural scenes)",
        "config": {"workload": f"batched SPPM, {len(work)} scenes", "triangles_total": sum(tris),
                   "triangles_min": min(tris), "triangles_max": max(tris),
                   "resolutions": sorted({w["cam"]["width"] for w in work}),
                   "photons_per_pass_total": sum(w["photons"] for w in work), "alpha": 0.7,
                   "autotune": not args.no_autotune, "tuned": choices if tuned else None},
        "sequential_eager_photons_per_s": round(photons * (W + K) / K / seq, 1),
        "batch_speedup_vs_sequential": round(seq * K / (W + K) / el, 3),
    }
    batch.close()
    for r in rs:
        r.close()
    print(json.dumps(line), flush=True)


def caustic_mode(args):
    """Aggregate photons/s of K passes over the caustic workload's camera
    variants, one graph launch per pass (each camera a context and a graph
    branch), with plain emission and with the guided proposal (128 x 6 fp16
    network, mixture 0.5, one Adam step per pass)."""
    import numpy as np
    import scenes
    import sppm_amd
    t0 = time.perf_counter()
    scene = scenes.caustic_scene()
    cams = scenes.caustic_cameras(args.cameras, args.caustic_res, args.caustic_res)
    N = args.caustic_photons
    alpha, r0 = 0.7, 0.01 * scene.diagonal()
    log(f"caustic scene: {scene.nfacets} triangles, generated in {time.perf_counter() - t0:.1f} s")
    K, W = args.steps, args.warmup
    out = {}
    for mode in ("plain", "guided"):
        rs = []
        for k, cam in enumerate(cams):
            r = sppm_amd.Renderer(scene, cam, alpha, r0, N, bvh_width=8)
            if mode == "guided":
                r.guide_set(True, mix_plain=0.5, learning_rate=1e-3, train_steps=1, seed=11 + k, history_decay=0.95,
                            hidden=128, layers=6, fp16=True)
            rs.append(r)
        log(f"{mode}: {len(rs)} contexts set up")
        batch = sppm_amd.Batch(rs, mode=BATCH_MODES[args.batch_mode])
        used_mode = BATCH_MODE_NAMES[batch.mode]
        for _ in range(W):
            batch.pass_()
        batch.sync()
        tb = time.perf_counter()
        for k in range(K):
            batch.pass_()
            if (k + 1) % 5 == 0:
                batch.sync()
                log(f"{mode}: {k + 1}/{K} passes")
        batch.sync()
        el = time.perf_counter() - tb
        c = rs[0].counters()
        img = rs[0].reconstruct()
        out[mode] = dict(photons_per_s=round(len(rs) * N * K / el, 1), ms_per_pass_all_cameras=round(el / K * 1e3, 3),
                         camera0_hit_points_per_pass=c["hit_points"] // max(1, c["passes"]),
                         camera0_neighbours_per_pass=c["neighbours"] // max(1, c["passes"]),
                         camera0_image_finite=bool(np.isfinite(img).all()))
        log(f"{mode}: {out[mode]['photons_per_s']:.4g} photons/s, {out[mode]['ms_per_pass_all_cameras']} ms per pass")
        batch.close()
        for r in rs:
            r.close()
    line = {
        "metric": "photons/s (caustic workload, aggregate over cameras, one graph launch per pass)",
        "batch_mode": f"{args.batch_mode} ({used_mode})" if args.batch_mode == "auto" else used_mode,
        "value": out["plain"]["photons_per_s"], "unit": "photons/s", "n_gpus": 1, "steps": K, "warmup": W,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (MLP fp16 MFMA)",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "caustic", "triangles": scene.nfacets, "cameras": len(cams),
                   "resolution": f"{args.caustic_res}x{args.caustic_res}", "photons_per_pass_per_camera": N,
                   "alpha": alpha, "r0": round(r0, 6), "bvh_width": 8, "network": "128x6 fp16",
                   "mix_plain": 0.5},
        "plain": out["plain"], "guided": out["guided"],
    }
    print(json.dumps(line), flush=True)


def guide_train_mode(args):
    """Wall time per Adam step of the proposal network (training forward with
    the activations kept, per-layer backward, gradient reduction, Adam) for
    the default and the wide network; FLOPs counted as 6 x parameters x cells."""
    import numpy as np
    import scenes
    import sppm_amd
    box = scenes.box_scene()
    rng = np.random.default_rng(1)
    hist = (rng.gamma(0.3, 1.0, sppm_amd.GUIDE_CELLS) * (rng.random(sppm_amd.GUIDE_CELLS) < 0.05)).astype(np.float32)
    steps = max(1, args.steps if args.steps != 255 else 20)
    nets = {}
    for hid, lay, f16 in ((64, 4, False), (128, 6, True)):
        r = sppm_amd.Renderer(box, scenes.pinhole_camera(32, 32), 0.9, 0.015, 2000, bvh_width=4)
        r.guide_set(True, mix_plain=0.5, learning_rate=1e-3, train_steps=0, seed=7, hidden=hid, layers=lay, fp16=f16)
        r.guide_train(hist, 2)
        r.sync()
        t0 = time.perf_counter()
        r.guide_train(hist, steps)
        r.sync()
        ms = (time.perf_counter() - t0) / steps * 1e3
        n = sppm_amd.guide_nparams(hid, lay)
        tf = 6.0 * n * sppm_amd.GUIDE_CELLS / (ms * 1e-3) / 1e12
        nets[f"{hid}x{lay}_{'fp16' if f16 else 'bf16'}"] = dict(params=n, ms_per_step=round(ms, 4),
                                                                tflops=round(tf, 2), frac_of_2500=round(tf / 2500, 4))
        log(f"network {hid}x{lay} {'fp16' if f16 else 'bf16'}: {ms:.3f} ms per Adam step, {tf:.1f} TFLOP/s")
        r.close()
    print(json.dumps({"metric": "proposal training: ms per Adam step over 262,144 cells", "unit": "ms",
                      "higher_is_better": False, "n_gpus": 1, "steps": steps, "data": "synthetic histogram",
                      "networks": nets}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(no multi-GPU path is built: photon and eye paths would split by index, the stored photons "
              f"would be all-gathered)", file=sys.stderr)
        sys.exit(2)
    if args.guide_train:
        guide_train_mode(args)
        return
    if args.caustic:
        caustic_mode(args)
        return
    if args.rmse:
        rmse_mode(args)
        return
    if args.batch:
        batch_mode(args)
        return
    import numpy as np  # noqa: F401
    import sppm_amd

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    scene, cam, cfg = workload(args.workload)
    if args.bvh:
        cfg["bvh"] = int(args.bvh.rstrip("l")) | (sppm_amd.BVH_LBVH if args.bvh.endswith("l") else 0)
    pk = photon_kernel_args(args)
    log(f"scene ready: {cfg['name']}, {cfg['triangles']} triangles")
    r = sppm_amd.Renderer(scene, cam, cfg["alpha"], cfg["r0"], cfg["photons"], bvh_width=cfg["bvh"], **pk)
    for _ in range(args.warmup):
        r.pass_(stats=False)
    r.sync()
    log(f"warm-up done ({args.warmup} passes); timing {args.steps} passes")
    c0 = r.counters()
    r.kernel_ms("reset")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.pass_(stats=False)
    r.sync()
    el = time.perf_counter() - t0
    log(f"timed passes done: {el:.3f} s")
    c1 = r.counters()
    ms = {k: r.kernel_ms(k) for k in ("photons", "build", "eye", "gather")}
    K = args.steps
    hits = (c1["hit_points"] - c0["hit_points"]) / K
    nbs = (c1["neighbours"] - c0["neighbours"]) / K
    stored = (c1["stored"] - c0["stored"]) / K
    cands = (c1["candidates"] - c0["candidates"]) / K
    prays = (c1["photon_rays"] - c0["photon_rays"]) / K
    erays = (c1["eye_rays"] - c0["eye_rays"]) / K
    algo = {
        "photons": stored * STORE_BYTES_PER_PHOTON,
        "build": stored * BUILD_BYTES_PER_PHOTON,
        "gather": hits * GATHER_BYTES_PER_HIT + nbs * GATHER_BYTES_PER_NEIGHBOUR,
    }
    r.close()
    # In a whole pass the eye pass runs beside the photon pass and the build,
    # and the next pass's photon pass beside the gather, so phases_ms are the
    # spans of phases that share the chip.  A second leg, after the timed
    # region, renders the same passes again with the split steps (one stream,
    # one step after another, HIP events around each) for the phases' own
    # times.
    r = sppm_amd.Renderer(scene, cam, cfg["alpha"], cfg["r0"], cfg["photons"], bvh_width=cfg["bvh"], **pk)
    for p in range(args.warmup + K):
        if p == args.warmup:
            r.sync()
            r.kernel_ms("reset")
            a0 = r.counters()
        r.photons(p + 1)
        r.build()
        r.eye(p + 1)
        r.gather()
    r.sync()
    a1 = r.counters()
    ms_alone = {k: r.kernel_ms(k) for k in ("photons", "build", "eye", "gather")}
    r.close()
    # A third leg: the same render with its passes captured as graphs
    # (sppm_graph_capture) and replayed, with the same warm-up and timed
    # passes; the graph-replayed passes overlap as the eager ones do.
    graph_leg = None
    if not args.no_graph_leg:
        r = sppm_amd.Renderer(scene, cam, cfg["alpha"], cfg["r0"], cfg["photons"], bvh_width=cfg["bvh"], **pk)
        r.graph_capture()
        for _ in range(args.warmup):
            r.pass_(stats=False)
        r.sync()
        tg = time.perf_counter()
        for _ in range(K):
            r.pass_(stats=False)
        r.sync()
        elg = time.perf_counter() - tg
        g1 = r.counters()
        r.close()
        graph_leg = {"value": round(cfg["photons"] * K / elg, 1), "unit": "photons/s",
                     "ms_per_step": round(elg / K * 1e3, 3),
                     "same_counters_as_eager": all(g1[k] == c1[k] for k in ("hit_points", "neighbours", "stored",
                                                                              "photon_rays", "eye_rays")),
                     "note": "the same passes captured with sppm_graph_capture and replayed (photon pass beside "
                             "the previous gather, eye pass beside the photon pass, each step a graph launch)"}
        log(f"graph-replayed passes: {graph_leg['ms_per_step']} ms per pass")
    a_hits = (a1["hit_points"] - a0["hit_points"]) / K
    a_nbs = (a1["neighbours"] - a0["neighbours"]) / K
    a_gather = a_hits * GATHER_BYTES_PER_HIT + a_nbs * GATHER_BYTES_PER_NEIGHBOUR
    # the roofline is given for the gather, the HBM-bound phase with the
    # largest time; the photon and eye phases are ray traversal (VALU / latency
    # bound, no HBM roofline) and are reported in phases_ms
    dom = max(("photons", "build", "eye", "gather"), key=lambda k: ms_alone[k])
    rk = "gather"
    achieved = algo[rk] / (ms[rk] * 1e-3) / 1e9
    achieved_alone = a_gather / (ms_alone[rk] * 1e-3) / 1e9
    traffic = read_traffic(os.path.join(ROOT, "profiles", f"traffic_{rk}_{args.workload}.md")) \
        if args.workload == "full" else None
    value = cfg["photons"] * K / el
    # the photon pass (the longest phase) is neither HBM nor MFMA bound: its
    # VALU rate, from per-launch counter totals (committed for the BVH8
    # megakernel; the photon pass's work does not depend on the radius)
    valu = read_valu(os.path.join(ROOT, "profiles", "valu_pass_kernels_full_material.md")) \
        if args.workload == "full" and cfg["bvh"] == 8 and args.photon_kernel in (None, "megakernel") else {}
    valu_line = None
    if "k_photon" in valu:
        lane_ops, slots = valu["k_photon"]
        t = ms_alone["photons"] * 1e-3
        valu_line = {"kernel": "k_photon (photon pass alone)", "bound": "valu",
                     "achieved": round(lane_ops / t / 1e12, 3), "peak": VALU_PEAK_T, "unit": "T lane-op/s",
                     "frac": round(lane_ops / t / 1e12 / VALU_PEAK_T, 4),
                     "issued_frac": round(slots / t / 1e12 / VALU_PEAK_T, 4),
                     "lane_ops_per_launch": lane_ops, "issued_slots_per_launch": slots,
                     "source": "profiles/valu_pass_kernels_full_material.md (counter totals per launch) / the "
                               "photon phase's own duration (phases_alone_ms.photons)"}
    out = {
        "metric": "photons/s + seconds to RMSE<=0.01, guided vs plain SPPM, 1 MI355X",
        "value": round(value, 1),
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(el / K * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": cfg["name"], "resolution": f"{cfg['width']}x{cfg['height']}",
                   "photons_per_pass": cfg["photons"], "triangles": cfg["triangles"], "bvh_width": cfg["bvh"] & 0xff,
                   "bvh_builder": "lbvh" if cfg["bvh"] & 0x100 else "sah",
                   "alpha": cfg["alpha"], "r0": round(cfg["r0"], 6), "camera": cam["type"] and "thin-lens" or "pinhole",
                   "photon_kernel": args.photon_kernel or "default", "wave_bounces": args.wave_bounces or "default",
                   "wave_lanes": args.wave_lanes or "default", "parallelism": "single GPU"},
        "roofline": {"kernel": rk, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": int(algo[rk]), "dominant_phase": dom,
                     "traffic_source": "profiles/traffic_gather_full.md (passes 2-4)" if traffic else None,
                     "duration": "mean span of the gather phase in the timed passes, beside the next photon pass"},
        "roofline_alone": {"kernel": rk, "bound": "hbm", "achieved": round(achieved_alone, 1),
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved_alone / HBM_PEAK_GBS, 4),
                           "algorithmic_bytes_per_launch": int(a_gather),
                           "duration": "gather phase alone: the same passes rendered again with the split steps"},
        "phases_ms": {k: round(v, 3) for k, v in ms.items()},
        "phases_alone_ms": {k: round(v, 3) for k, v in ms_alone.items()},
        "roofline_valu": valu_line,
        "graph_replay": graph_leg,
        "per_pass": {"hit_points": int(hits), "neighbours": int(nbs), "stored_photons": int(stored),
                     "gather_candidates": int(cands), "photon_rays": int(prays), "eye_rays": int(erays),
                     "hit_point_updates_per_s": round(hits / (ms_alone["gather"] * 1e-3), 1),
                     "photon_stage_photons_per_s": round(cfg["photons"] / (ms_alone["photons"] * 1e-3), 1),
                     "photon_stage_rays_per_s": round(prays / (ms_alone["photons"] * 1e-3), 1),
                     "eye_stage_rays_per_s": round(erays / (ms_alone["eye"] * 1e-3), 1),
                     "stage_rates": "per phase run alone (phases_alone_ms)"},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scene, cam, cfg, args.cpu_seconds)
    if not args.no_rmse and args.workload == "full":
        # the metric's second half: seconds to RMSE, guided vs plain, on the
        # occluded-light workload (1024^2, 1M photons, 297,700 triangles)
        w = rmse_workload(args)
        log(f"time to RMSE: {w['name']}, {w['scene'].nfacets} triangles, {w['res']}x{w['res']}")
        out["time_to_rmse"] = time_to_rmse(args, w, reference_image(args, w))
        tr = out["time_to_rmse"]
        tr["emission_bound"] = emission_gain_bound(w, tr["plain"]["ms_per_pass"], tr["plain"]["phases_ms"]["gather"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
