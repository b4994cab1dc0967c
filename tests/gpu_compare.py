"""Shared comparisons of the GPU tests: the HIP path (through the C ABI)
against the CPU oracle.

Bar: integer outputs bit-exact -- flattened photon lists (count, order,
positions, directions, IOR), hit point types and facets, per-hit-point
neighbour index sets and counts, N and NumGather; flux, PDFs, weights and the
image within 1e-4 relative RMSE (FP32).
"""
import numpy as np

import pyoracle

REL_TOL = 1e-4


def rel_rmse(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.sqrt(np.mean(b * b))
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(den, 1e-30))


def say(msg):
    print(msg, flush=True)


def compare_pass(gpu, ref, stored_ref, label, record=True):
    """One whole pass: photon list, hit points, counts (and neighbour sets
    when recorded), DirectLight, shared hit points."""
    pg = gpu.read_photons()
    pr = ref.read_photons(stored_ref)
    assert pg["pos"].shape[0] == stored_ref, f"{label}: stored photons {pg['pos'].shape[0]} != {stored_ref}"
    np.testing.assert_array_equal(pg["pos"], pr["pos"])
    np.testing.assert_array_equal(pg["indir"], pr["indir"])
    np.testing.assert_array_equal(pg["index"], pr["index"])
    assert rel_rmse(pg["flux"], pr["flux"]) <= REL_TOL
    assert rel_rmse(pg["pdf"], pr["pdf"]) <= REL_TOL
    del pg, pr
    hg, hr = gpu.read_hitpoints(), ref.read_hitpoints()
    np.testing.assert_array_equal(hg["type"], hr["type"])
    np.testing.assert_array_equal(hg["facet"], hr["facet"])
    hit = hr["type"] == 1
    np.testing.assert_array_equal(hg["pos"][hit], hr["pos"][hit])
    np.testing.assert_array_equal(hg["normal"][hit], hr["normal"][hit])
    np.testing.assert_array_equal(hg["outdir"][hit], hr["outdir"][hit])
    assert rel_rmse(hg["weight"][hit], hr["weight"][hit]) <= REL_TOL
    del hg, hr
    np.testing.assert_array_equal(gpu.read_counts(), ref.read_counts())
    if record:
        og, ig = gpu.read_neighbours()
        orf, irf = ref.read_neighbours()
        np.testing.assert_array_equal(og, orf)
        np.testing.assert_array_equal(ig, irf)
        say(f"  {label}: {len(ig)} neighbour indices identical")
        del og, ig, orf, irf
    dg, dr = gpu.read_direct(), ref.read_direct()
    np.testing.assert_array_equal(dg["num"], dr["num"])
    assert rel_rmse(dg["hit"], dr["hit"]) <= REL_TOL
    assert rel_rmse(dg["shared"], dr["shared"]) <= REL_TOL
    sg, sr = gpu.read_shared(), ref.read_shared()
    np.testing.assert_array_equal(sg["num_photon"], sr["num_photon"])
    np.testing.assert_array_equal(sg["num_gather"], sr["num_gather"])
    assert rel_rmse(sg["radius"], sr["radius"]) <= 1e-6
    e = rel_rmse(sg["radiance"], sr["radiance"])
    say(f"  {label}: {stored_ref} photons; radiance rel RMSE {e:.2e}")
    assert e <= REL_TOL, f"{label}: radiance rel RMSE {e}"
    return e


def oracle_pass(ref, record=False):
    """The oracle's whole pass; returns its stored photon count."""
    ref.full_pass(record=record)
    return pyoracle.orc().orc_flatten(ref.h)


def spot_check(gpu, orender, step, npix, npaths, seed, sampler_id=0, extra_pixels=()):
    """Spot parity of a pass at full size: chosen pixels (eye paths, with the
    shared camera stream advanced to their exact draw) and photon paths
    against the oracle's single-pixel / single-path restatement."""
    rng = np.random.default_rng(seed)
    NP = gpu.W * gpu.H
    pix = np.concatenate([rng.choice(NP, npix, replace=False), np.asarray(extra_pixels, np.int64)]).astype(np.int64)
    smp = pyoracle.uniform_samples_at(sampler_id, (step - 1) * NP + pix)
    ref = orender.eye_pixels(step, pix, smp)
    hg = gpu.read_hitpoints()
    np.testing.assert_array_equal(hg["type"][pix], ref["type"])
    np.testing.assert_array_equal(hg["facet"][pix], ref["facet"])
    hit = ref["type"] == 1
    for k in ("pos", "normal", "outdir"):
        np.testing.assert_array_equal(hg[k][pix][hit], ref[k][hit])
    np.testing.assert_allclose(hg["weight"][pix][hit], ref["weight"][hit], rtol=1e-5, equal_nan=True)
    del hg
    counts = gpu.read_path_counts()
    off = np.concatenate([[0], np.cumsum(counts, dtype=np.int64)])
    ph = gpu.read_photons()
    for p in rng.choice(gpu.N, npaths, replace=False):
        r = orender.photon_path(step, p)
        n = len(r["pos"])
        assert counts[p] in (n, n + 1) if n > 1 or r["index"].size == 0 else counts[p] >= 1
        sl = slice(off[p], off[p] + n)
        np.testing.assert_array_equal(ph["pos"][sl], r["pos"])
        np.testing.assert_array_equal(ph["indir"][sl], r["indir"])
        np.testing.assert_array_equal(ph["index"][sl], r["index"])
        np.testing.assert_allclose(ph["flux"][sl], r["flux"], rtol=1e-5)
        np.testing.assert_allclose(ph["pdf"][sl], r["pdf"], rtol=1e-5)
    return ref
