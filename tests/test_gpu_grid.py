"""The photon grid, the gather structure that takes the place of the
reference's photon kd-tree (kdtree<Photon>::CreateTree, R/kdtree.h:417-431),
read through the C ABI: the cell ranges partition the stored photons, every
photon sits in the cell of its position, and inside a cell the photons keep
their flattened (reference) order -- the contract of the stable radix sort
that lays the grid out.  Checked on a pass of a few hundred thousand stored
photons (tens of sort tiles) and on passes of a few million (hundreds)."""
import numpy as np
import pytest

import scenes
import sppm_amd
from gpu_compare import say

pytestmark = pytest.mark.gpu


def check_grid(r, label):
    g = r.read_grid()
    ph = r.read_photons()
    n = g["nphotons"]
    assert n == ph["pos"].shape[0] and n > 0
    start, order = g["start"].astype(np.int64), g["order"].astype(np.int64)
    assert start[0] == 0 and start[-1] == n and np.all(np.diff(start) >= 0)
    assert np.array_equal(np.sort(order), np.arange(n))
    cell_of_slot = np.repeat(np.arange(g["ncells"], dtype=np.int64), np.diff(start))
    # cells in increasing order, flattened indices increasing inside a cell
    assert np.all(np.diff(cell_of_slot * n + order) > 0)
    # the cell of each photon's position (float32, as the grid computes it)
    p = ph["pos"][order]
    inv = np.float32(1) / g["cell"]
    c = []
    for a, m in enumerate((g["nx"], g["ny"], g["nz"])):
        f = np.floor((p[:, a] - g["bmin"][a]) * inv)
        c.append(np.clip(f, 0, m - 1).astype(np.int64))
    cell = (c[2] * g["ny"] + c[1]) * g["nx"] + c[0]
    np.testing.assert_array_equal(cell, cell_of_slot)
    occupied = int(np.count_nonzero(np.diff(start)))
    say(f"{label}: {n} photons in {occupied} of {g['ncells']} cells, largest cell {int(np.diff(start).max())}")


def test_grid_layout_box(box_scene):
    r = sppm_amd.Renderer(box_scene, scenes.pinhole_camera(64, 64), 0.7, 0.01, 100_000, bvh_width=4)
    for p in range(2):
        r.pass_(stats=False)
        check_grid(r, f"box pass {p + 1}")
    r.close()


def test_grid_layout_full_workload():
    scene = scenes.full_scene(freq=100, tex_size=512)
    r = sppm_amd.Renderer(scene, scenes.thinlens_camera(256, 256), 0.7, 0.02 * scene.diagonal(), 1_000_000,
                          bvh_width=8)
    for p in range(2):
        r.pass_(stats=False)
        check_grid(r, f"full-material pass {p + 1}")
    r.close()
