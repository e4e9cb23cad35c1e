"""Counts the algorithmic work of the build workloads with the CPU oracle.

TEST INFRASTRUCTURE: it runs the oracle's nearest search, which follows the
reference's traversal order (KdTree.h:244-328), and counts the point-triangle
and point-box tests it makes per voxel.  The algorithmic flops of a voxel are
80 per triangle test plus 12 per box test.  The result is written to
profiles/algorithmic_work_counts.json, which bench.py reads to price its build
rooflines.

    python oracle/algorithmic_work.py
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[:0] = [HERE, os.path.join(ROOT, "sdf-contact-module_amd")]

import oracle as O  # noqa: E402
from sdfo import meshes, scenes  # noqa: E402

FLOPS_PER_TRI = 80.0
FLOPS_PER_BOX = 12.0


def _rates(r):
    t = float(np.mean(r["tri_tests"]))
    b = float(np.mean(r["box_tests"]))
    return dict(tri_tests_per_voxel=round(t, 3), box_tests_per_voxel=round(b, 3),
                flops_per_voxel=round(FLOPS_PER_TRI * t + FLOPS_PER_BOX * b, 1))


def torus128():
    v, t = scenes.ring_mesh(256, 64)
    g = O.setup_grid(v, 0.1, 0.0, 128, 128)
    r = O.signed_nearest(v, t, g.centers())
    out = dict(voxels=g.total, sample="every voxel", **_rates(r))
    out["flops"] = out["flops_per_voxel"] * g.total
    return out


def tube512():
    """Active bricks from the activation rule on the brick centres (one full
    signed search each, as the reference does), then every voxel of 2000
    seeded active bricks."""
    v, t = meshes.hollow_cylinder(0.04, 0.03, 0.05, 1024, 64)
    g = O.setup_grid(v, 0.1, 0.0, 512, 512)
    h = g.voxel.max()
    act_dist = 4 * h + 2 * h + 0.5 * np.sqrt(3.0) * 8 * h
    nb = [(s + 7) // 8 for s in g.shape]
    cells = np.arange(int(np.prod(nb)))
    b3 = np.stack([cells % nb[0], (cells // nb[0]) % nb[1], cells // (nb[0] * nb[1])], 1)
    lo = b3 * 8
    hi = np.minimum(lo + 8, np.array(g.shape))
    centre = g.bmin + g.voxel * (0.5 * (lo + hi))
    rb = O.signed_nearest(v, t, centre)
    active = np.flatnonzero(np.abs(rb["sd"]) <= act_dist)
    rng = np.random.default_rng(0)
    pick = rng.choice(active, 2000, replace=False)
    L = np.arange(512)
    off = np.stack([L % 8, (L // 8) % 8, L // 64], 1)
    ijk = (b3[pick][:, None, :] * 8 + off[None]).reshape(-1, 3)
    ijk = ijk[np.all(ijk < np.array(g.shape), 1)]
    r = O.signed_nearest(v, t, g.bmin + g.voxel * (ijk + 0.5))
    fill = _rates(r)
    act = _rates(rb)
    n_vox = len(active) * 512
    return dict(active_bricks=int(len(active)), active_voxels=n_vox, bricks=int(len(cells)),
                sample="every voxel of 2000 seeded active bricks; activation on every brick centre",
                fill=fill, activation=act,
                flops=fill["flops_per_voxel"] * n_vox + act["flops_per_voxel"] * len(cells))


def gears64():
    """4096 seeded voxels of each of the 64 gear grids, scaled by its voxel count."""
    batch = scenes.gear_batch(64, seed=5)
    rng = np.random.default_rng(1)
    total_vox, total_flops = 0, 0.0
    for v, t, res in batch:
        g = O.setup_grid(v, 0.1, 0.0, res, res)
        idx = rng.choice(g.total, min(4096, g.total), replace=False)
        nx, ny = g.shape[0], g.shape[1]
        ijk = np.stack([idx % nx, (idx // nx) % ny, idx // (nx * ny)], 1)
        r = O.signed_nearest(v, t, g.bmin + g.voxel * (ijk + 0.5))
        total_vox += g.total
        total_flops += _rates(r)["flops_per_voxel"] * g.total
    return dict(meshes=len(batch), voxels=int(total_vox), flops=total_flops,
                flops_per_voxel=round(total_flops / total_vox, 1),
                sample="4096 seeded voxels per grid, scaled by the grid's voxel count")


def main():
    t0 = time.time()
    out = dict(
        flop_model="80 flops per point-triangle test + 12 per point-box test, tests counted "
                   "along the reference's nearest search (KdTree.h:244-328); the ray-parity "
                   "sign walk is not counted",
        command="python oracle/algorithmic_work.py",
        torus_256x64_dense_128=torus128(), tube_1024x64_sparse_512=tube512(),
        gears_64_batched=gears64())
    path = os.path.join(ROOT, "profiles", "algorithmic_work_counts.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))
    print(f"{time.time() - t0:.1f} s", file=sys.stderr)


if __name__ == "__main__":
    main()
