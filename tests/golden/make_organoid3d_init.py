"""Writes tests/golden/organoid3d_init.npz: the explicit initial lattice of
notebooks/3d_organoid_demo.ipynb (cells 5 and 6) — 20 seed voxels drawn
uniformly in a ball of radius 21 at the centre of a 64^3 grid from
numpy's default_rng(11), grown breadth-first (6-connected, inside the ball)
until 5% of the ball is filled — and the shuffled type of each cell.
Stored sparsely: the flat row-major index and id of each occupied voxel.

This script restates the placement as the notebook describes it; nothing
from the reference is read or run."""
import os
from collections import deque

import numpy as np

GRID = (64, 64, 64)
N_CELLS = 20
RADIUS = 21.0
FILL = 0.05


def place(rng):
    centre = np.array(GRID) / 2.0
    seeds = set()
    while len(seeds) < N_CELLS:
        p = rng.uniform(-RADIUS, RADIUS, size=(3,))
        if np.linalg.norm(p) > RADIUS:
            continue
        v = tuple(np.round(p + centre).astype(int))
        if all(0 <= a < g for a, g in zip(v, GRID)):
            seeds.add(v)
    seeds = list(seeds)  # the set's iteration order numbers the cells
    idx = np.indices(GRID).transpose(1, 2, 3, 0).astype(np.float64)
    ball = np.linalg.norm(idx - centre, axis=-1) <= RADIUS
    lab = np.zeros(GRID, dtype=np.uint32)
    todo = deque()
    for cid, v in enumerate(seeds, start=1):
        lab[v] = cid
        todo.append(v)
    budget = int(FILL * ball.sum())
    placed = len(seeds)
    steps = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    while todo and placed < budget:
        x = todo.popleft()
        cid = lab[x]
        for s in steps:
            y = tuple(a + b for a, b in zip(x, s))
            if not all(0 <= a < g for a, g in zip(y, GRID)):
                continue
            if not ball[y] or lab[y]:
                continue
            lab[y] = cid
            todo.append(y)
            placed += 1
            if placed >= budget:
                break
    return lab


def main():
    rng = np.random.default_rng(11)
    lab = place(rng)
    types = np.repeat(np.arange(4), 5)
    rng.shuffle(types)
    flat = lab.ravel()
    occ = np.nonzero(flat)[0]
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "organoid3d_init.npz")
    np.savez_compressed(out, index=occ.astype(np.uint32), id=flat[occ].astype(np.uint32),
                        cell_type_of=types.astype(np.uint32))
    print(f"{len(occ)} occupied voxels, {int(flat.max())} cells -> {out}")


if __name__ == "__main__":
    main()
