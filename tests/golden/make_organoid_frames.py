"""Writes tests/golden/organoid2d_frames.npz from the reference's 2D organoid
notebook (notebooks/2d_organoid_demo.ipynb).

The notebook runs a 70x70 fixed-boundary, four-type tissue (seed 11, six
cells per type, scatter-and-grow initialisation, burn-in 0, readout
400000 MCS sampled every 2500; cell 3 and cell 5) and embeds every one of
its 160 sampled lattices as a PNG frame of the animation in cell 12's
output.  Each frame draws the lattice twice with nearest-neighbour
interpolation: left panel coloured by cell id (tab20 palette, colour
(id - 1) mod 20; cell 9), right panel coloured by type (cell 9's
type_colors).  Cell ids are assigned type by type in add_cells order
(crates/cpm-core/src/initialization.rs:202-206), so type t holds ids
6t+1 .. 6t+6 and the pair (type colour, id colour) names the id of every
site uniquely.

This script samples each lattice site at the centre of its block of
pixels in both panels, checks that the 3x3 pixels around each centre are
one colour (those at least two pixels inside the panel's frame
line, which is anti-aliased) and that both panels agree on medium, and stores the recovered
cell-id lattices (uint8, 160 x 70 x 70, row-major as Sample.lattice) with
the MCS of each frame (2500 * (k + 1)).  Only data leaves the notebook:
the lattices and the notebook's printed results (cell 5).

Usage (needs the reference checkout at /root/reference and Pillow):
    python tests/golden/make_organoid_frames.py
"""

from __future__ import annotations

import base64
import io
import json
import os
import re

import numpy as np

NOTEBOOK = "/root/reference/notebooks/2d_organoid_demo.ipynb"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "organoid2d_frames.npz")
GRID = 70
CELLS_PER_TYPE = 6

# cell 9: type_colors (medium, stem, progenitor, differentiated, surface),
# as 8-bit RGB in the rendered PNG
TYPE_RGB = {
    (255, 255, 255): -1,
    (216, 51, 51): 0,
    (242, 153, 25): 1,
    (51, 140, 216): 2,
    (89, 178, 89): 3,
}
# matplotlib's tab20 (the Vega category20 palette), in order
TAB20 = [
    (31, 119, 180), (174, 199, 232), (255, 127, 14), (255, 187, 120), (44, 160, 44),
    (152, 223, 138), (214, 39, 40), (255, 152, 150), (148, 103, 189), (197, 176, 213),
    (140, 86, 75), (196, 156, 148), (227, 119, 194), (247, 182, 210), (127, 127, 127),
    (199, 199, 199), (188, 189, 34), (219, 219, 141), (23, 190, 207), (158, 218, 229),
]
CELL_RGB = {c: i for i, c in enumerate(TAB20)}
CELL_RGB[(255, 255, 255)] = -1


def _panel_box(img: np.ndarray):
    """Pixel columns of the two panels' black frames and the frame rows."""
    dark = img.sum(axis=2) < 100
    cols = np.where(dark.sum(axis=0) > 250)[0]
    rows = np.where(dark.sum(axis=1) > 250)[0]
    assert len(cols) == 4 and len(rows) == 2, (cols, rows)
    return (cols[0], cols[1]), (cols[2], cols[3]), (rows[0], rows[1])


def _sample(img, box_x, box_y, table):
    x0, x1 = box_x
    y0, y1 = box_y
    out = np.empty((GRID, GRID), dtype=np.int16)
    wx = (x1 - x0) / GRID
    wy = (y1 - y0) / GRID
    for r in range(GRID):
        py = int(y0 + (r + 0.5) * wy)
        for c in range(GRID):
            px = int(x0 + (c + 0.5) * wx)
            ylo, yhi = max(py - 1, y0 + 2), min(py + 2, y1 - 1)
            xlo, xhi = max(px - 1, x0 + 2), min(px + 2, x1 - 1)
            block = img[ylo:yhi, xlo:xhi].reshape(-1, 3)
            assert (block == block[0]).all(), ("mixed colours at a site centre", r, c)
            out[r, c] = table[tuple(int(v) for v in block[0])]
    return out


def decode_frame(png: bytes) -> np.ndarray:
    from PIL import Image

    img = np.array(Image.open(io.BytesIO(png)).convert("RGB")).astype(np.int32)
    by_cell_x, by_type_x, ys = _panel_box(img)
    types = _sample(img, by_type_x, ys, TYPE_RGB)
    colours = _sample(img, by_cell_x, ys, CELL_RGB)
    assert ((types < 0) == (colours < 0)).all(), "panels disagree on medium"
    ids = np.zeros((GRID, GRID), dtype=np.uint8)
    n_ids = 4 * CELLS_PER_TYPE
    for t in range(4):
        for cid in range(CELLS_PER_TYPE * t + 1, CELLS_PER_TYPE * (t + 1) + 1):
            ids[(types == t) & (colours == (cid - 1) % 20)] = cid
    occupied = types >= 0
    assert (ids[occupied] > 0).all(), "a site's id colour does not belong to its type"
    assert ids.max() <= n_ids
    return ids


def main() -> None:
    nb = json.load(open(NOTEBOOK))
    html = "".join(nb["cells"][12]["outputs"][0]["data"]["text/html"])
    frames = re.findall(r'frames\[(\d+)\] = "data:image/png;base64,([^"]+)"', html)
    printed = "".join(nb["cells"][5]["outputs"][0]["text"])
    lattices = []
    for k, (idx, b64) in enumerate(frames):
        assert int(idx) == k
        lattices.append(decode_frame(base64.b64decode(b64.replace("\\n", ""))))
    lattices = np.stack(lattices)
    mcs = 2500 * (np.arange(len(lattices), dtype=np.uint64) + 1)
    np.savez_compressed(OUT, lattices=lattices, mcs=mcs,
                        printed=np.array(printed))
    print(f"{len(lattices)} frames -> {OUT}")
    print(printed)


if __name__ == "__main__":
    main()
