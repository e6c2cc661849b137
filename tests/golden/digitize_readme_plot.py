"""Digitise the crests and troughs of WarpVBD's README tip-displacement plot.

Source: the reference's asset/imgs/displacement_cantilever.png, written by
examples/run_cantilever.py:135-159 (10x3x3 voxels of 1 cm, SNH E=150 kPa,
nu=0.45, rho=1070, x = 0 face pinned, dt = 1 ms, 1500 steps, dx_tol 1e-9,
max_iter 100; mean z of the x = L face plotted against step index x dt).

The axes are calibrated on the plot's own grid: the vertical grid lines at
t = 0.25 ... 1.25 s and the left axis (t = 0), the horizontal grid lines at
z = 0, 0.5e-2, 1.0e-2 and 1.5e-2 m.  The curve is the matplotlib default blue
(31, 119, 180).  A crest is taken from the top edge of the drawn line over a
run of columns where it is lowest in a +-15 column window, a trough from the
bottom edge likewise, each moved by half the line thickness (7 px) to its
centre.  One pixel is 2.15 ms and 3.9e-5 m.

Run in a container that holds the reference:
    python tests/golden/digitize_readme_plot.py /path/to/displacement_cantilever.png
It writes readme_displacement_cantilever_extrema.json next to this script.
"""
import json
import os
import sys

import numpy as np
from PIL import Image

HERE = os.path.dirname(os.path.abspath(__file__))


def main(path):
    im = np.asarray(Image.open(path).convert("RGB")).astype(int)
    blue = (np.abs(im[:, :, 0] - 31) < 40) & (np.abs(im[:, :, 1] - 119) < 40) & (np.abs(im[:, :, 2] - 180) < 40)
    gray = np.all(np.abs(im - 176) < 12, axis=2)
    black = im.sum(2) < 100

    def centres(mask):
        idx = np.nonzero(mask)[0]
        groups = np.split(idx, np.nonzero(np.diff(idx) > 1)[0] + 1)
        return [float(g.mean()) for g in groups if len(g)]

    vcols = centres(gray.sum(0) > 150)          # t = 0.25, 0.5, 0.75, 1.0, 1.25
    hrows = centres(gray.sum(1) > 300)          # z = 1.5e-2, 1.0e-2, 0.5e-2, 0
    axis_col = centres(black.sum(0) > 300)[0]   # t = 0
    assert len(vcols) == 5 and len(hrows) == 4, (vcols, hrows)
    px_per_s = (vcols[-1] - vcols[0]) / 1.0
    px_per_m = (hrows[-1] - hrows[0]) / 1.5e-2
    assert abs((vcols[0] - axis_col) - 0.25 * px_per_s) < 2.0

    cols = np.arange(int(axis_col) + 2, int(axis_col + 1.5 * px_per_s) - 1)
    top = np.array([np.nonzero(blue[:, c])[0].min() for c in cols])
    bot = np.array([np.nonzero(blue[:, c])[0].max() for c in cols])
    thick = float(np.median((bot - top)[np.abs(np.gradient(top)) < 0.5]))
    out = []
    for kind, edge, pick in (("crest", top, np.min), ("trough", bot, np.max)):
        hits = [i for i in range(15, len(cols) - 15) if edge[i] == pick(edge[i - 15:i + 16])]
        runs = np.split(np.array(hits), np.nonzero(np.diff(hits) > 2)[0] + 1)
        for r in runs:
            c = float(cols[r].mean())
            row = edge[r[0]] + (thick / 2 if kind == "crest" else -thick / 2)
            out.append({"kind": kind, "t": (c - axis_col) / px_per_s, "z": (hrows[-1] - row) / px_per_m})
    out.sort(key=lambda e: e["t"])
    doc = {"source": "WarpVBD asset/imgs/displacement_cantilever.png (examples/run_cantilever.py:135-159)",
           "scene": {"voxels": [10, 3, 3], "h": 0.01, "youngs_modulus": 150e3, "poissons_ratio": 0.45,
                     "density": 1070.0, "dt": 1e-3, "steps": 1500, "dx_tol": 1e-9, "max_iter": 100},
           "time_axis": "step index x dt (the first plotted point is the state after step 1, at t = 0)",
           "pixel": {"s": 1.0 / px_per_s, "m": 1.0 / px_per_m}, "line_thickness_px": thick,
           "extrema": out}
    with open(os.path.join(HERE, "readme_displacement_cantilever_extrema.json"), "w") as f:
        json.dump(doc, f, indent=1)
    for e in out:
        print(f"{e['kind']:6s} t={e['t']:.4f} s  z={e['z']:.4e} m")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/asset/imgs/displacement_cantilever.png")
