"""Write golden vectors of the reference's coarse-to-fine NCC search.

Run once, in the build container only (the reference is not available anywhere else):

    PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden_pyramid.py

It imports ``al_dvc`` 1.2.1 from ``/root/reference/src`` and runs ``pyramid_search``
(``solver/integer_search.py:617-694``; numba absent, so its NCC maps come from the reference's FFT
engine, and the cleaning between levels from its median test and spring fill) and
``phase_correlation_shift`` (``:246-294``) on the large-shift case of the reference's own test
(``tests/test_integer_search.py:74-90``): an 80 x 72 x 72 bead pair translated by (11.3, -9.6, 6.4),
a 4 x 4 x 4 node lattice, 16^3 templates, radius 4, levels chosen automatically, without and with the
global shift. Inputs are this repository's seeded generator output, stored next to the outputs in
``tests/golden/golden_pyramid.npz``. Nothing here is imported by the tests.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, "/root/reference/src")
sys.dont_write_bytecode = True

from al_dvc.solver import integer_search as ris  # noqa: E402

from pyaldvc_amd import synthetic as syn  # noqa: E402

SHIFT = (11.3, -9.6, 6.4)


def main() -> None:
    f_raw, g_raw = syn.bead_pair_np((80, 72, 72), 5000, 2.0, syn.translation(SHIFT), seed=12, noise=0.0)
    f, g = syn.normalize_np(f_raw), syn.normalize_np(g_raw)
    ax = np.arange(24, 50, 8)
    gz, gy, gx = np.meshgrid(ax, ax, ax, indexing="ij")
    coords = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1).astype(np.int64)
    out = {"f": f, "g": g, "coords": coords, "grid": np.array([4, 4, 4]), "shift": np.array(SHIFT)}
    out["levels_auto"] = np.array(ris.auto_pyramid_levels(f.shape, (16, 16, 16), (4, 4, 4), f=f))
    gs = ris.phase_correlation_shift(f, g)
    out["global_shift"] = np.asarray(gs, dtype=np.float64)
    for key, shift in (("plain", None), ("gshift", gs)):
        r = ris.pyramid_search(f, g, coords, (4, 4, 4), (16, 16, 16), (4, 4, 4), 0, global_shift=shift)
        for k in ("disp", "cc", "pce", "clipped", "ok"):
            out[f"{key}_{k}"] = np.asarray(r[k])
        out[f"{key}_levels"] = np.array(r["levels"])
        out[f"{key}_radii"] = np.array([p["radius"] for p in r["per_level"]])
        out[f"{key}_n_ok"] = np.array([p["n_ok"] for p in r["per_level"]])
        out[f"{key}_expansions"] = np.array(r["expansions"])
    path = os.path.join(HERE, "golden_pyramid.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path))


if __name__ == "__main__":
    main()
