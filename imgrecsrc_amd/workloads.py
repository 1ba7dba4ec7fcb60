"""Seeded synthetic workloads of the benchmark configurations.

* ``ragged_geoms``: 64 independent slices of uneven size and view count, mixed parallel
  and fan beam (the batched configuration).
* ``STACK512``: one slice of the 512-slice parallel stack (512^2, 720 views over 180
  degrees, 729 bins); ``stack_truth`` builds its seeded phantoms.

Geometry tuples are (n, prjWidth, np, prjFull, dSize, effectiveRate, d), the fields of
the reference's prjConf (prj/prj.h:42-57); d = 0 is parallel beam.
"""
from __future__ import annotations

import numpy as np

FAN1024 = (1024, 1024, 360, 360, 1.0, 1.0, 4096.0)
PAR256 = (256, 367, 180, 360, 1.0, 1.0, 0.0)
FAN512 = (512, 729, 360, 360, 1.0, 1.0, 2000.0)
STACK512 = (512, 729, 720, 1440, 1.0, 1.0, 0.0)


def ragged_geoms(seed: int = 4, count: int = 64):
    """`count` slices: n from {128, 256, 384, 512, 768, 1024}, views U{90..720}, parallel
    or fan with equal odds, bins = ceil(n sqrt 2) + 4; fan d = 4n over 360 degrees
    (prjFull = views), parallel over 180 degrees (prjFull = 2 views)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        n = int(rng.choice([128, 256, 384, 512, 768, 1024]))
        na = int(rng.integers(90, 721))
        fan = bool(rng.integers(0, 2))
        W = int(np.ceil(n * np.sqrt(2))) + 4
        out.append((n, W, na, na, 1.0, 1.0, 4.0 * n) if fan else (n, W, na, 2 * na, 1.0, 1.0, 0.0))
    return out


def stack_truth(nslices: int, n: int = 512):
    """Phantoms of a slice stack as one float32 array (nslices, n, n): eight seeded
    ellipse phantoms (seeds 500..507), slice i is phantom i % 8 shifted down by
    (i // 8) % 32 rows (cyclically)."""
    from .phantom import ellipse_params, rasterise
    base = [rasterise(ellipse_params(500 + k, n), n).astype(np.float32) for k in range(8)]
    return np.stack([np.roll(base[i % 8], (i // 8) % 32, axis=0) for i in range(nslices)])
