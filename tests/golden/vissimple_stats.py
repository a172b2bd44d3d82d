"""Statistics of a z-slice in the column format of the reference's analysis/vissimple samples
(x y d vx vy p phi X1, one row per cell): the numbers two Kelvin-Helmholtz runs started from
different random perturbations can share, since the turbulence itself differs run to run.

Per column d, vx, vy, p, X1: min, max, mean, standard deviation and the 5/25/50/75/95 %
quantiles; the slice's mixing measure <X1 (1 - X1)> and sqrt(<vy^2>) (unweighted: the cells
of the slice are equal).
"""
from __future__ import annotations

import numpy as np

COLUMNS = {"d": 2, "vx": 3, "vy": 4, "p": 5, "X1": 7}
QUANTILES = (5, 25, 50, 75, 95)


def read_slices(paths) -> np.ndarray:
    """rows of 8 numbers from one or more slice files (comment lines start with #)"""
    rows = []
    for path in paths:
        with open(path) as f:
            for line in f:
                if line.startswith("#") or not line.strip():
                    continue
                v = line.split()
                if len(v) == 8:
                    rows.append([float(x) for x in v])
    return np.array(rows)


def slice_stats(a: np.ndarray) -> dict:
    out = {"cells": int(a.shape[0])}
    for name, c in COLUMNS.items():
        col = a[:, c]
        q = np.percentile(col, QUANTILES)
        out[name] = {"min": float(col.min()), "max": float(col.max()), "mean": float(col.mean()),
                     "std": float(col.std()), **{f"q{p}": float(v) for p, v in zip(QUANTILES, q)}}
    X = a[:, COLUMNS["X1"]]
    out["mix"] = float((X * (1.0 - X)).mean())
    out["vy_rms"] = float(np.sqrt((a[:, COLUMNS["vy"]] ** 2).mean()))
    return out
