"""Relative-L1 parity measure shared by the GPU tests, smoke() and bench.

Per variable: sum |a - b| / sum |b| over interior cells.  Components of a vector
(v or m, B) are divided by the sum of the vector magnitude instead of their own, and psi
by that of B: a component that is near zero in this problem (vz, By, psi grow from
round-off) would otherwise turn round-off into an O(1) "relative" error.  A variable
whose scale is zero is measured absolutely (mean |a - b|).
"""
import numpy as np

PRIM_GROUPS = {1: "v", 2: "v", 3: "v", 7: "b", 8: "b", 9: "b", 10: "b"}
CONS_GROUPS = {1: "v", 2: "v", 3: "v", 5: "b", 6: "b", 7: "b", 8: "b"}


def rel_l1(a: np.ndarray, b: np.ndarray, kind: str = "prim") -> np.ndarray:
    """a, b: (nv, ...) arrays of interior cells; kind 'prim' (12 vars) or 'cons' (10 vars)."""
    nv = a.shape[0]
    a = a.reshape(nv, -1)
    b = b.reshape(nv, -1)
    groups = PRIM_GROUPS if kind == "prim" else CONS_GROUPS
    vi = [1, 2, 3]
    bi = [7, 8, 9] if kind == "prim" else [5, 6, 7]
    scale = {"v": np.sqrt((b[vi] ** 2).sum(0)).sum(), "b": np.sqrt((b[bi] ** 2).sum(0)).sum()}
    out = np.zeros(nv)
    for n in range(nv):
        num = np.abs(a[n] - b[n]).sum()
        den = scale[groups[n]] if n in groups else np.abs(b[n]).sum()
        out[n] = num / den if den > 1e-30 * b.shape[1] else num / b.shape[1]
    return out
