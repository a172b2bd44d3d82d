"""Magnetic fields whose discrete divergence is known in closed form (test-only, numpy).

The divB diagnostic is the second-order central difference of the cell-centred field,
  divB(i,j,k) = (B1[i+1]-B1[i-1])/(x1b[i+1]-x1b[i-1]) + (same in y, z),
taken on interior cells whose six neighbours are interior, and max|divB| is its largest
magnitude there.  On a non-uniform grid of centres x:
  * B = (a x, b y, c z)        -> a + b + c in every cell;
  * B = (x^2, y^2, z^2)        -> (x[i+1] + x[i-1]) + (y[j+1] + y[j-1]) + (z[k+1] + z[k-1]),
                                  largest at the last eligible cell of every axis;
  * B = ((2-x)^2, (2-y)^2, (2-z)^2) -> the same sums minus 12, largest in magnitude at the
                                  first eligible cell of every axis (coordinates below 2);
  * B = (f(y,z), g(z,x), h(x,y)) -> exactly 0 (each difference is of equal values).
On a uniform grid of spacing h:
  * B = (sin(w x), 0, 0)        -> cos(w x_i) sin(w h) / h.
A one-sided difference, a spacing taken from the faces or from another axis, or an
eligible range one cell too wide, each changes the quadratic or sine value.
"""
from __future__ import annotations

import numpy as np

NGH = 2


def nonuniform_faces(n: int, lo: float, length: float, seed: int) -> np.ndarray:
    """n + 5 face coordinates: n interior cells of random widths spanning [lo, lo + length],
    ghost cells as wide as the edge cells they mirror"""
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.5, 1.5, n)
    w *= length / w.sum()
    faces = np.empty(n + 5)
    faces[NGH] = lo
    faces[NGH + 1:NGH + n + 1] = lo + np.cumsum(w)
    faces[1] = lo - w[0]
    faces[0] = faces[1] - w[1]
    faces[n + 3] = faces[n + 2] + w[-1]
    faces[n + 4] = faces[n + 3] + w[-2]
    return faces


def centres(faces: np.ndarray) -> np.ndarray:
    b = np.zeros_like(faces)
    b[:-1] = 0.5 * (faces[1:] + faces[:-1])
    return b


def fill(P: np.ndarray, xb, yb, zb, kind: str):
    """P: (ktot, jtot, 12, itot) host-layout primitives; a smooth background state and the
    magnetic field `kind`"""
    z = zb[:, None, None]
    y = yb[None, :, None]
    x = xb[None, None, :]
    shape = (zb.size, yb.size, xb.size)
    P[...] = 0.0
    P[:, :, 0, :] = 1.0
    P[:, :, 4, :] = 1.5
    P[:, :, 5, :] = 1.0
    P[:, :, 6, :] = np.sqrt(5.0 / 3.0)
    P[:, :, 2, :] = np.broadcast_to(0.01 * np.sin(3.0 * x + 2.0 * y), shape)  # vy, for the sums
    P[:, :, 11, :] = np.broadcast_to(0.5 + 0.4 * np.tanh(y), shape)
    if kind == "linear":
        b = (0.7 * x, -0.3 * y, 1.1 * z)
    elif kind == "quadratic":
        b = (x * x, y * y, z * z)
    elif kind == "quadratic_low":
        b = ((2 - x) ** 2, (2 - y) ** 2, (2 - z) ** 2)
    elif kind == "sine":
        b = (np.sin(2 * np.pi * x), 0 * y, 0 * z)
    elif kind == "solenoidal":
        b = (np.sin(2 * np.pi * y) + z, np.cos(2 * np.pi * z) * (1 + x), np.sin(2 * np.pi * x) * y)
    else:
        raise ValueError(kind)
    for d in range(3):
        P[:, :, 7 + d, :] = np.broadcast_to(b[d], shape)


def expected_max_divb(kind: str, xb, yb, zb, n3) -> float:
    """closed-form max|divB| over the eligible cells (index NGH+1 .. NGH+n-2 on each axis)"""
    el = [slice(NGH + 1, NGH + n - 1) for n in n3]
    if kind == "linear":
        return abs(0.7 - 0.3 + 1.1)
    if kind == "solenoidal":
        return 0.0
    if kind in ("quadratic", "quadratic_low"):
        s = [c[NGH + 2:NGH + n] + c[NGH:NGH + n - 2] for c, n in zip((xb, yb, zb), n3)]
        if kind == "quadratic":
            return float(np.max(s[0]) + np.max(s[1]) + np.max(s[2]))
        return float(12.0 - np.min(s[0]) - np.min(s[1]) - np.min(s[2]))
    if kind == "sine":
        h = xb[NGH + 1] - xb[NGH]
        w = 2 * np.pi
        return float(np.max(np.abs(np.cos(w * xb[el[0]]))) * np.sin(w * h) / h)
    raise ValueError(kind)
