"""Snapshots in the reference's on-disk formats (src_cpp_kokkos_host/output.cpp), written on
the host from a downloaded state, so the reference's analysis tools and ParaView read GPU
results unchanged.

Files for snapshot index `nout` under `outdir`:
  bindata/unf%05d.dat      "# time dt" and the three global sizes      (output.cpp:87-97)
  bindata/grid{1,2,3}D.bin [2][N+1] f64: cell centres then faces        (output.cpp:101-165, 241-257)
  bindata/field%05d.bin    [nvars][z][y][x] f64, d v1 v2 v3 b1 b2 b3 bp p X1 (output.cpp:172-203, 262-280)
  bindata/field%05d.xmf    XDMF 2 3DRectMesh pointing into the two above (output.cpp:285-360)
  ascdata/snap000-%05d.csv x y d vx vy p X1 on the first interior z plane (output.cpp:365-407)

The state can come from Box.download_prim() or, without stalling the run, from
Box.snapshot_begin() ... Box.snapshot_end(): the copy to pinned host memory then travels on a
second stream while the following steps run.
"""
from __future__ import annotations

import os

import numpy as np

NGH = 2
VARNAMES = ["d", "v1", "v2", "v3", "b1", "b2", "b3", "bp", "p", "X1"]
# primitive indices of the packed variables (src_cpp_kokkos_host/output.cpp:264-280)
PACK = [0, 1, 2, 3, 7, 8, 9, 10, 5, 11]


def pack_fields(P: np.ndarray) -> np.ndarray:
    """(ktot, jtot, 12, itot) primitive array -> (10, nz, ny, nx) packed interior fields"""
    inner = P[NGH:P.shape[0] - 3, NGH:P.shape[1] - 3, :, NGH:P.shape[3] - 3]
    return np.ascontiguousarray(inner[:, :, PACK, :].transpose(2, 0, 1, 3))


def write_snapshot(outdir: str, nout: int, time: float, dt: float, P: np.ndarray, grid: list[np.ndarray],
                   ascii_out: bool = True) -> list[str]:
    """grid = [x1a, x1b, x2a, x2b, x3a, x3b] (reference GridArray order)"""
    bdir = os.path.join(outdir, "bindata")
    os.makedirs(bdir, exist_ok=True)
    fields = pack_fields(P)
    nvars, nz, ny, nx = fields.shape
    written = []
    p = os.path.join(bdir, f"unf{nout:05d}.dat")
    with open(p, "w") as f:
        f.write(f"# {time:.17g} {dt:.17g}\n# {nx} \n# {ny} \n# {nz} \n")
    written.append(p)
    x1a, x1b, x2a, x2b, x3a, x3b = grid
    for name, a, b, n in (("grid1D.bin", x1a, x1b, nx), ("grid2D.bin", x2a, x2b, ny), ("grid3D.bin", x3a, x3b, nz)):
        g = np.stack([b[NGH:NGH + n + 1], a[NGH:NGH + n + 1]]).astype("<f8")
        p = os.path.join(bdir, name)
        g.tofile(p)
        written.append(p)
    p = os.path.join(bdir, f"field{nout:05d}.bin")
    fields.astype("<f8").tofile(p)
    written.append(p)
    written.append(write_xdmf(bdir, nout, time, nx, ny, nz, nvars))
    if ascii_out:
        written.append(write_ascii(outdir, nout, time, dt, P, x1b, x2b, x3b))
    return written


def write_xdmf(bdir: str, nout: int, time: float, nx: int, ny: int, nz: int, nvars: int) -> str:
    fdata = f"field{nout:05d}.bin"
    per_field = nx * ny * nz * 8
    lines = ['<?xml version="1.0" ?>', '<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>', '<Xdmf Version="2.0">', "  <Domain>",
             '    <Grid Name="Grid" GridType="Uniform">', f'      <Time Value="{time:.16e}"/>',
             f'      <Topology TopologyType="3DRectMesh" Dimensions="{nz + 1} {ny + 1} {nx + 1}"/>',
             '      <Geometry GeometryType="VXVYVZ">']
    for fname, n in (("grid1D.bin", nx + 1), ("grid2D.bin", ny + 1), ("grid3D.bin", nz + 1)):
        lines.append(f'        <DataItem Dimensions="{n}" NumberType="Float" Precision="8" Format="Binary" '
                     f'Endian="Little" Seek="{n * 8}">{fname}</DataItem>')
    lines.append("      </Geometry>")
    for k in range(nvars):
        lines += [f'      <Attribute Name="{VARNAMES[k]}" AttributeType="Scalar" Center="Cell">',
                  f'        <DataItem Dimensions="{nz} {ny} {nx}" NumberType="Float" Precision="8" Format="Binary" '
                  f'Endian="Little" Seek="{k * per_field}">{fdata}</DataItem>', "      </Attribute>"]
    lines += ["    </Grid>", "  </Domain>", "</Xdmf>"]
    p = os.path.join(bdir, f"field{nout:05d}.xmf")
    with open(p, "w") as f:
        f.write("\n".join(lines) + "\n")
    return p


def write_ascii(outdir: str, nout: int, time: float, dt: float, P: np.ndarray, x1b, x2b, x3b) -> str:
    adir = os.path.join(outdir, "ascdata")
    os.makedirs(adir, exist_ok=True)
    nx, ny = P.shape[3] - 5, P.shape[1] - 5
    k = NGH
    p = os.path.join(adir, f"snap000-{nout:05d}.csv")
    with open(p, "w") as f:
        f.write(f"# {time:.6e} {dt:.6e}\n# {nx}\n# {ny}\n# {k} {x3b[k]:.6e}\n# x y d vx vy p X1\n")
        for j in range(NGH, NGH + ny):
            for i in range(NGH, NGH + nx):
                q = P[k, j, :, i]
                f.write(f"{x1b[i]:.6e} {x2b[j]:.6e} {q[0]:.6e} {q[1]:.6e} {q[2]:.6e} {q[5]:.6e} {q[11]:.6e}\n")
            f.write("\n")
    return p


def read_field(path: str, nx: int, ny: int, nz: int, nvars: int = 10) -> np.ndarray:
    return np.fromfile(path, dtype="<f8").reshape(nvars, nz, ny, nx)
