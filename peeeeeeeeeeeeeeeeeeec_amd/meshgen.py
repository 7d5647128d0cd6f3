"""Synthetic quad surface meshes for the PEEC workloads.

Conventions (they fix every branch's index and sign through the
first-appearance edge rule of the reference's topology.c:160-304):

* plate: nodes row-major, index i + (nx+1)*j; quads counter-clockwise
  (n[i,j], n[i+1,j], n[i+1,j+1], n[i,j+1]) seen from +z.
* box:   the surface nodes of an (nx+1)(ny+1)(nz+1) lattice numbered in
  lexicographic (k, j, i) order; quads face by face (z-, z+, y-, y+, x-, x+),
  each oriented with its normal pointing out of the box.
* cylinder: an O-grid; each flat cap is an m x m square lattice mapped onto
  the disc, the side is 4m x nz quads sharing the cap rim nodes.
"""
from __future__ import annotations

import numpy as np

MASK64 = (1 << 64) - 1


def plate(nx: int, ny: int, lx: float = 1.0, ly: float = 1.0):
    """Flat plate [-lx/2, lx/2] x [-ly/2, ly/2] x {0}."""
    i = np.arange(nx + 1)
    j = np.arange(ny + 1)
    X = -0.5 * lx + lx * i / nx
    Y = -0.5 * ly + ly * j / ny
    xyz = np.zeros(((nx + 1) * (ny + 1), 3))
    for jj in range(ny + 1):
        xyz[jj * (nx + 1):(jj + 1) * (nx + 1), 0] = X
        xyz[jj * (nx + 1):(jj + 1) * (nx + 1), 1] = Y[jj]
    quads = []
    for jj in range(ny):
        for ii in range(nx):
            n0 = ii + (nx + 1) * jj
            quads.append((n0, n0 + 1, n0 + 1 + (nx + 1), n0 + (nx + 1)))
    return xyz, np.array(quads, dtype=np.int32)


def _orient(quads, xyz, outward):
    out = []
    for q in quads:
        p = xyz[list(q)]
        n = np.cross(p[1] - p[0], p[2] - p[0]) + np.cross(p[2] - p[0], p[3] - p[0])
        out.append(q if np.dot(n, outward) > 0 else (q[0], q[3], q[2], q[1]))
    return out


def box(lx: float, ly: float, lz: float, nx: int, ny: int, nz: int):
    """Closed box [0,lx] x [0,ly] x [0,lz], outward-oriented surface quads."""
    index = -np.ones((nz + 1, ny + 1, nx + 1), dtype=np.int64)
    pts = []
    for k in range(nz + 1):
        for j in range(ny + 1):
            for i in range(nx + 1):
                if i in (0, nx) or j in (0, ny) or k in (0, nz):
                    index[k, j, i] = len(pts)
                    pts.append((lx * i / nx, ly * j / ny, lz * k / nz))
    xyz = np.array(pts, dtype=np.float64)
    quads = []
    faces = [
        ("z", 0, (0, 0, -1)), ("z", nz, (0, 0, 1)),
        ("y", 0, (0, -1, 0)), ("y", ny, (0, 1, 0)),
        ("x", 0, (-1, 0, 0)), ("x", nx, (1, 0, 0)),
    ]
    for axis, level, normal in faces:
        face = []
        if axis == "z":
            for j in range(ny):
                for i in range(nx):
                    face.append((index[level, j, i], index[level, j, i + 1],
                                 index[level, j + 1, i + 1], index[level, j + 1, i]))
        elif axis == "y":
            for k in range(nz):
                for i in range(nx):
                    face.append((index[k, level, i], index[k, level, i + 1],
                                 index[k + 1, level, i + 1], index[k + 1, level, i]))
        else:
            for k in range(nz):
                for j in range(ny):
                    face.append((index[k, j, level], index[k, j + 1, level],
                                 index[k + 1, j + 1, level], index[k + 1, j, level]))
        quads.extend(_orient(face, xyz, np.array(normal, dtype=float)))
    return xyz, np.array(quads, dtype=np.int32)


def box_by_size(lx, ly, lz, h):
    """Box with cell size close to h (whole divisions per side)."""
    nx, ny, nz = (max(1, int(round(s / h))) for s in (lx, ly, lz))
    return box(lx, ly, lz, nx, ny, nz)


def cylinder(radius: float, length: float, m: int, nz: int):
    """Closed cylinder along z, radius r, caps at z=0 and z=length."""
    # square lattice [-1,1]^2 with m cells per side, mapped to the disc
    t = np.linspace(-1.0, 1.0, m + 1)
    U, V = np.meshgrid(t, t, indexing="xy")
    Xd = U * np.sqrt(1.0 - 0.5 * V * V)
    Yd = V * np.sqrt(1.0 - 0.5 * U * U)
    rim = []  # lattice coordinates of the square boundary, counter-clockwise
    for i in range(m):
        rim.append((0, i))
    for j in range(m):
        rim.append((j, m))
    for i in range(m, 0, -1):
        rim.append((m, i))
    for j in range(m, 0, -1):
        rim.append((j, 0))
    pts = []
    cap_index = [dict(), dict()]
    for c, z in enumerate((0.0, length)):
        for j in range(m + 1):
            for i in range(m + 1):
                cap_index[c][(j, i)] = len(pts)
                pts.append((radius * Xd[j, i], radius * Yd[j, i], z))
    # interior rings of the side
    ring = []
    for kz in range(1, nz):
        z = length * kz / nz
        row = []
        for (j, i) in rim:
            p = pts[cap_index[0][(j, i)]]
            row.append(len(pts))
            pts.append((p[0], p[1], z))
        ring.append(row)
    xyz = np.array(pts)
    quads = []
    for c, normal in ((0, (0, 0, -1)), (1, (0, 0, 1))):
        face = []
        for j in range(m):
            for i in range(m):
                face.append((cap_index[c][(j, i)], cap_index[c][(j, i + 1)],
                             cap_index[c][(j + 1, i + 1)], cap_index[c][(j + 1, i)]))
        quads.extend(_orient(face, xyz, np.array(normal, dtype=float)))
    layers = [[cap_index[0][r] for r in rim]] + ring + [[cap_index[1][r] for r in rim]]
    nphi = len(rim)
    for kz in range(nz):
        a, b = layers[kz], layers[kz + 1]
        face = []
        for p in range(nphi):
            q = (a[p], a[(p + 1) % nphi], b[(p + 1) % nphi], b[p])
            mid = xyz[list(q)].mean(axis=0)
            face.extend(_orient([q], xyz, np.array([mid[0], mid[1], 0.0])))
        quads.extend(face)
    return xyz, np.array(quads, dtype=np.int32)


def splitmix64(state: int):
    """One step of splitmix64: returns (value, new_state)."""
    state = (state + 0x9E3779B97F4A7C15) & MASK64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31), state


class Rng:
    """Counter-based uniform generator (splitmix64 -> 53-bit doubles)."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def uniform(self, lo=0.0, hi=1.0):
        v, self.state = splitmix64(self.state)
        return lo + (hi - lo) * ((v >> 11) * (1.0 / (1 << 53)))


def ragged_batch(count: int = 64, seed: int = 20261203, cells=(300, 3000)):
    """Independent plates/boxes with sides U(0.3,3) m, aspect U(1,4),
    cell counts U(cells), 8 log-uniform frequencies in [1e6, 1e8] each."""
    rng = Rng(seed)
    out = []
    for _ in range(count):
        kind = "plate" if rng.uniform() < 0.5 else "box"
        a = rng.uniform(0.3, 3.0)
        b = a / rng.uniform(1.0, 4.0)
        c = a / rng.uniform(1.0, 4.0)
        target = rng.uniform(*cells)
        area = a * b if kind == "plate" else 2 * (a * b + a * c + b * c)
        h = np.sqrt(area / target)
        if kind == "plate":
            xyz, quads = plate(max(1, int(round(a / h))), max(1, int(round(b / h))), a, b)
        else:
            xyz, quads = box_by_size(a, b, c, h)
        freqs = [10 ** rng.uniform(6.0, 8.0) for _ in range(8)]
        out.append({"kind": kind, "xyz": xyz, "quads": quads, "frequencies": freqs})
    return out
