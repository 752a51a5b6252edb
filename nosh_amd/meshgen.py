"""Synthetic meshes and seeded inputs for the NLS workloads.

Everything here is input synthesis, not the solver path: it produces the
vertex coordinates, cell connectivity ("file order"), the magnetic vector
potential tag "A" and the seeded random states psi / phi that a caller would
otherwise read from a mesh file (reference: ``nosh::read``,
src/mesh_reader.cpp:19-162, and the "A"/"psi" tags read in
test/keo.cpp:31-35).

Random numbers come from a SplitMix64 stream so that every consumer (GPU
library, CPU oracle, benchmarks) regenerates identical inputs:
``state += 0x9E3779B97F4A7C15``, the standard SplitMix64 finaliser, and a
uniform double ``(z >> 11) * 2**-53``.
"""
from __future__ import annotations

import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64_uniform(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Return ``count`` uniforms in [0,1) from the SplitMix64 stream of ``seed``.

    Draw number k (0-based, after skipping ``offset`` draws) uses the state
    ``seed + (offset+k+1)*golden`` (mod 2**64).
    """
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def random_state(n: int, seed: int) -> np.ndarray:
    """Interleaved complex state of n vertices: |psi|~U(0.5,1), arg~U(0,2pi).

    Two draws per vertex in vertex order: first the modulus, then the phase.
    """
    u = splitmix64_uniform(seed, 2 * n).reshape(n, 2)
    r = 0.5 + 0.5 * u[:, 0]
    th = 2.0 * math.pi * u[:, 1]
    out = np.empty(2 * n)
    out[0::2] = r * np.cos(th)
    out[1::2] = r * np.sin(th)
    return out


def square_tri(ncell: int, lo: float = -5.0, hi: float = 5.0):
    """Structured square [lo,hi]^2 with ncell x ncell cells, two right triangles each.

    Vertex id = j*(ncell+1) + i at (lo + i*h, lo + j*h, 0). Cell (i,j), row-major
    in j, is split by its (i,j)-(i+1,j+1) diagonal into the counter-clockwise
    triangles (v00, v10, v11) and (v00, v11, v01).
    """
    m = ncell + 1
    h = (hi - lo) / ncell
    jj, ii = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    coords = np.zeros((m * m, 3))
    coords[:, 0] = lo + ii.ravel() * h
    coords[:, 1] = lo + jj.ravel() * h
    cj, ci = np.meshgrid(np.arange(ncell), np.arange(ncell), indexing="ij")
    v00 = (cj * m + ci).ravel()
    v10 = v00 + 1
    v01 = v00 + m
    v11 = v01 + 1
    cells = np.empty((2 * ncell * ncell, 3), dtype=np.int32)
    cells[0::2] = np.stack([v00, v10, v11], axis=1)
    cells[1::2] = np.stack([v00, v11, v01], axis=1)
    return coords, cells


# The six Kuhn tetrahedra of the unit cube along its main diagonal 000->111,
# one per axis permutation: path 000 -> e_a -> e_a+e_b -> 111.
_KUHN_PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


def kuhn_cube(ncell: int, lo: float = -5.0, hi: float = 5.0, jitter: float = 0.0,
              seed: int = 2):
    """Cube [lo,hi]^3 with ncell^3 cells, each split into 6 Kuhn tetrahedra.

    Vertex id = (k*(ncell+1) + j)*(ncell+1) + i. Cells are emitted cube by
    cube (k, then j, then i slowest to fastest), six tets per cube in the axis
    permutation order (x,y,z), (x,z,y), (y,x,z), (y,z,x), (z,x,y), (z,y,x),
    each as (v000, v_a, v_ab, v111).

    With ``jitter > 0`` every interior vertex is moved by U(-jitter*h,
    jitter*h) per coordinate, three draws per interior vertex in vertex order
    from the SplitMix64 stream of ``seed``.
    """
    m = ncell + 1
    h = (hi - lo) / ncell
    kk, jj, ii = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    coords = np.empty((m ** 3, 3))
    coords[:, 0] = lo + ii.ravel() * h
    coords[:, 1] = lo + jj.ravel() * h
    coords[:, 2] = lo + kk.ravel() * h
    if jitter > 0.0:
        interior = ((ii > 0) & (ii < ncell) & (jj > 0) & (jj < ncell)
                    & (kk > 0) & (kk < ncell)).ravel()
        nint = int(interior.sum())
        u = splitmix64_uniform(seed, 3 * nint).reshape(nint, 3)
        coords[interior] += (2.0 * u - 1.0) * jitter * h
    ck, cj, ci = np.meshgrid(np.arange(ncell), np.arange(ncell), np.arange(ncell),
                             indexing="ij")
    base = ((ck * m + cj) * m + ci).ravel().astype(np.int64)
    stride = (1, m, m * m)
    v111 = base + stride[0] + stride[1] + stride[2]
    cells = np.empty((6 * base.size, 4), dtype=np.int32)
    for t, (a, b, _c) in enumerate(_KUHN_PERMS):
        va = base + stride[a]
        vab = va + stride[b]
        cells[t::6] = np.stack([base, va, vab, v111], axis=1)
    return coords, cells


def disk(h: float, radius: float = 5.0, jitter: float = 0.2, seed: int = 3):
    """Delaunay triangulation of a disk: jittered hexagonal lattice plus a boundary ring.

    Lattice points (spacing h, rows h*sqrt(3)/2 apart) strictly inside radius
    - h/2 are moved by U(-jitter*h, jitter*h) per coordinate from the
    SplitMix64 stream of ``seed`` (two draws per point, in lattice order);
    round(2 pi radius / h) points are placed evenly on the boundary circle;
    scipy's Delaunay triangulation (Qhull) joins them. Vertex order: lattice
    points row by row, then the boundary ring. h = 0.0213 gives about 200 k
    vertices at radius 5.
    """
    from scipy.spatial import Delaunay
    dy = h * np.sqrt(3.0) / 2.0
    nrow = int(np.ceil(radius / dy)) + 1
    pts = []
    for r in range(-nrow, nrow + 1):
        y = r * dy
        x0 = 0.5 * h * (r & 1)
        ncol = int(np.ceil(radius / h)) + 1
        xs = x0 + h * np.arange(-ncol, ncol + 1)
        keep = xs * xs + y * y < (radius - 0.5 * h) ** 2
        pts.append(np.stack([xs[keep], np.full(keep.sum(), y)], axis=1))
    lat = np.concatenate(pts)
    u = splitmix64_uniform(seed, 2 * lat.shape[0]).reshape(-1, 2) * 2.0 - 1.0
    lat = lat + jitter * h * u
    nb = int(round(2.0 * np.pi * radius / h))
    th = 2.0 * np.pi * np.arange(nb) / nb
    ring = np.stack([radius * np.cos(th), radius * np.sin(th)], axis=1)
    xy = np.concatenate([lat, ring])
    tri = Delaunay(xy)
    cells = np.ascontiguousarray(tri.simplices, dtype=np.int32)
    coords = np.zeros((xy.shape[0], 3))
    coords[:, :2] = xy
    return coords, cells


def magnetic_potential(coords: np.ndarray) -> np.ndarray:
    """The per-vertex tag A = 0.5*(-y, x, 0): constant curl B = e_z."""
    A = np.zeros_like(coords)
    A[:, 0] = -0.5 * coords[:, 1]
    A[:, 1] = 0.5 * coords[:, 0]
    return A


def count_edges(cells: np.ndarray) -> int:
    """Number of distinct undirected edges of a triangle or tetrahedron mesh."""
    nv = cells.shape[1]
    pairs = [(0, 1), (1, 2), (2, 0)] if nv == 3 else \
        [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
    c = cells.astype(np.int64)
    n = int(c.max()) + 1
    keys = []
    for a, b in pairs:
        lo = np.minimum(c[:, a], c[:, b])
        hi = np.maximum(c[:, a], c[:, b])
        keys.append(lo * n + hi)
    return int(np.unique(np.concatenate(keys)).size)
