"""Pins of the oracle's mesh cache (edges, edge coefficients, control volumes).

Each check is independent of the oracle's own formulas: brute-force edge
enumeration, closed forms of the structured meshes, and the defining moment
identity of the edge coefficients.
"""
import numpy as np
import pytest

from nosh_amd import meshgen as mg
from oracle import oracle as O

TRI_EDGES = [(0, 1), (1, 2), (2, 0)]
TET_EDGES = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]


def brute_edges(cells):
    """First-appearance edge numbering, written as plainly as possible."""
    local = TRI_EDGES if cells.shape[1] == 3 else TET_EDGES
    ids, edges, cell_edges = {}, [], []
    for c in cells:
        ce = []
        for a, b in local:
            key = (min(c[a], c[b]), max(c[a], c[b]))
            if key not in ids:
                ids[key] = len(edges)
                edges.append(key)
            ce.append(ids[key])
        cell_edges.append(sorted(ce))
    return np.array(edges), np.array(cell_edges)


@pytest.mark.parametrize("gen", [lambda: mg.square_tri(5), lambda: mg.kuhn_cube(3),
                                 lambda: mg.kuhn_cube(3, jitter=0.2, seed=7)])
def test_edge_numbering_matches_brute_force(gen):
    coords, cells = gen()
    m = O.Mesh(coords, cells)
    e, ce = brute_edges(cells)
    np.testing.assert_array_equal(m.edges(), e)
    np.testing.assert_array_equal(m.cell_edges(), ce)
    assert np.all(m.edges()[:, 0] < m.edges()[:, 1])


@pytest.mark.parametrize("nc,E", [(64, 12416), (16, 800)])
def test_square_counts(nc, E):
    coords, cells = mg.square_tri(nc)
    m = O.Mesh(coords, cells)
    assert m.n == (nc + 1) ** 2 and cells.shape[0] == 2 * nc * nc and m.E == E


@pytest.mark.parametrize("nc,E", [(16, 31024), (24, 102024), (32, 238688)])
def test_kuhn_counts(nc, E):
    coords, cells = mg.kuhn_cube(nc)
    assert mg.count_edges(cells) == E


def test_square_closed_forms():
    """Right-triangle grid: alpha = 1 (interior axis edge), 1/2 (boundary axis
    edge), 0 (diagonal); control volume h^2, h^2/2 on edges, h^2/4 at corners."""
    nc = 64
    h = 10.0 / nc
    coords, cells = mg.square_tri(nc)
    m = O.Mesh(coords, cells)
    ln, cov = m.edge_data()
    alpha = cov / ln
    e = m.edges()
    d = coords[e[:, 1]] - coords[e[:, 0]]
    diag = (np.abs(d[:, 0]) > 1e-12) & (np.abs(d[:, 1]) > 1e-12)
    on_bnd = np.zeros(len(e), bool)
    for ax in (0, 1):
        for val in (-5.0, 5.0):
            on_bnd |= (np.abs(coords[e[:, 0], ax] - val) < 1e-12) & \
                      (np.abs(coords[e[:, 1], ax] - val) < 1e-12)
    expect = np.where(diag, 0.0, np.where(on_bnd, 0.5, 1.0))
    np.testing.assert_allclose(alpha, expect, rtol=0, atol=1e-13)
    cv = m.control_volumes()
    bx = (np.abs(np.abs(coords[:, 0]) - 5) < 1e-12).astype(int)
    by = (np.abs(np.abs(coords[:, 1]) - 5) < 1e-12).astype(int)
    expect_cv = h * h / 2.0 ** (bx + by)
    np.testing.assert_allclose(cv, expect_cv, rtol=1e-13)
    assert abs(cv.sum() - 100.0) < 1e-11


def test_kuhn_closed_forms():
    """Kuhn cube: per tet c = h/6 on the three axis edges and 0 on the face
    and body diagonals, so alpha_axis = (#tets containing the edge) * h/6 and
    alpha_diag = 0; control volumes h^3 / 2^(#boundary coordinates)."""
    nc = 6
    h = 10.0 / nc
    coords, cells = mg.kuhn_cube(nc)
    m = O.Mesh(coords, cells)
    ce = m.cell_edges()
    coef = m.cell_edge_coefficients()
    e = m.edges()
    d = np.abs(coords[e[:, 1]] - coords[e[:, 0]])
    axis = (d > 1e-12).sum(axis=1) == 1
    np.testing.assert_allclose(coef, np.where(axis[ce], h / 6.0, 0.0), rtol=0, atol=1e-12)
    count = np.bincount(ce.ravel(), minlength=m.E)
    ln, cov = m.edge_data()
    np.testing.assert_allclose(cov / ln, np.where(axis, count * h / 6.0, 0.0), rtol=0, atol=1e-12)
    # Interior axis edges lie in 6 tets: alpha = h.
    assert np.isclose(np.max(cov / ln), h)
    cv = m.control_volumes()
    nb = sum((np.abs(np.abs(coords[:, ax]) - 5) < 1e-12).astype(int) for ax in range(3))
    np.testing.assert_allclose(cv, h ** 3 / 2.0 ** nb, rtol=1e-12)
    assert abs(cv.sum() - 1000.0) < 1e-9


@pytest.mark.parametrize("gen", [lambda: mg.kuhn_cube(4, jitter=0.2, seed=2),
                                 lambda: mg.kuhn_cube(5)])
def test_edge_coefficient_moment_identity(gen):
    """sum_e c_{T,e} e e^T = |T| I for every tetrahedron (the defining equation
    of src/mesh_tetra.cpp:105-160, including its coplanar-edge volume fallback,
    which cell 0 exercises: its edges 0,1,2 form a face)."""
    coords, cells = gen()
    m = O.Mesh(coords, cells)
    ce = m.cell_edges()
    coef = m.cell_edge_coefficients()
    e = m.edges()
    vec = coords[e[:, 0]] - coords[e[:, 1]]
    x = coords[cells]
    vol = np.abs(np.einsum("ij,ij->i", x[:, 1] - x[:, 0],
                           np.cross(x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]))) / 6.0
    M = np.einsum("tk,tki,tkj->tij", coef, vec[ce], vec[ce])
    np.testing.assert_allclose(M, vol[:, None, None] * np.eye(3), rtol=0,
                               atol=1e-12 * np.max(vol))


def test_triangle_moment_identity_jittered():
    coords, cells = mg.square_tri(6)
    rng = np.random.default_rng(3)
    inner = (np.abs(np.abs(coords[:, 0]) - 5) > 1e-9) & (np.abs(np.abs(coords[:, 1]) - 5) > 1e-9)
    coords[inner, :2] += rng.uniform(-0.2, 0.2, (inner.sum(), 2)) * (10 / 6)
    m = O.Mesh(coords, cells)
    ce = m.cell_edges()
    coef = m.cell_edge_coefficients()
    e = m.edges()
    vec = (coords[e[:, 0]] - coords[e[:, 1]])[:, :2]
    x = coords[cells][:, :, :2]
    u0, w0 = x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]
    area = 0.5 * np.abs(u0[:, 0] * w0[:, 1] - u0[:, 1] * w0[:, 0])
    M = np.einsum("tk,tki,tkj->tij", coef, vec[ce], vec[ce])
    np.testing.assert_allclose(M, area[:, None, None] * np.eye(2), rtol=0, atol=1e-13)
    # 2D: c = cot(opposite angle) / 2.
    for t in range(cells.shape[0]):
        for k, ed in enumerate(ce[t]):
            a, b = e[ed]
            o = [v for v in cells[t] if v not in (a, b)][0]
            u, w = coords[a, :2] - coords[o, :2], coords[b, :2] - coords[o, :2]
            cot = np.dot(u, w) / abs(u[0] * w[1] - u[1] * w[0])
            assert abs(coef[t, k] - 0.5 * cot) < 1e-12


def test_jittered_control_volumes_sum_to_volume():
    coords, cells = mg.kuhn_cube(8, jitter=0.2, seed=2)
    m = O.Mesh(coords, cells)
    cv = m.control_volumes()
    assert abs(cv.sum() - 1000.0) < 1e-9
    # Per-tet circumcentric pyramids sum to the tet volume: checked through the
    # total, and positivity of the dual cells of this mildly jittered mesh.
    assert np.all(cv > 0)


def test_disk_delaunay_mesh():
    """The disk generator's triangulation (jittered hexagonal lattice + boundary
    ring, Delaunay): a topological disk (V - E + F = 1) covering the inscribed
    polygon; its edge coefficients satisfy the moment identity; Delaunay makes
    every interior edge weight alpha = sum of (cot of opposite angle)/2 >= 0; the
    control volumes sum to the polygon's area; numbering matches brute force."""
    coords, cells = mg.disk(0.5)
    n, T = coords.shape[0], cells.shape[0]
    E = mg.count_edges(cells)
    assert n - E + T == 1
    m = O.Mesh(coords, cells)
    e, ce = brute_edges(cells)
    np.testing.assert_array_equal(m.edges(), e)
    x = coords[cells][:, :, :2]
    u0, w0 = x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]
    area = 0.5 * np.abs(u0[:, 0] * w0[:, 1] - u0[:, 1] * w0[:, 0])
    nb = int(round(2 * np.pi * 5.0 / 0.5))
    poly = 0.5 * nb * 25.0 * np.sin(2 * np.pi / nb)
    assert abs(area.sum() - poly) < 1e-10
    coef = m.cell_edge_coefficients()
    vec = (coords[e[:, 0]] - coords[e[:, 1]])[:, :2]
    M = np.einsum("tk,tki,tkj->tij", coef, vec[ce], vec[ce])
    np.testing.assert_allclose(M, area[:, None, None] * np.eye(2), rtol=0, atol=1e-12)
    alpha = np.zeros(E)
    np.add.at(alpha, ce.ravel(), coef.ravel())
    boundary = np.zeros(E, bool)
    cnt = np.bincount(ce.ravel(), minlength=E)
    boundary[cnt == 1] = True
    assert np.all(alpha[~boundary] >= -1e-12)
    assert abs(m.control_volumes().sum() - poly) < 1e-10
