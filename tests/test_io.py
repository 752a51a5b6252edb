"""Host I/O: the observer / eigen-data CSV layout of the reference's
csv_writer, .vtu and legacy .vtk meshes and states, checkpoints. No GPU."""
import os

import numpy as np
import pytest

from nosh_amd import io as nio
from nosh_amd import meshgen as mg


def test_csv_layout_matches_reference_writer(tmp_path):
    """Header '# ' + labels, columns joined by delimiter + 2 spaces, doubles
    '%.15e' left-aligned in 22, ints in max(5, len(label)), rows indented by
    2 (src/csv_writer.cpp:22-141); literal strings worked out by hand."""
    w = nio.CsvWriter("")
    items = [("(0) step", 0), ("(1) mu", 0.5), ("(2) Gibbs energy", -1.0), ("n", 7)]
    assert w.header_line(items) == ("# (0) step   (1) mu                   "
                                    "(2) Gibbs energy         n    ")
    assert w.row_line(items) == ("  0          5.000000000000000e-01    "
                                 "-1.000000000000000e+00   7    ")
    with pytest.raises(TypeError):
        w.header_line([("flag", True)])


def test_observer_and_eigen_saver_files(tmp_path):
    f = str(tmp_path / "cont.dat")
    obs = nio.ContinuationObserver(f, "mu")
    for k in range(3):
        obs.observe(param=0.1 * k, gibbs=-1.0 + 0.01 * k, norm=1.0 - 0.001 * k)
    obs.close()
    labels, rows = nio.read_csv(f)
    assert labels == ["(0) step", "(1) mu", "(2) Gibbs energy", "(2) ||x||_2 scaled"]
    assert np.allclose(rows, [[0, 0.0, -1.0, 1.0], [1, 0.1, -0.99, 0.999], [2, 0.2, -0.98, 0.998]],
                       rtol=1e-15)
    e = nio.EigenDataSaver(str(tmp_path / "eig.dat"))
    assert e.classify([-1e-3, 0.0, 5e-6, 2e-5, -1e-5]) == (1, 3, 1)
    e.save([-1e-3, 0.0, 2e-5])
    e.save([1e-3, 2e-3, 3e-3])
    e.close()
    labels, rows = nio.read_csv(str(tmp_path / "eig.dat"))
    assert labels == ["##step", "#0unstable", "#1null", "#2stable", "00-0Re()", "01-0Re()",
                      "02-0Re()"]
    assert rows[0].tolist() == [0, 1, 1, 1, -1e-3, 0.0, 2e-5]
    assert rows[1].tolist() == [1, 3, 0, 0, 1e-3, 2e-3, 3e-3]


@pytest.mark.parametrize("kind", ["tri", "tet"])
def test_vtu_roundtrip_is_bit_exact(tmp_path, kind):
    c, t = mg.square_tri(6) if kind == "tri" else mg.kuhn_cube(3, jitter=0.1, seed=4)
    n = c.shape[0]
    psi = mg.random_state(n, 0)
    A = mg.magnetic_potential(c)
    f = str(tmp_path / "m.vtu")
    nio.write_vtu(f, c, t, {"psi": nio.psi_point_data(psi), "A": A}, {"mu": 0.125, "g": 1.5})
    c2, t2, pd, fd = nio.read_vtu(f)
    assert np.array_equal(c2[:, :c.shape[1]], c) and np.array_equal(t2, t)
    assert np.array_equal(pd["psi"].reshape(-1), psi) and np.array_equal(pd["A"], A)
    assert fd == {"mu": 0.125, "g": 1.5}


def test_vtu_ascii_arrays_are_read(tmp_path):
    f = tmp_path / "a.vtu"
    f.write_text('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="0.1" '
                 'byte_order="LittleEndian">\n<UnstructuredGrid>\n'
                 '<Piece NumberOfPoints="4" NumberOfCells="2">\n<Points>\n'
                 '<DataArray type="Float64" NumberOfComponents="3" format="ascii">'
                 '0 0 0 1 0 0 1 1 0 0 1 0</DataArray>\n</Points>\n<Cells>\n'
                 '<DataArray type="Int32" Name="connectivity" format="ascii">0 1 2 0 2 3'
                 '</DataArray>\n<DataArray type="Int32" Name="offsets" format="ascii">3 6'
                 '</DataArray>\n<DataArray type="UInt8" Name="types" format="ascii">5 5'
                 '</DataArray>\n</Cells>\n<PointData>\n<DataArray type="Float64" Name="V" '
                 'format="ascii">1 2 3 4</DataArray>\n</PointData>\n</Piece>\n'
                 '</UnstructuredGrid>\n</VTKFile>\n')
    c, t, pd, fd = nio.read_vtu(str(f))
    assert c.shape == (4, 3) and t.tolist() == [[0, 1, 2], [0, 2, 3]]
    assert pd["V"].tolist() == [1, 2, 3, 4] and fd == {}


def test_legacy_vtk_reader(tmp_path):
    f = tmp_path / "m.vtk"
    f.write_text("# vtk DataFile Version 3.0\nmesh\nASCII\nDATASET UNSTRUCTURED_GRID\n"
                 "POINTS 4 double\n0 0 0\n1 0 0\n0 1 0\n0 0 1\nCELLS 1 5\n4 0 1 2 3\n"
                 "CELL_TYPES 1\n10\nPOINT_DATA 4\nSCALARS thickness double\nLOOKUP_TABLE default\n"
                 "1 1 1 1\nSCALARS psi double 2\nLOOKUP_TABLE default\n1 0 0.5 0.5 0 1 1 1\n"
                 "VECTORS A double\n0 0 0 0 0.5 0 -0.5 0 0 0 0 0\n")
    c, t, pd = nio.read_mesh(str(f))
    assert c.shape == (4, 3) and t.tolist() == [[0, 1, 2, 3]]
    assert pd["thickness"].tolist() == [1, 1, 1, 1]
    assert pd["psi"].shape == (4, 2) and pd["psi"][1].tolist() == [0.5, 0.5]
    assert pd["A"][2].tolist() == [-0.5, 0, 0]
    with pytest.raises(ValueError):
        nio.read_mesh(str(tmp_path / "m.h5m"))


def test_mesh_read_back_keeps_edge_numbering(tmp_path):
    """Cells keep their file order, so the edge numbering (first appearance)
    of the re-read mesh equals the original's (oracle mesh cache)."""
    from oracle import oracle as O
    c, t = mg.kuhn_cube(3, jitter=0.2, seed=1)
    f = str(tmp_path / "cube.vtu")
    nio.write_vtu(f, c, t)
    c2, t2, _ = nio.read_mesh(f)
    a, b = O.Mesh(c, t), O.Mesh(c2, t2)
    assert np.array_equal(a.edges(), b.edges())
    assert np.array_equal(a.cell_edges(), b.cell_edges())
    assert np.array_equal(a.control_volumes(), b.control_volumes())


def test_data_saver_and_checkpoint(tmp_path):
    c, t = mg.square_tri(4)
    n = c.shape[0]
    saver = nio.ContinuationDataSaver(c, t, str(tmp_path), "mu")
    states = [mg.random_state(n, s) for s in range(3)]
    paths = [saver.save_solution(p, 0.1 * k) for k, p in enumerate(states)]
    assert [os.path.basename(p) for p in paths] == ["out0000.vtu", "out0001.vtu", "out0002.vtu"]
    for k, p in enumerate(paths):
        psi, fd = nio.read_state(p)
        assert np.array_equal(psi, states[k]) and fd["mu"] == 0.1 * k
    f = str(tmp_path / "ck.npz")
    nio.save_checkpoint(f, c, t, states[1], {"mu": 0.3, "g": 1.0}, step=7, tangent=states[2])
    c2, t2, p2, params, step, extra = nio.load_checkpoint(f)
    assert np.array_equal(c2, c) and np.array_equal(t2, t) and np.array_equal(p2, states[1])
    assert params == {"g": 1.0, "mu": 0.3} and step == 7
    assert np.array_equal(extra["tangent"], states[2])
