"""Exodus II mesh input (the format of the reference's test meshes, read there
through MOAB, src/mesh_reader.cpp:19-162): round trips through the writer,
a file in the older conventions written field by field with scipy's netCDF
writer, the edge numbering of a mesh read back, and the errors."""
import numpy as np
import pytest

from nosh_amd import io as nio
from nosh_amd import meshgen as mg


def test_round_trip_tetrahedra_with_vector_and_scalar_fields(tmp_path):
    c, t = mg.kuhn_cube(3, jitter=0.2, seed=5)
    A = mg.magnetic_potential(c)
    V = np.linspace(-1.0, 1.0, c.shape[0])
    path = str(tmp_path / "cube.e")
    nio.write_exodus(path, c, t, {"A": A, "V": V}, time=0.25)
    c2, t2, pd = nio.read_mesh(path)
    assert np.array_equal(c2, c)
    assert t2.dtype == np.int32 and np.array_equal(t2, t)
    assert set(pd) == {"A", "V"}
    assert np.array_equal(pd["A"], A) and np.array_equal(pd["V"], V)


def test_round_trip_triangles_in_two_dimensions(tmp_path):
    c, t = mg.square_tri(8)
    path = str(tmp_path / "square.exo")
    nio.write_exodus(path, c, t, {"psi": np.stack([c[:, 0], -c[:, 1]], axis=1)})
    from scipy.io import netcdf_file
    with netcdf_file(path, "r", mmap=False) as f:
        assert int(f.dimensions["num_dim"]) == 2
        assert nio._exo_text(f.variables["connect1"].elem_type) == "TRI3"
        assert int(f.variables["connect1"][:].min()) == 1  # 1-based in the file
    c2, t2, pd = nio.read_exodus(path)
    assert np.array_equal(c2[:, :2], c[:, :2]) and not np.any(c2[:, 2])
    assert np.array_equal(t2, t)
    assert np.array_equal(pd["psi"], np.stack([c[:, 0], -c[:, 1]], axis=1))


def test_older_conventions_two_blocks_two_time_steps(tmp_path):
    """One "coord" array, one 3-index "vals_nod_var", two element blocks and
    two time steps, written with scipy's netCDF writer (CDF-1)."""
    from scipy.io import netcdf_file
    c, t = mg.kuhn_cube(2)
    n = c.shape[0]
    half = t.shape[0] // 2
    path = str(tmp_path / "old.e")
    names = ["A_X", "A_Y", "A_Z", "thickness"]
    with netcdf_file(path, "w", version=1) as f:
        f.createDimension("time_step", None)
        f.createDimension("num_nodes", n)
        f.createDimension("num_dim", 3)
        f.createDimension("num_el_blk", 2)
        f.createDimension("len_name", 33)
        f.createDimension("num_nod_var", len(names))
        f.createVariable("coord", "d", ("num_dim", "num_nodes"))[:] = c.T
        for b, part in ((1, t[:half]), (2, t[half:])):
            f.createDimension(f"num_el_in_blk{b}", part.shape[0])
            f.createDimension(f"num_nod_per_el{b}", 4)
            v = f.createVariable(f"connect{b}", "i", (f"num_el_in_blk{b}", f"num_nod_per_el{b}"))
            v[:] = part + 1
            v.elem_type = "TETRA"
        nv = f.createVariable("name_nod_var", "c", ("num_nod_var", "len_name"))
        for i, nm in enumerate(names):
            nv[i, :] = np.array(list(nm.ljust(33, "\0")), dtype="S1")
        vals = f.createVariable("vals_nod_var", "d", ("time_step", "num_nod_var", "num_nodes"))
        for s in range(2):
            vals[s] = np.stack([c[:, 0] + s, c[:, 1] + s, c[:, 2] + s, np.full(n, 1.0 + s)])
    c2, t2, pd = nio.read_exodus(path)
    assert np.array_equal(c2, c) and np.array_equal(t2, t)
    assert np.array_equal(pd["A"], c + 1.0) and np.array_equal(pd["thickness"], np.full(n, 2.0))
    _, _, pd0 = nio.read_exodus(path, step=0)
    assert np.array_equal(pd0["A"], c)


def test_edges_of_a_mesh_read_back_are_numbered_as_the_original():
    """Cells keep their order, so the oracle's first-appearance edge
    numbering of the read mesh is the original's."""
    import tempfile, os
    from oracle import oracle as O
    c, t = mg.kuhn_cube(3, jitter=0.1, seed=1)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "m.e")
        nio.write_exodus(path, c, t)
        c2, t2, _ = nio.read_exodus(path)
    e1, e2 = O.Mesh(c, t), O.Mesh(c2, t2)
    assert np.array_equal(e1.edges(), e2.edges())
    assert np.array_equal(e1.control_volumes(), e2.control_volumes())


def test_unsupported_blocks_are_refused(tmp_path):
    from scipy.io import netcdf_file
    path = str(tmp_path / "hex.e")
    with netcdf_file(path, "w", version=2) as f:
        f.createDimension("num_nodes", 8)
        f.createDimension("num_dim", 3)
        f.createDimension("num_el_blk", 1)
        f.createDimension("num_el_in_blk1", 1)
        f.createDimension("num_nod_per_el1", 8)
        for nm in ("coordx", "coordy", "coordz"):
            f.createVariable(nm, "d", ("num_nodes",))[:] = np.arange(8.0)
        v = f.createVariable("connect1", "i", ("num_el_in_blk1", "num_nod_per_el1"))
        v[:] = np.arange(1, 9)
        v.elem_type = "HEX8"
    with pytest.raises(ValueError, match="HEX8"):
        nio.read_exodus(path)
    with pytest.raises(ValueError):
        nio.write_exodus(str(tmp_path / "bad.e"), np.zeros((4, 3)), np.zeros((1, 5), int))
    with pytest.raises(ValueError, match="h5m"):
        nio.read_mesh("mesh.h5m")
