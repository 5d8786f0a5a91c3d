"""QES-Winds workspace files (NetCDF classic) written and read by libqesmc's
host code, checked against scipy's independent NetCDF classic
implementation (scipy.io.netcdf_file).  CPU only: no device is touched."""
import numpy as np
import pytest
from scipy.io import netcdf_file


def problem(O, Q, seed=3, ncx=23, ncy=17, ncz=9):
    dz = Q.dz_geometric(ncz, 1.0, 1.02)
    blds = Q.random_buildings(seed, ncx, ncy, 2.0, 3, 4.0, 12.0, 3.0, 9.0, margin=1)
    p = O.build_city(ncx, ncy, 2.0, dz, blds, (1, 4.0, 10.0, 0.5, 233.0))
    rng = np.random.default_rng(seed)
    u, v, w = (rng.standard_normal(p.nface).astype(np.float32) for _ in range(3))
    return p, u, v, w


def write(Q, path, p, u, v, w):
    Q.workspace_write(path, p.nx, p.ny, p.nz, p.dx, p.dy, p.dz_array, p.coeffs, p.icellflag, p.u0, p.v0, p.w0,
                      u, v, w)


def test_written_file_reads_in_scipy_with_the_reference_layout(qesmc, oracle, tmp_path):
    p, u, v, w = problem(oracle, qesmc)
    path = str(tmp_path / "ws.nc")
    write(qesmc, path, p, u, v, w)
    with netcdf_file(path, "r", mmap=False) as nc:
        assert nc.version_byte == 2
        dims = nc.dimensions
        assert dims["t"] is None  # the unlimited (record) dimension
        assert (dims["x_face"], dims["y_face"], dims["z_face"]) == (p.nx, p.ny, p.nz)
        assert (dims["x"], dims["y"], dims["z"]) == (p.nx - 1, p.ny - 1, p.nz - 1)
        face = ("t", "z_face", "y_face", "x_face")
        cell = ("t", "z", "y", "x")
        for name, arr in (("u", u), ("v", v), ("w", w), ("u0", p.u0), ("v0", p.v0), ("w0", p.w0)):
            var = nc.variables[name]
            assert var.dimensions == face and var.data.dtype == np.dtype(">f4")
            assert np.array_equal(var.data[0].ravel(), arr), name
            assert var.units == b"m s-1"
        for name, arr in zip("efghmn", p.coeffs):
            assert nc.variables[name].dimensions == cell
            assert np.array_equal(nc.variables[name].data[0].ravel(), arr), name
        flag = nc.variables["icellflag"]
        assert flag.data.dtype == np.dtype(">i4") and np.array_equal(flag.data[0].ravel(), p.icellflag)
        assert np.array_equal(nc.variables["dz_array"].data, p.dz_array)
        assert nc.variables["t"].data.shape == (1,) and nc.variables["t"].data[0] == 0.0
        # coordinates (Domain.cpp:153-166, WINDSOutputWorkspace.cpp:50-88)
        x = nc.variables["x"].data
        assert np.array_equal(x, ((np.arange(p.nx - 1, dtype=np.float32) + np.float32(0.5)) * np.float32(p.dx)))
        assert nc.variables["x_face"].data[0] == x[0] - np.float32(p.dx)
        z_face = nc.variables["z_face"].data
        assert z_face[0] == -p.dz_array[0] and z_face[1] == 0.0
        assert np.allclose(np.diff(z_face[1:]), p.dz_array[1:], rtol=1e-5)


def test_round_trip_through_the_library(qesmc, oracle, tmp_path):
    p, u, v, w = problem(oracle, qesmc, seed=5, ncx=31, ncy=12, ncz=14)
    path = str(tmp_path / "ws.nc")
    write(qesmc, path, p, u, v, w)
    assert qesmc.workspace_dims(path) == (p.nx, p.ny, p.nz)
    r = qesmc.workspace_read(path)
    assert r["dx"] == p.dx and r["dy"] == p.dy
    assert np.array_equal(r["dz_array"], p.dz_array)
    for name, arr in zip("efghmn", p.coeffs):
        assert np.array_equal(r[name], arr), name
    assert np.array_equal(r["icellflag"], p.icellflag)
    for name, arr in (("u0", p.u0), ("v0", p.v0), ("w0", p.w0), ("u", u), ("v", v), ("w", w)):
        assert np.array_equal(r[name], arr), name


def scipy_workspace(path, p, version, with_dz=True, dz_double=False, uniform_dz=None, flag_type="i", flag_fill=None):
    """A workspace file written by scipy: other variable order, extra
    variables and attributes, two time records, optional dz_array;
    icellflag of netCDF type flag_type, optionally with flag_fill stored in
    one cell."""
    with netcdf_file(path, "w", version=version) as nc:
        nc.title = "written by scipy"
        nc.createDimension("t", None)
        nc.createDimension("z", p.nz - 1)
        nc.createDimension("y", p.ny - 1)
        nc.createDimension("x", p.nx - 1)
        nc.createDimension("x_face", p.nx)
        nc.createDimension("y_face", p.ny)
        nc.createDimension("z_face", p.nz)
        t = nc.createVariable("t", "d", ("t",))
        t[:] = [0.0, 60.0]
        x = nc.createVariable("x", "f", ("x",))
        x[:] = (np.arange(p.nx - 1) + 0.5) * p.dx
        y = nc.createVariable("y", "f", ("y",))
        y[:] = (np.arange(p.ny - 1) + 0.5) * p.dy
        z = nc.createVariable("z", "f", ("z",))
        dzu = uniform_dz or 1.0
        z[:] = (np.arange(p.nz - 1) - 0.5) * dzu
        extra = nc.createVariable("mixlength", "f", ("z", "y", "x"))
        extra[:] = 7.0
        extra.units = "m"
        if with_dz:
            d = nc.createVariable("dz_array", "d" if dz_double else "f", ("z",))
            d[:] = p.dz_array
        shape_c = (2, p.nz - 1, p.ny - 1, p.nx - 1)
        shape_f = (2, p.nz, p.ny, p.nx)
        flag = nc.createVariable("icellflag", flag_type, ("t", "z", "y", "x"))
        fl0 = p.icellflag.reshape(shape_c[1:]).astype(flag.data.dtype)
        if flag_fill is not None:
            fl0.flat[fl0.size // 2] = flag_fill
        flag[:] = np.stack([fl0, np.zeros(shape_c[1:], fl0.dtype)])
        for name, arr in zip("nmhgfe", p.coeffs[::-1]):
            var = nc.createVariable(name, "f", ("t", "z", "y", "x"))
            var[:] = np.stack([arr.reshape(shape_c[1:]), np.ones(shape_c[1:], np.float32)])
        for name, arr in (("w0", p.w0), ("u0", p.u0), ("v0", p.v0)):
            var = nc.createVariable(name, "f", ("t", "z_face", "y_face", "x_face"))
            var[:] = np.stack([arr.reshape(shape_f[1:]), np.zeros(shape_f[1:], np.float32)])


@pytest.mark.parametrize("version", [1, 2])
@pytest.mark.parametrize("dz_double", [False, True])
def test_reads_files_written_by_scipy(qesmc, oracle, tmp_path, version, dz_double):
    p, _, _, _ = problem(oracle, qesmc, seed=8)
    path = str(tmp_path / "scipy.nc")
    scipy_workspace(path, p, version, dz_double=dz_double)
    r = qesmc.workspace_read(path, velocity=False)
    assert (r["nx"], r["ny"], r["nz"]) == (p.nx, p.ny, p.nz)
    assert r["dx"] == np.float32(p.dx) and r["dy"] == np.float32(p.dy)
    assert np.array_equal(r["dz_array"], p.dz_array)
    for name, arr in zip("efghmn", p.coeffs):
        assert np.array_equal(r[name], arr), name  # record 0 of two
    assert np.array_equal(r["icellflag"], p.icellflag)
    assert np.array_equal(r["u0"], p.u0) and np.array_equal(r["w0"], p.w0)


def test_missing_dz_array_is_uniform_from_z(qesmc, oracle, tmp_path):
    """Domain(const std::string&) without dz_array: dz = z[1] - z[0]
    everywhere (Domain.cpp:86-93)."""
    p, _, _, _ = problem(oracle, qesmc, seed=9)
    path = str(tmp_path / "nodz.nc")
    scipy_workspace(path, p, 2, with_dz=False, uniform_dz=2.5)
    r = qesmc.workspace_read(path, velocity=False)
    assert np.all(r["dz_array"] == np.float32(2.5))


def test_errors(qesmc, oracle, tmp_path):
    p, u, v, w = problem(oracle, qesmc)
    with pytest.raises(qesmc.QesmcError) as ei:
        qesmc.workspace_read(str(tmp_path / "missing.nc"))
    assert ei.value.code == qesmc.ERR_IO
    # a NetCDF-4 (HDF5) file is not read: its signature is refused
    h5 = tmp_path / "ws4.nc"
    h5.write_bytes(b"\x89HDF\r\n\x1a\n" + bytes(512))
    with pytest.raises(qesmc.QesmcError) as ei:
        qesmc.workspace_dims(str(h5))
    assert ei.value.code == qesmc.ERR_IO
    # a truncated classic file
    path = str(tmp_path / "ws.nc")
    write(qesmc, path, p, u, v, w)
    data = open(path, "rb").read()
    short = tmp_path / "short.nc"
    short.write_bytes(data[:200])
    with pytest.raises(qesmc.QesmcError) as ei:
        qesmc.workspace_read(str(short))
    assert ei.value.code == qesmc.ERR_IO
    # a file without velocities: asking for them is an argument error
    nov = str(tmp_path / "nov.nc")
    scipy_workspace(nov, p, 2)
    with pytest.raises(qesmc.QesmcError) as ei:
        qesmc.workspace_read(nov, velocity=True)
    assert ei.value.code == qesmc.ERR_ARG


@pytest.mark.parametrize("flag_type,fill", [("f", float("nan")), ("d", 1e20), ("d", -3e9), ("f", float("inf"))])
def test_icellflag_values_beyond_int_are_refused(qesmc, oracle, tmp_path, flag_type, fill):
    """icellflag stored as floating point is converted to int; a NaN or a
    value outside int's range cannot be converted and the read fails."""
    p, _, _, _ = problem(oracle, qesmc, seed=4)
    path = str(tmp_path / "badflag.nc")
    scipy_workspace(path, p, 2, flag_type=flag_type, flag_fill=fill)
    with pytest.raises(qesmc.QesmcError) as ei:
        qesmc.workspace_read(path, velocity=False)
    assert ei.value.code == qesmc.ERR_IO
    # the same file with an integral value in range reads
    scipy_workspace(path, p, 2, flag_type=flag_type, flag_fill=2.0)
    r = qesmc.workspace_read(path, velocity=False)
    assert r["icellflag"][r["icellflag"].size // 2] == 2
