"""A QES case through a workspace file: solve on the device, write the
inputs and the corrected field as a QES-Winds workspace file (NetCDF
classic), read it back and solve again from the file's inputs: the same
field, bitwise, and equal to the oracle's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_solve_from_a_workspace_file(gpu_ctx, oracle, qesmc, tmp_path):
    dz = qesmc.dz_geometric(14, 1.0, 1.02)
    blds = qesmc.random_buildings(12, 48, 36, 2.0, 5, 6.0, 16.0, 4.0, 12.0, margin=1)
    p = oracle.build_city(48, 36, 2.0, dz, blds, (1, 4.0, 10.0, 0.5, 200.0))
    u, v, w, lam, res = gpu_ctx.solve(p, itermax=80)
    path = str(tmp_path / "case.nc")
    qesmc.workspace_write(path, p.nx, p.ny, p.nz, p.dx, p.dy, p.dz_array, p.coeffs, p.icellflag, p.u0, p.v0,
                          p.w0, u, v, w)
    r = qesmc.workspace_read(path)
    q = oracle.Problem(r["nx"], r["ny"], r["nz"], r["dx"], r["dy"], r["dz_array"], r["e"], r["f"], r["g"], r["h"],
                       r["m"], r["n"], r["icellflag"], r["u0"], r["v0"], r["w0"])
    u2, v2, w2, _, res2 = gpu_ctx.solve(q, itermax=80)
    assert res2["iterations"] == res["iterations"]
    for a, b, c in ((u, u2, r["u"]), (v, v2, r["v"]), (w, w2, r["w"])):
        assert np.array_equal(a, b) and np.array_equal(a, c)
    ou, ov, ow, *_ = oracle.solve_rb_f32(q, itermax=80)
    assert np.array_equal(u2, ou) and np.array_equal(v2, ov) and np.array_equal(w2, ow)
