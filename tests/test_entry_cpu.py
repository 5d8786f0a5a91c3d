"""The entry points the tooling calls, checked without a device: bench.py's
refusal of more than one GPU, its algorithmic byte counts, and the build /
smoke functions of __graft_entry__."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("n", [2, 8])
def test_bench_refuses_more_than_one_gpu(n):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_bench_bytes_per_update():
    sys.path.insert(0, ROOT)
    import bench
    Q = bench.Q
    assert bench.bytes_per_lup({"path": Q.PATH_SPLIT, "compressed": 1}) == 17.0
    assert bench.bytes_per_lup({"path": Q.PATH_SPLIT, "compressed": 0}) == 40.0
    assert bench.bytes_per_lup({"path": Q.PATH_FUSED, "compressed": 1}) == 13.0
    assert bench.bytes_per_lup({"path": Q.PATH_PLAIN, "compressed": 0}) == 64.0


def test_graft_entry_exposes_build_and_smoke():
    sys.path.insert(0, ROOT)
    import __graft_entry__ as g
    assert callable(g.build) and callable(g.smoke)


def _bench():
    sys.path.insert(0, ROOT)
    import bench
    return bench


def test_fixture_probe_faces_address_the_sampled_values():
    """c5_fixture_probes maps every sample of the FP64 fixture (x-z and y-z
    slices and random faces) to the face index whose value it holds: a
    synthetic fixture sampled from a field f(face) = face index reads back
    exactly its own indices."""
    import numpy as np
    b = _bench()
    nx, ny, nz, ks = 13, 11, 9, 6
    field = np.arange(nx * ny * nz, dtype=np.float64).reshape(nz, ny, nx)
    rows, cols = (2, 7), (3, 10)
    pts = np.array([0, 5, nx * ny * nz - 1, 400], np.int32)
    F = {"ks": np.int32(ks), "xz_rows": np.array(rows), "yz_cols": np.array(cols), "pts": pts}
    for tag, scale in ((270, 1.0), (180, 2.0)):
        comp = [field * scale, field * scale + 0.25, field * scale + 0.5]
        F["xz_%d" % tag] = np.stack([np.stack([A[:ks, j, :] for A in comp]) for j in rows])
        F["yz_%d" % tag] = np.stack([np.stack([A[:ks, :, i] for A in comp]) for i in cols])
        F["pts_%d" % tag] = np.stack([A.ravel()[pts] for A in comp])
    faces, ref = b.c5_fixture_probes(F, nx, ny)
    assert faces.size == len(rows) * ks * nx + len(cols) * ks * ny + pts.size
    assert np.array_equal(ref[270][0], faces.astype(np.float64))
    assert np.array_equal(ref[270][2], faces + 0.5)
    assert np.array_equal(ref[180][1], 2.0 * faces + 0.25)
    # superposition: 270 deg is the first basis, 180 deg the second, 90 deg is -270
    assert np.allclose(b.c5_reference(ref, 270.0), ref[270], atol=1e-12)
    assert np.allclose(b.c5_reference(ref, 180.0), ref[180], atol=1e-9)
    assert np.allclose(b.c5_reference(ref, 90.0), -ref[270], atol=1e-9)
    s = np.sqrt(0.5)
    assert np.allclose(b.c5_reference(ref, 45.0), -s * ref[270] - s * ref[180], atol=1e-9)


def test_roofline_arithmetic():
    """achieved = bytes per update x updates per launch / mean launch time;
    two split launches per iteration."""
    b = _bench()
    Q = b.Q
    res = [{"iterations": 500, "path": Q.PATH_SPLIT, "compressed": 1, "tile_v": 2, "ms_sor": 300.0}] * 3
    roof, ms = b.roofline_of(res, 148_404_596, "no-such-workload")
    launch_s = 300.0 / 1000 / 1000.0
    assert ms == 300.0 and roof["avg_launch_us"] == 300.0
    assert abs(roof["achieved"] - 17.0 * 148_404_596 / 2 / launch_s / 1e9) < 0.1
    assert abs(roof["at_36_bytes_per_lup"]["achieved"] - 36.0 * 148_404_596 / 2 / launch_s / 1e9) < 0.1
    assert roof["kernel"] == "k_sor_split<true, 2>" and roof["traffic"] is None


def test_roofline_arithmetic_of_the_march():
    """The two-colour march: one launch per iteration over every updated cell
    at 13 B per update, and its committed traffic figure on the benchmark
    city."""
    b = _bench()
    Q = b.Q
    res = [{"iterations": 500, "path": Q.PATH_FUSED, "compressed": 1, "tile_k": 8, "tile_v": 1,
            "ms_sor": 283.5}] * 3
    roof, ms = b.roofline_of(res, 148_404_596, "city1000x1000x150")
    launch_s = 283.5 / 500 / 1000.0
    assert roof["kernel"] == "k_sor_march<8>" and roof["bytes_per_lup"] == 13.0
    assert roof["updates_per_launch"] == 148_404_596 and roof["avg_launch_us"] == 567.0
    assert abs(roof["achieved"] - 13.0 * 148_404_596 / launch_s / 1e9) < 0.1
    assert roof["traffic"] == b.load_traffic("k_sor_march<8>", "city1000x1000x150")[0] > 0


def test_traffic_entries_name_their_summaries():
    """Every counter figure the bench line can carry comes from a committed
    profiles/ summary."""
    import json
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as fh:
        t = json.load(fh)
    for wl, kernels in t.items():
        if wl.startswith("_"):
            continue
        for kernel, ent in kernels.items():
            assert ent["bytes"] > 0
            assert os.path.exists(os.path.join(ROOT, "profiles", ent["summary"])), ent
            assert _bench().load_traffic(kernel, wl)[0] == ent["bytes"]


def test_physical_core_count_is_sane():
    phys, logical = _bench().physical_cores()
    assert 1 <= phys <= logical
