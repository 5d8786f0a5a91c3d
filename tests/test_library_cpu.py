"""libqesmc on the CPU: the C ABI loads and exports every symbol its header
declares, the ctypes layouts equal the C compiler's, and the host-side input
builder equals the oracle's restatement bit for bit.  No device calls."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qesmc.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qesmc_[a-z_0-9]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(qesmc):
    names = declared_functions()
    assert len(names) >= 14
    L = qesmc.lib()
    for name in names:
        assert hasattr(L, name), name
    assert sorted(qesmc.EXPORTS) == names


def test_status_strings(qesmc):
    for code in range(-8, 1):
        s = qesmc.status_string(code)
        assert s and s != "unknown status"
    assert qesmc.status_string(-99) == "unknown status"


def test_struct_layouts_match_c_compiler(qesmc, tmp_path):
    """sizeof/offsetof from gcc vs the ctypes mirrors used by the binding."""
    structs = {
        "qesmc_grid": (qesmc.Grid, ["nx", "dx", "dz_array"]),
        "qesmc_coeffs": (qesmc.Coeffs, ["e", "n", "icellflag"]),
        "qesmc_fields": (qesmc.Fields, ["u0", "w"]),
        "qesmc_params": (qesmc.Params, ["omega", "tol", "itermax", "path", "flags", "tile_k", "wg_per_cu",
                                         "slab_planes", "stop_extend", "refine_every", "chebyshev_rho",
                                         "chebyshev_start"]),
        "qesmc_result": (qesmc.Result, ["iterations", "status", "final_max_dlambda", "glups", "tile_k", "reserved"]),
        "qesmc_building": (qesmc.Building, ["x_start", "rotation_deg"]),
        "qesmc_inflow": (qesmc.Inflow, ["kind", "wind_dir_deg"]),
    }
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "qesmc.h"', "int main(void){"]
    for s, (_, fields) in structs.items():
        src.append(f'printf("{s} %zu\\n", sizeof({s}));')
        for f in fields:
            src.append(f'printf("{s}.{f} %zu\\n", offsetof({s}, {f}));')
    src.append("return 0;}")
    c = tmp_path / "abi.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "abi"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    for s, (cls, fields) in structs.items():
        assert int(out[s]) == C.sizeof(cls), s
        for f in fields:
            assert int(out[f"{s}.{f}"]) == getattr(cls, f).offset, (s, f)


def test_create_without_device_fails_loudly(qesmc):
    if qesmc.gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(qesmc.QesmcError) as ei:
        qesmc.Context(0)
    assert ei.value.code == qesmc.ERR_NO_DEVICE


@pytest.mark.parametrize("seed,kind,wind,rot", [(1, 1, 233.0, 0.0), (2, 2, 90.0, 0.0), (3, 0, 317.5, 0.0),
                                                (4, 1, 11.25, 30.0)])
def test_city_builder_equals_oracle(qesmc, oracle, seed, kind, wind, rot):
    ncx, ncy, ncz, dx = 90, 70, 30, 2.0
    dz = qesmc.dz_geometric(ncz, 1.0, 1.02)
    blds = qesmc.random_buildings(seed, ncx, ncy, dx, 25, 8.0, 30.0, 4.0, 40.0)
    if rot:
        blds = [(b[0], b[1], b[2], b[3], b[4], b[5], rot) for b in blds[:6]]
    prof = (kind, 4.0, 10.0, {0: 0.1, 1: 0.5, 2: 0.15}[kind], wind)
    c = qesmc.city_build(ncx, ncy, dx, dz, blds, prof)
    p = oracle.build_city(ncx, ncy, dx, dz, blds, prof)
    for name in ("e", "f", "g", "h", "m", "n", "icellflag", "u0", "v0", "w0"):
        a, b = getattr(c, name), getattr(p, name)
        assert np.array_equal(a, b), (name, int((a != b).sum()))
    assert (c.icellflag == 0).sum() > 0


def test_random_buildings_properties(qesmc):
    ncx, ncy, dx = 200, 150, 2.0
    a = qesmc.random_buildings(42, ncx, ncy, dx, 120, 10.0, 40.0, 10.0, 60.0, margin=2)
    b = qesmc.random_buildings(42, ncx, ncy, dx, 120, 10.0, 40.0, 10.0, 60.0, margin=2)
    assert a == b and len(a) == 120
    occ = np.zeros((ncy, ncx), np.int32)
    for (xs, ys, L, W, H, base, rot) in a:
        assert 10.0 <= L <= 40.0 and 10.0 <= W <= 40.0 and 10.0 <= H <= 60.0
        i0, i1 = int(np.floor(xs / dx)), int(np.ceil((xs + L) / dx))
        j0, j1 = int(np.floor(ys / dx)), int(np.ceil((ys + W) / dx))
        assert i0 >= 2 and j0 >= 2 and i1 <= ncx - 2 and j1 <= ncy - 2
        occ[max(j0 - 1, 0):j1 + 1, max(i0 - 1, 0):i1 + 1] += 1  # footprint + one-cell gap
        occ[j0:j1, i0:i1] += 10
    # a footprint cell is never inside another building's footprint or gap
    assert occ[occ >= 10].max() == 11
    c = qesmc.random_buildings(43, ncx, ncy, dx, 120, 10.0, 40.0, 10.0, 60.0)
    assert c != a
    ln = qesmc.random_buildings(5, ncx, ncy, 1.0, 80, 6.0, 16.0, 20.0, 0.6, lognormal=True, hcap=120.0)
    hs = np.array([x[4] for x in ln])
    assert hs.max() <= 120.0 and 12.0 < np.median(hs) < 32.0


def test_city_builder_rejects_bad_arguments(qesmc):
    L = qesmc.lib()
    dz = np.ones(5, np.float32)
    inf = qesmc.Inflow(1, 4.0, 10.0, 0.5, 270.0)
    arr = np.zeros(10, np.float32)
    fp = arr.ctypes.data_as(C.POINTER(C.c_float))
    ip = np.zeros(10, np.int32).ctypes.data_as(C.POINTER(C.c_int))
    rc = L.qesmc_city_build(1, 5, 6, 1.0, dz.ctypes.data_as(C.POINTER(C.c_float)), 0, None, C.byref(inf),
                            fp, fp, fp, fp, fp, fp, ip, fp, fp, fp)
    assert rc == qesmc.ERR_ARG
    bad = qesmc.Inflow(7, 4.0, 10.0, 0.5, 270.0)
    with pytest.raises(qesmc.QesmcError):
        qesmc.city_build(4, 4, 1.0, qesmc.dz_uniform(3, 1.0), [], (7, 4.0, 10.0, 0.5, 270.0))
    assert qesmc.lib().qesmc_city_random_buildings(1, 10, 10, 1.0, -1, 1.0, 2.0, 1.0, 2.0, 0, 1.0, 0,
                                                   None) == qesmc.ERR_ARG
    del bad


def test_resident_calls_without_context_are_rejected(qesmc):
    L = qesmc.lib()
    assert L.qesmc_run(None, None, None) == qesmc.ERR_ARG
    assert L.qesmc_set_lambda(None, None) == qesmc.ERR_ARG
    assert L.qesmc_destroy(None) == qesmc.OK


def test_c_program_links_every_entry_point(qesmc, tmp_path):
    """A C99 program that includes the header and takes the address of every
    function it declares links against libqesmc.so (each declaration has a
    C-linkage definition of that name), and the entry points refuse a NULL
    context with QESMC_ERR_ARG without touching a device."""
    hdr = open(os.path.join(ROOT, "include", "qesmc.h")).read()
    names = sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*(qesmc_\w+)\s*\(", hdr, re.M)))
    assert "qesmc_refine" in names and "qesmc_velocity_gradient_walls" in names and len(names) > 30
    src = ["#include <stdio.h>", '#include "qesmc.h"', "int main(void){", "int bad = 0;"]
    src += [f"printf(\"%d\\n\", (int)((void *)&{n} != (void *)0));" for n in names]
    src += ["qesmc_params p; qesmc_default_params(&p); qesmc_result r;",
            "bad |= qesmc_refine(0, &p, &r) != QESMC_ERR_ARG;",
            "bad |= qesmc_run_static(0, &p, &r) != QESMC_ERR_ARG;",
            "double rho; float om;",
            "bad |= qesmc_estimate_omega(0, 10, &rho, &om) != QESMC_ERR_ARG;",
            "bad |= qesmc_velocity_gradient_walls(0, 2, 2, 0.1f, 0, 0, 0, 0) != QESMC_ERR_ARG;",
            "bad |= qesmc_wall_log_bc(0, 0.1f, 0) != QESMC_ERR_ARG;",
            "printf(bad ? \"FAIL\\n\" : \"OK\\n\"); return bad;}"]
    c = tmp_path / "links.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "links"
    libdir = os.path.join(ROOT, "qes-public_amd", "qesmc")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe),
                    "-L", libdir, "-lqesmc", f"-Wl,-rpath,{libdir}"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().splitlines()[-1] == "OK"
