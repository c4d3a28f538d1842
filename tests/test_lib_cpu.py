"""CPU-side checks of the product library: it loads, exports the declared ABI,
its host-only helpers agree with the oracle, and it refuses to run without a GPU."""
import ctypes
import importlib
import os
import re

import numpy as np
import pytest

from conftest import ROOT, golden

HEADER = os.path.join(ROOT, "include", "tbem_fmm.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(tbem_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(tbem):
    lib = tbem.load()
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    binding = importlib.import_module("3bem_amd._lib")
    assert sorted(binding.EXPORTED) == names


def test_surface_host_helper_matches_oracle(tbem, orc):
    g = golden("surfaces")
    for key, v in g.items():
        if key.startswith("order_"):
            assert np.array_equal(tbem.surface_points(int(key[6:])), v)


def test_operators_fail_loudly_without_gpu(tbem):
    lib = tbem.load()
    if lib.tbem_device_available():
        pytest.skip("a GPU is present")
    n = 10
    d = tbem.NBodyData(np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3)), np.ones(n))
    with pytest.raises(tbem.TbemError):
        tbem.FMMOperator(tbem.LaplaceSingle(), d, tbem.FMMConfig(0.3, 5, 10, 0.05, False))
    with pytest.raises(tbem.TbemError):
        tbem.make_direct_nbody_operator(d, tbem.LaplaceSingle())


def test_invalid_descriptors_are_rejected_before_device_use(tbem):
    """Argument validation happens in host code and names the problem."""
    lib = tbem.load()
    L = importlib.import_module("3bem_amd._lib")
    h = ctypes.c_void_p()
    assert lib.tbem_fmm_create(None, ctypes.byref(h)) != 0
    desc = L.FmmDesc()
    desc.kernel = 99
    desc.order = 5
    desc.min_pts_per_cell = 10
    rc = lib.tbem_fmm_create(ctypes.byref(desc), ctypes.byref(h))
    assert rc != 0
    msg = lib.tbem_last_error().decode()
    assert "kernel" in msg or "HIP" in msg or "hip" in msg


def test_host_config_mirrors_reference(tbem):
    c = tbem.FMMConfig(0.3, 5, 100, 0.05, False)
    assert (c.mac, c.order, c.min_pts_per_cell, c.d, c.account_for_small_cells) == (0.3, 5, 100, 0.05, False)
    assert tbem.ElasticDisplacement(1.0, 0.25).n_rows == 3
    assert tbem.LaplaceSingle().n_cols == 1
    with pytest.raises(ValueError):
        tbem.NBodyData(np.zeros((3, 3)), np.zeros((2, 3)), np.zeros((3, 3)), np.zeros((3, 3)), np.ones(3))


def test_workloads_shapes(wl):
    g = wl.icosphere_surface_15k()
    assert g["locs"].shape == (15360, 3)
    assert np.allclose(np.linalg.norm(g["normals"], axis=1), 1)
    assert np.all(np.sum(g["normals"] * g["locs"], axis=1) > 0)  # outward
    assert abs(g["weights"].sum() - 4 * np.pi) < 0.02  # polyhedral area below the sphere's
    c2 = wl.fault_and_free_surface(1000)
    assert c2["locs"].shape == (1000, 3) and c2["x"].shape == (3000,)
    c4 = wl.fault_and_stations(n_tri=400, n_obs=500)
    assert c4["src"].shape == (1200, 3) and c4["obs"].shape == (500, 3)


def test_saved_plan_refused_before_touching_the_device(tbem):
    """A blob that is not a saved plan is refused by the host checks (no device needed)."""
    lib = tbem.load()
    binding = importlib.import_module("3bem_amd._lib")
    rng = np.random.default_rng(0)
    pts = rng.uniform(0, 1, (10, 3))
    P = ctypes.POINTER(ctypes.c_double)
    d = binding.FmmDesc()
    d.kernel, d.n_obs, d.n_src = 0, 10, 10
    d.obs_locs = d.src_locs = pts.ctypes.data_as(P)
    d.obs_normals = d.src_normals = pts.ctypes.data_as(P)
    w = np.ones(10)
    d.src_weights = w.ctypes.data_as(P)
    d.mac, d.order, d.min_pts_per_cell, d.d = 0.3, 5, 4, 0.05
    h = ctypes.c_void_p()
    for blob, why in ((b"garbage!" * 4, b"magic"), (b"TBEMPLN1" + bytes(16) + b"xyz", b"truncated"),
                      (b"TBEMPLN1" + (3).to_bytes(8, "little") + bytes(8) + b"xyz", b"checksum")):
        assert lib.tbem_fmm_create_from_plan(ctypes.byref(d), blob, len(blob), ctypes.byref(h)) == 1
        assert why in lib.tbem_last_error()


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with status 2 and says why,
    before it touches torch or a device."""
    import subprocess
    import sys
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert p.returncode == 2
    assert "one GPU only" in p.stderr and p.stdout == ""


def test_ctypes_structs_match_the_c_header(tmp_path):
    """The Python binding's descriptor and info structs have the C header's size and
    field offsets (compiled with the host C compiler; a field added on one side only
    would shift every field after it)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    binding = importlib.import_module("3bem_amd._lib")
    structs = {"tbem_fmm_desc": binding.FmmDesc, "tbem_fmm_info_t": binding.FmmInfo,
               "tbem_nbody_desc": binding.NbodyDesc}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "tbem_fmm.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    seen = {}
    for line in out:
        if line.strip():
            cname, field, value = line.split()
            seen[(cname, field)] = int(value)
    for cname, py in structs.items():
        assert seen[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert seen[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
