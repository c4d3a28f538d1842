"""The C++ surface (include/tbem/fmm.hpp) driven from a compiled C++ caller,
checked against the oracle: FMMOperator<3,R,C>::apply, clone, and the matrix-free
make_direct_nbody_operator."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_err

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "3bem_amd", "tbem_shim_example")


@pytest.mark.parametrize("kernel", [0, 3, 4])
def test_cpp_shim_matches_oracle(orc, tmp_path, kernel):
    if not os.path.exists(EXE):
        pytest.fail("tbem_shim_example not built (run __graft_entry__.build())")
    rng = np.random.default_rng(30 + kernel)
    n = 2000
    R = 1 if kernel == 0 else 3
    locs = rng.uniform(0, 1, (n, 3))
    nrm = rng.standard_normal((n, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    w = rng.uniform(0.5, 1.5, n)
    x = rng.standard_normal(R * n)
    order, leaf, mac, d = 6, 32, 0.3, 0.05
    inp = tmp_path / "in.bin"
    out = tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([kernel, n, order, leaf], dtype=np.int64).tofile(f)
        np.array([mac, d], dtype=np.float64).tofile(f)
        for a in (locs, nrm, w, x):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
    print("CASE cpp shim kernel", kernel, flush=True)
    r = subprocess.run([EXE, str(inp), str(out)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=120)
    assert r.returncode == 0, r.stdout.decode()
    res = np.fromfile(out, dtype=np.float64).reshape(3, R * n)
    y, direct, yc = res
    ref = orc.fmm(kernel, locs, nrm, locs, nrm, w, mac, order, leaf, d, False).apply(x)
    assert rel_err(y, ref) < 1e-12
    assert np.array_equal(yc, y)
    dref = orc.direct(kernel, locs, nrm, locs, nrm, w, x, parallel=True)
    assert rel_err(direct, dref) < 1e-13
