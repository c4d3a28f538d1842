"""The C++ surface (include/tbem/fmm.hpp) driven from compiled C++ callers,
checked against the oracle: FMMOperator<3,R,C>::apply, clone, the matrix-free
make_direct_nbody_operator, SparseOperator and IntegralOperator<3,1,1>."""
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


SPARSE_EXE = os.path.join(ROOT, "3bem_amd", "tbem_sparse_example")


def test_cpp_sparse_and_integral_operator(orc, tbem, tmp_path):
    """SparseOperator::csr_from_coo / apply / members / right_multiply and
    IntegralOperator<3,1,1> with a direct far field, driven from C++: the products are
    bit-exact with the oracle, and the integral operator equals n + (g + c) formed
    from the oracle's sparse products and the same GPU far field."""
    if not os.path.exists(SPARSE_EXE):
        pytest.fail("tbem_sparse_example not built (run __graft_entry__.build())")
    rng = np.random.default_rng(61)
    n_rows, n_cols, nnz, n_pts = 700, 500, 9000, 300
    rows, cols = rng.integers(0, n_rows, nnz), rng.integers(0, n_cols, nnz)
    vals, x = rng.standard_normal(nnz), rng.standard_normal(n_cols)
    locs = rng.uniform(0, 1, (n_pts, 3))
    nrm = rng.standard_normal((n_pts, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    w = rng.uniform(0.5, 1.5, n_pts)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([n_rows, n_cols, nnz, n_pts], dtype=np.int64).tofile(f)
        rows.astype(np.int64).tofile(f)
        cols.astype(np.int64).tofile(f)
        for a in (vals, x, locs, nrm, w):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
    print("CASE cpp sparse and integral operator", flush=True)
    r = subprocess.run([SPARSE_EXE, str(inp), str(out)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=120)
    assert r.returncode == 0, r.stdout.decode()
    raw = np.fromfile(out, dtype=np.uint8)
    y, y2, yi = (raw[i * 8 * n_rows:(i + 1) * 8 * n_rows].view(np.float64) for i in range(3))
    rest = raw[24 * n_rows:].view(np.int64)
    rp, ci, nnz_ata = rest[:n_rows + 1], rest[n_rows + 1:n_rows + 1 + nnz], int(rest[n_rows + 1 + nnz])
    cres = rest[n_rows + 2 + nnz:].view(np.float64)
    assert cres.tolist() == [4.0, 14.0, 14.0, 4.0, 14.0, 14.0, 14.0]  # unit_tests CondenseThenDistribute
    rp_o, ci_o, v_o, y_o = orc.sparse_coo(n_rows, n_cols, rows, cols, vals, x)
    assert np.array_equal(rp, rp_o) and np.array_equal(ci, ci_o)
    assert np.array_equal(y.view(np.uint64), y_o.view(np.uint64))
    assert np.array_equal(y2.view(np.uint64), y_o.view(np.uint64))
    # the integral operator, composed on the host
    ir = np.arange(n_pts)
    t1 = orc.sparse_coo(n_pts, n_cols, ir, ir % n_cols, np.full(n_pts, 0.5), x)[3]
    data = tbem.NBodyData(locs, nrm, locs, nrm, w)
    t2 = tbem.DirectOperator(data, tbem.LaplaceSingle()).apply(t1)
    gr = np.arange(n_rows)
    g = orc.sparse_coo(n_rows, n_pts, gr, gr % n_pts, 1.0 + 0.001 * gr, t2)[3]
    assert np.array_equal(yi.view(np.uint64), (y_o + (g + y_o)).view(np.uint64))
    at = (n_cols, n_rows) + orc.sparse_coo(n_cols, n_rows, cols, rows, vals)[:3]
    a = (n_rows, n_cols, rp_o, ci_o, v_o)
    assert nnz_ata == orc.sparse_right_multiply(at, a)[4].size
