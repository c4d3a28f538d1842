"""GPU CSR sparse operator and integral operator against the oracle.

* SparseOperator.apply equals the oracle's restatement of SparseOperator::apply
  (cpp/sparse_operator.cpp:19-28) bit for bit, on the reference's own cases
  (unit_tests/test_sparse.cpp) and on ragged matrices that exercise the row blocking:
  empty rows, row tiles cut by the 512-row and 4096-entry limits, rows longer than
  one LDS pass, duplicates, signed zeros, subnormals and wide magnitudes.
* csr_from_coo and right_multiply equal the oracle's CSR arrays bit for bit.
* IntegralOperator: with the direct far field, y equals N x + (G F(I x) + C x) formed
  on the host from the oracle's sparse products and the same far field, bit for bit;
  with the FMM far field, it matches the oracle FMM to 1e-11 and the dense direct
  sum to the FMM's accuracy.
"""
import importlib
import zlib

import numpy as np
import pytest

from conftest import golden, rel_err

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def coo_to_csr_oracle(orc, n_rows, n_cols, rows, cols, vals, x=None):
    return orc.sparse_coo(n_rows, n_cols, rows, cols, vals, x)


def test_reference_sparse_cases_on_gpu(tbem):
    S = tbem.SparseOperator
    print("CASE sparse reference cases", flush=True)
    e = S.csr_from_coo(2, 2, [], [], [])
    assert e.apply(np.array([1.0, 1.0])).tolist() == [0.0, 0.0] and e.nnz() == 0
    a = S.csr_from_coo(2, 2, [0, 1], [0, 0], [1.0, 2.0])
    assert a.apply(np.array([0.5, 0.0])).tolist() == [0.5, 1.0]
    assert a.row_ptrs.tolist()[-1] == 2 and a.nnz() == 2
    b = S.csr_from_coo(3, 3, [0, 1, 1], [0, 0, 1], [1.0, 2.0, 4.0])
    assert b.apply(np.array([0.5, 7.0, 1.0])).tolist() == [0.5, 29.0, 0.0]
    assert b.n_rows() == 3 and b.n_cols() == 3


def ragged_case(rng, n_rows, n_cols, lengths, special=False):
    rows = np.repeat(np.arange(n_rows), lengths)
    rows = rng.permutation(rows)  # rows interleaved: csr_from_coo must keep input order per row
    cols = rng.integers(0, n_cols, rows.size)
    vals = rng.standard_normal(rows.size) * np.exp(rng.uniform(-30, 30, rows.size))
    x = rng.standard_normal(n_cols) * np.exp(rng.uniform(-20, 20, n_cols))
    if special and rows.size > 10:
        vals[:5] = [0.0, -0.0, 5e-324, -2.5e-310, 1e300]
        x[:3] = [-0.0, 3e-320, 1e-300]
        cols[:5] = [0, 1, 2, 1, 0]
    return rows, cols, vals, x


CASES = {
    # name: (n_rows, n_cols, row lengths)
    "short_rows": (5000, 3000, lambda r, n: r.integers(0, 6, n)),
    "row_cap_512": (2100, 50, lambda r, n: np.ones(n, dtype=int)),
    "chunk_edges": (12, 700, lambda r, n: np.array([2048, 2049, 4096, 4097, 1, 4095, 0, 8192, 1, 10001, 511, 513])),
    "near_field_like": (3000, 9000, lambda r, n: r.integers(60, 180, n)),
    "mostly_empty": (20000, 100, lambda r, n: (r.uniform(0, 1, n) < 0.01) * r.integers(1, 40, n)),
    "special_values": (300, 64, lambda r, n: r.integers(0, 30, n)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_sparse_apply_bit_exact_vs_oracle(tbem, orc, name):
    n_rows, n_cols, lengths = CASES[name]
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    rows, cols, vals, x = ragged_case(rng, n_rows, n_cols, lengths(rng, n_rows), special=name == "special_values")
    print(f"CASE sparse apply {name}: {n_rows}x{n_cols}, nnz {rows.size}", flush=True)
    rp, ci, v, y_ref = coo_to_csr_oracle(orc, n_rows, n_cols, rows, cols, vals, x)
    A = tbem.SparseOperator.csr_from_coo(n_rows, n_cols, rows, cols, vals)
    assert np.array_equal(A.row_ptrs, rp) and np.array_equal(A.column_indices, ci)
    assert np.array_equal(bits(A.values), bits(v))
    y = A.apply(x)
    assert np.array_equal(bits(y), bits(y_ref))
    B = tbem.SparseOperator(n_rows, n_cols, v, ci, rp)  # from the CSR arrays
    assert np.array_equal(bits(B.apply(x)), bits(y_ref))
    assert np.array_equal(bits(A.clone().apply(x)), bits(y_ref))


def test_sparse_apply_torch_stream_and_to_dense(tbem, orc):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(7)
    n_rows, n_cols = 4000, 2500
    rows, cols, vals, x = ragged_case(rng, n_rows, n_cols, rng.integers(0, 40, n_rows))
    print("CASE sparse apply torch", flush=True)
    _, _, _, y_ref = coo_to_csr_oracle(orc, n_rows, n_cols, rows, cols, vals, x)
    A = tbem.SparseOperator.csr_from_coo(n_rows, n_cols, rows, cols, vals)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        xt = torch.tensor(x, device="cuda")
        yt = A.apply(xt)
    s.synchronize()
    assert np.array_equal(bits(yt.cpu().numpy()), bits(y_ref))
    small = tbem.SparseOperator.csr_from_coo(50, 40, rows[rows < 50], cols[rows < 50] % 40, vals[rows < 50])
    D = np.zeros((50, 40))
    np.add.at(D, (rows[rows < 50], cols[rows < 50] % 40), vals[rows < 50])
    assert np.allclose(small.to_dense(), D, rtol=1e-15, atol=0)


def test_sparse_right_multiply_matches_oracle(tbem, orc):
    rng = np.random.default_rng(8)
    print("CASE sparse right_multiply", flush=True)
    for m, k, n, na, nb in ((1, 1, 1, 0, 0), (300, 200, 150, 3000, 1500), (2000, 900, 700, 20000, 9000)):
        ra, ca, va = rng.integers(0, m, na), rng.integers(0, k, na), rng.standard_normal(na)
        rb, cb, vb = rng.integers(0, k, nb), rng.integers(0, n, nb), rng.standard_normal(nb)
        A = tbem.SparseOperator.csr_from_coo(m, k, ra, ca, va)
        B = tbem.SparseOperator.csr_from_coo(k, n, rb, cb, vb)
        oa = (m, k) + coo_to_csr_oracle(orc, m, k, ra, ca, va)[:3]
        ob = (k, n) + coo_to_csr_oracle(orc, k, n, rb, cb, vb)[:3]
        _, _, rp, ci, v = orc.sparse_right_multiply(oa, ob)
        C = A.right_multiply(B)
        assert (C.n_rows(), C.n_cols()) == (m, n)
        assert np.array_equal(C.row_ptrs, rp) and np.array_equal(C.column_indices, ci)
        assert np.array_equal(bits(C.values), bits(v))
        x = rng.standard_normal(n)
        _, _, _, y = orc.sparse_coo(m, n, np.repeat(np.arange(m), np.diff(rp)), ci, v, x)
        assert np.array_equal(bits(C.apply(x)), bits(y))


def test_sparse_shape_errors(tbem):
    A = tbem.SparseOperator.csr_from_coo(3, 4, [0], [3], [1.0])
    B = tbem.SparseOperator.csr_from_coo(5, 2, [0], [1], [1.0])
    with pytest.raises(tbem.TbemError):
        A.right_multiply(B)
    with pytest.raises(ValueError):
        A.apply(np.zeros(5))
    with pytest.raises(tbem.TbemError):
        tbem.SparseOperator.csr_from_coo(3, 4, [3], [0], [1.0])


# ------------------------------------------------------------------ integral operator


def bem_problem(tbem, orc, rng, kid):
    """Two bumpy triangle meshes (tests/golden/make_bem_golden.py's generator fixed in
    the fixture), a 3-point far rule, and random near-field / correction operators
    of the integral operator's shape."""
    bem = importlib.import_module("3bem_amd.bem")
    g = golden("bem_far_ops")
    obs_f, src_f = g["obs_facets"], g["src_facets"]
    q = (g["q3_x"], g["q3_w"])
    K = tbem.make_kernel(kid, 1.0, 0.25)
    n_rows, n_cols = K.n_rows * 3 * obs_f.shape[0], K.n_cols * 3 * src_f.shape[0]
    parts = {}
    for name, nnz in (("near", 12 * n_rows), ("corr", 5 * n_rows)):
        r, c, v = rng.integers(0, n_rows, nnz), rng.integers(0, n_cols, nnz), rng.standard_normal(nnz)
        parts[name] = (r, c, v, tbem.SparseOperator.csr_from_coo(n_rows, n_cols, r, c, v))
    return bem, obs_f, src_f, q, K, n_rows, n_cols, parts


@pytest.mark.parametrize("kid", [0, 3])
def test_integral_operator_direct_far_field_bit_exact(tbem, orc, kid):
    rng = np.random.default_rng(30 + kid)
    bem, obs_f, src_f, q, K, n_rows, n_cols, parts = bem_problem(tbem, orc, rng, kid)
    print(f"CASE integral operator, direct far field, kernel {kid}", flush=True)
    op = bem.integral_operator(parts["near"][3], parts["corr"][3], obs_f, src_f, q, q, K)
    assert (op.n_rows(), op.n_cols()) == (n_rows, n_cols)
    x = rng.standard_normal(n_cols)
    y = op.apply(x)
    # host composition: oracle sparse products, the same GPU far field
    I_coo = bem.interpolation_coo(K.n_cols, src_f, q)
    G_coo = bem.galerkin_coo(K.n_rows, obs_f, q)
    t1 = orc.sparse_coo(*I_coo, x)[3]
    t2 = op.farfield.apply(t1)
    g = orc.sparse_coo(*G_coo, t2)[3]
    nv = orc.sparse_coo(n_rows, n_cols, *parts["near"][:3], x)[3]
    cv = orc.sparse_coo(n_rows, n_cols, *parts["corr"][:3], x)[3]
    assert np.array_equal(bits(y), bits(nv + (g + cv)))
    # and the far field itself against the oracle's direct sum
    d = bem.nbody_data_from_bem(obs_f, src_f, q, q)
    t2_ref = orc.direct(kid, d.obs_locs, d.obs_normals, d.src_locs, d.src_normals, d.src_weights, t1)
    assert rel_err(t2, t2_ref) < 1e-13
    torch = pytest.importorskip("torch")
    yt = op.apply(torch.tensor(x, device="cuda"))
    assert np.array_equal(bits(yt.cpu().numpy()), bits(y))
    assert np.array_equal(bits(op.clone().apply(x)), bits(y))


@pytest.mark.parametrize("kid", [0, 4])
def test_integral_operator_fmm_far_field(tbem, orc, kid):
    rng = np.random.default_rng(40 + kid)
    bem, obs_f, src_f, q, K, n_rows, n_cols, parts = bem_problem(tbem, orc, rng, kid)
    print(f"CASE integral operator, FMM far field, kernel {kid}", flush=True)
    cfg = tbem.FMMConfig(0.3, 10, 10, 0.05, False)
    op = bem.integral_operator(parts["near"][3], parts["corr"][3], obs_f, src_f, q, q, K, cfg)
    x = rng.standard_normal(n_cols)
    y = op.apply(x)
    d = bem.nbody_data_from_bem(obs_f, src_f, q, q)
    t1 = orc.sparse_coo(*bem.interpolation_coo(K.n_cols, src_f, q), x)[3]
    G_coo = bem.galerkin_coo(K.n_rows, obs_f, q)
    nv = orc.sparse_coo(n_rows, n_cols, *parts["near"][:3], x)[3]
    cv = orc.sparse_coo(n_rows, n_cols, *parts["corr"][:3], x)[3]
    far_orc = orc.fmm(kid, d.obs_locs, d.obs_normals, d.src_locs, d.src_normals, d.src_weights,
                      0.3, 10, 10, 0.05, False).apply(t1)
    y_orc = nv + (orc.sparse_coo(*G_coo, far_orc)[3] + cv)
    assert rel_err(y, y_orc) < 1e-11
    far_dir = orc.direct(kid, d.obs_locs, d.obs_normals, d.src_locs, d.src_normals, d.src_weights, t1)
    g_far = orc.sparse_coo(*G_coo, far_dir)[3]
    assert rel_err(y - (nv + cv), g_far) < 1e-2  # the FMM's approximation at order 10


def test_integral_operator_shape_errors(tbem, orc):
    rng = np.random.default_rng(50)
    bem, obs_f, src_f, q, K, n_rows, n_cols, parts = bem_problem(tbem, orc, rng, 0)
    data = bem.nbody_data_from_bem(obs_f, src_f, q, q)
    far = tbem.DirectOperator(data, K)
    G = bem.make_galerkin_operator(1, obs_f, q)
    interp = bem.make_interpolation_operator(1, src_f, q)
    with pytest.raises(tbem.TbemError):  # galerkin and interpolation swapped
        tbem.IntegralOperator(parts["near"][3], parts["corr"][3], interp, far, G)
    with pytest.raises(TypeError):
        tbem.IntegralOperator(parts["near"][3], parts["corr"][3], G, parts["near"][3], interp)


def test_sparse_and_integral_empty_shapes(tbem):
    """Zero rows, zero columns and zero entries: the products return the right
    lengths (zeros where rows exist) without launching anything out of range."""
    print("CASE sparse empty shapes", flush=True)
    S = tbem.SparseOperator
    z = S.csr_from_coo(0, 5, [], [], [])
    assert z.apply(np.ones(5)).size == 0
    e = S.csr_from_coo(4, 0, [], [], [])
    assert e.apply(np.zeros(0)).tolist() == [0.0] * 4
    r = S(3, 3, [], [], [0, 0, 0, 0])
    assert r.apply(np.ones(3)).tolist() == [0.0] * 3 and r.nnz() == 0
    assert S.csr_from_coo(0, 0, [], [], []).right_multiply(S.csr_from_coo(0, 2, [], [], [])).n_cols() == 2


def test_integral_operator_mixed_far_field(tbem, orc):
    """The integral operator with the mixed-precision FMM as its far field agrees with
    the FP64 one to the mixed path's accuracy (the sparse parts are exact in both)."""
    rng = np.random.default_rng(60)
    bem, obs_f, src_f, q, K, n_rows, n_cols, parts = bem_problem(tbem, orc, rng, 3)
    print("CASE integral operator, mixed far field", flush=True)
    cfg = tbem.FMMConfig(0.3, 20, 10, 0.05, False)
    x = rng.standard_normal(n_cols)
    y64 = bem.integral_operator(parts["near"][3], parts["corr"][3], obs_f, src_f, q, q, K, cfg).apply(x)
    y32 = bem.integral_operator(parts["near"][3], parts["corr"][3], obs_f, src_f, q, q, K, cfg,
                                precision="mixed").apply(x)
    assert rel_err(y32, y64) < 1e-5
