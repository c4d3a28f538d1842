"""Pins the CPU oracle (oracle/fmm_oracle.cpp) before it is trusted as the checker.

Each check is one the oracle could fail with a plausible mistake: the reference's
own golden data and known-answer tests, closed forms, brute force, the compiled
reference, and invariants of the method.
"""
import numpy as np
import pytest

from conftest import ELASTIC_NAMES, elastic_rows, golden, golden_names, rel_err
from oracle.oracle import components


# --------------------------------------------------------------------------- kernels
def test_elastic_kernels_vs_reference_sympy_data(orc):
    """unit_tests/test_data_elastic3: 360 sympy rows, rel 1e-8 (test_elastic_kernels.cpp:54-103)."""
    rows = elastic_rows()
    assert len(rows) == 360
    for r in rows:
        K = orc.kernel(ELASTIC_NAMES[r["name"]], r["obs"], r["src"], r["nobs"], r["nsrc"], r["mu"], r["nu"])
        got = K[r["k"], r["j"]]
        if r["value"] == 0.0:
            assert got == 0.0
        else:
            err = abs(r["value"] - got) / ((abs(r["value"]) + abs(got)) / 2)
            assert err < 1e-8, (r["name"], r["k"], r["j"], got, r["value"])


def test_laplace_closed_forms(orc):
    rng = np.random.default_rng(1)
    for _ in range(50):
        o, s = rng.standard_normal(3), rng.standard_normal(3)
        ns = rng.standard_normal(3)
        r = np.linalg.norm(s - o)
        assert orc.kernel(0, o, s, ns, ns)[0, 0] == pytest.approx(1 / (4 * np.pi * r), rel=1e-14)
        assert orc.kernel(1, o, s, ns, ns)[0, 0] == pytest.approx(ns @ (s - o) / (4 * np.pi * r ** 3), rel=1e-12)
        assert orc.kernel(2, o, s, ns, ns)[0, 0] == 0.0


def test_kernel_cutoff_and_identity(orc):
    o = np.array([0.1, 0.2, 0.3])
    for k in range(9):
        R = components(k)
        assert np.all(orc.kernel(k, o, o + 5e-7, o, o) == 0.0)  # r^2 = 7.5e-13 < 1e-12
    assert np.array_equal(orc.kernel(8, o, o + 1, o, o), np.eye(3))
    assert orc.kernel(7, o, o + 1, o, o)[0, 0] == 1.0


def test_elastic_u_symmetric_and_adjoint_identity(orc):
    """U(x<-y) = U(x<-y)^T; A_kj(y<-x; n_y) = T_jk(x<-y; n_y) (cpp/elastic_kernels.h:229-311)."""
    rng = np.random.default_rng(2)
    for _ in range(40):
        x, y = rng.standard_normal(3), rng.standard_normal(3)
        ny, nx = rng.standard_normal(3), rng.standard_normal(3)
        U = orc.kernel(3, x, y, nx, ny, 1.3, 0.27)
        assert np.allclose(U, U.T, rtol=1e-13, atol=0)
        T = orc.kernel(4, x, y, nx, ny, 1.3, 0.27)
        A = orc.kernel(5, y, x, ny, nx, 1.3, 0.27)
        assert np.allclose(A, T.T, rtol=1e-11, atol=1e-14)


# ----------------------------------------------------------------------- direct sum
def test_direct_known_answer(orc):
    """test_nbody_operator.cpp:27-41 (only the first source weight is used there)."""
    obs = np.array([[0, 0, 0], [0, 2, 0], [0, 0, -1]], dtype=float)
    src = np.array([[1, 0, 0], [0, 1, 0]], dtype=float)
    y = orc.direct(0, obs, np.zeros((3, 3)), src, np.zeros((2, 3)), np.array([1.0, 1.0]), [4.0, 0.0])
    assert np.allclose(y, [0.3183, 0.1424, 0.2251], atol=1e-3)


def test_direct_vs_brute_force_numpy(orc):
    rng = np.random.default_rng(3)
    o, s = rng.uniform(size=(40, 3)), rng.uniform(size=(50, 3))
    w, x = rng.uniform(size=50), rng.standard_normal(50)
    r = np.linalg.norm(s[None, :, :] - o[:, None, :], axis=2)
    y_np = (w * x / (4 * np.pi * r)).sum(axis=1)
    assert rel_err(orc.direct(0, o, o, s, s, w, x), y_np) < 1e-14


def test_direct_parallel_and_adjoint(orc):
    rng = np.random.default_rng(4)
    o, s = rng.uniform(size=(70, 3)), rng.uniform(size=(60, 3))
    on = rng.standard_normal((70, 3)); sn = rng.standard_normal((60, 3))
    w = rng.uniform(size=60)
    for k in (0, 1, 3, 4, 5, 6):
        R = components(k)
        x, g = rng.standard_normal(R * 60), rng.standard_normal(R * 70)
        y = orc.direct(k, o, on, s, sn, w, x)
        assert np.array_equal(y, orc.direct(k, o, on, s, sn, w, x, parallel=True))
        xt = orc.direct_adjoint(k, o, on, s, sn, w, g)
        assert abs(y @ g - x @ xt) < 1e-12 * np.abs(y).sum() * np.abs(g).max()


# --------------------------------------------------------------- geometry / octree
def test_surface_closed_form_order5(orc):
    """order 5 -> 6 points: theta = pi/4, 3pi/4; phi = 2 pi k / 3 (cpp/fmm.cpp:22-45)."""
    s = orc.surface(5)
    exp = []
    for th in (np.pi / 4, 3 * np.pi / 4):
        for k in range(3):
            ph = 2 * np.pi * k / 3
            exp.append([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    assert np.allclose(s, exp, atol=1e-15)


def test_surface_counts_and_unit_norm(orc):
    # point counts of the reference's sphere construction (cpp/fmm.cpp:22-45), as the
    # compiled reference produces them: n_e is close to, not equal to, the order
    for order, n in ((5, 6), (6, 6), (10, 11), (20, 20), (35, 35), (50, 50), (75, 74), (100, 99), (150, 151)):
        s = orc.surface(order)
        assert s.shape[0] == n
        assert np.allclose(np.linalg.norm(s, axis=1), 1.0, atol=1e-15)


def test_surfaces_match_golden(orc):
    g = golden("surfaces")
    for key, v in g.items():
        if key.startswith("order_"):
            assert np.array_equal(orc.surface(int(key[6:])), v)


def test_pinv_known_answers(orc):
    """test_blas_wrapper.cpp:32-66."""
    p = orc.pinv(np.array([[2.0, 1.0], [-1.0, 0.5]]))
    assert np.allclose(p, [[0.25, -0.5], [0.5, 1.0]], atol=1e-14)
    M = np.array([[0.0238032718573239, 0.1524037864980028], [0.1524037864980028, 0.9762067281426762]])
    assert np.allclose(orc.pinv(M), [[97620.6728142285282956, -15240.3786497941800917],
                                     [-15240.3786497941782727, 2380.3271857314393856]], atol=1e-4)
    assert np.allclose(orc.pinv(M, 1e-4), [[0.0237935097924219, 0.1524053105511083],
                                           [0.1524053105511085, 0.9762064902075779]], atol=1e-12)


def test_pinv_matches_golden_and_numpy(orc):
    g = golden("surfaces")
    assert np.allclose(orc.pinv(g["pinv_A"], 1e-10), g["pinv_A_1e-10"], rtol=0, atol=1e-9)
    A = np.random.default_rng(5).standard_normal((9, 9))
    assert np.allclose(orc.pinv(A, 1e-10), np.linalg.pinv(A), atol=1e-10)


@pytest.mark.parametrize("case", ["random", "cluster", "line", "dup", "icosphere"])
def test_octree_matches_reference_golden(orc, case):
    g = golden("octrees")
    cells, idx = orc.octree(g[f"{case}_pts"], int(g[f"{case}_leaf"]))
    assert np.array_equal(cells, g[f"{case}_cells"])
    assert np.array_equal(idx, g[f"{case}_idx"])


def count_children(cells):
    """count_children of test_octree.cpp:127-142 on depth-first pre-order cells:
    a leaf counts 1, an inner cell sum(1 + count_children(child))."""
    pos = [0]

    def rec():
        c = cells[pos[0]]
        pos[0] += 1
        if c[7] == 1:
            return 1
        total = 0
        while pos[0] < len(cells) and cells[pos[0]][0] == c[0] + 1:
            total += 1 + rec()
        return total

    return rec()


def test_octree_invariants(orc):
    """test_octree.cpp: degenerate line -> 31 cells; children partition parents; leaves cover."""
    g = golden("octrees")
    cells, idx = orc.octree(g["line_pts"], 1)
    assert count_children(cells) == 31
    pts = np.random.default_rng(6).uniform(-1, 1, (1000, 3))
    cells, idx = orc.octree(pts, 4)
    leaves = cells[cells[:, 7] == 1]
    assert leaves[:, 8].sum() == 1000
    assert np.all(leaves[:, 8] <= 4)
    all_leaf_idx = np.concatenate([idx[int(c[9]):int(c[9] + c[8])] for c in leaves])
    assert np.array_equal(np.sort(all_leaf_idx), np.arange(1000))
    for c in cells:  # ascending index lists
        seg = idx[int(c[9]):int(c[9] + c[8])]
        assert np.all(np.diff(seg) > 0)


# ---------------------------------------------------------------------------- FMM
@pytest.mark.parametrize("name", golden_names("small_") + ["self_elastic_u", "icosphere_laplace_15k",
                                                           "fault_surface_small_elastic_u", "highorder_laplace_single_35",
                                                           "highorder_laplace_double_75"])
def test_oracle_fmm_matches_reference_golden(orc, name):
    """The restated FMM reproduces the compiled reference's output (single thread)."""
    g = golden(name)
    op = orc.fmm(int(g["kernel"]), g["obs"], g["obs_normals"], g["src"], g["src_normals"], g["weights"],
                 float(g["mac"]), int(g["order"]), int(g["leaf"]), float(g["d"]), bool(g["small"]))
    c = op.counts()
    assert [c[k] for k in g["count_names"]] == list(g["counts"])
    y = op.apply(g["x"])
    assert rel_err(y, g["y"]) < 1e-12, name
    if "uc2e" in g:
        for L in range(g["uc2e"].shape[0]):
            assert np.array_equal(op.c2e(False, L), g["uc2e"][L])
        for which in ("p2p", "m2l", "m2p", "p2l"):
            assert np.array_equal(op.pairs(which), g[which])


def test_oracle_fmm_accuracy_reference_tolerances(orc):
    """test_fmm.cpp:70-150: per-entry min(rel, rel-to-mean) error vs direct."""
    for name, tol in (("highorder_laplace_single_35", 1e-4), ("highorder_laplace_double_75", 1e-4)):
        g = golden(name)
        y, ex = g["y"], g["direct"]
        mean = np.abs(ex).mean()
        err = np.minimum(np.abs((y - ex) / ex), np.abs((y - ex) / mean))
        assert err.max() < tol
        assert np.array_equal(y, orc.fmm(int(g["kernel"]), g["obs"], g["obs_normals"], g["src"],
                                         g["src_normals"], g["weights"], 0.35, int(g["order"]), 1, 0.05,
                                         False).apply(g["x"])) or rel_err(y, g["y"]) < 1e-12


def test_oracle_identity_fmm_is_exact(orc):
    """IdentityScalar: every stage carries the count exactly (test_fmm.cpp:31-68, 3D)."""
    rng = np.random.default_rng(8)
    n = 500
    src = rng.uniform(-1, 1, (n, 3)); obs = rng.uniform(1, 2, (n, 3))
    y = orc.fmm(7, obs, obs, src, src, np.ones(n), 0.3, 1, 1, 0.05, False).apply(np.ones(n))
    assert np.allclose(y, n, rtol=1e-12)


def test_oracle_vs_compiled_reference_live(orc, ref):
    """Where oracle/_ref is built: live comparison on fresh inputs, all kernels, both small-cell modes."""
    rng = np.random.default_rng(9)
    n = 300
    src, obs = rng.random((n, 3)), rng.random((n, 3))
    sn, on = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    w = rng.random(n)
    for k in range(9):
        x = rng.standard_normal(components(k) * n)
        for small in (False, True):
            a = ref.fmm(k, obs, on, src, sn, w, 0.3, 10, 8, 0.05, small).apply(x)
            b = orc.fmm(k, obs, on, src, sn, w, 0.3, 10, 8, 0.05, small).apply(x)
            assert rel_err(b, a) < 1e-12 or np.array_equal(a, b)


def test_planted_mistake_is_caught(orc):
    """A wrong MAC (0.35 instead of 0.3) changes the task lists and the result."""
    g = golden("small_laplace_single_nosc")
    op = orc.fmm(0, g["obs"], g["obs_normals"], g["src"], g["src_normals"], g["weights"], 0.35, 6, 20, 0.05, False)
    assert rel_err(op.apply(g["x"]), g["y"]) > 1e-8


def _dense_from_coo(n_rows, n_cols, rows, cols, vals):
    A = np.zeros((n_rows, n_cols))
    np.add.at(A, (np.asarray(rows, dtype=np.int64), np.asarray(cols, dtype=np.int64)), vals)
    return A


def test_sparse_known_answers(orc):
    """The reference's own sparse cases (unit_tests/test_sparse.cpp): an empty 2x2
    matrix gives zeros; {(0,0,1),(1,0,2)} times {0.5, 0} is {0.5, 1}; the 3x3
    {(0,0,1),(1,0,2),(1,1,4)} times {0.5, 7, 1} is {0.5, 29, 0}; row_ptrs ends at nnz."""
    rp, ci, v, y = orc.sparse_coo(2, 2, [], [], [], np.array([1.0, 1.0]))
    assert rp.tolist() == [0, 0, 0] and ci.size == 0 and y.tolist() == [0.0, 0.0]
    rp, ci, v, y = orc.sparse_coo(2, 2, [0, 1], [0, 0], [1.0, 2.0], np.array([0.5, 0.0]))
    assert y.tolist() == [0.5, 1.0] and rp[-1] == 2 and v.size == 2
    rp, ci, v, y = orc.sparse_coo(3, 3, [0, 1, 1], [0, 0, 1], [1.0, 2.0, 4.0], np.array([0.5, 7.0, 1.0]))
    assert y.tolist() == [0.5, 29.0, 0.0]
    assert rp.tolist() == [0, 1, 3, 3] and ci.tolist() == [0, 0, 1] and v.tolist() == [1.0, 2.0, 4.0]


def test_sparse_csr_layout_and_product(orc):
    """Random COO with duplicates and empty rows: the CSR keeps each row's entries in
    input order (a stable sort by row), row_ptrs are the prefix sums of the row counts,
    and the product equals the dense product of the summed matrix."""
    rng = np.random.default_rng(11)
    n_rows, n_cols, n = 300, 170, 4000
    rows = rng.integers(0, n_rows, n)
    rows[rows % 7 == 3] = 0  # leave rows 3, 10, ... empty
    cols = rng.integers(0, n_cols, n)
    vals = rng.standard_normal(n)
    x = rng.standard_normal(n_cols)
    rp, ci, v, y = orc.sparse_coo(n_rows, n_cols, rows, cols, vals, x)
    order = np.argsort(rows, kind="stable")
    assert np.array_equal(rp, np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n_rows))]))
    assert np.array_equal(ci, cols[order]) and np.array_equal(v, vals[order])
    assert rel_err(y, _dense_from_coo(n_rows, n_cols, rows, cols, vals) @ x) < 1e-14
    # the sum runs over each row's entries in stored order, starting from 0.0
    for r in (0, 1, 17, n_rows - 1):
        s = 0.0
        for k in range(rp[r], rp[r + 1]):
            s += v[k] * x[ci[k]]
        assert y[r] == s


def test_sparse_oracle_matches_compiled_reference(orc, ref):
    rng = np.random.default_rng(12)
    for n_rows, n_cols, n in ((1, 1, 0), (5, 9, 40), (257, 64, 3000)):
        rows, cols = rng.integers(0, n_rows, n), rng.integers(0, n_cols, n)
        vals, x = rng.standard_normal(n), rng.standard_normal(n_cols)
        a = orc.sparse_coo(n_rows, n_cols, rows, cols, vals, x)
        b = ref.sparse_coo(n_rows, n_cols, rows, cols, vals, x)
        for u, w in zip(a, b):
            assert np.array_equal(u, w)


def _random_csr(orc, rng, n_rows, n_cols, n):
    rows, cols = rng.integers(0, n_rows, n), rng.integers(0, n_cols, n)
    rp, ci, v, _ = orc.sparse_coo(n_rows, n_cols, rows, cols, rng.standard_normal(n))
    return n_rows, n_cols, rp, ci, v


def _dense(m):
    n_rows, n_cols, rp, ci, v = m
    out = np.zeros((n_rows, n_cols))
    np.add.at(out, (np.repeat(np.arange(n_rows), np.diff(rp)), ci), v)
    return out


def test_sparse_right_multiply_oracle(orc):
    """A B: columns ascending in each row, the dense product's values, and every
    touched position kept, explicit zeros included (A = [1 1], B = [1; -1] gives one
    stored 0.0)."""
    rng = np.random.default_rng(13)
    A, B = _random_csr(orc, rng, 60, 45, 400), _random_csr(orc, rng, 45, 33, 250)
    n_rows, n_cols, rp, ci, v = orc.sparse_right_multiply(A, B)
    assert (n_rows, n_cols) == (60, 33)
    for r in range(n_rows):
        assert np.all(np.diff(ci[rp[r]:rp[r + 1]]) > 0)
    assert np.abs(_dense((n_rows, n_cols, rp, ci, v)) - _dense(A) @ _dense(B)).max() < 1e-13
    z = orc.sparse_right_multiply((1, 2, [0, 2], [0, 1], [1.0, 1.0]), (2, 1, [0, 1, 2], [0, 0], [1.0, -1.0]))
    assert z[2].tolist() == [0, 1] and z[3].tolist() == [0] and z[4].tolist() == [0.0]


def test_sparse_right_multiply_oracle_matches_compiled_reference(orc, ref):
    rng = np.random.default_rng(14)
    for shape in ((1, 1, 1, 0, 0), (30, 20, 10, 150, 80), (200, 90, 70, 2000, 900)):
        m, k, n, na, nb = shape
        A, B = _random_csr(orc, rng, m, k, na), _random_csr(orc, rng, k, n, nb)
        a, b = orc.sparse_right_multiply(A, B), ref.sparse_right_multiply(A, B)
        assert a[:2] == b[:2]
        for u, w in zip(a[2:], b[2:]):
            assert np.array_equal(u, w)
