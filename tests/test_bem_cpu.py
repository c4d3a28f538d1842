"""Host-side far-field operators of the BEM integral operator (3bem_amd/bem.py)
against the compiled reference's make_galerkin_operator, make_interpolation_operator
and nbody_data_from_bem (fixture tests/golden/bem_far_ops.npz, written by
tests/golden/make_bem_golden.py from oracle/_ref; bit for bit), and against what the
geometry fixes: basis sums, areas, exact interpolation of linear functions. Also the
C ABI's input checks for sparse operators, which run before any device call."""
import ctypes
import importlib

import numpy as np
import pytest

from conftest import golden

bem = importlib.import_module("3bem_amd.bem")


def quad(g, qn):
    return g[f"{qn}_x"], g[f"{qn}_w"]


def csr(orc, coo):
    n_rows, n_cols, rows, cols, vals = coo
    rp, ci, v, _ = orc.sparse_coo(n_rows, n_cols, rows, cols, vals)
    return n_rows, n_cols, rp, ci, v


@pytest.mark.parametrize("qn", ["q3", "q7"])
@pytest.mark.parametrize("which", ["galerkin", "interp"])
@pytest.mark.parametrize("nc", [1, 3])
def test_bem_sparse_builders_match_reference_fixture(orc, qn, which, nc):
    g = golden("bem_far_ops")
    if which == "galerkin":
        coo = bem.galerkin_coo(nc, g["obs_facets"], quad(g, qn))
    else:
        coo = bem.interpolation_coo(nc, g["src_facets"], quad(g, qn))
    n_rows, n_cols, rp, ci, v = csr(orc, coo)
    key = f"{which}_{nc}_{qn}"
    assert [n_rows, n_cols] == g[key + "_shape"].tolist()
    assert np.array_equal(rp, g[key + "_rp"])
    assert np.array_equal(ci, g[key + "_ci"])
    assert np.array_equal(v, g[key + "_v"])


@pytest.mark.parametrize("qn", ["q3", "q7"])
def test_nbody_data_from_bem_matches_reference_fixture(qn):
    g = golden("bem_far_ops")
    d = bem.nbody_data_from_bem(g["obs_facets"], g["src_facets"], quad(g, qn), quad(g, qn))
    for k in ("obs_locs", "obs_normals", "src_locs", "src_normals", "src_weights"):
        assert np.array_equal(getattr(d, k), g[f"nbody_{qn}_{k}"]), k


def test_bem_builders_match_live_reference(ref, orc):
    """Fresh inputs through the compiled reference where it is built."""
    rng = np.random.default_rng(21)
    facets = rng.standard_normal((40, 3, 3))
    q = (rng.uniform(0, 0.5, (6, 2)), rng.uniform(0, 1, 6))
    for which, fn in (("galerkin", bem.galerkin_coo), ("interp", bem.interpolation_coo)):
        for nc in (1, 3):
            a = csr(orc, fn(nc, facets, q))
            b = ref.bem_sparse(which, nc, facets, q)
            assert a[:2] == b[:2]
            for u, w in zip(a[2:], b[2:]):
                assert np.array_equal(u, w)
    d = bem.nbody_data_from_bem(facets[:10], facets, q, q)
    r = ref.nbody_from_bem(facets[:10], facets, q, q)
    for k, w in zip(("obs_locs", "obs_normals", "src_locs", "src_normals", "src_weights"), r):
        assert np.array_equal(getattr(d, k), w), k


def test_bem_geometry_closed_forms():
    """Unit normals orthogonal to the edges; jacobian = twice the area (Heron); with a
    rule whose weights sum to 1/2 the source weights add up to the mesh area, the
    Galerkin rows of one facet add up to its area, and interpolating the corner
    coordinates reproduces the quadrature points (the basis is exact on linear
    functions)."""
    g = golden("bem_far_ops")
    f = g["obs_facets"]
    n = bem.facet_normals(f)
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, rtol=0, atol=1e-15)
    for e in (f[:, 1] - f[:, 0], f[:, 2] - f[:, 0]):
        assert np.abs(np.einsum("ij,ij->i", n, e)).max() < 1e-14
    a, b, c = (np.linalg.norm(f[:, i] - f[:, j], axis=1) for i, j in ((0, 1), (1, 2), (2, 0)))
    s = 0.5 * (a + b + c)
    heron = np.sqrt(s * (s - a) * (s - b) * (s - c))
    assert np.allclose(bem.facet_jacobians(f), 2 * heron, rtol=1e-12, atol=0)
    for qn in ("q3", "q7"):
        q = quad(g, qn)
        assert abs(q[1].sum() - 0.5) < 1e-15
        _, _, sw = bem.nbody_src_from_bem(f, q)
        assert abs(sw.sum() - heron.sum()) < 1e-12 * heron.sum()
        n_rows, n_cols, rows, cols, vals = bem.galerkin_coo(1, f, q)
        G = np.zeros((n_rows, n_cols))
        np.add.at(G, (rows, cols), vals)
        assert np.allclose(G.sum(1).reshape(-1, 3).sum(1), heron, rtol=1e-12)
        n_rows, n_cols, rows, cols, vals = bem.interpolation_coo(3, f, q)
        interp = np.zeros((n_rows, n_cols))
        np.add.at(interp, (rows, cols), vals)
        corner_xyz = f.reshape(-1, 3).T.reshape(-1)  # component-major dof values
        pts, _ = bem.nbody_obs_from_bem(f, q)
        assert np.allclose(interp @ corner_xyz, pts.T.reshape(-1), rtol=0, atol=1e-14)


def test_sparse_abi_rejects_bad_csr_before_touching_the_device(tbem):
    """Bad CSR input returns TBEM_E_INVALID, and shapes past 32-bit indexing
    TBEM_E_UNSUPPORTED, from the host checks (no device needed)."""
    lib = tbem.load()
    SZ = ctypes.c_size_t
    h = ctypes.c_void_p()

    def create(n_rows, n_cols, vals, cols, rp):
        v = np.asarray(vals, dtype=np.float64)
        c = np.asarray(cols, dtype=np.uint64)
        r = np.asarray(rp, dtype=np.uint64)
        return lib.tbem_sparse_create(n_rows, n_cols, len(v), v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      c.ctypes.data_as(ctypes.POINTER(SZ)), r.ctypes.data_as(ctypes.POINTER(SZ)),
                                      ctypes.byref(h))

    assert create(2, 2, [1.0, 2.0], [0, 2], [0, 1, 2]) == 1      # column out of range
    assert create(2, 2, [1.0, 2.0], [0, 1], [0, 2, 1]) == 1      # row_ptrs decrease
    assert create(2, 2, [1.0, 2.0], [0, 1], [1, 1, 2]) == 1      # row_ptrs[0] != 0
    assert create(2, 2, [1.0, 2.0], [0, 1], [0, 1, 3]) == 1      # row_ptrs[n_rows] != nnz
    assert b"column" in lib.tbem_last_error() or b"row_ptrs" in lib.tbem_last_error()
    rows = np.array([0, 3], dtype=np.uint64)
    cols = np.array([0, 0], dtype=np.uint64)
    vals = np.array([1.0, 1.0])
    rc = lib.tbem_sparse_from_coo(2, 2, 2, rows.ctypes.data_as(ctypes.POINTER(SZ)),
                                  cols.ctypes.data_as(ctypes.POINTER(SZ)),
                                  vals.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(h))
    assert rc == 1 and b"row index" in lib.tbem_last_error()
    # 32-bit column indices and row counts on the device: larger shapes are refused
    # before row_ptrs is read past its first entries.
    assert create(1 << 31, 2, [], [], [0]) == 4
    assert create(2, 1 << 31, [], [], [0, 0, 0]) == 4
    assert b"2^31" in lib.tbem_last_error()
    rc = lib.tbem_sparse_from_coo(2, 1 << 31, 0, None, None, None, ctypes.byref(h))
    assert rc == 4 and b"2^31" in lib.tbem_last_error()
