"""ORACLE TEST INFRASTRUCTURE — ctypes access to the two CPU checkers.

* ``Oracle("orc")`` loads ``oracle/libtbem_oracle.so``: the from-scratch FP64
  restatement of 3bem's FMM (``oracle/fmm_oracle.cpp``).
* ``Oracle("ref")`` loads ``oracle/_ref/libtbem_ref.so``: the unmodified
  reference sources (cpp/fmm.cpp, cpp/octree.cpp, ...) behind
  ``oracle/ref_driver.cpp``; present only where /root/reference was there to
  compile it.

Only tests/, tests/golden/, the smoke check and bench.py's cpu_baseline leg use
this module. The product (3bem_amd/) never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBS = {
    "orc": os.path.join(HERE, "libtbem_oracle.so"),
    "ref": os.path.join(HERE, "_ref", "libtbem_ref.so"),
}

# Kernel ids (same numbering as include/tbem_fmm.h TBEM_K_*)
LAPLACE_SINGLE, LAPLACE_DOUBLE, LAPLACE_HYPERSINGULAR = 0, 1, 2
ELASTIC_U, ELASTIC_T, ELASTIC_A, ELASTIC_H = 3, 4, 5, 6
IDENTITY_SCALAR, IDENTITY_TENSOR = 7, 8
SCALAR_KERNELS = (0, 1, 2, 7)

COUNT_NAMES = ("n_e", "n_src_cells", "n_obs_cells", "n_up_levels", "n_down_levels",
               "p2m", "m2m", "p2p", "m2p", "p2l", "m2l", "l2l", "l2p")

_D = ctypes.POINTER(ctypes.c_double)
_L = ctypes.POINTER(ctypes.c_long)
_SZ = ctypes.c_size_t


def components(kernel):
    return 1 if kernel in SCALAR_KERNELS else 3


def available(prefix):
    return os.path.exists(LIBS[prefix])


def _p(a):
    return a.ctypes.data_as(_D) if a is not None else None


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    def __init__(self, prefix="orc"):
        self.prefix = prefix
        path = LIBS[prefix]
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run make -C oracle)")
        self.lib = ctypes.CDLL(path)
        f = self._f
        f("fmm_create").restype = ctypes.c_void_p
        f("fmm_create").argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                    _SZ, _D, _D, _SZ, _D, _D, _D,
                                    ctypes.c_double, _SZ, _SZ, ctypes.c_double, ctypes.c_int]
        f("fmm_apply").argtypes = [ctypes.c_void_p, _D, _D]
        f("fmm_counts").argtypes = [ctypes.c_void_p, _L]
        f("fmm_c2e").argtypes = [ctypes.c_void_p, ctypes.c_int, _SZ, _D]
        f("fmm_pairs").argtypes = [ctypes.c_void_p, ctypes.c_int, _D]
        f("fmm_destroy").argtypes = [ctypes.c_void_p]
        for name in ("direct_eval",):
            f(name).argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                _SZ, _D, _D, _SZ, _D, _D, _D, _D, _D]
        f("kernel_eval").argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     _D, _D, _D, _D, _D]
        f("octree").argtypes = [_D, _SZ, _SZ, _L, _L, _D, _L]
        f("surface").restype = ctypes.c_long
        f("surface").argtypes = [_SZ, _D]
        f("pinv").argtypes = [_D, ctypes.c_int, ctypes.c_double, _D]
        f("sparse_coo").argtypes = [_SZ, _SZ, _SZ, _L, _L, _D, _L, _L, _D, _D, _D]
        f("sparse_csr_apply").restype = ctypes.c_double
        f("sparse_csr_apply").argtypes = [_SZ, _SZ, _L, _L, _D, _D, _D, ctypes.c_int]
        f("sparse_right_multiply").restype = ctypes.c_long
        f("sparse_right_multiply").argtypes = [_SZ, _SZ, _L, _L, _D, _SZ, _L, _L, _D, ctypes.c_long, _L, _L, _D]
        for name in ("distribute", "condense"):
            f(name).argtypes = [ctypes.c_long, _L, _L, _L, _D, _D, _D, _SZ, _D]
        if prefix == "ref":
            self.lib.ref_from_constraints.restype = ctypes.c_long
            self.lib.ref_from_constraints.argtypes = [_SZ, _L, _L, _D, _D, _L, _L, _L, _D, _D, _L]
            self.lib.ref_condense_sparse.restype = ctypes.c_long
            self.lib.ref_condense_sparse.argtypes = [ctypes.c_long, _L, _L, _L, _D, _D,
                                                     ctypes.c_long, _L, _L, _L, _D, _D,
                                                     _SZ, _SZ, _L, _L, _D, ctypes.c_long, _L, _L, _D]
            self.lib.ref_bem_sparse.argtypes = [ctypes.c_int, _SZ, _SZ, _D, _SZ, _D, _D, _L, _L, _D,
                                                ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)]
            self.lib.ref_nbody_from_bem.argtypes = [_SZ, _D, _SZ, _D, _SZ, _D, _D, _SZ, _D, _D,
                                                    _D, _D, _D, _D, _D]
        if prefix == "orc":
            for name in ("direct_eval_par", "direct_adjoint"):
                f(name).argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                    _SZ, _D, _D, _SZ, _D, _D, _D, _D, _D]

    def _f(self, name):
        return getattr(self.lib, f"{self.prefix}_{name}")

    # ------------------------------------------------------------ FMM operator
    def fmm(self, kernel, obs_locs, obs_normals, src_locs, src_normals, src_weights,
            mac, order, leaf, d, small_cells=False, mu=1.0, nu=0.25):
        return OracleFMM(self, kernel, mu, nu, obs_locs, obs_normals, src_locs,
                         src_normals, src_weights, mac, order, leaf, d, small_cells)

    def _direct(self, name, kernel, obs_locs, obs_normals, src_locs, src_normals,
                src_weights, x, mu, nu, n_out):
        ol, on, sl, sn, sw, xv = (_c(a) for a in (obs_locs, obs_normals, src_locs,
                                                  src_normals, src_weights, x))
        out = np.zeros(n_out)
        self._f(name)(kernel, mu, nu, ol.shape[0], _p(ol), _p(on), sl.shape[0], _p(sl),
                      _p(sn), _p(sw), _p(xv), _p(out))
        return out

    def direct(self, kernel, obs_locs, obs_normals, src_locs, src_normals, src_weights,
               x, mu=1.0, nu=0.25, parallel=False):
        """FP64 direct sum (nbody_eval, cpp/nbody_operator.h:117-141)."""
        name = "direct_eval_par" if (parallel and self.prefix == "orc") else "direct_eval"
        return self._direct(name, kernel, obs_locs, obs_normals, src_locs, src_normals,
                            src_weights, x, mu, nu, components(kernel) * len(obs_locs))

    def direct_adjoint(self, kernel, obs_locs, obs_normals, src_locs, src_normals,
                       src_weights, g, mu=1.0, nu=0.25):
        return self._direct("direct_adjoint", kernel, obs_locs, obs_normals, src_locs,
                            src_normals, src_weights, g, mu, nu,
                            components(kernel) * len(src_locs))

    def kernel(self, kernel, obs, src, nobs, nsrc, mu=1.0, nu=0.25):
        R = components(kernel)
        out = np.zeros(R * R)
        args = [_c(a) for a in (obs, src, nobs, nsrc)]
        self._f("kernel_eval")(kernel, mu, nu, *[_p(a) for a in args], _p(out))
        return out.reshape(R, R)

    def octree(self, pts, leaf):
        """Cells in depth-first pre-order: (cells[n,10], indices)."""
        p = _c(pts)
        nc, ni = ctypes.c_long(), ctypes.c_long()
        self._f("octree")(_p(p), p.shape[0], leaf, ctypes.byref(nc), ctypes.byref(ni), None, None)
        cells = np.zeros((nc.value, 10))
        idx = np.zeros(ni.value, dtype=np.int64)
        self._f("octree")(_p(p), p.shape[0], leaf, ctypes.byref(nc), ctypes.byref(ni),
                          _p(cells), idx.ctypes.data_as(_L))
        return cells, idx

    def surface(self, order):
        n = self._f("surface")(order, None)
        out = np.zeros((n, 3))
        self._f("surface")(order, _p(out))
        return out

    def pinv(self, A, threshold=0.0):
        A = _c(A)
        n = A.shape[0]
        out = np.zeros((n, n))
        self._f("pinv")(_p(A), n, threshold, _p(out))
        return out


    def sparse_coo(self, n_rows, n_cols, rows, cols, vals, x=None):
        """CSR from COO triplets (SparseOperator::csr_from_coo, cpp/sparse_operator.cpp:140-172)
        and, with x, the CSR product (SparseOperator::apply, cpp/sparse_operator.cpp:19-28).
        Returns (row_ptrs, column_indices, values, y or None)."""
        r = np.ascontiguousarray(rows, dtype=np.int64)
        c = np.ascontiguousarray(cols, dtype=np.int64)
        v = _c(vals)
        n = r.shape[0]
        assert c.shape[0] == n and v.shape[0] == n
        assert n == 0 or (r.min() >= 0 and r.max() < n_rows and c.min() >= 0 and c.max() < n_cols)
        rp = np.zeros(n_rows + 1, dtype=np.int64)
        ci = np.zeros(n, dtype=np.int64)
        vv = np.zeros(n)
        xv = _c(x) if x is not None else None
        if xv is not None:
            assert xv.shape[0] == n_cols
        y = np.zeros(n_rows) if x is not None else None
        self._f("sparse_coo")(n_rows, n_cols, n, r.ctypes.data_as(_L), c.ctypes.data_as(_L), _p(v),
                              rp.ctypes.data_as(_L), ci.ctypes.data_as(_L), _p(vv), _p(xv), _p(y))
        return rp, ci, vv, y


    def sparse_csr_apply(self, n_rows, n_cols, row_ptrs, column_indices, values, x, reps=1):
        """SparseOperator::apply on CSR arrays, reps times: (y, mean seconds per apply)."""
        rp = np.ascontiguousarray(row_ptrs, dtype=np.int64)
        ci = np.ascontiguousarray(column_indices, dtype=np.int64)
        v, xv = _c(values), _c(x)
        assert rp.shape[0] == n_rows + 1 and xv.shape[0] == n_cols
        y = np.zeros(n_rows)
        t = self._f("sparse_csr_apply")(n_rows, n_cols, rp.ctypes.data_as(_L), ci.ctypes.data_as(_L), _p(v),
                                        _p(xv), _p(y), reps)
        return y, t

    def sparse_right_multiply(self, a, b):
        """A B (SparseOperator::right_multiply, cpp/sparse_operator.cpp:114-138) of two
        CSR triples (n_rows, n_cols, row_ptrs, column_indices, values): the same form."""
        (ar, ac, arp, aci, av), (br, bc, brp, bci, bv) = a, b
        assert ac == br
        arrs = [np.ascontiguousarray(x, dtype=t) for x, t in ((arp, np.int64), (aci, np.int64), (av, np.float64),
                                                               (brp, np.int64), (bci, np.int64), (bv, np.float64))]
        L = lambda x: x.ctypes.data_as(_L)  # noqa: E731
        args = [ar, ac, L(arrs[0]), L(arrs[1]), _p(arrs[2]), bc, L(arrs[3]), L(arrs[4]), _p(arrs[5])]
        nnz = self._f("sparse_right_multiply")(*args, -1, None, None, None)  # size query
        rp, ci, v = np.zeros(ar + 1, dtype=np.int64), np.zeros(nnz, dtype=np.int64), np.zeros(nnz)
        got = self._f("sparse_right_multiply")(*args, nnz, L(rp), L(ci), _p(v))
        assert got == nnz
        return ar, bc, rp, ci, v

    # --------------------------------------------------------------- constraints
    # Solved constraints travel as a tuple (cdof, tptr, tdof, tw, trhs): constrained
    # dofs ascending, x_cdof[i] = trhs[i] + sum_{k in tptr[i]..tptr[i+1]} tw[k] x_tdof[k].

    def from_constraints(self, constraints):
        """The compiled reference's from_constraints; constraints = [(terms [(dof, w)], rhs)].
        Reference library only."""
        ptr = [0]
        dof, w, rhs = [], [], []
        for terms, r in constraints:
            for d, x in terms:
                dof.append(d)
                w.append(x)
            ptr.append(len(dof))
            rhs.append(r)
        ptr, dof = np.asarray(ptr, dtype=np.int64), np.asarray(dof, dtype=np.int64)
        w, rhs = np.asarray(w, dtype=np.float64), np.asarray(rhs, dtype=np.float64)
        nt = ctypes.c_long()
        L = lambda x: x.ctypes.data_as(_L)  # noqa: E731
        args = (len(constraints), L(ptr), L(dof), _p(w), _p(rhs))
        n = self.lib.ref_from_constraints(*args, None, None, None, None, None, ctypes.byref(nt))
        out = (np.zeros(n, dtype=np.int64), np.zeros(n + 1, dtype=np.int64), np.zeros(nt.value, dtype=np.int64),
               np.zeros(nt.value), np.zeros(n))
        self.lib.ref_from_constraints(*args, L(out[0]), L(out[1]), L(out[2]), _p(out[3]), _p(out[4]), ctypes.byref(nt))
        return out

    @staticmethod
    def _cm_args(cm):
        cd, tp, td, tw, tr = (np.ascontiguousarray(a, dtype=t) for a, t in
                              zip(cm, (np.int64, np.int64, np.int64, np.float64, np.float64)))
        keep = (cd, tp, td, tw, tr)
        return keep, [cd.shape[0], cd.ctypes.data_as(_L), tp.ctypes.data_as(_L), td.ctypes.data_as(_L), _p(tw), _p(tr)]

    def distribute(self, cm, v, n_total):
        """distribute_vector (cpp/constraint_matrix.cpp:54-82) for solved constraints cm."""
        keep, a = self._cm_args(cm)
        v = _c(v)
        assert v.shape[0] == n_total - a[0]
        out = np.zeros(n_total)
        self._f("distribute")(*a, _p(v), n_total, _p(out))
        return out

    def condense(self, cm, v):
        """condense_vector (cpp/constraint_matrix.cpp:99-125) for solved constraints cm."""
        keep, a = self._cm_args(cm)
        v = _c(v)
        out = np.zeros(v.shape[0] - a[0])
        self._f("condense")(*a, _p(v), v.shape[0], _p(out))
        return out

    def condense_sparse(self, row_cm, col_cm, n_rows, n_cols, rp, ci, v):
        """The compiled reference's condense_matrix on a CSR matrix: (row_ptrs, column_indices,
        values). Reference library only."""
        k1, a1 = self._cm_args(row_cm)
        k2, a2 = self._cm_args(col_cm)
        rp, ci, v = np.ascontiguousarray(rp, dtype=np.int64), np.ascontiguousarray(ci, dtype=np.int64), _c(v)
        L = lambda x: x.ctypes.data_as(_L)  # noqa: E731
        args = (*a1, *a2, n_rows, n_cols, L(rp), L(ci), _p(v))
        nnz = self.lib.ref_condense_sparse(*args, -1, None, None, None)
        orp, oci, ov = np.zeros(n_rows - a1[0] + 1, dtype=np.int64), np.zeros(nnz, dtype=np.int64), np.zeros(nnz)
        assert self.lib.ref_condense_sparse(*args, nnz, L(orp), L(oci), _p(ov)) == nnz
        return orp, oci, ov

    def bem_sparse(self, which, n_components, facets, quad):
        """The compiled reference's make_galerkin_operator ("galerkin") or
        make_interpolation_operator ("interp") on a triangle mesh: (n_rows, n_cols,
        row_ptrs, column_indices, values). Reference library only."""
        f = _c(facets)
        qx, qw = _c(quad[0]), _c(quad[1])
        nnz = f.shape[0] * qw.shape[0] * 3 * n_components
        n_rows_max = n_components * f.shape[0] * max(3, qw.shape[0])
        rp = np.zeros(n_rows_max + 1, dtype=np.int64)
        ci = np.zeros(nnz, dtype=np.int64)
        v = np.zeros(nnz)
        nr, nc = _SZ(), _SZ()
        self.lib.ref_bem_sparse({"galerkin": 0, "interp": 1}[which], n_components, f.shape[0], _p(f),
                                qw.shape[0], _p(qx), _p(qw), rp.ctypes.data_as(_L), ci.ctypes.data_as(_L),
                                _p(v), ctypes.byref(nr), ctypes.byref(nc))
        return nr.value, nc.value, rp[:nr.value + 1], ci, v

    def nbody_from_bem(self, obs_facets, src_facets, obs_quad, src_quad):
        """The compiled reference's nbody_data_from_bem: (obs_locs, obs_normals,
        src_locs, src_normals, src_weights). Reference library only."""
        of, sf = _c(obs_facets), _c(src_facets)
        oqx, oqw, sqx, sqw = _c(obs_quad[0]), _c(obs_quad[1]), _c(src_quad[0]), _c(src_quad[1])
        no, ns = of.shape[0] * oqw.shape[0], sf.shape[0] * sqw.shape[0]
        out = [np.zeros((no, 3)), np.zeros((no, 3)), np.zeros((ns, 3)), np.zeros((ns, 3)), np.zeros(ns)]
        self.lib.ref_nbody_from_bem(of.shape[0], _p(of), sf.shape[0], _p(sf), oqw.shape[0], _p(oqx), _p(oqw),
                                    sqw.shape[0], _p(sqx), _p(sqw), *[_p(a) for a in out])
        return tuple(out)


class OracleFMM:
    def __init__(self, orc, kernel, mu, nu, obs_locs, obs_normals, src_locs, src_normals,
                 src_weights, mac, order, leaf, d, small_cells):
        self.o = orc
        self.kernel = kernel
        self.R = components(kernel)
        self.n_obs = len(obs_locs)
        self.n_src = len(src_locs)
        arrs = [_c(a) for a in (obs_locs, obs_normals, src_locs, src_normals, src_weights)]
        self._keep = arrs
        ol, on, sl, sn, sw = arrs
        self.h = orc._f("fmm_create")(kernel, mu, nu, self.n_obs, _p(ol), _p(on), self.n_src,
                                      _p(sl), _p(sn), _p(sw), mac, order, leaf, d,
                                      1 if small_cells else 0)

    def apply(self, x):
        x = _c(x)
        assert x.shape[0] == self.R * self.n_src
        out = np.zeros(self.R * self.n_obs)
        self.o._f("fmm_apply")(self.h, _p(x), _p(out))
        return out

    def counts(self):
        c = np.zeros(len(COUNT_NAMES), dtype=np.int64)
        self.o._f("fmm_counts")(self.h, c.ctypes.data_as(_L))
        return dict(zip(COUNT_NAMES, (int(v) for v in c)))

    def c2e(self, down, level):
        n = self.R * self.counts()["n_e"]
        out = np.zeros((n, n))
        self.o._f("fmm_c2e")(self.h, 1 if down else 0, level, _p(out))
        return out

    def pairs(self, which):
        key = {"p2p": (0, "p2p"), "m2l": (1, "m2l"), "m2p": (2, "m2p"), "p2l": (3, "p2l")}[which]
        n = self.counts()[key[1]]
        out = np.zeros((n, 14))
        if n:
            self.o._f("fmm_pairs")(self.h, key[0], _p(out))
        return out

    def __del__(self):
        try:
            self.o._f("fmm_destroy")(self.h)
        except Exception:
            pass
