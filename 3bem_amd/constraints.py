"""Linear constraints on the dofs, as tbempy exposes them (python_wrapper/constraints.cpp),
over the GPU library.

  * LinearTerm(dof, weight), ConstraintEQ(terms, rhs) -- cpp/constraint.h:21-52
  * boundary_condition, continuity_constraint, shift_constraints -- cpp/constraint.cpp
  * from_constraints(constraints) -> ConstraintMatrix -- cpp/constraint_matrix.cpp:35-52
  * distribute_vector(cm, v, n_total), condense_vector(cm, v) -- cpp/constraint_matrix.cpp:54-125
  * condense_matrix(row_cm, col_cm, sparse) -- cpp/constraint_matrix.cpp:252-292
  * homogenize_constraints(cm) -- cpp/constraint_matrix.cpp:294-303

from_constraints and condense_matrix run on the host at set-up time, as in the
reference; distribute_vector and condense_vector run on the GPU (NumPy arrays are
copied to and from the device, float64 CUDA tensors stay there) and reproduce the
reference's vectors bit for bit.
"""
import ctypes

import numpy as np

from . import _lib
from .ThreeD import SparseOperator, _dp, _is_torch, _szp, _torch_stream


class LinearTerm:
    __slots__ = ("dof", "weight")

    def __init__(self, dof, weight):
        self.dof, self.weight = int(dof), float(weight)

    def __eq__(self, other):
        return self.dof == other.dof and self.weight == other.weight

    def __repr__(self):
        return f"({self.dof}, {self.weight})"


class ConstraintEQ:
    """sum_k terms[k].weight * x[terms[k].dof] = rhs."""

    def __init__(self, terms, rhs):
        self.terms = [t if isinstance(t, LinearTerm) else LinearTerm(*t) for t in terms]
        self.rhs = float(rhs)

    def __eq__(self, other):
        return self.terms == other.terms and self.rhs == other.rhs

    def __repr__(self):
        return f"ConstraintEQ({self.terms}, {self.rhs})"


def boundary_condition(dof, value):
    return ConstraintEQ([LinearTerm(dof, 1.0)], value)


def continuity_constraint(dof1, dof2):
    return ConstraintEQ([LinearTerm(dof1, 1.0), LinearTerm(dof2, -1.0)], 0.0)


def shift_constraints(constraints, shift_dof):
    return [ConstraintEQ([LinearTerm(t.dof + shift_dof, t.weight) for t in c.terms], c.rhs) for c in constraints]


class ConstraintMatrix:
    """The solved constraints: x_c = rhs + sum_j w_j x_{d_j}, one per constrained dof c."""

    def __init__(self, handle, lib):
        self._h, self._lib = handle, lib

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.tbem_constraints_destroy(h)
            self._h = None

    def size(self):
        return int(self._lib.tbem_constraints_size(self._h))

    def rearranged(self):
        """{c: (terms [(dof, weight), ...], rhs)} for every constrained dof c."""
        n, nt = self.size(), int(self._lib.tbem_constraints_n_terms(self._h))
        cd, tp = np.zeros(n, dtype=np.uint64), np.zeros(n + 1, dtype=np.uint64)
        td, tw, rhs = np.zeros(nt, dtype=np.uint64), np.zeros(nt), np.zeros(n)
        _lib.check(self._lib.tbem_constraints_get(self._h, _szp(cd), _szp(tp), _szp(td), _dp(tw), _dp(rhs)))
        return {int(cd[i]): ([(int(td[k]), float(tw[k])) for k in range(int(tp[i]), int(tp[i + 1]))], float(rhs[i]))
                for i in range(n)}

    @property
    def map(self):
        return self.rearranged()


def from_constraints(constraints):
    """from_constraints (cpp/constraint_matrix.cpp:35-52)."""
    lib = _lib.load()
    ptr = np.zeros(len(constraints) + 1, dtype=np.uint64)
    dofs, weights, rhs = [], [], np.zeros(len(constraints))
    for e, c in enumerate(constraints):
        for t in c.terms:
            dofs.append(t.dof)
            weights.append(t.weight)
        ptr[e + 1] = len(dofs)
        rhs[e] = c.rhs
    d = np.asarray(dofs, dtype=np.uint64)
    w = np.asarray(weights, dtype=np.float64)
    h = ctypes.c_void_p()
    _lib.check(lib.tbem_constraints_create(len(constraints), _szp(ptr), _szp(d), _dp(w), _dp(rhs), ctypes.byref(h)))
    return ConstraintMatrix(h, lib)


def from_constraint_arrays(eq_ptr, term_dof, term_weight, rhs):
    """from_constraints on equations given as arrays: equation e has the terms
    eq_ptr[e] .. eq_ptr[e+1]-1 (term_dof, term_weight) and right-hand side rhs[e]."""
    lib = _lib.load()
    ptr = np.ascontiguousarray(eq_ptr, dtype=np.uint64)
    d = np.ascontiguousarray(term_dof, dtype=np.uint64)
    w = np.ascontiguousarray(term_weight, dtype=np.float64)
    r = np.ascontiguousarray(rhs, dtype=np.float64)
    if ptr.size != r.size + 1 or d.size != w.size or int(ptr[-1]) != d.size:
        raise ValueError("eq_ptr must have len(rhs) + 1 entries ending at len(term_dof) == len(term_weight)")
    h = ctypes.c_void_p()
    _lib.check(lib.tbem_constraints_create(r.size, _szp(ptr), _szp(d), _dp(w), _dp(r), ctypes.byref(h)))
    return ConstraintMatrix(h, lib)


def homogenize_constraints(cm):
    h = ctypes.c_void_p()
    _lib.check(cm._lib.tbem_constraints_homogenize(cm._h, ctypes.byref(h)))
    return ConstraintMatrix(h, cm._lib)


def _vec_op(cm, host_fn, dev_fn, v, n_arg, n_out):
    _lib.require_device()
    if _is_torch(v):
        import torch
        assert v.is_cuda and v.dtype == torch.float64 and v.is_contiguous()
        out = torch.empty(n_out, dtype=torch.float64, device=v.device)
        _lib.check(dev_fn(cm._h, ctypes.c_void_p(v.data_ptr()), n_arg, ctypes.c_void_p(out.data_ptr()),
                          _torch_stream()))
        return out
    v = np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1))
    out = np.zeros(n_out)
    _lib.check(host_fn(cm._h, _dp(v), n_arg, _dp(out)))
    return out


def distribute_vector(cm, v, n_total):
    """The reduced vector v (n_total - cm.size() entries) to the full vector."""
    n_in = v.numel() if _is_torch(v) else np.asarray(v).size
    if n_in != n_total - cm.size():
        raise ValueError(f"distribute_vector: {n_in} reduced values for {n_total} dofs and {cm.size()} constraints")
    return _vec_op(cm, cm._lib.tbem_distribute_vector, cm._lib.tbem_distribute_vector_dev, v, n_total, n_total)


def condense_vector(cm, v):
    """The full vector v to the reduced one (len(v) - cm.size() entries)."""
    n = v.numel() if _is_torch(v) else np.asarray(v).size
    if n < cm.size():
        raise ValueError(f"condense_vector: {n} values for {cm.size()} constraints")
    return _vec_op(cm, cm._lib.tbem_condense_vector, cm._lib.tbem_condense_vector_dev, v, n, n - cm.size())


def condense_matrix(row_cm, col_cm, matrix):
    """condense_matrix for a SparseOperator (cpp/constraint_matrix.cpp:252-292)."""
    h = ctypes.c_void_p()
    _lib.check(row_cm._lib.tbem_condense_sparse(row_cm._h, col_cm._h, matrix._h, ctypes.byref(h)))
    o = SparseOperator.__new__(SparseOperator)
    o._h, o._lib = h, row_cm._lib
    return o
