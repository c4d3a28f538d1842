"""Writes tests/golden/constraints.npz from the compiled reference (oracle/_ref):
from_constraints, distribute_vector, condense_vector and the sparse condense_matrix
on three constraint sets:

  * grid_all   -- continuity between every ordered pair of coincident dofs on a 6 x 5
                  cell triangulated sheet, 3 components, boundary values 0.7 at one vertex;
  * grid_chain -- consecutive continuity on a 7 x 4 sheet, 1 component, boundary -2.0;
  * mixed      -- the reference unit tests' own sets (unit_tests/test_constraint_matrix.cpp)
                  shifted apart and joined: chains, a + b + c, zero weights, circular
                  pairs, the field-average equation.

    make -C oracle ref && python tests/golden/make_constraints_golden.py
"""
import importlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle.oracle import Oracle  # noqa: E402

wl = importlib.import_module("3bem_amd.workloads")


def as_list(eq_ptr, td, tw, rhs):
    return [([(int(td[k]), float(tw[k])) for k in range(eq_ptr[e], eq_ptr[e + 1])], float(rhs[e]))
            for e in range(len(rhs))]


def mixed_set():
    cont = lambda a, b: ([(a, 1.0), (b, -1.0)], 0.0)  # noqa: E731
    bc = lambda d, v: ([(d, 1.0)], v)  # noqa: E731
    eqs = [cont(0, 1), cont(1, 2),                                # a = b = c
           cont(10, 11), cont(12, 11), cont(12, 13),              # a = b = c = d
           ([(20, 1.0), (21, -1.0), (22, -1.0)], 0.0), cont(22, 23),  # a = b + c, c = d
           ([(30, 1.0), (31, 0.0)], 0.0),                         # zero weight
           ([(35, 0.0)], 0.0),                                    # empty
           cont(40, 41), cont(41, 40),                            # circular
           bc(51, 4.0), cont(50, 52), cont(52, 53), cont(53, 54),  # CondenseRecurse set
           ([(60, 2.0), (60, 1.0), (61, -1.5), (62, 0.25)], 1.5),  # duplicate dof combined
           cont(70, 71), cont(72, 73), cont(74, 75), cont(76, 77),
           ([(70 + i, 1.0) for i in range(8)], 0.0)]              # field average
    n_total = 80
    return eqs, n_total


def main():
    ref = Oracle("ref")
    rng = np.random.default_rng(17)
    out = {}
    cases = {}
    ep, td, tw, rhs, n = wl.grid_continuity_constraints(6, 5, 3, all_pairs=True, bc_value=0.7)
    cases["grid_all"] = (as_list(ep, td, tw, rhs), n)
    ep, td, tw, rhs, n = wl.grid_continuity_constraints(7, 4, 1, all_pairs=False, bc_value=-2.0)
    cases["grid_chain"] = (as_list(ep, td, tw, rhs), n)
    cases["mixed"] = mixed_set()
    for name, (eqs, n_total) in cases.items():
        ptr = np.cumsum([0] + [len(t) for t, _ in eqs])
        out[f"{name}_eq_ptr"] = ptr
        out[f"{name}_eq_dof"] = np.array([d for t, _ in eqs for d, _ in t], dtype=np.int64)
        out[f"{name}_eq_w"] = np.array([w for t, _ in eqs for _, w in t])
        out[f"{name}_eq_rhs"] = np.array([r for _, r in eqs])
        out[f"{name}_n_total"] = np.array(n_total)
        cm = ref.from_constraints(eqs)
        for k, a in zip(("cdof", "tptr", "tdof", "tw", "trhs"), cm):
            out[f"{name}_{k}"] = a
        red = rng.standard_normal(n_total - len(cm[0]))
        full = rng.standard_normal(n_total)
        out[f"{name}_red"], out[f"{name}_full"] = red, full
        out[f"{name}_distributed"] = ref.distribute(cm, red, n_total)
        out[f"{name}_condensed"] = ref.condense(cm, full)
        nnz = 8 * n_total
        rows, cols = rng.integers(0, n_total, nnz), rng.integers(0, n_total, nnz)
        o = Oracle("orc")
        rp, ci, v, _ = o.sparse_coo(n_total, n_total, rows, cols, rng.standard_normal(nnz))
        out[f"{name}_A_rp"], out[f"{name}_A_ci"], out[f"{name}_A_v"] = rp, ci, v
        orp, oci, ov = ref.condense_sparse(cm, cm, n_total, n_total, rp, ci, v)
        out[f"{name}_CA_rp"], out[f"{name}_CA_ci"], out[f"{name}_CA_v"] = orp, oci, ov
        print(name, "equations", len(eqs), "constrained", len(cm[0]), "condensed nnz", ov.size)
    path = os.path.join(HERE, "constraints.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
