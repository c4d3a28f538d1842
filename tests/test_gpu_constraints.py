"""Constraints on the GPU: distribute_vector and condense_vector equal the compiled
reference's vectors bit for bit (fixture tests/golden/constraints.npz) and the
reference's own unit-test values (unit_tests/test_constraint_matrix.cpp); the sparse
condense_matrix equals the reference's CSR arrays; the constrained GMRES driver of
acceptance_tests/dislocation.py:27-39 solves a constrained system to its tolerance."""
import importlib

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

C = importlib.import_module("3bem_amd.constraints")
CASES = ("grid_all", "grid_chain", "mixed")


def fixture_cm(g, name):
    ptr, dof, w, rhs = (g[f"{name}_eq_{k}"] for k in ("ptr", "dof", "w", "rhs"))
    eqs = [C.ConstraintEQ([(int(dof[k]), float(w[k])) for k in range(ptr[e], ptr[e + 1])], float(rhs[e]))
           for e in range(len(rhs))]
    return C.from_constraints(eqs)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("name", CASES)
def test_distribute_condense_bit_exact_vs_reference(tbem, name):
    g = golden("constraints")
    print(f"CASE constraints {name}", flush=True)
    cm = fixture_cm(g, name)
    n = int(g[f"{name}_n_total"])
    assert np.array_equal(bits(C.distribute_vector(cm, g[f"{name}_red"], n)), bits(g[f"{name}_distributed"]))
    assert np.array_equal(bits(C.condense_vector(cm, g[f"{name}_full"])), bits(g[f"{name}_condensed"]))
    A = tbem.SparseOperator(n, n, g[f"{name}_A_v"], g[f"{name}_A_ci"], g[f"{name}_A_rp"])
    CA = C.condense_matrix(cm, cm, A)
    assert np.array_equal(CA.row_ptrs, g[f"{name}_CA_rp"])
    assert np.array_equal(CA.column_indices, g[f"{name}_CA_ci"])
    assert np.array_equal(bits(CA.values), bits(g[f"{name}_CA_v"]))


def test_distribute_condense_torch_and_oracle(tbem, orc):
    torch = pytest.importorskip("torch")
    wl = importlib.import_module("3bem_amd.workloads")
    ep, td, tw, rhs, n = wl.grid_continuity_constraints(20, 15, 3, all_pairs=True, bc_value=1.25)
    eqs = [C.ConstraintEQ([(int(td[k]), float(tw[k])) for k in range(ep[e], ep[e + 1])], float(rhs[e]))
           for e in range(len(rhs))]
    print(f"CASE constraints torch, {len(eqs)} equations", flush=True)
    cm = C.from_constraints(eqs)
    r = cm.rearranged()
    cd = np.array(sorted(r), dtype=np.int64)
    solved = (cd, np.cumsum([0] + [len(r[c][0]) for c in cd]), np.array([d for c in cd for d, _ in r[c][0]]),
              np.array([w for c in cd for _, w in r[c][0]]), np.array([r[c][1] for c in cd]))
    rng = np.random.default_rng(4)
    red, full = rng.standard_normal(n - cm.size()), rng.standard_normal(n)
    d_t = C.distribute_vector(cm, torch.tensor(red, device="cuda"), n).cpu().numpy()
    c_t = C.condense_vector(cm, torch.tensor(full, device="cuda")).cpu().numpy()
    assert np.array_equal(bits(d_t), bits(orc.distribute(solved, red, n)))
    assert np.array_equal(bits(c_t), bits(orc.condense(solved, full)))
    assert np.array_equal(bits(C.distribute_vector(cm, red, n)), bits(d_t))


def test_reference_unit_values_on_gpu(tbem):
    cont, bc = C.continuity_constraint, C.boundary_condition
    assert C.distribute_vector(C.from_constraints([cont(0, 1)]), [1.0], 2).tolist() == [1.0, 1.0]
    assert C.distribute_vector(C.from_constraints([cont(0, 1), cont(2, 1), cont(2, 3)]), [1.0], 4).tolist() == [1.0] * 4
    abc = C.ConstraintEQ([(0, 1.0), (1, -1.0), (2, -1.0)], 0.0)
    assert C.distribute_vector(C.from_constraints([abc]), [1.0, 1.0], 3).tolist() == [1.0, 1.0, 0.0]
    assert C.distribute_vector(C.from_constraints([abc, cont(2, 3)]), [1.0, 0.5], 4).tolist() == [1.0, 0.5, 0.5, 0.5]
    assert C.condense_vector(C.from_constraints([cont(0, 2)]), [0.0, -1.0, 0.0]).tolist() == [0.0, -1.0]
    assert C.condense_vector(C.from_constraints([bc(0, 4.0)]), [2.0]).size == 0
    cm = C.from_constraints([bc(1, 4.0), cont(0, 2), cont(2, 3), cont(3, 4)])
    assert C.condense_vector(cm, [0.0, 0.0, 0.0, 0.0, 4.0]).tolist() == [4.0]
    red = C.condense_vector(cm, [2.0, 4.0, 4.0, 4.0, 4.0])
    assert red.tolist() == [14.0]
    assert np.allclose(C.distribute_vector(cm, red, 5), [14.0, 4.0, 14.0, 14.0, 14.0], rtol=0, atol=1e-13)
    cm2 = C.from_constraints([bc(2, 2.0), cont(0, 2), cont(2, 3), cont(3, 4)])
    assert C.condense_vector(cm2, [2.0, 4.0, 4.0, 4.0, 4.0]).tolist() == [4.0]
    avg = [cont(0, 1), cont(2, 3), cont(4, 5), cont(6, 7), C.ConstraintEQ([(i, 1.0) for i in range(8)], 0.0)]
    assert abs(C.distribute_vector(C.from_constraints(avg), [1.0, 1.0, 1.0], 8).sum()) < 1e-12
    s = tbem.SparseOperator.csr_from_coo(3, 3, [0, 1, 2], [0, 1, 2], [1.0, 1.0, 1.0])
    cm3 = C.from_constraints([cont(0, 2), cont(1, 2)])
    r = C.condense_matrix(cm3, cm3, s)
    assert r.nnz() == 1 and r.values.tolist() == [3.0] and r.column_indices.tolist() == [0]
    big = tbem.SparseOperator.csr_from_coo(5, 5, [0, 0, 1, 2, 2, 4, 4, 3], [0, 2, 1, 2, 0, 0, 3, 4],
                                           [1, 4, 1, 1, 1, -1, 1, 1])
    rb = C.condense_matrix(C.from_constraints([cont(1, 2)]), C.from_constraints([cont(1, 2)]), big)
    want = tbem.SparseOperator.csr_from_coo(4, 4, [0, 0, 1, 1, 3, 3, 2], [0, 1, 0, 1, 0, 2, 3],
                                            [1, 4, 1, 2, -1, 1, 1])
    assert rb.values.tolist() == want.values.tolist()
    assert rb.column_indices.tolist() == want.column_indices.tolist()
    assert rb.row_ptrs.tolist() == want.row_ptrs.tolist()


def test_constrained_gmres_solves_the_condensed_system(tbem):
    """(alpha I + S) on the dof points of a 6 x 5 sheet (S the Laplace single-layer
    direct operator), continuity and boundary constraints, b = condense(f): the driver's
    solution matches a dense solve of the condensed system and meets its tolerance."""
    torch = pytest.importorskip("torch")
    gm = importlib.import_module("3bem_amd.gmres")
    wl = importlib.import_module("3bem_amd.workloads")
    facets = wl.rough_fault_mesh(6, 5)
    pts = facets.reshape(-1, 3) + 1e-3 * np.random.default_rng(2).standard_normal((facets.shape[0] * 3, 3))
    n = pts.shape[0]
    K = tbem.LaplaceSingle()
    S = tbem.DirectOperator(tbem.NBodyData(pts, np.zeros((n, 3)), pts, np.zeros((n, 3)), np.full(n, 1.0 / n)), K)
    ep, td, tw, rhs, n_total = wl.grid_continuity_constraints(6, 5, 1, all_pairs=False, bc_value=0.3)
    assert n_total == n
    eqs = [C.ConstraintEQ([(int(td[k]), float(tw[k])) for k in range(ep[e], ep[e + 1])], float(rhs[e]))
           for e in range(len(rhs))]
    cm = C.from_constraints(eqs)
    hom = C.homogenize_constraints(cm)
    alpha = 2.0
    print("CASE constrained gmres", flush=True)

    def A(v):
        return alpha * v + S.apply(v)
    f = torch.tensor(np.random.default_rng(3).standard_normal(n), device="cuda")
    # inhomogeneous constraints: solve for the homogeneous part around distribute(0)
    x0 = C.distribute_vector(cm, torch.zeros(n - cm.size(), dtype=torch.float64, device="cuda"), n)
    b = C.condense_vector(hom, (f - A(x0)).contiguous())
    x, info = gm.solve_constrained(A, hom, b, n, tol=1e-12, restart=60, maxiter=600)
    assert info["converged"], info["residual"]
    mv = gm.constrained_operator(A, hom, n)
    m = n - cm.size()
    D = np.stack([mv(torch.tensor(np.eye(m)[j], device="cuda")).cpu().numpy() for j in range(m)], axis=1)
    x_dense = np.linalg.solve(D, b.cpu().numpy())
    assert np.linalg.norm(x.cpu().numpy() - x_dense) <= 1e-10 * np.linalg.norm(x_dense)
    full = C.distribute_vector(hom, x, n) + x0
    fv = full.cpu().numpy()
    two = [e for e in range(len(rhs)) if ep[e + 1] - ep[e] == 2]
    assert max(abs(fv[td[ep[e]]] - fv[td[ep[e] + 1]]) for e in two) < 1e-12  # continuity holds
    one = [td[ep[e]] for e in range(len(rhs)) if ep[e + 1] - ep[e] == 1]
    assert np.allclose(fv[one], 0.3, rtol=0, atol=1e-14)  # boundary values hold


def test_constraints_edge_cases(tbem):
    """No constraints (distribute and condense are the identity), every dof
    constrained (empty reduced vector), and refusals for lengths that do not fit."""
    print("CASE constraints edge cases", flush=True)
    none = C.from_constraints([])
    v = np.arange(5.0)
    assert C.distribute_vector(none, v, 5).tolist() == v.tolist()
    assert C.condense_vector(none, v).tolist() == v.tolist()
    allc = C.from_constraints([C.boundary_condition(d, 1.0 + d) for d in range(3)])
    assert C.distribute_vector(allc, np.zeros(0), 3).tolist() == [1.0, 2.0, 3.0]
    assert C.condense_vector(allc, [5.0, 6.0, 7.0]).size == 0
    with pytest.raises(ValueError):
        C.distribute_vector(allc, np.zeros(1), 3)
    with pytest.raises(ValueError):
        C.condense_vector(allc, [1.0, 2.0])  # fewer values than constraints
    far = C.from_constraints([C.boundary_condition(0, 1.0), C.boundary_condition(5, 1.0)])
    with pytest.raises(tbem.TbemError, match="beyond"):
        C.condense_vector(far, [1.0, 2.0, 3.0])  # a constrained dof beyond the vector
