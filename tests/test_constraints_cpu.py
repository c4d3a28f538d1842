"""Constraints on the CPU: the library's from_constraints (host code, no device call)
against the compiled reference (fixture tests/golden/constraints.npz from
tests/golden/make_constraints_golden.py, and live where oracle/_ref is built) and
the reference's own unit-test cases (unit_tests/test_constraint_matrix.cpp); the
oracle's distribute/condense restatement against the same fixture and cases."""
import importlib

import numpy as np
import pytest

from conftest import golden

C = importlib.import_module("3bem_amd.constraints")
CASES = ("grid_all", "grid_chain", "mixed")


def eqs_from_fixture(g, name):
    ptr, dof, w, rhs = (g[f"{name}_eq_{k}"] for k in ("ptr", "dof", "w", "rhs"))
    return [C.ConstraintEQ([(int(dof[k]), float(w[k])) for k in range(ptr[e], ptr[e + 1])], float(rhs[e]))
            for e in range(len(rhs))]


def solved_arrays(cm):
    r = cm.rearranged()
    cd = np.array(sorted(r), dtype=np.int64)
    tp = np.cumsum([0] + [len(r[c][0]) for c in cd]).astype(np.int64)
    td = np.array([d for c in cd for d, _ in r[c][0]], dtype=np.int64)
    tw = np.array([w for c in cd for _, w in r[c][0]])
    tr = np.array([r[c][1] for c in cd])
    return cd, tp, td, tw, tr


@pytest.mark.parametrize("name", CASES)
def test_from_constraints_matches_reference_fixture(name):
    g = golden("constraints")
    got = solved_arrays(C.from_constraints(eqs_from_fixture(g, name)))
    for k, a in zip(("cdof", "tptr", "tdof", "tw", "trhs"), got):
        ref = g[f"{name}_{k}"]
        assert a.shape == ref.shape and np.array_equal(a.view(np.uint64) if a.dtype == float else a,
                                                       ref.view(np.uint64) if ref.dtype == float else ref), k


def test_from_constraints_reference_unit_cases():
    cont, bc = C.continuity_constraint, C.boundary_condition
    assert C.from_constraints([cont(0, 1), cont(1, 2)]).size() == 2
    z = C.from_constraints([C.ConstraintEQ([(0, 1.0), (1, 0.0)], 0.0)]).rearranged()
    assert z[0][0] == []
    assert C.from_constraints([C.ConstraintEQ([(1, 0.0)], 0.0)]).size() == 0
    assert C.from_constraints([cont(0, 1), cont(1, 0)]).size() == 1
    h = C.homogenize_constraints(C.from_constraints([bc(1, 2.0), cont(0, 2), cont(2, 3), cont(3, 4)]))
    assert all(r == 0.0 for _, r in h.rearranged().values())
    assert [(t.dof, t.weight) for t in C.shift_constraints([cont(0, 1)], 5)[0].terms] == [(5, 1.0), (6, -1.0)]


def test_from_constraints_matches_live_reference(ref):
    """Random one- and two-term equations with random weights and right-hand sides
    (every solved constraint then keeps at most one term, which stays inside what the
    reference's substitution defines), repeated dofs and cycles included."""
    rng = np.random.default_rng(23)
    for trial in range(6):
        n = 60
        eqs = []
        for _ in range(90):
            k = 1 if rng.uniform() < 0.15 else 2
            dofs = rng.choice(n, k, replace=False)
            w = rng.uniform(-2, 2, k)
            eqs.append(([(int(d), float(x)) for d, x in zip(dofs, w)], float(rng.uniform(-1, 1))))
        want = ref.from_constraints(eqs)
        got = solved_arrays(C.from_constraints([C.ConstraintEQ(t, r) for t, r in eqs]))
        for a, b in zip(got, want):
            a, b = np.asarray(a), np.asarray(b)
            assert a.shape == b.shape and np.array_equal(a, b)


@pytest.mark.parametrize("name", CASES)
def test_oracle_distribute_condense_match_reference_fixture(orc, name):
    g = golden("constraints")
    cm = tuple(g[f"{name}_{k}"] for k in ("cdof", "tptr", "tdof", "tw", "trhs"))
    n = int(g[f"{name}_n_total"])
    assert np.array_equal(orc.distribute(cm, g[f"{name}_red"], n).view(np.uint64),
                          g[f"{name}_distributed"].view(np.uint64))
    assert np.array_equal(orc.condense(cm, g[f"{name}_full"]).view(np.uint64), g[f"{name}_condensed"].view(np.uint64))


def test_oracle_reference_unit_values(ref, orc):
    """CondenseRecurse, CondenseThenDistribute, CondenseWithFullyDeterminedSubset,
    AEqualsBPlusCAndCEqualsD and the field-average case of the reference's tests."""
    cont = lambda a, b: ([(a, 1.0), (b, -1.0)], 0.0)  # noqa: E731
    bc = lambda d, v: ([(d, 1.0)], v)  # noqa: E731
    cm = ref.from_constraints([bc(1, 4.0), cont(0, 2), cont(2, 3), cont(3, 4)])
    assert orc.condense(cm, [0.0, 0.0, 0.0, 0.0, 4.0]).tolist() == [4.0]
    red = orc.condense(cm, [2.0, 4.0, 4.0, 4.0, 4.0])
    assert red.tolist() == [14.0]
    assert np.allclose(orc.distribute(cm, red, 5), [14.0, 4.0, 14.0, 14.0, 14.0], rtol=0, atol=1e-13)
    cm2 = ref.from_constraints([bc(2, 2.0), cont(0, 2), cont(2, 3), cont(3, 4)])
    assert orc.condense(cm2, [2.0, 4.0, 4.0, 4.0, 4.0]).tolist() == [4.0]
    cm3 = ref.from_constraints([([(0, 1.0), (1, -1.0), (2, -1.0)], 0.0), cont(2, 3)])
    assert orc.distribute(cm3, [1.0, 0.5], 4).tolist() == [1.0, 0.5, 0.5, 0.5]
    avg = [cont(0, 1), cont(2, 3), cont(4, 5), cont(6, 7), ([(i, 1.0) for i in range(8)], 0.0)]
    v = orc.distribute(ref.from_constraints(avg), [1.0, 1.0, 1.0], 8)
    assert abs(v.sum()) < 1e-12
