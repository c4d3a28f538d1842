"""GMRES and FP64-residual refinement (3bem_amd/gmres.py) on CPU tensors."""
import importlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")
gm = importlib.import_module("3bem_amd.gmres")


def dense_system(n, seed, shift=3.0):
    rng = np.random.default_rng(seed)
    M = np.eye(n) * shift + rng.standard_normal((n, n)) / np.sqrt(n)
    return M, rng.standard_normal(n)


@pytest.mark.parametrize("restart", [5, 20, 80])
def test_gmres_matches_dense_solve(restart):
    M, b = dense_system(80, 1)
    Mt = torch.tensor(M)
    x, info = gm.gmres(lambda v: Mt @ v, torch.tensor(b), tol=1e-13, restart=restart, maxiter=2000)
    ref = np.linalg.solve(M, b)
    assert info["converged"]
    assert np.linalg.norm(x.numpy() - ref) / np.linalg.norm(ref) < 1e-11
    # the reported residual is the true one
    assert abs(info["residual"] - np.linalg.norm(M @ x.numpy() - b) / np.linalg.norm(b)) < 1e-14


def test_gmres_full_krylov_space_solves_exactly():
    # restart = n: Krylov space of full dimension, one cycle suffices
    M, b = dense_system(30, 2, shift=0.5)
    x, info = gm.gmres(lambda v: torch.tensor(M) @ v, torch.tensor(b), tol=1e-12, restart=30, maxiter=30)
    assert info["converged"] and info["iterations"] <= 30


def test_gmres_zero_rhs_and_initial_guess():
    M, b = dense_system(20, 3)
    Mt = torch.tensor(M)
    x, info = gm.gmres(lambda v: Mt @ v, torch.zeros(20))
    assert info["converged"] and float(torch.linalg.norm(x)) == 0.0
    ref = torch.tensor(np.linalg.solve(M, b))
    x, info = gm.gmres(lambda v: Mt @ v, torch.tensor(b), x0=ref.clone(), tol=1e-10)
    assert info["iterations"] == 0 and info["converged"]


def test_refinement_recovers_fp64_accuracy_from_an_inexact_inner_operator():
    """Inner operator = the matrix with 1e-4 relative perturbations (a stand-in for a
    lower-precision matvec): plain GMRES on it stalls at ~1e-4; refinement with the
    exact residual converges to 1e-12."""
    M, b = dense_system(100, 5)
    rng = np.random.default_rng(6)
    Mp = M * (1 + 1e-4 * rng.standard_normal(M.shape))
    Mt, Mpt, bt = torch.tensor(M), torch.tensor(Mp), torch.tensor(b)
    x_in, _ = gm.gmres(lambda v: Mpt @ v, bt, tol=1e-13, restart=50)
    r_in = float(torch.linalg.norm(bt - Mt @ x_in) / torch.linalg.norm(bt))
    assert r_in > 1e-6
    x, info = gm.refine(lambda v: Mt @ v, lambda v: Mpt @ v, bt, tol=1e-12, inner_tol=1e-6, restart=50)
    assert info["converged"] and info["residual"] <= 1e-12
    assert info["outer"] <= 4
    ref = np.linalg.solve(M, b)
    assert np.linalg.norm(x.numpy() - ref) / np.linalg.norm(ref) < 1e-10


def test_refinement_with_oracle_fmm_as_the_fp64_operator(orc, wl):
    """The FP64 operator is the oracle FMM (CPU restatement of the reference): a
    second-kind system alpha x + A x solved with refinement around a perturbed copy."""
    g = wl.fault_and_free_surface(3000)
    op = orc.fmm(0, g["locs"], g["normals"], g["locs"], g["normals"], g["weights"], 0.3, 5, 50, 0.05)
    n = len(g["locs"])
    apply64 = lambda v: torch.tensor(op.apply(v.numpy()))
    v = torch.ones(n, dtype=torch.float64)
    for _ in range(10):
        v = apply64(v)
        v = v / torch.linalg.norm(v)
    alpha = 2.0 * float(torch.linalg.norm(apply64(v)))
    A = lambda x: alpha * x + apply64(x)
    rng = np.random.default_rng(8)
    scale = torch.tensor(1 + 1e-5 * rng.standard_normal(n))
    Ainexact = lambda x: alpha * x + scale * apply64(x)
    b = torch.tensor(g["x"][:n])
    x, info = gm.refine(A, Ainexact, b, tol=1e-10, inner_tol=1e-5, restart=30)
    assert info["converged"] and info["residual"] <= 1e-10
