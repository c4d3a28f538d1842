"""GMRES on the GPU with the FMM inside, and FP64-residual iterative refinement
around the mixed-precision FMM (the solve loop of acceptance_tests/dislocation.py:27-39,
there scipy's GMRES on host vectors).

Operator: the elastic hypersingular FMM on the rough-fault geometry, made second
kind, A x = alpha x + H x with alpha = 2 ||H|| (power-iteration estimate), which is
how a collocation system with a jump term looks to the solver; right-hand side:
the traction sigma.n of a uniform unit remote stress. Reference solution: plain
FP64 GMRES with the FP64 FMM to 1e-12. The configuration asks for a 1e-8 relative
residual; refinement is run to 1e-11 here so that it is clearly below where plain
GMRES on the mixed operator stalls (~3e-9)."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system(tbem, wl):
    torch = pytest.importorskip("torch")
    g = wl.rough_fault_points(20000)
    n = len(g["locs"])
    K = tbem.ElasticHypersingular(1.0, 0.25)
    data = tbem.NBodyData(g["locs"], g["normals"], g["locs"], g["normals"], g["weights"])
    cfg = tbem.FMMConfig(0.3, 50, 64, 0.05, False)
    op64 = tbem.FMMOperator(K, data, cfg, precision="fp64")
    op32 = tbem.FMMOperator(K, data, cfg, precision="mixed")
    assert op32.info()["m2l_classes"] > 0
    v = torch.randn(3 * n, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(1))
    for _ in range(20):
        v = op64.apply(v)
        v = v / torch.linalg.norm(v)
    alpha = 2.0 * float(torch.linalg.norm(op64.apply(v)))
    b = torch.tensor(g["normals"].T.reshape(-1).copy(), device="cuda")
    return dict(torch=torch, op64=op64, op32=op32, alpha=alpha, b=b, n=n)


def test_gmres_fp64_converges(system):
    gm = importlib.import_module("3bem_amd.gmres")
    s = system
    A = lambda x: s["alpha"] * x + s["op64"].apply(x)
    print("CASE gmres fp64", flush=True)
    x, info = gm.gmres(A, s["b"], tol=1e-12, restart=40, maxiter=400)
    r = float(s["torch"].linalg.norm(s["b"] - A(x)) / s["torch"].linalg.norm(s["b"]))
    print(f"  fp64 GMRES: {info['iterations']} iterations, true residual {r:.2e}")
    assert info["converged"] and r < 1e-11
    s["x64"] = x


def test_refinement_with_mixed_fmm_reaches_fp64_residual(system):
    gm = importlib.import_module("3bem_amd.gmres")
    s = system
    torch = s["torch"]
    A64 = lambda x: s["alpha"] * x + s["op64"].apply(x)
    A32 = lambda x: s["alpha"] * x + s["op32"].apply(x)
    print("CASE refinement mixed", flush=True)
    x, info = gm.refine(A64, A32, s["b"], tol=1e-11, inner_tol=1e-4, restart=40)
    r = float(torch.linalg.norm(s["b"] - A64(x)) / torch.linalg.norm(s["b"]))
    print(f"  refinement: {info['outer']} outer steps, {info['inner_iterations']} inner iterations, "
          f"FP64 residual {r:.2e}, history {['%.1e' % h for h in info['history']]}")
    assert info["converged"] and r <= 1e-11
    # each outer step gains about max(inner tolerance, operator mismatch ~1e-7 relative)
    assert info["outer"] <= 4
    if "x64" in s:
        err = float(torch.linalg.norm(x - s["x64"]) / torch.linalg.norm(s["x64"]))
        print(f"  vs FP64 GMRES solution: {err:.2e}")
        assert err < 1e-9
    # the mixed operator alone stalls at its own accuracy: plain GMRES on it, however
    # tight its own tolerance, leaves an FP64 residual of the operator mismatch
    # (this is what the refinement loop is for)
    x32, _ = gm.gmres(A32, s["b"], tol=1e-13, restart=40, maxiter=300)
    r32 = float(torch.linalg.norm(s["b"] - A64(x32)) / torch.linalg.norm(s["b"]))
    print(f"  plain GMRES on the mixed operator: FP64 residual {r32:.2e}")
    assert r32 > 1e-10


def test_gmres_small_dense_against_numpy(system):
    """GMRES itself (restarts, Givens rotations) on a small nonsymmetric dense system."""
    gm = importlib.import_module("3bem_amd.gmres")
    torch = system["torch"]
    rng = np.random.default_rng(4)
    M = np.eye(60) * 3 + rng.standard_normal((60, 60)) / np.sqrt(60)
    rhs = rng.standard_normal(60)
    Mt = torch.tensor(M, device="cuda")
    x, info = gm.gmres(lambda v: Mt @ v, torch.tensor(rhs, device="cuda"), tol=1e-13, restart=7, maxiter=500)
    assert info["converged"]
    assert np.linalg.norm(x.cpu().numpy() - np.linalg.solve(M, rhs)) / np.linalg.norm(np.linalg.solve(M, rhs)) < 1e-11
