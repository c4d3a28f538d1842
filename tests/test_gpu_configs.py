"""The benchmark configurations exercised on the GPU.

* Reference fixtures (tests/golden/make_config_golden.py, the compiled reference
  single-threaded): the cube-and-clusters Laplace workload (30,000 points, order 75,
  leaf 100), the rough-fault hypersingular workload (20,000 points, order 6) and the
  fault + free-surface elastic workload at order 50 (4,000 points, displacement and
  hypersingular kernels). At orders 50 and 75 the FMM output depends on the last
  bits of the check-to-equivalent pseudo-inverses (LAPACK's SVD there, a Jacobi SVD
  here; the method amplifies a 1e-16 change by ~1e4), so the FP64 path runs on the
  reference's own UC2E/DC2E, stored in the fixture, and must then match to 1e-8;
  with its own C2E it must be as close to the direct sum as the reference is.
* Order 100 on the fault + free-surface workload, against the compiled reference
  run live on the host (its C2E is ~6 MB per fixture at this order).
* Full sizes: the 4M-point cube-and-clusters workload (mixed and FP64 against each
  other and against the GPU direct sum on 10,000 targets), the 2M-point rough-fault
  hypersingular solve (FP64-residual refinement around the mixed FMM to a 1e-8
  relative residual) and the mixed MFMA path at order 100 on 200,000 points
  against the direct sum.
"""
import importlib

import numpy as np
import pytest

from conftest import golden, rel_err

pytestmark = pytest.mark.gpu


def case(*a):
    print("CASE", *a, flush=True)


def fixture_op(tbem, g, precision="fp64", inject=False, translations="auto"):
    K = tbem.make_kernel(int(g["kernel"]), 1.0, 0.25)
    data = tbem.NBodyData(g["obs"], g["obs_normals"], g["src"], g["src_normals"], g["weights"])
    cfg = tbem.FMMConfig(float(g["mac"]), int(g["order"]), int(g["leaf"]), float(g["d"]), bool(g["small"]))
    kw = dict(up_c2e=g["uc2e"], down_c2e=g["dc2e"]) if inject else {}
    return tbem.FMMOperator(K, data, cfg, precision, translations=translations, **kw)


@pytest.mark.parametrize("name", ["cube_clusters_laplace_o75", "fault_surface_elastic_u_o50",
                                  "fault_surface_elastic_h_o50"])
def test_high_order_fixture(tbem, name):
    g = golden(name)
    y_ref, direct = g["y"], g["direct"]
    e_ref = rel_err(y_ref, direct)
    case("config fixture", name, f"reference vs direct {e_ref:.2e}")
    op_inj = fixture_op(tbem, g, inject=True)
    assert op_inj.info()["m2l_classes"] > 0  # FP64 M2L on class operators
    e_inj = rel_err(op_inj.apply(g["x"]), y_ref)
    e_inj_surf = rel_err(fixture_op(tbem, g, inject=True, translations="surfaces").apply(g["x"]), y_ref)
    y_own = fixture_op(tbem, g).apply(g["x"])
    y_mix = fixture_op(tbem, g, "mixed").apply(g["x"])
    print(f"  fp64 (reference C2E) vs reference: class M2L {e_inj:.2e}, surface M2L {e_inj_surf:.2e}; "
          f"fp64 (own C2E) vs direct {rel_err(y_own, direct):.2e}; mixed vs direct {rel_err(y_mix, direct):.2e}, "
          f"vs fp64 {rel_err(y_mix, y_own):.2e}")
    assert e_inj < 1e-8 and e_inj_surf < 1e-8
    assert rel_err(y_own, direct) < 1.2 * e_ref + 1e-9
    assert rel_err(y_mix, direct) < 1.2 * e_ref + 1e-6
    assert rel_err(y_mix, y_own) < 1e-6


def test_rough_fault_hypersingular_fixture(tbem):
    g = golden("rough_fault_elastic_h_o6")
    case("config fixture rough fault hypersingular order 6")
    y = fixture_op(tbem, g).apply(g["x"])
    assert rel_err(y, g["y"]) < 1e-12
    assert rel_err(fixture_op(tbem, g, "mixed").apply(g["x"]), g["y"]) < 1e-5


@pytest.mark.parametrize("kid", [3, 6])
def test_order_100_against_live_reference(tbem, wl, ref, kid):
    """Elastic displacement and hypersingular kernels at order 100 (the metric's
    order) on 4,000 fault + free-surface points: FP64 on the reference's C2E against
    the compiled reference, and the mixed MFMA path against the direct sum."""
    g = wl.fault_and_free_surface(4_000)
    loc, nrm, w = g["locs"], g["normals"], g["weights"]
    case("order 100 live reference", kid)
    rop = ref.fmm(kid, loc, nrm, loc, nrm, w, 0.3, 100, 64, 0.05, False)
    y_ref = rop.apply(g["x"])
    c = rop.counts()
    up = np.stack([rop.c2e(False, L) for L in range(c["n_up_levels"])])
    down = np.stack([rop.c2e(True, L) for L in range(c["n_down_levels"])])
    direct = ref.direct(kid, loc, nrm, loc, nrm, w, g["x"])
    K = tbem.make_kernel(kid, 1.0, 0.25)
    data = tbem.NBodyData(loc, nrm, loc, nrm, w)
    cfg = tbem.FMMConfig(0.3, 100, 64, 0.05, False)
    y_inj = tbem.FMMOperator(K, data, cfg, up_c2e=up, down_c2e=down).apply(g["x"])
    y_inj_surf = tbem.FMMOperator(K, data, cfg, up_c2e=up, down_c2e=down, translations="surfaces").apply(g["x"])
    op_mix = tbem.FMMOperator(K, data, cfg, precision="mixed")
    assert op_mix.info()["m2l_classes"] > 0  # the MFMA M2L runs
    y_mix = op_mix.apply(g["x"])
    e_ref = rel_err(y_ref, direct)
    print(f"  reference vs direct {e_ref:.2e}; fp64 (reference C2E) vs reference: class M2L "
          f"{rel_err(y_inj, y_ref):.2e}, surface M2L {rel_err(y_inj_surf, y_ref):.2e}; "
          f"mixed vs direct {rel_err(y_mix, direct):.2e}")
    assert rel_err(y_inj, y_ref) < 1e-8 and rel_err(y_inj_surf, y_ref) < 1e-8
    assert rel_err(y_mix, direct) < 1.2 * e_ref + 1e-6


def subsample_direct(tbem, g, K, sub, x):
    dsub = tbem.NBodyData(g["locs"][sub], g["normals"][sub], g["locs"], g["normals"], g["weights"])
    return tbem.make_direct_nbody_operator(dsub, K).apply(x)


def test_cube_and_clusters_4M_mixed_and_fp64(tbem, wl):
    """The full cube-and-clusters workload: 4M points, Laplace single layer, order 75,
    leaf 100, MAC 0.3. Mixed (FP32 MFMA M2L) against FP64, and both against the FP64
    GPU direct sum on 10,000 seeded targets (the direct operator is itself pinned to
    the oracle in test_gpu_fmm.py)."""
    torch = pytest.importorskip("torch")
    n = 4_000_000
    g = wl.cube_and_clusters(n)
    K = tbem.LaplaceSingle()
    data = tbem.NBodyData(g["locs"], g["normals"], g["locs"], g["normals"], g["weights"])
    cfg = tbem.FMMConfig(0.3, 75, 100, 0.05, False)
    x = torch.tensor(g["x"], device="cuda")
    case("cube and clusters 4M")
    y64 = tbem.FMMOperator(K, data, cfg).apply(x).cpu().numpy()
    op32 = tbem.FMMOperator(K, data, cfg, precision="mixed")
    assert op32.info()["m2l_classes"] > 0
    y32 = op32.apply(x).cpu().numpy()
    del op32
    sub = np.sort(np.random.default_rng(321).choice(n, 10_000, replace=False))
    yd = subsample_direct(tbem, g, K, sub, g["x"])
    e64, e32, e_mix = rel_err(y64[sub], yd), rel_err(y32[sub], yd), rel_err(y32, y64)
    print(f"  fp64 vs direct {e64:.2e}, mixed vs direct {e32:.2e}, mixed vs fp64 {e_mix:.2e}")
    assert e64 < 1e-6 and e32 < 1e-6 and e_mix < 1e-6
    torch.cuda.empty_cache()


def test_rough_fault_2M_refinement_to_1e8(tbem, wl):
    """The full rough-fault solve: elastic hypersingular FMM on 2M points (order 50,
    leaf 128, MAC 0.3) as (alpha I + H) x = b, alpha = 2 ||H|| (power iterations),
    b the traction of a uniform unit remote stress; FP64-residual refinement around
    the mixed FMM reaches a 1e-8 relative FP64 residual, recomputed here."""
    torch = pytest.importorskip("torch")
    gm = importlib.import_module("3bem_amd.gmres")
    n = 2_000_000
    g = wl.rough_fault_points(n)
    K = tbem.ElasticHypersingular(1.0, 0.25)
    data = tbem.NBodyData(g["locs"], g["normals"], g["locs"], g["normals"], g["weights"])
    cfg = tbem.FMMConfig(0.3, 50, 128, 0.05, False)
    case("rough fault 2M refinement")
    op64 = tbem.FMMOperator(K, data, cfg)
    op32 = tbem.FMMOperator(K, data, cfg, precision="mixed")
    assert op32.info()["m2l_classes"] > 0
    v = torch.randn(3 * n, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(1))
    for _ in range(8):
        v = op32.apply(v)
        v = v / torch.linalg.norm(v)
    alpha = 2.0 * float(torch.linalg.norm(op64.apply(v)))
    b = torch.tensor(g["normals"].T.reshape(-1).copy(), device="cuda")
    assert rel_err(op32.apply(b).cpu().numpy(), op64.apply(b).cpu().numpy()) < 1e-6
    A64 = lambda x: alpha * x + op64.apply(x)
    A32 = lambda x: alpha * x + op32.apply(x)
    x, info = gm.refine(A64, A32, b, tol=1e-8, inner_tol=1e-4, restart=50)
    r = float(torch.linalg.norm(b - A64(x)) / torch.linalg.norm(b))
    print(f"  refinement: {info['outer']} outer, {info['inner_iterations']} inner, FP64 residual {r:.2e}")
    assert r <= 1e-8
    del op64, op32
    torch.cuda.empty_cache()


def test_mixed_mfma_order_100_200k_against_direct(tbem, wl):
    """The mixed path with the MFMA M2L (blk 297) at the metric's order on 200,000
    fault + free-surface points (leaf 256) against the FP64 direct sum on 5,000
    seeded targets, beside the FP64 FMM's own error."""
    torch = pytest.importorskip("torch")
    n = 200_000
    g = wl.fault_and_free_surface(n)
    K = tbem.ElasticDisplacement(1.0, 0.25)
    data = tbem.NBodyData(g["locs"], g["normals"], g["locs"], g["normals"], g["weights"])
    cfg = tbem.FMMConfig(0.3, 100, 256, 0.05, False)
    x = torch.tensor(g["x"], device="cuda")
    case("mixed MFMA order 100 200k")
    op32 = tbem.FMMOperator(K, data, cfg, precision="mixed")
    assert op32.info()["m2l_classes"] > 0
    y32 = op32.apply(x).cpu().numpy()
    y64 = tbem.FMMOperator(K, data, cfg).apply(x).cpu().numpy()
    sub = np.sort(np.random.default_rng(5).choice(n, 5_000, replace=False))
    idx = np.concatenate([q * n + sub for q in range(3)])
    yd = subsample_direct(tbem, g, K, sub, g["x"])
    e32, e64 = rel_err(y32[idx], yd), rel_err(y64[idx], yd)
    print(f"  mixed vs direct {e32:.2e}, fp64 vs direct {e64:.2e}")
    assert e32 < 1e-6 and e64 < 1e-6
