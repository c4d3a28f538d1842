"""The FP64 path's M2L on per-class FP64 operators (v_mfma_f64_16x16x4f64).

By default the FP64 FMM translates each M2L pair with the FP64 operator of its
translation class (level pair, integer centre offset) instead of evaluating the
kernel between the two surfaces (translations="surfaces", the reference's way). Both are FP64
throughout and differ in rounding only, so they must agree far below the FMM's own
error, for every kernel, both directions (apply, apply_adjoint) and the observation
gradient. The reference fixtures of test_gpu_configs.py pin the class path against
the compiled reference itself.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def case(*a):
    print("CASE", *a, flush=True)


def cloud(n, seed):
    rng = np.random.default_rng(seed)
    locs = np.concatenate([rng.uniform(-1, 1, (n // 2, 3)), rng.normal(0.3, 0.2, (n - n // 2, 3))])
    nrm = rng.standard_normal((n, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    return locs, nrm, rng.uniform(0.5, 1.5, n)


def pair(tbem, kid, n, order, leaf, seed=0, obs=None):
    locs, nrm, w = cloud(n, seed)
    K = tbem.make_kernel(kid, 1.0, 0.25)
    if obs is None:
        data = tbem.NBodyData(locs, nrm, locs, nrm, w)
        kw = {}
    else:
        data = tbem.NBodyData(obs[0], obs[1], locs, nrm, w)
        kw = dict(tree_boxes="common")
    cfg = tbem.FMMConfig(0.3, order, leaf, 0.05, False)
    a = tbem.FMMOperator(K, data, cfg, translations="auto", **kw)
    b = tbem.FMMOperator(K, data, cfg, translations="surfaces", **kw)
    return K, data, a, b


# Orders at which blk = R * n_e >= 48 (the class path's lower bound), with the
# agreement required. The two paths round differently (summation order, and the
# class operator is built on the class's first pair's centres), and the
# check-to-equivalent pseudo-inverses amplify such last-bit differences more the
# higher the order (two FP64 implementations on the same C2E matrices differ by
# 7.9e-9 at order 100, test_gpu_configs.py); measured: <= 1e-12 at orders 16-20,
# 4e-11 at order 50 (Laplace), 2.4e-10 at order 40 (elastic). The direct-sum test
# below checks that the class path loses no accuracy.
@pytest.mark.parametrize("kid,order,tol", [(0, 50, 1e-9), (1, 50, 1e-9), (3, 16, 1e-10), (4, 20, 1e-10),
                                           (5, 20, 1e-10), (6, 16, 1e-10), (3, 40, 3e-9)])
def test_class_operators_match_surface_evaluation(tbem, kid, order, tol):
    case("fp64 class M2L vs surfaces", kid, order)
    K, data, a, b = pair(tbem, kid, 20_000, order, 64, seed=kid)
    assert a.info()["m2l_classes"] > 0, "the class path must run"
    assert b.info()["m2l_classes"] == 0
    rng = np.random.default_rng(1)
    x = rng.standard_normal(K.dim * data.n_src)
    g = rng.standard_normal(K.dim * data.n_obs)
    ya, yb = a.apply(x), b.apply(x)
    ta, tb = a.apply_adjoint(g), b.apply_adjoint(g)
    e_fwd, e_adj = rel_err(ya, yb), rel_err(ta, tb)
    dual = abs(g @ ya - ta @ x) / (np.linalg.norm(g) * np.linalg.norm(ya))
    print(f"  classes {a.info()['m2l_classes']:.0f}: forward {e_fwd:.2e}, adjoint {e_adj:.2e}, "
          f"<g,Ax> - <A^T g,x> {dual:.2e}")
    assert e_fwd < tol and e_adj < tol
    assert dual < 1e-12
    # deterministic: a second apply gives the same bits
    assert np.array_equal(a.apply(x), ya)


def test_class_operators_observation_gradient(tbem):
    case("fp64 class M2L observation gradient")
    K, data, a, b = pair(tbem, 3, 12_000, 20, 64, seed=5)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(3 * data.n_src)
    g = rng.standard_normal(3 * data.n_obs)
    ga, gb = a.obs_grad(x, g), b.obs_grad(x, g)
    e = rel_err(ga, gb)
    print(f"  gradient class vs surfaces {e:.2e}")
    assert e < 1e-9  # order 20


def test_class_operators_distinct_trees_on_common_box(tbem):
    """Distinct observation points on the common box: the FP64 class path serves
    them too, and agrees with the surface evaluation on the same trees."""
    case("fp64 class M2L distinct trees, common box")
    rng = np.random.default_rng(9)
    obs = rng.uniform(-1.2, 1.2, (9_000, 3))
    on = rng.standard_normal((9_000, 3))
    on /= np.linalg.norm(on, axis=1)[:, None]
    K, data, a, b = pair(tbem, 4, 15_000, 20, 64, seed=3, obs=(obs, on))
    assert a.info()["m2l_classes"] > 0
    x = rng.standard_normal(3 * data.n_src)
    g = rng.standard_normal(3 * data.n_obs)
    e_f = rel_err(a.apply(x), b.apply(x))
    e_a = rel_err(a.apply_adjoint(g), b.apply_adjoint(g))
    print(f"  forward {e_f:.2e}, adjoint {e_a:.2e}")
    assert e_f < 1e-10 and e_a < 1e-10


def test_class_operators_against_direct_sum(tbem):
    """Order 40 elastic on 30,000 points: the class path is as close to the FP64
    direct sum as the surface path."""
    case("fp64 class M2L vs direct sum")
    K, data, a, b = pair(tbem, 3, 30_000, 40, 128, seed=11)
    rng = np.random.default_rng(4)
    x = rng.standard_normal(3 * data.n_src)
    sub = np.sort(rng.choice(data.n_obs, 2_000, replace=False))
    dsub = tbem.NBodyData(data.obs_locs[sub], data.obs_normals[sub], data.src_locs, data.src_normals,
                          data.src_weights)
    yd = tbem.make_direct_nbody_operator(dsub, K).apply(x)
    idx = np.concatenate([q * data.n_obs + sub for q in range(3)])
    ea, eb = rel_err(a.apply(x)[idx], yd), rel_err(b.apply(x)[idx], yd)
    print(f"  class {ea:.3e}, surfaces {eb:.3e}")
    assert ea < 1.01 * eb + 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sliced_per_pair_output(tbem, precision):
    """A small m2l_output_budget runs the class M2L in slices of whole target cells:
    the forward is the same bit for bit (each target's pairs stay in one slice and are
    summed in the same order); the transpose adds its per-source partial sums slice
    by slice, so it agrees to rounding, which the transposed C2E products then
    amplify (measured 7.8e-13 at order 20)."""
    case("sliced class M2L", precision)
    locs, nrm, w = cloud(20_000, 21)
    K = tbem.make_kernel(3, 1.0, 0.25)
    data = tbem.NBodyData(locs, nrm, locs, nrm, w)
    cfg = tbem.FMMConfig(0.3, 20, 64, 0.05, False)
    whole = tbem.FMMOperator(K, data, cfg, precision)
    sliced = tbem.FMMOperator(K, data, cfg, precision, m2l_output_budget=1 << 20)
    bw, bs = whole.info()["m2l_output_bytes"], sliced.info()["m2l_output_bytes"]
    print(f"  per-pair output {bw / 1e6:.1f} MB whole, {bs / 1e6:.2f} MB sliced")
    assert whole.info()["m2l_classes"] > 0 and bs <= (1 << 20) < bw
    rng = np.random.default_rng(6)
    x = rng.standard_normal(3 * data.n_src)
    g = rng.standard_normal(3 * data.n_obs)
    assert np.array_equal(sliced.apply(x), whole.apply(x))
    e = rel_err(sliced.apply_adjoint(g), whole.apply_adjoint(g))
    print(f"  transpose sliced vs whole {e:.2e}")
    assert e < 1e-11
