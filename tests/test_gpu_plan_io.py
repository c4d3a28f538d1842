"""Saved FMM plans: an operator rebuilt from save_plan() bytes gives the same outputs
bit for bit as the one that saved it (same octrees, task lists, C2E matrices and, in
mixed precision, the same modal factors), for one shared tree and for distinct trees;
plans for other inputs or configurations, corrupt or truncated plans, and plans saved
after the observation points moved are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def problem(tbem, seed, n_obs, n_src):
    rng = np.random.default_rng(seed)
    src = rng.uniform(0, 1, (n_src, 3))
    sn = rng.standard_normal((n_src, 3))
    sn /= np.linalg.norm(sn, axis=1)[:, None]
    w = rng.uniform(0.5, 1.5, n_src)
    obs = src if n_obs is None else rng.uniform(-0.2, 1.2, (n_obs, 3))
    on = sn if n_obs is None else sn[: len(obs)] if len(obs) <= n_src else np.tile(sn, (2, 1))[: len(obs)]
    return tbem.NBodyData(obs, on, src, sn, w), rng.standard_normal(3 * n_src)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("shared", [True, False])
def test_saved_plan_reproduces_operator(tbem, prec, shared):
    data, x = problem(tbem, 5, None if shared else 15_000, 20_000)
    K = tbem.ElasticDisplacement(1.0, 0.25)
    cfg = tbem.FMMConfig(0.3, 30, 64, 0.05, False)
    print(f"CASE plan io {prec} shared={shared}", flush=True)
    op = tbem.FMMOperator(K, data, cfg, precision=prec)
    blob = op.save_plan()
    assert blob[:8] == b"TBEMPLN1"
    op2 = tbem.FMMOperator(K, data, cfg, precision=prec, plan=blob)
    y, y2 = op.apply(x), op2.apply(x)
    assert np.array_equal(y.view(np.uint64), y2.view(np.uint64))
    i1, i2 = op.info(), op2.info()
    for k in ("n_e", "n_src_cells", "n_obs_cells", "p2p", "m2l", "m2p", "p2l", "l2l", "m2l_classes"):
        assert i1[k] == i2[k], k
    for which in ("src", "obs"):
        c1, p1 = op.octree(which)
        c2, p2 = op2.octree(which)
        assert np.array_equal(c1, c2) and np.array_equal(p1, p2)
    assert np.array_equal(op.c2e(False, 1), op2.c2e(False, 1))
    assert op2.save_plan() == blob


def test_saved_plan_refusals(tbem):
    data, x = problem(tbem, 6, None, 8_000)
    K = tbem.LaplaceSingle()
    cfg = tbem.FMMConfig(0.3, 10, 64, 0.05, False)
    print("CASE plan io refusals", flush=True)
    op = tbem.FMMOperator(K, data, cfg)
    blob = op.save_plan()
    w2 = data.src_weights.copy()
    w2[17] *= 1.0000001
    other = tbem.NBodyData(data.obs_locs, data.obs_normals, data.src_locs, data.src_normals, w2)
    for args, why in (((K, other, cfg), "points, normals or weights"),
                      ((K, data, tbem.FMMConfig(0.3, 12, 64, 0.05, False)), "FMMConfig"),
                      ((tbem.LaplaceDouble(), data, cfg), "FMMConfig")):
        with pytest.raises(tbem.TbemError, match=why):
            tbem.FMMOperator(*args, plan=blob)
    bad = bytearray(blob)
    bad[len(bad) // 2] ^= 0x40
    with pytest.raises(tbem.TbemError, match="checksum"):
        tbem.FMMOperator(K, data, cfg, plan=bytes(bad))
    with pytest.raises(tbem.TbemError, match="truncated"):
        tbem.FMMOperator(K, data, cfg, plan=blob[:-8])
    with pytest.raises(tbem.TbemError, match="magic"):
        tbem.FMMOperator(K, data, cfg, plan=b"NOTAPLAN" + blob[8:])
    op.set_obs_locs(data.obs_locs + 1e-3)
    with pytest.raises(tbem.TbemError, match="moved"):
        op.save_plan()
