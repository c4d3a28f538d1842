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
    if prec == "mixed":
        assert i1["m2l_classes"] > 0  # distinct trees too: both on the union box
    if not shared:
        # a plan's trees were laid out for its precision's boxes: the other layout refuses it
        other = "fp64" if prec == "mixed" else "mixed"
        with pytest.raises(tbem.TbemError, match="other boxes"):
            tbem.FMMOperator(K, data, cfg, precision=other, plan=blob)
        op3 = tbem.FMMOperator(K, data, cfg, precision=other, tree_boxes="common" if prec == "mixed" else "own",
                               plan=blob)
        assert op3.info()["n_obs_cells"] == i1["n_obs_cells"] and op3.info()["p2p"] == i1["p2p"]


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
    op.set_obs_locs(data.obs_locs)  # a move, even to the same coordinates
    with pytest.raises(tbem.TbemError, match="moved"):
        op.save_plan()


M64 = (1 << 64) - 1


def _rotl(v, r):
    return ((v << r) | (v >> (64 - r))) & M64


def _mix_word(h, w):
    w = (w * 0xC2B2AE3D27D4EB4F) & M64
    w = _rotl(w, 31)
    x = h ^ ((w * 0x9E3779B185EBCA87) & M64)
    return (_rotl(x, 27) * 5 + 0x52DCE729) & M64


def plan_hash(b, h=0x9E3779B97F4A7C15):
    """The saved plan's 64-bit payload checksum (four multiply-rotate lanes per 32
    bytes, then the tail words and bytes, then a final avalanche), written out here
    so that a test can re-seal a payload it has edited."""
    n = len(b)
    words = np.frombuffer(b[: n - n % 8], dtype="<u8").tolist()
    lane = [h, h + 1, h + 2, h + 3]
    i = 0
    while 4 * (i + 1) <= len(words) and 32 * (i + 1) <= n:
        for k in range(4):
            lane[k] = _mix_word(lane[k], words[4 * i + k])
        i += 1
    h = _mix_word(_mix_word(_mix_word(_mix_word(h ^ n, lane[0]), lane[1]), lane[2]), lane[3])
    for w in words[4 * i:]:
        h = _mix_word(h, w)
    for c in b[len(words) * 8:]:
        h = ((h ^ c) * 0x100000001B3) & M64
    h ^= h >> 33
    h = (h * 0xFF51AFD7ED558CCD) & M64
    h ^= h >> 33
    return h


def test_saved_plan_with_bad_parent_link_is_refused(tbem):
    """A payload whose checksum is right but whose octree names a wrong parent for
    one cell (parent -1 below the root, or a cell of the same level) is refused while
    the plan is read on the host; nothing is applied."""
    data, _ = problem(tbem, 7, None, 6_000)
    K = tbem.LaplaceSingle()
    cfg = tbem.FMMConfig(0.3, 8, 64, 0.05, False)
    print("CASE plan io bad parent", flush=True)
    blob = tbem.FMMOperator(K, data, cfg).save_plan()
    payload = bytearray(blob[24:])
    assert plan_hash(bytes(payload)) == int.from_bytes(blob[16:24], "little")  # the checksum as written
    off = 84 + 12  # header fields, then n_pts, n_cells, n_levels of the source tree
    vecs = []
    for _ in range(7):  # level, begin, count, first_child, n_child, parent, octant
        cnt = int.from_bytes(payload[off:off + 8], "little")
        vecs.append((off + 8, cnt))
        off += 8 + 4 * cnt
    (lvl_off, n_cells), (par_off, _) = vecs[0], vecs[5]
    level = np.frombuffer(bytes(payload[lvl_off:lvl_off + 4 * n_cells]), dtype="<i4")
    parent = np.frombuffer(bytes(payload[par_off:par_off + 4 * n_cells]), dtype="<i4")
    c = int(np.flatnonzero(level == 2)[0])
    same_level = int(np.flatnonzero(level == 2)[1])
    for bad in (-1, same_level, int(parent[c]) + 1 if level[int(parent[c]) + 1] == 1 else 0):
        if bad == parent[c]:
            continue
        p = bytearray(payload)
        p[par_off + 4 * c:par_off + 4 * c + 4] = int(bad).to_bytes(4, "little", signed=True)
        sealed = blob[:16] + plan_hash(bytes(p)).to_bytes(8, "little") + bytes(p)
        with pytest.raises(tbem.TbemError, match="parent and child links"):
            tbem.FMMOperator(K, data, cfg, plan=sealed)


def test_threaded_plan_build_equals_one_thread(tbem, wl, monkeypatch):
    """The plan's host work (the dual-tree traversal, the task lists and their merged
    point ranges) runs on host threads. With one thread it is the plain recursion and
    plain loops; the threaded build must produce the same task lists in the same
    order: byte-identical saved plans and bit-identical outputs, with the MFMA M2L and
    the adjoint lists (built on first use) in play."""
    torch = pytest.importorskip("torch")
    g = wl.cube_and_clusters(300_000)
    data = tbem.NBodyData(g["locs"], g["normals"], g["locs"], g["normals"], g["weights"])
    cfg = tbem.FMMConfig(0.3, 50, 64, 0.05, False)
    x = torch.tensor(g["x"], device="cuda")
    print("CASE threaded plan build", flush=True)
    out = {}
    for threads in ("1", "16"):
        monkeypatch.setenv("OMP_NUM_THREADS", threads)
        op = tbem.FMMOperator(tbem.LaplaceSingle(), data, cfg, precision="mixed")
        info = op.info()
        assert info["m2l"] > (1 << 20) and info["m2l_classes"] > 0  # long lists; the MFMA M2L runs
        out[threads] = (op.save_plan(), op.apply(x).cpu().numpy(), op.apply_adjoint(x).cpu().numpy())
        del op
    assert out["1"][0] == out["16"][0]
    for k in (1, 2):
        assert np.array_equal(out["1"][k].view(np.uint64), out["16"][k].view(np.uint64))
