"""tbem_nbody_apply_noalloc, the direct operator's apply that allocates nothing,
called through the C ABI: equal bit for bit to tbem_nbody_apply_dev on two calls
in a row (the second with fewer points, on the same static scratch), and refused
for more source densities than its scratch holds."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def handle(tbem, lib, kid, n_obs, n_src, seed):
    rng = np.random.default_rng(seed)
    nrm = lambda m: (lambda v: v / np.linalg.norm(v, axis=1)[:, None])(rng.standard_normal((m, 3)))
    obs, src = rng.uniform(0, 1, (n_obs, 3)), rng.uniform(0, 1, (n_src, 3))
    data = tbem.NBodyData(obs, nrm(n_obs), src, nrm(n_src), rng.uniform(0.5, 1.5, n_src))
    op = tbem.make_direct_nbody_operator(data, tbem.make_kernel(kid, 1.0, 0.25))
    return op, rng


def run(lib, fn, op, x, y, stream):
    rc = fn(op._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), stream)
    if rc != 0:
        raise RuntimeError(lib.tbem_last_error().decode())


@pytest.mark.parametrize("kid", [3, 0])
def test_noalloc_matches_apply_dev(tbem, kid):
    torch = pytest.importorskip("torch")
    from importlib import import_module
    lib = import_module("3bem_amd._lib").require_device()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    R = 3 if kid == 3 else 1
    print(f"CASE noalloc kid={kid}", flush=True)
    for n_obs, n_src, seed in ((30_000, 40_000, 1), (7_000, 9_001, 2)):  # second call: smaller input
        op, rng = handle(tbem, lib, kid, n_obs, n_src, seed)
        x = torch.tensor(rng.standard_normal(R * n_src), device="cuda")
        y_dev = torch.full((R * n_obs,), np.nan, dtype=torch.float64, device="cuda")
        y_na = torch.full((R * n_obs,), np.nan, dtype=torch.float64, device="cuda")
        run(lib, lib.tbem_nbody_apply_dev, op, x, y_dev, stream)
        run(lib, lib.tbem_nbody_apply_noalloc, op, x, y_na, stream)
        torch.cuda.synchronize()
        a, b = y_dev.cpu().numpy(), y_na.cpu().numpy()
        assert np.all(np.isfinite(a))
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (n_obs, n_src)


def test_noalloc_refuses_more_than_its_scratch(tbem):
    torch = pytest.importorskip("torch")
    from importlib import import_module
    lib = import_module("3bem_amd._lib").require_device()
    n_src = (4 << 20) // 3 + 1  # 3 * n_src = 4,194,306 > 4,194,304 doubles
    print("CASE noalloc refusal", flush=True)
    op, rng = handle(tbem, lib, 3, 10, n_src, 3)
    x = torch.zeros(3 * n_src, dtype=torch.float64, device="cuda")
    y = torch.zeros(30, dtype=torch.float64, device="cuda")
    rc = lib.tbem_nbody_apply_noalloc(op._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc != 0 and b"scratch" in lib.tbem_last_error()
