"""Differentiable forward model (slip -> surface displacement through the FMM,
elastic traction / displacement kernels): torch autograd gradients against
frozen-plan central differences and against the FP64 direct-sum gradient."""
import importlib

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def make_problem(tbem, wl, K, order=20):
    torch = pytest.importorskip("torch")
    g = wl.fault_and_stations(n_tri=400, n_obs=1500)
    data = tbem.NBodyData(g["obs"], g["obs_normals"], g["src"], g["src_normals"], g["src_weights"])
    op = tbem.FMMOperator(K, data, tbem.FMMConfig(0.3, order, 32, 0.05, False))
    ref_u = tbem.FMMOperator(K, data, tbem.FMMConfig(0.3, order, 32, 0.05, False)).apply(g["x_true"])
    return dict(g=g, K=K, op=op, u_obs=torch.tensor(ref_u, device="cuda"), torch=torch)


@pytest.fixture(scope="module")
def problem(tbem, wl):
    return make_problem(tbem, wl, tbem.ElasticTraction(1.0, 0.25))


@pytest.fixture(scope="module")
def problem_u(tbem, wl):
    return make_problem(tbem, wl, tbem.ElasticDisplacement(1.0, 0.25), order=75)


def loss_and_grads(tbem, p, obs):
    torch = p["torch"]
    ag = importlib.import_module("3bem_amd.autograd")
    slip = torch.tensor(p["g"]["x"], device="cuda", requires_grad=True)
    o = torch.tensor(obs, device="cuda", requires_grad=True)
    L = ag.forward_model_loss(p["op"], slip, p["u_obs"], o)
    L.backward()
    return L.item(), slip.grad.cpu().numpy(), o.grad.cpu().numpy()


def test_autograd_matches_frozen_plan_finite_differences(tbem, problem):
    p = problem
    torch = p["torch"]
    ag = importlib.import_module("3bem_amd.autograd")
    obs = p["g"]["obs"].copy()
    print("CASE autograd FD", flush=True)
    L, gs, go = loss_and_grads(tbem, p, obs)

    def loss(slip, o):
        with torch.no_grad():
            return ag.forward_model_loss(p["op"], torch.tensor(slip, device="cuda"), p["u_obs"],
                                         torch.tensor(o, device="cuda")).item()

    rng = np.random.default_rng(3)
    slip0 = p["g"]["x"]
    for j in rng.choice(slip0.size, 10, replace=False):  # linear in slip: FD exact up to rounding
        h = 1e-4
        sp, sm = slip0.copy(), slip0.copy()
        sp[j] += h
        sm[j] -= h
        fd = (loss(sp, obs) - loss(sm, obs)) / (2 * h)
        assert abs(fd - gs[j]) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(gs).max()), (j, fd, gs[j])
    for i in rng.choice(obs.shape[0], 12, replace=False):
        for m in range(3):
            h = 1e-5
            op_, om = obs.copy(), obs.copy()
            op_[i, m] += h
            om[i, m] -= h
            fd = (loss(slip0, op_) - loss(slip0, om)) / (2 * h)
            assert abs(fd - go[i, m]) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(go).max()), (i, m, fd, go[i, m])
    p["op"].set_obs_locs(obs)


def test_autograd_matches_fp64_direct_sum_gradient(tbem, orc, problem_u):
    """Gradients of the direct-sum loss (oracle) agree with the FMM's to its accuracy.

    Order 75 (forward error ~6e-6 against the direct sum on this geometry), so a
    1e-3 bound separates FMM approximation error from a gradient that is wrong in
    structure (a missing transpose, a dropped list, a sign), which is off by O(1).
    Displacement kernel: for the traction kernel with these station normals the
    reference's FMM is inaccurate at every order tried, leaving nothing to compare."""
    p = problem_u
    g = p["g"]
    obs = g["obs"]
    n = obs.shape[0]
    args = (g["obs_normals"], g["src"], g["src_normals"], g["src_weights"])
    print("CASE autograd vs direct", flush=True)
    L, gs, go = loss_and_grads(tbem, p, obs)
    u_obs = p["u_obs"].cpu().numpy()
    u = orc.direct(3, obs, *args, g["x"], parallel=True)
    assert rel_err(p["op"].apply(g["x"]), u) < 1e-4
    r = u - u_obs
    gs_ref = orc.direct_adjoint(3, obs, *args, r)
    assert rel_err(gs, gs_ref) < 1e-3, rel_err(gs, gs_ref)
    # observation gradient of the direct sum: each row depends only on its own
    # observation point, so one central difference per axis covers every point
    go_ref = np.zeros_like(go)
    h = 1e-6
    for m in range(3):
        op_, om = obs.copy(), obs.copy()
        op_[:, m] += h
        om[:, m] -= h
        du = (orc.direct(3, op_, *args, g["x"], parallel=True) - orc.direct(3, om, *args, g["x"], parallel=True))
        du /= 2 * h
        go_ref[:, m] = sum(r[q * n:(q + 1) * n] * du[q * n:(q + 1) * n] for q in range(3))
    assert rel_err(go, go_ref) < 1e-3, rel_err(go, go_ref)


def test_autograd_backward_uses_its_own_forward_stations(tbem, problem):
    """loss = f(obs1) + f(obs2) on one operator: the second forward moves the shared
    stations, so the first term's backward must move them back. Both gradients
    equal the sums of the single-call gradients, bit for bit."""
    p = problem
    torch = p["torch"]
    ag = importlib.import_module("3bem_amd.autograd")
    obs1 = p["g"]["obs"].copy()
    obs2 = obs1.copy()
    obs2[:, 2] += 1e-6 * np.sin(np.arange(obs1.shape[0]))  # stays inside every leaf cell
    print("CASE autograd two forwards", flush=True)
    L1, gs1, go1 = loss_and_grads(tbem, p, obs1)
    L2, gs2, go2 = loss_and_grads(tbem, p, obs2)
    slip = torch.tensor(p["g"]["x"], device="cuda", requires_grad=True)
    o1 = torch.tensor(obs1, device="cuda", requires_grad=True)
    o2 = torch.tensor(obs2, device="cuda", requires_grad=True)
    L = ag.forward_model_loss(p["op"], slip, p["u_obs"], o1) + ag.forward_model_loss(p["op"], slip, p["u_obs"], o2)
    L.backward()
    assert abs(L.item() - (L1 + L2)) <= 1e-14 * abs(L1 + L2)
    np.testing.assert_array_equal(o1.grad.cpu().numpy(), go1)
    np.testing.assert_array_equal(o2.grad.cpu().numpy(), go2)
    assert rel_err(slip.grad.cpu().numpy(), gs1 + gs2) < 1e-15
    p["op"].set_obs_locs(obs1)


def test_set_obs_locs_device_and_leaf_box_check(tbem, problem):
    """Moving stations from a device tensor equals moving them from the host; a
    station pushed out of its leaf cell is refused and nothing moves."""
    p = problem
    torch = p["torch"]
    op = p["op"]
    obs = p["g"]["obs"].copy()
    x = p["g"]["x"]
    print("CASE set_obs_locs device + box check", flush=True)
    moved = obs.copy()
    moved[:, 0] += 1e-7 * np.cos(np.arange(obs.shape[0]))
    v0 = op.obs_version()
    op.set_obs_locs(moved)
    y_host = op.apply(x)
    op.set_obs_locs(obs)
    op.set_obs_locs(torch.tensor(moved, device="cuda"))
    assert op.obs_version() == v0 + 3
    np.testing.assert_array_equal(op.apply(x), y_host)
    cells, _ = op.octree("obs")
    span = cells[0, 4:7]  # root half widths
    far = moved.copy()
    far[7, :] += 4.0 * span  # outside the root cell, hence outside its leaf
    for arg in (far, torch.tensor(far, device="cuda")):
        with pytest.raises(tbem.TbemError, match="outside their leaf cells"):
            op.set_obs_locs(arg)
        np.testing.assert_array_equal(op.apply(x), y_host)  # still at `moved`
    assert op.obs_version() == v0 + 3
    op.set_obs_locs(obs)
