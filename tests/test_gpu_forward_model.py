"""The differentiable elastic forward model at full size: 40k fault triangles (120k
quadrature points, slip densities N(0,1)) to 200k surface stations, loss
0.5 ||u - u_obs||^2 with u_obs from a seeded random slip. Slip to displacement is
the elastic traction (double-layer) kernel with the fault normals, at order 150
(forward 3.4e-6 from the direct sum on this geometry); the displacement kernel at
order 75 (1.3e-5) is the second case. Backward through the FMM (autograd: adjoint matvec for dL/dslip,
observation-gradient kernels for dL/dobs) against

  * the FP64 direct sum: dL/dslip = D^T r with the direct operator's exact
    transpose, dL/dobs by central differences of the direct sum along each axis
    (each row depends only on its own station);
  * central finite differences of the frozen-plan FMM loss itself, step 1e-4 with
    one Richardson step for the station coordinates (from the moved station's own
    loss terms: no other output changes; for the mixed operator, whose FP32 near
    field makes its own difference quotients noisy at ~1e-3, the FP64 model's). The
    loss is exactly quadratic in slip, so the slip
    check uses a unit step (a central difference of a quadratic is exact), which
    keeps the loss's own rounding out of the comparison: a 1e-4 step leaves 1e-3
    (displacement, order 75) to 0.4 (traction, order 150) relative noise from the
    FMM's rounding, against gradients that match the direct sum to 3e-6;
  * the FP64 operator's gradients, for the mixed-precision operator. With distinct
    observation and source points the mixed operator builds both trees on the
    union box, so that its M2L runs as FP32 MFMA translation classes; the FP64
    operator it is compared with is built on the same box (tree_boxes="common"),
    which leaves the precision as the only difference. The stations are at least
    0.1 from the fault, so no leaf pair is near enough for P2P."""
import importlib

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


# kernel, order, leaf, and bounds (forward, dL/dslip, dL/dobs) against the direct sum
CASES = {"traction": (4, 150, 256, (2e-5, 2e-5, 1e-4)), "displacement": (3, 75, 128, (5e-5, 2e-5, 1e-3))}


@pytest.fixture(scope="module", params=list(CASES))
def model(request, tbem, wl):
    torch = pytest.importorskip("torch")
    kid, order, leaf, bounds = CASES[request.param]
    g = wl.fault_and_stations()
    obs, src = g["obs"], g["src"]
    assert len(src) == 120_000 and len(obs) == 200_000
    K = tbem.make_kernel(kid, 1.0, 0.25)

    def data(o):
        return tbem.NBodyData(o, g["obs_normals"], src, g["src_normals"], g["src_weights"])

    D = tbem.make_direct_nbody_operator(data(obs), K)
    u_true = D.apply(g["x_true"])
    u_dir = D.apply(g["x"])
    r = u_dir - u_true
    n = len(obs)
    go_ref = np.zeros((n, 3))
    h = 1e-6
    for m in range(3):
        op_, om = obs.copy(), obs.copy()
        op_[:, m] += h
        om[:, m] -= h
        du = (tbem.make_direct_nbody_operator(data(op_), K).apply(g["x"]) -
              tbem.make_direct_nbody_operator(data(om), K).apply(g["x"])) / (2 * h)
        go_ref[:, m] = sum(r[q * n:(q + 1) * n] * du[q * n:(q + 1) * n] for q in range(3))
    cfg = tbem.FMMConfig(0.3, order, leaf, 0.05, False)
    ops = {p: tbem.FMMOperator(K, data(obs), cfg, precision=p) for p in ("fp64", "mixed")}
    assert ops["mixed"].info()["m2l_classes"] > 0  # the MFMA M2L runs
    ops["fp64_common"] = tbem.FMMOperator(K, data(obs), cfg, tree_boxes="common")
    return dict(torch=torch, g=g, obs=obs, u_dir=u_dir, u_obs=torch.tensor(u_true, device="cuda"),
                gs_ref=D.apply_adjoint(r), go_ref=go_ref, ops=ops, grads={}, name=request.param, bounds=bounds)


def grads(p, prec):
    if prec not in p["grads"]:
        torch = p["torch"]
        ag = importlib.import_module("3bem_amd.autograd")
        slip = torch.tensor(p["g"]["x"], device="cuda", requires_grad=True)
        o = torch.tensor(p["obs"], device="cuda", requires_grad=True)
        ag.forward_model_loss(p["ops"][prec], slip, p["u_obs"], o).backward()
        p["grads"][prec] = (slip.grad.cpu().numpy(), o.grad.cpu().numpy())
    return p["grads"][prec]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_forward_model_gradients_match_fp64_direct_sum(model, prec):
    p = model
    print(f"CASE forward model full size vs direct sum, {p['name']}, {prec}", flush=True)
    gs, go = grads(p, prec)
    e_u = rel_err(p["ops"][prec].apply(p["g"]["x"]), p["u_dir"])
    e_s, e_o = rel_err(gs, p["gs_ref"]), rel_err(go, p["go_ref"])
    print(f"  forward {e_u:.2e}, dL/dslip {e_s:.2e}, dL/dobs {e_o:.2e}")
    b_u, b_s, b_o = p["bounds"]  # a few times the FMM approximation error at these orders
    assert e_u < b_u
    assert e_s < b_s
    assert e_o < b_o


def test_forward_model_mixed_gradients_match_fp64_operator(model):
    p = model
    print(f"CASE forward model mixed vs fp64 gradients, {p['name']}", flush=True)
    gs64, go64 = grads(p, "fp64_common")
    gs32, go32 = grads(p, "mixed")
    e_s, e_o = rel_err(gs32, gs64), rel_err(go32, go64)
    print(f"  dL/dslip {e_s:.2e}, dL/dobs {e_o:.2e}")
    assert e_s < 1e-6 and e_o < 1e-5


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_forward_model_gradients_match_finite_differences(model, prec):
    p = model
    torch = p["torch"]
    ag = importlib.import_module("3bem_amd.autograd")
    op = p["ops"][prec]
    obs, slip0 = p["obs"], p["g"]["x"]
    print(f"CASE forward model finite differences, {p['name']}, {prec}", flush=True)
    gs, go = grads(p, prec)

    def loss(slip, o):
        with torch.no_grad():
            return ag.forward_model_loss(op, torch.tensor(slip, device="cuda"), p["u_obs"],
                                         torch.tensor(o, device="cuda")).item()

    rng = np.random.default_rng(5)
    for j in rng.choice(slip0.size, 6, replace=False):
        sp, sm = slip0.copy(), slip0.copy()
        sp[j] += 1.0
        sm[j] -= 1.0
        fd = (loss(sp, obs) - loss(sm, obs)) / 2.0
        assert abs(fd - gs[j]) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(gs).max()), (j, fd, gs[j])
    # Moving station i changes only station i's outputs (frozen plan, sums in a fixed
    # order), so the loss difference is the difference of station i's own terms;
    # taking it from those terms keeps the ~1e-16 L rounding of the full 600k-term
    # sum out of a difference that is ~1e-11 here.
    # The mixed operator's FP32 parts make its own difference quotients over a 5e-5
    # step noisy at ~1e-3, so its station gradient is checked against the differences
    # of the FP64 model on the same trees (the common box).
    n = obs.shape[0]
    u_obs = p["u_obs"].cpu().numpy()
    x = torch.tensor(slip0, device="cuda")
    op_fd = p["ops"]["fp64" if prec == "fp64" else "fp64_common"]

    def station_loss(o, i):
        op_fd.set_obs_locs(o)
        u = op_fd.apply(x).cpu().numpy()
        return 0.5 * sum((u[q * n + i] - u_obs[q * n + i]) ** 2 for q in range(3))

    def central(i, m, h):
        op_, om = obs.copy(), obs.copy()
        op_[i, m] += h
        om[i, m] -= h
        return (station_loss(op_, i) - station_loss(om, i)) / (2 * h)

    # At order 150 the local expansion's field varies on millimetre scales near a
    # leaf's equivalent sphere, so the plain step-1e-4 difference carries an O(h^2)
    # truncation error of ~1e-4; one Richardson step (h and h/2) removes it.
    # relative 1e-5, with an absolute floor for components far below the largest: 1e-8
    # of the largest for FP64; 1e-7 for the mixed gradient, whose FP32 near field puts
    # it 2.6e-7 (in norm) from the FP64 gradient these differences describe
    floor = 1e-3 if prec == "fp64" else 1e-2
    for i in rng.choice(n, 8, replace=False):
        for m in range(3):
            h = 1e-4
            fd = (4.0 * central(i, m, h / 2) - central(i, m, h)) / 3.0
            assert abs(fd - go[i, m]) <= 1e-5 * max(abs(fd), floor * np.abs(go).max()), (i, m, fd, go[i, m])
    op_fd.set_obs_locs(obs)
    op.set_obs_locs(obs)
