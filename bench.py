"""Benchmark: FMM matvec throughput on one MI355X.

Default workload: plain FP64 elastic-displacement (Kelvin, mu=1, nu=0.25) FMM
matvec on 1M points (500k uniform on the free surface [-1/2,1/2]^2 x {0}, 500k
on a 0.5 x 0.5 fault dipping 45 degrees centred at depth 0.3), 3-component N(0,1)
densities, leaf 64, order 6, MAC 0.3, d 0.05.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config fault_surface|icosphere|cube_clusters|rough_fault] [--n N]

One JSON line on stdout. `value` is FP64 Mpts/s = N_obs / apply time with inputs
resident in HBM (device pointers). Beside it: the mixed FP32 path, relative L2
error of both against an FP64 direct sum on a 10k-target subsample (GPU FP64
direct operator, itself checked against the CPU oracle in tests/), the
backward/forward time ratio (adjoint matvec + observation-coordinate gradient vs
forward), the roofline of the dominant kernel (the fused P2P+M2P+L2P leaf stage)
timed with HIP events on its stream, and the compiled reference CPU FMM timed on
this host on a bounded sample. With the default workload, `target_fp32_mfma_vs_fp64`
adds the precision comparison at the accuracy target: 2M points, order 100 (FP64
FMM ~4e-7 from the direct sum), FP64 against mixed precision (FP32 MFMA M2L, FP32
point-to-point), with the roofline of the MFMA M2L kernel.
"""
import argparse
import importlib
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FMM matvec Mpts/s at stated rel err, FP32-MFMA vs FP64; bwd/fwd time ratio"
# algorithmic flops per source-target pair (FMA = 2, rsqrt = 1), per kernel id
FLOPS_PER_PAIR = {0: 11, 1: 19, 2: 0, 3: 28, 4: 60, 5: 60, 6: 120, 7: 2, 8: 6}
PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 vector (= FP64 matrix) peak
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector peak (packed)
PEAK_HBM_GBS = 8000.0     # MI355X HBM3E
# HBM bytes per launch of the FP64 leaf stage on the default workload: rocprofv3
# --pmc FETCH_SIZE and WRITE_SIZE of the same command, scaled by what the counters
# read for a plain device copy (profiles/k_leaf_stage_elastic_u_1M_order6.md).
# Other workloads have no measured figure and report null.
LEAF_TRAFFIC_BYTES = {("fault_surface", 1_000_000): 644_165_984}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name, n):
    wl = importlib.import_module("3bem_amd.workloads")
    if name == "icosphere":
        g = wl.icosphere_surface_15k()
        return dict(kernel=0, locs=g["locs"], normals=g["normals"], weights=g["weights"], x=g["x"],
                    cfg=(0.3, 5, 100, 0.05, False),
                    desc="Laplace single-layer FMM matvec, icosphere 5120 triangles x 3 points = 15,360 "
                         "points, order 5, leaf 100, MAC 0.3")
    if name == "fault_surface":
        g = wl.fault_and_free_surface(n or 1_000_000)
        return dict(kernel=3, locs=g["locs"], normals=g["normals"], weights=g["weights"], x=g["x"],
                    cfg=(0.3, 6, 64, 0.05, False),
                    desc=f"elastic displacement (Kelvin) FMM matvec, {len(g['locs'])} points (half free "
                         "surface, half 45-degree fault), mu=1 nu=0.25, order 6, leaf 64, MAC 0.3")
    if name == "cube_clusters":
        g = wl.cube_and_clusters(n or 4_000_000)
        return dict(kernel=0, locs=g["locs"], normals=g["normals"], weights=g["weights"], x=g["x"],
                    cfg=(0.3, 75, 100, 0.05, False),
                    desc=f"Laplace single-layer FMM matvec, {len(g['locs'])} points (cube + 8 clusters), "
                         "order 75, leaf 100, MAC 0.3")
    if name == "rough_fault":
        g = wl.rough_fault_points(n or 2_000_000)
        x = np.random.default_rng(wl.SEED + 1).standard_normal(3 * len(g["locs"]))
        return dict(kernel=6, locs=g["locs"], normals=g["normals"], weights=g["weights"], x=x,
                    cfg=(0.3, 6, 64, 0.05, False),
                    desc=f"elastic hypersingular FMM matvec, {len(g['locs'])} points on a rough fault, "
                         "order 6, leaf 64, MAC 0.3")
    raise SystemExit(f"unknown config {name}")


def time_steps(fn, steps, warmup, torch, after=None):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    per = []
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
        if after is not None:
            per.append(after())
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, per


def backward_times(op, x, g, steps, torch):
    """The backward of y = op.apply(x) as autograd runs it: after an untimed forward,
    the adjoint A^T g and the observation gradient, which reuses that forward's
    expansions (forward token). Returns (adjoint s, obs_grad s, obs_grad s with the
    expansions recomputed, the path taken when another apply intervened), medians
    over steps after one warmup."""
    adj, grad, recomp = [], [], []
    for _ in range(steps + 1):
        op.apply(x)
        tok = op.forward_token()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.apply_adjoint(g)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        op.obs_grad(x, g, token=tok)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        op.obs_grad(x, g)
        torch.cuda.synchronize()
        adj.append(t1 - t0)
        grad.append(t2 - t1)
        recomp.append(time.perf_counter() - t2)
    return tuple(statistics.median(v[1:]) for v in (adj, grad, recomp))


def cpu_baseline(kernel, cfg, budget_n):
    """Compiled reference FMM (oracle/_ref) on a bounded sample of the same geometry;
    falls back to the oracle restatement when the reference build is absent."""
    from oracle.oracle import Oracle, available
    wl = importlib.import_module("3bem_amd.workloads")
    kind = "reference" if available("ref") else "port"
    o = Oracle("ref" if kind == "reference" else "orc")
    g = wl.fault_and_free_surface(budget_n) if kernel == 3 else wl.uniform_cube(budget_n, dim=3 if kernel >= 3 else 1)
    t0 = time.perf_counter()
    op = o.fmm(kernel, g["locs"], g["normals"], g["locs"], g["normals"], g["weights"], *cfg)
    t_plan = time.perf_counter() - t0
    times = []
    for _ in range(2):
        t0 = time.perf_counter()
        op.apply(g["x"][: (3 if kernel >= 3 else 1) * budget_n])
        times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    return {"value": round(budget_n / t / 1e6, 5), "unit": "Mpts/s", "cores": cores, "kind": kind,
            "sample": f"same geometry family at {budget_n} points, FMMOperator::apply median of 2 runs "
                      f"({t:.2f} s each; plan build {t_plan:.2f} s excluded), OMP threads={cores}"}


def compare_run(tbem, torch, steps, g, kid, cfg_t, n_sub, desc):
    """FP32/MFMA against FP64 on one workload, both precisions on the same plan
    parameters: forward time, backward (adjoint + observation gradient) time, and
    relative L2 errors against an FP64 direct sum on an n_sub-target subsample."""
    n = len(g["locs"])
    R = 1 if kid in (0, 1, 2, 7) else 3
    K = tbem.make_kernel(kid, 1.0, 0.25)
    data = tbem.NBodyData(g["locs"], g["normals"], g["locs"], g["normals"], g["weights"])
    cfg = tbem.FMMConfig(*cfg_t)
    x = torch.tensor(g["x"], device="cuda")
    rng = np.random.default_rng(321)
    sub = np.sort(rng.choice(n, n_sub, replace=False))
    dsub = tbem.NBodyData(g["locs"][sub], g["normals"][sub], g["locs"], g["normals"], g["weights"])
    yd = tbem.make_direct_nbody_operator(dsub, K).apply(g["x"])
    idx = np.concatenate([q * n + sub for q in range(R)])
    out = {"workload": desc}
    ys = {}
    for prec in ("fp64", "mixed"):
        t0 = time.perf_counter()
        op = tbem.FMMOperator(K, data, cfg, precision=prec)
        plan = time.perf_counter() - t0
        dt, m2l = time_steps(lambda: op.apply(x), steps, 1, torch, after=lambda: op.m2l_kernel_ms())
        ys[prec] = op.apply(x).cpu().numpy()
        info = op.info()
        stage_ms = {k: round(v, 3) for k, v in op.stage_times().items()}
        gv = torch.randn(R * n, dtype=torch.float64, device="cuda")
        dt_adj, dt_grad, dt_grad_re = backward_times(op, x, gv, 2, torch)
        out[prec] = {"ms_per_step": round(dt * 1e3, 3), "Mpts_per_s": round(n / dt / 1e6, 3),
                     "bwd_fwd_ratio": round((dt_adj + dt_grad) / dt, 3),
                     "bwd_ms": {"adjoint": round(dt_adj * 1e3, 2), "obs_grad": round(dt_grad * 1e3, 2),
                                "obs_grad_recomputing": round(dt_grad_re * 1e3, 2)},
                     "plan_build_s": round(plan, 2), "m2l_kernel_ms": round(float(np.mean(m2l)), 3),
                     "stage_ms": stage_ms,
                     "rel_err_vs_direct": float(np.linalg.norm(ys[prec][idx] - yd) / np.linalg.norm(yd))}
        # the same operator rebuilt from its saved plan (octrees, task lists, C2E)
        blob = op.save_plan()
        t0 = time.perf_counter()
        op_l = tbem.FMMOperator(K, data, cfg, precision=prec, plan=blob)
        out[prec]["plan_from_saved_s"] = round(time.perf_counter() - t0, 3)
        out[prec]["saved_plan_MB"] = round(len(blob) / 2 ** 20, 1)
        out[prec]["saved_plan_same_output"] = bool(np.array_equal(op_l.apply(x).cpu().numpy(), ys[prec]))
        del op_l, blob
        if prec == "mixed" and info["m2l_classes"] > 0:
            blk = R * info["n_e"]
            flop = 2.0 * info["m2l"] * blk * blk
            ach = flop / (float(np.mean(m2l)) * 1e-3) / 1e12
            out["m2l_mfma_roofline"] = {"bound": "mfma", "kernel": "k_m2l_mfma (FP32 v_mfma_f32_16x16x4f32)",
                                        "achieved": round(ach, 2), "peak": 157.3, "unit": "TFLOP/s",
                                        "frac": round(ach / 157.3, 4), "algorithmic_flop": flop,
                                        "classes": info["m2l_classes"], "pairs": info["m2l"]}
        log(f"[bench] {desc[:40]} {prec}: {dt * 1e3:.2f} ms, rel err {out[prec]['rel_err_vs_direct']:.2e}, "
            f"m2l kernel {out[prec]['m2l_kernel_ms']:.2f} ms")
        del op
        torch.cuda.empty_cache()
    out["rel_err_mixed_vs_fp64"] = float(np.linalg.norm(ys["mixed"] - ys["fp64"]) / np.linalg.norm(ys["fp64"]))
    out["speedup_mixed_vs_fp64"] = round(out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"], 3)
    return out


def target_run(tbem, torch, steps):
    """The FP32/MFMA-vs-FP64 comparison at the accuracy target: elastic displacement
    FMM matvec on 2M points of the fault + free-surface geometry, order 100, leaf 256,
    MAC 0.3 (FP64 FMM ~4e-7 from the direct sum), errors on a 5k-target subsample."""
    wl = importlib.import_module("3bem_amd.workloads")
    return compare_run(tbem, torch, steps, wl.fault_and_free_surface(2_000_000), 3, (0.3, 100, 256, 0.05, False),
                       5000, "elastic displacement FMM matvec, 2,000,000 points (fault + free surface), order 100, "
                             "leaf 256, MAC 0.3")


def laplace_run(tbem, torch, steps):
    """The mixed-precision Laplace workload: single-layer FMM matvec on 4M points
    (half uniform in the unit cube, half in 8 Gaussian clusters, sigma 0.02), order
    75, leaf 100, MAC 0.3, errors on a 10k-target subsample."""
    wl = importlib.import_module("3bem_amd.workloads")
    return compare_run(tbem, torch, steps, wl.cube_and_clusters(4_000_000), 0, (0.3, 75, 100, 0.05, False),
                       10_000, "Laplace single-layer FMM matvec, 4,000,000 points (cube + 8 clusters), order 75, "
                               "leaf 100, MAC 0.3")


def forward_model_run(tbem, torch, steps):
    """The differentiable elastic forward model: 40k fault triangles (120k points)
    to 200k surface stations through the elastic traction (double-layer) kernel
    with the fault normals, loss 0.5 ||u - u_obs||^2, order 150, leaf 256, MAC 0.3.
    Forward and backward (adjoint + observation gradient) times, and the autograd
    gradients against the FP64 direct sum (dL/dslip = D^T r; dL/dobs by central
    differences of the direct sum, step 1e-6)."""
    wl = importlib.import_module("3bem_amd.workloads")
    ag = importlib.import_module("3bem_amd.autograd")
    g = wl.fault_and_stations()
    obs, src = g["obs"], g["src"]
    n = len(obs)
    K = tbem.ElasticTraction(1.0, 0.25)

    def data(o):
        return tbem.NBodyData(o, g["obs_normals"], src, g["src_normals"], g["src_weights"])

    D = tbem.make_direct_nbody_operator(data(obs), K)
    u_true = D.apply(g["x_true"])
    r = D.apply(g["x"]) - u_true
    gs_ref = D.apply_adjoint(r)
    go_ref = np.zeros((n, 3))
    for m in range(3):
        op_, om = obs.copy(), obs.copy()
        op_[:, m] += 1e-6
        om[:, m] -= 1e-6
        du = (tbem.make_direct_nbody_operator(data(op_), K).apply(g["x"]) -
              tbem.make_direct_nbody_operator(data(om), K).apply(g["x"])) / 2e-6
        go_ref[:, m] = sum(r[q * n:(q + 1) * n] * du[q * n:(q + 1) * n] for q in range(3))
    u_obs = torch.tensor(u_true, device="cuda")
    x = torch.tensor(g["x"], device="cuda")
    gv = torch.tensor(r, device="cuda")
    out = {"workload": "differentiable elastic forward model: 40,000 fault triangles (120,000 points) to 200,000 "
                       "surface stations, elastic traction kernel, loss 0.5||u - u_obs||^2, order 150, leaf 256, "
                       "MAC 0.3"}
    for prec in ("fp64", "mixed"):
        op = tbem.FMMOperator(K, data(obs), tbem.FMMConfig(0.3, 150, 256, 0.05, False), precision=prec)
        dt, _ = time_steps(lambda: op.apply(x), steps, 1, torch)
        dt_adj, dt_grad, _ = backward_times(op, x, gv, steps, torch)
        dtb = dt_adj + dt_grad
        slip = torch.tensor(g["x"], device="cuda", requires_grad=True)
        o = torch.tensor(obs, device="cuda", requires_grad=True)
        ag.forward_model_loss(op, slip, u_obs, o).backward()
        gs, go = slip.grad.cpu().numpy(), o.grad.cpu().numpy()
        out[prec] = {"fwd_ms": round(dt * 1e3, 3), "bwd_ms": round(dtb * 1e3, 3),
                     "bwd_fwd_ratio": round(dtb / dt, 3),
                     "rel_err_dL_dslip_vs_direct": float(np.linalg.norm(gs - gs_ref) / np.linalg.norm(gs_ref)),
                     "rel_err_dL_dobs_vs_direct": float(np.linalg.norm(go - go_ref) / np.linalg.norm(go_ref))}
        log(f"[bench] forward model {prec}: fwd {dt * 1e3:.2f} ms, bwd {dtb * 1e3:.2f} ms, grad errors "
            f"{out[prec]['rel_err_dL_dslip_vs_direct']:.2e} / {out[prec]['rel_err_dL_dobs_vs_direct']:.2e}")
        del op
        torch.cuda.empty_cache()
    return out


def bem_solve_run(tbem, torch):
    """The BEM solve: elastic hypersingular FMM on the rough fault (2M points, order
    50, leaf 128, MAC 0.3) as the second-kind system (alpha I + H) x = b, alpha =
    2 ||H|| from 8 power iterations, b the traction of a uniform unit remote stress.
    Plain FP64 GMRES(50) to a 1e-8 relative residual against FP64-residual iterative
    refinement around the mixed-precision FMM (inner GMRES to 1e-4), both timed end
    to end with the operators already built."""
    wl = importlib.import_module("3bem_amd.workloads")
    gm = importlib.import_module("3bem_amd.gmres")
    n = 2_000_000
    g = wl.rough_fault_points(n)
    K = tbem.ElasticHypersingular(1.0, 0.25)
    data = tbem.NBodyData(g["locs"], g["normals"], g["locs"], g["normals"], g["weights"])
    cfg = tbem.FMMConfig(0.3, 50, 128, 0.05, False)
    op64 = tbem.FMMOperator(K, data, cfg, precision="fp64")
    op32 = tbem.FMMOperator(K, data, cfg, precision="mixed")
    v = torch.randn(3 * n, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(1))
    for _ in range(8):
        v = op32.apply(v)
        v = v / torch.linalg.norm(v)
    alpha = 2.0 * float(torch.linalg.norm(op64.apply(v)))
    b = torch.tensor(g["normals"].T.reshape(-1).copy(), device="cuda")
    out = {"workload": "elastic hypersingular BEM solve, 2,000,000 points on the rough fault, order 50, leaf 128, "
                       "MAC 0.3, (alpha I + H) x = b, alpha = 2||H||, to a 1e-8 relative FP64 residual"}
    for name, op in (("fp64", op64), ("mixed", op32)):
        dt, _ = time_steps(lambda: op.apply(b), 3, 1, torch)
        out[f"{name}_matvec_ms"] = round(dt * 1e3, 2)
    A64 = lambda x: alpha * x + op64.apply(x)
    A32 = lambda x: alpha * x + op32.apply(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x64, i64 = gm.gmres(A64, b, tol=1e-8, restart=50, maxiter=500)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    x, info = gm.refine(A64, A32, b, tol=1e-8, inner_tol=1e-4, restart=50)
    torch.cuda.synchronize()
    tr = time.perf_counter() - t0
    bn = float(torch.linalg.norm(b))
    out["fp64_gmres"] = {"s": round(t64, 3), "iterations": i64["iterations"],
                         "residual": float(torch.linalg.norm(b - A64(x64))) / bn}
    out["mixed_refinement"] = {"s": round(tr, 3), "outer": info["outer"], "inner_iterations": info["inner_iterations"],
                               "residual": float(torch.linalg.norm(b - A64(x))) / bn,
                               "rel_diff_vs_fp64_solution": float(torch.linalg.norm(x - x64) / torch.linalg.norm(x64))}
    out["speedup_refinement_vs_fp64_gmres"] = round(t64 / tr, 3)
    log(f"[bench] BEM solve: fp64 GMRES {t64:.2f} s ({i64['iterations']} it), refinement {tr:.2f} s "
        f"({info['outer']} outer, {info['inner_iterations']} inner)")
    del op64, op32
    torch.cuda.empty_cache()
    return out


def near_field_run(tbem, torch, steps, warmup, skip_cpu):
    """The integral operator y = N x + (G F(I x) + C x) on a triangulated rough fault
    (300 x 300 grid cells, 180,000 facets; far field on the 3-point rule's 540,000
    points, elastic hypersingular FMM, order 50, leaf 128, MAC 0.3) with a near field N
    and a correction C of a Galerkin near-field operator's sparsity (every dof row
    holds the 3 x 3 x 18 dofs of the facets in the 3 x 3 grid cells around its own:
    1,620,000 rows, 261M entries). Times the CSR apply kernel alone (HIP events on the
    stream it runs on) and the whole integral operator; the CPU reference
    SparseOperator::apply runs on a bounded sample of the rows."""
    wl = importlib.import_module("3bem_amd.workloads")
    bem = importlib.import_module("3bem_amd.bem")
    nx = ny = 300
    t0 = time.perf_counter()
    n, m, rp, ci, v = wl.near_field_csr(nx, ny)
    N = tbem.SparseOperator(n, m, v, ci, rp)
    C = tbem.SparseOperator(n, m, -0.1 * v, ci, rp)
    t_build = time.perf_counter() - t0
    nnz = int(rp[-1])
    x = torch.rand(m, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(3)) - 0.5
    for _ in range(warmup):
        y = N.apply(x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()  # the apply runs on torch's current stream: these events bracket its launches
    for _ in range(steps):
        y = N.apply(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    alg_bytes = 12 * nnz + 16 * n + 8 * m
    gbs = alg_bytes / (ms * 1e-3) / 1e9
    out = {"workload": f"CSR near-field apply, {n} x {m}, {nnz} entries (18-facet Galerkin pattern on a "
                       f"{nx} x {ny}-cell triangulated rough fault), FP64",
           "csr_build_s": round(t_build, 2), "csr_apply_ms": round(ms, 4),
           "roofline": {"bound": "hbm", "kernel": "k_csr_apply", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS,
                        "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 4), "algorithmic_bytes": alg_bytes,
                        "bytes_rule": "12 B per entry (FP64 value, int32 column), 16 B per row (int64 row "
                                      "pointer, FP64 output), 8 B per column (x)"}}
    log(f"[bench] near field: {nnz} entries, k_csr_apply {ms:.3f} ms = {gbs:.0f} GB/s")
    if not skip_cpu:
        from oracle.oracle import Oracle, available
        kind = "reference" if available("ref") else "port"
        o = Oracle("ref" if kind == "reference" else "orc")
        rows = int(np.searchsorted(rp, 40_000_000))
        xh = x.cpu().numpy()
        y_cpu, t_cpu = o.sparse_csr_apply(rows, m, rp[:rows + 1], ci[:rp[rows]], v[:rp[rows]], xh, reps=3)
        y_gpu = y.cpu().numpy()[:rows]
        out["cpu_baseline"] = {"value": round(int(rp[rows]) / t_cpu / 1e9, 4), "unit": "G entries/s", "cores": 1,
                               "kind": kind, "sample": f"first {rows} rows ({int(rp[rows])} entries), "
                                                       "SparseOperator::apply mean of 3"}
        out["gpu_g_entries_per_s"] = round(nnz / (ms * 1e-3) / 1e9, 3)
        out["bit_exact_vs_cpu_sample"] = bool(np.array_equal(y_gpu.view(np.uint64), y_cpu.view(np.uint64)))
        log(f"[bench] near field CPU ({kind}) {out['cpu_baseline']['value']} G entries/s; "
            f"bit-exact {out['bit_exact_vs_cpu_sample']}")
    del rp, ci, v
    # continuity constraints of the mesh's discontinuous dofs: what wraps every matvec
    # of the constrained solve (distribute before, condense after)
    cons = importlib.import_module("3bem_amd.constraints")
    t0 = time.perf_counter()
    ep, td, tw, rhs, n_total = wl.grid_continuity_constraints(nx, ny, 3, all_pairs=False)
    cm = cons.from_constraint_arrays(ep, td, tw, rhs)
    t_cm = time.perf_counter() - t0
    assert n_total == m
    red = torch.rand(m - cm.size(), dtype=torch.float64, device="cuda")
    full = torch.rand(m, dtype=torch.float64, device="cuda")
    cons.distribute_vector(cm, red, m)
    cons.condense_vector(cm, full)
    torch.cuda.synchronize()
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record()
    for _ in range(steps):
        cons.distribute_vector(cm, red, m)
    e1.record()
    for _ in range(steps):
        cons.condense_vector(cm, full)
    e2.record()
    torch.cuda.synchronize()
    out["constraints"] = {"dofs": m, "equations": int(len(rhs)), "constrained": cm.size(),
                          "from_constraints_s": round(t_cm, 2),
                          "distribute_ms": round(e0.elapsed_time(e1) / steps, 4),
                          "condense_ms": round(e1.elapsed_time(e2) / steps, 4)}
    log(f"[bench] constraints: {cm.size()} of {m} dofs constrained; distribute "
        f"{out['constraints']['distribute_ms']:.3f} ms, condense {out['constraints']['condense_ms']:.3f} ms")
    del cm, red, full
    facets = wl.rough_fault_mesh(nx, ny)
    q3 = (np.array([[1 / 6, 1 / 6], [2 / 3, 1 / 6], [1 / 6, 2 / 3]]), np.full(3, 1 / 6))
    K = tbem.ElasticHypersingular(1.0, 0.25)
    cfg = tbem.FMMConfig(0.3, 50, 128, 0.05, False)
    for prec in ("fp64", "mixed"):
        op = bem.integral_operator(N, C, facets, facets, q3, q3, K, cfg, precision=prec)
        dt_far, _ = time_steps(lambda: op.farfield.apply(torch.zeros(op.farfield.n_cols(), dtype=torch.float64,
                                                                     device="cuda")), steps, warmup, torch)
        dt, _ = time_steps(lambda: op.apply(x), steps, warmup, torch)
        out[f"integral_apply_ms_{prec}"] = round(dt * 1e3, 3)
        out[f"far_field_fmm_ms_{prec}"] = round(dt_far * 1e3, 3)
        log(f"[bench] integral operator ({prec}): {dt * 1e3:.2f} ms (FMM far field alone {dt_far * 1e3:.2f} ms)")
        del op
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="fault_surface")
    ap.add_argument("--n", type=int, default=0)
    ap.add_argument("--cpu-n", type=int, default=100_000)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-extra", action="store_true", help="FP64 forward only")
    ap.add_argument("--skip-target", action="store_true",
                    help="skip the 2M-point order-100 FP32/MFMA vs FP64 comparison")
    ap.add_argument("--skip-laplace", action="store_true",
                    help="skip the 4M-point Laplace FP32/MFMA-vs-FP64 comparison")
    ap.add_argument("--skip-forward-model", action="store_true",
                    help="skip the differentiable forward model (120k fault points, 200k stations)")
    ap.add_argument("--skip-bem-solve", action="store_true",
                    help="skip the 2M-point hypersingular GMRES solve")
    ap.add_argument("--skip-near", action="store_true",
                    help="skip the CSR near field / integral operator section")
    ap.add_argument("--only-near", action="store_true",
                    help="run only the headline FP64 step and the near field / integral operator section")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 256 MiB device copy (HBM counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.only_near:
        args.skip_target = args.skip_laplace = args.skip_forward_model = args.skip_bem_solve = True
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for --gpus {args.gpus}); the FMM "
              "sharding over several GPUs is not implemented", file=sys.stderr)
        raise SystemExit(2)

    import torch
    tbem = importlib.import_module("3bem_amd")
    torch.cuda.set_device(0)
    if args.calibrate_copy:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")  # 256 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
            dst.add_(src)  # elementwise kernel: reads 512 MiB, writes 256 MiB
        torch.cuda.synchronize()
        log("[bench] calibration: 3 x (256 MiB copy, 256 MiB add_)")
        return
    w = workload(args.config, args.n)
    kid = w["kernel"]
    R = 1 if kid in (0, 1, 2, 7) else 3
    n = len(w["locs"])
    K = tbem.make_kernel(kid, 1.0, 0.25)
    data = tbem.NBodyData(w["locs"], w["normals"], w["locs"], w["normals"], w["weights"])
    cfg = tbem.FMMConfig(*w["cfg"])
    log(f"[bench] {w['desc']}")

    t0 = time.perf_counter()
    op64 = tbem.FMMOperator(K, data, cfg, precision="fp64")
    plan64 = time.perf_counter() - t0
    info = op64.info()
    log(f"[bench] fp64 plan {plan64:.2f}s: " + ", ".join(f"{k}={info[k]}" for k in
        ("n_e", "n_src_cells", "n_up_levels", "p2p", "m2l", "p2p_pairs")))
    x = torch.tensor(w["x"], device="cuda")

    leaf_ms = []
    dt64, leaf_ms = time_steps(lambda: op64.apply(x), args.steps, args.warmup, torch,
                               after=lambda: op64.stage_times()["leaf"])
    stages = op64.stage_times()
    value = n / dt64 / 1e6
    log(f"[bench] fp64 apply {dt64 * 1e3:.2f} ms -> {value:.2f} Mpts/s; stages {stages}")

    leaf_avg_ms = float(np.mean(leaf_ms))
    alg_flops = FLOPS_PER_PAIR[kid] * (info["p2p_pairs"] + n * info["n_e"])
    achieved = alg_flops / (leaf_avg_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "pipe": "fp64 VALU (MI355X FP64 vector peak = FP64 matrix peak)",
                "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
                "traffic": LEAF_TRAFFIC_BYTES.get((args.config, n)), "traffic_unit": "bytes per launch",
                "kernel": "k_leaf_stage (P2P+M2P+L2P)", "kernel_avg_ms": round(leaf_avg_ms, 4),
                "algorithmic_flop": alg_flops}

    out = {"metric": METRIC, "value": round(value, 3), "unit": "Mpts/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(dt64 * 1e3, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": w["desc"], "n_points": n, "kernel": type(K).__name__,
                      "mac": cfg.mac, "order": cfg.order, "n_e": info["n_e"], "min_pts_per_cell": cfg.min_pts_per_cell,
                      "plan_build_s": round(plan64, 3), "stage_ms": {k: round(v, 4) for k, v in stages.items()}},
           "roofline": roofline}

    if not args.skip_extra and not args.only_near:
        # accuracy vs FP64 direct sum on a 10k-target subsample
        rng = np.random.default_rng(123)
        sub = np.sort(rng.choice(n, min(n, 10_000), replace=False))
        dsub = tbem.NBodyData(w["locs"][sub], w["normals"][sub], w["locs"], w["normals"], w["weights"])
        direct = tbem.make_direct_nbody_operator(dsub, K)
        t0 = time.perf_counter()
        yd = direct.apply(w["x"])
        t_direct = time.perf_counter() - t0
        idx = np.concatenate([q * n + sub for q in range(R)])
        y64 = op64.apply(x).cpu().numpy()
        e64 = float(np.linalg.norm(y64[idx] - yd) / np.linalg.norm(yd))

        t0 = time.perf_counter()
        op32 = tbem.FMMOperator(K, data, cfg, precision="mixed")
        plan32 = time.perf_counter() - t0
        dt32, _ = time_steps(lambda: op32.apply(x), args.steps, args.warmup, torch)
        y32 = op32.apply(x).cpu().numpy()
        e32 = float(np.linalg.norm(y32[idx] - yd) / np.linalg.norm(yd))
        e32_64 = float(np.linalg.norm(y32 - y64) / np.linalg.norm(y64))

        g = torch.randn(R * n, dtype=torch.float64, device="cuda")
        dt_adj, dt_grad, dt_grad_re = backward_times(op64, x, g, max(2, args.steps // 2), torch)
        out.update({
            "rel_err_fp64_vs_direct": e64,
            "mixed": {"value": round(n / dt32 / 1e6, 3), "ms_per_step": round(dt32 * 1e3, 4), "rel_err_vs_direct": e32,
                      "rel_err_vs_fp64_fmm": e32_64, "speedup_vs_fp64": round(dt64 / dt32, 3),
                      "plan_build_s": round(plan32, 3)},
            "bwd_fwd_ratio": round((dt_adj + dt_grad) / dt64, 3),
            "bwd_ms": {"adjoint": round(dt_adj * 1e3, 3), "obs_grad": round(dt_grad * 1e3, 3),
                       "obs_grad_recomputing": round(dt_grad_re * 1e3, 3)},
            "direct_subsample_s": round(t_direct, 3),
        })
        log(f"[bench] rel err fp64 {e64:.3e}, mixed {e32:.3e}; mixed {dt32 * 1e3:.2f} ms; "
            f"adjoint {dt_adj * 1e3:.2f} ms, obs_grad {dt_grad * 1e3:.2f} ms")

    if not args.skip_extra and not args.skip_target and args.config == "fault_surface":
        out["target_fp32_mfma_vs_fp64"] = target_run(tbem, torch, max(3, args.steps // 2))
    if not args.skip_extra and not args.skip_laplace and args.config == "fault_surface":
        out["laplace_4M_fp32_mfma_vs_fp64"] = laplace_run(tbem, torch, 3)
    if not args.skip_extra and not args.skip_forward_model and args.config == "fault_surface":
        out["forward_model_fault_stations"] = forward_model_run(tbem, torch, 5)
    if not args.skip_extra and not args.skip_bem_solve and args.config == "fault_surface":
        out["bem_solve_rough_fault"] = bem_solve_run(tbem, torch)
    if not args.skip_extra and not args.skip_near and args.config == "fault_surface":
        out["integral_operator_near_field"] = near_field_run(tbem, torch, args.steps, args.warmup, args.skip_cpu)

    if not args.skip_cpu:
        cfg_t = (cfg.mac, cfg.order, cfg.min_pts_per_cell, cfg.d, cfg.account_for_small_cells)
        cpu_n = min(args.cpu_n, n)
        log(f"[bench] CPU baseline on {cpu_n} points ...")
        out["cpu_baseline"] = cpu_baseline(kid, cfg_t, cpu_n)
        log(f"[bench] CPU baseline {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
