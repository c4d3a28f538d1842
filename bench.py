"""Benchmark: FMM matvec throughput on one MI355X.

Headline workload (the metric's configuration): plain FP64 elastic-displacement
(Kelvin, mu=1, nu=0.25) FMM matvec on 2M points (1M uniform on the free surface
[-1/2,1/2]^2 x {0}, 1M on a 0.5 x 0.5 fault dipping 45 degrees centred at depth
0.3), 3-component N(0,1) densities, order 100, leaf 256, MAC 0.3, d 0.05.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME] [--n N]

One JSON line on stdout. `value` is FP64 Mpts/s = N_obs / apply time with inputs
resident in HBM (device pointers). Beside it, on the same workload: the mixed path
(FP32 MFMA M2L, FP32 point-to-point) with its speed-up, the relative L2 error of
both against an FP64 direct sum on a 10k-target subsample (GPU FP64 direct
operator, itself checked against the CPU oracle in tests/), the backward/forward
time ratio (adjoint matvec + observation-coordinate gradient vs forward), and the
rooflines of the leaf stage and the M2L kernel of each precision, timed with HIP
events on the stream they run on; `roofline` is the FP64 step's dominant kernel.
Further sections: the 1M-point order-6 version of the workload, the 4M-point
cube-and-clusters Laplace workload, the differentiable forward model (fault to
stations), the rough-fault BEM solve and the near-field CSR / integral operator.
`cpu_baseline` times the compiled reference CPU FMM on this host on a bounded sample.
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
# a pair evaluated once for both directions (k_leaf_sym), per kernel id: Kelvin U 12
# shared (displacement, r^2, rsqrt, r^-3, C2/r) + 18 per direction (d.v, its factor,
# 3 x 2 FMA); Laplace single layer 9 shared (displacement, r^2, rsqrt) + 2 per direction
FLOPS_PER_SYM_PAIR = {0: 13, 3: 48}
PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 vector (= FP64 matrix) peak
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector peak (packed)
PEAK_HBM_GBS = 8000.0     # MI355X HBM3E
# HBM bytes per launch of a workload's leaf and M2L kernels: rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE of the same command (mean over its dispatches; the leaf
# stage's are summed over its kernels, k_leaf_wave + k_leaf_parts + k_leaf_stage
# (+ k_leaf_sym + k_sym_gather where pairs are evaluated once for both directions),
# for the fused launches the rooflines time), scaled by what the counters read for
# a plain device copy (factors 2.0 and 1.0; the profiles under profiles/ give the
# runs). Keyed by (workload, points, precision, kernel); other workloads report null.
TRAFFIC_BYTES = {
    ("fault_surface", 1_000_000, "fp64", "leaf"): 334_100_000,
    ("fault_surface_o100", 2_000_000, "fp64", "m2l"): 15_485_747_200,
    ("fault_surface_o100", 2_000_000, "fp64", "leaf"): 4_213_500_000,
    ("fault_surface_o100", 2_000_000, "mixed", "leaf"): 4_216_000_000,
    ("fault_surface_o100", 2_000_000, "mixed", "m2l"): 7_941_620_238,
    ("cube_clusters", 4_000_000, "fp64", "leaf"): 14_155_000_000,
}
TRAFFIC_UNIT = "HBM bytes per launch (rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0)"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name, n):
    w = _workload(name, n)
    w["name"] = name
    return w


def _workload(name, n):
    wl = importlib.import_module("3bem_amd.workloads")
    if name == "fault_surface_o100":
        g = wl.fault_and_free_surface(n or 2_000_000)
        return dict(kernel=3, locs=g["locs"], normals=g["normals"], weights=g["weights"], x=g["x"],
                    cfg=(0.3, 100, 256, 0.05, False),
                    desc=f"elastic displacement (Kelvin) FMM matvec, {len(g['locs'])} points (half free "
                         "surface, half 45-degree fault), mu=1 nu=0.25, order 100, leaf 256, MAC 0.3")
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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(kernel, cfg, n_sample, runs):
    """The compiled reference FMM (oracle/_ref; the FP64 restatement when the
    reference build is absent) on an n_sample-point sample of the headline's geometry
    family with the same kernel and FMMConfig: median of `runs` warm applies at the
    OpenMP thread count of this host's environment and again at one thread, plan
    build excluded and reported."""
    import ctypes
    from oracle.oracle import Oracle, available
    wl = importlib.import_module("3bem_amd.workloads")
    kind = "reference" if available("ref") else "port"
    o = Oracle("ref" if kind == "reference" else "orc")
    g = wl.fault_and_free_surface(n_sample) if kernel >= 3 else wl.uniform_cube(n_sample, dim=1)
    R = 3 if kernel >= 3 else 1
    x = g["x"][: R * n_sample]
    t0 = time.perf_counter()
    op = o.fmm(kernel, g["locs"], g["normals"], g["locs"], g["normals"], g["weights"], *cfg)
    t_plan = time.perf_counter() - t0
    gomp = ctypes.CDLL("libgomp.so.1")
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    med = {}
    for nt in (threads, 1):
        gomp.omp_set_num_threads(nt)
        op.apply(x)  # warm
        times = []
        for _ in range(runs):
            t0 = time.perf_counter()
            op.apply(x)
            times.append(time.perf_counter() - t0)
        med[nt] = statistics.median(times)
    gomp.omp_set_num_threads(threads)
    return {"value": round(n_sample / med[threads] / 1e6, 5), "unit": "Mpts/s", "cores": threads, "kind": kind,
            "value_1_thread": round(n_sample / med[1] / 1e6, 5), "cpu_model": cpu_model(),
            "sample": f"{n_sample} points of the same geometry (fault + free surface), same kernel and FMMConfig; "
                      f"FMMOperator::apply median of {runs} warm runs at {threads} OpenMP threads "
                      f"({med[threads]:.3f} s) and at 1 thread ({med[1]:.3f} s); plan build {t_plan:.2f} s excluded"}


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


def fmm_rooflines(info, n, kid, prec, leaf_ms, m2l_ms, wname=None):
    """Rooflines of the two kernels that dominate a matvec: the fused P2P+M2P+L2P
    leaf stage (algorithmic flops F_K per point pair) and the M2L translation kernel:
    with translation classes 2 blk^2 flops per cell pair (FP64 or FP32 MFMA), without
    them F_K per surface-point pair, n_e^2 per cell pair (kernel on the fly)."""
    n_e, R = info["n_e"], (1 if kid in (0, 1, 2, 7) else 3)
    out = {}
    flop = FLOPS_PER_PAIR[kid] * (info["p2p_pairs"] + n * n_e)
    ach = flop / (leaf_ms * 1e-3) / 1e12
    # the leaf stage's kernels: one wave per small leaf or chunk of a large one (k_leaf_wave,
    # k_leaf_parts adding the chunks), one workgroup per remaining leaf (k_leaf_stage), and
    # where Kelvin U's pairs of neighbouring leaves run once for both directions,
    # k_leaf_sym and k_sym_gather (the reverse sums)
    t = "double" if prec == "fp64" else "float"
    sym = info.get("symmetric_pairs", 0.0)
    parts = ([f"k_leaf_wave<{t}>"] if info["leaf_wave_items"] else []) + \
        (["k_leaf_parts"] if info["leaf_chunked"] else []) + ([f"k_leaf_stage<{t}>"] if info["leaf_block"] else []) + \
        (["k_leaf_sym", "k_sym_gather"] if sym else [])
    split = {"wave_items": info["leaf_wave_items"], "chunked_leaves": info["leaf_chunked"],
             "workgroup_leaves": info["leaf_block"], "symmetric_leaves": info.get("leaf_symmetric", 0),
             "symmetric_pairs": sym}
    if prec == "fp64":
        out["leaf"] = {"bound": "mfma", "pipe": "FP64 VALU (MI355X FP64 vector peak = FP64 matrix peak)",
                       "kernel": " + ".join(parts) + " (P2P+M2P+L2P)", "peak": PEAK_FP64_TFLOPS}
    else:
        out["leaf"] = {"bound": "mfma", "pipe": "FP32 VALU; the published 157.3 TFLOP/s counts packed FP32, which "
                                                "this build leaves off, so 78.6 TFLOP/s is its issue ceiling",
                       "kernel": " + ".join(parts) + " (P2P in FP32, M2P+L2P FP64)", "peak": PEAK_FP32_TFLOPS}
    out["leaf"]["split"] = split
    out["leaf"].update({"achieved": round(ach, 3), "unit": "TFLOP/s", "frac": round(ach / out["leaf"]["peak"], 4),
                        "kernel_avg_ms": round(leaf_ms, 4), "algorithmic_flop": flop,
                        "flop_rule": f"{FLOPS_PER_PAIR[kid]} flop per (target, source) pair: P2P pairs + L2P pairs"})
    if sym:
        # the arithmetic the kernels execute: each symmetric pair once for both directions
        ex = flop - FLOPS_PER_PAIR[kid] * 2 * sym + FLOPS_PER_SYM_PAIR[kid] * sym
        out["leaf"].update({"executed_flop": ex, "executed_frac": round(ex / (leaf_ms * 1e-3) / 1e12 /
                                                                          out["leaf"]["peak"], 4),
                            "executed_rule": f"{FLOPS_PER_SYM_PAIR[kid]} flop per symmetric pair (both directions), "
                                             f"{FLOPS_PER_PAIR[kid]} per other pair"})
    if m2l_ms > 0:
        if info["m2l_classes"] == 0:
            flop = FLOPS_PER_PAIR[kid] * info["m2l"] * n_e * n_e
            ach = flop / (m2l_ms * 1e-3) / 1e12
            out["m2l"] = {"bound": "mfma", "pipe": "FP64 VALU", "kernel": "k_surf_stage (FP64 M2L, kernel on the fly)",
                          "peak": PEAK_FP64_TFLOPS, "flop_rule": f"{FLOPS_PER_PAIR[kid]} flop per surface-point pair, "
                                                                 "n_e^2 pairs per M2L cell pair"}
        elif prec == "fp64":
            blk = R * n_e
            flop = 2.0 * info["m2l"] * blk * blk
            ach = flop / (m2l_ms * 1e-3) / 1e12
            out["m2l"] = {"bound": "mfma", "pipe": "FP64 MFMA (v_mfma_f64_16x16x4f64)",
                          "kernel": "k_m2l_mfma64 (FP64 MFMA M2L, translation classes)", "peak": PEAK_FP64_TFLOPS,
                          "flop_rule": "2 blk^2 flop per M2L cell pair", "classes": info["m2l_classes"]}
        else:
            blk = R * n_e
            flop = 2.0 * info["m2l"] * blk * blk
            ach = flop / (m2l_ms * 1e-3) / 1e12
            out["m2l"] = {"bound": "mfma", "pipe": "FP32 MFMA (v_mfma_f32_16x16x4f32)",
                          "kernel": "k_m2l_mfma (FP32 MFMA M2L, translation classes)", "peak": PEAK_FP32_TFLOPS,
                          "flop_rule": "2 blk^2 flop per M2L cell pair", "classes": info["m2l_classes"]}
        out["m2l"].update({"achieved": round(ach, 3), "unit": "TFLOP/s", "frac": round(ach / out["m2l"]["peak"], 4),
                           "kernel_avg_ms": round(m2l_ms, 4), "algorithmic_flop": flop, "pairs": info["m2l"]})
    for k, r in out.items():
        r["traffic"] = TRAFFIC_BYTES.get((wname, n, prec, k))
        r["traffic_unit"] = TRAFFIC_UNIT
    return out


def run_precision(tbem, torch, K, data, cfg, x, prec, steps, warmup, bwd_steps, wname=None):
    """Builds the operator, times `steps` matvecs after `warmup` (HIP events on the
    handle's stream give each launch's leaf-stage and M2L-kernel time), then the
    backward as autograd runs it. Returns (result dict, operator, output)."""
    t0 = time.perf_counter()
    op = tbem.FMMOperator(K, data, cfg, precision=prec)
    plan = time.perf_counter() - t0
    per = []
    # the timed steps (the P2P beside the far field, as the operator runs by default);
    # then the kernels' own times for the rooflines from fused applies, where each
    # kernel has the device to itself
    dt, _ = time_steps(lambda: op.apply(x), steps, warmup, torch)
    op.set_overlap(False)
    _, per = time_steps(lambda: op.apply(x), max(3, steps // 2), 1, torch,
                        after=lambda: (op.stage_times()["leaf"], op.m2l_kernel_ms()))
    stages = op.stage_times()  # of a fused apply: each stage alone on the device
    op.set_overlap(True)
    y = op.apply(x)
    info = op.info()
    R = K.dim
    n_obs = data.n_obs
    gv = torch.randn(R * n_obs, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(7))
    dt_adj, dt_grad, dt_grad_re = backward_times(op, x, gv, bwd_steps, torch)
    leaf_ms = float(np.mean([p[0] for p in per]))
    m2l_ms = float(np.mean([p[1] for p in per]))
    res = {"ms_per_step": round(dt * 1e3, 4), "Mpts_per_s": round(n_obs / dt / 1e6, 3),
           "bwd_fwd_ratio": round((dt_adj + dt_grad) / dt, 3),
           "bwd_ms": {"adjoint": round(dt_adj * 1e3, 3), "obs_grad": round(dt_grad * 1e3, 3),
                      "obs_grad_recomputing": round(dt_grad_re * 1e3, 3)},
           "plan_build_s": round(plan, 3), "stage_ms": {k: round(v, 4) for k, v in stages.items()},
           "rooflines": fmm_rooflines(info, n_obs, K.kernel_id, prec, leaf_ms, m2l_ms, wname),
           "m2l_output_GB": round(info["m2l_output_bytes"] / 1e9, 3), "work_GB": round(info["work_bytes"] / 1e9, 3)}
    return res, op, y, info


def workload_pair(tbem, torch, w, steps, warmup, n_sub, seed=123):
    """FP64 and mixed precision on one workload: both timed, both against the FP64
    GPU direct sum on an n_sub-target subsample and against each other."""
    kid = w["kernel"]
    n = len(w["locs"])
    K = tbem.make_kernel(kid, 1.0, 0.25)
    R = K.dim
    data = tbem.NBodyData(w["locs"], w["normals"], w["locs"], w["normals"], w["weights"])
    cfg = tbem.FMMConfig(*w["cfg"])
    x = torch.tensor(w["x"], device="cuda")
    sub = np.sort(np.random.default_rng(seed).choice(n, min(n, n_sub), replace=False))
    dsub = tbem.NBodyData(w["locs"][sub], w["normals"][sub], w["locs"], w["normals"], w["weights"])
    t0 = time.perf_counter()
    yd = tbem.make_direct_nbody_operator(dsub, K).apply(w["x"])
    t_direct = time.perf_counter() - t0
    idx = np.concatenate([q * n + sub for q in range(R)])
    out, ys, infos = {}, {}, {}
    for prec in ("fp64", "mixed"):
        res, op, y, infos[prec] = run_precision(tbem, torch, K, data, cfg, x, prec, steps, warmup,
                                                max(2, steps // 4), w.get("name"))
        ys[prec] = y.cpu().numpy()
        res["rel_err_vs_direct"] = float(np.linalg.norm(ys[prec][idx] - yd) / np.linalg.norm(yd))
        out[prec] = res
        log(f"[bench] {w['name']}, {n} points, {prec}: {res['ms_per_step']:.2f} ms ({res['Mpts_per_s']:.1f} Mpts/s), "
            f"rel err {res['rel_err_vs_direct']:.2e}, bwd/fwd {res['bwd_fwd_ratio']}, "
            f"leaf {res['rooflines']['leaf']['kernel_avg_ms']:.2f} ms, stages {res['stage_ms']}")
        del op
        torch.cuda.empty_cache()
    out["rel_err_mixed_vs_fp64"] = float(np.linalg.norm(ys["mixed"] - ys["fp64"]) / np.linalg.norm(ys["fp64"]))
    out["speedup_mixed_vs_fp64"] = round(out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"], 3)
    out["direct_subsample"] = {"targets": int(len(sub)), "s": round(t_direct, 3)}
    out["info"] = {k: infos["fp64"][k] for k in ("n_e", "n_src_cells", "n_up_levels", "p2p", "m2l", "p2p_pairs")}
    out["info"]["m2l_classes_mixed"] = infos["mixed"]["m2l_classes"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="fault_surface_o100",
                    help="headline workload: fault_surface_o100 (default), fault_surface, icosphere, "
                         "cube_clusters, rough_fault")
    ap.add_argument("--n", type=int, default=0)
    ap.add_argument("--cpu-n", type=int, default=8_000)
    ap.add_argument("--cpu-runs", type=int, default=10)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-extra", action="store_true", help="the headline workload only")
    ap.add_argument("--skip-cfg2", action="store_true",
                    help="skip the 1M-point order-6 fault + free-surface workload")
    ap.add_argument("--skip-laplace", action="store_true",
                    help="skip the 4M-point Laplace FP32/MFMA-vs-FP64 comparison")
    ap.add_argument("--skip-forward-model", action="store_true",
                    help="skip the differentiable forward model (120k fault points, 200k stations)")
    ap.add_argument("--skip-bem-solve", action="store_true",
                    help="skip the 2M-point hypersingular GMRES solve")
    ap.add_argument("--skip-near", action="store_true",
                    help="skip the CSR near field / integral operator section")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 256 MiB device copy (HBM counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
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
    n = len(w["locs"])
    K = tbem.make_kernel(kid, 1.0, 0.25)
    data = tbem.NBodyData(w["locs"], w["normals"], w["locs"], w["normals"], w["weights"])
    cfg = tbem.FMMConfig(*w["cfg"])
    log(f"[bench] {w['desc']}")
    x = torch.tensor(w["x"], device="cuda")

    # the headline: plain FP64 matvec, inputs resident in HBM
    res64, op64, y64, info = run_precision(tbem, torch, K, data, cfg, x, "fp64", args.steps, args.warmup,
                                           max(2, args.steps // 4), args.config)
    dt64 = res64["ms_per_step"] * 1e-3
    value = n / dt64 / 1e6
    log(f"[bench] fp64 {res64['ms_per_step']:.2f} ms -> {value:.2f} Mpts/s; stages {res64['stage_ms']}")
    rl = res64["rooflines"]
    dom = "m2l" if "m2l" in rl and rl["m2l"]["kernel_avg_ms"] > rl["leaf"]["kernel_avg_ms"] else "leaf"
    roofline = dict(rl[dom])
    out = {"metric": METRIC, "value": round(value, 3), "unit": "Mpts/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": res64["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": w["desc"], "n_points": n, "kernel": type(K).__name__, "mac": cfg.mac,
                      "order": cfg.order, "n_e": info["n_e"], "min_pts_per_cell": cfg.min_pts_per_cell,
                      "plan_build_s": res64["plan_build_s"], "stage_ms": res64["stage_ms"]},
           "roofline": roofline}
    del op64
    torch.cuda.empty_cache()

    if not args.skip_extra:
        # the same workload in mixed precision, errors of both against the direct sum
        R = K.dim
        sub = np.sort(np.random.default_rng(123).choice(n, min(n, 10_000), replace=False))
        dsub = tbem.NBodyData(w["locs"][sub], w["normals"][sub], w["locs"], w["normals"], w["weights"])
        t0 = time.perf_counter()
        yd = tbem.make_direct_nbody_operator(dsub, K).apply(w["x"])
        t_direct = time.perf_counter() - t0
        idx = np.concatenate([q * n + sub for q in range(R)])
        y64h = y64.cpu().numpy()
        e64 = float(np.linalg.norm(y64h[idx] - yd) / np.linalg.norm(yd))
        res32, op32, y32, info32 = run_precision(tbem, torch, K, data, cfg, x, "mixed", args.steps, args.warmup,
                                                 max(2, args.steps // 4), args.config)
        y32h = y32.cpu().numpy()
        res32["rel_err_vs_direct"] = float(np.linalg.norm(y32h[idx] - yd) / np.linalg.norm(yd))
        res32["rel_err_vs_fp64_fmm"] = float(np.linalg.norm(y32h - y64h) / np.linalg.norm(y64h))
        res32["speedup_vs_fp64"] = round(res64["ms_per_step"] / res32["ms_per_step"], 3)
        res32["m2l_classes"] = info32["m2l_classes"]
        del op32
        torch.cuda.empty_cache()
        res64["rel_err_vs_direct"] = e64
        out["fp64"] = res64
        out["mixed"] = res32
        out["rel_err_fp64_vs_direct"] = e64
        out["bwd_fwd_ratio"] = res64["bwd_fwd_ratio"]
        out["speedup_mixed_vs_fp64"] = res32["speedup_vs_fp64"]
        out["direct_subsample"] = {"targets": int(len(sub)), "s": round(t_direct, 3)}
        log(f"[bench] mixed {res32['ms_per_step']:.2f} ms ({res32['speedup_vs_fp64']}x), rel err fp64 {e64:.2e}, "
            f"mixed {res32['rel_err_vs_direct']:.2e}; bwd/fwd fp64 {res64['bwd_fwd_ratio']}, mixed "
            f"{res32['bwd_fwd_ratio']}")
    del y64
    torch.cuda.empty_cache()

    extra = not args.skip_extra and args.config == "fault_surface_o100"
    if extra and not args.skip_cfg2:
        out["fault_surface_1M_order6"] = workload_pair(tbem, torch, workload("fault_surface", 1_000_000),
                                                       args.steps, args.warmup, 10_000)
    if extra and not args.skip_laplace:
        out["laplace_4M_fp32_mfma_vs_fp64"] = workload_pair(tbem, torch, workload("cube_clusters", 4_000_000), 3, 1,
                                                            10_000)
    if extra and not args.skip_forward_model:
        out["forward_model_fault_stations"] = forward_model_run(tbem, torch, 5)
    if extra and not args.skip_bem_solve:
        out["bem_solve_rough_fault"] = bem_solve_run(tbem, torch)
    if extra and not args.skip_near:
        out["integral_operator_near_field"] = near_field_run(tbem, torch, args.steps, args.warmup, args.skip_cpu)

    if not args.skip_cpu:
        cfg_t = (cfg.mac, cfg.order, cfg.min_pts_per_cell, cfg.d, cfg.account_for_small_cells)
        log(f"[bench] CPU baseline on {args.cpu_n} points ...")
        out["cpu_baseline"] = cpu_baseline(kid, cfg_t, min(args.cpu_n, n), args.cpu_runs)
        log(f"[bench] CPU baseline {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
