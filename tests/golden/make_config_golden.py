"""Writes the configuration-shaped fixtures from the compiled reference
(oracle/_ref/libtbem_ref.so): subsets of the cube-and-clusters Laplace workload and
the rough-fault hypersingular workload, and the fault + free-surface elastic
workload at order 50 (displacement and hypersingular kernels), each with the
reference's FMMOperator::apply output, the FP64 direct sum and, where the order
makes the output depend on the check-to-equivalent matrices' last bits, the
reference's per-level UC2E/DC2E so that a test can run the GPU path on them.

Run here, where /root/reference exists to build oracle/_ref:
    make -C oracle ref && OMP_NUM_THREADS=1 python tests/golden/make_config_golden.py [name ...]

Single-threaded, as tests/golden/make_golden.py, for deterministic sums.
"""
import importlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle.oracle import Oracle, components  # noqa: E402

wl = importlib.import_module("3bem_amd.workloads")


def write(name, kernel, g, x, mac, order, leaf, with_c2e):
    ref = Oracle("ref")
    t0 = time.time()
    loc, nrm, w = g["locs"], g["normals"], g["weights"]
    op = ref.fmm(kernel, loc, nrm, loc, nrm, w, mac, order, leaf, 0.05, False)
    y = op.apply(x)
    c = op.counts()
    out = dict(kernel=kernel, obs=loc, obs_normals=nrm, src=loc, src_normals=nrm, weights=w, x=x, mac=mac,
               order=order, leaf=leaf, d=0.05, small=0, y=y, counts=np.array([c[k] for k in sorted(c)]),
               count_names=np.array(sorted(c)))
    if with_c2e:
        out["uc2e"] = np.stack([op.c2e(False, L) for L in range(c["n_up_levels"])])
        out["dc2e"] = np.stack([op.c2e(True, L) for L in range(c["n_down_levels"])])
    t1 = time.time()
    out["direct"] = ref.direct(kernel, loc, nrm, loc, nrm, w, x)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(f"{name}: fmm {t1 - t0:.1f} s, direct {time.time() - t1:.1f} s, "
          f"rel err vs direct {np.linalg.norm(y - out['direct']) / np.linalg.norm(out['direct']):.3e}", flush=True)


def cube_clusters():
    # the cube-and-clusters Laplace workload at 30,000 points, order 75, leaf 100
    g = wl.cube_and_clusters(30_000)
    write("cube_clusters_laplace_o75", 0, g, g["x"], 0.3, 75, 100, True)


def rough_fault():
    # the rough-fault hypersingular workload at 20,000 points, order 6, leaf 64
    g = wl.rough_fault_points(20_000)
    x = np.random.default_rng(wl.SEED + 1).standard_normal(3 * 20_000)
    write("rough_fault_elastic_h_o6", 6, g, x, 0.3, 6, 64, False)


def fault_surface(kernel, tag):
    # the fault + free-surface elastic workload at 4,000 points, order 50, leaf 64
    g = wl.fault_and_free_surface(4_000)
    write(f"fault_surface_elastic_{tag}_o50", kernel, g, g["x"], 0.3, 50, 64, True)


CASES = {"cube_clusters": cube_clusters, "rough_fault": rough_fault,
         "fault_surface_u": lambda: fault_surface(3, "u"), "fault_surface_h": lambda: fault_surface(6, "h")}

if __name__ == "__main__":
    assert os.environ.get("OMP_NUM_THREADS") == "1", "run with OMP_NUM_THREADS=1"
    for n in (sys.argv[1:] or list(CASES)):
        CASES[n]()
