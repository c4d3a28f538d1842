"""Writes tests/golden/bem_far_ops.npz from the compiled reference (oracle/_ref):
make_galerkin_operator, make_interpolation_operator (1 and 3 components) and
nbody_data_from_bem on a small perturbed triangle mesh with two quadrature rules.

    make -C oracle ref && python tests/golden/make_bem_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle.oracle import Oracle  # noqa: E402


def mesh(n=4, seed=3):
    """An n x n grid of squares, two triangles each, on a bumpy sheet, tilted."""
    rng = np.random.default_rng(seed)
    g = np.linspace(-1.0, 1.0, n + 1)
    X, Y = np.meshgrid(g, g, indexing="ij")
    Z = 0.3 * rng.standard_normal(X.shape)
    V = np.stack([X, Y, Z], axis=-1)
    f = []
    for i in range(n):
        for j in range(n):
            f.append([V[i, j], V[i + 1, j], V[i + 1, j + 1]])
            f.append([V[i, j], V[i + 1, j + 1], V[i, j + 1]])
    c, s = np.cos(0.4), np.sin(0.4)
    R = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]]) @ np.array([[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]])
    return np.array(f) @ R.T + np.array([0.2, -0.1, 0.5])


def quads(seed=4):
    three = (np.array([[1 / 6, 1 / 6], [2 / 3, 1 / 6], [1 / 6, 2 / 3]]), np.full(3, 1 / 6))
    rng = np.random.default_rng(seed)
    a = rng.uniform(0, 1, (7, 2))
    a[a.sum(1) > 1] = 1 - a[a.sum(1) > 1]  # fold into the triangle
    w = rng.uniform(0.5, 1.5, 7)
    return {"q3": three, "q7": (a, w * 0.5 / w.sum())}


def main():
    ref = Oracle("ref")
    obs, src = mesh(4, 3), mesh(3, 5)
    out = {"obs_facets": obs, "src_facets": src}
    for qn, q in quads().items():
        out[f"{qn}_x"], out[f"{qn}_w"] = q
        for which in ("galerkin", "interp"):
            for nc in (1, 3):
                f = obs if which == "galerkin" else src
                nr, ncol, rp, ci, v = ref.bem_sparse(which, nc, f, q)
                key = f"{which}_{nc}_{qn}"
                out[key + "_shape"] = np.array([nr, ncol])
                out[key + "_rp"], out[key + "_ci"], out[key + "_v"] = rp, ci, v
        for k, a in zip(("obs_locs", "obs_normals", "src_locs", "src_normals", "src_weights"),
                        ref.nbody_from_bem(obs, src, q, q)):
            out[f"nbody_{qn}_{k}"] = a
    path = os.path.join(HERE, "bem_far_ops.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
