"""Synthetic point sets of the benchmark configurations (seeded, deterministic).

Conventions (the reference has no generator for any of them):
  * geometry from numpy.random.default_rng(seed), densities from default_rng(seed + 1);
  * surface points carry unit normals and quadrature weights (area / point), as
    nbody_src_from_bem forms them (weight = rule weight x Jacobian,
    cpp/nbody_operator.h:64-73);
  * volume points (cube, clusters) get random unit normals and weight 1.
"""
import numpy as np

SEED = 20261202


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def icosphere(refinements=4):
    """Unit icosahedron subdivided `refinements` times by edge midpoints pushed to the sphere."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [_unit(np.array(p, dtype=np.float64)) for p in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(refinements):
        cache = {}
        new = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                verts.append(_unit((verts[a] + verts[b]) / 2.0))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = new
    return np.array(verts), np.array(faces, dtype=np.int64)


def tri_three_point(verts, tris, outward_from=None):
    """3-point rule at barycentric (2/3,1/6,1/6) and permutations, weight = area/3."""
    p0, p1, p2 = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
    cr = np.cross(p1 - p0, p2 - p0)
    area = 0.5 * np.linalg.norm(cr, axis=1)
    n = _unit(cr)
    if outward_from is not None:
        cen = (p0 + p1 + p2) / 3.0
        flip = np.sum(n * (cen - outward_from), axis=1) < 0
        n[flip] *= -1
    bary = np.array([[2 / 3, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6], [1 / 6, 1 / 6, 2 / 3]])
    pts = np.einsum("qk,tkd->tqd", bary, np.stack([p0, p1, p2], axis=1)).reshape(-1, 3)
    normals = np.repeat(n, 3, axis=0)
    weights = np.repeat(area / 3.0, 3)
    return pts, normals, weights


def icosphere_surface_15k(seed=SEED):
    """15,360 points: 5,120-triangle icosphere, 3 points per triangle, outward normals."""
    v, f = icosphere(4)
    pts, nrm, w = tri_three_point(v, f, outward_from=np.zeros(3))
    x = np.random.default_rng(seed + 1).standard_normal(len(pts))
    return dict(locs=pts, normals=nrm, weights=w, x=x)


def fault_and_free_surface(n_total=1_000_000, seed=SEED):
    """Half the points uniform on the free surface [-1/2,1/2]^2 x {0}, half uniform on a
    0.5 x 0.5 fault plane dipping 45 degrees about the x axis, centred at depth 0.3
    (centre (0, 0, -0.3); along-strike x, down-dip direction (0, cos45, -sin45))."""
    rng = np.random.default_rng(seed)
    n1 = n_total // 2
    n2 = n_total - n1
    s = np.column_stack([rng.uniform(-0.5, 0.5, n1), rng.uniform(-0.5, 0.5, n1), np.zeros(n1)])
    sn = np.tile([0.0, 0.0, 1.0], (n1, 1))
    sw = np.full(n1, 1.0 / n1)
    a, b = rng.uniform(-0.25, 0.25, n2), rng.uniform(-0.25, 0.25, n2)
    c45 = np.sqrt(0.5)
    f = np.column_stack([a, c45 * b, -0.3 - c45 * b])
    fn = np.tile([0.0, c45, c45], (n2, 1))
    fw = np.full(n2, 0.25 / n2)
    locs = np.vstack([s, f])
    x = np.random.default_rng(seed + 1).standard_normal(3 * n_total)
    return dict(locs=locs, normals=np.vstack([sn, fn]), weights=np.concatenate([sw, fw]), x=x)


def cube_and_clusters(n_total=4_000_000, seed=SEED):
    """Half uniform in [0,1]^3, half from 8 Gaussian clusters (sigma 0.02, centres uniform)."""
    rng = np.random.default_rng(seed)
    n1 = n_total // 2
    n2 = n_total - n1
    u = rng.uniform(0, 1, (n1, 3))
    centres = rng.uniform(0, 1, (8, 3))
    lab = rng.integers(0, 8, n2)
    g = centres[lab] + 0.02 * rng.standard_normal((n2, 3))
    locs = np.vstack([u, g])
    normals = _unit(rng.standard_normal((n_total, 3)))
    x = np.random.default_rng(seed + 1).standard_normal(n_total)
    return dict(locs=locs, normals=normals, weights=np.ones(n_total), x=x)


def fault_and_stations(n_tri=40_000, n_obs=200_000, seed=SEED):
    """Planar rectangular fault (1 x 0.5, dipping 45 degrees, top at depth 0.1) meshed into
    n_tri triangles (3 points each) and n_obs surface stations uniform in [-1,1]^2 x {0}."""
    rng = np.random.default_rng(seed)
    nx = int(np.sqrt(n_tri / 4.0))
    ny = max(1, n_tri // (2 * 2 * nx))
    while 2 * nx * ny < n_tri:
        nx += 1
    xs = np.linspace(-0.5, 0.5, nx + 1)
    ys = np.linspace(0.0, 0.5, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    c45 = np.sqrt(0.5)
    verts = np.column_stack([X.ravel(), c45 * Y.ravel(), -0.1 - c45 * Y.ravel()])
    tris = []
    for i in range(nx):
        for j in range(ny):
            a = i * (ny + 1) + j
            b, c, d = a + ny + 1, a + 1, a + ny + 2
            tris += [(a, b, c), (c, b, d)]
    tris = np.array(tris[:n_tri], dtype=np.int64)
    src, sn, sw = tri_three_point(verts, tris)
    obs = np.column_stack([rng.uniform(-1, 1, n_obs), rng.uniform(-1, 1, n_obs), np.zeros(n_obs)])
    on = np.tile([0.0, 0.0, 1.0], (n_obs, 1))
    slip = np.random.default_rng(seed + 1).standard_normal(3 * len(src))
    slip_true = np.random.default_rng(seed + 2).standard_normal(3 * len(src))
    return dict(src=src, src_normals=sn, src_weights=sw, obs=obs, obs_normals=on, x=slip,
                x_true=slip_true)


def rough_fault_points(n_total=2_000_000, seed=SEED):
    """n_total points on z = 0.05 sin(6x) cos(4y) + 0.005 * N(0,1) roughness over [0,1]^2."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, n_total)
    y = rng.uniform(0, 1, n_total)
    z = 0.05 * np.sin(6 * x) * np.cos(4 * y) + 0.005 * rng.standard_normal(n_total)
    dzdx = 0.3 * np.cos(6 * x) * np.cos(4 * y)
    dzdy = -0.2 * np.sin(6 * x) * np.sin(4 * y)
    n = _unit(np.column_stack([-dzdx, -dzdy, np.ones(n_total)]))
    w = np.sqrt(1 + dzdx ** 2 + dzdy ** 2) / n_total
    return dict(locs=np.column_stack([x, y, z]), normals=n, weights=w)


def uniform_cube(n, seed=SEED, dim=1):
    rng = np.random.default_rng(seed)
    locs = rng.uniform(0, 1, (n, 3))
    normals = _unit(rng.standard_normal((n, 3)))
    weights = rng.uniform(0.5, 1.5, n)
    x = np.random.default_rng(seed + 1).standard_normal(dim * n)
    return dict(locs=locs, normals=normals, weights=weights, x=x)


def rough_fault_mesh(nx=300, ny=300, seed=SEED):
    """Triangulated rough fault: an nx x ny grid over [0,1]^2, two triangles per cell,
    vertices on z = 0.05 sin(6x) cos(4y) + 0.005 N(0,1). Facets (2 nx ny, 3, 3), the
    two triangles of cell (i, j) at 2 (i ny + j) and 2 (i ny + j) + 1."""
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(np.linspace(0, 1, nx + 1), np.linspace(0, 1, ny + 1), indexing="ij")
    Z = 0.05 * np.sin(6 * X) * np.cos(4 * Y) + 0.005 * rng.standard_normal(X.shape)
    V = np.stack([X, Y, Z], axis=-1)
    a, b, c, d = V[:-1, :-1], V[1:, :-1], V[:-1, 1:], V[1:, 1:]
    t0 = np.stack([a, b, d], axis=2)
    t1 = np.stack([a, d, c], axis=2)
    return np.stack([t0, t1], axis=2).reshape(-1, 3, 3)


def near_field_csr(nx=300, ny=300, components=3, seed=SEED):
    """Sparsity of a Galerkin near-field operator on rough_fault_mesh(nx, ny): the row
    of dof (d, facet f, corner b), at d n_dofs + 3 f + b, holds every (d', g, b') with g
    one of the up to 18 facets of the 3 x 3 grid cells around f's cell, columns
    ascending (d' n_dofs + 3 g + b', the order right_multiply leaves them in).
    Values uniform in [-0.5, 0.5). Returns (n_rows, n_cols, row_ptrs, column_indices,
    values) with n_rows = n_cols = components * 3 * 2 nx ny."""
    F = 2 * nx * ny
    n_dofs = 3 * F
    ci_, cj_ = np.divmod(np.arange(F) // 2, ny)
    near = np.full((F, 18), -1, dtype=np.int64)
    k = 0
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            i2, j2 = ci_ + di, cj_ + dj
            ok = (i2 >= 0) & (i2 < nx) & (j2 >= 0) & (j2 < ny)
            base = 2 * (i2 * ny + j2)
            for t in (0, 1):
                near[:, k] = np.where(ok, base + t, -1)
                k += 1
    near.sort(axis=1)  # invalid (-1) first, then ascending facets
    valid = near >= 0
    count = valid.sum(1)
    # columns of one facet's rows: (d', g, b') ascending
    cols = (np.arange(components)[:, None, None, None] * n_dofs + 3 * near[None, :, :, None].clip(0)
            + np.arange(3)[None, None, None, :])                     # (C, F, 18, 3)
    cols = np.transpose(cols, (1, 0, 2, 3)).reshape(F, -1)              # (F, C*18*3)
    vmask = np.broadcast_to(valid[:, None, :, None], (F, components, 18, 3)).reshape(F, -1)
    per_row = components * 3 * count                                    # entries of each row of facet f
    row_len = np.broadcast_to(per_row[None, :, None], (components, F, 3)).reshape(-1)
    row_ptrs = np.zeros(row_len.size + 1, dtype=np.int64)
    np.cumsum(row_len, out=row_ptrs[1:])
    packed = cols[vmask]                                                # facet-major, valid only
    fstart = np.zeros(F + 1, dtype=np.int64)
    np.cumsum(per_row, out=fstart[1:])
    # row (d, f, b) copies facet f's packed columns
    rows_f = np.broadcast_to(np.arange(F)[None, :, None], (components, F, 3)).reshape(-1)
    idx = np.repeat(fstart[rows_f] - row_ptrs[:-1], row_len) + np.arange(row_ptrs[-1])
    column_indices = packed[idx]
    values = np.random.default_rng(seed + 1).random(row_ptrs[-1]) - 0.5
    n = components * n_dofs
    return n, n, row_ptrs, column_indices, values


def grid_continuity_constraints(nx, ny, components=3, all_pairs=True, bc_value=0.0):
    """Continuity constraints of rough_fault_mesh(nx, ny) with discontinuous linear dofs
    (dof 3 f + k at corner k of facet f, n_dofs = 6 nx ny per component): for every
    vertex, continuity_constraint(p, q) between the dofs that meet there, for every
    ordered pair p != q when all_pairs (as an overlap map of coincident vertices lists
    them), else for consecutive ones; each component shifted by n_dofs; plus
    boundary_condition(dof, bc_value) on the dofs at vertex (0, 0). Returns
    (eq_ptr, term_dof, term_weight, rhs, n_total)."""
    F = 2 * nx * ny
    n_dofs = 3 * F
    vid = lambda i, j: i * (ny + 1) + j  # noqa: E731
    at = [[] for _ in range((nx + 1) * (ny + 1))]
    for i in range(nx):
        for j in range(ny):
            f = 2 * (i * ny + j)
            for k, v in enumerate((vid(i, j), vid(i + 1, j), vid(i + 1, j + 1))):
                at[v].append(3 * f + k)
            for k, v in enumerate((vid(i, j), vid(i + 1, j + 1), vid(i, j + 1))):
                at[v].append(3 * (f + 1) + k)
    pairs = []
    for dofs in at:
        if all_pairs:
            pairs += [(p, q) for p in dofs for q in dofs if p != q]
        else:
            pairs += list(zip(dofs[:-1], dofs[1:]))
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    bc = np.asarray(at[0], dtype=np.int64)
    eq_dofs, eq_w, rhs = [], [], []
    for c in range(components):
        s = c * n_dofs
        eq_dofs.append(pairs + s)
        eq_w.append(np.tile([1.0, -1.0], (len(pairs), 1)))
        rhs.append(np.zeros(len(pairs)))
    n_pair_eqs = components * len(pairs)
    term_dof = np.concatenate([np.concatenate(eq_dofs).reshape(-1), np.concatenate([bc + c * n_dofs for c in range(components)])])
    term_weight = np.concatenate([np.concatenate(eq_w).reshape(-1), np.ones(components * len(bc))])
    rhs = np.concatenate(rhs + [np.full(components * len(bc), bc_value)])
    eq_ptr = np.concatenate([np.arange(0, 2 * n_pair_eqs + 1, 2),
                             2 * n_pair_eqs + np.arange(1, components * len(bc) + 1)]).astype(np.int64)
    return eq_ptr, term_dof, term_weight, rhs, components * n_dofs
