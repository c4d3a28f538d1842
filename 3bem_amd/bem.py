"""Boundary-element plumbing around the FMM: the far-field operators of a Galerkin
integral operator on a triangle mesh, and their composition with the GPU FMM.

Reference (cpp/integral_operator.h:67-98, boundary_operator): the far field of
A = G(F I + N + C) is

  * I, make_interpolation_operator (cpp/interpolation_operator.cpp:11-41): the
    linear basis of each facet evaluated at the source quadrature points;
  * F, an N-body operator on nbody_data_from_bem (cpp/nbody_operator.h:38-85):
    quadrature points, facet normals and quadrature weight times facet jacobian;
  * G, make_galerkin_operator (cpp/galerkin_operator.cpp:11-45): the observation
    quadrature weighted by basis function and jacobian.

The reference builds F as a dense direct operator and leaves its FMM commented out
(cpp/integral_operator.h:82-87); integral_operator() here puts the GPU FMM there.

A mesh is an array of facets, shape (n_facets, 3, 3): three corners of xyz each,
as Mesh<3>::facets (cpp/mesh.h:14-19). A quadrature rule is a pair (x_hat, w) of
points on the reference triangle (0,0)-(1,0)-(0,1), shape (n_q, 2), and weights,
shape (n_q,), as QuadRule<2> (cpp/quad_rule.h:8-15). Degrees of freedom are the
three corners of every facet (n_dofs = 3 n_facets, cpp/mesh.cpp:30-33); vectors
are component-major. The arithmetic follows the reference step for step, so the
operators and point data come out identical to the reference's.
"""
import numpy as np

from .ThreeD import FMMOperator, DirectOperator, IntegralOperator, NBodyData, SparseOperator


def _facets(f):
    f = np.ascontiguousarray(np.asarray(f, dtype=np.float64))
    if f.ndim != 3 or f.shape[1:] != (3, 3):
        raise ValueError("facets must have shape (n_facets, 3, 3)")
    return f


def _quad(q):
    x_hat, w = q
    x_hat = np.ascontiguousarray(np.asarray(x_hat, dtype=np.float64).reshape(-1, 2))
    w = np.ascontiguousarray(np.asarray(w, dtype=np.float64).reshape(-1))
    if x_hat.shape[0] != w.shape[0]:
        raise ValueError("quadrature points and weights differ in number")
    return x_hat, w


def _unscaled_normal(f):
    """cross(c2 - c0, c2 - c1) (cpp/geometry.h:41-48, :113-117)."""
    a = f[:, 2] - f[:, 0]
    b = f[:, 2] - f[:, 1]
    return np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1],
                     a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                     a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], axis=1)


def _norm(n):
    """sqrt((n0^2 + n1^2) + n2^2) (hypot, cpp/geometry.h:50-66)."""
    return np.sqrt((n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1]) + n[:, 2] * n[:, 2])


def facet_normals(facets):
    """Unit normals (facet_normal, cpp/geometry.h:128-133): the unscaled normal
    times the reciprocal of its length (cpp/vec_ops.h:151-155)."""
    n = _unscaled_normal(_facets(facets))
    return n * (1 / _norm(n))[:, None]


def facet_jacobians(facets):
    """2 x triangle area (facet_jacobian<3>, cpp/numerics.h:34-36, cpp/geometry.h:135-138)."""
    return 2.0 * (0.5 * _norm(_unscaled_normal(_facets(facets))))


def linear_basis(x_hat):
    """(1 - x - y, x, y) at each point (cpp/numerics.h:69-72): shape (n_q, 3)."""
    x_hat = np.asarray(x_hat, dtype=np.float64).reshape(-1, 2)
    return np.stack([(1 - x_hat[:, 0]) - x_hat[:, 1], x_hat[:, 0], x_hat[:, 1]], axis=1)


def ref_to_real(x_hat, facets):
    """Points x_hat mapped onto every facet (cpp/numerics.h:91-100): shape
    (n_facets, n_q, 3), each coordinate (b0 c0 + b1 c1) + b2 c2."""
    b = linear_basis(x_hat)
    f = _facets(facets)
    return ((b[None, :, 0, None] * f[:, None, 0, :] + b[None, :, 1, None] * f[:, None, 1, :]) +
            b[None, :, 2, None] * f[:, None, 2, :])


def nbody_obs_from_bem(obs_facets, obs_quad):
    """Observation points and normals, point i*n_q + q on facet i (cpp/nbody_operator.h:38-53)."""
    f = _facets(obs_facets)
    x_hat, _ = _quad(obs_quad)
    locs = ref_to_real(x_hat, f).reshape(-1, 3)
    normals = np.repeat(facet_normals(f), x_hat.shape[0], axis=0)
    return locs, normals


def nbody_src_from_bem(src_facets, src_quad):
    """Source points, normals and weights w_q * jacobian (cpp/nbody_operator.h:56-74)."""
    f = _facets(src_facets)
    x_hat, w = _quad(src_quad)
    locs = ref_to_real(x_hat, f).reshape(-1, 3)
    normals = np.repeat(facet_normals(f), x_hat.shape[0], axis=0)
    weights = (w[None, :] * facet_jacobians(f)[:, None]).reshape(-1)
    return locs, normals, weights


def nbody_data_from_bem(obs_facets, src_facets, obs_quad, src_quad):
    """NBodyData of a mesh pair (nbody_data_from_bem, cpp/nbody_operator.h:77-85)."""
    ol, on = nbody_obs_from_bem(obs_facets, obs_quad)
    sl, sn, sw = nbody_src_from_bem(src_facets, src_quad)
    return NBodyData(ol, on, sl, sn, sw)


def _entry_grid(n_facets, n_q, n_components):
    """Index grids in the reference's loop order: facet, quadrature point, basis
    function, component."""
    return np.meshgrid(np.arange(n_facets), np.arange(n_q), np.arange(3), np.arange(n_components),
                       indexing="ij")


def galerkin_coo(n_components, obs_facets, obs_quad):
    """Triplets of make_galerkin_operator (cpp/galerkin_operator.cpp:11-45) in its
    push order: row d*n_dofs + 3i + b, column d*n_quad + i*n_q + q, value
    basis_b(x_q) * (jacobian_i * w_q). Returns (n_rows, n_cols, rows, cols, values)."""
    f = _facets(obs_facets)
    x_hat, w = _quad(obs_quad)
    nf, nq = f.shape[0], x_hat.shape[0]
    n_dofs, n_quad = 3 * nf, nf * nq
    i, q, b, d = _entry_grid(nf, nq, n_components)
    weight = facet_jacobians(f)[:, None] * w[None, :]
    vals = linear_basis(x_hat)[q, b] * weight[i, q]
    return (n_components * n_dofs, n_components * n_quad, (d * n_dofs + 3 * i + b).reshape(-1),
            (d * n_quad + i * nq + q).reshape(-1), vals.reshape(-1))


def interpolation_coo(n_components, src_facets, src_quad):
    """Triplets of make_interpolation_operator (cpp/interpolation_operator.cpp:11-41)
    in its push order: row d*n_quad + i*n_q + q, column d*n_dofs + 3i + b, value
    basis_b(x_q). Returns (n_rows, n_cols, rows, cols, values)."""
    f = _facets(src_facets)
    x_hat, _ = _quad(src_quad)
    nf, nq = f.shape[0], x_hat.shape[0]
    n_dofs, n_quad = 3 * nf, nf * nq
    i, q, b, d = _entry_grid(nf, nq, n_components)
    vals = linear_basis(x_hat)[q, b]
    return (n_components * n_quad, n_components * n_dofs, (d * n_quad + i * nq + q).reshape(-1),
            (d * n_dofs + 3 * i + b).reshape(-1), vals.reshape(-1))


def make_galerkin_operator(n_components, obs_facets, obs_quad):
    """make_galerkin_operator (cpp/galerkin_operator.cpp:11-45) as a GPU SparseOperator."""
    return SparseOperator.csr_from_coo(*galerkin_coo(n_components, obs_facets, obs_quad))


def make_interpolation_operator(n_components, src_facets, src_quad):
    """make_interpolation_operator (cpp/interpolation_operator.cpp:11-41) as a GPU SparseOperator."""
    return SparseOperator.csr_from_coo(*interpolation_coo(n_components, src_facets, src_quad))


def integral_operator(nearfield, farfield_correction, obs_facets, src_facets, obs_far_quad, src_far_quad, K,
                      fmm_config=None, precision="fp64"):
    """The far-field half of boundary_operator (cpp/integral_operator.h:67-98) with the
    GPU FMM as the far field: returns IntegralOperator(nearfield, farfield_correction,
    G, F, I) with G = make_galerkin_operator(R, obs, obs_far_quad), I =
    make_interpolation_operator(C, src, src_far_quad) and F = FMMOperator(K,
    nbody_data_from_bem(...), fmm_config) (the direct N-body operator when fmm_config
    is None, as the reference's dense far field). nearfield and farfield_correction
    are the assembled near-field sparse operators, already multiplied by their
    Galerkin operators (near_galerkin.right_multiply(nearfield) and
    far_galerkin.right_multiply(far_correction) in the reference)."""
    data = nbody_data_from_bem(obs_facets, src_facets, obs_far_quad, src_far_quad)
    far = DirectOperator(data, K, precision) if fmm_config is None else FMMOperator(K, data, fmm_config, precision)
    galerkin = make_galerkin_operator(K.n_rows, obs_facets, obs_far_quad)
    interp = make_interpolation_operator(K.n_cols, src_facets, src_far_quad)
    return IntegralOperator(nearfield, farfield_correction, galerkin, far, interp)
