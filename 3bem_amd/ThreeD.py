"""Python host layer mirroring tbempy.ThreeD's FMM-facing surface on the MI355X library.

Reference surface (python_wrapper/*.cpp, cpp/fmm.h, cpp/nbody_operator.h):
  * FMMConfig(mac, order, min_pts_per_cell, d, account_for_small_cells)
    -- python_wrapper/fmm_config.cpp:7, cpp/fmm.h:73-96
  * kernels LaplaceSingle/Double/Hypersingular(), ElasticDisplacement/Traction/
    AdjointTraction/Hypersingular(shear_modulus, poisson_ratio)
    -- python_wrapper/kernels.cpp:11-23
  * operators with apply / n_rows / n_cols / clone -- python_wrapper/op_wrap.h:4-11
  * FMMOperator(K, NBodyData, FMMConfig) -- cpp/fmm.h:146-147
  * make_direct_nbody_operator(data, K), nbody_eval(K, data, x) -- cpp/nbody_operator.h:117-152

apply() takes a NumPy array (copied to and from the device) or a torch CUDA
tensor (used in place on torch's current stream). Densities and outputs are
component-major, x[d*N_src + j], as in the reference.
"""
import ctypes

import numpy as np

from . import _lib

# ------------------------------------------------------------------------ kernels


class Kernel:
    kernel_id = -1
    dim = 1  # R = C
    shear_modulus = 1.0
    poisson_ratio = 0.25

    @property
    def n_rows(self):
        return self.dim

    @property
    def n_cols(self):
        return self.dim


class LaplaceSingle(Kernel):
    kernel_id, dim = 0, 1


class LaplaceDouble(Kernel):
    kernel_id, dim = 1, 1


class LaplaceHypersingular(Kernel):
    """The reference's 3D Laplace hypersingular kernel is a stub returning 0
    (cpp/laplace_kernels.h:50-64); kept identical."""
    kernel_id, dim = 2, 1


class _Elastic(Kernel):
    dim = 3

    def __init__(self, shear_modulus, poisson_ratio):
        self.shear_modulus = float(shear_modulus)
        self.poisson_ratio = float(poisson_ratio)


class ElasticDisplacement(_Elastic):
    kernel_id = 3


class ElasticTraction(_Elastic):
    kernel_id = 4


class ElasticAdjointTraction(_Elastic):
    kernel_id = 5


class ElasticHypersingular(_Elastic):
    kernel_id = 6


class IdentityScalar(Kernel):
    """Test kernel (cpp/identity_kernels.h:9-39)."""
    kernel_id, dim = 7, 1


class IdentityTensor(Kernel):
    kernel_id, dim = 8, 3


KERNELS = {k.kernel_id: k for k in (LaplaceSingle, LaplaceDouble, LaplaceHypersingular,
                                   ElasticDisplacement, ElasticTraction, ElasticAdjointTraction,
                                   ElasticHypersingular, IdentityScalar, IdentityTensor)}


def make_kernel(kernel_id, shear_modulus=1.0, poisson_ratio=0.25):
    cls = KERNELS[kernel_id]
    if issubclass(cls, _Elastic):
        return cls(shear_modulus, poisson_ratio)
    return cls()


# ------------------------------------------------------------------ config / data


class FMMConfig:
    """FMMConfig(mac, order, min_pts_per_cell, d, account_for_small_cells) (cpp/fmm.h:73-96).

    `order` is the approximate number of points on each translation sphere."""

    def __init__(self, mac, order, min_pts_per_cell, d, account_for_small_cells):
        self.mac = float(mac)
        self.order = int(order)
        self.min_pts_per_cell = int(min_pts_per_cell)
        self.d = float(d)
        self.account_for_small_cells = bool(account_for_small_cells)

    def __repr__(self):
        return (f"FMMConfig({self.mac}, {self.order}, {self.min_pts_per_cell}, {self.d}, "
                f"{self.account_for_small_cells})")


def _aos(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, 3))
    if n is not None and a.shape[0] != n:
        raise ValueError(f"expected {n} points, got {a.shape[0]}")
    return a


class NBodyData:
    """NBodyData{obs_locs, obs_normals, src_locs, src_normals, src_weights}
    (cpp/nbody_operator.h:29-36); arrays of shape (N, 3) / (N,)."""

    def __init__(self, obs_locs, obs_normals, src_locs, src_normals, src_weights):
        self.obs_locs = _aos(obs_locs)
        self.obs_normals = _aos(obs_normals, self.obs_locs.shape[0])
        self.src_locs = _aos(src_locs)
        self.src_normals = _aos(src_normals, self.src_locs.shape[0])
        self.src_weights = np.ascontiguousarray(np.asarray(src_weights, dtype=np.float64).reshape(-1))
        if self.src_weights.shape[0] != self.src_locs.shape[0]:
            raise ValueError("src_weights must have one entry per source")

    @property
    def n_obs(self):
        return self.obs_locs.shape[0]

    @property
    def n_src(self):
        return self.src_locs.shape[0]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None and a.size else None


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _torch_stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


PRECISIONS = {"fp64": _lib.FP64, "mixed": _lib.MIXED, "fp32": _lib.MIXED}

# ----------------------------------------------------------------------- operators


class FMMOperator:
    """FMMOperator<3,R,C>(K, data, config) on the GPU (cpp/fmm.h:131-237).

    precision="fp64": plain FP64 path (parity with the reference FMM).
    precision="mixed": FP32 P2P/surface stages with compensated sums; C2E in FP64.
    up_c2e / down_c2e: optional per-level check-to-equivalent matrices, shape
    (levels, n, n), used instead of the GPU-computed ones (parity runs).
    plan: bytes from save_plan() of an operator built on the same inputs and config.
    tree_boxes: for distinct observation and source points, "auto" (each tree on its
    own points' box as the reference, except mixed precision at orders the MFMA M2L
    serves, where both share the union box), "own" or "common" to force either.
    translations: the translations between surfaces (M2M, L2L, and the FP64 path's
    M2L), "auto" (one FP64 operator per translation class on the FP64 MFMA units
    where the classes exist, else as "surfaces") or "surfaces" (the kernel between
    the two surfaces of every pair, as the reference).
    m2l_output_budget: bytes the class M2L's per-pair output may take (None: 6 GiB);
    beyond it the pairs run in slices.
    point_pairs: "auto" (the forward evaluates each pair of points in neighbouring
    leaves once for both directions where the kernel is symmetric: elastic
    displacement or Laplace single layer, one point set on both sides, leaves of
    64 (Laplace: 16) to 256 points; forward and
    transpose) or "ordered" (every ordered pair, as the reference).
    """

    TREE_BOXES = {"auto": 0, "own": 1, "common": 2}
    TRANSLATIONS = {"auto": 0, "surfaces": 1}
    POINT_PAIRS = {"auto": 0, "ordered": 1}

    def __init__(self, K, data, config, precision="fp64", up_c2e=None, down_c2e=None, plan=None,
                 tree_boxes="auto", translations="auto",
                 m2l_output_budget=None, point_pairs="auto"):
        lib = _lib.require_device()
        self.K = K
        self.data = data
        self.config = config
        self.precision = precision
        self._up = None if up_c2e is None else np.ascontiguousarray(up_c2e, dtype=np.float64)
        self._down = None if down_c2e is None else np.ascontiguousarray(down_c2e, dtype=np.float64)
        desc = _lib.FmmDesc()
        desc.kernel = K.kernel_id
        desc.shear_modulus = K.shear_modulus
        desc.poisson_ratio = K.poisson_ratio
        desc.n_obs = data.n_obs
        desc.obs_locs, desc.obs_normals = _dp(data.obs_locs), _dp(data.obs_normals)
        desc.n_src = data.n_src
        desc.src_locs, desc.src_normals = _dp(data.src_locs), _dp(data.src_normals)
        desc.src_weights = _dp(data.src_weights)
        desc.mac, desc.order = config.mac, config.order
        desc.min_pts_per_cell, desc.d = config.min_pts_per_cell, config.d
        desc.account_for_small_cells = 1 if config.account_for_small_cells else 0
        desc.precision = PRECISIONS[precision]
        desc.tree_boxes = self.TREE_BOXES[tree_boxes]
        desc.translations = self.TRANSLATIONS[translations]
        desc.m2l_output_budget = int(m2l_output_budget or 0)
        desc.point_pairs = self.POINT_PAIRS[point_pairs]
        if self._up is not None:
            desc.up_check_to_equiv, desc.n_up_levels = _dp(self._up), self._up.shape[0]
        if self._down is not None:
            desc.down_check_to_equiv, desc.n_down_levels = _dp(self._down), self._down.shape[0]
        h = ctypes.c_void_p()
        if plan is None:
            _lib.check(lib.tbem_fmm_create(ctypes.byref(desc), ctypes.byref(h)))
        else:
            buf = plan if isinstance(plan, bytes) else bytes(plan)
            _lib.check(lib.tbem_fmm_create_from_plan(ctypes.byref(desc), buf, len(buf), ctypes.byref(h)))
        self._h = h
        self._lib = lib

    @classmethod
    def _wrap(cls, proto, handle):
        o = cls.__new__(cls)
        o.__dict__.update({k: v for k, v in proto.__dict__.items() if k != "_h"})
        o._h = handle
        return o

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.tbem_fmm_destroy(h)
            self._h = None

    def save_plan(self):
        """The plan as bytes (octrees, task lists, C2E): FMMOperator(K, data, config,
        plan=...) rebuilds the operator from them without the octree builds, the
        traversal and the C2E SVDs. Refused for other inputs or configuration."""
        n = ctypes.c_size_t()
        _lib.check(self._lib.tbem_fmm_plan_save(self._h, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _lib.check(self._lib.tbem_fmm_plan_save(self._h, buf, n.value, ctypes.byref(n)))
        return buf.raw[:n.value]

    def n_rows(self):
        return int(self._lib.tbem_fmm_n_rows(self._h))

    def n_cols(self):
        return int(self._lib.tbem_fmm_n_cols(self._h))

    def clone(self):
        """OperatorI::clone: shares the immutable device plan."""
        h = ctypes.c_void_p()
        _lib.check(self._lib.tbem_fmm_clone(self._h, ctypes.byref(h)))
        return FMMOperator._wrap(self, h)

    def _run(self, host_fn, dev_fn, x, n_in, n_out):
        if _is_torch(x):
            import torch
            assert x.is_cuda and x.dtype == torch.float64 and x.is_contiguous() and x.numel() == n_in
            y = torch.empty(n_out, dtype=torch.float64, device=x.device)
            _lib.check(dev_fn(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                              _torch_stream()))
            return y
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
        if x.shape[0] != n_in:
            raise ValueError(f"x has {x.shape[0]} entries, operator expects {n_in}")
        y = np.zeros(n_out)
        _lib.check(host_fn(self._h, _dp(x), _dp(y)))
        return y

    def apply(self, x):
        """y = A x (FMMOperator::apply, cpp/fmm.cpp:550-558)."""
        return self._run(self._lib.tbem_fmm_apply, self._lib.tbem_fmm_apply_dev, x,
                         self.n_cols(), self.n_rows())

    def apply_adjoint(self, g):
        """x = A^T g, the exact transpose of apply()."""
        return self._run(self._lib.tbem_fmm_apply_adjoint, self._lib.tbem_fmm_apply_adjoint_dev, g,
                         self.n_rows(), self.n_cols())

    def forward_token(self):
        """Token of the far field the operator holds after its last apply (0: none)."""
        return int(self._lib.tbem_fmm_forward_token(self._h))

    def obs_grad(self, x, g, token=None):
        """(N_obs, 3) array: d/d obs_i of sum_k g_k(i) u_k(i), u = apply(x); plan held fixed.
        token: forward_token() of the apply(x) this gradient belongs to; while no other
        apply has run on this operator since, its far field is reused instead of
        recomputed (device tensors only; x must be that apply's input)."""
        if _is_torch(x):
            import torch
            out = torch.empty(3 * self.data.n_obs, dtype=torch.float64, device=x.device)
            if token:
                _lib.check(self._lib.tbem_fmm_obs_grad_reuse_dev(
                    self._h, ctypes.c_uint64(token), ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(g.data_ptr()),
                    ctypes.c_void_p(out.data_ptr()), _torch_stream()))
            else:
                _lib.check(self._lib.tbem_fmm_obs_grad_dev(
                    self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(g.data_ptr()),
                    ctypes.c_void_p(out.data_ptr()), _torch_stream()))
            return out.view(-1, 3)
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        g = np.ascontiguousarray(g, dtype=np.float64).reshape(-1)
        out = np.zeros(3 * self.data.n_obs)
        _lib.check(self._lib.tbem_fmm_obs_grad(self._h, _dp(x), _dp(g), _dp(out)))
        return out.reshape(-1, 3)

    def set_obs_locs(self, obs_locs):
        """Move the observation points without rebuilding the plan (shared with clones).
        obs_locs: (N_obs, 3) NumPy array, or a float64 CUDA tensor (moved on the device,
        on torch's current stream). Every point must stay inside the ball (centre,
        radius |half width|) of the leaf it was sorted into; otherwise this raises and
        nothing moves."""
        if _is_torch(obs_locs):
            import torch
            t = obs_locs.detach()
            assert t.is_cuda and t.dtype == torch.float64 and t.numel() == 3 * self.data.n_obs
            t = t.contiguous()
            _lib.check(self._lib.tbem_fmm_set_obs_locs_dev(self._h, ctypes.c_void_p(t.data_ptr()),
                                                           _torch_stream()))
            return
        a = _aos(obs_locs, self.data.n_obs)
        _lib.check(self._lib.tbem_fmm_set_obs_locs(self._h, _dp(a)))

    def obs_version(self):
        """Number of observation-point moves of the plan so far (shared with clones)."""
        return int(self._lib.tbem_fmm_obs_version(self._h))

    def info(self):
        inf = _lib.FmmInfo()
        _lib.check(self._lib.tbem_fmm_info(self._h, ctypes.byref(inf)))
        return {name: getattr(inf, name) for name, _ in _lib.FmmInfo._fields_}

    def c2e(self, down, level):
        n = self.K.dim * self.info()["n_e"]
        out = np.zeros((n, n))
        _lib.check(self._lib.tbem_fmm_c2e(self._h, 1 if down else 0, level, _dp(out)))
        return out

    def pairs(self, which):
        w = {"p2p": 0, "m2l": 1, "m2p": 2, "p2l": 3}[which]
        n = ctypes.c_size_t()
        _lib.check(self._lib.tbem_fmm_pairs(self._h, w, None, ctypes.byref(n)))
        out = np.zeros((n.value, 14))
        if n.value:
            _lib.check(self._lib.tbem_fmm_pairs(self._h, w, _dp(out), ctypes.byref(n)))
        return out

    def octree(self, which="src"):
        w = 0 if which == "src" else 1
        nc, ni = ctypes.c_size_t(), ctypes.c_size_t()
        _lib.check(self._lib.tbem_fmm_octree(self._h, w, None, None, ctypes.byref(nc), ctypes.byref(ni)))
        cells = np.zeros((nc.value, 10))
        idx = np.zeros(ni.value, dtype=np.int64)
        _lib.check(self._lib.tbem_fmm_octree(self._h, w, _dp(cells),
                                             idx.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
                                             ctypes.byref(nc), ctypes.byref(ni)))
        return cells, idx

    def stage_times(self):
        """Device milliseconds of the last apply: gather, p2m, m2m, m2l, l2l, leaf, scatter."""
        out = np.zeros(7)
        _lib.check(self._lib.tbem_fmm_stage_times(self._h, _dp(out)))
        return dict(zip(("gather", "p2m", "m2m", "m2l", "l2l", "leaf", "scatter"), out.tolist()))

    def set_overlap(self, on=True):
        """Run the forward's point-to-point pairs on a second stream beside the far
        field (the default), or as one fused leaf launch after it (each kernel then
        has the device to itself, for timing). Same results to rounding."""
        _lib.check(self._lib.tbem_fmm_set_overlap(self._h, 1 if on else 0))

    def m2l_kernel_ms(self):
        """Device milliseconds of the M2L translation kernel alone in the last apply."""
        out = np.zeros(1)
        _lib.check(self._lib.tbem_fmm_m2l_ms(self._h, _dp(out)))
        return float(out[0])


class DirectOperator:
    """Matrix-free direct N-body operator on the GPU: replaces
    make_direct_nbody_operator(data, K) + DenseOperator::apply (cpp/nbody_operator.h:143-152)."""

    def __init__(self, data, K, precision="fp64"):
        lib = _lib.require_device()
        self.K, self.data, self.precision = K, data, precision
        desc = _lib.NbodyDesc()
        desc.kernel = K.kernel_id
        desc.shear_modulus, desc.poisson_ratio = K.shear_modulus, K.poisson_ratio
        desc.n_obs = data.n_obs
        desc.obs_locs, desc.obs_normals = _dp(data.obs_locs), _dp(data.obs_normals)
        desc.n_src = data.n_src
        desc.src_locs, desc.src_normals = _dp(data.src_locs), _dp(data.src_normals)
        desc.src_weights = _dp(data.src_weights)
        desc.precision = PRECISIONS[precision]
        h = ctypes.c_void_p()
        _lib.check(lib.tbem_nbody_create(ctypes.byref(desc), ctypes.byref(h)))
        self._h, self._lib = h, lib

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.tbem_nbody_destroy(h)
            self._h = None

    def n_rows(self):
        return int(self._lib.tbem_nbody_n_rows(self._h))

    def n_cols(self):
        return int(self._lib.tbem_nbody_n_cols(self._h))

    def clone(self):
        return DirectOperator(self.data, self.K, self.precision)

    def apply(self, x):
        if _is_torch(x):
            import torch
            y = torch.empty(self.n_rows(), dtype=torch.float64, device=x.device)
            _lib.check(self._lib.tbem_nbody_apply_dev(self._h, ctypes.c_void_p(x.data_ptr()),
                                                      ctypes.c_void_p(y.data_ptr()), _torch_stream()))
            return y
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
        if x.shape[0] != self.n_cols():
            raise ValueError(f"x has {x.shape[0]} entries, operator expects {self.n_cols()}")
        y = np.zeros(self.n_rows())
        _lib.check(self._lib.tbem_nbody_apply(self._h, _dp(x), _dp(y)))
        return y

    def apply_noalloc(self, x, out):
        """out = A x into a caller-owned float64 CUDA tensor, allocating nothing
        (tbem_nbody_apply_noalloc: R * n_src <= 4,194,304; calls that overlap on one
        GPU race on its static scratch)."""
        import torch
        assert x.is_cuda and x.dtype == torch.float64 and x.is_contiguous() and x.numel() == self.n_cols()
        assert out.is_cuda and out.dtype == torch.float64 and out.is_contiguous() and out.numel() == self.n_rows()
        _lib.check(self._lib.tbem_nbody_apply_noalloc(self._h, ctypes.c_void_p(x.data_ptr()),
                                                      ctypes.c_void_p(out.data_ptr()), _torch_stream()))
        return out

    def apply_adjoint(self, g):
        g = np.ascontiguousarray(np.asarray(g, dtype=np.float64).reshape(-1))
        y = np.zeros(self.n_cols())
        _lib.check(self._lib.tbem_nbody_apply_adjoint(self._h, _dp(g), _dp(y)))
        return y


def make_direct_nbody_operator(data, K, precision="fp64"):
    return DirectOperator(data, K, precision)


def _apply_dev_or_host(lib, h, host_fn, dev_fn, x, n_in, n_out):
    if _is_torch(x):
        import torch
        assert x.is_cuda and x.dtype == torch.float64 and x.is_contiguous() and x.numel() == n_in
        y = torch.empty(n_out, dtype=torch.float64, device=x.device)
        _lib.check(dev_fn(h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), _torch_stream()))
        return y
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
    if x.shape[0] != n_in:
        raise ValueError(f"x has {x.shape[0]} entries, operator expects {n_in}")
    y = np.zeros(n_out)
    _lib.check(host_fn(h, _dp(x), _dp(y)))
    return y


def _szp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_size_t)) if a.size else None


class SparseOperator:
    """SparseOperator (cpp/sparse_operator.h:21-60) held on the GPU in CSR form.

    SparseOperator(n_rows, n_cols, values, column_indices, row_ptrs) takes the CSR
    arrays (cpp/sparse_operator.cpp:7-17); SparseOperator.csr_from_coo(n_rows, n_cols,
    rows, cols, values) builds them from triplets (cpp/sparse_operator.cpp:140-172).
    apply() equals the reference's SparseOperator::apply bit for bit.
    """

    def __init__(self, n_rows, n_cols, values, column_indices, row_ptrs):
        lib = _lib.require_device()
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        c = np.ascontiguousarray(column_indices, dtype=np.uint64).reshape(-1)
        r = np.ascontiguousarray(row_ptrs, dtype=np.uint64).reshape(-1)
        if c.shape[0] != v.shape[0] or r.shape[0] != n_rows + 1:
            raise ValueError("need len(column_indices) == len(values) and len(row_ptrs) == n_rows + 1")
        h = ctypes.c_void_p()
        _lib.check(lib.tbem_sparse_create(n_rows, n_cols, v.shape[0], _dp(v), _szp(c), _szp(r), ctypes.byref(h)))
        self._h, self._lib = h, lib

    @classmethod
    def csr_from_coo(cls, n_rows, n_cols, rows, cols, values):
        lib = _lib.require_device()
        r = np.ascontiguousarray(rows, dtype=np.uint64).reshape(-1)
        c = np.ascontiguousarray(cols, dtype=np.uint64).reshape(-1)
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if not (r.shape == c.shape == v.shape):
            raise ValueError("rows, cols and values must have the same length")
        h = ctypes.c_void_p()
        _lib.check(lib.tbem_sparse_from_coo(n_rows, n_cols, v.shape[0], _szp(r), _szp(c), _dp(v), ctypes.byref(h)))
        o = cls.__new__(cls)
        o._h, o._lib = h, lib
        return o

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.tbem_sparse_destroy(h)
            self._h = None

    def n_rows(self):
        return int(self._lib.tbem_sparse_n_rows(self._h))

    def n_cols(self):
        return int(self._lib.tbem_sparse_n_cols(self._h))

    def nnz(self):
        return int(self._lib.tbem_sparse_nnz(self._h))

    def _csr(self):
        nnz = self.nnz()
        v, c = np.zeros(nnz), np.zeros(nnz, dtype=np.uint64)
        r = np.zeros(self.n_rows() + 1, dtype=np.uint64)
        _lib.check(self._lib.tbem_sparse_csr(self._h, _dp(v), _szp(c), _szp(r)))
        return v, c.astype(np.int64), r.astype(np.int64)

    @property
    def values(self):
        return self._csr()[0]

    @property
    def column_indices(self):
        return self._csr()[1]

    @property
    def row_ptrs(self):
        return self._csr()[2]

    def to_dense(self):
        """SparseOperator::to_dense (cpp/sparse_operator.cpp:31-40) as an (n_rows, n_cols) array."""
        v, c, r = self._csr()
        out = np.zeros((self.n_rows(), self.n_cols()))
        np.add.at(out, (np.repeat(np.arange(self.n_rows()), np.diff(r)), c), v)
        return out

    def clone(self):
        h = ctypes.c_void_p()
        _lib.check(self._lib.tbem_sparse_clone(self._h, ctypes.byref(h)))
        o = SparseOperator.__new__(SparseOperator)
        o._h, o._lib = h, self._lib
        return o

    def right_multiply(self, other):
        """The sparse product self @ other (SparseOperator::right_multiply,
        cpp/sparse_operator.cpp:114-138), formed on the host like the reference's."""
        h = ctypes.c_void_p()
        _lib.check(self._lib.tbem_sparse_right_multiply(self._h, other._h, ctypes.byref(h)))
        o = SparseOperator.__new__(SparseOperator)
        o._h, o._lib = h, self._lib
        return o

    def apply(self, x):
        """y = A x (SparseOperator::apply, cpp/sparse_operator.cpp:19-28)."""
        return _apply_dev_or_host(self._lib, self._h, self._lib.tbem_sparse_apply, self._lib.tbem_sparse_apply_dev,
                                  x, self.n_cols(), self.n_rows())


class IntegralOperator:
    """IntegralOperator (cpp/integral_operator.h:26-64): y = N x + (G F(I x) + C x)
    with the far field F an FMMOperator or a DirectOperator on the GPU. Arguments in
    the order of the reference's constructor: nearfield, farfield_correction,
    galerkin, farfield, interp."""

    def __init__(self, nearfield, farfield_correction, galerkin, farfield, interp):
        lib = _lib.require_device()
        if isinstance(farfield, FMMOperator):
            kind = _lib.FARFIELD_FMM
        elif isinstance(farfield, DirectOperator):
            kind = _lib.FARFIELD_DIRECT
        else:
            raise TypeError("farfield must be an FMMOperator or a DirectOperator")
        # the library borrows these handles: keep the operators alive with this one
        self.nearfield, self.farfield_correction, self.galerkin = nearfield, farfield_correction, galerkin
        self.farfield, self.interp = farfield, interp
        h = ctypes.c_void_p()
        _lib.check(lib.tbem_integral_create(nearfield._h, farfield_correction._h, galerkin._h, kind,
                                            farfield._h, interp._h, ctypes.byref(h)))
        self._h, self._lib = h, lib

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.tbem_integral_destroy(h)
            self._h = None

    def n_rows(self):
        return int(self._lib.tbem_integral_n_rows(self._h))

    def n_cols(self):
        return int(self._lib.tbem_integral_n_cols(self._h))

    def clone(self):
        return IntegralOperator(self.nearfield, self.farfield_correction, self.galerkin, self.farfield.clone(),
                                self.interp)

    def apply(self, x):
        return _apply_dev_or_host(self._lib, self._h, self._lib.tbem_integral_apply,
                                  self._lib.tbem_integral_apply_dev, x, self.n_cols(), self.n_rows())


def nbody_eval(K, data, x):
    """nbody_eval(K, data, x) (cpp/nbody_operator.h:117-141), on the GPU."""
    return DirectOperator(data, K).apply(x)


def kernel_eval(K, obs, src, nobs, nsrc):
    """K(obs_i <- src_i) for n pairs on the GPU: (n, R, C) array."""
    lib = _lib.require_device()
    arrs = [_aos(a) for a in (obs, src, nobs, nsrc)]
    n = arrs[0].shape[0]
    out = np.zeros(n * K.dim * K.dim)
    _lib.check(lib.tbem_kernel_eval(K.kernel_id, K.shear_modulus, K.poisson_ratio, n,
                                    *[_dp(a) for a in arrs], _dp(out)))
    return out.reshape(n, K.dim, K.dim)


def surface_points(order):
    """Unit translation-sphere points for `order` (cpp/fmm.cpp:22-45); host code, no GPU."""
    lib = _lib.load()
    n = lib.tbem_surface(order, None)
    out = np.zeros((n, 3))
    lib.tbem_surface(order, _dp(out))
    return out


def pinv(A, threshold=1e-10):
    """Thresholded FP64 pseudo-inverse on the GPU (one-sided Jacobi SVD)."""
    lib = _lib.require_device()
    A = np.ascontiguousarray(A, dtype=np.float64)
    out = np.zeros_like(A)
    _lib.check(lib.tbem_pinv(_dp(A), A.shape[0], threshold, _dp(out)))
    return out
