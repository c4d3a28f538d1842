/*
 * tbem_fmm.h — C ABI of the MI355X (gfx950) kernel-independent FMM matvec.
 *
 * Drop-in boundary for 3bem's FMM and direct N-body operators. Plain C types
 * only: pointers, sizes, ints and doubles. The C++ shim in include/tbem/fmm.hpp
 * rebuilds the reference's operator surface on top of these entry points.
 *
 * Layouts follow the reference:
 *   - point arrays are AoS xyz, as NBodyData's std::array<double,3>
 *     (cpp/nbody_operator.h:29-36);
 *   - densities x and outputs y are component-major, x[d*N_src + j] and
 *     y[d*N_obs + i] (cpp/fmm.cpp:189, :289).
 *
 * Errors: every int-returning call gives 0 on success and a nonzero TBEM_E_*
 * code otherwise; tbem_last_error() then describes the failure (thread-local).
 * The reference only asserts (cpp/fmm.cpp:462-526, cpp/blas_wrapper.cpp:87).
 *
 * Threading: each handle owns one HIP stream and one set of work buffers. Host-
 * pointer calls run on that stream and return synchronised; _dev calls run on the
 * stream given (NULL = the default stream). Calls on one handle are ordered, even
 * across streams; separate handles (clones) are independent. The reference apply()
 * is const and re-entrant (cpp/fmm.cpp:550-558).
 */
#ifndef TBEM_FMM_H
#define TBEM_FMM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel ids. Each maps to one reference kernel class:
 *   TBEM_K_LAPLACE_SINGLE        LaplaceSingle<3>            cpp/laplace_kernels.h:18-32
 *   TBEM_K_LAPLACE_DOUBLE        LaplaceDouble<3>            cpp/laplace_kernels.h:34-48
 *   TBEM_K_LAPLACE_HYPERSINGULAR LaplaceHypersingular<3> (0) cpp/laplace_kernels.h:50-64
 *   TBEM_K_ELASTIC_DISPLACEMENT  ElasticDisplacement<3>      cpp/elastic_kernels.h:191-227
 *   TBEM_K_ELASTIC_TRACTION      ElasticTraction<3>          cpp/elastic_kernels.h:229-269
 *   TBEM_K_ELASTIC_ADJOINT_TRACTION ElasticAdjointTraction<3> cpp/elastic_kernels.h:271-311
 *   TBEM_K_ELASTIC_HYPERSINGULAR ElasticHypersingular<3>     cpp/elastic_kernels.h:313-365
 *   TBEM_K_IDENTITY_SCALAR       IdentityScalar<3>           cpp/identity_kernels.h:9-39
 *   TBEM_K_IDENTITY_TENSOR       IdentityTensor<3,3,3>       cpp/identity_kernels.h:9-39
 */
enum {
    TBEM_K_LAPLACE_SINGLE = 0,
    TBEM_K_LAPLACE_DOUBLE = 1,
    TBEM_K_LAPLACE_HYPERSINGULAR = 2,
    TBEM_K_ELASTIC_DISPLACEMENT = 3,
    TBEM_K_ELASTIC_TRACTION = 4,
    TBEM_K_ELASTIC_ADJOINT_TRACTION = 5,
    TBEM_K_ELASTIC_HYPERSINGULAR = 6,
    TBEM_K_IDENTITY_SCALAR = 7,
    TBEM_K_IDENTITY_TENSOR = 8
};

/* Arithmetic of the apply path.
 *   TBEM_FP64  : every stage in FP64 (the plain path, parity with the reference FMM).
 *   TBEM_MIXED : P2P and the point<->surface stages in FP32 with compensated
 *                accumulation; M2L as per-offset FP32 MFMA GEMMs; the check-to-
 *                equivalent pseudo-inverses (UC2E/DC2E) stay FP64. */
enum { TBEM_FP64 = 0, TBEM_MIXED = 1 };

enum {
    TBEM_OK = 0,
    TBEM_E_INVALID = 1,   /* bad argument (null pointer, zero order, wrong size, ...) */
    TBEM_E_HIP = 2,       /* a HIP runtime call failed */
    TBEM_E_NOMEM = 3,     /* device or host allocation failed */
    TBEM_E_UNSUPPORTED = 4,
    TBEM_E_NUMERIC = 5    /* e.g. the SVD did not converge */
};

typedef struct tbem_fmm_s* tbem_fmm_t;
typedef struct tbem_nbody_s* tbem_nbody_t;

/* Everything FMMOperator<3,R,C>(K, NBodyData, FMMConfig) receives
 * (cpp/fmm.h:146-147, cpp/fmm.h:73-96, cpp/nbody_operator.h:29-36). */
typedef struct {
    int kernel;              /* TBEM_K_* */
    double shear_modulus;    /* elastic kernels only (cpp/elastic_kernels.h:194-204) */
    double poisson_ratio;

    size_t n_obs;
    const double* obs_locs;      /* 3*n_obs, AoS */
    const double* obs_normals;   /* 3*n_obs, AoS */
    size_t n_src;
    const double* src_locs;      /* 3*n_src, AoS */
    const double* src_normals;   /* 3*n_src, AoS */
    const double* src_weights;   /* n_src */

    /* FMMConfig{mac, order, min_pts_per_cell, d, account_for_small_cells} */
    double mac;
    size_t order;
    size_t min_pts_per_cell;
    double d;
    int account_for_small_cells;

    int precision;               /* TBEM_FP64 or TBEM_MIXED */

    /* Optional: per-level check-to-equivalent matrices to use instead of the
     * ones computed on the GPU (row-major (R*n_e)^2 blocks, levels
     * concatenated). NULL = compute them. Used for bit-level parity runs. */
    const double* up_check_to_equiv;
    size_t n_up_levels;
    const double* down_check_to_equiv;
    size_t n_down_levels;

    /* Boxes of distinct observation and source trees (no reference counterpart;
     * ignored when one point set serves both sides): TBEM_TREES_AUTO (0, the
     * default) builds them as the reference does, each on its own points' bounding
     * box (cpp/fmm.cpp:88-89), except in TBEM_MIXED precision with an operator the
     * MFMA M2L serves (48 <= R*n_e <= 512), where both go on the union box so that
     * their cells lie on one grid and the M2L pairs fall into translation classes;
     * TBEM_TREES_OWN_BOXES and TBEM_TREES_COMMON_BOX force either layout. */
    int tree_boxes;

    /* The translations between surfaces (no reference counterpart):
     * TBEM_TRANSLATIONS_AUTO (0, the default) applies one FP64 operator per
     * translation class (level pair and integer centre offset) on the FP64 MFMA units
     * wherever the classes exist: M2M and L2L always (the cells of one tree lie on one
     * grid), the FP64 path's M2L when the cells of both sides do (one shared tree, or
     * tree_boxes TBEM_TREES_COMMON_BOX). Elsewhere, and everywhere with
     * TBEM_TRANSLATIONS_SURFACES, the kernel is evaluated between the two surfaces of
     * every pair, as cpp/fmm.cpp:204-251, 362-413 do. Both are FP64 throughout and
     * differ in rounding only. TBEM_MIXED's M2L always runs on its FP32 modal
     * operators. */
    int translations;

    /* Bytes the class M2L's per-pair output may take (no reference counterpart):
     * 0 = 6 GiB. One buffer per plan, shared by its clones; when the pairs need more,
     * they are processed in slices of whole target cells that fit, with the same
     * forward result bit for bit. */
    size_t m2l_output_budget;

    /* The point pairs of the near field (no reference counterpart):
     * TBEM_POINT_PAIRS_AUTO (0, the default) evaluates each pair of points in
     * neighbouring leaves once for both directions where the kernel is symmetric
     * (TBEM_K_ELASTIC_DISPLACEMENT and TBEM_K_LAPLACE_SINGLE, one point set on both
     * sides, leaves of 64 (Laplace: 16) to 256 points; forward and transpose, not the
     * observation gradient; FP32 pair arithmetic in TBEM_MIXED), where
     * cpp/fmm.cpp:254-294 evaluates it twice; TBEM_POINT_PAIRS_ORDERED evaluates every
     * ordered pair, as the reference does. They differ in rounding only. */
    int point_pairs;
} tbem_fmm_desc;

enum { TBEM_TREES_AUTO = 0, TBEM_TREES_OWN_BOXES = 1, TBEM_TREES_COMMON_BOX = 2 };
enum { TBEM_TRANSLATIONS_AUTO = 0, TBEM_TRANSLATIONS_SURFACES = 1 };
enum { TBEM_POINT_PAIRS_AUTO = 0, TBEM_POINT_PAIRS_ORDERED = 1 };

/* Plan statistics (tbem_fmm_info). */
typedef struct {
    size_t n_e;             /* points per translation surface (cpp/fmm.cpp:22-45) */
    size_t n_src_cells, n_obs_cells;
    size_t n_up_levels, n_down_levels;
    size_t p2m, m2m, p2p, m2p, p2l, m2l, l2l, l2p;   /* task counts as FMMTasks (cpp/fmm.h:98-120) */
    double p2p_pairs;        /* sum over P2P tasks of n_obs*n_src */
    double m2l_classes;      /* distinct translation classes (MIXED, or FP64 with class operators) */
    double plan_seconds;     /* wall time of tbem_fmm_create */
    double m2l_output_bytes; /* the class M2L's per-pair output (FP32 for MIXED, FP64 for FP64):
                                one buffer per plan, shared by its clones, whose M2L sections
                                take turns on it */
    double work_bytes;       /* device buffers of this handle's own (each clone has its own;
                                the adjoint and gradient add theirs on first use) */
    /* How the forward point-target stage (P2P + M2P + L2P) runs its observation leaves:
       leaf_wave_items one-wave work items (whole small leaves and chunks), of which the
       chunks of leaf_chunked leaves too large for one wave; leaf_block leaves on the
       one-workgroup kernel (more than 64 points with more pairs than one wave takes).
       leaf_symmetric leaves evaluate symmetric_pairs point pairs of their neighbours'
       once for both directions (TBEM_POINT_PAIRS_AUTO); the counts above then describe
       the rest of the pairs. */
    int leaf_wave_items, leaf_chunked, leaf_block, leaf_symmetric;
    double symmetric_pairs;
} tbem_fmm_info_t;

/* ---------------------------------------------------------------- FMM operator */

/* Replaces FMMOperator<3,R,C>::FMMOperator (cpp/fmm.cpp:71-101): copies the
 * inputs to the device, builds both octrees, the C2E matrices and the task lists. */
int tbem_fmm_create(const tbem_fmm_desc* desc, tbem_fmm_t* out);

/* Plan serialization (no reference counterpart: 3bem rebuilds its plan on every
 * construction). tbem_fmm_plan_save writes the handle's plan (both octrees, the task
 * lists, the per-level C2E matrices and, for TBEM_MIXED, their SVD factors) into buf
 * and sets *bytes to its size; with buf NULL it only sets *bytes. The format starts
 * with "TBEMPLN1" and a checksum and names the kernel, FMMConfig and a hash of every
 * input array. A plan cannot be saved after tbem_fmm_set_obs_locs (TBEM_E_UNSUPPORTED).
 * tbem_fmm_create_from_plan builds a handle from the same descriptor plus a saved
 * plan, skipping the octree builds, the traversal and the C2E SVDs; it refuses
 * (TBEM_E_INVALID) a plan saved for other inputs or configuration, a corrupt one, or
 * one whose indices do not fit its trees. Injected C2E matrices in desc are ignored. */
int tbem_fmm_plan_save(tbem_fmm_t h, void* buf, size_t cap, size_t* bytes);
int tbem_fmm_create_from_plan(const tbem_fmm_desc* desc, const void* plan, size_t plan_bytes,
                              tbem_fmm_t* out);

/* Replaces FMMOperator::apply (cpp/fmm.cpp:550-558). Host pointers:
 * x has n_cols() doubles, y receives n_rows() doubles. */
int tbem_fmm_apply(tbem_fmm_t h, const double* x, double* y);

/* Same on device pointers, enqueued on `stream` (a hipStream_t; NULL = the
 * default stream). No host synchronisation. Successive calls on one handle are
 * ordered even across streams (they share the handle's work buffers). */
int tbem_fmm_apply_dev(tbem_fmm_t h, const double* d_x, double* d_y, void* stream);

/* Exact transpose of the FMM operator (every stage transposed, roles of the
 * two trees swapped): g has n_rows() doubles, y receives n_cols(). No reference
 * counterpart: 3bem has no adjoint; it is the backward of apply. */
int tbem_fmm_apply_adjoint(tbem_fmm_t h, const double* g, double* y);
int tbem_fmm_apply_adjoint_dev(tbem_fmm_t h, const double* d_g, double* d_y, void* stream);

/* Gradient of 0.5*|y|^2-style losses w.r.t. observation coordinates: with
 * u = A(obs) x and upstream g = dL/du (n_rows()), writes
 * grad[3*i + m] = sum_k g[k*N_obs+i] d u[k*N_obs+i] / d obs_i[m] (AoS, 3*n_obs).
 * The plan geometry (trees, surfaces) is held fixed. No reference counterpart. */
int tbem_fmm_obs_grad(tbem_fmm_t h, const double* x, const double* g, double* grad);
int tbem_fmm_obs_grad_dev(tbem_fmm_t h, const double* d_x, const double* d_g, double* d_grad,
                          void* stream);

/* The backward's reuse of the forward's upward pass (no reference counterpart:
 * 3bem has no backward). tbem_fmm_forward_token returns a token for the far field
 * (multipole and local expansions) that the handle's work buffers hold after its
 * last forward (0: none). tbem_fmm_obs_grad_reuse_dev is tbem_fmm_obs_grad_dev
 * that, while the token is still the handle's current one (no forward or
 * recomputing gradient on this handle since; adjoints do not count), reuses that
 * far field instead of recomputing it; otherwise it recomputes, as
 * tbem_fmm_obs_grad_dev. d_x must hold the same values as that forward's input.
 * Costs no memory: the expansions are the handle's own work buffers. */
uint64_t tbem_fmm_forward_token(tbem_fmm_t h);
int tbem_fmm_obs_grad_reuse_dev(tbem_fmm_t h, uint64_t token, const double* d_x, const double* d_g,
                                double* d_grad, void* stream);

/* Move observation points without rebuilding the plan (same tree cells, same
 * surfaces, same lists); for the differentiable forward model and for
 * finite-difference checks of tbem_fmm_obs_grad. No reference counterpart (3bem
 * rebuilds FMMOperator for new points, cpp/fmm.cpp:71-101). obs_locs is AoS,
 * 3*n_obs, in the caller's order. Every moved point must stay inside the ball of
 * the leaf it was sorted into (centre, radius |half width|: the cell as the
 * dual-tree MAC test sees it, cpp/fmm.cpp:577-586): otherwise the call returns
 * TBEM_E_INVALID and changes nothing. The move applies to every handle that shares
 * the plan (clones): it waits for their calls in flight and blocks new ones until
 * it returns. _dev takes a device pointer and runs on the caller's stream. */
int tbem_fmm_set_obs_locs(tbem_fmm_t h, const double* obs_locs);
int tbem_fmm_set_obs_locs_dev(tbem_fmm_t h, const double* d_obs_locs, void* stream);

/* Counter of the plan's observation-point moves (0 for a new plan; shared by
 * clones): a caller that keeps a result can tell whether the points it was
 * computed at are still the plan's. No reference counterpart. */
uint64_t tbem_fmm_obs_version(tbem_fmm_t h);

/* OperatorI::n_rows / n_cols (cpp/fmm.h:235-236). */
size_t tbem_fmm_n_rows(tbem_fmm_t h);
size_t tbem_fmm_n_cols(tbem_fmm_t h);

/* OperatorI::clone (cpp/fmm.cpp:103-109). The reference rebuilds the plan;
 * here the immutable device plan is shared by reference count. */
int tbem_fmm_clone(tbem_fmm_t h, tbem_fmm_t* out);

void tbem_fmm_destroy(tbem_fmm_t h);

int tbem_fmm_info(tbem_fmm_t h, tbem_fmm_info_t* out);

/* Copy of the level-`level` UC2E (which=0) or DC2E (which=1) matrix,
 * (R*n_e)^2 doubles, row-major — FMMOperator::up/down_check_to_equiv. */
int tbem_fmm_c2e(tbem_fmm_t h, int which, size_t level, double* out);

/* Interaction lists for parity checks: which = 0 p2p, 1 m2l, 2 m2p, 3 p2l.
 * 14 doubles per pair: obs cell (level, center xyz, half width xyz), then the
 * src cell likewise. Pass out=NULL to get the count in *n. */
int tbem_fmm_pairs(tbem_fmm_t h, int which, double* out, size_t* n);

/* Octree of the source (which=0) or observation (which=1) points, cells in
 * depth-first pre-order (children 0..7): 10 doubles per cell (level, center
 * xyz, half width xyz, is_leaf, n_indices, offset) and concatenated index
 * lists, as Octree<3> (cpp/octree.h:14-29). NULL outputs return sizes only. */
int tbem_fmm_octree(tbem_fmm_t h, int which, double* cells, long* indices,
                    size_t* n_cells, size_t* n_indices);

/* Per-stage device times of the last apply, milliseconds, in the order
 * gather, p2m, m2m, m2l, l2l, p2p(+m2p+l2p), scatter (7 doubles). */
int tbem_fmm_stage_times(tbem_fmm_t h, double* ms7);

/* Device time of the M2L translation kernel alone in the last apply (ms): the
 * grouped MFMA GEMM in mixed precision, the FP64 M2L stage otherwise. No
 * reference counterpart: a profiling aid for roofline reports. */
int tbem_fmm_m2l_ms(tbem_fmm_t h, double* ms);

/* on != 0: the point-to-point pairs (P2P, and the transpose's P2P^T) run on a second
 * stream of the handle beside the far field, and the surface sources (M2P, L2P) are
 * added after both; on == 0: one fused leaf launch after the far field, as the
 * reference orders the stages (cpp/fmm.cpp:550-558). Same results to rounding. The
 * default is on where the operator size R * n_e is at least 48 (a large far field)
 * and off below. No reference counterpart: it also lets a profiler time each kernel
 * on an idle device. */
int tbem_fmm_set_overlap(tbem_fmm_t h, int on);

/* ---------------------------------------------------------- direct N-body operator */

/* Matrix-free replacement for make_direct_nbody_operator(data, K)
 * (cpp/nbody_operator.h:143-152) + DenseOperator::apply (cpp/dense_operator.cpp:38-46)
 * and for nbody_eval (cpp/nbody_operator.h:117-141): y_i = sum_j w_j K(x_i, y_j) x_j
 * evaluated on the GPU without forming the dense matrix. */
typedef struct {
    int kernel;
    double shear_modulus, poisson_ratio;
    size_t n_obs;
    const double* obs_locs;
    const double* obs_normals;
    size_t n_src;
    const double* src_locs;
    const double* src_normals;
    const double* src_weights;
    int precision;           /* TBEM_FP64 or TBEM_MIXED */
} tbem_nbody_desc;

int tbem_nbody_create(const tbem_nbody_desc* desc, tbem_nbody_t* out);
int tbem_nbody_apply(tbem_nbody_t h, const double* x, double* y);
int tbem_nbody_apply_dev(tbem_nbody_t h, const double* d_x, double* d_y, void* stream);
/* tbem_nbody_apply_dev that allocates nothing (no hipMalloc, no hipFree, no
 * allocating library call). No counterpart in the reference: it is for callers that
 * cannot allocate device memory, such as code under HIP graph capture, tight
 * serving loops and harnesses that forbid allocation. The weighted source
 * densities are staged in a static device array of 4,194,304 doubles (32 MiB) that
 * each GPU the library runs on holds until the process ends: R * n_src above that
 * returns TBEM_E_UNSUPPORTED. The current device must be the handle's. Every call
 * on one GPU shares that array, so calls that overlap on one GPU (on different
 * streams, or through different handles) race; calls ordered on one stream do
 * not. Same outputs as tbem_nbody_apply_dev, bit for bit. */
int tbem_nbody_apply_noalloc(tbem_nbody_t h, const double* d_x, double* d_y, void* stream);
/* y_j = w_j sum_i K(x_i, y_j)^T g_i — transpose of the direct operator. */
int tbem_nbody_apply_adjoint(tbem_nbody_t h, const double* g, double* y);
size_t tbem_nbody_n_rows(tbem_nbody_t h);
size_t tbem_nbody_n_cols(tbem_nbody_t h);
void tbem_nbody_destroy(tbem_nbody_t h);

/* ------------------------------------------------------- CSR sparse operator */

/* SparseOperator (cpp/sparse_operator.h:21-60) on the GPU. The product reproduces
 * SparseOperator::apply (cpp/sparse_operator.cpp:19-28) bit for bit: y_i sums
 * values[k] * x[column_indices[k]] over row i's entries in stored order, from 0.0,
 * each product rounded on its own. Limits: n_rows, n_cols < 2^31. */
typedef struct tbem_sparse_s* tbem_sparse_t;

/* SparseOperator(n_rows, n_cols, values, column_indices, row_ptrs)
 * (cpp/sparse_operator.cpp:7-17). row_ptrs has n_rows + 1 entries, starts at 0,
 * does not decrease and ends at nnz; every column index is below n_cols
 * (TBEM_E_INVALID otherwise). The arrays are copied. */
int tbem_sparse_create(size_t n_rows, size_t n_cols, size_t nnz, const double* values,
                       const size_t* column_indices, const size_t* row_ptrs, tbem_sparse_t* out);
/* SparseOperator::csr_from_coo (cpp/sparse_operator.cpp:140-172): the triplets
 * (rows[i], cols[i], values[i]) in CSR, each row's entries in input order,
 * duplicates kept. */
int tbem_sparse_from_coo(size_t n_rows, size_t n_cols, size_t n_entries, const size_t* rows,
                         const size_t* cols, const double* values, tbem_sparse_t* out);
int tbem_sparse_apply(tbem_sparse_t h, const double* x, double* y);
int tbem_sparse_apply_dev(tbem_sparse_t h, const double* d_x, double* d_y, void* stream);
/* The CSR arrays (values nnz, column_indices nnz, row_ptrs n_rows + 1), any may be
 * NULL: the values / column_indices / row_ptrs members (python_wrapper/linalg.cpp:25-37). */
int tbem_sparse_csr(tbem_sparse_t h, double* values, size_t* column_indices, size_t* row_ptrs);
int tbem_sparse_clone(tbem_sparse_t h, tbem_sparse_t* out);
/* SparseOperator::right_multiply (cpp/sparse_operator.cpp:114-138): the sparse
 * product a b, each row's columns ascending, every touched entry kept (explicit
 * zeros included), the sums formed in the reference's order. Computed on the host
 * at set-up time, like the reference; the result lives on the GPU. */
int tbem_sparse_right_multiply(tbem_sparse_t a, tbem_sparse_t b, tbem_sparse_t* out);
size_t tbem_sparse_n_rows(tbem_sparse_t h);
size_t tbem_sparse_n_cols(tbem_sparse_t h);
size_t tbem_sparse_nnz(tbem_sparse_t h);   /* SparseOperator::nnz (cpp/sparse_operator.h:37) */
void tbem_sparse_destroy(tbem_sparse_t h);

/* ------------------------------------------------------- integral operator */

/* IntegralOperator<3,R,C> (cpp/integral_operator.h:26-64) with the far field on the
 * GPU: y = N x + (G F(I x) + C x), summed per row as n_i + (g_i + c_i)
 * (cpp/integral_operator.h:55-63). The far field F is an FMM operator (the FMM
 * that boundary_operator leaves commented out, cpp/integral_operator.h:82-87) or
 * the direct N-body operator (the dense far field it uses). The handles are
 * borrowed: they must outlive the integral operator, and no other host thread may
 * apply them while it runs. */
typedef struct tbem_integral_s* tbem_integral_t;
enum { TBEM_FARFIELD_FMM = 0, TBEM_FARFIELD_DIRECT = 1 };

/* Argument order of the IntegralOperator constructor (cpp/integral_operator.h:34-44);
 * farfield is a tbem_fmm_t when farfield_kind is TBEM_FARFIELD_FMM, a tbem_nbody_t
 * when it is TBEM_FARFIELD_DIRECT. Shapes: nearfield and farfield_correction n_rows x n_cols,
 * galerkin n_rows x F.n_rows, interp F.n_cols x n_cols. */
int tbem_integral_create(tbem_sparse_t nearfield, tbem_sparse_t farfield_correction,
                         tbem_sparse_t galerkin, int farfield_kind, void* farfield,
                         tbem_sparse_t interp, tbem_integral_t* out);
int tbem_integral_apply(tbem_integral_t h, const double* x, double* y);
int tbem_integral_apply_dev(tbem_integral_t h, const double* d_x, double* d_y, void* stream);
size_t tbem_integral_n_rows(tbem_integral_t h);  /* galerkin.n_rows() */
size_t tbem_integral_n_cols(tbem_integral_t h);  /* nearfield.n_cols() */
void tbem_integral_destroy(tbem_integral_t h);

/* ------------------------------------------------------- linear constraints */

/* ConstraintMatrix (cpp/constraint_matrix.h:17-25): linear constraints on the dofs,
 * each solved for its highest dof. Set up on the host; distribute and condense run
 * on the GPU and equal the reference's vectors bit for bit. */
typedef struct tbem_constraints_s* tbem_constraints_t;

/* from_constraints (cpp/constraint_matrix.cpp:35-52) of n_eq ConstraintEQs
 * (cpp/constraint.h:33-52): equation e has the terms eq_ptr[e] .. eq_ptr[e+1]-1
 * (term_dof[k], term_weight[k]) and right-hand side rhs[e]. */
int tbem_constraints_create(size_t n_eq, const size_t* eq_ptr, const size_t* term_dof,
                            const double* term_weight, const double* rhs, tbem_constraints_t* out);
/* ConstraintMatrix::size (cpp/constraint_matrix.h:20-22): the number of constrained dofs. */
size_t tbem_constraints_size(tbem_constraints_t h);
size_t tbem_constraints_n_terms(tbem_constraints_t h);
/* The solved constraints (RearrangedConstraintEQ, cpp/constraint.h:60-70), constrained
 * dofs ascending: x_{constrained_dof[i]} = rhs[i] + sum over term_ptr[i] .. term_ptr[i+1]-1
 * of term_weight[k] x_{term_dof[k]}. Sizes: tbem_constraints_size and _n_terms (+1 for
 * term_ptr). Any pointer may be NULL. */
int tbem_constraints_get(tbem_constraints_t h, size_t* constrained_dof, size_t* term_ptr,
                         size_t* term_dof, double* term_weight, double* rhs);
/* homogenize_constraints (cpp/constraint_matrix.cpp:294-303): every rhs set to 0. */
int tbem_constraints_homogenize(tbem_constraints_t h, tbem_constraints_t* out);
/* distribute_vector (cpp/constraint_matrix.cpp:54-82): the reduced vector (n_total -
 * size() entries) to the full one (n_total entries). */
int tbem_distribute_vector(tbem_constraints_t h, const double* in, size_t n_total, double* out);
int tbem_distribute_vector_dev(tbem_constraints_t h, const double* d_in, size_t n_total, double* d_out,
                               void* stream);
/* condense_vector (cpp/constraint_matrix.cpp:99-125): the full vector (n_all entries)
 * to the reduced one (n_all - size() entries). */
int tbem_condense_vector(tbem_constraints_t h, const double* all, size_t n_all, double* out);
int tbem_condense_vector_dev(tbem_constraints_t h, const double* d_all, size_t n_all, double* d_out,
                             void* stream);
/* condense_matrix for a sparse operator (cpp/constraint_matrix.cpp:252-292), on the host
 * like the reference; the result lives on the GPU. */
int tbem_condense_sparse(tbem_constraints_t row_cm, tbem_constraints_t col_cm, tbem_sparse_t matrix,
                         tbem_sparse_t* out);
void tbem_constraints_destroy(tbem_constraints_t h);

/* ----------------------------------------------------------------------- misc */

/* sqrt(a[i]) and a[i] / b[i] evaluated on the GPU exactly as the device code of
 * the reference-order kernel tensors evaluates them. No reference counterpart: a
 * parity aid that lets a test check that gfx950's FP64 square root and division
 * are correctly rounded (the reference's C2E entries depend on it bit for bit). */
int tbem_fp64_sqrt_div(size_t n, const double* a, const double* b, double* sqrt_a, double* a_div_b);

/* One kernel evaluation K(obs <- src) on the GPU, R*C doubles row-major
 * (Kernel<3,R,C>::operator(), cpp/kernel.h:19-33). For kernel parity tests. */
int tbem_kernel_eval(int kernel, double shear_modulus, double poisson_ratio, size_t n,
                     const double* obs, const double* src, const double* nobs,
                     const double* nsrc, double* out);

/* Translation-surface points for `order` (unit sphere), n_e*3 doubles
 * (TranslationSurface<3>::make_surrounding_surface, cpp/fmm.cpp:22-51).
 * Returns n_e; out may be NULL. */
long tbem_surface(size_t order, double* out);

/* Thresholded FP64 pseudo-inverse of a row-major n x n matrix on the GPU
 * (one-sided Jacobi SVD); singular values <= threshold*sigma_max are dropped.
 * Replaces svd_decompose + set_threshold + svd_pseudoinverse
 * (cpp/blas_wrapper.cpp:67-127). */
int tbem_pinv(const double* A, int n, double threshold, double* out);

const char* tbem_last_error(void);

/* 1 if a HIP device is usable, else 0 (no error raised). */
int tbem_device_available(void);

#ifdef __cplusplus
}
#endif

#endif /* TBEM_FMM_H */
