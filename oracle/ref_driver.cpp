// ORACLE TEST INFRASTRUCTURE — not product code.
//
// extern "C" driver around the UNMODIFIED reference FMM sources compiled from
// /root/reference/cpp by oracle/Makefile into oracle/_ref/libtbem_ref.so.
// It only calls the reference's public C++ API:
//   FMMOperator<3,R,C> ctor / apply / tasks / C2E lists   (cpp/fmm.h:131-237)
//   nbody_eval, make_direct_nbody_operator                 (cpp/nbody_operator.h:117-152)
//   make_octree                                            (cpp/octree.cpp:198-223)
//   TranslationSurface<3>::make_surrounding_surface        (cpp/fmm.cpp:22-51)
//   svd_decompose / set_threshold / svd_pseudoinverse      (cpp/blas_wrapper.cpp:67-127)
//   Kernel<3,R,C>::operator()                              (cpp/kernel.h:19-33)
// Only tests/, tests/golden/ and bench.py's cpu_baseline leg load it.

#include <chrono>
#include "fmm.h"
#include "nbody_operator.h"
#include "laplace_kernels.h"
#include "elastic_kernels.h"
#include "identity_kernels.h"
#include "octree.h"
#include "geometry.h"
#include "blas_wrapper.h"
#include "sparse_operator.h"
#include "constraint_matrix.h"
#include "mesh.h"
#include "quad_rule.h"
#include "galerkin_operator.h"
#include "interpolation_operator.h"

#include <memory>
#include <vector>
#include <cstring>

using namespace tbem;

namespace {

// Kernel ids shared with include/tbem_fmm.h (TBEM_K_*).
enum { K_LS = 0, K_LD = 1, K_LH = 2, K_EU = 3, K_ET = 4, K_EA = 5, K_EH = 6, K_ID1 = 7, K_ID3 = 8 };

bool is_scalar(int k) { return k == K_LS || k == K_LD || k == K_LH || k == K_ID1; }

std::unique_ptr<Kernel<3,1,1>> make_k1(int k) {
    switch (k) {
        case K_LS: return std::unique_ptr<Kernel<3,1,1>>(new LaplaceSingle<3>());
        case K_LD: return std::unique_ptr<Kernel<3,1,1>>(new LaplaceDouble<3>());
        case K_LH: return std::unique_ptr<Kernel<3,1,1>>(new LaplaceHypersingular<3>());
        default:   return std::unique_ptr<Kernel<3,1,1>>(new IdentityScalar<3>());
    }
}

std::unique_ptr<Kernel<3,3,3>> make_k3(int k, double mu, double nu) {
    switch (k) {
        case K_EU: return std::unique_ptr<Kernel<3,3,3>>(new ElasticDisplacement<3>(mu, nu));
        case K_ET: return std::unique_ptr<Kernel<3,3,3>>(new ElasticTraction<3>(mu, nu));
        case K_EA: return std::unique_ptr<Kernel<3,3,3>>(new ElasticAdjointTraction<3>(mu, nu));
        case K_EH: return std::unique_ptr<Kernel<3,3,3>>(new ElasticHypersingular<3>(mu, nu));
        default:   return std::unique_ptr<Kernel<3,3,3>>(new IdentityTensor<3,3,3>());
    }
}

std::vector<Vec3<double>> to_vec3(const double* p, size_t n) {
    std::vector<Vec3<double>> out(n);
    for (size_t i = 0; i < n; i++) {
        out[i] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    }
    return out;
}

NBodyData<3> make_data(size_t n_obs, const double* obs_locs, const double* obs_normals,
                       size_t n_src, const double* src_locs, const double* src_normals,
                       const double* src_weights) {
    NBodyData<3> d;
    d.obs_locs = to_vec3(obs_locs, n_obs);
    d.obs_normals = to_vec3(obs_normals, n_obs);
    d.src_locs = to_vec3(src_locs, n_src);
    d.src_normals = to_vec3(src_normals, n_src);
    d.src_weights.assign(src_weights, src_weights + n_src);
    return d;
}

struct Handle {
    virtual ~Handle() {}
    virtual void apply(const double* x, double* out) = 0;
    virtual void counts(long* c) = 0;
    virtual void c2e(int down, size_t level, double* out) = 0;
    virtual void pairs(int which, double* out) = 0;
};

template <size_t R>
struct HandleT: Handle {
    std::unique_ptr<Kernel<3,R,R>> K;
    FMMOperator<3,R,R> op;
    HandleT(std::unique_ptr<Kernel<3,R,R>> k, const NBodyData<3>& data, const FMMConfig& cfg):
        K(std::move(k)), op(*K, data, cfg) {}

    void apply(const double* x, double* out) override {
        std::vector<double> xv(x, x + op.n_cols());
        auto y = op.apply(xv);
        std::memcpy(out, y.data(), y.size() * sizeof(double));
    }
    void counts(long* c) override {
        c[0] = (long)op.up_equiv_surface.pts.size();
        c[1] = (long)(1 + op.src_oct.n_children());
        c[2] = (long)(1 + op.obs_oct.n_children());
        c[3] = (long)op.up_check_to_equiv.size();
        c[4] = (long)op.down_check_to_equiv.size();
        c[5] = (long)op.tasks.p2ms.size();
        c[6] = (long)op.tasks.m2ms.size();
        c[7] = (long)op.tasks.p2ps.size();
        c[8] = (long)op.tasks.m2ps.size();
        c[9] = (long)op.tasks.p2ls.size();
        c[10] = (long)op.tasks.m2ls.size();
        c[11] = (long)op.tasks.l2ls.size();
        c[12] = (long)op.tasks.l2ps.size();
    }
    void c2e(int down, size_t level, double* out) override {
        auto& v = down ? op.down_check_to_equiv[level] : op.up_check_to_equiv[level];
        std::memcpy(out, v.data(), v.size() * sizeof(double));
    }
    static void put_cell(const Octree<3>& c, double* o) {
        o[0] = (double)c.level;
        for (int d = 0; d < 3; d++) {
            o[1 + d] = c.bounds.center[d];
            o[4 + d] = c.bounds.half_width[d];
        }
    }
    // which: 0 p2p, 1 m2l, 2 m2p, 3 p2l; 14 doubles per pair (obs cell, src cell)
    void pairs(int which, double* out) override {
        const std::vector<typename FMMTasks<3>::CellPairTask>* v =
            which == 0 ? &op.tasks.p2ps : which == 1 ? &op.tasks.m2ls :
            which == 2 ? &op.tasks.m2ps : &op.tasks.p2ls;
        for (size_t i = 0; i < v->size(); i++) {
            put_cell((*v)[i].obs_cell, out + 14 * i);
            put_cell((*v)[i].src_cell, out + 14 * i + 7);
        }
    }
};

void dfs_cells(const Octree<3>& c, std::vector<double>& cells, std::vector<long>& idx) {
    // per cell: level, center xyz, hw xyz, is_leaf, n_indices, offset
    cells.push_back((double)c.level);
    for (int d = 0; d < 3; d++) cells.push_back(c.bounds.center[d]);
    for (int d = 0; d < 3; d++) cells.push_back(c.bounds.half_width[d]);
    cells.push_back(c.is_leaf() ? 1.0 : 0.0);
    cells.push_back((double)c.indices.size());
    cells.push_back((double)idx.size());
    for (auto i: c.indices) idx.push_back((long)i);
    for (auto& ch: c.children) {
        if (ch != nullptr) dfs_cells(*ch, cells, idx);
    }
}

} // namespace

extern "C" {

void* ref_fmm_create(int kernel, double mu, double nu,
                     size_t n_obs, const double* obs_locs, const double* obs_normals,
                     size_t n_src, const double* src_locs, const double* src_normals,
                     const double* src_weights,
                     double mac, size_t order, size_t leaf, double d, int small_cells) {
    auto data = make_data(n_obs, obs_locs, obs_normals, n_src, src_locs, src_normals, src_weights);
    FMMConfig cfg(mac, order, leaf, d, small_cells != 0);
    if (is_scalar(kernel)) {
        return (Handle*)new HandleT<1>(make_k1(kernel), data, cfg);
    }
    return (Handle*)new HandleT<3>(make_k3(kernel, mu, nu), data, cfg);
}

void ref_fmm_apply(void* h, const double* x, double* out) { ((Handle*)h)->apply(x, out); }
void ref_fmm_counts(void* h, long* c) { ((Handle*)h)->counts(c); }
void ref_fmm_c2e(void* h, int down, size_t level, double* out) { ((Handle*)h)->c2e(down, level, out); }
void ref_fmm_pairs(void* h, int which, double* out) { ((Handle*)h)->pairs(which, out); }
void ref_fmm_destroy(void* h) { delete (Handle*)h; }

// nbody_eval (cpp/nbody_operator.h:117-141), the reference's direct sum.
void ref_direct_eval(int kernel, double mu, double nu,
                     size_t n_obs, const double* obs_locs, const double* obs_normals,
                     size_t n_src, const double* src_locs, const double* src_normals,
                     const double* src_weights, const double* x, double* out) {
    auto data = make_data(n_obs, obs_locs, obs_normals, n_src, src_locs, src_normals, src_weights);
    std::vector<double> y;
    if (is_scalar(kernel)) {
        y = nbody_eval(*make_k1(kernel), data, x);
    } else {
        y = nbody_eval(*make_k3(kernel, mu, nu), data, x);
    }
    std::memcpy(out, y.data(), y.size() * sizeof(double));
}

// make_direct_nbody_operator(...).apply (dense matrix + dgemv), small sizes only.
void ref_dense_apply(int kernel, double mu, double nu,
                     size_t n_obs, const double* obs_locs, const double* obs_normals,
                     size_t n_src, const double* src_locs, const double* src_normals,
                     const double* src_weights, const double* x, double* out) {
    auto data = make_data(n_obs, obs_locs, obs_normals, n_src, src_locs, src_normals, src_weights);
    std::vector<double> y;
    if (is_scalar(kernel)) {
        auto op = make_direct_nbody_operator(data, *make_k1(kernel));
        y = op.apply(std::vector<double>(x, x + op.n_cols()));
    } else {
        auto op = make_direct_nbody_operator(data, *make_k3(kernel, mu, nu));
        y = op.apply(std::vector<double>(x, x + op.n_cols()));
    }
    std::memcpy(out, y.data(), y.size() * sizeof(double));
}

// One kernel evaluation K(obs, src, nobs, nsrc) -> R*C row-major.
void ref_kernel_eval(int kernel, double mu, double nu, const double* obs, const double* src,
                     const double* nobs, const double* nsrc, double* out) {
    Vec3<double> o{obs[0], obs[1], obs[2]}, s{src[0], src[1], src[2]};
    Vec3<double> no{nobs[0], nobs[1], nobs[2]}, ns{nsrc[0], nsrc[1], nsrc[2]};
    if (is_scalar(kernel)) {
        out[0] = (*make_k1(kernel))(o, s, no, ns)[0][0];
    } else {
        auto v = (*make_k3(kernel, mu, nu))(o, s, no, ns);
        for (int k = 0; k < 3; k++)
            for (int j = 0; j < 3; j++) out[3 * k + j] = v[k][j];
    }
}

// Octree cells in DFS pre-order (children 0..7), 10 doubles per cell, plus the
// concatenated per-cell index lists. Call with null outputs to get sizes.
void ref_octree(const double* pts, size_t n, size_t leaf, long* n_cells, long* n_idx,
                double* cells_out, long* idx_out) {
    auto oct = make_octree(to_vec3(pts, n), leaf);
    std::vector<double> cells;
    std::vector<long> idx;
    dfs_cells(oct, cells, idx);
    *n_cells = (long)(cells.size() / 10);
    *n_idx = (long)idx.size();
    if (cells_out) std::memcpy(cells_out, cells.data(), cells.size() * sizeof(double));
    if (idx_out) std::memcpy(idx_out, idx.data(), idx.size() * sizeof(long));
}

long ref_surface(size_t order, double* out) {
    auto s = TranslationSurface<3>::make_surrounding_surface(order);
    if (out) {
        for (size_t i = 0; i < s.pts.size(); i++)
            for (int d = 0; d < 3; d++) out[3 * i + d] = s.pts[i][d];
    }
    return (long)s.pts.size();
}

// Thresholded SVD pseudo-inverse of a row-major n x n matrix.
// SparseOperator::csr_from_coo (cpp/sparse_operator.cpp:140-172) of n_entries COO
// triplets; writes the CSR arrays (row_ptrs n_rows + 1, column_indices and values
// n_entries) and, with x, SparseOperator::apply (cpp/sparse_operator.cpp:19-28).
void ref_sparse_coo(size_t n_rows, size_t n_cols, size_t n_entries, const long* rows, const long* cols,
                    const double* vals, long* row_ptrs, long* column_indices, double* values, const double* x,
                    double* y) {
    std::vector<MatrixEntry> e(n_entries);
    for (size_t i = 0; i < n_entries; i++) e[i] = MatrixEntry{{(size_t)rows[i], (size_t)cols[i]}, vals[i]};
    auto op = SparseOperator::csr_from_coo(n_rows, n_cols, e);
    for (size_t i = 0; i <= n_rows; i++) row_ptrs[i] = (long)op.row_ptrs[i];
    for (size_t i = 0; i < n_entries; i++) {
        column_indices[i] = (long)op.column_indices[i];
        values[i] = op.values[i];
    }
    if (x) {
        auto out = op.apply(std::vector<double>(x, x + n_cols));
        std::memcpy(y, out.data(), n_rows * sizeof(double));
    }
}

// SparseOperator::right_multiply (cpp/sparse_operator.cpp:114-138) of A (a_rows x
// a_cols) and B (a_cols x b_cols) in CSR; writes at most cap entries; returns nnz.
long ref_sparse_right_multiply(size_t a_rows, size_t a_cols, const long* a_rp, const long* a_ci, const double* a_v,
                               size_t b_cols, const long* b_rp, const long* b_ci, const double* b_v,
                               long cap, long* rp, long* ci, double* v) {
    auto mk = [](size_t nr, size_t nc, const long* rp_, const long* ci_, const double* v_) {
        const size_t nnz = (size_t)rp_[nr];
        return SparseOperator(nr, nc, std::vector<double>(v_, v_ + nnz),
                              std::vector<size_t>(ci_, ci_ + nnz), std::vector<size_t>(rp_, rp_ + nr + 1));
    };
    auto out = mk(a_rows, a_cols, a_rp, a_ci, a_v).right_multiply(mk(a_cols, b_cols, b_rp, b_ci, b_v));
    const long nnz = (long)out.values.size();
    if (nnz > cap) return nnz;
    for (size_t i = 0; i <= a_rows; i++) rp[i] = (long)out.row_ptrs[i];
    for (long k = 0; k < nnz; k++) {
        ci[k] = (long)out.column_indices[k];
        v[k] = out.values[k];
    }
    return nnz;
}

// SparseOperator::apply (cpp/sparse_operator.cpp:19-28) on CSR arrays, reps times;
// returns the mean seconds per apply (construction excluded) and leaves y.
double ref_sparse_csr_apply(size_t n_rows, size_t n_cols, const long* rp, const long* ci, const double* v,
                            const double* x, double* y, int reps) {
    const size_t nnz = (size_t)rp[n_rows];
    SparseOperator A(n_rows, n_cols, std::vector<double>(v, v + nnz), std::vector<size_t>(ci, ci + nnz),
                     std::vector<size_t>(rp, rp + n_rows + 1));
    std::vector<double> xv(x, x + n_cols), out;
    auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < reps; r++) out = A.apply(xv);
    auto t1 = std::chrono::steady_clock::now();
    std::memcpy(y, out.data(), n_rows * sizeof(double));
    return std::chrono::duration<double>(t1 - t0).count() / reps;
}

// from_constraints (cpp/constraint_matrix.cpp:35-52) of n_eq equations (terms
// eq_ptr[e] .. eq_ptr[e+1]-1). Writes the solved constraints, constrained dofs
// ascending (when cdof is not NULL); returns their number and sets *n_terms.
long ref_from_constraints(size_t n_eq, const long* eq_ptr, const long* dof, const double* w, const double* rhs,
                          long* cdof, long* tptr, long* tdof, double* tw, double* trhs, long* n_terms) {
    std::vector<ConstraintEQ> eqs;
    for (size_t e = 0; e < n_eq; e++) {
        std::vector<LinearTerm> terms;
        for (long k = eq_ptr[e]; k < eq_ptr[e + 1]; k++) terms.push_back(LinearTerm{(size_t)dof[k], w[k]});
        eqs.push_back(ConstraintEQ{terms, rhs[e]});
    }
    auto cm = from_constraints(eqs);
    std::vector<size_t> keys;
    for (auto& kv : cm.map) keys.push_back(kv.first);
    std::sort(keys.begin(), keys.end());
    long k = 0;
    if (tptr) tptr[0] = 0;
    for (size_t i = 0; i < keys.size(); i++) {
        const auto& c = cm.map.at(keys[i]);
        if (cdof) {
            cdof[i] = (long)keys[i];
            trhs[i] = c.rhs;
        }
        for (const auto& t : c.terms) {
            if (cdof) {
                tdof[k] = (long)t.dof;
                tw[k] = t.weight;
            }
            k++;
        }
        if (cdof) tptr[i + 1] = k;
    }
    *n_terms = k;
    return (long)keys.size();
}

static ConstraintMatrix ref_cm(long n_c, const long* cdof, const long* tptr, const long* tdof, const double* tw,
                               const double* trhs) {
    ConstraintMatrixData m;
    for (long i = 0; i < n_c; i++) {
        std::vector<LinearTerm> terms;
        for (long k = tptr[i]; k < tptr[i + 1]; k++) terms.push_back(LinearTerm{(size_t)tdof[k], tw[k]});
        m.insert(std::make_pair((size_t)cdof[i], RearrangedConstraintEQ{(size_t)cdof[i], terms, trhs[i]}));
    }
    return ConstraintMatrix{m};
}

// distribute_vector / condense_vector (cpp/constraint_matrix.cpp:54-125) for solved
// constraints given as arrays; in has n_total - n_c entries (distribute) or n_total
// (condense).
void ref_distribute(long n_c, const long* cdof, const long* tptr, const long* tdof, const double* tw,
                    const double* trhs, const double* in, size_t n_total, double* out) {
    auto cm = ref_cm(n_c, cdof, tptr, tdof, tw, trhs);
    auto v = distribute_vector(cm, std::vector<double>(in, in + (n_total - n_c)), n_total);
    std::memcpy(out, v.data(), n_total * sizeof(double));
}

void ref_condense(long n_c, const long* cdof, const long* tptr, const long* tdof, const double* tw,
                  const double* trhs, const double* all, size_t n_total, double* out) {
    auto cm = ref_cm(n_c, cdof, tptr, tdof, tw, trhs);
    auto v = condense_vector(cm, std::vector<double>(all, all + n_total));
    std::memcpy(out, v.data(), v.size() * sizeof(double));
}

// condense_matrix for a CSR matrix (cpp/constraint_matrix.cpp:252-292), row and
// column constraints given as solved arrays; writes at most cap entries, returns nnz.
long ref_condense_sparse(long rn, const long* rcd, const long* rtp, const long* rtd, const double* rtw,
                         const double* rrhs, long cn, const long* ccd, const long* ctp, const long* ctd,
                         const double* ctw, const double* crhs, size_t n_rows, size_t n_cols, const long* rp,
                         const long* ci, const double* v, long cap, long* orp, long* oci, double* ov) {
    auto rcm = ref_cm(rn, rcd, rtp, rtd, rtw, rrhs);
    auto ccm = ref_cm(cn, ccd, ctp, ctd, ctw, crhs);
    const size_t nnz = (size_t)rp[n_rows];
    SparseOperator A(n_rows, n_cols, std::vector<double>(v, v + nnz), std::vector<size_t>(ci, ci + nnz),
                     std::vector<size_t>(rp, rp + n_rows + 1));
    auto out = condense_matrix(rcm, ccm, A);
    const long onnz = (long)out.values.size();
    if (onnz > cap) return onnz;
    for (size_t i = 0; i <= out.n_rows(); i++) orp[i] = (long)out.row_ptrs[i];
    for (long k = 0; k < onnz; k++) {
        oci[k] = (long)out.column_indices[k];
        ov[k] = out.values[k];
    }
    return onnz;
}

static Mesh<3> ref_mesh(size_t n_facets, const double* f) {
    std::vector<Facet<3>> facets(n_facets);
    for (size_t i = 0; i < n_facets; i++)
        for (int v = 0; v < 3; v++)
            for (int d = 0; d < 3; d++) facets[i][v][d] = f[9 * i + 3 * v + d];
    return Mesh<3>(facets);
}

static QuadRule<2> ref_quad(size_t nq, const double* qx, const double* qw) {
    QuadRule<2> q;
    for (size_t i = 0; i < nq; i++) q.push_back(QuadPt<2>{{{qx[2 * i], qx[2 * i + 1]}}, qw[i]});
    return q;
}

// make_galerkin_operator (which 0, cpp/galerkin_operator.cpp:11-45) or
// make_interpolation_operator (which 1, cpp/interpolation_operator.cpp:11-41) on a
// triangle mesh (facets: n_facets x 3 vertices x xyz) and a quadrature rule on the
// reference triangle; writes the CSR arrays (nnz = n_facets * nq * 3 * n_components).
void ref_bem_sparse(int which, size_t n_components, size_t n_facets, const double* facets, size_t nq,
                    const double* qx, const double* qw, long* row_ptrs, long* column_indices, double* values,
                    size_t* n_rows, size_t* n_cols) {
    auto mesh = ref_mesh(n_facets, facets);
    auto quad = ref_quad(nq, qx, qw);
    auto op = which == 0 ? make_galerkin_operator(n_components, mesh, quad)
                         : make_interpolation_operator(n_components, mesh, quad);
    *n_rows = op.n_rows();
    *n_cols = op.n_cols();
    for (size_t i = 0; i < op.row_ptrs.size(); i++) row_ptrs[i] = (long)op.row_ptrs[i];
    for (size_t i = 0; i < op.values.size(); i++) {
        column_indices[i] = (long)op.column_indices[i];
        values[i] = op.values[i];
    }
}

// nbody_data_from_bem (cpp/nbody_operator.h:38-85): quadrature points, facet
// normals and, for the sources, quadrature weight times facet jacobian.
void ref_nbody_from_bem(size_t n_obs_facets, const double* obs_facets, size_t n_src_facets,
                        const double* src_facets, size_t nq_obs, const double* qx_obs, const double* qw_obs,
                        size_t nq_src, const double* qx_src, const double* qw_src, double* obs_locs,
                        double* obs_normals, double* src_locs, double* src_normals, double* src_weights) {
    auto data = nbody_data_from_bem(ref_mesh(n_obs_facets, obs_facets), ref_mesh(n_src_facets, src_facets),
                                    ref_quad(nq_obs, qx_obs, qw_obs), ref_quad(nq_src, qx_src, qw_src));
    for (size_t i = 0; i < data.obs_locs.size(); i++)
        for (int d = 0; d < 3; d++) {
            obs_locs[3 * i + d] = data.obs_locs[i][d];
            obs_normals[3 * i + d] = data.obs_normals[i][d];
        }
    for (size_t i = 0; i < data.src_locs.size(); i++) {
        for (int d = 0; d < 3; d++) {
            src_locs[3 * i + d] = data.src_locs[i][d];
            src_normals[3 * i + d] = data.src_normals[i][d];
        }
        src_weights[i] = data.src_weights[i];
    }
}

void ref_pinv(const double* A, int n, double threshold, double* out) {
    std::vector<double> m(A, A + (size_t)n * n);
    auto svd = svd_decompose(m);
    if (threshold > 0) set_threshold(svd, threshold);
    auto p = svd_pseudoinverse(svd);
    std::memcpy(out, p.data(), p.size() * sizeof(double));
}

} // extern "C"
