// tbem/fmm.hpp — header-only C++ surface of 3bem's FMM, direct N-body, sparse and integral operators
// on top of the C ABI in tbem_fmm.h (link -ltbem_fmm).
//
// Mirrors (reference file:line):
//   OperatorI {n_rows, n_cols, apply, clone}           cpp/operator.h:14-19
//   FMMConfig(mac, order, min_pts_per_cell, d, small)  cpp/fmm.h:73-96
//   NBodyData<3>{obs_locs, obs_normals, src_locs, src_normals, src_weights}
//                                                      cpp/nbody_operator.h:29-36
//   kernels LaplaceSingle<3> ... ElasticHypersingular<3> with shear_modulus /
//   poisson_ratio members                              cpp/laplace_kernels.h, cpp/elastic_kernels.h
//   FMMOperator<3,R,C>(K, data, config).apply(x)       cpp/fmm.h:131-237
//   make_direct_nbody_operator(data, K), nbody_eval    cpp/nbody_operator.h:117-152
//   MatrixEntry, SparseOperator {apply, nnz, right_multiply, csr_from_coo}
//                                                      cpp/sparse_operator.h:12-60
//   IntegralOperator<3,R,C>(near, correction, galerkin, farfield, interp)
//                                                      cpp/integral_operator.h:26-64
//   LinearTerm, ConstraintEQ, boundary_condition, continuity_constraint,
//   shift_constraints, from_constraints, homogenize_constraints,
//   distribute_vector, condense_vector                 cpp/constraint.h, cpp/constraint_matrix.h
// Errors (nonzero status) become std::runtime_error carrying tbem_last_error().
#ifndef TBEM_FMM_HPP
#define TBEM_FMM_HPP

#include <array>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../tbem_fmm.h"

namespace tbem {

template <typename T, size_t dim>
using Vec = std::array<T, dim>;

struct OperatorI {
    virtual ~OperatorI() {}
    virtual size_t n_rows() const = 0;
    virtual size_t n_cols() const = 0;
    virtual std::vector<double> apply(const std::vector<double>& x) const = 0;
    virtual std::unique_ptr<OperatorI> clone() const = 0;
};

struct FMMConfig {
    const double mac;
    const size_t order;
    const size_t min_pts_per_cell;
    const double d;
    const bool account_for_small_cells;
    FMMConfig(double mac, size_t order, size_t min_pts_per_cell, double d, bool account_for_small_cells)
        : mac(mac), order(order), min_pts_per_cell(min_pts_per_cell), d(d),
          account_for_small_cells(account_for_small_cells) {}
};

template <size_t dim>
struct NBodyData {
    std::vector<Vec<double, dim>> obs_locs;
    std::vector<Vec<double, dim>> obs_normals;
    std::vector<Vec<double, dim>> src_locs;
    std::vector<Vec<double, dim>> src_normals;
    std::vector<double> src_weights;
};

// Kernel tags: each names its C-ABI kernel id and block size R = C.
template <size_t dim, size_t R, size_t C>
struct Kernel {
    const int id;
    const double shear_modulus;
    const double poisson_ratio;
    Kernel(int id, double mu = 1.0, double nu = 0.25) : id(id), shear_modulus(mu), poisson_ratio(nu) {}
};

template <size_t dim> struct LaplaceSingle;
template <size_t dim> struct LaplaceDouble;
template <size_t dim> struct LaplaceHypersingular;
template <size_t dim> struct ElasticDisplacement;
template <size_t dim> struct ElasticTraction;
template <size_t dim> struct ElasticAdjointTraction;
template <size_t dim> struct ElasticHypersingular;

template <> struct LaplaceSingle<3> : Kernel<3, 1, 1> { LaplaceSingle() : Kernel(TBEM_K_LAPLACE_SINGLE) {} };
template <> struct LaplaceDouble<3> : Kernel<3, 1, 1> { LaplaceDouble() : Kernel(TBEM_K_LAPLACE_DOUBLE) {} };
template <> struct LaplaceHypersingular<3> : Kernel<3, 1, 1> {
    LaplaceHypersingular() : Kernel(TBEM_K_LAPLACE_HYPERSINGULAR) {}
};
template <> struct ElasticDisplacement<3> : Kernel<3, 3, 3> {
    ElasticDisplacement(double mu, double nu) : Kernel(TBEM_K_ELASTIC_DISPLACEMENT, mu, nu) {}
};
template <> struct ElasticTraction<3> : Kernel<3, 3, 3> {
    ElasticTraction(double mu, double nu) : Kernel(TBEM_K_ELASTIC_TRACTION, mu, nu) {}
};
template <> struct ElasticAdjointTraction<3> : Kernel<3, 3, 3> {
    ElasticAdjointTraction(double mu, double nu) : Kernel(TBEM_K_ELASTIC_ADJOINT_TRACTION, mu, nu) {}
};
template <> struct ElasticHypersingular<3> : Kernel<3, 3, 3> {
    ElasticHypersingular(double mu, double nu) : Kernel(TBEM_K_ELASTIC_HYPERSINGULAR, mu, nu) {}
};

inline void check(int code) {
    if (code != TBEM_OK) throw std::runtime_error(std::string("tbem: ") + tbem_last_error());
}

namespace detail {
inline const double* flat(const std::vector<Vec<double, 3>>& v) { return v.empty() ? nullptr : v[0].data(); }
}  // namespace detail

// FMMOperator<3,R,C>: plan built on the GPU at construction, apply() on the GPU.
template <size_t dim, size_t R, size_t C>
struct FMMOperator : public OperatorI {
    static_assert(dim == 3, "the MI355X FMM is three-dimensional");
    std::shared_ptr<tbem_fmm_s> h;

    FMMOperator(const Kernel<dim, R, C>& K, const NBodyData<dim>& data, const FMMConfig& config,
                int precision = TBEM_FP64) {
        tbem_fmm_desc d{};
        d.kernel = K.id;
        d.shear_modulus = K.shear_modulus;
        d.poisson_ratio = K.poisson_ratio;
        d.n_obs = data.obs_locs.size();
        d.obs_locs = detail::flat(data.obs_locs);
        d.obs_normals = detail::flat(data.obs_normals);
        d.n_src = data.src_locs.size();
        d.src_locs = detail::flat(data.src_locs);
        d.src_normals = detail::flat(data.src_normals);
        d.src_weights = data.src_weights.empty() ? nullptr : data.src_weights.data();
        d.mac = config.mac;
        d.order = config.order;
        d.min_pts_per_cell = config.min_pts_per_cell;
        d.d = config.d;
        d.account_for_small_cells = config.account_for_small_cells ? 1 : 0;
        d.precision = precision;
        tbem_fmm_t raw = nullptr;
        check(tbem_fmm_create(&d, &raw));
        h.reset(raw, tbem_fmm_destroy);
    }
    explicit FMMOperator(tbem_fmm_t raw) : h(raw, tbem_fmm_destroy) {}

    size_t n_rows() const override { return tbem_fmm_n_rows(h.get()); }
    size_t n_cols() const override { return tbem_fmm_n_cols(h.get()); }

    std::vector<double> apply(const std::vector<double>& x) const override {
        if (x.size() != n_cols()) throw std::invalid_argument("FMMOperator::apply: x has the wrong length");
        std::vector<double> y(n_rows());
        check(tbem_fmm_apply(h.get(), x.data(), y.data()));
        return y;
    }

    std::vector<double> apply_adjoint(const std::vector<double>& g) const {
        if (g.size() != n_rows()) throw std::invalid_argument("apply_adjoint: g has the wrong length");
        std::vector<double> y(n_cols());
        check(tbem_fmm_apply_adjoint(h.get(), g.data(), y.data()));
        return y;
    }

    std::unique_ptr<OperatorI> clone() const override {
        tbem_fmm_t raw = nullptr;
        check(tbem_fmm_clone(h.get(), &raw));
        return std::unique_ptr<OperatorI>(new FMMOperator(raw));
    }
};

// Matrix-free direct N-body operator (make_direct_nbody_operator + DenseOperator::apply).
struct DirectNBodyOperator : public OperatorI {
    std::shared_ptr<tbem_nbody_s> h;
    NBodyData<3> data;
    int kernel;
    double mu, nu;
    template <size_t R, size_t C>
    DirectNBodyOperator(const NBodyData<3>& data, const Kernel<3, R, C>& K)
        : data(data), kernel(K.id), mu(K.shear_modulus), nu(K.poisson_ratio) {
        build();
    }
    void build() {
        tbem_nbody_desc d{};
        d.kernel = kernel;
        d.shear_modulus = mu;
        d.poisson_ratio = nu;
        d.n_obs = data.obs_locs.size();
        d.obs_locs = detail::flat(data.obs_locs);
        d.obs_normals = detail::flat(data.obs_normals);
        d.n_src = data.src_locs.size();
        d.src_locs = detail::flat(data.src_locs);
        d.src_normals = detail::flat(data.src_normals);
        d.src_weights = data.src_weights.empty() ? nullptr : data.src_weights.data();
        d.precision = TBEM_FP64;
        tbem_nbody_t raw = nullptr;
        check(tbem_nbody_create(&d, &raw));
        h.reset(raw, tbem_nbody_destroy);
    }
    size_t n_rows() const override { return tbem_nbody_n_rows(h.get()); }
    size_t n_cols() const override { return tbem_nbody_n_cols(h.get()); }
    std::vector<double> apply(const std::vector<double>& x) const override {
        if (x.size() != n_cols()) throw std::invalid_argument("apply: x has the wrong length");
        std::vector<double> y(n_rows());
        check(tbem_nbody_apply(h.get(), x.data(), y.data()));
        return y;
    }
    std::unique_ptr<OperatorI> clone() const override {
        return std::unique_ptr<OperatorI>(new DirectNBodyOperator(*this));
    }
};

template <size_t R, size_t C>
DirectNBodyOperator make_direct_nbody_operator(const NBodyData<3>& data, const Kernel<3, R, C>& K) {
    return DirectNBodyOperator(data, K);
}

template <size_t R, size_t C>
std::vector<double> nbody_eval(const Kernel<3, R, C>& K, const NBodyData<3>& data, const double* x) {
    auto op = make_direct_nbody_operator(data, K);
    return op.apply(std::vector<double>(x, x + op.n_cols()));
}

// MatrixEntry (cpp/sparse_operator.h:12-20).
struct MatrixEntry {
    size_t loc[2];
    double value;
    size_t row() const { return loc[0]; }
    size_t col() const { return loc[1]; }
};

struct OperatorShape {
    size_t n_rows;
    size_t n_cols;
};

// SparseOperator (cpp/sparse_operator.h:21-60): the CSR arrays as the reference's
// members, the matrix on the GPU; apply() equals SparseOperator::apply bit for bit.
struct SparseOperator : public OperatorI {
    std::shared_ptr<tbem_sparse_s> h;
    const OperatorShape shape;
    const std::vector<double> values;
    const std::vector<size_t> column_indices;
    const std::vector<size_t> row_ptrs;

    SparseOperator(size_t n_rows, size_t n_cols, const std::vector<double>& values,
                   const std::vector<size_t>& column_indices, const std::vector<size_t>& row_ptrs)
        : shape{n_rows, n_cols}, values(values), column_indices(column_indices), row_ptrs(row_ptrs) {
        if (row_ptrs.size() != n_rows + 1 || column_indices.size() != values.size())
            throw std::invalid_argument("SparseOperator: need n_rows + 1 row_ptrs and one column per value");
        tbem_sparse_t raw = nullptr;
        check(tbem_sparse_create(n_rows, n_cols, values.size(), values.data(), column_indices.data(),
                                 row_ptrs.data(), &raw));
        h.reset(raw, tbem_sparse_destroy);
    }

    size_t n_rows() const override { return shape.n_rows; }
    size_t n_cols() const override { return shape.n_cols; }
    size_t nnz() const { return row_ptrs.back(); }

    std::vector<double> apply(const std::vector<double>& x) const override {
        if (x.size() != n_cols()) throw std::invalid_argument("SparseOperator::apply: x has the wrong length");
        std::vector<double> y(n_rows());
        check(tbem_sparse_apply(h.get(), x.data(), y.data()));
        return y;
    }

    std::unique_ptr<OperatorI> clone() const override {
        return std::unique_ptr<OperatorI>(new SparseOperator(*this));
    }

    // SparseOperator::right_multiply (cpp/sparse_operator.cpp:114-138), bit for bit.
    SparseOperator right_multiply(const SparseOperator& other) const {
        tbem_sparse_t raw = nullptr;
        check(tbem_sparse_right_multiply(h.get(), other.h.get(), &raw));
        return from_handle(raw);
    }

    // SparseOperator::csr_from_coo (cpp/sparse_operator.cpp:140-172).
    static SparseOperator csr_from_coo(size_t n_rows, size_t n_cols, const std::vector<MatrixEntry>& entries) {
        std::vector<size_t> r(entries.size()), c(entries.size());
        std::vector<double> v(entries.size());
        for (size_t i = 0; i < entries.size(); i++) {
            r[i] = entries[i].row();
            c[i] = entries[i].col();
            v[i] = entries[i].value;
        }
        tbem_sparse_t raw = nullptr;
        check(tbem_sparse_from_coo(n_rows, n_cols, entries.size(), r.data(), c.data(), v.data(), &raw));
        return from_handle(raw);
    }

private:
    SparseOperator(std::shared_ptr<tbem_sparse_s> handle, std::vector<double> v, std::vector<size_t> c,
                   std::vector<size_t> r, size_t n_rows, size_t n_cols)
        : h(std::move(handle)), shape{n_rows, n_cols}, values(std::move(v)), column_indices(std::move(c)),
          row_ptrs(std::move(r)) {}

    static SparseOperator from_handle(tbem_sparse_t raw) {
        std::shared_ptr<tbem_sparse_s> handle(raw, tbem_sparse_destroy);
        const size_t nr = tbem_sparse_n_rows(raw), nc = tbem_sparse_n_cols(raw), nnz = tbem_sparse_nnz(raw);
        std::vector<double> v(nnz);
        std::vector<size_t> c(nnz), r(nr + 1);
        check(tbem_sparse_csr(raw, v.data(), c.data(), r.data()));
        return SparseOperator(std::move(handle), std::move(v), std::move(c), std::move(r), nr, nc);
    }
};

// IntegralOperator<3,R,C> (cpp/integral_operator.h:26-64): y = N x + (G F(I x) + C x)
// on the GPU. The far field must be an FMMOperator<3,R,C> or a DirectNBodyOperator.
template <size_t dim, size_t R, size_t C>
struct IntegralOperator : public OperatorI {
    const SparseOperator nearfield;
    const SparseOperator farfield_correction;
    const SparseOperator galerkin;
    const std::shared_ptr<OperatorI> farfield;
    const SparseOperator interp;
    std::shared_ptr<tbem_integral_s> h;

    IntegralOperator(const SparseOperator& nearfield, const SparseOperator& farfield_correction,
                     const SparseOperator& galerkin, const std::shared_ptr<OperatorI>& farfield,
                     const SparseOperator& interp)
        : nearfield(nearfield), farfield_correction(farfield_correction), galerkin(galerkin), farfield(farfield),
          interp(interp) {
        int kind;
        void* far;
        if (auto f = dynamic_cast<FMMOperator<dim, R, C>*>(farfield.get())) {
            kind = TBEM_FARFIELD_FMM;
            far = f->h.get();
        } else if (auto d = dynamic_cast<DirectNBodyOperator*>(farfield.get())) {
            kind = TBEM_FARFIELD_DIRECT;
            far = d->h.get();
        } else {
            throw std::invalid_argument("IntegralOperator: the far field must be an FMMOperator or a "
                                        "DirectNBodyOperator");
        }
        tbem_integral_t raw = nullptr;
        check(tbem_integral_create(this->nearfield.h.get(), this->farfield_correction.h.get(),
                                   this->galerkin.h.get(), kind, far, this->interp.h.get(), &raw));
        h.reset(raw, tbem_integral_destroy);
    }

    size_t n_rows() const override { return galerkin.n_rows(); }
    size_t n_cols() const override { return nearfield.n_cols(); }

    std::vector<double> apply(const std::vector<double>& x) const override {
        if (x.size() != n_cols()) throw std::invalid_argument("IntegralOperator::apply: x has the wrong length");
        std::vector<double> y(n_rows());
        check(tbem_integral_apply(h.get(), x.data(), y.data()));
        return y;
    }

    std::unique_ptr<OperatorI> clone() const override {
        return std::unique_ptr<OperatorI>(new IntegralOperator(nearfield, farfield_correction, galerkin,
                                                               std::shared_ptr<OperatorI>(farfield->clone()),
                                                               interp));
    }
};

// LinearTerm, ConstraintEQ (cpp/constraint.h:21-52) and the helpers of cpp/constraint.cpp.
struct LinearTerm {
    size_t dof;
    double weight;
    bool operator==(const LinearTerm& o) const { return dof == o.dof && weight == o.weight; }
};

struct ConstraintEQ {
    std::vector<LinearTerm> terms;
    double rhs;
};

inline ConstraintEQ boundary_condition(size_t dof, double value) { return {{LinearTerm{dof, 1.0}}, value}; }
inline ConstraintEQ continuity_constraint(size_t dof1, size_t dof2) {
    return {{LinearTerm{dof1, 1.0}, LinearTerm{dof2, -1.0}}, 0.0};
}
inline std::vector<ConstraintEQ> shift_constraints(const std::vector<ConstraintEQ>& cs, size_t shift_dof) {
    std::vector<ConstraintEQ> out;
    for (const auto& c : cs) {
        ConstraintEQ e{{}, c.rhs};
        for (const auto& t : c.terms) e.terms.push_back(LinearTerm{t.dof + shift_dof, t.weight});
        out.push_back(e);
    }
    return out;
}

// ConstraintMatrix (cpp/constraint_matrix.h:17-25): built on the host by
// from_constraints; distribute_vector / condense_vector run on the GPU and equal the
// reference's bit for bit.
struct ConstraintMatrix {
    std::shared_ptr<tbem_constraints_s> h;
    size_t size() const { return tbem_constraints_size(h.get()); }
};

inline ConstraintMatrix from_constraints(const std::vector<ConstraintEQ>& constraints) {
    std::vector<size_t> ptr{0}, dof;
    std::vector<double> w, rhs;
    for (const auto& c : constraints) {
        for (const auto& t : c.terms) {
            dof.push_back(t.dof);
            w.push_back(t.weight);
        }
        ptr.push_back(dof.size());
        rhs.push_back(c.rhs);
    }
    tbem_constraints_t raw = nullptr;
    check(tbem_constraints_create(constraints.size(), ptr.data(), dof.data(), w.data(), rhs.data(), &raw));
    return ConstraintMatrix{std::shared_ptr<tbem_constraints_s>(raw, tbem_constraints_destroy)};
}

inline ConstraintMatrix homogenize_constraints(const ConstraintMatrix& cm) {
    tbem_constraints_t raw = nullptr;
    check(tbem_constraints_homogenize(cm.h.get(), &raw));
    return ConstraintMatrix{std::shared_ptr<tbem_constraints_s>(raw, tbem_constraints_destroy)};
}

inline std::vector<double> distribute_vector(const ConstraintMatrix& cm, const std::vector<double>& in,
                                             size_t n_total_dofs) {
    if (in.size() + cm.size() != n_total_dofs)
        throw std::invalid_argument("distribute_vector: wrong number of reduced values");
    std::vector<double> out(n_total_dofs);
    check(tbem_distribute_vector(cm.h.get(), in.data(), n_total_dofs, out.data()));
    return out;
}

inline std::vector<double> condense_vector(const ConstraintMatrix& cm, const std::vector<double>& all) {
    std::vector<double> out(all.size() - cm.size());
    check(tbem_condense_vector(cm.h.get(), all.data(), all.size(), out.data()));
    return out;
}

}  // namespace tbem

#endif
