// tbem/fmm.hpp — header-only C++ surface of 3bem's FMM and direct N-body operators
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

}  // namespace tbem

#endif
