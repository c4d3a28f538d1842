// See fmm_mi355x.h. Instantiated for the reference's 3D FMM operators, <3,1,1> and
// <3,3,3> (cpp/fmm.cpp:653-654); make_fmm_operator also for its 2D ones.
#include "fmm_mi355x.h"

#include <stdexcept>
#include <string>
#include <typeinfo>

#include "elastic_kernels.h"
#include "identity_kernels.h"
#include "laplace_kernels.h"
#include "tbem_fmm.h"

namespace tbem {

static void check(int rc) {
    if (rc != TBEM_OK) throw std::runtime_error(std::string("libtbem_fmm: ") + tbem_last_error());
}

// Exact types only: a user subclass of one of these kernels may override call(), so
// it stays with the reference's own operator.
template <size_t R, size_t C>
int mi355x_kernel_id(const Kernel<3, R, C>& K, double* mu, double* nu) {
    const std::type_info& t = typeid(K);
    if (t == typeid(LaplaceSingle<3>)) return TBEM_K_LAPLACE_SINGLE;
    if (t == typeid(LaplaceDouble<3>)) return TBEM_K_LAPLACE_DOUBLE;
    if (t == typeid(LaplaceHypersingular<3>)) return TBEM_K_LAPLACE_HYPERSINGULAR;
    if (t == typeid(IdentityTensor<3, R, C>)) return R == 1 ? TBEM_K_IDENTITY_SCALAR : TBEM_K_IDENTITY_TENSOR;
    const int id = t == typeid(ElasticDisplacement<3>)      ? TBEM_K_ELASTIC_DISPLACEMENT
                   : t == typeid(ElasticTraction<3>)        ? TBEM_K_ELASTIC_TRACTION
                   : t == typeid(ElasticAdjointTraction<3>) ? TBEM_K_ELASTIC_ADJOINT_TRACTION
                   : t == typeid(ElasticHypersingular<3>)   ? TBEM_K_ELASTIC_HYPERSINGULAR
                                                            : -1;
    if (id < 0) return -1;
    // the four elastic kernels keep mu and nu in the same public members
    // (cpp/elastic_kernels.h:194-204, 233-243, 275-285, 317-329)
    // (dynamic_cast: the scalar instantiation has these types only as unrelated classes)
    if (auto e = dynamic_cast<const ElasticDisplacement<3>*>(&K)) {
        *mu = e->shear_modulus;
        *nu = e->poisson_ratio;
    } else if (auto e = dynamic_cast<const ElasticTraction<3>*>(&K)) {
        *mu = e->shear_modulus;
        *nu = e->poisson_ratio;
    } else if (auto e = dynamic_cast<const ElasticAdjointTraction<3>*>(&K)) {
        *mu = e->shear_modulus;
        *nu = e->poisson_ratio;
    } else if (auto e = dynamic_cast<const ElasticHypersingular<3>*>(&K)) {
        *mu = e->shear_modulus;
        *nu = e->poisson_ratio;
    }
    return id;
}

// constructor body of cpp/fmm.cpp:71-101: the NBodyData vectors of std::array<double,3>
// are contiguous, so they go over as flat AoS arrays
template <size_t R, size_t C>
GPUFMMOperator<R, C>::GPUFMMOperator(const Kernel<3, R, C>& K, const NBodyData<3>& data,
                                     const FMMConfig& config) {
    tbem_fmm_desc d{};
    d.shear_modulus = 1.0;
    d.poisson_ratio = 0.25;
    d.kernel = mi355x_kernel_id(K, &d.shear_modulus, &d.poisson_ratio);
    if (d.kernel < 0) throw std::invalid_argument("GPUFMMOperator: kernel not served by libtbem_fmm");
    d.n_obs = data.obs_locs.size();
    d.obs_locs = d.n_obs ? data.obs_locs[0].data() : nullptr;
    d.obs_normals = d.n_obs ? data.obs_normals[0].data() : nullptr;
    d.n_src = data.src_locs.size();
    d.src_locs = d.n_src ? data.src_locs[0].data() : nullptr;
    d.src_normals = d.n_src ? data.src_normals[0].data() : nullptr;
    d.src_weights = d.n_src ? data.src_weights.data() : nullptr;
    d.mac = config.mac;
    d.order = config.order;
    d.min_pts_per_cell = config.min_pts_per_cell;
    d.d = config.d;
    d.account_for_small_cells = config.account_for_small_cells ? 1 : 0;
    d.precision = TBEM_FP64;
    check(tbem_fmm_create(&d, &h));
}

template <size_t R, size_t C>
GPUFMMOperator<R, C>::~GPUFMMOperator() {
    tbem_fmm_destroy(h);
}

template <size_t R, size_t C>
size_t GPUFMMOperator<R, C>::n_rows() const {
    return tbem_fmm_n_rows(h);
}

template <size_t R, size_t C>
size_t GPUFMMOperator<R, C>::n_cols() const {
    return tbem_fmm_n_cols(h);
}

// FMMOperator::apply (cpp/fmm.cpp:550-558): component-major x in, y out
template <size_t R, size_t C>
std::vector<double> GPUFMMOperator<R, C>::apply(const std::vector<double>& x) const {
    if (x.size() != n_cols()) throw std::invalid_argument("GPUFMMOperator::apply: x has the wrong length");
    std::vector<double> y(n_rows());
    check(tbem_fmm_apply(h, x.data(), y.data()));
    return y;
}

template <size_t R, size_t C>
std::unique_ptr<OperatorI> GPUFMMOperator<R, C>::clone() const {
    std::unique_ptr<GPUFMMOperator<R, C>> c(new GPUFMMOperator<R, C>());
    check(tbem_fmm_clone(h, &c->h));
    return std::move(c);
}

namespace {
template <size_t dim, size_t R, size_t C>
struct Pick {
    static std::unique_ptr<OperatorI> make(const Kernel<dim, R, C>& K, const NBodyData<dim>& data,
                                           const FMMConfig& config) {
        return std::unique_ptr<OperatorI>(new FMMOperator<dim, R, C>(K, data, config));
    }
};
template <size_t R, size_t C>
struct Pick<3, R, C> {
    static std::unique_ptr<OperatorI> make(const Kernel<3, R, C>& K, const NBodyData<3>& data,
                                           const FMMConfig& config) {
        double mu, nu;
        if (mi355x_kernel_id(K, &mu, &nu) >= 0)
            return std::unique_ptr<OperatorI>(new GPUFMMOperator<R, C>(K, data, config));
        return std::unique_ptr<OperatorI>(new FMMOperator<3, R, C>(K, data, config));
    }
};
}  // namespace

template <size_t dim, size_t R, size_t C>
std::unique_ptr<OperatorI> make_fmm_operator(const Kernel<dim, R, C>& K, const NBodyData<dim>& data,
                                             const FMMConfig& config) {
    return Pick<dim, R, C>::make(K, data, config);
}

template struct GPUFMMOperator<1, 1>;
template struct GPUFMMOperator<3, 3>;
template std::unique_ptr<OperatorI> make_fmm_operator(const Kernel<2, 1, 1>&, const NBodyData<2>&, const FMMConfig&);
template std::unique_ptr<OperatorI> make_fmm_operator(const Kernel<2, 2, 2>&, const NBodyData<2>&, const FMMConfig&);
template std::unique_ptr<OperatorI> make_fmm_operator(const Kernel<3, 1, 1>&, const NBodyData<3>&, const FMMConfig&);
template std::unique_ptr<OperatorI> make_fmm_operator(const Kernel<3, 3, 3>&, const NBodyData<3>&, const FMMConfig&);

}  // namespace tbem
