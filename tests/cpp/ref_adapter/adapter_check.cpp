// Checks the 3bem-side adapter (fmm_mi355x.h) inside a build against the
// reference's own sources: for every kernel the GPU serves, make_fmm_operator must
// hand back a GPUFMMOperator whose apply and clone match the reference's CPU
// FMMOperator to 1e-12 (relative L2); a user kernel and a 2D kernel must come back
// as the reference's FMMOperator. Prints one line per case, exits non-zero on a
// mismatch. Run with OMP_NUM_THREADS=1 (the reference's sums are then
// deterministic).
#include <cmath>
#include <cstdio>
#include <random>

#include "elastic_kernels.h"
#include "fmm_mi355x.h"
#include "identity_kernels.h"
#include "laplace_kernels.h"

using namespace tbem;

// a user kernel the GPU library does not know: served by the reference's FMMOperator
struct ScaledLaplace : LaplaceSingle<3> {
    Vec1<Vec1<double>> call(double r2, const Vec<double, 3>& d, const Vec<double, 3>& no,
                            const Vec<double, 3>& ns) const override {
        auto v = LaplaceSingle<3>::call(r2, d, no, ns);
        v[0][0] *= 2.0;
        return v;
    }
    std::unique_ptr<Kernel<3, 1, 1>> clone() const override {
        return std::unique_ptr<Kernel<3, 1, 1>>(new ScaledLaplace());
    }
};

static double rel(const std::vector<double>& a, const std::vector<double>& b) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); i++) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return std::sqrt(num / (den > 0 ? den : 1.0));
}

template <size_t dim>
static NBodyData<dim> points(size_t n, std::mt19937_64& g) {
    std::uniform_real_distribution<double> u(0, 1), w(0.5, 1.5);
    std::normal_distribution<double> z;
    NBodyData<dim> d;
    for (size_t i = 0; i < n; i++) {
        Vec<double, dim> p, q;
        double s = 0;
        for (size_t k = 0; k < dim; k++) {
            p[k] = u(g);
            q[k] = z(g);
            s += q[k] * q[k];
        }
        for (size_t k = 0; k < dim; k++) q[k] /= std::sqrt(s);
        d.obs_locs.push_back(p);
        d.obs_normals.push_back(q);
        d.src_locs.push_back(p);
        d.src_normals.push_back(q);
        d.src_weights.push_back(w(g));
    }
    return d;
}

static int failures = 0;

template <size_t dim, size_t R, size_t C>
static void compare(const char* name, const Kernel<dim, R, C>& K, const NBodyData<dim>& data, bool want_gpu,
                    std::mt19937_64& g) {
    FMMConfig cfg(0.3, 6, 64, 0.05, false);
    std::normal_distribution<double> z;
    std::vector<double> x(C * data.src_locs.size());
    for (auto& v : x) v = z(g);
    auto op = make_fmm_operator(K, data, cfg);
    bool is_cpu = dynamic_cast<FMMOperator<dim, R, C>*>(op.get()) != nullptr;
    FMMOperator<dim, R, C> ref(K, data, cfg);
    auto y_ref = ref.apply(x);
    auto y = op->apply(x);
    auto yc = op->clone()->apply(x);
    double e = rel(y, y_ref), ec = rel(yc, y_ref);
    bool ok = is_cpu != want_gpu && op->n_rows() == ref.n_rows() && op->n_cols() == ref.n_cols() && e < 1e-12 &&
              ec < 1e-12;
    std::printf("%-28s %s  rel err vs reference FMMOperator %.3e, clone %.3e  %s\n", name, is_cpu ? "cpu" : "gpu", e,
                ec, ok ? "OK" : "FAIL");
    failures += ok ? 0 : 1;
}

int main() {
    std::mt19937_64 g(20261203);
    auto d3 = points<3>(3000, g);
    auto d2 = points<2>(1500, g);
    compare("LaplaceSingle<3>", LaplaceSingle<3>(), d3, true, g);
    compare("LaplaceDouble<3>", LaplaceDouble<3>(), d3, true, g);
    compare("ElasticDisplacement<3>", ElasticDisplacement<3>(1.3, 0.27), d3, true, g);
    compare("ElasticTraction<3>", ElasticTraction<3>(1.0, 0.25), d3, true, g);
    compare("ElasticAdjointTraction<3>", ElasticAdjointTraction<3>(1.0, 0.25), d3, true, g);
    compare("ElasticHypersingular<3>", ElasticHypersingular<3>(1.0, 0.25), d3, true, g);
    compare("IdentityTensor<3,3,3>", IdentityTensor<3, 3, 3>(), d3, true, g);
    compare("user kernel (3D)", ScaledLaplace(), d3, false, g);
    compare("LaplaceSingle<2>", LaplaceSingle<2>(), d2, false, g);
    std::printf("%s\n", failures ? "ADAPTER FAIL" : "ADAPTER OK");
    return failures ? 1 : 0;
}
