// C++ caller of the MI355X FMM through include/tbem/fmm.hpp (the reference's
// FMMOperator<3,R,C> / make_direct_nbody_operator surface on the C ABI).
//
// usage: shim_example <in.bin> <out.bin>
// in.bin : int64 kernel, int64 n, int64 order, int64 leaf, double mac, double d,
//          then n*3 locs, n*3 normals, n weights, R*n densities (obs = src points)
// out.bin: R*n FMM outputs, R*n direct outputs, R*n outputs of a clone
#include <cstdint>
#include <cstdio>
#include <vector>

#include "tbem/fmm.hpp"

using namespace tbem;

template <size_t R>
static int run(const Kernel<3, R, R>& K, const NBodyData<3>& data, const FMMConfig& cfg,
               const std::vector<double>& x, FILE* out) {
    FMMOperator<3, R, R> op(K, data, cfg);
    auto y = op.apply(x);
    auto direct = make_direct_nbody_operator(data, K).apply(x);
    auto c = op.clone();
    auto yc = c->apply(x);
    if (y.size() != op.n_rows() || direct.size() != y.size() || yc.size() != y.size()) return 3;
    fwrite(y.data(), sizeof(double), y.size(), out);
    fwrite(direct.data(), sizeof(double), direct.size(), out);
    fwrite(yc.data(), sizeof(double), yc.size(), out);
    return 0;
}

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    FILE* in = fopen(argv[1], "rb");
    if (!in) return 2;
    int64_t hdr[4];
    double md[2];
    if (fread(hdr, sizeof(int64_t), 4, in) != 4 || fread(md, sizeof(double), 2, in) != 2) return 2;
    const int kernel = (int)hdr[0];
    const size_t n = (size_t)hdr[1];
    const size_t R = (kernel == TBEM_K_LAPLACE_SINGLE || kernel == TBEM_K_LAPLACE_DOUBLE) ? 1 : 3;
    std::vector<double> locs(3 * n), nrm(3 * n), w(n), x(R * n);
    if (fread(locs.data(), 8, 3 * n, in) != 3 * n || fread(nrm.data(), 8, 3 * n, in) != 3 * n ||
        fread(w.data(), 8, n, in) != n || fread(x.data(), 8, R * n, in) != R * n)
        return 2;
    fclose(in);
    NBodyData<3> data;
    data.obs_locs.resize(n);
    data.obs_normals.resize(n);
    for (size_t i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            data.obs_locs[i][d] = locs[3 * i + d];
            data.obs_normals[i][d] = nrm[3 * i + d];
        }
    data.src_locs = data.obs_locs;
    data.src_normals = data.obs_normals;
    data.src_weights = w;
    FMMConfig cfg(md[0], (size_t)hdr[2], (size_t)hdr[3], md[1], false);
    FILE* out = fopen(argv[2], "wb");
    if (!out) return 2;
    int rc;
    try {
        switch (kernel) {
        case TBEM_K_LAPLACE_SINGLE: rc = run<1>(LaplaceSingle<3>(), data, cfg, x, out); break;
        case TBEM_K_LAPLACE_DOUBLE: rc = run<1>(LaplaceDouble<3>(), data, cfg, x, out); break;
        case TBEM_K_ELASTIC_DISPLACEMENT: rc = run<3>(ElasticDisplacement<3>(1.0, 0.25), data, cfg, x, out); break;
        case TBEM_K_ELASTIC_TRACTION: rc = run<3>(ElasticTraction<3>(1.0, 0.25), data, cfg, x, out); break;
        default: rc = run<3>(ElasticHypersingular<3>(1.0, 0.25), data, cfg, x, out); break;
        }
    } catch (const std::exception& e) {
        fprintf(stderr, "shim_example: %s\n", e.what());
        rc = 1;
    }
    fclose(out);
    return rc;
}
