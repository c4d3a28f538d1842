// C++ caller of the GPU sparse and integral operators through include/tbem/fmm.hpp
// (the reference's SparseOperator / IntegralOperator surface on the C ABI).
//
// usage: sparse_example <in.bin> <out.bin>
// in.bin : int64 n_rows, n_cols, nnz, n_pts; then nnz int64 rows, nnz int64 cols,
//          nnz double values, n_cols double x, n_pts*3 locs, n_pts*3 normals, n_pts weights
// out.bin: n_rows y = A x (A from csr_from_coo); n_rows y of the SparseOperator rebuilt
//          from A's CSR members; n_rows y of the integral operator
//          N = A, C = A, G (n_rows x n_pts, G[i, i % n_pts] = 1 + 0.001 i),
//          I (n_pts x n_cols, I[j, j % n_cols] = 0.5), F = Laplace single-layer direct
//          operator on the points; then A's row_ptrs (n_rows + 1 int64) and
//          column_indices (nnz int64), the nnz of A^T A (A^T built by csr_from_coo), and 7
//          doubles: the size, condensed value and distributed vector of the reference's
//          CondenseThenDistribute constraint case.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "tbem/fmm.hpp"

using namespace tbem;

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    FILE* in = fopen(argv[1], "rb");
    if (!in) return 2;
    int64_t hdr[4];
    if (fread(hdr, 8, 4, in) != 4) return 2;
    const size_t n_rows = (size_t)hdr[0], n_cols = (size_t)hdr[1], nnz = (size_t)hdr[2], n_pts = (size_t)hdr[3];
    std::vector<int64_t> rows(nnz), cols(nnz);
    std::vector<double> vals(nnz), x(n_cols), locs(3 * n_pts), nrm(3 * n_pts), w(n_pts);
    if (fread(rows.data(), 8, nnz, in) != nnz || fread(cols.data(), 8, nnz, in) != nnz ||
        fread(vals.data(), 8, nnz, in) != nnz || fread(x.data(), 8, n_cols, in) != n_cols ||
        fread(locs.data(), 8, 3 * n_pts, in) != 3 * n_pts || fread(nrm.data(), 8, 3 * n_pts, in) != 3 * n_pts ||
        fread(w.data(), 8, n_pts, in) != n_pts)
        return 2;
    fclose(in);
    FILE* out = fopen(argv[2], "wb");
    if (!out) return 2;
    try {
        std::vector<MatrixEntry> e(nnz), et(nnz);
        for (size_t k = 0; k < nnz; k++) {
            e[k] = MatrixEntry{{(size_t)rows[k], (size_t)cols[k]}, vals[k]};
            et[k] = MatrixEntry{{(size_t)cols[k], (size_t)rows[k]}, vals[k]};
        }
        auto A = SparseOperator::csr_from_coo(n_rows, n_cols, e);
        auto y = A.apply(x);
        SparseOperator A2(A.n_rows(), A.n_cols(), A.values, A.column_indices, A.row_ptrs);
        auto y2 = A2.clone()->apply(x);

        std::vector<MatrixEntry> g, ie;
        for (size_t i = 0; i < n_rows; i++) g.push_back(MatrixEntry{{i, i % n_pts}, 1.0 + 0.001 * (double)i});
        for (size_t j = 0; j < n_pts; j++) ie.push_back(MatrixEntry{{j, j % n_cols}, 0.5});
        NBodyData<3> data;
        data.obs_locs.resize(n_pts);
        data.obs_normals.resize(n_pts);
        for (size_t i = 0; i < n_pts; i++)
            for (int d = 0; d < 3; d++) {
                data.obs_locs[i][d] = locs[3 * i + d];
                data.obs_normals[i][d] = nrm[3 * i + d];
            }
        data.src_locs = data.obs_locs;
        data.src_normals = data.obs_normals;
        data.src_weights = w;
        std::shared_ptr<OperatorI> far(new DirectNBodyOperator(make_direct_nbody_operator(data, LaplaceSingle<3>())));
        IntegralOperator<3, 1, 1> op(A, A, SparseOperator::csr_from_coo(n_rows, n_pts, g), far,
                                     SparseOperator::csr_from_coo(n_pts, n_cols, ie));
        auto yi = op.clone()->apply(x);
        auto At = SparseOperator::csr_from_coo(n_cols, n_rows, et);
        auto AtA = At.right_multiply(A);
        fwrite(y.data(), 8, n_rows, out);
        fwrite(y2.data(), 8, n_rows, out);
        fwrite(yi.data(), 8, n_rows, out);
        std::vector<int64_t> rp(A.row_ptrs.begin(), A.row_ptrs.end()), ci(A.column_indices.begin(), A.column_indices.end());
        fwrite(rp.data(), 8, rp.size(), out);
        fwrite(ci.data(), 8, ci.size(), out);
        const int64_t nnz_ata = (int64_t)AtA.nnz();
        fwrite(&nnz_ata, 8, 1, out);
        // constraints: the reference's CondenseThenDistribute case
        auto cm = from_constraints({boundary_condition(1, 4.0), continuity_constraint(0, 2),
                                    continuity_constraint(2, 3), continuity_constraint(3, 4)});
        auto red = condense_vector(cm, {2.0, 4.0, 4.0, 4.0, 4.0});
        auto full = distribute_vector(cm, red, 5);
        const double cres[7] = {(double)cm.size(), red[0], full[0], full[1], full[2], full[3], full[4]};
        fwrite(cres, 8, 7, out);
    } catch (const std::exception& ex) {
        fprintf(stderr, "sparse_example: %s\n", ex.what());
        fclose(out);
        return 1;
    }
    fclose(out);
    return 0;
}
