// CSR sparse operator and the integral-operator composite on the GPU.
//
// SparseOperator (cpp/sparse_operator.h:21-60, cpp/sparse_operator.cpp:7-28): an
// n_rows x n_cols CSR matrix, y_i = sum over row i's entries in stored order of
// values[k] * x[column_indices[k]], accumulated from 0.0 (cpp/sparse_operator.cpp:19-28).
// The product here reproduces that sum bit for bit: every product is rounded on
// its own (no fused multiply-add) and each row is summed by one lane in stored order.
//
// IntegralOperator::apply (cpp/integral_operator.h:55-63): y = N x + (G F(I x) + C x)
// with N the near field, C the far-field correction, G the Galerkin operator, I the
// interpolation operator (all CSR) and F the far field, here the GPU FMM or the
// direct N-body operator. The final sum is formed in the reference's order,
// y_i = n_i + (g_i + c_i).
//
// Layout in HBM: values as FP64, column indices as int32 (n_cols < 2^31), row
// pointers as int64, plus a 16-byte descriptor per row tile.
//
// Kernel (k_csr_apply): one workgroup of 512 threads per row tile, a run of at most
// 512 consecutive rows holding at most kChunk = 4096 entries (a longer row is a tile
// of its own and is streamed in chunks of kChunk). The tile's first entry comes with
// the tile, so the entry loads start without waiting for a row pointer. The block
// loads its entries with coalesced reads, forms the products values[k] * x[col[k]]
// and stores them in LDS; then thread t adds up row r0 + t from LDS in stored order.
// HBM traffic per apply: 12 B per entry, 16 B per row (row pointer, output), plus
// the gathered x.
#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>

#include "common.h"

namespace tbem {

void set_last_error(const std::string& m);

constexpr int kCsrBlock = 512;   // threads per workgroup, also the most rows per tile
constexpr int kChunk = 4096;     // entries staged in LDS per pass (32 KB of products)
constexpr int kPerThread = kChunk / kCsrBlock;

// A row tile: rows r0 .. r0 + nr - 1, whose entries start at k0. The tile list ends
// with {n_rows, 0, nnz}, so tile b's entries end at tiles[b + 1].k0.
struct CsrTile {
    int r0, nr;
    long k0;
};

// One row tile per workgroup. out_a and out_b are optional addends: with both
// given, y[r] = out_a[r] + (sum_r + out_b[r]).
__global__ __launch_bounds__(kCsrBlock) void k_csr_apply(const CsrTile* __restrict__ tiles,
                                                       const long* __restrict__ rp,
                                                       const int* __restrict__ ci,
                                                       const double* __restrict__ val,
                                                       const double* __restrict__ x,
                                                       const double* __restrict__ out_a,
                                                       const double* __restrict__ out_b,
                                                       double* __restrict__ y) {
    __shared__ double prod[kChunk];
    const int t = threadIdx.x;
    const CsrTile tl = tiles[blockIdx.x];
    const int r0 = uni(tl.r0);
    const int nr = uni(tl.nr);  // 1..kCsrBlock
    const long k0 = tl.k0;
    const long total = tiles[blockIdx.x + 1].k0 - k0;
    const int n_pass = uni((int)((total + kChunk - 1) / kChunk));
    long mb = 0, me = 0;
    if (t < nr) {
        mb = rp[r0 + t];
        me = rp[r0 + t + 1];
    }
    double s = 0.0;
    for (int pass = 0; pass < n_pass; pass++) {
        const long c0 = k0 + (long)pass * kChunk;
        const int n = (int)std::min<long>(kChunk, k0 + total - c0);
        int col[kPerThread];
        double v[kPerThread];
        // entries are read once: non-temporal loads keep them from displacing x in the caches
#pragma unroll
        for (int j = 0; j < kPerThread; j++) {
            const int i = t + j * kCsrBlock;
            col[j] = i < n ? __builtin_nontemporal_load(ci + c0 + i) : 0;
            v[j] = i < n ? __builtin_nontemporal_load(val + c0 + i) : 0.0;
        }
#pragma unroll
        for (int j = 0; j < kPerThread; j++) {
            const int i = t + j * kCsrBlock;
            if (i < n) prod[i] = __dmul_rn(v[j], x[col[j]]);
        }
        __syncthreads();
        const long lo = std::max(mb, c0), hi = std::min(me, c0 + n);
        const double* p = prod - c0;
        long k = lo;
        for (; k + 4 <= hi; k += 4) {
            const double a0 = p[k], a1 = p[k + 1], a2 = p[k + 2], a3 = p[k + 3];
            s = __dadd_rn(s, a0);
            s = __dadd_rn(s, a1);
            s = __dadd_rn(s, a2);
            s = __dadd_rn(s, a3);
        }
        for (; k < hi; k++) s = __dadd_rn(s, p[k]);
        __syncthreads();
    }
    if (t < nr) {
        const int r = r0 + t;
        y[r] = out_a ? __dadd_rn(out_a[r], __dadd_rn(s, out_b[r])) : s;
    }
}

// Row tiles: runs of consecutive rows, at most kCsrBlock rows and kChunk entries
// each; a row longer than kChunk forms a tile alone. Ends with {n_rows, 0, nnz}.
static std::vector<CsrTile> row_tiles(const std::vector<long>& rp) {
    const int n_rows = (int)rp.size() - 1;
    std::vector<CsrTile> tiles;
    int r = 0;
    while (r < n_rows) {
        const int r_start = r;
        const long k_start = rp[r];
        r++;
        while (r < n_rows && r - r_start < kCsrBlock && rp[r + 1] - k_start <= kChunk) r++;
        tiles.push_back(CsrTile{r_start, r - r_start, k_start});
    }
    tiles.push_back(CsrTile{n_rows, 0, rp[n_rows]});
    return tiles;
}

}  // namespace tbem

struct tbem_sparse_s {
    size_t n_rows = 0, n_cols = 0, nnz = 0;
    std::vector<double> values;   // host copies for tbem_sparse_csr / clone
    std::vector<long> row_ptrs;
    std::vector<int> cols;
    tbem::DevBuf<double> d_val;
    tbem::DevBuf<int> d_col;
    tbem::DevBuf<long> d_rp;
    tbem::DevBuf<tbem::CsrTile> d_tiles;
    int n_blocks = 0;
    tbem::DevBuf<double> xb, yb;  // staging for the host entry point
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::mutex mu;
    ~tbem_sparse_s() {
        if (done) (void)hipEventDestroy(done);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

struct tbem_integral_s {
    tbem_sparse_t near = nullptr, corr = nullptr, gal = nullptr, interp = nullptr;
    int far_kind = 0;
    void* far = nullptr;
    size_t far_rows = 0, far_cols = 0;
    tbem::DevBuf<double> t_interp, t_far, t_near, t_corr, xb, yb;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::mutex mu;
    ~tbem_integral_s() {
        if (done) (void)hipEventDestroy(done);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace tbem {

template <class F>
static int sp_guard(F&& f) {
    try {
        f();
        return TBEM_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return TBEM_E_NOMEM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return TBEM_E_INVALID;
    }
}

// Checks the CSR arrays (row pointers from 0, non-decreasing, ending at nnz;
// every column index below n_cols) and uploads them with the row blocks.
static std::unique_ptr<tbem_sparse_s> make_sparse(size_t n_rows, size_t n_cols, std::vector<double> values,
                                                  std::vector<long> row_ptrs, std::vector<int> cols) {
    const size_t nnz = values.size();
    TBEM_REQUIRE(n_rows < (1u << 31) && n_cols < (1u << 31), TBEM_E_UNSUPPORTED,
                 "n_rows and n_cols must be below 2^31");
    TBEM_REQUIRE(row_ptrs.size() == n_rows + 1 && row_ptrs[0] == 0 && (size_t)row_ptrs[n_rows] == nnz,
                 TBEM_E_INVALID, "row_ptrs must start at 0 and end at nnz");
    for (size_t r = 0; r < n_rows; r++)
        TBEM_REQUIRE(row_ptrs[r] <= row_ptrs[r + 1], TBEM_E_INVALID, "row_ptrs must be non-decreasing");
    auto h = std::unique_ptr<tbem_sparse_s>(new tbem_sparse_s);
    h->n_rows = n_rows;
    h->n_cols = n_cols;
    h->nnz = nnz;
    TBEM_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    TBEM_HIP(hipEventCreateWithFlags(&h->done, hipEventDisableTiming));
    const std::vector<CsrTile> tiles = row_tiles(row_ptrs);
    h->n_blocks = (int)tiles.size() - 1;
    h->d_val.upload(values, h->stream);
    h->d_col.upload(cols, h->stream);
    h->d_rp.upload(row_ptrs, h->stream);
    h->d_tiles.upload(tiles, h->stream);
    TBEM_HIP(hipStreamSynchronize(h->stream));
    h->values = std::move(values);
    h->row_ptrs = std::move(row_ptrs);
    h->cols = std::move(cols);
    TBEM_HIP(hipEventRecord(h->done, h->stream));
    return h;
}

static void check_cols(const size_t* c, size_t nnz, size_t n_cols, std::vector<int>& out) {
    out.resize(nnz);
    for (size_t k = 0; k < nnz; k++) {
        TBEM_REQUIRE(c[k] < n_cols, TBEM_E_INVALID, "column index out of range");
        out[k] = (int)c[k];
    }
}

// y = A x on stream s; with a and b, y = a + (A x + b) elementwise.
static void sparse_run(tbem_sparse_s& A, const double* x, const double* a, const double* b, double* y,
                       hipStream_t s) {
    if (A.n_blocks == 0) return;
    TBEM_REQUIRE(y && (x || A.nnz == 0), TBEM_E_INVALID, "null device buffer");
    k_csr_apply<<<A.n_blocks, kCsrBlock, 0, s>>>(A.d_tiles.p, A.d_rp.p, A.d_col.p, A.d_val.p, x, a, b, y);
    TBEM_LAUNCH_CHECK();
}

static void sparse_apply(tbem_sparse_s& A, const double* x, double* y, hipStream_t s) {
    TBEM_HIP(hipStreamWaitEvent(s, A.done, 0));
    sparse_run(A, x, nullptr, nullptr, y, s);
    TBEM_HIP(hipEventRecord(A.done, s));
}

static void far_apply(tbem_integral_s& I, const double* x, double* y, hipStream_t s) {
    int rc = I.far_kind == TBEM_FARFIELD_FMM ? tbem_fmm_apply_dev((tbem_fmm_t)I.far, x, y, s)
                                              : tbem_nbody_apply_dev((tbem_nbody_t)I.far, x, y, s);
    if (rc != TBEM_OK) throw Error(rc, std::string("far field: ") + tbem_last_error());
}

// y = N x + (G F(I x) + C x), all on stream s: the near field and the correction
// first, then the interpolation and the far field, then the Galerkin product with
// both addends in its epilogue. (Running the sparse products on a second stream
// beside the FMM was slower: they contend with its kernels for the compute units and
// HBM, and the FMM's time grows by more than theirs.)
static void integral_apply(tbem_integral_s& I, const double* x, double* y, hipStream_t s) {
    TBEM_HIP(hipStreamWaitEvent(s, I.done, 0));
    for (tbem_sparse_t A : {I.near, I.corr, I.gal, I.interp}) TBEM_HIP(hipStreamWaitEvent(s, A->done, 0));
    sparse_run(*I.near, x, nullptr, nullptr, I.t_near.p, s);
    sparse_run(*I.corr, x, nullptr, nullptr, I.t_corr.p, s);
    sparse_run(*I.interp, x, nullptr, nullptr, I.t_interp.p, s);
    far_apply(I, I.t_interp.p, I.t_far.p, s);
    sparse_run(*I.gal, I.t_far.p, I.t_near.p, I.t_corr.p, y, s);
    for (tbem_sparse_t A : {I.near, I.corr, I.gal, I.interp}) TBEM_HIP(hipEventRecord(A->done, s));
    TBEM_HIP(hipEventRecord(I.done, s));
}

// For the other translation units: a CSR operator from host arrays, and a read-only
// view of one's host copy.
tbem_sparse_s* sparse_from_host(size_t n_rows, size_t n_cols, std::vector<double> values, std::vector<long> row_ptrs,
                                std::vector<int> cols) {
    return make_sparse(n_rows, n_cols, std::move(values), std::move(row_ptrs), std::move(cols)).release();
}

void sparse_host_view(tbem_sparse_s* h, size_t* n_rows, size_t* n_cols, const std::vector<double>** values,
                      const std::vector<long>** row_ptrs, const std::vector<int>** cols) {
    *n_rows = h->n_rows;
    *n_cols = h->n_cols;
    *values = &h->values;
    *row_ptrs = &h->row_ptrs;
    *cols = &h->cols;
}

template <class H, class Run>
static void host_roundtrip(H& h, hipStream_t st, const double* x, size_t nin, double* y, size_t nout, Run run) {
    TBEM_REQUIRE((x || nin == 0) && (y || nout == 0), TBEM_E_INVALID, "null argument");
    if (h.xb.n < nin) h.xb.alloc(nin);
    if (h.yb.n < nout) h.yb.alloc(nout);
    h.xb.upload(x, nin, st);
    run(h.xb.p, h.yb.p, st);
    h.yb.download(y, nout, st);
    TBEM_HIP(hipStreamSynchronize(st));
}

}  // namespace tbem

using namespace tbem;

extern "C" {

int tbem_sparse_create(size_t n_rows, size_t n_cols, size_t nnz, const double* values,
                       const size_t* column_indices, const size_t* row_ptrs, tbem_sparse_t* out) {
    return sp_guard([&] {
        TBEM_REQUIRE(out && row_ptrs && (nnz == 0 || (values && column_indices)), TBEM_E_INVALID,
                     "null argument");
        *out = nullptr;
        TBEM_REQUIRE(n_rows < (1u << 31) && n_cols < (1u << 31), TBEM_E_UNSUPPORTED,
                     "n_rows and n_cols must be below 2^31");
        std::vector<int> cols;
        check_cols(column_indices, nnz, n_cols, cols);
        std::vector<long> rp(n_rows + 1);
        for (size_t r = 0; r <= n_rows; r++) {
            TBEM_REQUIRE(row_ptrs[r] <= nnz, TBEM_E_INVALID, "row_ptrs must start at 0 and end at nnz");
            rp[r] = (long)row_ptrs[r];
        }
        *out = make_sparse(n_rows, n_cols, std::vector<double>(values, values + nnz), std::move(rp),
                           std::move(cols))
                   .release();
    });
}

int tbem_sparse_from_coo(size_t n_rows, size_t n_cols, size_t n_entries, const size_t* rows, const size_t* cols,
                         const double* values, tbem_sparse_t* out) {
    return sp_guard([&] {
        TBEM_REQUIRE(out && (n_entries == 0 || (rows && cols && values)), TBEM_E_INVALID, "null argument");
        *out = nullptr;
        TBEM_REQUIRE(n_rows < (1u << 31) && n_cols < (1u << 31), TBEM_E_UNSUPPORTED,
                     "n_rows and n_cols must be below 2^31");
        std::vector<long> rp(n_rows + 1, 0);
        for (size_t i = 0; i < n_entries; i++) {
            TBEM_REQUIRE(rows[i] < n_rows, TBEM_E_INVALID, "row index out of range");
            TBEM_REQUIRE(cols[i] < n_cols, TBEM_E_INVALID, "column index out of range");
            rp[rows[i] + 1]++;
        }
        for (size_t r = 0; r < n_rows; r++) rp[r + 1] += rp[r];
        std::vector<long> fill(rp.begin(), rp.end() - 1);
        std::vector<double> v(n_entries);
        std::vector<int> c(n_entries);
        for (size_t i = 0; i < n_entries; i++) {  // input order within each row
            const long slot = fill[rows[i]]++;
            v[slot] = values[i];
            c[slot] = (int)cols[i];
        }
        *out = make_sparse(n_rows, n_cols, std::move(v), std::move(rp), std::move(c)).release();
    });
}

int tbem_sparse_apply_dev(tbem_sparse_t h, const double* d_x, double* d_y, void* stream) {
    return sp_guard([&] {
        TBEM_REQUIRE(h, TBEM_E_INVALID, "null handle");
        std::lock_guard<std::mutex> lk(h->mu);
        sparse_apply(*h, d_x, d_y, (hipStream_t)stream);
    });
}

int tbem_sparse_apply(tbem_sparse_t h, const double* x, double* y) {
    return sp_guard([&] {
        TBEM_REQUIRE(h, TBEM_E_INVALID, "null handle");
        std::lock_guard<std::mutex> lk(h->mu);
        host_roundtrip(*h, h->stream, x, h->n_cols, y, h->n_rows,
                       [&](const double* dx, double* dy, hipStream_t s) { sparse_apply(*h, dx, dy, s); });
    });
}

int tbem_sparse_csr(tbem_sparse_t h, double* values, size_t* column_indices, size_t* row_ptrs) {
    return sp_guard([&] {
        TBEM_REQUIRE(h, TBEM_E_INVALID, "null handle");
        if (values) std::memcpy(values, h->values.data(), h->nnz * sizeof(double));
        if (column_indices)
            for (size_t k = 0; k < h->nnz; k++) column_indices[k] = (size_t)h->cols[k];
        if (row_ptrs)
            for (size_t r = 0; r <= h->n_rows; r++) row_ptrs[r] = (size_t)h->row_ptrs[r];
    });
}

int tbem_sparse_clone(tbem_sparse_t h, tbem_sparse_t* out) {
    return sp_guard([&] {
        TBEM_REQUIRE(h && out, TBEM_E_INVALID, "null argument");
        *out = make_sparse(h->n_rows, h->n_cols, h->values, h->row_ptrs, h->cols).release();
    });
}

int tbem_sparse_right_multiply(tbem_sparse_t a, tbem_sparse_t b, tbem_sparse_t* out) {
    return sp_guard([&] {
        TBEM_REQUIRE(a && b && out, TBEM_E_INVALID, "null argument");
        *out = nullptr;
        TBEM_REQUIRE(a->n_cols == b->n_rows, TBEM_E_INVALID, "inner dimensions differ");
        // Row i of A B: for each entry (i, j) of A in stored order and each entry
        // (j, c) of B in stored order, acc[c] += b_jc * a_ij, each acc starting from
        // 0.0 at its first touch; the touched columns are emitted in ascending order,
        // explicit zeros included.
        std::vector<double> acc(b->n_cols, 0.0);
        std::vector<char> seen(b->n_cols, 0);
        std::vector<int> touched;
        std::vector<long> rp(a->n_rows + 1, 0);
        std::vector<int> cols;
        std::vector<double> vals;
        for (size_t i = 0; i < a->n_rows; i++) {
            touched.clear();
            for (long ka = a->row_ptrs[i]; ka < a->row_ptrs[i + 1]; ka++) {
                const int j = a->cols[ka];
                const double av = a->values[ka];
                for (long kb = b->row_ptrs[j]; kb < b->row_ptrs[j + 1]; kb++) {
                    const int c = b->cols[kb];
                    if (!seen[c]) {
                        seen[c] = 1;
                        acc[c] = 0.0;
                        touched.push_back(c);
                    }
                    acc[c] += b->values[kb] * av;
                }
            }
            std::sort(touched.begin(), touched.end());
            for (int c : touched) {
                cols.push_back(c);
                vals.push_back(acc[c]);
                seen[c] = 0;
            }
            rp[i + 1] = (long)cols.size();
        }
        *out = make_sparse(a->n_rows, b->n_cols, std::move(vals), std::move(rp), std::move(cols)).release();
    });
}

size_t tbem_sparse_n_rows(tbem_sparse_t h) { return h ? h->n_rows : 0; }
size_t tbem_sparse_n_cols(tbem_sparse_t h) { return h ? h->n_cols : 0; }
size_t tbem_sparse_nnz(tbem_sparse_t h) { return h ? h->nnz : 0; }

void tbem_sparse_destroy(tbem_sparse_t h) {
    if (!h) return;
    (void)hipEventSynchronize(h->done);
    (void)hipStreamSynchronize(h->stream);
    delete h;
}

int tbem_integral_create(tbem_sparse_t nearfield, tbem_sparse_t farfield_correction, tbem_sparse_t galerkin,
                         int farfield_kind, void* farfield, tbem_sparse_t interp, tbem_integral_t* out) {
    return sp_guard([&] {
        TBEM_REQUIRE(out && nearfield && farfield_correction && galerkin && farfield && interp, TBEM_E_INVALID,
                     "null argument");
        *out = nullptr;
        TBEM_REQUIRE(farfield_kind == TBEM_FARFIELD_FMM || farfield_kind == TBEM_FARFIELD_DIRECT, TBEM_E_INVALID,
                     "farfield_kind must be TBEM_FARFIELD_FMM or TBEM_FARFIELD_DIRECT");
        auto h = std::unique_ptr<tbem_integral_s>(new tbem_integral_s);
        h->near = nearfield;
        h->corr = farfield_correction;
        h->gal = galerkin;
        h->interp = interp;
        h->far_kind = farfield_kind;
        h->far = farfield;
        if (farfield_kind == TBEM_FARFIELD_FMM) {
            h->far_rows = tbem_fmm_n_rows((tbem_fmm_t)farfield);
            h->far_cols = tbem_fmm_n_cols((tbem_fmm_t)farfield);
        } else {
            h->far_rows = tbem_nbody_n_rows((tbem_nbody_t)farfield);
            h->far_cols = tbem_nbody_n_cols((tbem_nbody_t)farfield);
        }
        const size_t n_rows = galerkin->n_rows, n_cols = nearfield->n_cols;
        TBEM_REQUIRE(nearfield->n_rows == n_rows && farfield_correction->n_rows == n_rows &&
                         farfield_correction->n_cols == n_cols,
                     TBEM_E_INVALID, "nearfield, farfield_correction and galerkin must have the same row count");
        TBEM_REQUIRE(galerkin->n_cols == h->far_rows, TBEM_E_INVALID,
                     "galerkin columns must match the far field's rows");
        TBEM_REQUIRE(interp->n_rows == h->far_cols && interp->n_cols == n_cols, TBEM_E_INVALID,
                     "interp must map the operator's columns to the far field's columns");
        TBEM_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        TBEM_HIP(hipEventCreateWithFlags(&h->done, hipEventDisableTiming));
        h->t_interp.alloc(h->far_cols);
        h->t_far.alloc(h->far_rows);
        h->t_near.alloc(n_rows);
        h->t_corr.alloc(n_rows);
        TBEM_HIP(hipEventRecord(h->done, h->stream));
        *out = h.release();
    });
}

int tbem_integral_apply_dev(tbem_integral_t h, const double* d_x, double* d_y, void* stream) {
    return sp_guard([&] {
        TBEM_REQUIRE(h, TBEM_E_INVALID, "null handle");
        std::lock_guard<std::mutex> lk(h->mu);
        integral_apply(*h, d_x, d_y, (hipStream_t)stream);
    });
}

int tbem_integral_apply(tbem_integral_t h, const double* x, double* y) {
    return sp_guard([&] {
        TBEM_REQUIRE(h, TBEM_E_INVALID, "null handle");
        std::lock_guard<std::mutex> lk(h->mu);
        host_roundtrip(*h, h->stream, x, h->near->n_cols, y, h->gal->n_rows,
                       [&](const double* dx, double* dy, hipStream_t s) { integral_apply(*h, dx, dy, s); });
    });
}

size_t tbem_integral_n_rows(tbem_integral_t h) { return h ? h->gal->n_rows : 0; }
size_t tbem_integral_n_cols(tbem_integral_t h) { return h ? h->near->n_cols : 0; }

void tbem_integral_destroy(tbem_integral_t h) {
    if (!h) return;
    (void)hipEventSynchronize(h->done);
    (void)hipStreamSynchronize(h->stream);
    delete h;
}

}  // extern "C"
