// Linear constraints on the dofs: from_constraints on the host, distribute_vector and
// condense_vector on the GPU, condense_matrix for sparse operators on the host.
//
// Reference: cpp/constraint.{h,cpp} and cpp/constraint_matrix.{h,cpp}.
//   * from_constraints (constraint_matrix.cpp:35-52): each equation has its duplicate
//     dofs combined (constraint.cpp:170-183), terms with |w| <= 1e-15 dropped
//     (constraint.cpp:160-168), every already-constrained highest dof substituted away
//     (make_lower_triangular, constraint_matrix.cpp:17-33; substitute,
//     constraint.cpp:123-158), and the highest dof isolated (isolate_term_on_lhs,
//     constraint.cpp:66-84). Reproduced step for step, so the rearranged weights are
//     the reference's bit for bit.
//   * distribute_vector (constraint_matrix.cpp:54-82): free dofs take the reduced
//     vector in order; constrained dof c, in ascending order, takes
//     rhs_c + sum_j w_j out[d_j] (d_j < c), summed in term order from rhs_c.
//   * condense_vector (constraint_matrix.cpp:99-125): dofs visited in descending
//     order; a constrained dof c passes w_j (all[c] + acc[c]) to acc[d_j]; a free dof
//     u ends as 0.0 + (all[u] + acc[u]).
//
// On the GPU both are level-scheduled gathers. For distribute, level(c) is one more
// than the highest level among c's constrained terms (0 if all its terms are free);
// each level is one launch of one thread per constrained dof, summing in the
// reference's term order. For condense, the transposed terms of dof u are the pairs
// (c, w) with u among c's terms, ordered by c descending and then term order: the
// order in which the reference's descending sweep adds them to acc[u]. A constrained
// dof's level is one more than the highest level among the constraints that name it.
// Levels run top down, then one launch finishes the free dofs. Every product is rounded
// on its own and every sum follows the reference's order, so both vectors equal the
// reference's bit for bit. Continuity chains give a handful of levels (one per link).
#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>

#include "common.h"

struct tbem_sparse_s;
namespace tbem {
void set_last_error(const std::string& m);
}

namespace {

struct Term {
    size_t dof;
    double weight;
    bool operator==(const Term& o) const { return dof == o.dof && weight == o.weight; }
};

struct Eq {
    std::vector<Term> terms;
    double rhs;
};

// A constraint solved for its highest dof: x_c = rhs + sum w_j x_{d_j}.
struct Solved {
    std::vector<Term> terms;
    double rhs;
};

using SolvedMap = std::unordered_map<size_t, Solved>;

// Duplicate dofs merged into the first occurrence (constraint.cpp:170-183).
Eq combine_terms(const Eq& c) {
    Eq out{{}, c.rhs};
    for (const Term& t : c.terms) {
        auto it = std::find_if(out.terms.begin(), out.terms.end(), [&](const Term& o) { return o.dof == t.dof; });
        if (it == out.terms.end()) out.terms.push_back(t);
        else it->weight += t.weight;
    }
    return out;
}

// Terms with |w| <= eps dropped (constraint.cpp:160-168).
Eq filter_zero_terms(const Eq& c, double eps = 1e-15) {
    Eq out{{}, c.rhs};
    for (const Term& t : c.terms)
        if (std::fabs(t.weight) > eps) out.terms.push_back(t);
    return out;
}

// Index of the first term with the highest dof (constraint.cpp:86-94).
size_t last_dof_index(const Eq& c) {
    size_t best = 0;
    for (size_t i = 1; i < c.terms.size(); i++)
        if (c.terms[best].dof < c.terms[i].dof) best = i;
    return best;
}

// x_k eliminated from c with x_k = s (constraint.cpp:123-158). The list of unused
// terms of s loses, for each term of c that s also holds, the entry at that term's
// POSITION in s (not its value), as the reference's remove() does; a position past
// the list's end is where the reference's erase is undefined, so it is refused.
Eq substitute(const Eq& c, size_t k_index, const Solved& s) {
    const double factor = c.terms[k_index].weight;
    Eq out{{}, 0.0};
    std::vector<size_t> unused(s.terms.size());
    for (size_t i = 0; i < unused.size(); i++) unused[i] = i;
    for (size_t i = 0; i < c.terms.size(); i++) {
        if (i == k_index) continue;
        const Term& t = c.terms[i];
        auto it = std::find_if(s.terms.begin(), s.terms.end(), [&](const Term& o) { return o.dof == t.dof; });
        if (it == s.terms.end()) {
            out.terms.push_back(t);
        } else {
            out.terms.push_back(Term{t.dof, t.weight + it->weight * factor});
            const size_t pos = (size_t)(it - s.terms.begin());
            TBEM_REQUIRE(pos < unused.size(), TBEM_E_UNSUPPORTED,
                         "from_constraints: substitution pattern outside what the reference defines");
            unused.erase(unused.begin() + pos);
        }
    }
    for (size_t u : unused) out.terms.push_back(Term{s.terms[u].dof, s.terms[u].weight * factor});
    out.rhs = c.rhs - factor * s.rhs;
    return out;
}

// Substitute already-constrained highest dofs until the highest dof is free
// (constraint_matrix.cpp:17-33).
Eq make_lower_triangular(Eq c, const SolvedMap& m) {
    while (!c.terms.empty()) {
        const size_t k = last_dof_index(c);
        auto it = m.find(c.terms[k].dof);
        if (it == m.end()) break;
        c = filter_zero_terms(substitute(c, k, it->second));
    }
    return c;
}

// x_k isolated: the other terms negated and divided by w_k (constraint.cpp:66-84).
// Terms equal to the isolated one (same dof and weight) are skipped, as there.
Solved isolate(const Eq& c, size_t k) {
    const Term ct = c.terms[k];
    Solved out{{}, c.rhs / ct.weight};
    for (const Term& t : c.terms) {
        if (t == ct) continue;
        out.terms.push_back(Term{t.dof, -t.weight / ct.weight});
    }
    return out;
}

// One level-scheduled step: thread i sums row i of (ptr, col, w) over x from init[i]
// (or 0.0), then writes add[dof[i]] + s (or s), plus 0.0 in front when finish is set,
// to out[out_idx[i]] (or out[dof[i]]).
__global__ __launch_bounds__(256) void k_cons_rows(int n, const int* __restrict__ dof, const int* __restrict__ ptr,
                                                   const int* __restrict__ col, const double* __restrict__ w,
                                                   const double* __restrict__ init, const double* x,
                                                   const double* __restrict__ add, int finish,
                                                   const int* __restrict__ out_idx, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = init ? init[i] : 0.0;
    for (int k = ptr[i]; k < ptr[i + 1]; k++) s = __dadd_rn(s, __dmul_rn(w[k], x[col[k]]));
    const int d = dof[i];
    double v = add ? __dadd_rn(add[d], s) : s;
    if (finish) v = __dadd_rn(0.0, v);
    out[out_idx ? out_idx[i] : d] = v;
}

__global__ __launch_bounds__(256) void k_cons_scatter(int n, const int* __restrict__ free_dof,
                                                      const double* __restrict__ in, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[free_dof[i]] = in[i];
}

// Rows of one launch, on the device.
struct RowSet {
    int n = 0;
    tbem::DevBuf<int> dof, ptr, col, out_idx;
    tbem::DevBuf<double> w, init;
};

struct HostRows {
    std::vector<int> dof, ptr{0}, col, out_idx;
    std::vector<double> w, init;
    void upload(RowSet& r, hipStream_t s) const {
        r.n = (int)dof.size();
        r.dof.upload(dof, s);
        r.ptr.upload(ptr, s);
        r.col.upload(col, s);
        r.w.upload(w, s);
        if (!init.empty()) r.init.upload(init, s);
        if (!out_idx.empty()) r.out_idx.upload(out_idx, s);
    }
};

// Device plan for vectors of n_total dofs.
struct ConsPlan {
    size_t n_total = 0;
    int n_free = 0;
    tbem::DevBuf<int> free_dof;
    std::vector<RowSet> dist;   // distribute levels, bottom up
    std::vector<RowSet> cond;   // condense levels for constrained dofs, top down
    RowSet cond_free;           // condense: the free dofs
    tbem::DevBuf<double> acc;   // condense: values of the constrained dofs
};

}  // namespace

// Calls on one handle are ordered, even across streams: each waits for the previous
// call's device work (done) before using the handle's plan and scratch, and marks its
// own end.
struct tbem_constraints_s {
    SolvedMap map;
    std::unique_ptr<ConsPlan> plan;
    tbem::DevBuf<double> xb, yb;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::mutex mu;
    void ready() {
        if (!stream) TBEM_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        if (!done) {
            TBEM_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
            TBEM_HIP(hipEventRecord(done, stream));
        }
    }
    ~tbem_constraints_s() {
        if (done) (void)hipEventDestroy(done);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace tbem {
// defined in sparse.hip: CSR operator from host arrays
tbem_sparse_s* sparse_from_host(size_t n_rows, size_t n_cols, std::vector<double> values, std::vector<long> row_ptrs,
                                std::vector<int> cols);
void sparse_host_view(tbem_sparse_s* h, size_t* n_rows, size_t* n_cols, const std::vector<double>** values,
                      const std::vector<long>** row_ptrs, const std::vector<int>** cols);
}  // namespace tbem

namespace {

template <class F>
int cs_guard(F&& f) {
    try {
        f();
        return TBEM_OK;
    } catch (const tbem::Error& e) {
        tbem::set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        tbem::set_last_error("host allocation failed");
        return TBEM_E_NOMEM;
    } catch (const std::exception& e) {
        tbem::set_last_error(e.what());
        return TBEM_E_INVALID;
    }
}

std::vector<size_t> sorted_dofs(const SolvedMap& m) {
    std::vector<size_t> d;
    d.reserve(m.size());
    for (const auto& kv : m) d.push_back(kv.first);
    std::sort(d.begin(), d.end());
    return d;
}

void build_plan(tbem_constraints_s& h, size_t n_total) {
    TBEM_REQUIRE(n_total < (1u << 31), TBEM_E_UNSUPPORTED, "at most 2^31 - 1 dofs");
    const std::vector<size_t> cdofs = sorted_dofs(h.map);
    TBEM_REQUIRE(cdofs.empty() || cdofs.back() < n_total, TBEM_E_INVALID,
                 "a constrained dof lies beyond the vector's length");
    hipStream_t s = h.stream;
    auto P = std::unique_ptr<ConsPlan>(new ConsPlan);
    P->n_total = n_total;
    std::vector<char> constrained(n_total, 0);
    for (size_t c : cdofs) constrained[c] = 1;
    std::vector<int> free_dof, reduced(n_total, -1);
    for (size_t d = 0; d < n_total; d++)
        if (!constrained[d]) {
            reduced[d] = (int)free_dof.size();
            free_dof.push_back((int)d);
        }
    P->n_free = (int)free_dof.size();
    P->free_dof.upload(free_dof, s);

    // distribute: levels bottom up, ascending dofs within a level
    std::vector<int> lev(n_total, -1);
    int n_lev = 0;
    for (size_t c : cdofs) {
        int l = 0;
        for (const Term& t : h.map.at(c).terms) {
            TBEM_REQUIRE(t.dof < c, TBEM_E_INVALID, "a constraint names a dof above its own");
            if (constrained[t.dof]) l = std::max(l, lev[t.dof] + 1);
        }
        lev[c] = l;
        n_lev = std::max(n_lev, l + 1);
    }
    std::vector<HostRows> dl(n_lev);
    for (size_t c : cdofs) {
        HostRows& r = dl[lev[c]];
        const Solved& e = h.map.at(c);
        r.dof.push_back((int)c);
        for (const Term& t : e.terms) {
            r.col.push_back((int)t.dof);
            r.w.push_back(t.weight);
        }
        r.ptr.push_back((int)r.col.size());
        r.init.push_back(e.rhs);
    }
    P->dist.resize(n_lev);
    for (int l = 0; l < n_lev; l++) dl[l].upload(P->dist[l], s);

    // condense: transposed terms, ordered by naming constraint descending, then term order
    std::vector<std::vector<std::pair<int, double>>> tr(n_total);
    for (auto it = cdofs.rbegin(); it != cdofs.rend(); ++it)
        for (const Term& t : h.map.at(*it).terms) tr[t.dof].push_back({(int)*it, t.weight});
    std::vector<int> top(n_total, -1);
    int n_top = 0;
    for (auto it = cdofs.rbegin(); it != cdofs.rend(); ++it) {
        int l = 0;
        for (const auto& p : tr[*it]) l = std::max(l, top[p.first] + 1);
        top[*it] = l;
        n_top = std::max(n_top, l + 1);
    }
    std::vector<HostRows> cl(n_top);
    for (auto it = cdofs.rbegin(); it != cdofs.rend(); ++it) {
        HostRows& r = cl[top[*it]];
        r.dof.push_back((int)*it);
        for (const auto& p : tr[*it]) {
            r.col.push_back(p.first);
            r.w.push_back(p.second);
        }
        r.ptr.push_back((int)r.col.size());
    }
    P->cond.resize(n_top);
    for (int l = 0; l < n_top; l++) cl[l].upload(P->cond[l], s);
    HostRows fr;
    for (int d : free_dof) {
        fr.dof.push_back(d);
        fr.out_idx.push_back(reduced[d]);
        for (const auto& p : tr[d]) {
            fr.col.push_back(p.first);
            fr.w.push_back(p.second);
        }
        fr.ptr.push_back((int)fr.col.size());
    }
    fr.upload(P->cond_free, s);
    P->acc.alloc(n_total);
    TBEM_HIP(hipStreamSynchronize(s));
    h.plan = std::move(P);
}

ConsPlan& plan_for(tbem_constraints_s& h, size_t n_total) {
    if (!h.plan || h.plan->n_total != n_total) {
        TBEM_HIP(hipEventSynchronize(h.done));  // no earlier call may still read the old plan
        build_plan(h, n_total);
    }
    return *h.plan;
}

void launch_rows(const RowSet& r, const double* x, const double* add, int finish, bool use_init, double* out,
                 hipStream_t s) {
    if (r.n == 0) return;
    k_cons_rows<<<tbem::div_up(r.n, 256), 256, 0, s>>>(r.n, r.dof.p, r.ptr.p, r.col.p, r.w.p,
                                                        use_init ? r.init.p : nullptr, x, add, finish,
                                                        r.out_idx.p, out);
    TBEM_LAUNCH_CHECK();
}

void distribute(tbem_constraints_s& h, const double* in, size_t n_total, double* out, hipStream_t s) {
    ConsPlan& P = plan_for(h, n_total);
    TBEM_HIP(hipStreamWaitEvent(s, h.done, 0));
    if (P.n_free) {
        k_cons_scatter<<<tbem::div_up(P.n_free, 256), 256, 0, s>>>(P.n_free, P.free_dof.p, in, out);
        TBEM_LAUNCH_CHECK();
    }
    for (const RowSet& r : P.dist) launch_rows(r, out, nullptr, 0, true, out, s);
    TBEM_HIP(hipEventRecord(h.done, s));
}

void condense(tbem_constraints_s& h, const double* all, size_t n_all, double* out, hipStream_t s) {
    ConsPlan& P = plan_for(h, n_all);
    TBEM_HIP(hipStreamWaitEvent(s, h.done, 0));
    for (const RowSet& r : P.cond) launch_rows(r, P.acc.p, all, 0, false, P.acc.p, s);
    launch_rows(P.cond_free, P.acc.p, all, 1, false, out, s);
    TBEM_HIP(hipEventRecord(h.done, s));
}

// add_entry_with_constraints (constraint_matrix.cpp:127-165): an entry on a free
// row and column lands in the map; otherwise it is passed on, weight times value,
// along the constrained row's (or else column's) terms.
void add_entry(const SolvedMap& rows, const SolvedMap& cols, size_t n_cols, std::map<size_t, double>& dok, size_t r,
               size_t c, double v) {
    auto ri = rows.find(r);
    if (ri == rows.end()) {
        auto ci = cols.find(c);
        if (ci == cols.end()) {
            dok[r * n_cols + c] += v;
            return;
        }
        for (const Term& t : ci->second.terms) add_entry(rows, cols, n_cols, dok, r, t.dof, t.weight * v);
        return;
    }
    for (const Term& t : ri->second.terms) add_entry(rows, cols, n_cols, dok, t.dof, c, t.weight * v);
}

}  // namespace

using namespace tbem;

extern "C" {

int tbem_constraints_create(size_t n_eq, const size_t* eq_ptr, const size_t* term_dof, const double* term_weight,
                            const double* rhs, tbem_constraints_t* out) {
    return cs_guard([&] {
        TBEM_REQUIRE(out && (n_eq == 0 || (eq_ptr && rhs)), TBEM_E_INVALID, "null argument");
        *out = nullptr;
        TBEM_REQUIRE(n_eq == 0 || eq_ptr[0] == 0, TBEM_E_INVALID, "eq_ptr must start at 0");
        TBEM_REQUIRE(n_eq == 0 || eq_ptr[n_eq] == 0 || (term_dof && term_weight), TBEM_E_INVALID, "null terms");
        auto h = std::unique_ptr<tbem_constraints_s>(new tbem_constraints_s);
        for (size_t e = 0; e < n_eq; e++) {
            TBEM_REQUIRE(eq_ptr[e] <= eq_ptr[e + 1], TBEM_E_INVALID, "eq_ptr must not decrease");
            Eq c{{}, rhs[e]};
            for (size_t k = eq_ptr[e]; k < eq_ptr[e + 1]; k++) c.terms.push_back(Term{term_dof[k], term_weight[k]});
            Eq lt = make_lower_triangular(filter_zero_terms(combine_terms(c)), h->map);
            if (lt.terms.empty()) continue;
            const size_t k = last_dof_index(lt);
            h->map.insert(std::make_pair(lt.terms[k].dof, isolate(lt, k)));
        }
        *out = h.release();
    });
}

size_t tbem_constraints_size(tbem_constraints_t h) { return h ? h->map.size() : 0; }

size_t tbem_constraints_n_terms(tbem_constraints_t h) {
    if (!h) return 0;
    size_t n = 0;
    for (const auto& kv : h->map) n += kv.second.terms.size();
    return n;
}

int tbem_constraints_get(tbem_constraints_t h, size_t* constrained_dof, size_t* term_ptr, size_t* term_dof,
                         double* term_weight, double* rhs) {
    return cs_guard([&] {
        TBEM_REQUIRE(h, TBEM_E_INVALID, "null handle");
        size_t k = 0, i = 0;
        if (term_ptr) term_ptr[0] = 0;
        for (size_t c : sorted_dofs(h->map)) {
            const Solved& e = h->map.at(c);
            if (constrained_dof) constrained_dof[i] = c;
            if (rhs) rhs[i] = e.rhs;
            for (const Term& t : e.terms) {
                if (term_dof) term_dof[k] = t.dof;
                if (term_weight) term_weight[k] = t.weight;
                k++;
            }
            i++;
            if (term_ptr) term_ptr[i] = k;
        }
    });
}

int tbem_constraints_homogenize(tbem_constraints_t h, tbem_constraints_t* out) {
    return cs_guard([&] {
        TBEM_REQUIRE(h && out, TBEM_E_INVALID, "null argument");
        auto o = std::unique_ptr<tbem_constraints_s>(new tbem_constraints_s);
        for (const auto& kv : h->map) o->map.insert(std::make_pair(kv.first, Solved{kv.second.terms, 0.0}));
        *out = o.release();
    });
}

int tbem_distribute_vector_dev(tbem_constraints_t h, const double* d_in, size_t n_total, double* d_out,
                               void* stream) {
    return cs_guard([&] {
        TBEM_REQUIRE(h && (d_out || n_total == 0), TBEM_E_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(h->mu);
        h->ready();
        distribute(*h, d_in, n_total, d_out, (hipStream_t)stream);
    });
}

int tbem_condense_vector_dev(tbem_constraints_t h, const double* d_all, size_t n_all, double* d_out, void* stream) {
    return cs_guard([&] {
        TBEM_REQUIRE(h && (d_all || n_all == 0), TBEM_E_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(h->mu);
        h->ready();
        condense(*h, d_all, n_all, d_out, (hipStream_t)stream);
    });
}

int tbem_distribute_vector(tbem_constraints_t h, const double* in, size_t n_total, double* out) {
    return cs_guard([&] {
        TBEM_REQUIRE(h && (out || n_total == 0), TBEM_E_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(h->mu);
        h->ready();
        ConsPlan& P = plan_for(*h, n_total);
        TBEM_REQUIRE(in || P.n_free == 0, TBEM_E_INVALID, "null argument");
        if (h->xb.n < (size_t)P.n_free) h->xb.alloc(P.n_free);
        if (h->yb.n < n_total) h->yb.alloc(n_total);
        h->xb.upload(in, P.n_free, h->stream);
        distribute(*h, h->xb.p, n_total, h->yb.p, h->stream);
        h->yb.download(out, n_total, h->stream);
        TBEM_HIP(hipStreamSynchronize(h->stream));
    });
}

int tbem_condense_vector(tbem_constraints_t h, const double* all, size_t n_all, double* out) {
    return cs_guard([&] {
        TBEM_REQUIRE(h && (all || n_all == 0), TBEM_E_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(h->mu);
        h->ready();
        ConsPlan& P = plan_for(*h, n_all);
        TBEM_REQUIRE(out || P.n_free == 0, TBEM_E_INVALID, "null argument");
        if (h->xb.n < n_all) h->xb.alloc(n_all);
        if (h->yb.n < (size_t)P.n_free) h->yb.alloc(P.n_free);
        h->xb.upload(all, n_all, h->stream);
        condense(*h, h->xb.p, n_all, h->yb.p, h->stream);
        h->yb.download(out, P.n_free, h->stream);
        TBEM_HIP(hipStreamSynchronize(h->stream));
    });
}

int tbem_condense_sparse(tbem_constraints_t row_cm, tbem_constraints_t col_cm, tbem_sparse_t matrix,
                         tbem_sparse_t* out) {
    return cs_guard([&] {
        TBEM_REQUIRE(row_cm && col_cm && matrix && out, TBEM_E_INVALID, "null argument");
        *out = nullptr;
        size_t n_rows, n_cols;
        const std::vector<double>* vals;
        const std::vector<long>* rp;
        const std::vector<int>* cols;
        sparse_host_view(matrix, &n_rows, &n_cols, &vals, &rp, &cols);
        for (const auto& kv : row_cm->map)
            TBEM_REQUIRE(kv.first < n_rows, TBEM_E_INVALID, "a row constraint lies beyond the matrix");
        for (const auto& kv : col_cm->map)
            TBEM_REQUIRE(kv.first < n_cols, TBEM_E_INVALID, "a column constraint lies beyond the matrix");
        // rows descending, each row's entries descending (constraint_matrix.cpp:262-281)
        std::map<size_t, double> dok;
        for (size_t r = n_rows; r-- > 0;)
            for (long k = (*rp)[r + 1]; k-- > (*rp)[r];)
                add_entry(row_cm->map, col_cm->map, n_cols, dok, r, (size_t)(*cols)[k], (*vals)[k]);
        // renumber onto the free rows and columns, in key order (constraint_matrix.cpp:215-241)
        std::vector<long> row_map(n_rows, -1), col_map(n_cols, -1);
        long nr = 0, nc = 0;
        for (size_t r = 0; r < n_rows; r++)
            if (!row_cm->map.count(r)) row_map[r] = nr++;
        for (size_t c = 0; c < n_cols; c++)
            if (!col_cm->map.count(c)) col_map[c] = nc++;
        std::vector<long> out_rp(nr + 1, 0);
        std::vector<int> out_cols;
        std::vector<double> out_vals;
        out_cols.reserve(dok.size());
        out_vals.reserve(dok.size());
        // keys ascend as (row, column), so the renumbered entries are already in CSR order
        for (const auto& kv : dok) {
            const size_t c = kv.first % n_cols, r = kv.first / n_cols;
            out_rp[row_map[r] + 1]++;
            out_cols.push_back((int)col_map[c]);
            out_vals.push_back(kv.second);
        }
        for (long r = 0; r < nr; r++) out_rp[r + 1] += out_rp[r];
        *out = sparse_from_host((size_t)nr, (size_t)nc, std::move(out_vals), std::move(out_rp), std::move(out_cols));
    });
}

void tbem_constraints_destroy(tbem_constraints_t h) {
    if (!h) return;
    if (h->done) (void)hipEventSynchronize(h->done);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    delete h;
}

}  // extern "C"
