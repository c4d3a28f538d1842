// FMM plan + apply orchestration and the C ABI (include/tbem_fmm.h).
//
// Reference: FMMOperator<3,R,C> (cpp/fmm.h:131-237, cpp/fmm.cpp:71-109, 444-558).
// The plan holds both octrees (breadth-first cells, points permuted so each cell is
// a contiguous range), the per-level C2E matrices, and every task list turned into
// a CSR list keyed by the cell that accumulates. apply() then runs
//   gather -> P2M -> M2M (deepest level first) -> M2L, P2L -> L2L (root first)
//   -> P2P + M2P + L2P fused per observation leaf -> scatter
// on the handle's stream. Every stage owns its outputs (no atomics), so the FP64
// result is deterministic. The reference order (P2P before M2L etc.,
// cpp/fmm.cpp:459-545) only changes how the same terms are summed.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstddef>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <type_traits>

#include "c2e.h"
#include "m2l_mfma.h"
#include "stages.h"

namespace tbem {

static thread_local std::string g_last_error;
void set_last_error(const std::string& m) { g_last_error = m; }

// ----------------------------------------------------------------------- the plan
// Default bytes the class M2L's per-pair output may take (one buffer per plan); more
// pairs are processed in slices that fit.
constexpr size_t kM2lOutputBudget = size_t(6) << 30;

// The point-target stage's leaf positions split between the one-wave kernel (leaves of
// at most kWaveLeafMax points: one work item each, or chunks of a large one) and the
// one-workgroup kernel (the rest), both ordered by decreasing work (split_leaves).
struct LeafSplit {
    DevBuf<WaveItem> items;
    DevBuf<WaveLeafParts> parts;
    DevBuf<int> block;
    int n_items = 0, n_parts = 0, n_block = 0;  // entries of items, parts, block (with padding)
    int n_work = 0, n_block_leaves = 0;         // work items and workgroup leaves without it
    size_t part_size = 0;  // doubles of chunk sums between the kernels (Work::wpart)
};

struct FmmPlan {
    int kid = 0, R = 1, precision = TBEM_FP64;
    KConst kc{};
    size_t n_obs = 0, n_src = 0;
    double mac = 0, d = 0;
    size_t order = 0, leaf = 0;
    bool small_cells = false;
    LeafSplit leaf_split;  // of the observation leaves (forward)
    // Symmetric point pairs (Kelvin U or the Laplace single layer, one shared tree;
    // forward and transpose): the pairs of neighbouring leaves of kSymLeafMin ..
    // kSymLeafMax points, listed both ways, are evaluated once for both directions
    // (k_leaf_sym, in the precision's pair arithmetic), by the leaf that comes first in
    // point order. ord: leaf_p2p without them, with its own split; own: per leaf
    // position, the ranges of the leaves it evaluates for both; base: its first slot
    // of the reverse sums; g_*: per receiving position, the slots addressed to it.
    struct SymP2P {
        bool on = false;
        DevList ord, own;
        LeafSplit split;
        DevBuf<int> base, order, g_pos, g_ptr, g_off;
        int n_order = 0, n_g = 0, n_leaves = 0;
        size_t n_rev = 0;
        double pairs = 0;  // unordered point pairs
        std::vector<int> part_ptr, part;  // per position, its partner leaves (sorted): the
                                          // adjoint's lists leave them out too
    } sym;
    Surfaces surf;
    int n_e = 0, blk = 0;
    HostTree st, ot;
    DevTree dst, dot;
    DevPoints sp, op;
    DevBuf<double> d_unit, d_inner, d_outer;
    DevBuf<double> uc2e, dc2e;
    std::vector<double> h_uc2e, h_dc2e;
    TaskLists tl;
    // forward lists
    DevList p2m, m2l, p2l, leaf_p2p, leaf_m2p;
    // every M2L pair's surfaces lie farther apart than the kernels' r2 < 1e-12
    // cutoff (m2l_clear_of_cutoff), so the FP64 M2L runs without the test
    bool m2l_no_cutoff = false;
    std::vector<DevList> m2m;    // by source level: parents at L, entries = children
    std::vector<DevList> l2l;    // by observation level L >= 1: cells at L, entry = parent
    DevBuf<int> obs_leaves, src_leaves;
    // adjoint lists: built on the first adjoint call (adjoint_lists), since a
    // forward-only caller never needs them and they cost a second pass over every
    // task list (0.9 s of a 4M-point plan)
    struct AdjLists {
        DevList l2p;                 // obs leaves, entry = itself
        std::vector<DevList> l2l;    // by obs level: parents at L, entries = children
        DevList m2l, m2p;            // items = source cells
        std::vector<DevList> m2m;    // by src level L >= 1: cells at L, entry = parent
        DevList leaf_p2p, leaf_p2l;  // indexed by source leaf position
        LeafSplit leaf_split;        // of the source leaves
        DevList sym_ord;             // leaf_p2p without the symmetric pairs (FmmPlan::sym)
        LeafSplit sym_split;
        std::vector<CellTiles> t_slev;  // every cell of each source level
    };
    mutable std::mutex adj_mu;
    mutable std::unique_ptr<AdjLists> adj;
    const AdjLists& adjoint_lists(hipStream_t s) const;
    M2lMfma m2l_mfma;            // M2L as grouped MFMA GEMMs: FP32 modal (mixed) or FP64 (fp64)
    // M2M (by source level: parents at L) and L2L (by observation level: children at L)
    // on per-class FP64 operators (8 octant classes per level), where enabled
    std::vector<M2lMfma> m2m_ops, l2l_ops;
    size_t tr_y = 0;             // their largest per-pair output, in doubles (a per-handle buffer)
    // GEMM tiles of the C2E products: the cells each stage writes
    CellTiles t_p2m, t_m2l, t_p2l;
    std::vector<CellTiles> t_m2m;          // by source level: the M2M parents
    std::vector<CellTiles> t_olev;         // every cell of each observation level
    double p2p_pairs = 0;
    double plan_seconds = 0;
    // serialization: the inputs' hash, the mixed path's C2E SVD factors (host copies),
    // and whether set_obs_locs has moved the observation points since
    uint64_t input_hash = 0;
    bool same_tree = false;        // one octree serves both sides
    // tree layout: kTreesOwnBoxes (the reference's: each tree on its points' tight
    // box), kTreeShared (one tree for both sides), or kTreesCommonBox (mixed precision
    // with distinct point sets: both trees on the union box, so that their cells lie
    // on one grid and the MFMA M2L's translation classes exist)
    int tree_layout = 0;
    std::vector<double> h_fac[4];  // up W, up V, down W, down V
    bool obs_moved = false;
    // Moving the observation points (set_obs_locs): every apply of any handle that
    // shares the plan holds pts_mu shared while it enqueues; set_obs_locs holds it
    // alone, waits for the device, checks the moved points against their frozen
    // leaf cells and overwrites the sorted coordinates in place.
    mutable std::shared_mutex pts_mu;
    uint64_t obs_version = 0;       // bumped by every set_obs_locs
    DevBuf<int> d_flag;             // set_obs_locs: count of points outside their leaf cell
    // The MFMA M2L's per-pair output, one buffer for all handles that share the
    // plan: the M2L sections of their calls are ordered on the device by y_done
    // and enqueued one at a time under y_mu.
    mutable std::mutex y_mu;
    DevBuf<float> m2l_y;
    DevBuf<double> m2l_y64;         // the FP64 path's class M2L output
    hipEvent_t y_done = nullptr;
    FmmPlan() = default;
    FmmPlan(const FmmPlan&) = delete;
    FmmPlan& operator=(const FmmPlan&) = delete;
    ~FmmPlan() {
        if (y_done) (void)hipEventDestroy(y_done);
    }
};

// Holds the plan's shared MFMA M2L output buffer for one section of a call on
// stream s: waits (on the device) for the previous section that used it, on
// whichever stream, and marks the end of this one.
class YLease {
  public:
    YLease(const FmmPlan& P, hipStream_t s) : P_(P), s_(s), lk_(P.y_mu) {
        TBEM_HIP(hipStreamWaitEvent(s_, P_.y_done, 0));
    }
    ~YLease() { (void)hipEventRecord(P_.y_done, s_); }
    YLease(const YLease&) = delete;
    YLease& operator=(const YLease&) = delete;

  private:
    const FmmPlan& P_;
    hipStream_t s_;
    std::unique_lock<std::mutex> lk_;
};

struct Work {
    hipStream_t stream = nullptr;
    DevBuf<double> dens, mult, locals, out_sorted, xbuf, ybuf;
    DevBuf<double> gsort, lbar, cbar, mbar, ebar, grad_sorted;
    DevBuf<float> chk_up;   // modal up-check potential per source cell (input of the MFMA M2L)
    DevBuf<float> m2l_w;    // transpose: modal input per observation cell
    DevBuf<double> e2;      // transpose: modal M2L^T contribution to ebar per source cell
    DevBuf<double> fld;     // raw check potentials of one stage, blk per cell (input of the C2E GEMMs)
    DevBuf<double> ty;      // per-pair output of the M2M / L2L class operators
    DevBuf<double> wpart;   // chunk sums of the one-wave leaf kernel (LeafSplit::part_size)
    DevBuf<double> wrev;    // reverse sums of the symmetric pairs (R rows of FmmPlan::sym.n_rev)
    hipEvent_t ev[12];  // 0..7 stage boundaries; 8, 9 around the M2L translation kernel; 10, 11
                        // around the point-pair (P2P) launch on the side stream
    hipEvent_t done = nullptr;  // end of the last call's work, on whichever stream it ran
    bool timed = false;
    // The forward's P2P needs only the gathered densities, so it can run on a side
    // stream beside the far field (P2M ... L2L: MFMA and small launches beside the
    // P2P's VALU work), joined before the surface sources (M2P, L2P) are added. On by
    // default where the far field is large (R * n_e >= 48: 2M points, order 100, FP64
    // 30.0 -> 27.9 ms); at low order the extra surface launch costs more than the
    // overlap gains (1M points, order 6: 4.8 -> 5.1 ms).
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    bool overlap = true;
    bool split_last = false;  // the last forward ran split (stage times add both leaf launches)
    // far-field state: dens (of x), mult and locals hold the far field of the call
    // that set far_token (a forward or a recomputing gradient; 0: none). The adjoint
    // leaves mult and locals alone but overwrites dens.
    uint64_t far_token = 0, next_token = 0;
    explicit Work(const FmmPlan& P) {
        TBEM_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        for (auto& e: ev) TBEM_HIP(hipEventCreate(&e));
        TBEM_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
        TBEM_HIP(hipEventRecord(done, stream));
        TBEM_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        overlap = P.blk >= 48;
        TBEM_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        TBEM_HIP(hipEventCreateWithFlags(&join, hipEventDisableTiming));
        dens.alloc(P.R * P.n_src);
        mult.alloc((size_t)P.dst.n_cells * P.blk);
        locals.alloc((size_t)P.dot.n_cells * P.blk);
        out_sorted.alloc(P.R * P.n_obs);
        fld.alloc((size_t)std::max(P.dst.n_cells, P.dot.n_cells) * P.blk);
        ty.alloc(P.tr_y);
        wpart.alloc(std::max(P.leaf_split.part_size, P.sym.split.part_size));
        if (P.sym.on) wrev.alloc((size_t)P.R * P.sym.n_rev);
        if (P.m2l_mfma.enabled && !P.m2l_mfma.fp64) {
            chk_up.alloc((size_t)P.dst.n_cells * P.blk);
            chk_up.zero(stream);  // cells without a P2M/M2M stage keep a zero potential
            TBEM_HIP(hipStreamSynchronize(stream));
        }
    }
    // The work buffers belong to the handle: a call on stream s first waits for
    // the previous call (possibly on another stream) and marks its own end.
    void begin(hipStream_t s) { TBEM_HIP(hipStreamWaitEvent(s, done, 0)); }
    void end(hipStream_t s) { TBEM_HIP(hipEventRecord(done, s)); }
    ~Work() {
        if (done) (void)hipEventDestroy(done);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        for (auto& e: ev) (void)hipEventDestroy(e);
        if (side) (void)hipStreamDestroy(side);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

}  // namespace tbem

struct tbem_fmm_s {
    std::shared_ptr<tbem::FmmPlan> plan;
    std::unique_ptr<tbem::Work> work;
    std::mutex mu;
};

namespace tbem {

// ------------------------------------------------------------------ list builders
// items = cells that accumulate; pairs (a, b): the item is pair[col], the entry pair[1-col].
// Entries of an item keep the pairs' order. Host threads take contiguous chunks of
// the pairs: per-chunk counts, then each chunk fills its slots after the earlier
// chunks' (the same result as one pass in order).
static HostList list_from_pairs(const std::vector<int>& items, int n_cells_item_tree,
                                const std::vector<std::array<int, 2>>& pairs, int col,
                                const HostTree* pts_tree) {
    HostList l;
    l.items = items;
    const size_t ni = items.size(), np = pairs.size();
    std::vector<int> pos(n_cells_item_tree, -1);
    for (size_t k = 0; k < ni; k++) pos[items[k]] = (int)k;
    const size_t nch = np < (size_t(1) << 20) ? 1 : (size_t)plan_threads();
    const size_t per = (np + nch - 1) / std::max<size_t>(nch, 1);
    std::vector<std::vector<int>> cnt(nch, std::vector<int>(ni, 0));
    parallel_for(nch, [&](size_t c) {
        for (size_t i = c * per; i < std::min(np, (c + 1) * per); i++) {
            const int k = pos[pairs[i][col]];
            if (k >= 0) cnt[c][k]++;
        }
    });
    l.ptr.assign(ni + 1, 0);
    for (size_t k = 0; k < ni; k++) {
        int run = l.ptr[k];
        for (size_t c = 0; c < nch; c++) {
            const int v = cnt[c][k];
            cnt[c][k] = run;  // chunk c's first slot of item k
            run += v;
        }
        l.ptr[k + 1] = run;
    }
    l.cell.resize(l.ptr.back());
    parallel_for(nch, [&](size_t c) {
        std::vector<int>& fill = cnt[c];
        for (size_t i = c * per; i < std::min(np, (c + 1) * per); i++) {
            const int k = pos[pairs[i][col]];
            if (k >= 0) l.cell[fill[k]++] = pairs[i][1 - col];
        }
    });
    if (pts_tree) cells_to_ranges(l, *pts_tree);
    return l;
}

static std::vector<int> cells_with(const std::vector<std::array<int, 2>>& pairs, int col, int n_cells) {
    std::vector<char> m(n_cells, 0);
    for (auto& p: pairs) m[p[col]] = 1;
    std::vector<int> out;
    for (int c = 0; c < n_cells; c++)
        if (m[c]) out.push_back(c);
    return out;
}

// children lists of the non-leaf cells at level L
static HostList children_list(const HostTree& t, int L) {
    HostList l;
    l.ptr.assign(1, 0);
    for (int c = t.level_begin[L]; c < t.level_begin[L + 1]; c++) {
        if (t.leaf[c]) continue;
        l.items.push_back(c);
        for (int k = 0; k < t.n_child[c]; k++) l.cell.push_back(t.first_child[c] + k);
        l.ptr.push_back((int)l.cell.size());
    }
    return l;
}

// parent lists of all cells at level L >= 1
static HostList parent_list(const HostTree& t, int L) {
    HostList l;
    l.ptr.assign(1, 0);
    for (int c = t.level_begin[L]; c < t.level_begin[L + 1]; c++) {
        l.items.push_back(c);
        l.cell.push_back(t.parent[c]);
        l.ptr.push_back((int)l.cell.size());
    }
    return l;
}

static HostList self_list(const HostTree& t, const std::vector<int>& cells) {
    HostList l;
    l.items = cells;
    l.ptr.assign(1, 0);
    for (int c: cells) {
        l.cell.push_back(c);
        l.ptr.push_back((int)l.cell.size());
    }
    cells_to_ranges(l, t);
    return l;
}

// Whether every M2L pair (t, s) keeps its surfaces (the inner ones: target down-check
// and source up-equivalent, radius rin * level_r about each center) farther apart
// than the kernels' 1e-6 cutoff radius with a wide margin:
//   |c_t - c_s| - rin (r_t + r_s) > max(1e-5, 1e-9 * |c|max).
// The margin covers the rounding of the FP64 displacements (a few ulp of the
// coordinates), so that no computed r2 can fall below 1e-12 and the test in the
// kernels never fires: leaving it out changes no result.
static bool m2l_clear_of_cutoff(const HostList& l, const HostTree& ot, const HostTree& st, const Surfaces& surf) {
    double rin = 0;
    for (int i = 0; i < surf.n_e; i++) {
        const double* u = &surf.inner[3 * i];
        rin = std::max(rin, std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
    }
    double cmax = 0;
    for (const HostTree* t : {&ot, &st})
        for (int c = 0; c < t->n_cells; c++)
            cmax = std::max({cmax, std::fabs(t->cx[c]), std::fabs(t->cy[c]), std::fabs(t->cz[c])});
    const double need = std::max(1e-5, 1e-9 * cmax);
    const int n = (int)l.items.size();
    std::atomic<int> bad{0};
    parallel_for((size_t)n, [&](size_t k) {
        const int t = l.items[k];
        const double rt = rin * ot.level_r[ot.level[t]];
        for (int p = l.ptr[k]; p < l.ptr[k + 1]; p++) {
            const int c = l.cell[p];
            const double dx = ot.cx[t] - st.cx[c], dy = ot.cy[t] - st.cy[c], dz = ot.cz[t] - st.cz[c];
            const double gap = std::sqrt(dx * dx + dy * dy + dz * dz) - rt - rin * st.level_r[st.level[c]];
            if (!(gap > need)) bad.store(1, std::memory_order_relaxed);
        }
    });
    return bad.load() == 0;
}

static void validate(const tbem_fmm_desc* d) {
    TBEM_REQUIRE(d != nullptr, TBEM_E_INVALID, "null descriptor");
    TBEM_REQUIRE(kernel_valid(d->kernel), TBEM_E_INVALID, "unknown kernel id");
    TBEM_REQUIRE(d->order > 0, TBEM_E_INVALID, "FMMConfig.order must be > 0");
    TBEM_REQUIRE(d->min_pts_per_cell > 0, TBEM_E_INVALID, "FMMConfig.min_pts_per_cell must be > 0");
    TBEM_REQUIRE(d->precision == TBEM_FP64 || d->precision == TBEM_MIXED, TBEM_E_INVALID,
                 "precision must be TBEM_FP64 or TBEM_MIXED");
    TBEM_REQUIRE(d->tree_boxes >= TBEM_TREES_AUTO && d->tree_boxes <= TBEM_TREES_COMMON_BOX, TBEM_E_INVALID,
                 "tree_boxes must be TBEM_TREES_AUTO, _OWN_BOXES or _COMMON_BOX");
    TBEM_REQUIRE(d->translations == TBEM_TRANSLATIONS_AUTO || d->translations == TBEM_TRANSLATIONS_SURFACES,
                 TBEM_E_INVALID, "translations must be TBEM_TRANSLATIONS_AUTO or TBEM_TRANSLATIONS_SURFACES");
    TBEM_REQUIRE(d->n_obs < (1u << 30) && d->n_src < (1u << 30), TBEM_E_UNSUPPORTED,
                 "more than 2^30 points per side");
    TBEM_REQUIRE(d->n_obs == 0 || (d->obs_locs && d->obs_normals), TBEM_E_INVALID, "null observation arrays");
    TBEM_REQUIRE(d->n_src == 0 || (d->src_locs && d->src_normals && d->src_weights), TBEM_E_INVALID,
                 "null source arrays");
}

// ------------------------------------------------------------- plan serialization
// A saved plan: both octrees, the task lists, the per-level C2E matrices and, for
// the mixed path, their SVD factors, under a header that names the inputs. Layout
// (little endian): the 8 bytes "TBEMPLN1"; u64 payload length; u64 hash64 of the
// payload; then the payload: the header fields (kernel, mu, nu, n_obs, n_src, mac,
// order, min_pts_per_cell, d, account_for_small_cells, tree layout, input hash) and
// the sections, each array as a u64 count followed by its elements.
constexpr int kTreesOwnBoxes = 0, kTreeShared = 1, kTreesCommonBox = 2;

struct SavedPlan {
    bool same = false;
    int layout = kTreesOwnBoxes;
    HostTree st, ot;
    std::vector<double> uc2e, dc2e, fac[4];
    TaskLists tl;
};

// A 64-bit integrity hash, four independent 8-byte lanes per 32-byte step (several
// GB/s): each word is multiplied, rotated and folded into its lane; the lanes, the
// tail words and the tail bytes are then combined. It guards against corrupt files
// and mismatched inputs, not against deliberate collisions.
static inline uint64_t mix_word(uint64_t h, uint64_t w) {
    w *= 0xC2B2AE3D27D4EB4Full;
    w = (w << 31) | (w >> 33);
    const uint64_t x = h ^ (w * 0x9E3779B185EBCA87ull);
    return ((x << 27) | (x >> 37)) * 5 + 0x52DCE729;
}

static uint64_t hash64(const void* p, size_t n, uint64_t h = 0x9E3779B97F4A7C15ull) {
    const unsigned char* b = (const unsigned char*)p;
    uint64_t lane[4] = {h, h + 1, h + 2, h + 3};
    size_t i = 0;
    for (; i + 32 <= n; i += 32)
        for (int k = 0; k < 4; k++) {
            uint64_t w;
            std::memcpy(&w, b + i + 8 * k, 8);
            lane[k] = mix_word(lane[k], w);
        }
    h = mix_word(mix_word(mix_word(mix_word(h ^ n, lane[0]), lane[1]), lane[2]), lane[3]);
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        std::memcpy(&w, b + i, 8);
        h = mix_word(h, w);
    }
    for (; i < n; i++) h = (h ^ b[i]) * 0x100000001B3ull;
    h ^= h >> 33;
    h *= 0xFF51AFD7ED558CCDull;
    h ^= h >> 33;
    return h;
}

// The hash of what the plan depends on: counts and every input array.
static uint64_t input_hash(const tbem_fmm_desc* d) {
    uint64_t h = hash64(&d->n_obs, sizeof(d->n_obs));
    h = hash64(&d->n_src, sizeof(d->n_src), h);
    if (d->n_obs) {
        h = hash64(d->obs_locs, 3 * d->n_obs * sizeof(double), h);
        h = hash64(d->obs_normals, 3 * d->n_obs * sizeof(double), h);
    }
    if (d->n_src) {
        h = hash64(d->src_locs, 3 * d->n_src * sizeof(double), h);
        h = hash64(d->src_normals, 3 * d->n_src * sizeof(double), h);
        h = hash64(d->src_weights, d->n_src * sizeof(double), h);
    }
    return h;
}

struct PlanWriter {
    std::vector<char> b;
    template <class T> void pod(const T& v) {
        const char* p = (const char*)&v;
        b.insert(b.end(), p, p + sizeof(T));
    }
    template <class T> void vec(const std::vector<T>& v) {
        pod((uint64_t)v.size());
        const char* p = (const char*)v.data();
        b.insert(b.end(), p, p + v.size() * sizeof(T));
    }
};

struct PlanReader {
    const char* p;
    size_t n, off;
    template <class T> T pod() {
        TBEM_REQUIRE(sizeof(T) <= n - off, TBEM_E_INVALID, "saved plan: truncated");
        T v;
        std::memcpy(&v, p + off, sizeof(T));
        off += sizeof(T);
        return v;
    }
    template <class T> std::vector<T> vec() {
        const uint64_t k = pod<uint64_t>();
        TBEM_REQUIRE(k <= (n - off) / sizeof(T), TBEM_E_INVALID, "saved plan: truncated");
        std::vector<T> v(k);
        if (k) std::memcpy((void*)v.data(), p + off, k * sizeof(T));
        off += k * sizeof(T);
        return v;
    }
};

static void put_tree(PlanWriter& w, const HostTree& t) {
    w.pod(t.n_pts);
    w.pod(t.n_cells);
    w.pod(t.n_levels);
    for (const auto* v : {&t.level, &t.begin, &t.count, &t.first_child, &t.n_child, &t.parent, &t.octant})
        w.vec(*v);
    w.vec(t.leaf);
    for (const auto* v : {&t.cx, &t.cy, &t.cz, &t.level_hw, &t.level_r}) w.vec(*v);
    for (const auto* v : {&t.level_begin, &t.perm, &t.leaves, &t.dfs}) w.vec(*v);
}

// A tree read back is checked so that every index the kernels follow stays inside
// its arrays: a corrupt or foreign file is refused instead of being run.
static HostTree get_tree(PlanReader& r, int n_pts) {
    HostTree t;
    t.n_pts = r.pod<int>();
    t.n_cells = r.pod<int>();
    t.n_levels = r.pod<int>();
    for (auto* v : {&t.level, &t.begin, &t.count, &t.first_child, &t.n_child, &t.parent, &t.octant})
        *v = r.vec<int>();
    t.leaf = r.vec<char>();
    for (auto* v : {&t.cx, &t.cy, &t.cz, &t.level_hw, &t.level_r}) *v = r.vec<double>();
    for (auto* v : {&t.level_begin, &t.perm, &t.leaves, &t.dfs}) *v = r.vec<int>();
    bool ok = t.n_pts == n_pts && t.n_cells >= 0 && t.n_levels >= 0 && t.n_levels <= kMaxKeyLevels + 1;
    const size_t nc = ok ? (size_t)t.n_cells : 0, nl = ok ? (size_t)t.n_levels : 0;
    for (const auto* v : {&t.level, &t.begin, &t.count, &t.first_child, &t.n_child, &t.parent, &t.octant})
        ok = ok && v->size() == nc;
    ok = ok && t.leaf.size() == nc && t.cx.size() == nc && t.cy.size() == nc && t.cz.size() == nc;
    ok = ok && t.level_hw.size() == 3 * nl && t.level_r.size() == nl && t.level_begin.size() == nl + 1;
    ok = ok && t.perm.size() == (size_t)n_pts && t.leaves.size() <= nc && t.dfs.size() == nc;
    TBEM_REQUIRE(ok, TBEM_E_INVALID, "saved plan: octree arrays do not fit together");
    ok = t.level_begin[0] == 0 && t.level_begin[nl] == t.n_cells;
    for (size_t L = 0; ok && L < nl; L++) ok = t.level_begin[L] <= t.level_begin[L + 1];
    for (size_t c = 0; ok && c < nc; c++) {
        ok = t.level[c] >= 0 && t.level[c] < t.n_levels && t.begin[c] >= 0 && t.count[c] >= 0 &&
             (long)t.begin[c] + t.count[c] <= n_pts && t.n_child[c] >= 0 && t.n_child[c] <= 8 &&
             (t.n_child[c] == 0 || (t.first_child[c] >= 0 && (long)t.first_child[c] + t.n_child[c] <= t.n_cells)) &&
             t.parent[c] >= -1 && t.parent[c] < t.n_cells && (long)c >= t.level_begin[t.level[c]] &&
             (long)c < t.level_begin[t.level[c] + 1] && (t.leaf[c] == 0 || t.leaf[c] == 1);
    }
    std::vector<char> seen(n_pts, 0);
    for (int i : t.perm) {
        ok = ok && i >= 0 && i < n_pts && !seen[i];
        if (ok) seen[i] = 1;
    }
    for (const auto* v : {&t.leaves, &t.dfs})
        for (int c : *v) ok = ok && c >= 0 && c < t.n_cells;
    for (int c : t.leaves) ok = ok && t.leaf[c];
    TBEM_REQUIRE(ok, TBEM_E_INVALID, "saved plan: octree indices out of range");
    // structure: one root; every other cell's parent lies one level up and lists it
    // among its children; children lie one level down (the M2M/L2L stages and
    // their transposes follow these links)
    ok = nc == 0 || (nl >= 1 && t.level_begin[1] == 1);
    for (size_t c = 0; ok && c < nc; c++) {
        const int p = t.parent[c];
        if (t.level[c] == 0) {
            ok = p == -1;
        } else {
            ok = p >= 0 && t.level[p] == t.level[c] - 1 && t.first_child[p] >= 0 &&
                 (long)c >= t.first_child[p] && (long)c < (long)t.first_child[p] + t.n_child[p];
        }
        for (int k = 0; ok && k < t.n_child[c]; k++) ok = t.parent[t.first_child[c] + k] == (int)c;
        ok = ok && (t.leaf[c] ? t.n_child[c] == 0 : t.n_child[c] > 0);
    }
    TBEM_REQUIRE(ok, TBEM_E_INVALID, "saved plan: octree parent and child links disagree");
    return t;
}

static std::vector<std::array<int, 2>> read_pairs(PlanReader& r, int n_obs_cells, int n_src_cells) {
    const uint64_t k = r.pod<uint64_t>();
    TBEM_REQUIRE(k % 2 == 0 && k / 2 <= (r.n - r.off) / sizeof(std::array<int, 2>), TBEM_E_INVALID,
                 "saved plan: truncated");
    std::vector<std::array<int, 2>> v(k / 2);
    if (k) std::memcpy((void*)v.data(), r.p + r.off, k * sizeof(int));
    r.off += k * sizeof(int);
    bool ok = true;
    for (const auto& q : v) ok &= q[0] >= 0 && q[0] < n_obs_cells && q[1] >= 0 && q[1] < n_src_cells;
    TBEM_REQUIRE(ok, TBEM_E_INVALID, "saved plan: task list names a cell outside the trees");
    return v;
}

static std::vector<int> read_cells(PlanReader& r, int n_cells) {
    std::vector<int> v = r.vec<int>();
    for (int c : v)
        TBEM_REQUIRE(c >= 0 && c < n_cells, TBEM_E_INVALID, "saved plan: task list names a cell outside the trees");
    return v;
}

static std::vector<char> save_plan(const FmmPlan& P) {
    TBEM_REQUIRE(!P.obs_moved, TBEM_E_UNSUPPORTED,
                 "the observation points have moved since the plan was built: save a freshly built plan");
    PlanWriter w;
    w.pod(P.kid);
    w.pod(P.kc.mu);
    w.pod(P.kc.nu);
    w.pod((uint64_t)P.n_obs);
    w.pod((uint64_t)P.n_src);
    w.pod(P.mac);
    w.pod((uint64_t)P.order);
    w.pod((uint64_t)P.leaf);
    w.pod(P.d);
    w.pod((int)P.small_cells);
    const int layout = P.tree_layout;
    w.pod(layout);
    w.pod(P.input_hash);
    put_tree(w, P.st);
    if (layout != kTreeShared) put_tree(w, P.ot);
    w.vec(P.h_uc2e);
    w.vec(P.h_dc2e);
    for (const auto& f : P.h_fac) w.vec(f);
    for (const auto* v : {&P.tl.p2p, &P.tl.m2l, &P.tl.m2p, &P.tl.p2l}) {
        w.pod((uint64_t)(2 * v->size()));  // as flat (obs cell, src cell) ints
        const char* q = (const char*)v->data();
        w.b.insert(w.b.end(), q, q + v->size() * sizeof((*v)[0]));
    }
    for (const auto* v : {&P.tl.p2m, &P.tl.m2m, &P.tl.l2l, &P.tl.l2p}) w.vec(*v);
    std::vector<char> out(24);
    std::memcpy(out.data(), "TBEMPLN1", 8);
    const uint64_t len = w.b.size(), hash = hash64(w.b.data(), w.b.size());
    std::memcpy(out.data() + 8, &len, 8);
    std::memcpy(out.data() + 16, &hash, 8);
    out.insert(out.end(), w.b.begin(), w.b.end());
    return out;
}

static SavedPlan load_plan(const tbem_fmm_desc* d, const void* buf, size_t n) {
    TBEM_REQUIRE(buf && n >= 24 && std::memcmp(buf, "TBEMPLN1", 8) == 0, TBEM_E_INVALID,
                 "not a saved FMM plan (bad magic)");
    uint64_t len, hash;
    std::memcpy(&len, (const char*)buf + 8, 8);
    std::memcpy(&hash, (const char*)buf + 16, 8);
    TBEM_REQUIRE(len == n - 24, TBEM_E_INVALID, "saved plan: truncated");
    TBEM_REQUIRE(hash64((const char*)buf + 24, len) == hash, TBEM_E_INVALID, "saved plan: checksum mismatch");
    PlanReader r{(const char*)buf + 24, (size_t)len, 0};
    const int kid = r.pod<int>();
    const double mu = r.pod<double>(), nu = r.pod<double>();
    const uint64_t n_obs = r.pod<uint64_t>(), n_src = r.pod<uint64_t>();
    const double mac = r.pod<double>();
    const uint64_t order = r.pod<uint64_t>(), leaf = r.pod<uint64_t>();
    const double dd = r.pod<double>();
    const int small = r.pod<int>(), layout = r.pod<int>();
    const uint64_t ih = r.pod<uint64_t>();
    const KConst kc = make_kconst(d->shear_modulus, d->poisson_ratio);
    TBEM_REQUIRE(kid == d->kernel && mu == kc.mu && nu == kc.nu && n_obs == d->n_obs && n_src == d->n_src &&
                     mac == d->mac && order == d->order && leaf == d->min_pts_per_cell && dd == d->d &&
                     small == (d->account_for_small_cells != 0),
                 TBEM_E_INVALID, "saved plan: built for another kernel or FMMConfig");
    TBEM_REQUIRE(ih == input_hash(d), TBEM_E_INVALID, "saved plan: built for other points, normals or weights");
    TBEM_REQUIRE(layout == kTreesOwnBoxes || layout == kTreeShared || layout == kTreesCommonBox, TBEM_E_INVALID,
                 "saved plan: unknown tree layout");
    TBEM_REQUIRE(layout != kTreeShared || n_obs == n_src, TBEM_E_INVALID,
                 "saved plan: shared tree with different counts");
    SavedPlan sv;
    sv.same = layout == kTreeShared;
    sv.layout = layout;
    sv.st = get_tree(r, (int)n_src);
    sv.ot = sv.same ? sv.st : get_tree(r, (int)n_obs);
    sv.uc2e = r.vec<double>();
    sv.dc2e = r.vec<double>();
    for (auto& f : sv.fac) f = r.vec<double>();
    for (auto* v : {&sv.tl.p2p, &sv.tl.m2l, &sv.tl.m2p, &sv.tl.p2l}) *v = read_pairs(r, sv.ot.n_cells, sv.st.n_cells);
    sv.tl.p2m = read_cells(r, sv.st.n_cells);
    sv.tl.m2m = read_cells(r, sv.st.n_cells);
    sv.tl.l2l = read_cells(r, sv.ot.n_cells);
    sv.tl.l2p = read_cells(r, sv.ot.n_cells);
    TBEM_REQUIRE(r.off == r.n, TBEM_E_INVALID, "saved plan: trailing bytes");
    const size_t blk = (size_t)kernel_dim(d->kernel) * make_surfaces(d->order, d->d).n_e, bb = blk * blk;
    TBEM_REQUIRE(sv.uc2e.size() == bb * sv.st.n_levels && sv.dc2e.size() == bb * sv.ot.n_levels, TBEM_E_INVALID,
                 "saved plan: C2E matrices of the wrong size");
    for (int k = 0; k < 4; k++)
        TBEM_REQUIRE(sv.fac[k].empty() || sv.fac[k].size() == bb * (k < 2 ? sv.st.n_levels : sv.ot.n_levels),
                     TBEM_E_INVALID, "saved plan: C2E factors of the wrong size");
    TBEM_REQUIRE(sv.fac[0].empty() == sv.fac[1].empty() && sv.fac[2].empty() == sv.fac[3].empty(), TBEM_E_INVALID,
                 "saved plan: incomplete C2E factors");
    return sv;
}

// Phase times of a plan build, printed to stderr when TBEM_PLAN_TIMING is set.
struct PhaseClock {
    bool on;
    hipStream_t s;
    std::chrono::steady_clock::time_point t;
    std::string line;
    explicit PhaseClock(hipStream_t st) : on(std::getenv("TBEM_PLAN_TIMING") != nullptr), s(st),
                                          t(std::chrono::steady_clock::now()) {}
    void mark(const char* what) {
        if (!on) return;
        (void)hipStreamSynchronize(s);
        const auto now = std::chrono::steady_clock::now();
        line += std::string(" ") + what + " " +
                std::to_string(std::chrono::duration<double, std::milli>(now - t).count()).substr(0, 7) + " ms;";
        t = now;
    }
    ~PhaseClock() {
        if (on) std::fprintf(stderr, "[tbem plan]%s\n", line.c_str());
    }
};

// The point-target stage leaves the r2 < 1e-12 cutoff test out of source tiles whose
// ranges lie farther than this from the leaf's points (same result; see
// launch_range_close). The one-wave kernel checks a tile with one ballot; the
// 512-thread one with a ballot per wave into LDS (measured with flags on every leaf
// against none: 4M-point Laplace, leaf 100, FP64 leaf stage 40.8 -> 38.1 ms; 1M-point
// elastic, leaf 64, and 2M points, leaf 256: within 1%). kCloseMargin (stages.h).

// l: a point-range list indexed by the leaves `leaves` of tree tt (points tp), over
// the source points sp: (re)computes its per-range close flags.
static void close_flags(DevList& l, const int* leaves, const DevTree& tt, const DevPoints& tp, const DevPoints& sp,
                        hipStream_t s) {
    const size_t n = l.rbeg.n;
    if (l.rclose.n != n) l.rclose.alloc(n);
    if (n) launch_range_close(l.n_items, leaves, view(tt), view(tp), l.ptr.p, l.rbeg.p, l.rcnt.p, view(sp),
                              kCloseMargin, l.rclose.p, s);
}

// Splits the positions of a point-target list (items = leaves of tree tt; point
// ranges pts, surface cells surf, and each leaf's own surface): leaves of at most
// kWaveLeafMax points go to the one-wave kernel, one item each when they have at most
// kWavePairsMax pairs (points x sources); larger ones of at most kWaveChunkLeaf points
// are cut into chunks of about equal sources (points cut anywhere, surface cells
// whole, the own surface with the last chunk), whose sums k_leaf_parts adds in chunk
// order; the rest go to the one-workgroup kernel. Both lists are laid out by
// xcd_layout: per XCD a run of neighbouring leaves, the most work first.
// (kWaveChunkLeaf: a wave holds ceil(m/2) targets per source slice, so above 64
// points one slice is left and down to half of its lanes idle, where the 512-thread
// kernel keeps 96% busy.)
constexpr int kWaveChunkLeaf = 64;

// Lays out work items (work, item), given in tree order, for a launch whose workgroup
// b runs on XCD b % 8 (the hardware deals workgroups to the XCDs in turn), g items per
// workgroup. The items are cut into 8 runs of about equal work, one per XCD, so that
// neighbouring leaves, which share most of their sources, meet in one XCD's L2. Within
// a run the items go by work in power-of-two classes, largest first (so that the
// longest start first), tree order within a class; workgroup b takes group b / 8 of
// run b % 8, and the shorter runs are padded with `empty`.
template <class I>
static std::vector<I> xcd_layout(const std::vector<std::pair<long, I>>& items, int g, const I& empty) {
    constexpr int kXcds = 8;
    long total = 0;
    for (const auto& it: items) total += it.first;
    std::vector<std::vector<std::pair<int, I>>> run(kXcds);
    long acc = 0;
    for (const auto& it: items) {
        const int x = total > 0 ? (int)std::min<long>(kXcds - 1, (acc * kXcds) / total) : 0;
        int cls = 0;
        for (long w = it.first; w > 1; w >>= 1) cls++;
        run[x].push_back({cls, it.second});
        acc += it.first;
    }
    size_t groups = 0;
    for (auto& r: run) {
        std::stable_sort(r.begin(), r.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
        groups = std::max(groups, (r.size() + g - 1) / g);
    }
    std::vector<I> out(groups * kXcds * g, empty);
    for (int x = 0; x < kXcds; x++)
        for (size_t i = 0; i < run[x].size(); i++) out[((i / g) * kXcds + x) * g + i % g] = run[x][i].second;
    return out;
}

static void split_leaves(const HostTree& tt, const std::vector<int>& leaves, const HostList& pts,
                         const HostList& surf, int n_e, LeafSplit& out, hipStream_t s) {
    std::vector<std::pair<long, WaveItem>> w;
    std::vector<std::pair<long, int>> b;
    std::vector<WaveLeafParts> parts;
    long part = 0;
    for (size_t k = 0; k < leaves.size(); k++) {
        const int pos = (int)k;
        const long m = tt.count[leaves[k]];
        if (m == 0) continue;  // no targets, nothing to write
        const int p0 = pts.ptr.empty() ? 0 : pts.ptr[k], p1 = pts.ptr.empty() ? 0 : pts.ptr[k + 1];
        const long np = p0 < p1 ? (long)pts.rpre[p1 - 1] + pts.rcnt[p1 - 1] : 0;
        const int c0 = surf.ptr.empty() ? 0 : surf.ptr[k], c1 = surf.ptr.empty() ? 0 : surf.ptr[k + 1];
        const long src = np + (long)n_e * (c1 - c0 + 1);
        const long nc = (m * src + kWavePairsMax - 1) / kWavePairsMax;
        if (m <= kWaveLeafMax && nc <= 1) {
            w.push_back({m * src, WaveItem{pos, p0, p1, 0, (int)np, c0, c1, 1, -1}});
            continue;
        }
        if (m > kWaveChunkLeaf || part + (nc + 1) * kPartRows * m > INT32_MAX) {
            b.push_back({m * src, pos});
            continue;
        }
        const long per = (src + nc - 1) / nc;
        std::vector<std::pair<long, WaveItem>> ch;
        const auto pre = pts.rpre.begin();
        for (long v = 0; v < np; v += per) {
            const int v0 = (int)v, v1 = (int)std::min(np, v + per);
            const int f0 = (int)(std::upper_bound(pre + p0, pre + p1, v0) - pre) - 1;
            const int f1 = (int)(std::lower_bound(pre + p0, pre + p1, v1) - pre);
            ch.push_back({m * (v1 - v0), WaveItem{pos, f0, f1, v0, v1, 0, 0, 0, 0}});
        }
        const int cpc = (int)std::max(1L, per / std::max(1, n_e));
        for (int c = c0; c < c1; c += cpc) {
            const int ce = std::min(c1, c + cpc);
            ch.push_back({m * n_e * (ce - c), WaveItem{pos, 0, 0, 0, 0, c, ce, 0, 0}});
        }
        if (ch.empty()) ch.push_back({0, WaveItem{pos, 0, 0, 0, 0, 0, 0, 0, 0}});
        ch.back().second.self = 1;
        ch.back().first += m * n_e;
        for (size_t j = 0; j < ch.size(); j++) {
            ch[j].second.part = (int)(part + (long)j * kPartRows * m);
            w.push_back(ch[j]);
        }
        parts.push_back(WaveLeafParts{pos, (int)part, (int)ch.size()});
        part += (long)ch.size() * kPartRows * m;
    }
    const std::vector<WaveItem> wi =
        xcd_layout(w, kWaveBlock / 64, WaveItem{-1, 0, 0, 0, 0, 0, 0, 0, -1});
    const std::vector<int> bl = xcd_layout(b, 1, -1);
    out.n_work = (int)w.size();
    out.n_block_leaves = (int)b.size();
    out.n_items = (int)wi.size();
    out.n_parts = (int)parts.size();
    out.n_block = (int)bl.size();
    out.part_size = (size_t)part;
    out.items.upload(wi, s);
    out.parts.upload(parts, s);
    out.block.upload(bl, s);
}

// FmmPlan::sym from the P2P pairs of the shared tree T (forward lists by leaf
// position; lm: the surface lists the split also weighs). A pair (A, B) of leaves
// A != B goes to the symmetric evaluation when both hold lo .. kSymLeafMax points
// (lo below) and the pairs list it both ways; the leaf with the smaller first point owns
// it, and B's points take slots base[A] + (their offset in A's owned stream); B's
// slots are listed by owner position.
static void build_sym(FmmPlan& P, const HostTree& T, const std::vector<std::array<int, 2>>& p2p,
                      const HostList& lm, hipStream_t s, PhaseClock& clk) {
    FmmPlan::SymP2P& Y = P.sym;
    const std::vector<int>& leaves = T.leaves;
    const int nl = (int)leaves.size();
    std::vector<int> pos_of(T.n_cells, -1);
    for (int k = 0; k < nl; k++) pos_of[leaves[k]] = k;
    // Below 64 points a leaf's pairs pay for the Laplace single layer (4M points, cube +
    // clusters, leaf 100: FP64 matvec 71.1 -> 66.7 ms with leaves from 16 points, mixed
    // 47.4 -> 45.6) but not for Kelvin U (1M points, fault + surface, leaf 64: FP64
    // 2.90 -> 3.15 ms, mixed 2.30 -> 2.75; from 32 points 3.09 and 2.51), whose one-wave
    // ordered kernel serves small leaves better.
    const int lo = P.kid == TBEM_K_LAPLACE_SINGLE ? kSymLeafMin : 64;
    auto fits = [&](int k) {
        const int m = T.count[leaves[k]];
        return m >= lo && m <= kSymLeafMax;
    };
    const HostList lc = list_from_pairs(leaves, T.n_cells, p2p, 0, nullptr);
    std::vector<int> srt = lc.cell;  // each position's cells, sorted, for the mirror test
    parallel_for((size_t)nl, [&](size_t k) { std::sort(srt.begin() + lc.ptr[k], srt.begin() + lc.ptr[k + 1]); });
    auto listed = [&](int k, int c) {
        return std::binary_search(srt.begin() + lc.ptr[k], srt.begin() + lc.ptr[k + 1], c);
    };
    // two passes over chunks of positions: mark and count, then fill at the prefix offsets
    constexpr int kChunk = 1024;
    const size_t nch = ((size_t)nl + kChunk - 1) / kChunk;
    std::vector<char> both(lc.cell.size());
    std::vector<int> n_ord(nl), n_own(nl), n_part(nl);
    parallel_for(nch, [&](size_t ch) {
        for (int k = (int)ch * kChunk; k < std::min(nl, (int)(ch + 1) * kChunk); k++) {
            const int ca = leaves[k];
            int o = 0, w = 0, q = 0;
            for (int e = lc.ptr[k]; e < lc.ptr[k + 1]; e++) {
                const int c = lc.cell[e], j = pos_of[c];
                const bool b = c != ca && j >= 0 && fits(k) && fits(j) && listed(j, ca);
                both[e] = b;
                if (!b) o++;
                else {
                    q++;
                    if (T.begin[ca] < T.begin[c]) w++;
                }
            }
            n_ord[k] = o;
            n_own[k] = w;
            n_part[k] = q;
        }
    });
    HostList ord, own;
    ord.items = own.items = leaves;
    ord.ptr.assign(nl + 1, 0);
    own.ptr.assign(nl + 1, 0);
    Y.part_ptr.assign(nl + 1, 0);
    for (int k = 0; k < nl; k++) {
        ord.ptr[k + 1] = ord.ptr[k] + n_ord[k];
        own.ptr[k + 1] = own.ptr[k] + n_own[k];
        Y.part_ptr[k + 1] = Y.part_ptr[k] + n_part[k];
    }
    if (own.ptr[nl] == 0) return;
    ord.cell.resize(ord.ptr[nl]);
    own.cell.resize(own.ptr[nl]);
    Y.part.resize(Y.part_ptr[nl]);
    parallel_for(nch, [&](size_t ch) {
        for (int k = (int)ch * kChunk; k < std::min(nl, (int)(ch + 1) * kChunk); k++) {
            const int ca = leaves[k];
            int o = ord.ptr[k], w = own.ptr[k], q = Y.part_ptr[k];
            for (int e = lc.ptr[k]; e < lc.ptr[k + 1]; e++) {
                const int c = lc.cell[e];
                if (!both[e]) {
                    ord.cell[o++] = c;
                    continue;
                }
                Y.part[q++] = c;
                if (T.begin[ca] < T.begin[c]) own.cell[w++] = c;
            }
            std::sort(Y.part.begin() + Y.part_ptr[k], Y.part.begin() + Y.part_ptr[k + 1]);
        }
    });
    clk.mark("sym: pairs");
    const std::vector<int> own_cells = own.cell, own_ptr = own.ptr;
    cells_to_ranges(ord, T);
    cells_to_ranges(own, T);
    clk.mark("sym: ranges");
    std::vector<int> base(nl, 0);
    std::vector<std::pair<long, int>> work;
    size_t n_rev = 0;
    for (int k = 0; k < nl; k++) {
        base[k] = (int)n_rev;
        const int e0 = own.ptr[k], e1 = own.ptr[k + 1];
        if (e0 == e1) continue;
        const long n = (long)own.rpre[e1 - 1] + own.rcnt[e1 - 1];
        work.push_back({n * T.count[leaves[k]], k});
        Y.pairs += (double)n * T.count[leaves[k]];
        n_rev += (size_t)n;
        TBEM_REQUIRE(n_rev <= (size_t)INT32_MAX, TBEM_E_UNSUPPORTED, "symmetric pairs: too many slots");
    }
    // the slot of each owned leaf's first point (in parallel), then the slots addressed
    // to each position, by owner position (counted, then filled in owner order)
    std::vector<int> slot(own_cells.size());
    parallel_for(nch, [&](size_t ch) {
        for (int k = (int)ch * kChunk; k < std::min(nl, (int)(ch + 1) * kChunk); k++) {
            const int e0 = own.ptr[k], e1 = own.ptr[k + 1];
            for (int q = own_ptr[k]; q < own_ptr[k + 1]; q++) {
                const int c = own_cells[q];
                const int e = (int)(std::upper_bound(own.rbeg.begin() + e0, own.rbeg.begin() + e1, T.begin[c]) -
                                    own.rbeg.begin()) - 1;
                slot[q] = base[k] + own.rpre[e] + (T.begin[c] - own.rbeg[e]);
            }
        }
    });
    std::vector<int> g_cnt(nl, 0);
    for (size_t q = 0; q < own_cells.size(); q++) g_cnt[pos_of[own_cells[q]]]++;
    std::vector<int> g_pos, g_ptr(1, 0), g_at(nl, -1);
    for (int k = 0; k < nl; k++) {
        if (g_cnt[k] == 0) continue;
        g_at[k] = g_ptr.back();
        g_pos.push_back(k);
        g_ptr.push_back(g_ptr.back() + g_cnt[k]);
    }
    std::vector<int> g_off(g_ptr.back());
    for (int k = 0; k < nl; k++)
        for (int q = own_ptr[k]; q < own_ptr[k + 1]; q++) g_off[g_at[pos_of[own_cells[q]]]++] = slot[q];
    const std::vector<int> order = xcd_layout(work, 1, -1);
    clk.mark("sym: slots");
    Y.ord.upload(ord, s);
    Y.own.upload(own, s);
    Y.base.upload(base, s);
    Y.order.upload(order, s);
    Y.g_pos.upload(g_pos, s);
    Y.g_ptr.upload(g_ptr, s);
    Y.g_off.upload(g_off, s);
    Y.n_order = (int)order.size();
    Y.n_leaves = (int)work.size();
    Y.n_g = (int)g_pos.size();
    Y.n_rev = n_rev;
    clk.mark("sym: upload");
    split_leaves(T, leaves, ord, lm, P.n_e, Y.split, s);
    clk.mark("sym: split");
    close_flags(Y.ord, P.obs_leaves.p, P.dot, P.op, P.sp, s);
    Y.on = true;
}

static std::shared_ptr<FmmPlan> build_plan(const tbem_fmm_desc* d, hipStream_t s, const SavedPlan* sv = nullptr) {
    auto t0 = std::chrono::steady_clock::now();
    PhaseClock clk(s);
    validate(d);
    auto P = std::make_shared<FmmPlan>();
    P->kid = d->kernel;
    P->R = kernel_dim(d->kernel);
    P->precision = d->precision;
    P->kc = make_kconst(d->shear_modulus, d->poisson_ratio);
    P->n_obs = d->n_obs;
    P->n_src = d->n_src;
    P->mac = d->mac;
    P->d = d->d;
    P->order = d->order;
    P->leaf = d->min_pts_per_cell;
    P->small_cells = d->account_for_small_cells != 0;
    P->surf = make_surfaces(d->order, d->d);
    P->n_e = P->surf.n_e;
    P->blk = P->R * P->n_e;
    const int leaf = (int)std::min<size_t>(d->min_pts_per_cell, (size_t)1 << 30);

    P->input_hash = input_hash(d);
    const bool same = d->n_obs == d->n_src &&
                      (d->obs_locs == d->src_locs ||
                       (d->n_obs > 0 && std::memcmp(d->obs_locs, d->src_locs, 3 * sizeof(double) * d->n_obs) == 0));
    P->same_tree = same;
    const bool common = d->tree_boxes == TBEM_TREES_COMMON_BOX ||
                        (d->tree_boxes == TBEM_TREES_AUTO && P->precision == TBEM_MIXED && m2l_mfma_size_ok(P->blk));
    P->tree_layout = same ? kTreeShared : (common && d->n_obs > 0 && d->n_src > 0) ? kTreesCommonBox : kTreesOwnBoxes;
    if (sv) {
        TBEM_REQUIRE(sv->same == same, TBEM_E_INVALID, "saved plan: tree sharing differs from the inputs'");
        TBEM_REQUIRE(sv->layout == P->tree_layout, TBEM_E_INVALID,
                     "saved plan: its trees were laid out on other boxes (precision or tree_boxes differ)");
        P->st = sv->st;
        P->ot = sv->ot;
    } else if (P->tree_layout == kTreesCommonBox) {
        double a[6], b[6];
        point_bounds(d->src_locs, (int)d->n_src, a);
        point_bounds(d->obs_locs, (int)d->n_obs, b);
        for (int k = 0; k < 3; k++) {
            a[k] = std::min(a[k], b[k]);
            a[3 + k] = std::max(a[3 + k], b[3 + k]);
        }
        P->st = build_octree(d->src_locs, (int)d->n_src, leaf, s, a);
        P->ot = build_octree(d->obs_locs, (int)d->n_obs, leaf, s, a);
    } else {
        P->st = build_octree(d->src_locs, (int)d->n_src, leaf, s);
        P->ot = same ? P->st : build_octree(d->obs_locs, (int)d->n_obs, leaf, s);
    }
    clk.mark("octrees");
    P->dst.upload(P->st, s);
    P->dot.upload(P->ot, s);
    P->sp.upload(P->st, d->src_locs, d->src_normals, d->src_weights, s);
    P->op.upload(P->ot, d->obs_locs, d->obs_normals, nullptr, s);
    P->d_unit.upload(P->surf.unit, s);
    P->d_inner.upload(P->surf.inner, s);
    P->d_outer.upload(P->surf.outer, s);

    clk.mark("uploads");
    // check-to-equivalent operators
    const size_t bb = (size_t)P->blk * P->blk;
    // the mixed path's modal factors come from the C2E pseudo-inverses' own SVDs
    const bool keep_fac = P->precision == TBEM_MIXED;
    DevBuf<double> up_fac[2], down_fac[2];
    if (sv) {
        P->uc2e.upload(sv->uc2e, s);
        P->dc2e.upload(sv->dc2e, s);
        if (keep_fac) {
            for (int k = 0; k < 2; k++) {
                if (!sv->fac[k].empty()) up_fac[k].upload(sv->fac[k], s);
                if (!sv->fac[2 + k].empty()) down_fac[k].upload(sv->fac[2 + k], s);
            }
        }
    } else if (d->up_check_to_equiv) {
        TBEM_REQUIRE(d->n_up_levels >= (size_t)P->st.n_levels, TBEM_E_INVALID,
                     "injected UC2E has fewer levels than the source tree");
        P->uc2e.upload(d->up_check_to_equiv, bb * P->st.n_levels, s);
    } else {
        compute_c2e(P->kid, P->kc, P->st, P->surf, true, P->uc2e, s, keep_fac ? &up_fac[0] : nullptr,
                    keep_fac ? &up_fac[1] : nullptr);
    }
    if (sv) {
    } else if (d->down_check_to_equiv) {
        TBEM_REQUIRE(d->n_down_levels >= (size_t)P->ot.n_levels, TBEM_E_INVALID,
                     "injected DC2E has fewer levels than the observation tree");
        P->dc2e.upload(d->down_check_to_equiv, bb * P->ot.n_levels, s);
    } else {
        compute_c2e(P->kid, P->kc, P->ot, P->surf, false, P->dc2e, s, keep_fac ? &down_fac[0] : nullptr,
                    keep_fac ? &down_fac[1] : nullptr);
    }
    P->h_uc2e = P->uc2e.to_host(s);
    P->h_dc2e = P->dc2e.to_host(s);
    if (keep_fac) {
        for (int k = 0; k < 2; k++) {
            if (up_fac[k].n) P->h_fac[k] = up_fac[k].to_host(s);
            if (down_fac[k].n) P->h_fac[2 + k] = down_fac[k].to_host(s);
        }
    }

    clk.mark("c2e");
    // task lists (cpp/fmm.cpp:560-648)
    if (sv) P->tl = sv->tl;
    else dual_tree_traversal(P->ot, P->st, P->mac, P->n_e, P->small_cells, P->tl);
    clk.mark("traversal");
    const TaskLists& tl = P->tl;
    for (auto& p: tl.p2p) P->p2p_pairs += (double)P->ot.count[p[0]] * (double)P->st.count[p[1]];

    const HostTree& ST = P->st;
    const HostTree& OT = P->ot;

    P->p2m.upload(self_list(ST, ST.leaves), s);
    P->t_p2m = make_cell_tiles(ST.leaves, ST, s);
    // M2M and L2L on class operators: the cells of one tree always lie on one grid
    const bool tr_classes = d->translations == TBEM_TRANSLATIONS_AUTO && m2l_fp64_size_ok(P->blk);
    auto note_y = [&](const M2lMfma& M) {
        if (M.enabled) P->tr_y = std::max(P->tr_y, (size_t)M.y_pairs * M.ystride);
    };
    P->m2m.resize(ST.n_levels);
    P->t_m2m.resize(ST.n_levels);
    P->m2m_ops.resize(ST.n_levels);
    for (int L = 0; L + 1 < ST.n_levels; L++) {
        const HostList l = children_list(ST, L);
        P->m2m[L].upload(l, s);
        P->t_m2m[L] = make_cell_tiles(l.items, ST, s);
        if (tr_classes) {
            build_m2l_mfma(P->m2m_ops[L], P->kid, P->kc, ST, ST, P->dst, P->dst, true, true, kM2lOutputBudget, l,
                           P->m2m[L], P->surf, nullptr, nullptr, s, nullptr, nullptr, &P->surf.outer, &P->surf.inner);
            note_y(P->m2m_ops[L]);
        }
    }
    const HostList m2l_h = list_from_pairs(cells_with(tl.m2l, 0, OT.n_cells), OT.n_cells, tl.m2l, 0, nullptr);
    P->m2l.upload(m2l_h, s);
    P->t_m2l = make_cell_tiles(m2l_h.items, OT, s);
    P->m2l_no_cutoff = m2l_clear_of_cutoff(m2l_h, OT, ST, P->surf);
    clk.mark("lists");
    // FP64: class operators where the classes exist (the builder also requires the
    // operators to fit an eighth of the device memory)
    const bool fp64_classes = P->precision == TBEM_FP64 && d->translations == TBEM_TRANSLATIONS_AUTO &&
                              P->tree_layout != kTreesOwnBoxes && m2l_fp64_size_ok(P->blk);
    if (P->precision == TBEM_MIXED || fp64_classes)
        build_m2l_mfma(P->m2l_mfma, P->kid, P->kc, OT, ST, P->dot, P->dst, P->tree_layout != kTreesOwnBoxes,
                       fp64_classes, d->m2l_output_budget ? d->m2l_output_budget : kM2lOutputBudget, m2l_h, P->m2l,
                       P->surf, P->uc2e.p, P->dc2e.p, s, up_fac, down_fac);
    clk.mark("m2l_mfma");
    TBEM_HIP(hipEventCreateWithFlags(&P->y_done, hipEventDisableTiming));
    TBEM_HIP(hipEventRecord(P->y_done, s));
    if (P->m2l_mfma.enabled) {
        const size_t ny = (size_t)P->m2l_mfma.y_pairs * P->m2l_mfma.ystride;
        if (P->m2l_mfma.fp64) P->m2l_y64.alloc(ny);
        else P->m2l_y.alloc(ny);
    }
    P->d_flag.alloc(1);
    const std::vector<int> p2l_items = cells_with(tl.p2l, 0, OT.n_cells);
    P->p2l.upload(list_from_pairs(p2l_items, OT.n_cells, tl.p2l, 0, &ST), s);
    P->t_p2l = make_cell_tiles(p2l_items, OT, s);
    P->t_olev.resize(OT.n_levels);
    for (int L = 0; L < OT.n_levels; L++) P->t_olev[L] = make_level_tiles(OT, L, s);
    P->l2l.resize(OT.n_levels);
    P->l2l_ops.resize(OT.n_levels);
    for (int L = 1; L < OT.n_levels; L++) {
        const HostList l = parent_list(OT, L);
        P->l2l[L].upload(l, s);
        if (tr_classes) {
            build_m2l_mfma(P->l2l_ops[L], P->kid, P->kc, OT, OT, P->dot, P->dot, true, true, kM2lOutputBudget, l,
                           P->l2l[L], P->surf, nullptr, nullptr, s, nullptr, nullptr, &P->surf.inner, &P->surf.outer);
            note_y(P->l2l_ops[L]);
        }
    }
    P->obs_leaves.upload(OT.leaves, s);
    P->src_leaves.upload(ST.leaves, s);
    {
        const HostList lp = list_from_pairs(OT.leaves, OT.n_cells, tl.p2p, 0, &ST);
        const HostList lm = list_from_pairs(OT.leaves, OT.n_cells, tl.m2p, 0, nullptr);
        P->leaf_p2p.upload(lp, s);
        P->leaf_m2p.upload(lm, s);
        split_leaves(OT, OT.leaves, lp, lm, P->n_e, P->leaf_split, s);
        clk.mark("leaf lists");
        if (P->same_tree && (P->kid == TBEM_K_ELASTIC_DISPLACEMENT || P->kid == TBEM_K_LAPLACE_SINGLE) &&
            d->point_pairs == TBEM_POINT_PAIRS_AUTO)
            build_sym(*P, OT, tl.p2p, lm, s, clk);
        clk.mark("symmetric pairs");
    }
    close_flags(P->leaf_p2p, P->obs_leaves.p, P->dot, P->op, P->sp, s);

    clk.mark("rest");
    TBEM_HIP(hipStreamSynchronize(s));
    P->plan_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return P;
}

// The transposed lists of the adjoint, built once per plan on the first call that
// needs them (on that call's stream, which is then synchronized: another handle of
// the plan may use them on another stream right after).
const FmmPlan::AdjLists& FmmPlan::adjoint_lists(hipStream_t s) const {
    std::lock_guard<std::mutex> lk(adj_mu);
    if (adj) return *adj;
    std::unique_ptr<AdjLists> A(new AdjLists);
    const HostTree& ST = st;
    const HostTree& OT = ot;
    A->l2p.upload(self_list(OT, OT.leaves), s);
    A->l2l.resize(OT.n_levels);
    for (int L = 0; L + 1 < OT.n_levels; L++) A->l2l[L].upload(children_list(OT, L), s);
    A->m2l.upload(list_from_pairs(cells_with(tl.m2l, 1, ST.n_cells), ST.n_cells, tl.m2l, 1, nullptr), s);
    A->m2p.upload(list_from_pairs(cells_with(tl.m2p, 1, ST.n_cells), ST.n_cells, tl.m2p, 1, &OT), s);
    A->m2m.resize(ST.n_levels);
    for (int L = 1; L < ST.n_levels; L++) A->m2m[L].upload(parent_list(ST, L), s);
    {
        const HostList lp = list_from_pairs(ST.leaves, ST.n_cells, tl.p2p, 1, &OT);
        const HostList ll = list_from_pairs(ST.leaves, ST.n_cells, tl.p2l, 1, nullptr);
        A->leaf_p2p.upload(lp, s);
        A->leaf_p2l.upload(ll, s);
        split_leaves(ST, ST.leaves, lp, ll, n_e, A->leaf_split, s);
        if (sym.on) {  // one shared tree: source leaf positions are the forward's
            HostList ord = list_from_pairs(ST.leaves, ST.n_cells, tl.p2p, 1, nullptr);
            const size_t nl = ord.ptr.size() - 1;
            std::vector<char> keep(ord.cell.size());
            std::vector<int> ptr(nl + 1, 0), cell;
            parallel_for((nl + 1023) / 1024, [&](size_t ch) {  // mark, then fill at the prefix offsets
                for (size_t k = ch * 1024; k < std::min(nl, (ch + 1) * 1024); k++) {
                    const auto b = sym.part.begin() + sym.part_ptr[k], e = sym.part.begin() + sym.part_ptr[k + 1];
                    int n = 0;
                    for (int q = ord.ptr[k]; q < ord.ptr[k + 1]; q++) {
                        keep[q] = !std::binary_search(b, e, ord.cell[q]);
                        n += keep[q];
                    }
                    ptr[k + 1] = n;
                }
            });
            for (size_t k = 0; k < nl; k++) ptr[k + 1] += ptr[k];
            cell.resize(ptr[nl]);
            parallel_for((nl + 1023) / 1024, [&](size_t ch) {
                for (size_t k = ch * 1024; k < std::min(nl, (ch + 1) * 1024); k++) {
                    int o = ptr[k];
                    for (int q = ord.ptr[k]; q < ord.ptr[k + 1]; q++)
                        if (keep[q]) cell[o++] = ord.cell[q];
                }
            });
            ord.ptr = std::move(ptr);
            ord.cell = std::move(cell);
            cells_to_ranges(ord, ST);
            A->sym_ord.upload(ord, s);
            split_leaves(ST, ST.leaves, ord, ll, n_e, A->sym_split, s);
            close_flags(A->sym_ord, src_leaves.p, dst, sp, op, s);
        }
    }
    close_flags(A->leaf_p2p, src_leaves.p, dst, sp, op, s);
    A->t_slev.resize(ST.n_levels);
    for (int L = 0; L < ST.n_levels; L++) A->t_slev[L] = make_level_tiles(ST, L, s);
    TBEM_HIP(hipStreamSynchronize(s));
    adj = std::move(A);
    return *adj;
}

// --------------------------------------------------------------------- stage glue
static SurfStage surf_base(const FmmPlan& P) {
    SurfStage a{};
    a.unit = P.d_unit.p;
    a.n_e = P.n_e;
    a.blk = P.blk;
    return a;
}

static void set_list(SurfStage& a, const DevList& l) {
    a.n_items = l.n_items;
    a.items = l.items.p;
    a.ptr = l.ptr.p;
    a.cell = l.cell.p;
    a.rbeg = l.rbeg.p;
    a.rcnt = l.rcnt.p;
    a.rpre = l.rpre.p;
}

// gather, upward pass, M2L/P2L and L2L: everything the leaf stage consumes. FP64,
// except the M2L when mfma is set (mixed precision, MFMA operators built).
// after_gather (optional) runs once the densities are gathered (on s), before the
// upward pass is enqueued; without it (F = nullptr_t) nothing is called.
template <class F = std::nullptr_t>
static void farfield(const FmmPlan& P, Work& W, const double* x, hipStream_t s, bool mfma, F&& after_gather = nullptr) {
    const int R = P.R;
    const int Ns = (int)P.n_src;
    TBEM_HIP(hipEventRecord(W.ev[0], s));
    launch_gather_dens(x, P.sp.perm.p, P.sp.w.p, Ns, R, W.dens.p, s);
    TBEM_HIP(hipEventRecord(W.ev[1], s));
    if constexpr (!std::is_same<typename std::decay<F>::type, std::nullptr_t>::value) after_gather();

    // the up-check potentials of a stage's cells -> equivalent densities, and, for
    // the MFMA M2L, the modal FP32 input Bu * check
    auto up_c2e = [&](const CellTiles& T) {
        launch_level_gemm(T, P.blk, P.uc2e.p, false, W.fld.p, nullptr, W.mult.p, nullptr, false, s);
        if (mfma) launch_level_gemm(T, P.blk, P.m2l_mfma.Bu.p, false, W.fld.p, nullptr, nullptr, W.chk_up.p, false, s);
    };
    // down-check potentials -> added to the down-equivalent densities
    auto down_c2e = [&](const CellTiles& T) {
        launch_level_gemm(T, P.blk, P.dc2e.p, false, W.fld.p, nullptr, W.locals.p, nullptr, true, s);
    };

    // P2M (cpp/fmm.cpp:162-202): points of each source leaf -> up-check -> UC2E
    SurfStage a = surf_base(P);
    set_list(a, P.p2m);
    a.tt = view(P.dst); a.t_surf = P.d_outer.p;
    a.st = view(P.dst); a.sp = view(P.sp); a.dens = W.dens.p; a.n_src_pts = Ns;
    a.out = W.fld.p;
    launch_surf_stage(P.kid, false, false, false, a, P.kc, s);
    up_c2e(P.t_p2m);
    TBEM_HIP(hipEventRecord(W.ev[2], s));

    // M2M (cpp/fmm.cpp:204-251), deepest parents first
    for (int L = P.st.n_levels - 2; L >= 0; L--) {
        if (P.m2m_ops[L].enabled) {
            m2l_forward64(P.m2m_ops[L], P.m2m[L].items.p, P.m2m[L].ptr.p, P.m2m[L].cell.p, W.mult.p, W.ty.p, W.fld.p,
                          s, nullptr);
        } else {
            SurfStage m = surf_base(P);
            set_list(m, P.m2m[L]);
            m.tt = view(P.dst); m.t_surf = P.d_outer.p;
            m.st = view(P.dst); m.s_surf = P.d_inner.p; m.s_coeff = W.mult.p;
            m.out = W.fld.p;
            launch_surf_stage(P.kid, true, false, false, m, P.kc, s);
        }
        up_c2e(P.t_m2m[L]);
    }
    TBEM_HIP(hipEventRecord(W.ev[3], s));

    // M2L (cpp/fmm.cpp:362-382) and P2L (cpp/fmm.cpp:296-330) into locals
    launch_fill(W.locals.p, W.locals.n, 0.0, s);
    TBEM_HIP(hipEventRecord(W.ev[8], s));
    if (mfma) {
        // Vd M_c Bu check = DC2E K_c UC2E check = the FP64 stage's DC2E K_c mult
        YLease y(P, s);
        m2l_forward32(P.m2l_mfma, P.m2l.items.p, P.m2l.ptr.p, P.m2l.cell.p, P.t_m2l, W.chk_up.p, P.m2l_y.p, W.fld.p,
                      W.locals.p, s, W.ev[9]);
    } else if (P.m2l_mfma.enabled && P.m2l_mfma.fp64) {
        // FP64 class operators: check potential = sum over pairs of K_c mult, then DC2E
        YLease y(P, s);
        m2l_forward64(P.m2l_mfma, P.m2l.items.p, P.m2l.ptr.p, P.m2l.cell.p, W.mult.p, P.m2l_y64.p, W.fld.p, s,
                      W.ev[9]);
        down_c2e(P.t_m2l);
    } else {
        SurfStage m = surf_base(P);
        set_list(m, P.m2l);
        m.tt = view(P.dot); m.t_surf = P.d_inner.p;
        m.st = view(P.dst); m.s_surf = P.d_inner.p; m.s_coeff = W.mult.p;
        m.out = W.fld.p;
        m.no_cutoff = P.m2l_no_cutoff;
        launch_surf_stage(P.kid, true, false, false, m, P.kc, s);
        TBEM_HIP(hipEventRecord(W.ev[9], s));
        down_c2e(P.t_m2l);
    }
    SurfStage pl = surf_base(P);
    set_list(pl, P.p2l);
    pl.tt = view(P.dot); pl.t_surf = P.d_inner.p;
    pl.st = view(P.dst); pl.sp = view(P.sp); pl.dens = W.dens.p; pl.n_src_pts = Ns;
    pl.out = W.fld.p;
    launch_surf_stage(P.kid, false, false, false, pl, P.kc, s);
    down_c2e(P.t_p2l);
    TBEM_HIP(hipEventRecord(W.ev[4], s));

    // L2L (cpp/fmm.cpp:384-413), root first
    for (int L = 1; L < P.ot.n_levels; L++) {
        if (P.l2l_ops[L].enabled) {
            m2l_forward64(P.l2l_ops[L], P.l2l[L].items.p, P.l2l[L].ptr.p, P.l2l[L].cell.p, W.locals.p, W.ty.p,
                          W.fld.p, s, nullptr);
        } else {
            SurfStage l = surf_base(P);
            set_list(l, P.l2l[L]);
            l.tt = view(P.dot); l.t_surf = P.d_inner.p;
            l.st = view(P.dot); l.s_surf = P.d_outer.p; l.s_coeff = W.locals.p;
            l.out = W.fld.p;
            launch_surf_stage(P.kid, true, false, false, l, P.kc, s);
        }
        down_c2e(P.t_olev[L]);
    }
    TBEM_HIP(hipEventRecord(W.ev[5], s));
}

static void set_split(LeafStage& f, const LeafSplit& sp, Work& W) {
    if (W.wpart.n < sp.part_size) {
        // (the transpose's split on its first call) the previous call's leaf kernels may
        // still read the old buffer: wait for them before it is freed
        TBEM_HIP(hipEventSynchronize(W.done));
        W.wpart.alloc(sp.part_size);
    }
    f.wave_items = sp.items.p; f.n_wave = sp.n_items;
    f.wave_parts = sp.parts.p; f.n_parts = sp.n_parts; f.part_buf = W.wpart.p;
    f.block_pos = sp.block.p; f.n_block = sp.n_block;
}

// The forward's point pairs through the symmetric evaluation, where the plan has it
// and the observation points have not moved off the sources.
// In the transpose (A non-null) U^T = U: the same pairs, each sum times the source
// weight of the point it goes to (t_weight).
static void use_sym(LeafStage& f, const FmmPlan& P, Work& W, const FmmPlan::AdjLists* A = nullptr) {
    const FmmPlan::SymP2P& Y = P.sym;
    if (!Y.on || P.obs_moved) return;
    const DevList& o = A ? A->sym_ord : Y.ord;
    f.rp_ptr = o.ptr.p; f.rp_beg = o.rbeg.p; f.rp_cnt = o.rcnt.p;
    f.rp_pre = o.rpre.p; f.rp_close = o.rclose.p;
    set_split(f, A ? A->sym_split : Y.split, W);
    f.sy_ptr = Y.own.ptr.p; f.sy_beg = Y.own.rbeg.p; f.sy_cnt = Y.own.rcnt.p; f.sy_pre = Y.own.rpre.p;
    f.sy_base = Y.base.p; f.sy_order = Y.order.p; f.n_sy_order = Y.n_order;
    f.sg_pos = Y.g_pos.p; f.sg_ptr = Y.g_ptr.p; f.sg_off = Y.g_off.p; f.n_sg = Y.n_g;
    f.rev = W.wrev.p; f.n_rev = Y.n_rev;
}

// P2P + M2P + L2P per observation leaf (cpp/fmm.cpp:254-294, 332-360, 415-441)
static LeafStage leaf_stage_args(const FmmPlan& P, Work& W) {
    const int Ns = (int)P.n_src, No = (int)P.n_obs;
    LeafStage f{};
    f.n_leaves = (int)P.ot.leaves.size();
    f.leaves = P.obs_leaves.p;
    f.tt = view(P.dot); f.tp = view(P.op); f.t_weight = nullptr; f.n_tgt = No;
    f.rp_ptr = P.leaf_p2p.ptr.p; f.rp_beg = P.leaf_p2p.rbeg.p; f.rp_cnt = P.leaf_p2p.rcnt.p;
    f.rp_pre = P.leaf_p2p.rpre.p; f.rp_close = P.leaf_p2p.rclose.p;
    set_split(f, P.leaf_split, W);
    f.ps = view(P.sp); f.pdens = W.dens.p; f.n_ps = Ns;
    f.sl_ptr = P.leaf_m2p.ptr.p; f.sl_cell = P.leaf_m2p.cell.p;
    f.slt = view(P.dst); f.sl_surf = P.d_inner.p; f.sl_coeff = W.mult.p;
    f.self_surf = P.d_outer.p; f.self_coeff = W.locals.p;
    f.unit = P.d_unit.p; f.n_e = P.n_e; f.blk = P.blk;
    f.out_sorted = W.out_sorted.p;
    return f;
}

template <class T>
static void forward_T(const FmmPlan& P, Work& W, const double* x, double* y, hipStream_t s) {
    const int R = P.R;
    const int Ns = (int)P.n_src, No = (int)P.n_obs;
    if (No == 0) return;
    if (Ns == 0) {
        launch_fill(y, (size_t)R * No, 0.0, s);
        return;
    }
    const bool mfma = std::is_same<T, float>::value && P.m2l_mfma.enabled;
    W.split_last = W.overlap;
    if (W.overlap) {
        // P2P on the side stream once the densities are gathered, beside the far field
        farfield(P, W, x, s, mfma, [&] {
            TBEM_HIP(hipEventRecord(W.fork, s));
            TBEM_HIP(hipStreamWaitEvent(W.side, W.fork, 0));
            LeafStage f = leaf_stage_args(P, W);
            f.sl_ptr = nullptr;
            f.sl_cell = nullptr;
            f.self_coeff = nullptr;
            use_sym(f, P, W);
            TBEM_HIP(hipEventRecord(W.ev[10], W.side));
            launch_leaf_stage<T>(P.kid, false, f, P.kc, W.side, true);
            TBEM_HIP(hipEventRecord(W.ev[11], W.side));
            TBEM_HIP(hipEventRecord(W.join, W.side));
        });
        W.far_token = ++W.next_token;
        // then M2P + L2P (FP64) added to the P2P output
        TBEM_HIP(hipStreamWaitEvent(s, W.join, 0));
        LeafStage f = leaf_stage_args(P, W);
        f.rp_ptr = nullptr;
        f.accumulate = 1;
        launch_leaf_stage<double>(P.kid, false, f, P.kc, s, true);
    } else {
        farfield(P, W, x, s, mfma);
        W.far_token = ++W.next_token;
        LeafStage f = leaf_stage_args(P, W);
        use_sym(f, P, W);
        launch_leaf_stage<T>(P.kid, false, f, P.kc, s, true);
    }
    TBEM_HIP(hipEventRecord(W.ev[6], s));
    launch_scatter_out(W.out_sorted.p, P.op.perm.p, No, R, y, s);
    TBEM_HIP(hipEventRecord(W.ev[7], s));
    W.timed = true;
}

__global__ void k_scatter_aos3(const double* __restrict__ in, const int* __restrict__ perm, int n,
                               double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t g = (size_t)perm[i];
    for (int m = 0; m < 3; m++) out[3 * g + m] = in[3 * (size_t)i + m];
}

// grad[3*i+m] = sum_q g_q(i) d u_q(i) / d obs_i[m], plan geometry held fixed. The
// far field is the operator's own; in mixed precision the gradient leaf stage runs
// its point sources in FP32.
// reuse_token: a far_token of this handle; while it is still current, the far
// field (mult, locals) of that call is reused and only the densities of x are
// gathered again (x must then hold that call's input).
static void obs_grad(const FmmPlan& P, Work& W, const double* x, const double* g, double* grad,
                     hipStream_t s, uint64_t reuse_token = 0) {
    const int R = P.R;
    const int Ns = (int)P.n_src, No = (int)P.n_obs;
    if (No == 0) return;
    if (Ns == 0) {
        launch_fill(grad, 3 * (size_t)No, 0.0, s);
        return;
    }
    // the mixed far field (MFMA M2L) agrees with the FP64 one to ~2e-7
    if (reuse_token != 0 && reuse_token == W.far_token) {
        launch_gather_dens(x, P.sp.perm.p, P.sp.w.p, Ns, R, W.dens.p, s);
    } else {
        farfield(P, W, x, s, P.precision == TBEM_MIXED && P.m2l_mfma.enabled);
        W.far_token = ++W.next_token;
    }
    if (W.gsort.n < (size_t)R * No) W.gsort.alloc((size_t)R * No);
    if (W.grad_sorted.n < 3 * (size_t)No) W.grad_sorted.alloc(3 * (size_t)No);
    launch_gather_dens(g, P.op.perm.p, nullptr, No, R, W.gsort.p, s);
    LeafStage f = leaf_stage_args(P, W);
    f.grad_g = W.gsort.p;
    f.grad_out = W.grad_sorted.p;
    launch_leaf_grad(P.kid, f, P.kc, s, P.precision == TBEM_MIXED);
    hipLaunchKernelGGL(k_scatter_aos3, dim3(div_up(No, 256)), dim3(256), 0, s, W.grad_sorted.p, P.op.perm.p,
                       No, grad);
    TBEM_LAUNCH_CHECK();
}

// Exact transpose of forward_T: every stage transposed, run in reverse order.
template <class T>
static void adjoint_T(const FmmPlan& P, Work& W, const double* g, double* xbar, hipStream_t s) {
    const int R = P.R;
    const int Ns = (int)P.n_src, No = (int)P.n_obs;
    if (Ns == 0) return;
    if (No == 0) {
        launch_fill(xbar, (size_t)R * Ns, 0.0, s);
        return;
    }
    const size_t nco = (size_t)P.dot.n_cells * P.blk, ncs = (size_t)P.dst.n_cells * P.blk;
    if (W.gsort.n < (size_t)R * No) W.gsort.alloc((size_t)R * No);
    if (W.lbar.n < nco) { W.lbar.alloc(nco); W.cbar.alloc(nco); }
    if (W.mbar.n < ncs) { W.mbar.alloc(ncs); W.ebar.alloc(ncs); }
    if (W.dens.n < (size_t)R * Ns) W.dens.alloc((size_t)R * Ns);
    const bool mfma = std::is_same<T, float>::value && P.m2l_mfma.enabled;
    const FmmPlan::AdjLists& A = P.adjoint_lists(s);
    if (mfma && W.e2.n < ncs) { W.e2.alloc(ncs); W.m2l_w.alloc(nco); }
    launch_gather_dens(g, P.op.perm.p, nullptr, No, R, W.gsort.p, s);
    // P2P^T + P2L^T + P2M^T on every source leaf, times the source weights
    auto leaf_args = [&] {
        LeafStage f{};
        f.n_leaves = (int)P.st.leaves.size();
        f.leaves = P.src_leaves.p;
        f.tt = view(P.dst); f.tp = view(P.sp); f.t_weight = P.sp.w.p; f.n_tgt = Ns;
        f.rp_ptr = A.leaf_p2p.ptr.p; f.rp_beg = A.leaf_p2p.rbeg.p; f.rp_cnt = A.leaf_p2p.rcnt.p;
        f.rp_pre = A.leaf_p2p.rpre.p; f.rp_close = A.leaf_p2p.rclose.p;
        set_split(f, A.leaf_split, W);
        f.ps = view(P.op); f.pdens = W.gsort.p; f.n_ps = No;
        f.sl_ptr = A.leaf_p2l.ptr.p; f.sl_cell = A.leaf_p2l.cell.p;
        f.slt = view(P.dot); f.sl_surf = P.d_inner.p; f.sl_coeff = W.cbar.p;
        f.self_surf = P.d_outer.p; f.self_coeff = W.ebar.p;
        f.unit = P.d_unit.p; f.n_e = P.n_e; f.blk = P.blk;
        f.out_sorted = W.dens.p;
        return f;
    };
    const bool split = W.overlap;
    if (split) {  // P2P^T needs only the gathered g: on the side stream, beside the rest
        TBEM_HIP(hipEventRecord(W.fork, s));
        TBEM_HIP(hipStreamWaitEvent(W.side, W.fork, 0));
        LeafStage f = leaf_args();
        f.sl_ptr = nullptr;
        f.sl_cell = nullptr;
        f.self_coeff = nullptr;
        use_sym(f, P, W, &A);
        launch_leaf_stage<T>(P.kid, true, f, P.kc, W.side, true);
        TBEM_HIP(hipEventRecord(W.join, W.side));
    }
    launch_fill(W.lbar.p, nco, 0.0, s);
    launch_fill(W.cbar.p, nco, 0.0, s);
    launch_fill(W.mbar.p, ncs, 0.0, s);
    launch_fill(W.ebar.p, ncs, 0.0, s);

    // L2P^T: observation points -> down-equivalent surface of their leaf
    SurfStage a = surf_base(P);
    set_list(a, A.l2p);
    a.tt = view(P.dot); a.t_surf = P.d_outer.p;
    a.st = view(P.dot); a.sp = view(P.op); a.dens = W.gsort.p; a.n_src_pts = No;
    a.out = W.lbar.p;
    launch_surf_stage(P.kid, false, false, true, a, P.kc, s);

    // L2L^T, deepest level first: cbar = DC2E^T lbar, then push to parents
    for (int L = P.ot.n_levels - 1; L >= 0; L--) {
        launch_level_gemm(P.t_olev[L], P.blk, P.dc2e.p, true, W.lbar.p, nullptr, W.cbar.p, nullptr, false, s);
        if (L == 0) break;
        if (P.l2l_ops[L].enabled) {  // lbar[parent] += sum over its children of K^T cbar[child]
            m2l_adjoint64(P.l2l_ops[L], W.cbar.p, W.ty.p, W.lbar.p, s);
            continue;
        }
        SurfStage l = surf_base(P);
        set_list(l, A.l2l[L - 1]);
        l.tt = view(P.dot); l.t_surf = P.d_outer.p;
        l.st = view(P.dot); l.s_surf = P.d_inner.p; l.s_coeff = W.cbar.p;
        l.out = W.lbar.p;
        launch_surf_stage(P.kid, true, true, true, l, P.kc, s);
    }

    // M2L^T into mbar (FP64), or, on MFMA, e2 = Bu^T M_c^T Vd^T lbar, added to ebar below
    if (mfma) {
        launch_fill(W.e2.p, ncs, 0.0, s);
        YLease y(P, s);
        m2l_adjoint32(P.m2l_mfma, P.t_m2l, W.lbar.p, W.m2l_w.p, P.m2l_y.p, W.fld.p, W.e2.p, s);
    } else if (P.m2l_mfma.enabled && P.m2l_mfma.fp64) {
        // FP64 class operators: mbar[s] += sum over the pairs of s of K_c^T cbar[t]
        YLease y(P, s);
        m2l_adjoint64(P.m2l_mfma, W.cbar.p, P.m2l_y64.p, W.mbar.p, s);
    } else {
        SurfStage m = surf_base(P);
        set_list(m, A.m2l);
        m.tt = view(P.dst); m.t_surf = P.d_inner.p;
        m.st = view(P.dot); m.s_surf = P.d_inner.p; m.s_coeff = W.cbar.p;
        m.out = W.mbar.p;
        m.no_cutoff = P.m2l_no_cutoff;
        launch_surf_stage(P.kid, true, true, true, m, P.kc, s);
    }
    // M2P^T into mbar
    SurfStage mp = surf_base(P);
    set_list(mp, A.m2p);
    mp.tt = view(P.dst); mp.t_surf = P.d_inner.p;
    mp.st = view(P.dot); mp.sp = view(P.op); mp.dens = W.gsort.p; mp.n_src_pts = No;
    mp.out = W.mbar.p;
    launch_surf_stage(P.kid, false, true, true, mp, P.kc, s);

    // M2M^T, root first: ebar = UC2E^T mbar, then push to children
    for (int L = 0; L < P.st.n_levels; L++) {
        launch_level_gemm(A.t_slev[L], P.blk, P.uc2e.p, true, W.mbar.p, mfma ? W.e2.p : nullptr, W.ebar.p, nullptr,
                          false, s);
        if (L + 1 >= P.st.n_levels) break;
        if (P.m2m_ops[L].enabled) {  // mbar[child] += K^T ebar[parent]
            m2l_adjoint64(P.m2m_ops[L], W.ebar.p, W.ty.p, W.mbar.p, s);
            continue;
        }
        SurfStage c = surf_base(P);
        set_list(c, A.m2m[L + 1]);
        c.tt = view(P.dst); c.t_surf = P.d_inner.p;
        c.st = view(P.dst); c.s_surf = P.d_outer.p; c.s_coeff = W.ebar.p;
        c.out = W.mbar.p;
        launch_surf_stage(P.kid, true, true, true, c, P.kc, s);
    }

    if (split) {  // P2L^T + P2M^T (FP64) added to the side stream's P2P^T
        TBEM_HIP(hipStreamWaitEvent(s, W.join, 0));
        LeafStage f = leaf_args();
        f.rp_ptr = nullptr;
        f.accumulate = 1;
        launch_leaf_stage<double>(P.kid, true, f, P.kc, s, true);
    } else {
        LeafStage f = leaf_args();
        use_sym(f, P, W, &A);
        launch_leaf_stage<T>(P.kid, true, f, P.kc, s, true);
    }
    launch_scatter_out(W.dens.p, P.sp.perm.p, Ns, R, xbar, s);
}

// ------------------------------------------------- moving the observation points
// Counts the moved observation points (AoS, caller's order) that lie outside the
// ball of the leaf they were sorted into: centre c, radius |half width|, the
// radius the dual-tree MAC test gives the cell (cpp/fmm.cpp:577-586), with a 1e-9
// relative allowance for the rounding of the subdivided centres. (A box test
// would refuse every move off a flat point set, whose cells have zero extent
// along its normal.) NaN coordinates count as outside. One workgroup per
// observation leaf.
__global__ void k_obs_check(const double* __restrict__ locs, const int* __restrict__ perm,
                            const int* __restrict__ leaves, TreeView t, int* __restrict__ n_out) {
    const int leaf = leaves[blockIdx.x];
    const int b = t.begin[leaf], m = t.count[leaf];
    const double r = t.level_r[t.level[leaf]];
    const double c[3] = {t.cx[leaf], t.cy[leaf], t.cz[leaf]};
    int bad = 0;
    for (int i = b + threadIdx.x; i < b + m; i += blockDim.x) {
        const size_t g = (size_t)perm[i];
        const double dx = locs[3 * g] - c[0], dy = locs[3 * g + 1] - c[1], dz = locs[3 * g + 2] - c[2];
        bad += !(dx * dx + dy * dy + dz * dz <= r * r * (1.0 + 1e-9));
    }
    if (bad) atomicAdd(n_out, bad);
}

// sorted SoA coordinates <- AoS points in the caller's order
__global__ void k_obs_move(const double* __restrict__ locs, const int* __restrict__ perm, int n,
                           double* __restrict__ x, double* __restrict__ y, double* __restrict__ z) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t g = (size_t)perm[i];
    x[i] = locs[3 * g];
    y[i] = locs[3 * g + 1];
    z[i] = locs[3 * g + 2];
}

// Overwrites the plan's observation coordinates with d_locs (device, AoS, 3*n_obs)
// after checking that every point stays inside its leaf's ball; on failure nothing
// changes. The caller holds P.pts_mu alone, so no handle can enqueue meanwhile.
static void move_obs(FmmPlan& P, const double* d_locs, hipStream_t s) {
    const int No = (int)P.n_obs;
    TBEM_HIP(hipDeviceSynchronize());  // no call of any handle sharing the plan is in flight
    TBEM_HIP(hipMemsetAsync(P.d_flag.p, 0, sizeof(int), s));
    const int n_leaves = (int)P.ot.leaves.size();
    hipLaunchKernelGGL(k_obs_check, dim3(n_leaves), dim3(256), 0, s, d_locs, P.op.perm.p, P.obs_leaves.p,
                       view(P.dot), P.d_flag.p);
    TBEM_LAUNCH_CHECK();
    int n_out = 0;
    TBEM_HIP(hipMemcpyAsync(&n_out, P.d_flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    TBEM_HIP(hipStreamSynchronize(s));
    TBEM_REQUIRE(n_out == 0, TBEM_E_INVALID,
                 std::to_string(n_out) + " observation points moved outside their leaf cells: the frozen "
                 "plan no longer holds them; build a new operator");
    hipLaunchKernelGGL(k_obs_move, dim3(div_up(No, 256)), dim3(256), 0, s, d_locs, P.op.perm.p, No, P.op.x.p,
                       P.op.y.p, P.op.z.p);
    TBEM_LAUNCH_CHECK();
    // the point-target stages' close flags follow the moved points
    close_flags(P.leaf_p2p, P.obs_leaves.p, P.dot, P.op, P.sp, s);
    {
        std::lock_guard<std::mutex> lk(P.adj_mu);
        if (P.adj) close_flags(P.adj->leaf_p2p, P.src_leaves.p, P.dst, P.sp, P.op, s);
    }
    TBEM_HIP(hipStreamSynchronize(s));
    P.obs_moved = true;
    P.obs_version++;
}

static void forward(const FmmPlan& P, Work& W, const double* x, double* y, hipStream_t s) {
    W.begin(s);
    if (P.precision == TBEM_MIXED) forward_T<float>(P, W, x, y, s);
    else forward_T<double>(P, W, x, y, s);
    W.end(s);
}
static void adjoint(const FmmPlan& P, Work& W, const double* g, double* xb, hipStream_t s) {
    W.begin(s);
    if (P.precision == TBEM_MIXED) adjoint_T<float>(P, W, g, xb, s);
    else adjoint_T<double>(P, W, g, xb, s);
    W.end(s);
}

// -------------------------------------------------------------------- C boundary
template <class F>
static int guard(F&& f) {
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

static void host_roundtrip(tbem_fmm_s* h, bool adj, const double* in, double* out) {
    FmmPlan& P = *h->plan;
    Work& W = *h->work;
    const size_t nin = adj ? P.R * P.n_obs : P.R * P.n_src;
    const size_t nout = adj ? P.R * P.n_src : P.R * P.n_obs;
    if (W.xbuf.n < nin) W.xbuf.alloc(nin);
    if (W.ybuf.n < nout) W.ybuf.alloc(nout);
    W.xbuf.upload(in, nin, W.stream);
    if (adj) adjoint(P, W, W.xbuf.p, W.ybuf.p, W.stream);
    else forward(P, W, W.xbuf.p, W.ybuf.p, W.stream);
    W.ybuf.download(out, nout, W.stream);
    TBEM_HIP(hipStreamSynchronize(W.stream));
}

}  // namespace tbem

using namespace tbem;

extern "C" {

const char* tbem_last_error(void) { return g_last_error.c_str(); }

int tbem_device_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n > 0 ? 1 : 0;
}

int tbem_fmm_create(const tbem_fmm_desc* desc, tbem_fmm_t* out) {
    return guard([&] {
        TBEM_REQUIRE(out != nullptr, TBEM_E_INVALID, "null output handle");
        *out = nullptr;
        hipStream_t s;
        TBEM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        std::shared_ptr<FmmPlan> plan;
        try {
            plan = build_plan(desc, s);
        } catch (...) {
            (void)hipStreamDestroy(s);
            throw;
        }
        (void)hipStreamDestroy(s);
        auto h = std::unique_ptr<tbem_fmm_s>(new tbem_fmm_s);
        h->plan = plan;
        h->work.reset(new Work(*plan));
        *out = h.release();
    });
}

int tbem_fmm_create_from_plan(const tbem_fmm_desc* desc, const void* plan, size_t plan_bytes, tbem_fmm_t* out) {
    return guard([&] {
        TBEM_REQUIRE(out != nullptr, TBEM_E_INVALID, "null output handle");
        *out = nullptr;
        validate(desc);
        const SavedPlan sv = load_plan(desc, plan, plan_bytes);
        hipStream_t s;
        TBEM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        std::shared_ptr<FmmPlan> p;
        try {
            p = build_plan(desc, s, &sv);
        } catch (...) {
            (void)hipStreamDestroy(s);
            throw;
        }
        (void)hipStreamDestroy(s);
        auto h = std::unique_ptr<tbem_fmm_s>(new tbem_fmm_s);
        h->plan = p;
        h->work.reset(new Work(*p));
        *out = h.release();
    });
}

int tbem_fmm_plan_save(tbem_fmm_t h, void* buf, size_t cap, size_t* bytes) {
    return guard([&] {
        TBEM_REQUIRE(h && bytes, TBEM_E_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(h->mu);
        std::shared_lock<std::shared_mutex> pl(h->plan->pts_mu);
        const std::vector<char> b = save_plan(*h->plan);
        *bytes = b.size();
        if (buf) {
            TBEM_REQUIRE(cap >= b.size(), TBEM_E_INVALID, "buffer too small for the saved plan");
            std::memcpy(buf, b.data(), b.size());
        }
    });
}

int tbem_fmm_apply(tbem_fmm_t h, const double* x, double* y) {
    return guard([&] {
        TBEM_REQUIRE(h && (x || h->plan->n_src == 0) && (y || h->plan->n_obs == 0), TBEM_E_INVALID,
                     "null argument");
        std::lock_guard<std::mutex> lk(h->mu);
        std::shared_lock<std::shared_mutex> pl(h->plan->pts_mu);
        host_roundtrip(h, false, x, y);
    });
}

int tbem_fmm_apply_dev(tbem_fmm_t h, const double* d_x, double* d_y, void* stream) {
    return guard([&] {
        TBEM_REQUIRE(h, TBEM_E_INVALID, "null handle");
        std::lock_guard<std::mutex> lk(h->mu);
        std::shared_lock<std::shared_mutex> pl(h->plan->pts_mu);
        forward(*h->plan, *h->work, d_x, d_y, (hipStream_t)stream);
    });
}

int tbem_fmm_apply_adjoint(tbem_fmm_t h, const double* g, double* y) {
    return guard([&] {
        TBEM_REQUIRE(h && (g || h->plan->n_obs == 0) && (y || h->plan->n_src == 0), TBEM_E_INVALID,
                     "null argument");
        std::lock_guard<std::mutex> lk(h->mu);
        std::shared_lock<std::shared_mutex> pl(h->plan->pts_mu);
        host_roundtrip(h, true, g, y);
    });
}

int tbem_fmm_apply_adjoint_dev(tbem_fmm_t h, const double* d_g, double* d_y, void* stream) {
    return guard([&] {
        TBEM_REQUIRE(h, TBEM_E_INVALID, "null handle");
        std::lock_guard<std::mutex> lk(h->mu);
        std::shared_lock<std::shared_mutex> pl(h->plan->pts_mu);
        adjoint(*h->plan, *h->work, d_g, d_y, (hipStream_t)stream);
    });
}

int tbem_fmm_obs_grad_dev(tbem_fmm_t h, const double* d_x, const double* d_g, double* d_grad, void* stream) {
    return guard([&] {
        TBEM_REQUIRE(h, TBEM_E_INVALID, "null handle");
        std::lock_guard<std::mutex> lk(h->mu);
        std::shared_lock<std::shared_mutex> pl(h->plan->pts_mu);
        hipStream_t s = (hipStream_t)stream;
        h->work->begin(s);
        obs_grad(*h->plan, *h->work, d_x, d_g, d_grad, s);
        h->work->end(s);
    });
}

uint64_t tbem_fmm_forward_token(tbem_fmm_t h) {
    if (!h) return 0;
    std::lock_guard<std::mutex> lk(h->mu);
    return h->work->far_token;
}

int tbem_fmm_obs_grad_reuse_dev(tbem_fmm_t h, uint64_t token, const double* d_x, const double* d_g,
                                double* d_grad, void* stream) {
    return guard([&] {
        TBEM_REQUIRE(h, TBEM_E_INVALID, "null handle");
        std::lock_guard<std::mutex> lk(h->mu);
        std::shared_lock<std::shared_mutex> pl(h->plan->pts_mu);
        hipStream_t s = (hipStream_t)stream;
        h->work->begin(s);
        obs_grad(*h->plan, *h->work, d_x, d_g, d_grad, s, token);
        h->work->end(s);
    });
}

int tbem_fmm_obs_grad(tbem_fmm_t h, const double* x, const double* g, double* grad) {
    return guard([&] {
        TBEM_REQUIRE(h, TBEM_E_INVALID, "null handle");
        std::lock_guard<std::mutex> lk(h->mu);
        std::shared_lock<std::shared_mutex> pl(h->plan->pts_mu);
        FmmPlan& P = *h->plan;
        Work& W = *h->work;
        const size_t nx = P.R * P.n_src, ng = P.R * P.n_obs;
        TBEM_REQUIRE((x || nx == 0) && (g || ng == 0) && (grad || P.n_obs == 0), TBEM_E_INVALID,
                     "null argument");
        DevBuf<double> dx, dg, dgr(3 * P.n_obs);
        W.begin(W.stream);
        dx.upload(x, nx, W.stream);
        dg.upload(g, ng, W.stream);
        obs_grad(P, W, dx.p, dg.p, dgr.p, W.stream);
        W.end(W.stream);
        dgr.download(grad, 3 * P.n_obs, W.stream);
        TBEM_HIP(hipStreamSynchronize(W.stream));
    });
}

// Moves the observation points of the (shared) plan; cells, surfaces and lists stay.
int tbem_fmm_set_obs_locs(tbem_fmm_t h, const double* obs_locs) {
    return guard([&] {
        TBEM_REQUIRE(h && (obs_locs || h->plan->n_obs == 0), TBEM_E_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(h->mu);
        FmmPlan& P = *h->plan;
        if (P.n_obs == 0) return;
        std::unique_lock<std::shared_mutex> pl(P.pts_mu);
        DevBuf<double> d;
        d.upload(obs_locs, 3 * P.n_obs, h->work->stream);
        move_obs(P, d.p, h->work->stream);
    });
}

int tbem_fmm_set_obs_locs_dev(tbem_fmm_t h, const double* d_obs_locs, void* stream) {
    return guard([&] {
        TBEM_REQUIRE(h && (d_obs_locs || h->plan->n_obs == 0), TBEM_E_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(h->mu);
        FmmPlan& P = *h->plan;
        if (P.n_obs == 0) return;
        std::unique_lock<std::shared_mutex> pl(P.pts_mu);
        move_obs(P, d_obs_locs, (hipStream_t)stream);
    });
}

uint64_t tbem_fmm_obs_version(tbem_fmm_t h) {
    if (!h) return 0;
    std::shared_lock<std::shared_mutex> pl(h->plan->pts_mu);
    return h->plan->obs_version;
}

size_t tbem_fmm_n_rows(tbem_fmm_t h) { return h ? h->plan->R * h->plan->n_obs : 0; }
size_t tbem_fmm_n_cols(tbem_fmm_t h) { return h ? h->plan->R * h->plan->n_src : 0; }

int tbem_fmm_clone(tbem_fmm_t h, tbem_fmm_t* out) {
    return guard([&] {
        TBEM_REQUIRE(h && out, TBEM_E_INVALID, "null argument");
        auto c = std::unique_ptr<tbem_fmm_s>(new tbem_fmm_s);
        c->plan = h->plan;
        c->work.reset(new Work(*h->plan));
        *out = c.release();
    });
}

void tbem_fmm_destroy(tbem_fmm_t h) {
    if (!h) return;
    if (h->work) (void)hipStreamSynchronize(h->work->stream);
    delete h;
}

int tbem_fmm_info(tbem_fmm_t h, tbem_fmm_info_t* o) {
    return guard([&] {
        TBEM_REQUIRE(h && o, TBEM_E_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(h->mu);
        const FmmPlan& P = *h->plan;
        std::memset(o, 0, sizeof(*o));
        o->n_e = P.n_e;
        o->n_src_cells = P.st.n_cells;
        o->n_obs_cells = P.ot.n_cells;
        o->n_up_levels = P.st.n_levels;
        o->n_down_levels = P.ot.n_levels;
        o->p2m = P.tl.p2m.size();
        o->m2m = P.tl.m2m.size();
        o->p2p = P.tl.p2p.size();
        o->m2p = P.tl.m2p.size();
        o->p2l = P.tl.p2l.size();
        o->m2l = P.tl.m2l.size();
        o->l2l = P.tl.l2l.size();
        o->l2p = P.tl.l2p.size();
        o->p2p_pairs = P.p2p_pairs;
        o->m2l_classes = P.m2l_mfma.enabled ? P.m2l_mfma.n_classes : 0;
        o->plan_seconds = P.plan_seconds;
        o->m2l_output_bytes = (double)P.m2l_y.n * sizeof(float) + (double)P.m2l_y64.n * sizeof(double);
        const Work& W = *h->work;
        double wb = 0;
        for (const DevBuf<double>* b : {&W.dens, &W.mult, &W.locals, &W.out_sorted, &W.xbuf, &W.ybuf, &W.gsort,
                                        &W.lbar, &W.cbar, &W.mbar, &W.ebar, &W.grad_sorted, &W.e2, &W.fld, &W.ty,
                                        &W.wpart, &W.wrev})
            wb += (double)b->n * sizeof(double);
        for (const DevBuf<float>* b : {&W.chk_up, &W.m2l_w}) wb += (double)b->n * sizeof(float);
        o->work_bytes = wb;
        o->leaf_wave_items = P.leaf_split.n_work;
        o->leaf_chunked = P.leaf_split.n_parts;
        o->leaf_block = P.leaf_split.n_block_leaves;
        o->leaf_symmetric = P.sym.n_leaves;
        o->symmetric_pairs = P.sym.pairs;
    });
}

int tbem_fmm_c2e(tbem_fmm_t h, int which, size_t level, double* out) {
    return guard([&] {
        TBEM_REQUIRE(h && out, TBEM_E_INVALID, "null argument");
        const FmmPlan& P = *h->plan;
        const auto& v = which ? P.h_dc2e : P.h_uc2e;
        const size_t bb = (size_t)P.blk * P.blk;
        TBEM_REQUIRE((level + 1) * bb <= v.size(), TBEM_E_INVALID, "level out of range");
        std::memcpy(out, v.data() + level * bb, bb * sizeof(double));
    });
}

int tbem_fmm_pairs(tbem_fmm_t h, int which, double* out, size_t* n) {
    return guard([&] {
        TBEM_REQUIRE(h && n && which >= 0 && which <= 3, TBEM_E_INVALID, "bad argument");
        const FmmPlan& P = *h->plan;
        const auto& v = which == 0 ? P.tl.p2p : which == 1 ? P.tl.m2l : which == 2 ? P.tl.m2p : P.tl.p2l;
        *n = v.size();
        if (!out) return;
        for (size_t i = 0; i < v.size(); i++) {
            const HostTree* ts[2] = {&P.ot, &P.st};
            for (int k = 0; k < 2; k++) {
                const HostTree& t = *ts[k];
                const int c = v[i][k];
                double* o = out + 14 * i + 7 * k;
                o[0] = t.level[c];
                o[1] = t.cx[c]; o[2] = t.cy[c]; o[3] = t.cz[c];
                for (int dd = 0; dd < 3; dd++) o[4 + dd] = t.level_hw[3 * t.level[c] + dd];
            }
        }
    });
}

int tbem_fmm_octree(tbem_fmm_t h, int which, double* cells, long* indices, size_t* n_cells, size_t* n_idx) {
    return guard([&] {
        TBEM_REQUIRE(h && n_cells && n_idx, TBEM_E_INVALID, "null argument");
        const HostTree& t = which ? h->plan->ot : h->plan->st;
        // depth-first pre-order over the breadth-first cells
        std::vector<int> order;
        order.reserve(t.n_cells);
        std::vector<int> stack{0};
        while (!stack.empty() && t.n_cells > 0) {
            int c = stack.back();
            stack.pop_back();
            order.push_back(c);
            for (int k = t.n_child[c] - 1; k >= 0; k--) stack.push_back(t.first_child[c] + k);
        }
        size_t ni = 0;
        for (int c: order) ni += t.count[c];
        *n_cells = order.size();
        *n_idx = ni;
        size_t off = 0;
        for (size_t k = 0; k < order.size(); k++) {
            const int c = order[k];
            if (cells) {
                double* o = cells + 10 * k;
                o[0] = t.level[c];
                o[1] = t.cx[c]; o[2] = t.cy[c]; o[3] = t.cz[c];
                for (int dd = 0; dd < 3; dd++) o[4 + dd] = t.level_hw[3 * t.level[c] + dd];
                o[7] = t.leaf[c] ? 1.0 : 0.0;
                o[8] = t.count[c];
                o[9] = (double)off;
            }
            if (indices) {
                // ascending original index inside every cell (leaves are; internal cells re-sorted)
                std::vector<int> ids(t.perm.begin() + t.begin[c], t.perm.begin() + t.begin[c] + t.count[c]);
                std::sort(ids.begin(), ids.end());
                for (size_t q = 0; q < ids.size(); q++) indices[off + q] = ids[q];
            }
            off += t.count[c];
        }
    });
}

int tbem_fmm_stage_times(tbem_fmm_t h, double* ms) {
    return guard([&] {
        TBEM_REQUIRE(h && ms, TBEM_E_INVALID, "null argument");
        Work& W = *h->work;
        TBEM_REQUIRE(W.timed, TBEM_E_INVALID, "no timed apply yet");
        TBEM_HIP(hipEventSynchronize(W.ev[7]));
        for (int k = 0; k < 7; k++) {
            float t = 0;
            TBEM_HIP(hipEventElapsedTime(&t, W.ev[k], W.ev[k + 1]));
            ms[k] = t;
        }
        if (W.split_last) {  // the leaf stage ran as P2P (side stream) + M2P/L2P
            float t = 0;
            TBEM_HIP(hipEventElapsedTime(&t, W.ev[10], W.ev[11]));
            ms[5] += t;
        }
    });
}

int tbem_fmm_set_overlap(tbem_fmm_t h, int on) {
    return guard([&] {
        TBEM_REQUIRE(h, TBEM_E_INVALID, "null handle");
        std::lock_guard<std::mutex> lk(h->mu);
        h->work->overlap = on != 0;
    });
}

int tbem_fmm_m2l_ms(tbem_fmm_t h, double* ms) {
    return guard([&] {
        TBEM_REQUIRE(h && ms, TBEM_E_INVALID, "null argument");
        Work& W = *h->work;
        TBEM_REQUIRE(W.timed, TBEM_E_INVALID, "no timed apply yet");
        TBEM_HIP(hipEventSynchronize(W.ev[7]));
        float t = 0;
        TBEM_HIP(hipEventElapsedTime(&t, W.ev[8], W.ev[9]));
        *ms = t;
    });
}

long tbem_surface(size_t order, double* out) {
    if (order == 0) return 0;
    auto p = sphere_points(order);
    if (out) std::memcpy(out, p.data(), p.size() * sizeof(double));
    return (long)(p.size() / 3);
}

int tbem_pinv(const double* A, int n, double threshold, double* out) {
    return guard([&] {
        TBEM_REQUIRE(A && out && n > 0, TBEM_E_INVALID, "bad argument");
        DevBuf<double> dA, dO((size_t)n * n);
        dA.upload(A, (size_t)n * n);
        gpu_pinv_batched(dA.p, 1, n, threshold > 0 ? threshold : 1e-16, dO.p, 0);
        dO.download(out, (size_t)n * n);
        TBEM_HIP(hipDeviceSynchronize());
    });
}

}  // extern "C"

// ------------------------------------------------------------------ kernel parity
namespace tbem {
__global__ void k_kernel_eval(int kid, KConst kc, int n, const double* obs, const double* src,
                              const double* nobs, const double* nsrc, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int R = kernel_dim(kid);
    tensor_ref_dispatch(kid, obs + 3 * i, src + 3 * i, nobs + 3 * i, nsrc + 3 * i, kc, out + (size_t)R * R * i);
}
}  // namespace tbem

namespace tbem {
// sqrt(a) and a / b as the device code of tensor_ref evaluates them
__global__ void k_fp64_sqrt_div(int n, const double* __restrict__ a, const double* __restrict__ b,
                                double* __restrict__ sq, double* __restrict__ q) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sq[i] = sqrt(a[i]);
    q[i] = a[i] / b[i];
}
}  // namespace tbem

extern "C" int tbem_fp64_sqrt_div(size_t n, const double* a, const double* b, double* sqrt_a, double* a_div_b) {
    return guard([&] {
        TBEM_REQUIRE(a && b && sqrt_a && a_div_b, TBEM_E_INVALID, "null argument");
        TBEM_REQUIRE(n < (size_t(1) << 30), TBEM_E_UNSUPPORTED, "too many values");
        if (n == 0) return;
        DevBuf<double> da, db, ds(n), dq(n);
        da.upload(a, n);
        db.upload(b, n);
        hipLaunchKernelGGL(k_fp64_sqrt_div, dim3(div_up(n, 256)), dim3(256), 0, 0, (int)n, da.p, db.p, ds.p, dq.p);
        TBEM_LAUNCH_CHECK();
        ds.download(sqrt_a, n);
        dq.download(a_div_b, n);
        TBEM_HIP(hipDeviceSynchronize());
    });
}

extern "C" int tbem_kernel_eval(int kernel, double mu, double nu, size_t n, const double* obs,
                                const double* src, const double* nobs, const double* nsrc, double* out) {
    return guard([&] {
        TBEM_REQUIRE(kernel_valid(kernel), TBEM_E_INVALID, "unknown kernel id");
        TBEM_REQUIRE(obs && src && nobs && nsrc && out, TBEM_E_INVALID, "null argument");
        if (n == 0) return;
        const int R = kernel_dim(kernel);
        DevBuf<double> a, b, c, d, o((size_t)R * R * n);
        a.upload(obs, 3 * n); b.upload(src, 3 * n); c.upload(nobs, 3 * n); d.upload(nsrc, 3 * n);
        hipLaunchKernelGGL(k_kernel_eval, dim3(div_up(n, 256)), dim3(256), 0, 0, kernel,
                           make_kconst(mu, nu), (int)n, a.p, b.p, c.p, d.p, o.p);
        TBEM_LAUNCH_CHECK();
        o.download(out, (size_t)R * R * n);
        TBEM_HIP(hipDeviceSynchronize());
    });
}
