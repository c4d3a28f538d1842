// Matrix-free direct N-body operator on the GPU.
//
// Replaces make_direct_nbody_operator + DenseOperator::apply
// (cpp/nbody_operator.h:143-152, cpp/dense_operator.cpp:38-46) and nbody_eval
// (cpp/nbody_operator.h:117-141): y_i = sum_j w_j K(x_i <- y_j) x_j without the
// O(N^2) matrix. Points keep their input order and are cut into flat "cells":
// 128-point target chunks, each streaming every source chunk through the same
// leaf-stage kernel the FMM's P2P uses (LDS tiles, slice-split lanes).
#include <algorithm>
#include <cstring>
#include <mutex>

#include "stages.h"

namespace tbem {

static constexpr int kTargetChunk = 128;

// Scratch of tbem_nbody_apply_noalloc: the weighted source densities in sorted
// order, R * n_src doubles, at most kNoallocDoubles (32 MiB). Every GPU the library
// runs on holds one copy until the process ends. In host code `scratch` names a
// host-side stand-in, not device memory: kernels receive the device address that
// hipGetSymbolAddress gives for the current device (noalloc_scratch()).
constexpr size_t kNoallocDoubles = size_t(4) << 20;
static __device__ __align__(256) double scratch[kNoallocDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on the first
// call per device (with that device current) and kept.
static double* noalloc_scratch(int dev) {
    TBEM_REQUIRE(dev >= 0 && dev < kMaxDevices, TBEM_E_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        TBEM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        TBEM_REQUIRE(ptr != nullptr, TBEM_E_HIP, "no device address for the no-allocation scratch");
        g_scratch[dev] = static_cast<double*>(ptr);
    }
    return g_scratch[dev];
}

// a flat tree of fixed-size chunks over n points in input order
static HostTree flat_tree(const double* locs, int n, int chunk) {
    HostTree t;
    t.n_pts = n;
    const int nc = std::max(1, (n + chunk - 1) / chunk);
    t.n_cells = nc;
    t.n_levels = 1;
    t.level_begin = {0, nc};
    t.level_hw = {0, 0, 0};
    t.level_r = {0};
    t.perm.resize(n);
    for (int i = 0; i < n; i++) t.perm[i] = i;
    for (int c = 0; c < nc; c++) {
        const int b = c * chunk, e = std::min(n, b + chunk);
        double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (int i = b; i < e; i++)
            for (int d = 0; d < 3; d++) {
                const double v = locs[3 * (size_t)i + d];
                lo[d] = (i == b) ? v : std::min(lo[d], v);
                hi[d] = (i == b) ? v : std::max(hi[d], v);
            }
        t.level.push_back(0);
        t.begin.push_back(b);
        t.count.push_back(std::max(0, e - b));
        t.first_child.push_back(-1);
        t.n_child.push_back(0);
        t.parent.push_back(-1);
        t.octant.push_back(0);
        t.leaf.push_back(1);
        t.cx.push_back(0.5 * (lo[0] + hi[0]));
        t.cy.push_back(0.5 * (lo[1] + hi[1]));
        t.cz.push_back(0.5 * (lo[2] + hi[2]));
        t.leaves.push_back(c);
    }
    return t;
}

// every target chunk streams the whole source array as one range
static HostList all_pairs(const HostTree& tgt, int n_src) {
    HostList l;
    l.items = tgt.leaves;
    l.ptr.assign(1, 0);
    for (size_t k = 0; k < tgt.leaves.size(); k++) {
        if (n_src > 0) {
            l.rbeg.push_back(0);
            l.rcnt.push_back(n_src);
            l.rpre.push_back(0);
        }
        l.ptr.push_back((int)l.rbeg.size());
    }
    return l;
}

struct NbodyPlan {
    int kid = 0, R = 1, precision = TBEM_FP64;
    KConst kc{};
    int n_obs = 0, n_src = 0;
    HostTree obs_t, src_t;  // 128-point target chunks of each side
    DevTree d_obs_t, d_src_t;
    DevPoints op, sp;
    DevList fwd, adj;
    DevBuf<int> obs_leaves, src_leaves;
    DevBuf<double> dens, gsort, xb, yb, outs;
    int device = 0;  // the device the handle's buffers live on
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;  // end of the last call's work (buffers are per handle)
    ~NbodyPlan() {
        if (done) (void)hipEventDestroy(done);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

}  // namespace tbem

struct tbem_nbody_s {
    tbem::NbodyPlan p;
};

namespace tbem {
void set_last_error(const std::string& m);

template <class F>
static int nb_guard(F&& f) {
    try {
        f();
        return TBEM_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return TBEM_E_INVALID;
    }
}

static void nb_forward_impl(const NbodyPlan& P, const double* x, double* y, double* dens, hipStream_t s);
static void nb_adjoint_impl(NbodyPlan& P, const double* g, double* y, hipStream_t s);

static void nb_forward(NbodyPlan& P, const double* x, double* y, hipStream_t s) {
    TBEM_HIP(hipStreamWaitEvent(s, P.done, 0));
    nb_forward_impl(P, x, y, P.dens.p, s);
    TBEM_HIP(hipEventRecord(P.done, s));
}

static void nb_adjoint(NbodyPlan& P, const double* g, double* y, hipStream_t s) {
    TBEM_HIP(hipStreamWaitEvent(s, P.done, 0));
    nb_adjoint_impl(P, g, y, s);
    TBEM_HIP(hipEventRecord(P.done, s));
}

// dens: R * n_src doubles of scratch for the weighted densities
static void nb_forward_impl(const NbodyPlan& P, const double* x, double* y, double* dens, hipStream_t s) {
    if (P.n_obs == 0) return;
    if (P.n_src == 0) {
        launch_fill(y, (size_t)P.R * P.n_obs, 0.0, s);
        return;
    }
    launch_gather_dens(x, P.sp.perm.p, P.sp.w.p, P.n_src, P.R, dens, s);
    LeafStage f{};
    f.n_leaves = (int)P.obs_t.leaves.size();
    f.leaves = P.obs_leaves.p;
    f.tt = view(P.d_obs_t); f.tp = view(P.op); f.n_tgt = P.n_obs;
    f.rp_ptr = P.fwd.ptr.p; f.rp_beg = P.fwd.rbeg.p; f.rp_cnt = P.fwd.rcnt.p; f.rp_pre = P.fwd.rpre.p;
    f.ps = view(P.sp); f.pdens = dens; f.n_ps = P.n_src;
    f.out_sorted = y;  // identity permutation: write straight into y
    if (P.precision == TBEM_MIXED) launch_leaf_stage<float>(P.kid, false, f, P.kc, s, false);
    else launch_leaf_stage<double>(P.kid, false, f, P.kc, s, false);
}

static void nb_adjoint_impl(NbodyPlan& P, const double* g, double* y, hipStream_t s) {
    if (P.n_src == 0) return;
    if (P.n_obs == 0) {
        launch_fill(y, (size_t)P.R * P.n_src, 0.0, s);
        return;
    }
    LeafStage f{};
    f.n_leaves = (int)P.src_t.leaves.size();
    f.leaves = P.src_leaves.p;
    f.tt = view(P.d_src_t); f.tp = view(P.sp); f.t_weight = P.sp.w.p; f.n_tgt = P.n_src;
    f.rp_ptr = P.adj.ptr.p; f.rp_beg = P.adj.rbeg.p; f.rp_cnt = P.adj.rcnt.p; f.rp_pre = P.adj.rpre.p;
    f.ps = view(P.op); f.pdens = g; f.n_ps = P.n_obs;
    f.out_sorted = y;
    if (P.precision == TBEM_MIXED) launch_leaf_stage<float>(P.kid, true, f, P.kc, s, false);
    else launch_leaf_stage<double>(P.kid, true, f, P.kc, s, false);
}

}  // namespace tbem

using namespace tbem;

extern "C" {

int tbem_nbody_create(const tbem_nbody_desc* d, tbem_nbody_t* out) {
    return nb_guard([&] {
        TBEM_REQUIRE(d && out, TBEM_E_INVALID, "null argument");
        *out = nullptr;
        TBEM_REQUIRE(kernel_valid(d->kernel), TBEM_E_INVALID, "unknown kernel id");
        TBEM_REQUIRE(d->precision == TBEM_FP64 || d->precision == TBEM_MIXED, TBEM_E_INVALID,
                     "precision must be TBEM_FP64 or TBEM_MIXED");
        TBEM_REQUIRE(d->n_obs == 0 || (d->obs_locs && d->obs_normals), TBEM_E_INVALID, "null observation arrays");
        TBEM_REQUIRE(d->n_src == 0 || (d->src_locs && d->src_normals && d->src_weights), TBEM_E_INVALID,
                     "null source arrays");
        TBEM_REQUIRE(d->n_obs < (1u << 30) && d->n_src < (1u << 30), TBEM_E_UNSUPPORTED, "too many points");
        auto h = std::unique_ptr<tbem_nbody_s>(new tbem_nbody_s);
        NbodyPlan& P = h->p;
        TBEM_HIP(hipGetDevice(&P.device));
        TBEM_HIP(hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking));
        TBEM_HIP(hipEventCreateWithFlags(&P.done, hipEventDisableTiming));
        hipStream_t s = P.stream;
        P.kid = d->kernel;
        P.R = kernel_dim(d->kernel);
        P.precision = d->precision;
        P.kc = make_kconst(d->shear_modulus, d->poisson_ratio);
        P.n_obs = (int)d->n_obs;
        P.n_src = (int)d->n_src;
        P.obs_t = flat_tree(d->obs_locs, P.n_obs, kTargetChunk);
        P.src_t = flat_tree(d->src_locs, P.n_src, kTargetChunk);
        P.d_obs_t.upload(P.obs_t, s);
        P.d_src_t.upload(P.src_t, s);
        P.op.upload(P.obs_t, d->obs_locs, d->obs_normals, nullptr, s);
        P.sp.upload(P.src_t, d->src_locs, d->src_normals, d->src_weights, s);
        P.fwd.upload(all_pairs(P.obs_t, P.n_src), s);
        P.adj.upload(all_pairs(P.src_t, P.n_obs), s);
        P.obs_leaves.upload(P.obs_t.leaves, s);
        P.src_leaves.upload(P.src_t.leaves, s);
        P.dens.alloc((size_t)P.R * P.n_src);
        TBEM_HIP(hipEventRecord(P.done, s));
        TBEM_HIP(hipStreamSynchronize(s));
        *out = h.release();
    });
}

int tbem_nbody_apply_dev(tbem_nbody_t h, const double* d_x, double* d_y, void* stream) {
    return nb_guard([&] {
        TBEM_REQUIRE(h, TBEM_E_INVALID, "null handle");
        nb_forward(h->p, d_x, d_y, (hipStream_t)stream);
    });
}

// tbem_nbody_apply_dev without allocating: the weighted densities go to the
// device's static scratch instead of the handle's buffer. Calls that overlap on
// one device share that scratch and race.
int tbem_nbody_apply_noalloc(tbem_nbody_t h, const double* d_x, double* d_y, void* stream) {
    return nb_guard([&] {
        TBEM_REQUIRE(h, TBEM_E_INVALID, "null handle");
        const NbodyPlan& P = h->p;
        const size_t need = (size_t)P.R * P.n_src;
        TBEM_REQUIRE(need <= kNoallocDoubles, TBEM_E_UNSUPPORTED,
                     "tbem_nbody_apply_noalloc: " + std::to_string(need) +
                         " source density values exceed its scratch of " + std::to_string(kNoallocDoubles));
        TBEM_REQUIRE((d_x || need == 0) && (d_y || P.n_obs == 0), TBEM_E_INVALID, "null argument");
        int dev = -1;
        TBEM_HIP(hipGetDevice(&dev));
        TBEM_REQUIRE(dev == P.device, TBEM_E_INVALID,
                     "tbem_nbody_apply_noalloc: the current device is not the handle's");
        nb_forward_impl(P, d_x, d_y, noalloc_scratch(dev), (hipStream_t)stream);
    });
}

int tbem_nbody_apply(tbem_nbody_t h, const double* x, double* y) {
    return nb_guard([&] {
        TBEM_REQUIRE(h, TBEM_E_INVALID, "null handle");
        NbodyPlan& P = h->p;
        const size_t nin = (size_t)P.R * P.n_src, nout = (size_t)P.R * P.n_obs;
        TBEM_REQUIRE((x || nin == 0) && (y || nout == 0), TBEM_E_INVALID, "null argument");
        if (P.xb.n < nin) P.xb.alloc(nin);
        if (P.yb.n < nout) P.yb.alloc(nout);
        P.xb.upload(x, nin, P.stream);
        nb_forward(P, P.xb.p, P.yb.p, P.stream);
        P.yb.download(y, nout, P.stream);
        TBEM_HIP(hipStreamSynchronize(P.stream));
    });
}

int tbem_nbody_apply_adjoint(tbem_nbody_t h, const double* g, double* y) {
    return nb_guard([&] {
        TBEM_REQUIRE(h, TBEM_E_INVALID, "null handle");
        NbodyPlan& P = h->p;
        const size_t nin = (size_t)P.R * P.n_obs, nout = (size_t)P.R * P.n_src;
        TBEM_REQUIRE((g || nin == 0) && (y || nout == 0), TBEM_E_INVALID, "null argument");
        if (P.xb.n < nin) P.xb.alloc(nin);
        if (P.yb.n < nout) P.yb.alloc(nout);
        P.xb.upload(g, nin, P.stream);
        nb_adjoint(P, P.xb.p, P.yb.p, P.stream);
        P.yb.download(y, nout, P.stream);
        TBEM_HIP(hipStreamSynchronize(P.stream));
    });
}

size_t tbem_nbody_n_rows(tbem_nbody_t h) { return h ? (size_t)h->p.R * h->p.n_obs : 0; }
size_t tbem_nbody_n_cols(tbem_nbody_t h) { return h ? (size_t)h->p.R * h->p.n_src : 0; }

void tbem_nbody_destroy(tbem_nbody_t h) {
    if (!h) return;
    (void)hipStreamSynchronize(h->p.stream);
    delete h;
}

}  // extern "C"
