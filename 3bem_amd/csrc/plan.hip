// Plan construction: translation surfaces, device uploads, dual-tree traversal.
// (The octree itself is built on the GPU in octree.hip.)
#include <algorithm>
#include <cmath>
#include <cstring>

#include "plan.h"

namespace tbem {

// ----------------------------------------------------------------------- surfaces
std::vector<double> sphere_points(size_t order) {
#pragma clang fp contract(off)
    std::vector<double> pts;
    double a = 4 * kPi / order;
    double d = std::sqrt(a);
    size_t Mt = (size_t)std::round(kPi / d);
    double dt = kPi / Mt;
    double dp = a / dt;
    for (size_t m = 0; m < Mt; m++) {
        double th = kPi * (m + 0.5) / Mt;
        size_t Mp = (size_t)std::round(2 * kPi * std::sin(th) / dp);
        for (size_t k = 0; k < Mp; k++) {
            double ph = 2 * kPi * k / Mp;
            pts.push_back(std::sin(th) * std::cos(ph));
            pts.push_back(std::sin(th) * std::sin(ph));
            pts.push_back(std::cos(th));
        }
    }
    return pts;
}

Surfaces make_surfaces(size_t order, double d) {
#pragma clang fp contract(off)
    Surfaces s;
    s.unit = sphere_points(order);
    s.n_e = (int)(s.unit.size() / 3);
    const double r_far = 4.0 - std::sqrt(2.0) - 2 * d;
    s.inner.resize(s.unit.size());
    s.outer.resize(s.unit.size());
    for (size_t i = 0; i < s.unit.size(); i++) {
        s.inner[i] = 0.3 * s.unit[i];
        s.outer[i] = r_far * s.unit[i];
    }
    return s;
}

// ------------------------------------------------------------------------- uploads
void DevTree::upload(const HostTree& t, hipStream_t s) {
    n_cells = t.n_cells;
    n_levels = t.n_levels;
    cx.upload(t.cx, s); cy.upload(t.cy, s); cz.upload(t.cz, s);
    level_r.upload(t.level_r, s);
    level.upload(t.level, s); begin.upload(t.begin, s); count.upload(t.count, s);
}

__global__ void k_gather_points(const double* __restrict__ locs, const double* __restrict__ nrm,
                                const double* __restrict__ w, const int* __restrict__ perm, int n,
                                double* x, double* y, double* z, double* nx, double* ny, double* nz,
                                double* wo) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    size_t g = (size_t)perm[i];
    x[i] = locs[3 * g]; y[i] = locs[3 * g + 1]; z[i] = locs[3 * g + 2];
    nx[i] = nrm[3 * g]; ny[i] = nrm[3 * g + 1]; nz[i] = nrm[3 * g + 2];
    if (wo) wo[i] = w ? w[g] : 1.0;
}

void DevPoints::upload(const HostTree& t, const double* locs, const double* normals,
                       const double* wts, hipStream_t s) {
    int n = t.n_pts;
    perm.upload(t.perm, s);
    x.alloc(n); y.alloc(n); z.alloc(n); nx.alloc(n); ny.alloc(n); nz.alloc(n); w.alloc(n);
    if (n == 0) return;
    DevBuf<double> dl, dn, dw;
    dl.upload(locs, 3 * (size_t)n, s);
    dn.upload(normals, 3 * (size_t)n, s);
    if (wts) dw.upload(wts, n, s);
    hipLaunchKernelGGL(k_gather_points, dim3(div_up(n, 256)), dim3(256), 0, s, dl.p, dn.p,
                       wts ? dw.p : nullptr, perm.p, n, x.p, y.p, z.p, nx.p, ny.p, nz.p, w.p);
    TBEM_LAUNCH_CHECK();
    TBEM_HIP(hipStreamSynchronize(s));
}

void DevList::upload(const HostList& l, hipStream_t s) {
    n_items = (int)l.items.size();
    items.upload(l.items, s);
    ptr.upload(l.ptr, s);
    cell.upload(l.cell, s);
    rbeg.upload(l.rbeg, s);
    rcnt.upload(l.rcnt, s);
    rpre.upload(l.rpre, s);
}

void cells_to_ranges(HostList& l, const HostTree& t) {
    std::vector<int> ptr(1, 0);
    std::vector<std::pair<int, int>> seg;
    l.rbeg.clear();
    l.rcnt.clear();
    for (size_t k = 0; k + 1 < l.ptr.size(); k++) {
        seg.clear();
        for (int e = l.ptr[k]; e < l.ptr[k + 1]; e++) {
            const int c = l.cell[e];
            if (t.count[c] > 0) seg.push_back({t.begin[c], t.count[c]});
        }
        std::sort(seg.begin(), seg.end());
        for (auto& r: seg) {
            if (!l.rbeg.empty() && (int)l.rbeg.size() > ptr.back() && l.rbeg.back() + l.rcnt.back() == r.first) {
                l.rcnt.back() += r.second;
            } else {
                l.rbeg.push_back(r.first);
                l.rcnt.push_back(r.second);
            }
        }
        ptr.push_back((int)l.rbeg.size());
    }
    l.ptr = std::move(ptr);
    l.cell.clear();
    l.rpre.assign(l.rbeg.size(), 0);
    for (size_t k = 0; k + 1 < l.ptr.size(); k++) {
        int acc = 0;
        for (int e = l.ptr[k]; e < l.ptr[k + 1]; e++) {
            l.rpre[e] = acc;
            acc += l.rcnt[e];
        }
    }
}

// ------------------------------------------------------------------ dual traversal
namespace {
struct Traversal {
    const HostTree& ot;
    const HostTree& st;
    double mac;
    int n_e;
    bool small_cells;
    TaskLists& tl;

    // cpp/fmm.cpp:577-630
    void dual(int o, int s) {
#pragma clang fp contract(off)
        const double rs = st.level_r[st.level[s]], ro = ot.level_r[ot.level[o]];
        const double rmax = std::max(rs, ro), rmin = std::min(rs, ro);
        const double dx = ot.cx[o] - st.cx[s], dy = ot.cy[o] - st.cy[s], dz = ot.cz[o] - st.cz[s];
        const double sep = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (rmax + mac * rmin <= mac * sep) {
            const bool small_src = st.count[s] <= n_e, small_obs = ot.count[o] <= n_e;
            if (small_cells && small_obs && small_src) tl.p2p.push_back({o, s});
            else if (small_cells && small_obs) tl.m2p.push_back({o, s});
            else if (small_cells && small_src) tl.p2l.push_back({o, s});
            else tl.m2l.push_back({o, s});
            return;
        }
        const bool ol = ot.leaf[o], sl = st.leaf[s];
        if (ol && sl) {
            tl.p2p.push_back({o, s});
            return;
        }
        const bool split_src = (ot.level[o] > st.level[s] && !sl) || ol;
        if (split_src) {
            for (int k = 0; k < st.n_child[s]; k++) dual(o, st.first_child[s] + k);
        } else {
            for (int k = 0; k < ot.n_child[o]; k++) dual(ot.first_child[o] + k, s);
        }
    }
};
}  // namespace

void dual_tree_traversal(const HostTree& ot, const HostTree& st, double mac, int n_e,
                         bool small_cells, TaskLists& tl) {
    // upward (cpp/fmm.cpp:560-575) and downward (cpp/fmm.cpp:632-648) lists, as sets
    for (int c = 0; c < st.n_cells; c++) (st.leaf[c] ? tl.p2m : tl.m2m).push_back(c);
    for (int c = 0; c < ot.n_cells; c++) (ot.leaf[c] ? tl.l2p : tl.l2l).push_back(c);
    Traversal tr{ot, st, mac, n_e, small_cells, tl};
    tr.dual(0, 0);
}

}  // namespace tbem
