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

// Items are independent: host threads merge each item's ranges, then the lists are
// concatenated in item order.
void cells_to_ranges(HostList& l, const HostTree& t) {
    const size_t ni = l.ptr.size() - 1;
    std::vector<std::vector<std::pair<int, int>>> per(ni);
    parallel_for(ni, [&](size_t k) {
        std::vector<std::pair<int, int>>& seg = per[k];
        for (int e = l.ptr[k]; e < l.ptr[k + 1]; e++) {
            const int c = l.cell[e];
            if (t.count[c] > 0) seg.push_back({t.begin[c], t.count[c]});
        }
        std::sort(seg.begin(), seg.end());
        size_t w = 0;  // merge touching ranges in place
        for (size_t r = 0; r < seg.size(); r++) {
            if (w > 0 && seg[w - 1].first + seg[w - 1].second == seg[r].first) seg[w - 1].second += seg[r].second;
            else seg[w++] = seg[r];
        }
        seg.resize(w);
    });
    std::vector<int> ptr(ni + 1, 0);
    for (size_t k = 0; k < ni; k++) ptr[k + 1] = ptr[k] + (int)per[k].size();
    l.rbeg.assign(ptr[ni], 0);
    l.rcnt.assign(ptr[ni], 0);
    l.rpre.assign(ptr[ni], 0);
    parallel_for(ni, [&](size_t k) {
        int acc = 0;
        for (size_t r = 0; r < per[k].size(); r++) {
            const int e = ptr[k] + (int)r;
            l.rbeg[e] = per[k][r].first;
            l.rcnt[e] = per[k][r].second;
            l.rpre[e] = acc;
            acc += per[k][r].second;
        }
    });
    l.ptr = std::move(ptr);
    l.cell.clear();
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
    // pairs reached after `stop_depth` splits are not descended here but recorded as
    // slots, with the list sizes at that point (stop_depth < 0: descend everything)
    int stop_depth = -1;
    struct Slot {
        int o, s;
        size_t mark[4];  // p2p, m2l, m2p, p2l sizes when the slot was reached
    };
    std::vector<Slot>* slots = nullptr;

    // cpp/fmm.cpp:577-630
    void dual(int o, int s, int depth = 0) {
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
        if (depth == stop_depth) {
            slots->push_back({o, s, {tl.p2p.size(), tl.m2l.size(), tl.m2p.size(), tl.p2l.size()}});
            return;
        }
        const bool split_src = (ot.level[o] > st.level[s] && !sl) || ol;
        if (split_src) {
            for (int k = 0; k < st.n_child[s]; k++) dual(o, st.first_child[s] + k, depth + 1);
        } else {
            for (int k = 0; k < ot.n_child[o]; k++) dual(ot.first_child[o] + k, s, depth + 1);
        }
    }
};
}  // namespace

void dual_tree_traversal(const HostTree& ot, const HostTree& st, double mac, int n_e,
                         bool small_cells, TaskLists& tl) {
    // upward (cpp/fmm.cpp:560-575) and downward (cpp/fmm.cpp:632-648) lists, as sets
    for (int c = 0; c < st.n_cells; c++) (st.leaf[c] ? tl.p2m : tl.m2m).push_back(c);
    for (int c = 0; c < ot.n_cells; c++) (ot.leaf[c] ? tl.l2p : tl.l2l).push_back(c);
    // The recursion's first splits run here, the subtrees below them on host threads;
    // each list is then spliced back in the order one recursion would produce it.
    std::vector<Traversal::Slot> slots;
    Traversal top{ot, st, mac, n_e, small_cells, tl};
    top.stop_depth = plan_threads() > 1 ? 4 : -1;  // one thread: the plain recursion
    top.slots = &slots;
    top.dual(0, 0);
    if (slots.empty()) return;
    std::vector<TaskLists> part(slots.size());
    parallel_for(slots.size(), [&](size_t k) {
        Traversal tr{ot, st, mac, n_e, small_cells, part[k]};
        tr.dual(slots[k].o, slots[k].s);
    });
    auto splice = [&](int which, std::vector<std::array<int, 2>> TaskLists::*member) {
        std::vector<std::array<int, 2>>& direct = tl.*member;
        size_t total = direct.size();
        for (const auto& pt : part) total += (pt.*member).size();
        std::vector<std::array<int, 2>> out;
        out.reserve(total);
        size_t from = 0;
        for (size_t k = 0; k < slots.size(); k++) {
            out.insert(out.end(), direct.begin() + from, direct.begin() + slots[k].mark[which]);
            from = slots[k].mark[which];
            out.insert(out.end(), (part[k].*member).begin(), (part[k].*member).end());
        }
        out.insert(out.end(), direct.begin() + from, direct.end());
        direct = std::move(out);
    };
    splice(0, &TaskLists::p2p);
    splice(1, &TaskLists::m2l);
    splice(2, &TaskLists::m2p);
    splice(3, &TaskLists::p2l);
}

}  // namespace tbem
