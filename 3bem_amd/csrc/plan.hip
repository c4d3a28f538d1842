// Plan construction: GPU Morton-key octree, translation surfaces, dual-tree traversal.
//
// Octree semantics (cpp/octree.cpp:95-180, cpp/geometry.cpp:12-95):
//  * the root's bounds are the tight box expanded by 1e-5 * max half width, but its
//    children split the tight (unexpanded) box;
//  * a child's box halves the half width and moves the center by +-half width,
//    child id = [x > cx] << 2 | [y > cy] << 1 | [z > cz] (strict >, x most significant);
//  * a cell is a leaf when it holds <= min_pts_per_cell points, or (below the root)
//    when all of its points coincide.
// Every point's path is a pure function of its coordinates and of centers obtained
// by replaying those FP64 operations, so each point gets a 63-bit key (21 levels,
// 3 bits each) on the GPU; a stable radix sort of (key, index) puts every cell of
// the reference tree on a contiguous key run, and the host reads the cells off the
// runs. Point order inside a leaf is then restored to ascending original index.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "plan.h"

namespace tbem {

// ------------------------------------------------------------------ GPU octree keys
__global__ void k_bbox_partial(const double* __restrict__ pts, int n, double* __restrict__ part) {
    __shared__ double s[6][256];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int d = 0; d < 3; d++) {
            double v = pts[3 * (size_t)i + d];
            lo[d] = fmin(lo[d], v);
            hi[d] = fmax(hi[d], v);
        }
    }
    for (int d = 0; d < 3; d++) {
        s[d][threadIdx.x] = lo[d];
        s[3 + d][threadIdx.x] = hi[d];
    }
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int d = 0; d < 3; d++) {
                s[d][threadIdx.x] = fmin(s[d][threadIdx.x], s[d][threadIdx.x + w]);
                s[3 + d][threadIdx.x] = fmax(s[3 + d][threadIdx.x], s[3 + d][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 6; q++) part[6 * blockIdx.x + q] = s[q][0];
}

struct Box3 {
    double c[3], hw[3];
};

__global__ void k_morton_keys(const double* __restrict__ pts, int n, Box3 root,
                              unsigned long long* __restrict__ keys, int* __restrict__ idx) {
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double p[3] = {pts[3 * (size_t)i], pts[3 * (size_t)i + 1], pts[3 * (size_t)i + 2]};
    double c[3] = {root.c[0], root.c[1], root.c[2]};
    double h[3] = {root.hw[0], root.hw[1], root.hw[2]};
    unsigned long long key = 0;
    for (int l = 0; l < kMaxKeyLevels; l++) {
        int digit = 0;
        for (int d = 0; d < 3; d++) {
            int b = p[d] > c[d] ? 1 : 0;
            digit |= b << (2 - d);
            h[d] = h[d] * 0.5;
            c[d] = c[d] + ((double)b * 2 - 1) * h[d];
        }
        key = (key << 3) | (unsigned long long)digit;
    }
    keys[i] = key;
    idx[i] = i;
}

namespace {

struct TmpCell {
    int level, b, e, parent, octant;
    double c[3];
    int child[8];
    bool leaf;
};

struct Assembler {
    const std::vector<unsigned long long>& key;
    const std::vector<int>& perm;
    const double* pts;
    int leaf_size;
    std::vector<TmpCell> cells;  // depth-first pre-order

    bool coincident(int b, int e) const {
        if (key[b] != key[e - 1]) return false;  // different keys -> different points
        const double* p0 = pts + 3 * (size_t)perm[b];
        for (int k = b + 1; k < e; k++) {
            const double* p = pts + 3 * (size_t)perm[k];
            if (p[0] != p0[0] || p[1] != p0[1] || p[2] != p0[2]) return false;
        }
        return true;
    }

    int add(int level, int b, int e, int parent, int octant, const double* c) {
        TmpCell t;
        t.level = level; t.b = b; t.e = e; t.parent = parent; t.octant = octant;
        for (int d = 0; d < 3; d++) t.c[d] = c[d];
        for (int q = 0; q < 8; q++) t.child[q] = -1;
        t.leaf = true;
        cells.push_back(t);
        return (int)cells.size() - 1;
    }

    // split `id` whose splitting box is (c, hw)
    void split(int id, const double* c, const double* hw) {
#pragma clang fp contract(off)
        const int b = cells[id].b, e = cells[id].e, L = cells[id].level;
        if (e - b <= leaf_size) return;
        if (L + 1 > kMaxKeyLevels)
            throw Error(TBEM_E_UNSUPPORTED,
                        "octree deeper than 21 levels (near-coincident points); not supported");
        cells[id].leaf = false;
        const int shift = 63 - 3 * (L + 1);
        int pos = b;
        for (int oct = 0; oct < 8 && pos < e; oct++) {
            int hi = (int)(std::upper_bound(key.begin() + pos, key.begin() + e, oct,
                                            [shift](int o, unsigned long long k) {
                                                return o < (int)((k >> shift) & 7ull);
                                            }) - key.begin());
            if (hi > pos) {
                double hc[3], cc[3];
                for (int d = 0; d < 3; d++) {
                    int bit = (oct >> (2 - d)) & 1;
                    hc[d] = hw[d] * 0.5;
                    cc[d] = c[d] + ((double)bit * 2 - 1) * hc[d];
                }
                int ch = add(L + 1, pos, hi, id, oct, cc);
                cells[id].child[oct] = ch;
                if (!(hi - pos > leaf_size && coincident(pos, hi))) split(ch, cc, hc);
            }
            pos = hi;
        }
    }
};

}  // namespace

HostTree build_octree(const double* pts, int n, int leaf_size, hipStream_t s) {
#pragma clang fp contract(off)
    HostTree t;
    t.n_pts = n;
    // tight bounding box (Box::bounding_box, cpp/geometry.cpp:60-80)
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    std::vector<unsigned long long> keys;
    std::vector<int> perm;
    DevBuf<double> dpts;
    if (n > 0) {
        dpts.upload(pts, 3 * (size_t)n, s);
        int nb = std::min(1024, div_up(n, 256));
        DevBuf<double> part(6 * (size_t)nb);
        hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(256), 0, s, dpts.p, n, part.p);
        TBEM_LAUNCH_CHECK();
        auto hp = part.to_host(s);
        for (int d = 0; d < 3; d++) {
            lo[d] = hp[d];
            hi[d] = hp[3 + d];
        }
        for (int b = 1; b < nb; b++)
            for (int d = 0; d < 3; d++) {
                lo[d] = std::min(lo[d], hp[6 * b + d]);
                hi[d] = std::max(hi[d], hp[6 * b + 3 + d]);
            }
    }
    Box3 tight;
    for (int d = 0; d < 3; d++) {
        tight.c[d] = (hi[d] + lo[d]) * 0.5;
        tight.hw[d] = (hi[d] - lo[d]) * 0.5;
    }
    double mx = std::max(tight.hw[0], std::max(tight.hw[1], tight.hw[2]));
    double root_hw[3];
    for (int d = 0; d < 3; d++) root_hw[d] = tight.hw[d] + 1e-5 * mx;

    if (n > 0) {
        DevBuf<unsigned long long> k_in(n), k_out(n);
        DevBuf<int> i_in(n), i_out(n);
        hipLaunchKernelGGL(k_morton_keys, dim3(div_up(n, 256)), dim3(256), 0, s, dpts.p, n, tight,
                           k_in.p, i_in.p);
        TBEM_LAUNCH_CHECK();
        size_t tmp_bytes = 0;
        TBEM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k_in.p, k_out.p, i_in.p,
                                                    i_out.p, n, 0, 63, s));
        DevBuf<unsigned char> tmp(tmp_bytes);
        TBEM_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, k_in.p, k_out.p, i_in.p,
                                                    i_out.p, n, 0, 63, s));
        keys = k_out.to_host(s);
        perm = i_out.to_host(s);
    }

    Assembler as{keys, perm, pts, leaf_size, {}};
    as.cells.reserve(n > 0 ? 2 * (size_t)n / std::max(1, leaf_size) + 64 : 1);
    as.add(0, 0, n, -1, 0, tight.c);
    as.split(0, tight.c, tight.hw);  // root: no coincidence test (cpp/octree.cpp:198-214)

    // breadth-first renumbering (stable in depth-first order within a level)
    const int nc = (int)as.cells.size();
    int max_level = 0;
    for (auto& c: as.cells) max_level = std::max(max_level, c.level);
    t.n_levels = max_level + 1;
    t.n_cells = nc;
    t.level_begin.assign(t.n_levels + 1, 0);
    for (auto& c: as.cells) t.level_begin[c.level + 1]++;
    for (int L = 0; L < t.n_levels; L++) t.level_begin[L + 1] += t.level_begin[L];
    std::vector<int> fill(t.level_begin.begin(), t.level_begin.end() - 1);
    std::vector<int> bfs(nc);
    for (int i = 0; i < nc; i++) bfs[i] = fill[as.cells[i].level]++;

    t.level.resize(nc); t.begin.resize(nc); t.count.resize(nc); t.first_child.assign(nc, -1);
    t.n_child.assign(nc, 0); t.parent.resize(nc); t.octant.resize(nc); t.leaf.resize(nc);
    t.cx.resize(nc); t.cy.resize(nc); t.cz.resize(nc); t.dfs.resize(nc);
    for (int i = 0; i < nc; i++) {
        const TmpCell& c = as.cells[i];
        int j = bfs[i];
        t.dfs[i] = j;
        t.level[j] = c.level; t.begin[j] = c.b; t.count[j] = c.e - c.b;
        t.parent[j] = c.parent >= 0 ? bfs[c.parent] : -1;
        t.octant[j] = c.octant;
        t.leaf[j] = c.leaf ? 1 : 0;
        t.cx[j] = c.c[0]; t.cy[j] = c.c[1]; t.cz[j] = c.c[2];
        for (int q = 0; q < 8; q++) {
            if (c.child[q] < 0) continue;
            int cj = bfs[c.child[q]];
            if (t.first_child[j] < 0 || cj < t.first_child[j]) t.first_child[j] = cj;
            t.n_child[j]++;
        }
        if (c.leaf) t.leaves.push_back(j);
    }

    // per-level half widths: level 0 is the expanded root box, deeper levels halve the tight box
    t.level_hw.resize(3 * (size_t)t.n_levels);
    t.level_r.resize(t.n_levels);
    double h[3] = {tight.hw[0], tight.hw[1], tight.hw[2]};
    for (int L = 0; L < t.n_levels; L++) {
        for (int d = 0; d < 3; d++) {
            if (L > 0) h[d] = h[d] * 0.5;
            t.level_hw[3 * L + d] = (L == 0) ? root_hw[d] : h[d];
        }
        const double* q = &t.level_hw[3 * L];
        t.level_r[L] = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    }

    // ascending original index inside each leaf
    for (int j: t.leaves) std::sort(perm.begin() + t.begin[j], perm.begin() + t.begin[j] + t.count[j]);
    t.perm = std::move(perm);
    return t;
}

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
