// Plan data structures: octrees, surfaces, C2E operators and task lists.
#pragma once

#include <algorithm>
#include <array>
#include <atomic>
#include <cstdlib>
#include <exception>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace tbem {

constexpr int kMaxKeyLevels = 21;  // 3 bits per level in a 63-bit Morton key

// Host threads for plan construction: OMP_NUM_THREADS when set, else the hardware's,
// at most 32.
inline int plan_threads() {
    int t = (int)std::thread::hardware_concurrency();
    if (const char* e = std::getenv("OMP_NUM_THREADS")) t = std::atoi(e);
    return std::max(1, std::min(32, t));
}

// fn(i) for i in [0, n) on plan_threads() host threads, items taken in order from a
// shared counter; fn must only write data of its own i.
template <class F>
void parallel_for(size_t n, F&& fn) {
    const int nt = (int)std::min<size_t>(plan_threads(), n);
    if (nt <= 1) {
        for (size_t i = 0; i < n; i++) fn(i);
        return;
    }
    std::atomic<size_t> next(0);
    std::exception_ptr err;
    std::mutex err_mu;
    auto work = [&] {
        try {
            for (size_t i; (i = next.fetch_add(1)) < n;) fn(i);
        } catch (...) {
            std::lock_guard<std::mutex> lk(err_mu);
            if (!err) err = std::current_exception();
            next = n;
        }
    };
    std::vector<std::thread> th;
    for (int k = 1; k < nt; k++) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
    if (err) std::rethrow_exception(err);
}

// Octree in breadth-first numbering: level L's cells are [level_begin[L], level_begin[L+1]),
// Morton (= depth-first) order inside a level, so the children of a cell are contiguous.
// Points are permuted so that every cell is a contiguous range [begin, begin+count) and,
// inside a leaf, original indices ascend (the order of Octree::indices, cpp/octree.cpp:125-136).
struct HostTree {
    int n_pts = 0;
    int n_cells = 0;
    int n_levels = 0;
    std::vector<int> level, begin, count, first_child, n_child, parent, octant;
    std::vector<char> leaf;
    std::vector<double> cx, cy, cz;          // Box::center of each cell's bounds
    std::vector<double> level_hw;            // 3 per level: Box::half_width (level 0 = expanded root)
    std::vector<double> level_r;             // hypot(half_width) per level
    std::vector<int> level_begin;            // n_levels + 1
    std::vector<int> perm;                   // sorted position -> original index
    std::vector<int> leaves;                 // leaf cell ids in depth-first order
    std::vector<int> dfs;                    // all cell ids in depth-first pre-order
};

// Builds the octree of n AoS points on the GPU (Morton keys + radix sort), then
// assembles cells from the sorted key runs on the host. The root box is the points'
// tight bounding box (the reference's), or box_lo_hi (lo xyz, hi xyz) when given,
// which must contain every point: two trees built on one box put their cells on one
// grid (same centres and half widths level by level).
HostTree build_octree(const double* pts_host, int n, int min_pts_per_cell, hipStream_t s,
                      const double* box_lo_hi = nullptr);
// Tight bounding box of n AoS points: lo xyz, hi xyz.
void point_bounds(const double* pts_host, int n, double* lo_hi);

// Device primitives of the plan build (octree.hip). Stable LSD radix sorts of
// (key, value) pairs by the low `bits` bits of the keys, 8 bits per pass; the result
// ends in k, v (kt, vt: n-element scratch).
void radix_sort_pairs_u64(unsigned long long* k, int* v, unsigned long long* kt, int* vt, int n, int bits,
                          hipStream_t s);
void radix_sort_pairs_u32(unsigned int* k, int* v, unsigned int* kt, int* vt, int n, int bits, hipStream_t s);
// out = exclusive prefix sum of the n ints of a (device); returns the total.
int device_exclusive_scan(const int* a, int n, int* out, hipStream_t s);

// Translation surfaces (cpp/fmm.h:42-66, cpp/fmm.cpp:22-51).
struct Surfaces {
    int n_e = 0;
    std::vector<double> unit;   // unit sphere points = normals, 3*n_e
    std::vector<double> inner;  // 0.3 * unit            (up-equivalent, down-check)
    std::vector<double> outer;  // (4 - sqrt2 - 2d) * unit (up-check, down-equivalent)
};
Surfaces make_surfaces(size_t order, double d);
std::vector<double> sphere_points(size_t order);

// Device mirror of a HostTree (SoA).
struct DevTree {
    int n_cells = 0, n_levels = 0;
    DevBuf<double> cx, cy, cz, level_r;
    DevBuf<int> level, begin, count;
    void upload(const HostTree& t, hipStream_t s);
};

struct DevPoints {
    DevBuf<double> x, y, z, nx, ny, nz, w;  // sorted SoA
    DevBuf<int> perm;                       // sorted -> original
    void upload(const HostTree& t, const double* locs, const double* normals, const double* w,
                hipStream_t s);
};

// A CSR list "item -> entries". Surface lists: entries are cell ids of the
// source tree (cell[ptr[k] .. ptr[k+1]]). Point lists: entries are contiguous
// ranges of sorted source points (rbeg/rcnt[ptr[k] .. ptr[k+1]]), the cells'
// ranges merged where they touch.
// rpre[e] is the exclusive prefix of rcnt inside the item's list, so a list is
// a virtual array of points that tiles can cut anywhere.
struct HostList {
    std::vector<int> items, ptr, cell, rbeg, rcnt, rpre;
};
struct DevList {
    int n_items = 0;
    DevBuf<int> items, ptr, cell, rbeg, rcnt, rpre;
    DevBuf<int> rclose;  // point-range lists of the point-target stage: launch_range_close
    void upload(const HostList& l, hipStream_t s);
};

// Cells grouped into tiles of at most 64 cells of one level (make_cell_tiles).
struct CellTiles {
    int n = 0;            // number of tiles
    DevBuf<int> cells;    // the cells, grouped by level
    DevBuf<int> tiles;    // per tile: first index into cells, cell count, level
};

// Turn a cell list (ptr/cell) into merged point ranges of tree t (ptr/rbeg/rcnt).
void cells_to_ranges(HostList& l, const HostTree& t);

struct TaskLists {
    std::vector<std::array<int, 2>> p2p, m2l, m2p, p2l;  // (obs cell, src cell)
    std::vector<int> p2m, m2m, l2l, l2p;
};

void dual_tree_traversal(const HostTree& ot, const HostTree& st, double mac, int n_e,
                         bool small_cells, TaskLists& tl);

}  // namespace tbem
