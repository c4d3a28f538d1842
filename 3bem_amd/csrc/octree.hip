// Octree build on the GPU (cpp/octree.cpp:95-180, cpp/geometry.cpp:12-95).
//
// Semantics of the reference tree:
//  * the root's bounds are the tight box expanded by 1e-5 * max half width, but its
//    children split the tight (unexpanded) box;
//  * a child's box halves the half width and moves the centre by +-half width,
//    child id = [x > cx] << 2 | [y > cy] << 1 | [z > cz] (strict >, x most significant);
//  * a cell is a leaf when it holds <= min_pts_per_cell points, or (below the root)
//    when all of its points coincide;
//  * a leaf's points are listed in ascending original index.
// Every point's path is a pure function of its coordinates and of centres obtained
// by replaying those FP64 operations, so each point gets a 63-bit key (21 levels,
// 3 bits each). A stable radix sort of (key, index) puts every cell of the reference
// tree on a contiguous key run; the cells are then read off the runs level by level
// (a device scan of run starts per level), breadth first and in Morton order within
// a level. A second stable radix sort, of the original indices by the start of their
// leaf, orders each leaf's points by original index.
#include <algorithm>
#include <cmath>

#include "plan.h"

namespace tbem {

// ------------------------------------------------------------------ keys and bounds
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
    for (int w = 128; w > 0; w >>= 1) {  // launched with 256 threads
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

// ------------------------------------------------------------------------- scans
// Tiles of 2048 ints: 256 threads x 8 consecutive items each.
constexpr int kScanThreads = 256, kScanItems = 8, kScanTile = kScanThreads * kScanItems;

// inclusive scan of one value per thread across the 256 threads of the block
__device__ __forceinline__ int block_inclusive_scan(int v, int* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const int u = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += u;
        __syncthreads();
    }
    const int r = sh[t];
    __syncthreads();
    return r;
}

// sums[b] = sum of tile b of a
__global__ __launch_bounds__(kScanThreads) void k_tile_sums(const int* __restrict__ a, int n, int* __restrict__ sums) {
    __shared__ int sh[kScanThreads];
    const size_t b0 = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    int s = 0;
    for (int k = 0; k < kScanItems; k++)
        if (b0 + k < (size_t)n) s += a[b0 + k];
    const int incl = block_inclusive_scan(s, sh);
    if (threadIdx.x == kScanThreads - 1) sums[blockIdx.x] = incl;
}

// exclusive scan of m ints in place, one workgroup of 1024 threads (chunks per thread)
__global__ __launch_bounds__(1024) void k_scan_single(int* __restrict__ a, int m) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (m + 1023) / 1024;
    const int b = min(m, t * per), e = min(m, b + per);
    int s = 0;
    for (int i = b; i < e; i++) s += a[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int u = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += u;
        __syncthreads();
    }
    int run = t ? part[t - 1] : 0;
    for (int i = b; i < e; i++) {
        const int v = a[i];
        a[i] = run;
        run += v;
    }
}

// out = exclusive scan of a, given the exclusive scan of the tile sums
// (a and out may be the same array: each thread reads its elements before it writes them)
__global__ __launch_bounds__(kScanThreads) void k_tile_scan(const int* a, int n, const int* __restrict__ tile_off,
                                                            int* out) {
    __shared__ int sh[kScanThreads];
    const size_t b0 = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    int v[kScanItems], s = 0;
    for (int k = 0; k < kScanItems; k++) {
        v[k] = (b0 + k < (size_t)n) ? a[b0 + k] : 0;
        s += v[k];
    }
    int run = block_inclusive_scan(s, sh) - s + tile_off[blockIdx.x];
    for (int k = 0; k < kScanItems; k++) {
        if (b0 + k < (size_t)n) out[b0 + k] = run;
        run += v[k];
    }
}

// a = exclusive scan of a (n ints), in place and without a host synchronization:
// tile sums, one workgroup over the tile sums, then each tile (whose threads read
// their elements before writing them). sums: scratch of at least div_up(n,
// kScanTile) + 1 ints.
static void device_exclusive_scan_inplace(int* a, int n, int* sums, hipStream_t s) {
    if (n == 0) return;
    const int tiles = div_up(n, kScanTile);
    hipLaunchKernelGGL(k_tile_sums, dim3(tiles), dim3(kScanThreads), 0, s, a, n, sums);
    TBEM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(1024), 0, s, sums, tiles + 1);
    TBEM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tile_scan, dim3(tiles), dim3(kScanThreads), 0, s, a, n, sums, a);
    TBEM_LAUNCH_CHECK();
}

// out = exclusive scan of a (n ints); returns the total
int device_exclusive_scan(const int* a, int n, int* out, hipStream_t s) {
    if (n == 0) return 0;
    const int tiles = div_up(n, kScanTile);
    DevBuf<int> sums(tiles + 1);
    hipLaunchKernelGGL(k_tile_sums, dim3(tiles), dim3(kScanThreads), 0, s, a, n, sums.p);
    TBEM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(1024), 0, s, sums.p, tiles + 1);
    TBEM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tile_scan, dim3(tiles), dim3(kScanThreads), 0, s, a, n, sums.p, out);
    TBEM_LAUNCH_CHECK();
    int total = 0;
    TBEM_HIP(hipMemcpyAsync(&total, sums.p + tiles, sizeof(int), hipMemcpyDeviceToHost, s));
    TBEM_HIP(hipStreamSynchronize(s));
    return total;
}

// ---------------------------------------------------- stable LSD radix sort (8-bit)
// Tiles of 2048 elements, element i of tile b at b*2048 + r*256 + t (round r, thread
// t), so that (round, wave, lane) is the elements' order. Each pass: a digit
// histogram per tile, an exclusive scan over (digit, tile), and a scatter in which
// each wave ranks its 64 elements by digit with eight ballots, the waves of a round
// and the rounds of a tile taking their places in order: the sort is stable.
constexpr int kRsThreads = 256, kRsItems = 8, kRsTile = kRsThreads * kRsItems;

template <class K>
__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const K* __restrict__ keys, int n, int shift,
                                                        int* __restrict__ counts, int tiles) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kRsTile;
    for (int r = 0; r < kRsItems; r++) {
        const size_t i = base + (size_t)r * kRsThreads + threadIdx.x;
        if (i < (size_t)n) atomicAdd(&h[(int)((keys[i] >> shift) & 255)], 1);
    }
    __syncthreads();
    counts[(size_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

template <class K>
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(const K* __restrict__ kin, const int* __restrict__ vin,
                                                           int n, int shift, const int* __restrict__ offsets,
                                                           int tiles, K* __restrict__ kout, int* __restrict__ vout) {
    __shared__ int run[256];    // next free slot of each digit for this tile
    __shared__ int wc[4][256];  // digit counts of each wave in the current round
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    run[t] = offsets[(size_t)t * tiles + blockIdx.x];
    for (int w = 0; w < 4; w++) wc[w][t] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kRsTile;
    const unsigned long long below = (1ull << lane) - 1;
    for (int r = 0; r < kRsItems; r++) {
        const size_t i = base + (size_t)r * kRsThreads + t;
        const bool valid = i < (size_t)n;
        const K k = valid ? kin[i] : K(0);
        const int v = valid ? vin[i] : 0;
        const int d = (int)((k >> shift) & 255);
        unsigned long long m = __ballot(valid);
        for (int b = 0; b < 8; b++) {
            const unsigned long long bb = __ballot((d >> b) & 1);
            m &= ((d >> b) & 1) ? bb : ~bb;
        }
        const int rank = __popcll(m & below);
        if (valid && rank == 0) wc[wave][d] = __popcll(m);
        __syncthreads();
        if (valid) {
            int off = run[d] + rank;
            for (int w = 0; w < wave; w++) off += wc[w][d];
            kout[off] = k;
            vout[off] = v;
        }
        __syncthreads();
        run[t] += wc[0][t] + wc[1][t] + wc[2][t] + wc[3][t];
        for (int w = 0; w < 4; w++) wc[w][t] = 0;
        __syncthreads();
    }
}

// Stable sort of (key, value) pairs by the low `bits` bits of the keys, in place
// (the results end in k, v; kt, vt are n-element scratch).
template <class K>
static void radix_sort_pairs(K* k, int* v, K* kt, int* vt, int n, int bits, hipStream_t s) {
    if (n <= 1) return;
    const int tiles = div_up(n, kRsTile);
    DevBuf<int> counts((size_t)256 * tiles);
    DevBuf<int> sums(div_up(256 * tiles, kScanTile) + 1);
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist<K>, dim3(tiles), dim3(kRsThreads), 0, s, k, n, shift, counts.p, tiles);
        TBEM_LAUNCH_CHECK();
        device_exclusive_scan_inplace(counts.p, 256 * tiles, sums.p, s);
        hipLaunchKernelGGL(k_rs_scatter<K>, dim3(tiles), dim3(kRsThreads), 0, s, k, v, n, shift, counts.p, tiles,
                           kt, vt);
        TBEM_LAUNCH_CHECK();
        std::swap(k, kt);
        std::swap(v, vt);
    }
    if (((bits + 7) / 8) & 1) {  // an odd number of passes left the result in the scratch
        TBEM_HIP(hipMemcpyAsync(kt, k, (size_t)n * sizeof(K), hipMemcpyDeviceToDevice, s));
        TBEM_HIP(hipMemcpyAsync(vt, v, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, s));
    }
}

void radix_sort_pairs_u64(unsigned long long* k, int* v, unsigned long long* kt, int* vt, int n, int bits,
                          hipStream_t s) {
    radix_sort_pairs<unsigned long long>(k, v, kt, vt, n, bits, s);
}
void radix_sort_pairs_u32(unsigned int* k, int* v, unsigned int* kt, int* vt, int n, int bits, hipStream_t s) {
    radix_sort_pairs<unsigned int>(k, v, kt, vt, n, bits, s);
}

// -------------------------------------------------------------- cells, level by level
// Device arrays of one level's cells, in Morton (= depth-first) order.
struct DevLevel {
    int n = 0;
    DevBuf<int> begin, count, parent, octant;
    DevBuf<char> leaf;
    DevBuf<double> c;  // 3 per cell: the centre of the box the cell splits
};

// flag[i] = 1 where sorted point i starts a child run of a level-L cell that splits
// (key bits above `shift` differ from the previous point's, or i starts the cell)
__global__ void k_child_flags(int n, const unsigned long long* __restrict__ keys, const int* __restrict__ owner,
                              const int* __restrict__ cbeg, const char* __restrict__ cleaf, int shift,
                              int* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = owner[i];
    int f = 0;
    if (o >= 0 && !cleaf[o]) f = (i == cbeg[o] || (keys[i] >> shift) != (keys[i - 1] >> shift)) ? 1 : 0;
    flag[i] = f;
}

// the new cells' start, parent, octant and centre (parent centre +- child half width)
__global__ void k_child_make(int n, const unsigned long long* __restrict__ keys, const int* __restrict__ flag,
                             const int* __restrict__ pos, const int* __restrict__ owner,
                             const double* __restrict__ pc, int shift, double h0, double h1, double h2,
                             int* __restrict__ nbeg, int* __restrict__ npar, int* __restrict__ noct,
                             double* __restrict__ nc) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    const int c = pos[i], p = owner[i];
    const int oct = (int)((keys[i] >> shift) & 7ull);
    nbeg[c] = i;
    npar[c] = p;
    noct[c] = oct;
    const double h[3] = {h0, h1, h2};
    for (int d = 0; d < 3; d++) {
        const int bit = (oct >> (2 - d)) & 1;
        nc[3 * c + d] = pc[3 * p + d] + ((double)bit * 2 - 1) * h[d];
    }
}

// counts and leaf flags of the new cells: a cell is a leaf when it holds <= leaf
// points or when all of them coincide (equal keys, then equal coordinates); a
// non-leaf at the deepest key level cannot be split further (deep is set)
__global__ void k_child_finish(int m, const int* __restrict__ nbeg, const int* __restrict__ npar,
                               const int* __restrict__ pbeg, const int* __restrict__ pcnt,
                               const unsigned long long* __restrict__ keys, const int* __restrict__ perm,
                               const double* __restrict__ pts, int leaf, int last_level, int* __restrict__ ncnt,
                               char* __restrict__ nleaf, int* __restrict__ deep) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m) return;
    const int b = nbeg[c], p = npar[c];
    const int e = (c + 1 < m && npar[c + 1] == p) ? nbeg[c + 1] : pbeg[p] + pcnt[p];
    const int cnt = e - b;
    bool lf = cnt <= leaf;
    if (!lf && keys[b] == keys[e - 1]) {
        const double* q0 = pts + 3 * (size_t)perm[b];
        bool same = true;
        for (int k = b + 1; k < e && same; k++) {
            const double* q = pts + 3 * (size_t)perm[k];
            same = q[0] == q0[0] && q[1] == q0[1] && q[2] == q0[2];
        }
        lf = same;
    }
    ncnt[c] = cnt;
    nleaf[c] = lf ? 1 : 0;
    if (!lf && last_level) atomicOr(deep, 1);
}

// owner[i] <- the new cell holding point i (pos[i] + flag[i] - 1), -1 when point i's
// cell did not split
__global__ void k_owner_next(int n, const int* __restrict__ flag, const int* __restrict__ pos,
                             const char* __restrict__ cleaf, int* __restrict__ owner) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = owner[i];
    owner[i] = (o >= 0 && !cleaf[o]) ? pos[i] + flag[i] - 1 : -1;
}

// lkey[perm[i]] = start of the leaf holding sorted point i (one workgroup per leaf)
__global__ void k_leaf_keys(const int* __restrict__ lbeg, const int* __restrict__ lcnt, const int* __restrict__ perm,
                            unsigned int* __restrict__ lkey) {
    const int b = lbeg[blockIdx.x], c = lcnt[blockIdx.x];
    for (int k = threadIdx.x; k < c; k += blockDim.x) lkey[perm[b + k]] = (unsigned int)b;
}

__global__ void k_iota(int n, int* __restrict__ v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}

template <class T>
static std::vector<T> host_copy(const DevBuf<T>& b, size_t n, hipStream_t s) {
    std::vector<T> v(n);
    if (n) b.download(v.data(), n, s);
    return v;
}

void point_bounds(const double* pts, int n, double* lo_hi) {
    for (int d = 0; d < 3; d++) {
        lo_hi[d] = n > 0 ? pts[d] : 0.0;
        lo_hi[3 + d] = n > 0 ? pts[d] : 0.0;
    }
    for (size_t i = 1; i < (size_t)n; i++)
        for (int d = 0; d < 3; d++) {
            lo_hi[d] = std::min(lo_hi[d], pts[3 * i + d]);
            lo_hi[3 + d] = std::max(lo_hi[3 + d], pts[3 * i + d]);
        }
}

HostTree build_octree(const double* pts, int n, int leaf_size, hipStream_t s, const double* box_lo_hi) {
#pragma clang fp contract(off)
    HostTree t;
    t.n_pts = n;
    // tight bounding box (Box::bounding_box, cpp/geometry.cpp:60-80), or the given box
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    DevBuf<double> dpts;
    if (box_lo_hi) {
        for (int d = 0; d < 3; d++) {
            lo[d] = box_lo_hi[d];
            hi[d] = box_lo_hi[3 + d];
        }
    }
    if (n > 0) dpts.upload(pts, 3 * (size_t)n, s);
    if (n > 0 && !box_lo_hi) {
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

    // keys, sorted with their original indices
    DevBuf<unsigned long long> keys(std::max(n, 1)), ktmp(std::max(n, 1));
    DevBuf<int> perm(std::max(n, 1)), itmp(std::max(n, 1));
    if (n > 0) {
        hipLaunchKernelGGL(k_morton_keys, dim3(div_up(n, 256)), dim3(256), 0, s, dpts.p, n, tight, keys.p, perm.p);
        TBEM_LAUNCH_CHECK();
        radix_sort_pairs<unsigned long long>(keys.p, perm.p, ktmp.p, itmp.p, n, 63, s);
    }

    // cells, level by level; the root splits the tight box with no coincidence test
    std::vector<DevLevel> levels(1);
    {
        DevLevel& r = levels[0];
        r.n = 1;
        const std::vector<int> zero{0}, nn{n}, none{-1};
        r.begin.upload(zero, s);
        r.count.upload(nn, s);
        r.parent.upload(none, s);
        r.octant.upload(zero, s);
        const std::vector<char> lf{(char)(n <= leaf_size ? 1 : 0)};
        r.leaf.upload(lf, s);
        const std::vector<double> c0{tight.c[0], tight.c[1], tight.c[2]};
        r.c.upload(c0, s);
    }
    DevBuf<int> owner(std::max(n, 1)), flag(std::max(n, 1)), pos(std::max(n, 1)), deep(1);
    TBEM_HIP(hipMemsetAsync(owner.p, 0, (size_t)std::max(n, 1) * sizeof(int), s));
    TBEM_HIP(hipMemsetAsync(deep.p, 0, sizeof(int), s));
    double h[3] = {tight.hw[0], tight.hw[1], tight.hw[2]};
    const int gn = div_up(std::max(n, 1), 256);
    // (cells at the deepest key level are never split: a non-leaf there sets `deep`)
    for (int L = 0; n > leaf_size && L < kMaxKeyLevels; L++) {
        DevLevel& cur = levels[L];
        const int shift = 63 - 3 * (L + 1);
        hipLaunchKernelGGL(k_child_flags, dim3(gn), dim3(256), 0, s, n, keys.p, owner.p, cur.begin.p, cur.leaf.p,
                           shift, flag.p);
        TBEM_LAUNCH_CHECK();
        const int m = device_exclusive_scan(flag.p, n, pos.p, s);
        if (m == 0) break;
        for (int d = 0; d < 3; d++) h[d] = h[d] * 0.5;
        DevLevel nx;
        nx.n = m;
        nx.begin.alloc(m); nx.count.alloc(m); nx.parent.alloc(m); nx.octant.alloc(m); nx.leaf.alloc(m);
        nx.c.alloc(3 * (size_t)m);
        hipLaunchKernelGGL(k_child_make, dim3(gn), dim3(256), 0, s, n, keys.p, flag.p, pos.p, owner.p, cur.c.p,
                           shift, h[0], h[1], h[2], nx.begin.p, nx.parent.p, nx.octant.p, nx.c.p);
        TBEM_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_child_finish, dim3(div_up(m, 256)), dim3(256), 0, s, m, nx.begin.p, nx.parent.p,
                           cur.begin.p, cur.count.p, keys.p, perm.p, dpts.p, leaf_size,
                           (L + 1 == kMaxKeyLevels) ? 1 : 0, nx.count.p, nx.leaf.p, deep.p);
        TBEM_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_owner_next, dim3(gn), dim3(256), 0, s, n, flag.p, pos.p, cur.leaf.p, owner.p);
        TBEM_LAUNCH_CHECK();
        levels.push_back(std::move(nx));
    }
    int h_deep = 0;
    TBEM_HIP(hipMemcpyAsync(&h_deep, deep.p, sizeof(int), hipMemcpyDeviceToHost, s));
    TBEM_HIP(hipStreamSynchronize(s));
    TBEM_REQUIRE(h_deep == 0, TBEM_E_UNSUPPORTED,
                 "octree deeper than 21 levels (near-coincident points); not supported");

    // breadth-first numbering: level L's cells follow the previous levels', in Morton order
    t.n_levels = (int)levels.size();
    t.level_begin.assign(t.n_levels + 1, 0);
    for (int L = 0; L < t.n_levels; L++) t.level_begin[L + 1] = t.level_begin[L] + levels[L].n;
    const int nc = t.level_begin[t.n_levels];
    t.n_cells = nc;
    t.level.resize(nc); t.begin.resize(nc); t.count.resize(nc); t.first_child.assign(nc, -1);
    t.n_child.assign(nc, 0); t.parent.resize(nc); t.octant.resize(nc); t.leaf.resize(nc);
    t.cx.resize(nc); t.cy.resize(nc); t.cz.resize(nc);
    for (int L = 0; L < t.n_levels; L++) {
        const DevLevel& lv = levels[L];
        const int off = t.level_begin[L];
        auto b = host_copy(lv.begin, lv.n, s), c = host_copy(lv.count, lv.n, s), p = host_copy(lv.parent, lv.n, s),
             o = host_copy(lv.octant, lv.n, s);
        auto lf = host_copy(lv.leaf, lv.n, s);
        auto cc = host_copy(lv.c, 3 * (size_t)lv.n, s);
        TBEM_HIP(hipStreamSynchronize(s));
        for (int k = 0; k < lv.n; k++) {
            const int j = off + k;
            t.level[j] = L;
            t.begin[j] = b[k];
            t.count[j] = c[k];
            t.parent[j] = p[k] >= 0 ? t.level_begin[L - 1] + p[k] : -1;
            t.octant[j] = o[k];
            t.leaf[j] = lf[k];
            t.cx[j] = cc[3 * k];
            t.cy[j] = cc[3 * k + 1];
            t.cz[j] = cc[3 * k + 2];
        }
    }
    for (int j = 1; j < nc; j++) {
        const int p = t.parent[j];
        if (t.first_child[p] < 0) t.first_child[p] = j;
        t.n_child[p]++;
    }
    // depth-first pre-order: by start, a cell before its descendants (which share or
    // follow its start at deeper levels)
    std::vector<int> order(nc);
    for (int j = 0; j < nc; j++) order[j] = j;
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        return t.begin[a] != t.begin[b] ? t.begin[a] < t.begin[b] : t.level[a] < t.level[b];
    });
    t.dfs = order;
    for (int j : order)
        if (t.leaf[j]) t.leaves.push_back(j);

    // per-level half widths: level 0 is the expanded root box, deeper levels halve the tight box
    t.level_hw.resize(3 * (size_t)t.n_levels);
    t.level_r.resize(t.n_levels);
    double hh[3] = {tight.hw[0], tight.hw[1], tight.hw[2]};
    for (int L = 0; L < t.n_levels; L++) {
        for (int d = 0; d < 3; d++) {
            if (L > 0) hh[d] = hh[d] * 0.5;
            t.level_hw[3 * L + d] = (L == 0) ? root_hw[d] : hh[d];
        }
        const double* q = &t.level_hw[3 * L];
        t.level_r[L] = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    }

    // ascending original index inside each leaf: original indices (in ascending
    // order) stably sorted by the start of their leaf
    if (n > 0) {
        const int nl = (int)t.leaves.size();
        std::vector<int> lb(nl), lc(nl);
        for (int k = 0; k < nl; k++) {
            lb[k] = t.begin[t.leaves[k]];
            lc[k] = t.count[t.leaves[k]];
        }
        DevBuf<int> dlb, dlc;
        dlb.upload(lb, s);
        dlc.upload(lc, s);
        DevBuf<unsigned int> lkey(n), lktmp(n);
        hipLaunchKernelGGL(k_leaf_keys, dim3(nl), dim3(64), 0, s, dlb.p, dlc.p, perm.p, lkey.p);
        TBEM_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_iota, dim3(gn), dim3(256), 0, s, n, itmp.p);
        TBEM_LAUNCH_CHECK();
        int bits = 1;
        while (bits < 31 && (1ll << bits) < (long long)n) bits++;
        radix_sort_pairs<unsigned int>(lkey.p, itmp.p, lktmp.p, perm.p, n, bits, s);
        t.perm = host_copy(itmp, n, s);
        TBEM_HIP(hipStreamSynchronize(s));
    }
    return t;
}

}  // namespace tbem
