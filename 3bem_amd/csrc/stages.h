// Device views and launchers of the FMM stage kernels.
#pragma once

#include <type_traits>

#include "kernels.h"
#include "plan.h"

namespace tbem {

// threads per workgroup (multiples of the 64-lane wavefront)
constexpr int kSurfBlock = 128;  // surface-target stages (P2M, M2M, M2L, P2L, L2L)
constexpr int kLeafBlock = 512;  // point-target stage (P2P + M2P + L2P)
// Shape of the point-target stage per kernel and precision: workgroup width and
// the minimum resident waves per SIMD the compiler must allow (which caps the
// VGPRs at 512 / waves). Measured on 2M points:
//  * hypersingular (150-174 VGPRs uncapped) and the FP32 traction kernels (156-163,
//    which spill under a 128 cap): 256-thread workgroups at three waves (<= 168
//    VGPRs, 12 waves per CU instead of one 512-thread workgroup's 8);
//  * FP64 traction kernels: 512 threads at four waves (125 VGPRs, from 148-152; two
//    workgroups per CU): adjoint-traction forward 75 -> 65 ms, traction adjoint 84
//    -> 74 ms;
//  * the rest (<= 128 VGPRs already): 512 threads, no cap; a cap at four waves made
//    the FP64 elastic-U leaf 16% slower (101 VGPRs, less latency hiding);
//  * Laplace single layer, whose pair is cheap, at 256 threads (4M points, leaf 100,
//    the larger leaves on this kernel: FP64 leaf stage 38.2 -> 36.9 ms, mixed 27.9
//    -> 26.5 ms).
template <int KID, class T>
constexpr bool leaf_narrow() {
    return KID == 0 || KID == 6 || ((KID == 4 || KID == 5) && std::is_same<T, float>::value);
}
template <int KID, class T>
constexpr int leaf_block() {
    return leaf_narrow<KID, T>() ? 256 : kLeafBlock;
}
template <int KID, class T>
constexpr int leaf_waves() {
    return leaf_narrow<KID, T>() ? 3 : ((KID == 4 || KID == 5) ? 4 : 1);
}
constexpr int kGradBlock = 256;  // observation-coordinate gradient of the point-target stage
constexpr int kWaveBlock = 256;        // k_leaf_wave: four work items per workgroup
constexpr int kWaveLeafMax = 128;      // leaves of up to this many points go to k_leaf_wave ...
constexpr long kWavePairsMax = 65536;  // ... one wave per leaf, or per chunk of at most this
                                       // many pairs (points x sources) of a larger one
constexpr int kSymBlock = 256;         // k_leaf_sym: four waves per leaf
constexpr int kSymLeafMin = 16;        // k_leaf_sym takes leaves of kSymLeafMin .. kSymLeafMax
constexpr int kSymLeafMax = 256;       // points (the plan chooses which, per kernel)
constexpr int kSymSlots = kSymLeafMax + 64;  // k_leaf_sym's target slots per wave and component
constexpr int kSymSrc = 3;             // k_leaf_sym: sources per lane

// One work item of k_leaf_wave: the leaf at list position pos, whole or one chunk of
// its sources. Point sources [v0, v1) of the leaf's packed ranges (offsets into their
// concatenation; ranges f0 .. f1-1 hold them), surface cells [c0, c1) of its surface
// list, and its own surface when self != 0. part < 0: the item is the whole leaf and
// writes the output; else it writes its unscaled sums, R x m values (the gradient's
// 3 x m) at part_buf[part + q*m + t], and k_leaf_parts (k_leaf_grad_parts) adds the
// chunks of the leaf in order.
struct WaveItem {
    int pos, f0, f1, v0, v1, c0, c1, self, part;
};
// A leaf cut into n chunks: chunk c wrote part_buf[part + c*kPartRows*m ...] (rows:
// the R output components, or the gradient's 3).
constexpr int kPartRows = 3;
struct WaveLeafParts {
    int pos, part, n;
};

struct TreeView {
    const double *cx, *cy, *cz, *level_r;
    const int *level, *begin, *count;
};
struct PtsView {
    const double *x, *y, *z, *nx, *ny, *nz;
};

inline TreeView view(const DevTree& t) {
    return TreeView{t.cx.p, t.cy.p, t.cz.p, t.level_r.p, t.level.p, t.begin.p, t.count.p};
}
inline PtsView view(const DevPoints& p) { return PtsView{p.x.p, p.y.p, p.z.p, p.nx.p, p.ny.p, p.nz.p}; }

// Check-potential stage. For each item (a target cell of tree tt) accumulate, on
// the cell's target surface, the field of the item's sources:
//   out[cell] (= or +=) field
// The check-to-equivalent products that follow run batched per level
// (launch_level_gemm).
// Sources are either equivalent surfaces of cells of tree st carrying s_coeff
// (src_surf: entries cell[ptr[item] .. ptr[item+1]]) or contiguous ranges of the
// sorted source points carrying dens (!src_surf: ranges rbeg/rcnt[ptr[item] ..]).
// adj = true evaluates the transposed kernel with the geometric roles swapped:
// the target surface plays the kernel's source, each source point its observer.
struct SurfStage {
    int n_items;
    const int *items, *ptr, *cell, *rbeg, *rcnt, *rpre;
    TreeView tt;
    const double* t_surf;   // scaled target surface, 3*n_e
    TreeView st;
    const double* s_surf;   // scaled source surface (surface sources)
    const double* s_coeff;  // blk per source cell
    PtsView sp;
    const double* dens;     // point densities, component-major in sorted order
    int n_src_pts;
    const double* unit;     // unit sphere = surface normals
    int n_e, blk;
    double* out;            // blk per cell of tt: the raw check potentials (= or += with add)
    int no_cutoff;          // surface sources: every pair is known to lie beyond the r2 < 1e-12
                            // cutoff, so the kernels leave the test out (same result)
};

// Tiles of at most 64 cells of one level for the batched per-level products (the
// cells are grouped by level first).
CellTiles make_cell_tiles(const std::vector<int>& cells, const HostTree& t, hipStream_t s);
CellTiles make_level_tiles(const HostTree& t, int level, hipStream_t s);

// For every cell c of the tiles: out[c] (= or += when add) op(M[level c]) in[c]
// (+ addin[c] when given), op = identity or transpose, blk doubles per cell, M per
// level blk*blk row-major. FP64; out32 receives the result rounded to FP32 instead
// of out when given. A tiled GEMM: each level matrix is read once per 64 cells.
void launch_level_gemm(const CellTiles& T, int blk, const double* M, bool trans, const double* in,
                       const double* addin, double* out, float* out32, bool add, hipStream_t s);

// Point-target stage (FP64, or T = float for the point sources (a) with every
// displacement formed in FP64; the surface sources (b), (c) are always FP64).
// For each leaf of tree tt, for each of its points i:
//   out_sorted[q*n_tgt + i] = wscale_i * sum over
//     (a) the sorted source points in its ranges                (P2P,   or P2P^T)
//     (b) the surfaces (sl_surf) of the cells in its surface list (M2P,   or P2L^T)
//     (c) its own surface self_surf with self_coeff             (L2P,   or P2M^T)
// wscale_i = t_weight[i] when t_weight != nullptr, else 1; with accumulate the sums are
// added to out_sorted. A null rp_ptr leaves out (a), null sl_ptr (b), null self_coeff (c).
struct LeafStage {
    int n_leaves;
    const int* leaves;
    TreeView tt;
    PtsView tp;
    const double* t_weight;
    int n_tgt;
    const int *rp_ptr, *rp_beg, *rp_cnt, *rp_pre;   // point ranges, indexed by leaf position
    const int* rp_close;    // per range: 1 when its points may come within the cutoff radius of
                            // the leaf's (launch_range_close); null: every range
    PtsView ps;
    const double* pdens;
    int n_ps;
    const int *sl_ptr, *sl_cell;           // indexed by leaf position
    TreeView slt;
    const double* sl_surf;
    const double* sl_coeff;
    const double* self_surf;
    const double* self_coeff;
    const double* unit;
    int n_e, blk;
    double* out_sorted;
    int accumulate;         // out_sorted += (instead of =) the stage's sums
    // The leaf positions a launch serves (null: 0 .. n_leaves-1, one workgroup each).
    const int* order;
    int n_order;
    // Optional split of the positions (launch_leaf_stage): wave_items, leaves of at most
    // kWaveLeafMax points, whole or in chunks, one wave each, the most work first;
    // block_pos, the rest, one workgroup each. wave_parts: the leaves cut into chunks,
    // whose partial sums part_buf holds between the two kernels.
    const WaveItem* wave_items;
    const int* block_pos;
    int n_wave, n_block;
    const WaveLeafParts* wave_parts;
    int n_parts;
    double* part_buf;
    const double* grad_g;   // obs-gradient mode: upstream g (sorted, component-major)
    double* grad_out;       // obs-gradient mode: 3 per target (sorted, AoS)
    // Symmetric pairs (Kelvin U with one point set on both sides; null: none). Per leaf
    // position, the point ranges of the leaves whose pairs with it k_leaf_sym
    // evaluates once for both directions (the rp_* lists then leave those pairs
    // out): sy_base[pos] is the first slot of the leaf's reverse sums in rev (R
    // component rows of n_rev); sy_order: the positions that own pairs, one
    // workgroup each; sg_pos / sg_ptr / sg_off: per receiving position, the slot
    // offsets (in rev) of the sums addressed to its first point, added in list order.
    const int *sy_ptr, *sy_beg, *sy_cnt, *sy_pre, *sy_base, *sy_order;
    int n_sy_order;
    const int *sg_pos, *sg_ptr, *sg_off;
    int n_sg;
    double* rev;
    size_t n_rev;
};

// FP64 only: every check potential here feeds a check-to-equivalent pseudo-inverse
// (or is the transpose of such a step), whose conditioning FP32 input cannot survive.
void launch_surf_stage(int kid, bool src_surf, bool add, bool adj, const SurfStage& a,
                       const KConst& kc, hipStream_t s);
// fast: point-to-point pairs use the one-Newton-step FP64 rsqrt (relative error
// <= 9e-14 per pair; the FMM's parity bar is 1e-12). The direct N-body operator
// keeps the full one.
template <class T>
void launch_leaf_stage(int kid, bool adj, const LeafStage& a, const KConst& kc, hipStream_t s, bool fast);
// d/d(target coords) of the forward leaf stage contracted with grad_g; fp32_points:
// point sources in FP32 (mixed precision; kernels with closed-form gradients)
void launch_leaf_grad(int kid, const LeafStage& a, const KConst& kc, hipStream_t s, bool fp32_points);


// A source farther than this from the box of a leaf's points pairs with none of them
// inside the r2 < 1e-12 cutoff, so the test may be left out (same result).
constexpr double kCloseMargin = 1e-5;

// For the point-range list l of the targets' leaves (item k = leaf leaves[k] of tree
// tt, points tp; ranges of source points sp): close[e] = 1 when a point of range e
// lies within `margin` of the tight box of its item's target points, else 0. A range
// with close 0 holds no pair within that distance, so the point-target stage may
// leave the r2 < 1e-12 test out for it (margin >= 1e-5 keeps that exact).
void launch_range_close(int n_items, const int* leaves, const TreeView& tt, const PtsView& tp, const int* ptr,
                        const int* rbeg, const int* rcnt, const PtsView& sp, double margin, int* close,
                        hipStream_t s);

void launch_gather_dens(const double* x, const int* perm, const double* w, int n, int R,
                        double* dens, hipStream_t s);
void launch_scatter_out(const double* out_sorted, const int* perm, int n, int R, double* y,
                        hipStream_t s);
void launch_fill(double* p, size_t n, double v, hipStream_t s);

}  // namespace tbem
