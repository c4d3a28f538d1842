// M2L of the mixed-precision path as grouped FP32 MFMA GEMMs.
#pragma once

#include "plan.h"

namespace tbem {

// Plan part. Every M2L pair (target t, source s) of the same octree falls in a
// class (level(t), level(s), integer offset of the centers in units of the finer
// half-width); all pairs of a class share one operator. With the modal factors
// UC2E = Ru Bu and DC2E = Vd Ld of the C2E matrices (Jacobi SVD), the FP64 M2L
// contribution DC2E K_c UC2E check_s is computed as
//   Vd[level t] * M_c * (Bu[level s] * check_s),   M_c = Ld K_c Ru  (blk x blk)
// with K_c(check_i <- equiv_j) the kernel between the target's down-check and the
// source's up-equivalent surface. Only M_c and the modal input Bu * check_s are
// FP32: the FP32 data never pass through a pseudo-inverse, whose conditioning FP32
// input cannot survive, and in the modal basis they carry no cancellation. M_c is
// stored pre-swizzled into v_mfma_f32_16x16x4f32 A-operand order:
//   S[((c*mt + rt)*kt + ks)*64 + lane] = M_c[rt*16 + lane%16][ks*4 + lane/16].
//
// FP64 path (fp64): the same classes with K_c itself in FP64, S64[...] = K_c[...] in
// the same order (v_mfma_f64_16x16x4f64 shares the A-operand layout), no modal
// factors: Y[p] = K_c * mult[src(p)] in FP64, the target's check potential is the sum
// of its pairs' Y, and the FP64 DC2E follows as on the surface-evaluation path.
struct M2lMfma {
    bool enabled = false;
    bool fp64 = false;
    int n_classes = 0, blk = 0, mt = 0, kt = 0, ystride = 0;
    int n_units = 0, n_pairs = 0;
    DevBuf<float> S, ST;    // M_c and M_c^T, A-operand order
    DevBuf<double> S64, ST64;  // FP64 path: K_c and K_c^T, A-operand order
    DevBuf<double> Bu, Vd;  // per source / observation level, blk x blk row-major
    DevBuf<int> unit_class, unit_off, unit_cnt, colpair;
    DevBuf<int> pair_tgt;   // target cell of every pair
    int n_adj = 0;          // transpose: source cells with pairs, and their pairs (CSR)
    DevBuf<int> adj_items, adj_ptr, adj_pairs;
    CellTiles adj_tiles;    // adj_items as GEMM tiles
    double plan_seconds = 0;
};

// Whether the MFMA M2L serves operators of this size on the current device: blk
// from 48 (below, the FP64 stage is faster) up to what one workgroup stages in LDS.
bool m2l_mfma_size_ok(int blk);
// The same for the FP64 operators (32 staged FP64 columns per unit).
bool m2l_fp64_size_ok(int blk);
// Bytes of the per-pair output Y for n_pairs pairs (16-row padded blk per pair).
size_t m2l_output_bytes(int n_pairs, int blk, bool fp64);

// m2l: the forward M2L list (items = target cells, cell[p] = source of pair p).
// Leaves M.enabled false (the caller keeps the FP64 M2L) when the trees' cells are
// not on one grid (same_grid false: distinct trees on their own root boxes), an
// offset is not integral, or m2l_mfma_size_ok(blk) is false.
// up_fac, down_fac: optionally {W, V}, the SVD factors of UC2E and DC2E per level
// (from the pseudo-inverse's own SVD, gpu_pinv_batched); without them the C2E
// matrices are factored again.
// dot, dst: the trees on the device; dm2l: the device copy of m2l.
// fp64: build the FP64 path's operators (see above) instead of the mixed path's.
void build_m2l_mfma(M2lMfma& M, int kid, const KConst& kc, const HostTree& ot, const HostTree& st,
                    const DevTree& dot, const DevTree& dst, bool same_grid, bool fp64, const HostList& m2l,
                    const DevList& dm2l, const Surfaces& surf, const double* d_uc2e, const double* d_dc2e,
                    hipStream_t s, const DevBuf<double>* up_fac = nullptr, const DevBuf<double>* down_fac = nullptr);

// Y[p] (ystride floats per pair) = M_{class(p)} * chk_up[src(p)] for every pair,
// chk_up = Bu * up-check potential per source cell, rounded to FP32.
// transpose: Y[p] = M_{class(p)}^T * chk_up[tgt(p)] (pair_src is then unused).
void launch_m2l_mfma(const M2lMfma& M, const int* pair_src, const float* chk_up, float* Y, hipStream_t s,
                     bool transpose = false);

// FP64 path: Y[p] (ystride doubles per pair) = K_{class(p)} * x[src(p)] (x: blk
// doubles per source cell, the up-equivalent densities); transpose:
// Y[p] = K_{class(p)}^T * x[tgt(p)] (x per target cell).
void launch_m2l_mfma64(const M2lMfma& M, const int* pair_src, const double* x, double* Y, hipStream_t s,
                       bool transpose = false);
// FP64 path: out[items[k]] = sum of Y over the pairs ptr[k] .. ptr[k+1] (pair order).
void launch_m2l_sum64(const M2lMfma& M, int n_items, const int* items, const int* ptr, const double* Y, double* out,
                      hipStream_t s);
// FP64 path, transpose: out[s] += sum of Y over the pairs of source cell s.
void launch_m2l_sum64_adj(const M2lMfma& M, const double* Y, double* out, hipStream_t s);

// Transpose, input: w[t] = float(Vd[level t]^T lbar[t]) for every M2L target t
// (targets: the M2L targets as GEMM tiles).
void launch_m2l_modal_in(const M2lMfma& M, const CellTiles& targets, const double* lbar, float* w, hipStream_t s);

// Transpose, output: e2[s] = Bu[level s]^T * (sum over the pairs p of source s of
// Y[p]), FP64, for the sources with pairs; z (blk doubles per source cell) is scratch.
void launch_m2l_reduce_adj(const M2lMfma& M, const float* Y, double* z, double* e2, hipStream_t s);

// locals[t] += Vd[level t] * (sum over the pairs p of target t of Y[p]), in FP64
// (targets: items as GEMM tiles; z, blk doubles per target cell, is scratch).
void launch_m2l_reduce(const M2lMfma& M, int n_items, const int* items, const int* ptr, const CellTiles& targets,
                       const float* Y, double* z, double* locals, hipStream_t s);

}  // namespace tbem
