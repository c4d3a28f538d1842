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
    // The per-pair output Y holds one slice of the pairs at a time: slice q is the
    // M2L items slice_items[q] .. slice_items[q+1] (targets, whole), i.e. the pairs
    // slice_pairs[q] .. slice_pairs[q+1], computed by units slice_units[q] ..
    // slice_units[q+1]; Y holds y_pairs (the largest slice) * ystride values.
    int n_slices = 1, y_pairs = 0;
    std::vector<int> slice_items, slice_pairs, slice_units;
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
// fp64: build the FP64 path's operators (see above) instead of the mixed path's.
// y_budget: bytes the per-pair output may take; the pairs are sliced to fit.
// t_surf, s_surf: the target's check and the source's equivalent surface (unit
// radius; null: the inner one, as for the M2L). With the outer / inner pair the
// FP64 operators serve the M2M (target: the parent), with inner / outer the L2L.
void build_m2l_mfma(M2lMfma& M, int kid, const KConst& kc, const HostTree& ot, const HostTree& st,
                    const DevTree& dot, const DevTree& dst, bool same_grid, bool fp64, size_t y_budget,
                    const HostList& m2l,
                    const DevList& dm2l, const Surfaces& surf, const double* d_uc2e, const double* d_dc2e,
                    hipStream_t s, const DevBuf<double>* up_fac = nullptr, const DevBuf<double>* down_fac = nullptr,
                    const std::vector<double>* t_surf = nullptr, const std::vector<double>* s_surf = nullptr);

// Mixed path, forward: slice by slice, Y[p] = M_{class(p)} * chk_up[src(p)] (FP32
// MFMA; chk_up = Bu * up-check potential per source cell, FP32) and z[t] = sum of the
// target's Y in pair order (FP64); then locals[t] += Vd[level t] z[t] for the targets
// (items as GEMM tiles). kernel_end, if given, is recorded after the last slice's
// MFMA kernel. items/ptr/pair_src: the forward M2L list on the device.
void m2l_forward32(const M2lMfma& M, const int* items, const int* ptr, const int* pair_src, const CellTiles& targets,
                   const float* chk_up, float* Y, double* z, double* locals, hipStream_t s, hipEvent_t kernel_end);
// FP64 path, forward: slice by slice, Y[p] = K_{class(p)} * mult[src(p)] (FP64 MFMA)
// and fld[t] = sum of the target's Y in pair order.
void m2l_forward64(const M2lMfma& M, const int* items, const int* ptr, const int* pair_src, const double* mult,
                   double* Y, double* fld, hipStream_t s, hipEvent_t kernel_end);
// Mixed path, transpose: w[t] = float(Vd[level t]^T lbar[t]) for the M2L targets,
// then slice by slice Y[p] = M_c^T w[tgt(p)] and z[s] (+)= the source's Y in the
// slice, and e2[s] = Bu[level s]^T z[s] for the sources with pairs (z: scratch).
void m2l_adjoint32(const M2lMfma& M, const CellTiles& targets, const double* lbar, float* w, float* Y, double* z,
                   double* e2, hipStream_t s);
// FP64 path, transpose: mbar[s] += sum over the pairs of source s of K_c^T cbar[tgt(p)].
void m2l_adjoint64(const M2lMfma& M, const double* cbar, double* Y, double* mbar, hipStream_t s);

}  // namespace tbem
