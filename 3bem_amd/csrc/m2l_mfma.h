// M2L of the mixed-precision path as grouped FP32 MFMA GEMMs.
#pragma once

#include "plan.h"

namespace tbem {

// Plan part. Every M2L pair (target t, source s) of the same octree falls in a
// class (level(t), level(s), integer offset of the centers in units of the finer
// half-width); all pairs of a class share one operator
//   S_c = DC2E[level t] * K_c * UC2E[level s]        (blk x blk, formed in FP64)
// with K_c(check_i <- equiv_j) the kernel between the target's down-check and the
// source's up-equivalent surface. S_c maps the source's up-check potential (the
// field of its points on its check sphere, natural size) straight to the
// target's down-equivalent density, so the FP32 data never passes through a
// pseudo-inverse, whose conditioning FP32 input cannot survive. S_c is stored in
// FP32, pre-swizzled into v_mfma_f32_16x16x4f32 A-operand order:
//   S[((c*mt + rt)*kt + ks)*64 + lane] = S_c[rt*16 + lane%16][ks*4 + lane/16].
struct M2lMfma {
    bool enabled = false;
    int n_classes = 0, blk = 0, mt = 0, kt = 0, ystride = 0;
    int n_units = 0, n_pairs = 0;
    DevBuf<float> S;
    DevBuf<int> unit_class, unit_off, unit_cnt, colpair;
    double plan_seconds = 0;
};

// m2l: the forward M2L list (items = target cells, cell[p] = source of pair p).
// Leaves M.enabled false (the caller keeps the FP64 M2L) when the two trees
// differ, an offset is not integral, or blk exceeds what one workgroup stages.
void build_m2l_mfma(M2lMfma& M, int kid, const KConst& kc, const HostTree& ot, const HostTree& st,
                    bool same_tree, const HostList& m2l, const Surfaces& surf, const double* d_uc2e,
                    const double* d_dc2e, hipStream_t s);

// Y[p] (ystride floats per pair) = S_{class(p)} * float(chk_up[src(p)]) for every pair.
void launch_m2l_mfma(const M2lMfma& M, const int* pair_src, const double* chk_up, float* Y, hipStream_t s);

// locals[t] += sum over the pairs p of target t of Y[p] (FP64 sum), for every M2L target.
void launch_m2l_reduce(int n_items, const int* items, const int* ptr, const float* Y, int ystride, int blk,
                       double* locals, hipStream_t s);

}  // namespace tbem
