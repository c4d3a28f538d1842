#pragma once

#include "plan.h"

namespace tbem {

// Batched thresholded pseudo-inverse of nmat row-major n x n FP64 matrices (device
// memory). With fac_w and fac_v, also the pseudo-inverses' SVD factors in the
// convention of gpu_svd_batched (dropped modes have zero rows in fac_w).
void gpu_pinv_batched(const double* d_mats, int nmat, int n, double threshold, double* d_out, hipStream_t s,
                      DevBuf<double>* fac_w = nullptr, DevBuf<double>* fac_v = nullptr);

// One-sided Jacobi SVD of nmat row-major n x n FP64 matrices A: on return, row i of
// W[m] is column i of A V (= sigma_i u_i) and row i of V[m] is v_i, so that
// A = sum_i W_i V_i^T; the order of the singular triplets is unspecified.
void gpu_svd_batched(const double* d_mats, int nmat, int n, DevBuf<double>& W, DevBuf<double>& V,
                     hipStream_t s);

// Per-level UC2E (up = true: equivalent surface inner, check surface outer) or DC2E
// (up = false: equivalent outer, check inner) of tree t, levels concatenated.
// fac_w, fac_v: optionally the C2E matrices' SVD factors (as gpu_pinv_batched).
void compute_c2e(int kid, const KConst& kc, const HostTree& t, const Surfaces& surf, bool up,
                 DevBuf<double>& out, hipStream_t s, DevBuf<double>* fac_w = nullptr,
                 DevBuf<double>* fac_v = nullptr);

}  // namespace tbem
