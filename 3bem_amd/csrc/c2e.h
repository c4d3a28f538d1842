#pragma once

#include "plan.h"

namespace tbem {

// Batched thresholded pseudo-inverse of nmat row-major n x n FP64 matrices (device memory).
void gpu_pinv_batched(const double* d_mats, int nmat, int n, double threshold, double* d_out,
                      hipStream_t s);

// Per-level UC2E (up = true: equivalent surface inner, check surface outer) or DC2E
// (up = false: equivalent outer, check inner) of tree t, levels concatenated.
void compute_c2e(int kid, const KConst& kc, const HostTree& t, const Surfaces& surf, bool up,
                 DevBuf<double>& out, hipStream_t s);

}  // namespace tbem
