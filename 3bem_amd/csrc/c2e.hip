// Check-to-equivalent operators (UC2E / DC2E) in FP64 on the GPU.
//
// Reference: build_check_to_equiv / svd_inverse_nbody (cpp/fmm.cpp:112-160) take, for
// every tree level, the first depth-first cell of that level, form the
// (R n_e) x (R n_e) matrix K(check_i <- equiv_j) (nbody_matrix, weights 1, row
// d1*n_e+i, column d2*n_e+j) and invert it with a LAPACK SVD whose singular values
// <= 1e-10 * sigma_max are dropped (cpp/blas_wrapper.cpp:67-127).
//
// Here: one kernel assembles every level's matrix (entries bit-identical to the
// reference's: same moved points, same kernel expression order, no contraction);
// a second runs a one-sided (Hestenes) Jacobi SVD per matrix, one workgroup each,
// columns rotated in round-robin order so n/2 rotations proceed in parallel, and
// assembles pinv = V diag(1/sigma) U^T from the rotated columns.
#include <algorithm>
#include <cmath>

#include "c2e.h"

namespace tbem {

__global__ void k_c2e_assemble(int kid, KConst kc, int n_e, const double* __restrict__ centers,
                               const double* __restrict__ radii, const double* __restrict__ check,
                               const double* __restrict__ equiv, const double* __restrict__ unit,
                               double* __restrict__ mats) {
#pragma clang fp contract(off)
    const int m = blockIdx.y;
    const int R = (kid == K_LS || kid == K_LD || kid == K_LH || kid == K_ID1) ? 1 : 3;
    const int n = R * n_e;
    const double r = radii[m];
    const double* c = centers + 3 * m;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n_e * n_e; p += gridDim.x * blockDim.x) {
        const int i = p / n_e, j = p % n_e;
        double ck[3], eq[3];
        for (int d = 0; d < 3; d++) {
            ck[d] = r * check[3 * i + d] + c[d];
            eq[d] = r * equiv[3 * j + d] + c[d];
        }
        double kv[9];
        tensor_ref_dispatch(kid, ck, eq, unit + 3 * i, unit + 3 * j, kc, kv);
        double* M = mats + (size_t)m * n * n;
        for (int d1 = 0; d1 < R; d1++)
            for (int d2 = 0; d2 < R; d2++)
                M[(size_t)(d1 * n_e + i) * n + d2 * n_e + j] = 1.0 * kv[d1 * R + d2];
    }
}

// One workgroup per matrix. W holds the columns of A as rows (W = A^T), V the
// columns of V as rows; both row-major n x n in global memory.
__global__ __launch_bounds__(1024) void k_jacobi_pinv(int n, double threshold, const double* __restrict__ A,
                                                      double* __restrict__ Wb, double* __restrict__ Vb,
                                                      double* __restrict__ out, int* __restrict__ status) {
    const int m = blockIdx.x;
    const double* a = A + (size_t)m * n * n;
    double* W = Wb + (size_t)m * n * n;
    double* V = Vb + (size_t)m * n * n;
    double* P = out + (size_t)m * n * n;
    const int tid = threadIdx.x, nthr = blockDim.x;
    __shared__ int rotated;
    __shared__ double smax;
    extern __shared__ double sig[];  // n

    for (int p = tid; p < n * n; p += nthr) {
        const int i = p / n, k = p % n;
        W[p] = a[(size_t)k * n + i];
        V[p] = (i == k) ? 1.0 : 0.0;
    }
    __syncthreads();

    // squared Frobenius norm: columns below 1e-15 ||A||_F only carry rounding noise
    __shared__ double fro2;
    if (tid == 0) fro2 = 0;
    __syncthreads();
    {
        double s = 0;
        for (int p = tid; p < n * n; p += nthr) s += W[p] * W[p];
        atomicAdd(&fro2, s);
    }
    __syncthreads();
    const double tiny = 1e-30 * fro2;
    const double tol = fmax(1e-15, sqrt((double)n) * 2.220446049250313e-16);

    const int np = (n + 1) & ~1;  // even number of players (index n is a bye when n is odd)
    const int half = np / 2;
    int tpp = 1;                  // threads per pair (power of two, <= 64)
    while (tpp * 2 * half <= nthr && tpp < 64) tpp *= 2;
    const int team = tid / tpp, lane = tid % tpp;
    const int teams = nthr / tpp;

    int sweep = 0;
    for (; sweep < 60; sweep++) {
        if (tid == 0) rotated = 0;
        __syncthreads();
        for (int round = 0; round < np - 1; round++) {
            for (int k = team; k < half; k += teams) {
                int p, q;
                if (k == 0) {
                    p = np - 1;
                    q = round;
                } else {
                    p = (round + k) % (np - 1);
                    q = (round - k + (np - 1)) % (np - 1);
                }
                bool valid = p < n && q < n;
                double al = 0, be = 0, ga = 0;
                if (valid) {
                    const double* wp = W + (size_t)p * n;
                    const double* wq = W + (size_t)q * n;
                    for (int t = lane; t < n; t += tpp) {
                        double x = wp[t], y = wq[t];
                        al += x * x;
                        be += y * y;
                        ga += x * y;
                    }
                }
                for (int off = tpp / 2; off > 0; off >>= 1) {
                    al += __shfl_xor(al, off, tpp);
                    be += __shfl_xor(be, off, tpp);
                    ga += __shfl_xor(ga, off, tpp);
                }
                if (valid && al > tiny && be > tiny && fabs(ga) > tol * sqrt(al) * sqrt(be)) {
                    const double zeta = (be - al) / (2 * ga);
                    const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
                    const double cs = 1 / sqrt(1 + t * t), sn = cs * t;
                    double* wp = W + (size_t)p * n;
                    double* wq = W + (size_t)q * n;
                    double* vp = V + (size_t)p * n;
                    double* vq = V + (size_t)q * n;
                    for (int t2 = lane; t2 < n; t2 += tpp) {
                        double x = wp[t2], y = wq[t2];
                        wp[t2] = cs * x - sn * y;
                        wq[t2] = sn * x + cs * y;
                        x = vp[t2];
                        y = vq[t2];
                        vp[t2] = cs * x - sn * y;
                        vq[t2] = sn * x + cs * y;
                    }
                    if (lane == 0) rotated = 1;
                }
            }
            __syncthreads();
        }
        if (!uni(rotated)) break;
        __syncthreads();
    }
    if (tid == 0) status[m] = (sweep < 60) ? 0 : 1;
    if (out == nullptr) return;  // SVD only: W (rows = columns of A V) and V are the factors

    // singular values = column norms of A V
    for (int i = tid; i < n; i += nthr) {
        double s = 0;
        const double* w = W + (size_t)i * n;
        for (int k = 0; k < n; k++) s += w[k] * w[k];
        sig[i] = sqrt(s);
    }
    __syncthreads();
    if (tid == 0) {
        double mx = 0;
        for (int i = 0; i < n; i++) mx = fmax(mx, sig[i]);
        smax = mx;
    }
    __syncthreads();
    const double cutoff = threshold * smax;
    for (int i = tid; i < n; i += nthr) sig[i] = (sig[i] > cutoff) ? 1.0 / (sig[i] * sig[i]) : 0.0;
    __syncthreads();
    // pinv[r][c] = sum_i V[i][r] W[i][c] / sigma_i^2
    for (int p = tid; p < n * n; p += nthr) {
        const int r = p / n, c = p % n;
        double s = 0;
        for (int i = 0; i < n; i++) s += V[(size_t)i * n + r] * W[(size_t)i * n + c] * sig[i];
        P[p] = s;
    }
}

// The pseudo-inverse's own factors, in gpu_svd_batched's convention, from the SVD
// of A (rows of W = sigma_j u_j, rows of V = v_j): pinv(A) = sum_j (1/sigma_j) v_j
// u_j^T, so row j of Wp = v_j / sigma_j (zero for the modes the threshold drops)
// and row j of Vp = u_j = W_j / sigma_j.
__global__ void k_pinv_factors(int n, double threshold, const double* __restrict__ W, const double* __restrict__ V,
                               double* __restrict__ Wp, double* __restrict__ Vp) {
    const int m = blockIdx.x;
    const size_t off = (size_t)m * n * n;
    extern __shared__ double sig[];  // n, then the maximum
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        double s2 = 0;
        for (int k = 0; k < n; k++) s2 += W[off + (size_t)j * n + k] * W[off + (size_t)j * n + k];
        sig[j] = sqrt(s2);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double mx = 0;
        for (int j = 0; j < n; j++) mx = fmax(mx, sig[j]);
        sig[n] = mx;
    }
    __syncthreads();
    const double cutoff = threshold * sig[n];
    for (size_t e = threadIdx.x; e < (size_t)n * n; e += blockDim.x) {
        const int j = (int)(e / n);
        const double sj = sig[j];
        Wp[off + e] = (sj > cutoff) ? V[off + e] / sj : 0.0;
        Vp[off + e] = (sj > 0) ? W[off + e] / sj : 0.0;
    }
}

void gpu_pinv_batched(const double* d_mats, int nmat, int n, double threshold, double* d_out, hipStream_t s,
                      DevBuf<double>* fac_w, DevBuf<double>* fac_v) {
    if (nmat == 0 || n == 0) return;
    DevBuf<double> W((size_t)nmat * n * n), V((size_t)nmat * n * n);
    DevBuf<int> st(nmat);
    hipLaunchKernelGGL(k_jacobi_pinv, dim3(nmat), dim3(1024), n * sizeof(double), s, n, threshold,
                       d_mats, W.p, V.p, d_out, st.p);
    TBEM_LAUNCH_CHECK();
    auto hs = st.to_host(s);
    for (int v: hs) TBEM_REQUIRE(v == 0, TBEM_E_NUMERIC, "Jacobi SVD did not converge in 60 sweeps");
    if (fac_w && fac_v) {
        fac_w->alloc((size_t)nmat * n * n);
        fac_v->alloc((size_t)nmat * n * n);
        hipLaunchKernelGGL(k_pinv_factors, dim3(nmat), dim3(256), (n + 1) * sizeof(double), s, n, threshold, W.p,
                           V.p, fac_w->p, fac_v->p);
        TBEM_LAUNCH_CHECK();
        TBEM_HIP(hipStreamSynchronize(s));  // W and V are freed on return
    }
}

void gpu_svd_batched(const double* d_mats, int nmat, int n, DevBuf<double>& W, DevBuf<double>& V,
                     hipStream_t s) {
    W.alloc((size_t)nmat * n * n);
    V.alloc((size_t)nmat * n * n);
    if (nmat == 0 || n == 0) return;
    DevBuf<int> st(nmat);
    hipLaunchKernelGGL(k_jacobi_pinv, dim3(nmat), dim3(1024), n * sizeof(double), s, n, 0.0, d_mats, W.p, V.p,
                       (double*)nullptr, st.p);
    TBEM_LAUNCH_CHECK();
    auto hs = st.to_host(s);
    for (int v: hs) TBEM_REQUIRE(v == 0, TBEM_E_NUMERIC, "Jacobi SVD did not converge in 60 sweeps");
}

void compute_c2e(int kid, const KConst& kc, const HostTree& t, const Surfaces& surf, bool up,
                 DevBuf<double>& out, hipStream_t s, DevBuf<double>* fac_w, DevBuf<double>* fac_v) {
    const int R = kernel_dim(kid);
    const int n_e = surf.n_e, n = R * n_e, L = t.n_levels;
    std::vector<double> centers(3 * (size_t)L), radii(L);
    for (int l = 0; l < L; l++) {
        int cell = t.level_begin[l];  // first depth-first cell of the level
        centers[3 * l] = t.cx[cell];
        centers[3 * l + 1] = t.cy[cell];
        centers[3 * l + 2] = t.cz[cell];
        radii[l] = t.level_r[l];
    }
    // up: equivalent = inner (0.3), check = outer; down: equivalent = outer, check = inner
    const std::vector<double>& check = up ? surf.outer : surf.inner;
    const std::vector<double>& equiv = up ? surf.inner : surf.outer;
    DevBuf<double> dc, dr, dck, deq, du, mats((size_t)L * n * n);
    dc.upload(centers, s); dr.upload(radii, s);
    dck.upload(check, s); deq.upload(equiv, s); du.upload(surf.unit, s);
    dim3 grid(std::max(1, std::min(64, div_up((size_t)n_e * n_e, 256))), L);
    hipLaunchKernelGGL(k_c2e_assemble, grid, dim3(256), 0, s, kid, kc, n_e, dc.p, dr.p, dck.p, deq.p,
                       du.p, mats.p);
    TBEM_LAUNCH_CHECK();
    out.alloc((size_t)L * n * n);
    gpu_pinv_batched(mats.p, L, n, 1e-10, out.p, s, fac_w, fac_v);
    TBEM_HIP(hipStreamSynchronize(s));
}

}  // namespace tbem
