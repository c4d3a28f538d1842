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
// then a one-sided (Hestenes) Jacobi SVD, spread over many workgroups per matrix
// (block round-robin, k_bjac_round; a single-workgroup version, k_jacobi_pinv,
// serves sizes whose row blocks do not fit LDS), and pinv = V diag(1/sigma) U^T
// from the rotated columns as a tiled product.
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

// ------------------------------------------------------------ block one-sided Jacobi
// The same SVD spread over many workgroups per matrix. W's n rows (the columns of A)
// are cut into nb blocks of b rows (nb even; the last blocks may be short or empty).
// One launch per round of a round-robin tournament over the blocks: workgroup k of
// matrix m takes block pair k of the round, holds its <= 2b rows of W in LDS and
// runs one cyclic sweep over them (a round-robin over the 2b rows, b disjoint row
// pairs per step, 256/b threads per pair), accumulating the rotations in J (2b x 2b;
// W_new = J W_old for those rows). It then writes the rows back and replaces the
// same rows of V by J V. Over the nb - 1 rounds of a sweep every pair of rows meets
// in some workgroup; a sweep that rotates nothing ends the iteration. Rotation rule,
// tolerance and negligible-column test are those of k_jacobi_pinv.
constexpr int kBjThreads = 256;

__global__ void k_svd_init(int n, const double* __restrict__ A, double* __restrict__ W, double* __restrict__ V,
                           double* __restrict__ tiny) {
    const int m = blockIdx.y;
    const size_t off = (size_t)m * n * n;
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < (size_t)n * n;
         p += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(p / n), k = (int)(p % n);
        W[off + p] = A[off + (size_t)k * n + i];
        V[off + p] = (i == k) ? 1.0 : 0.0;
    }
    if (blockIdx.x == 0) {  // 1e-30 ||A||_F^2: columns below it only carry rounding noise
        __shared__ double red[256];
        double s = 0;
        for (size_t p = threadIdx.x; p < (size_t)n * n; p += blockDim.x) s += A[off + p] * A[off + p];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {  // launched with 256 threads
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) tiny[m] = 1e-30 * red[0];
    }
}

__device__ __forceinline__ void rr_pair(int players, int round, int k, int& p, int& q) {
    if (k == 0) {
        p = players - 1;
        q = round;
    } else {
        p = (round + k) % (players - 1);
        q = (round - k + (players - 1)) % (players - 1);
    }
}

template <int B>
__global__ __launch_bounds__(kBjThreads) void k_bjac_round(int n, int nb, int round, double* __restrict__ Wg,
                                                            double* __restrict__ Vg, const double* __restrict__ tiny,
                                                            int* __restrict__ flags) {
    constexpr int T = kBjThreads / B;  // threads per row pair
    constexpr int R2 = 2 * B;
    extern __shared__ double lds[];
    double* w = lds;                    // R2 x n
    double* J = lds + (size_t)R2 * n;   // R2 x R2
    double* vs = J + R2 * R2;           // R2 x 64: staged V columns
    __shared__ int rotated;
    const int m = blockIdx.y, tid = threadIdx.x;
    int P, Q;
    rr_pair(nb, round, blockIdx.x, P, Q);
    const int cP = max(0, min(n, (P + 1) * B) - P * B), cQ = max(0, min(n, (Q + 1) * B) - Q * B);
    double* W = Wg + (size_t)m * n * n;
    double* V = Vg + (size_t)m * n * n;
    auto grow = [&](int r) { return r < B ? P * B + r : Q * B + (r - B); };
    auto valid = [&](int r) { return r < B ? r < cP : r - B < cQ; };
    for (int e = tid; e < R2 * n; e += kBjThreads) {
        const int r = e / n, c = e - r * n;
        w[e] = valid(r) ? W[(size_t)grow(r) * n + c] : 0.0;
    }
    for (int e = tid; e < R2 * R2; e += kBjThreads) J[e] = (e / R2 == e % R2) ? 1.0 : 0.0;
    if (tid == 0) rotated = 0;
    __syncthreads();
    const double tol = fmax(1e-15, sqrt((double)n) * 2.220446049250313e-16);
    const double tn = tiny[m];
    const int team = tid / T, lane = tid % T;
    for (int rr = 0; rr < R2 - 1; rr++) {
        int p, q;
        rr_pair(R2, rr, team, p, q);
        const bool ok = valid(p) && valid(q);
        double al = 0, be = 0, ga = 0;
        if (ok) {
            const double* wp = w + (size_t)p * n;
            const double* wq = w + (size_t)q * n;
            for (int t = lane; t < n; t += T) {
                const double x = wp[t], y = wq[t];
                al += x * x;
                be += y * y;
                ga += x * y;
            }
        }
        for (int off = T / 2; off > 0; off >>= 1) {
            al += __shfl_xor(al, off, T);
            be += __shfl_xor(be, off, T);
            ga += __shfl_xor(ga, off, T);
        }
        if (ok && al > tn && be > tn && fabs(ga) > tol * sqrt(al) * sqrt(be)) {
            const double zeta = (be - al) / (2 * ga);
            const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
            const double cs = 1 / sqrt(1 + t * t), sn = cs * t;
            double* wp = w + (size_t)p * n;
            double* wq = w + (size_t)q * n;
            for (int t2 = lane; t2 < n; t2 += T) {
                const double x = wp[t2], y = wq[t2];
                wp[t2] = cs * x - sn * y;
                wq[t2] = sn * x + cs * y;
            }
            for (int k = lane; k < R2; k += T) {
                const double x = J[p * R2 + k], y = J[q * R2 + k];
                J[p * R2 + k] = cs * x - sn * y;
                J[q * R2 + k] = sn * x + cs * y;
            }
            if (lane == 0) rotated = 1;
        }
        __syncthreads();
    }
    for (int e = tid; e < R2 * n; e += kBjThreads) {
        const int r = e / n, c = e - r * n;
        if (valid(r)) W[(size_t)grow(r) * n + c] = w[e];
    }
    // V rows <- J V rows, 64 columns at a time
    for (int c0 = 0; c0 < n; c0 += 64) {
        const int nc = min(64, n - c0);
        for (int e = tid; e < R2 * 64; e += kBjThreads) {
            const int r = e / 64, c = e % 64;
            vs[e] = (valid(r) && c < nc) ? V[(size_t)grow(r) * n + c0 + c] : 0.0;
        }
        __syncthreads();
        for (int e = tid; e < R2 * 64; e += kBjThreads) {
            const int r = e / 64, c = e % 64;
            if (!valid(r) || c >= nc) continue;
            double acc = 0;
            for (int k = 0; k < R2; k++) acc += J[r * R2 + k] * vs[k * 64 + c];
            V[(size_t)grow(r) * n + c0 + c] = acc;
        }
        __syncthreads();
    }
    if (tid == 0 && rotated) atomicOr(flags + m, 1);
}

// sig[m][i] = |W_i| (the singular values), sig[m][n] = their maximum
__global__ __launch_bounds__(256) void k_svd_sigma(int n, const double* __restrict__ W, double* __restrict__ sig) {
    const int m = blockIdx.x;
    const double* w = W + (size_t)m * n * n;
    double* sg = sig + (size_t)m * (n + 1);
    __shared__ double red[256];
    double mx = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double s = 0;
        for (int k = 0; k < n; k++) s += w[(size_t)i * n + k] * w[(size_t)i * n + k];
        sg[i] = sqrt(s);
        mx = fmax(mx, sg[i]);
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {  // launched with 256 threads
        if ((int)threadIdx.x < s2) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s2]);
        __syncthreads();
    }
    if (threadIdx.x == 0) sg[n] = red[0];
}

// P[m] = V^T diag(g) W with g_i = 1 / sigma_i^2 where sigma_i > threshold * sigma_max,
// else 0 (pinv(A) = sum_i v_i u_i^T / sigma_i, with W_i = sigma_i u_i); a 64 x 64
// output tile per workgroup, 4 x 4 per thread.
__global__ __launch_bounds__(256) void k_pinv_product(int n, double threshold, const double* __restrict__ Wg,
                                                      const double* __restrict__ Vg, const double* __restrict__ sig,
                                                      double* __restrict__ Pg) {
    const int m = blockIdx.z;
    const double* W = Wg + (size_t)m * n * n;
    const double* V = Vg + (size_t)m * n * n;
    const double* sg = sig + (size_t)m * (n + 1);
    double* P = Pg + (size_t)m * n * n;
    const double cut = threshold * sg[n];
    const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
    __shared__ double As[16][65], Bs[16][65];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    double acc[4][4] = {};
    for (int i0 = 0; i0 < n; i0 += 16) {
        for (int e = threadIdx.x; e < 1024; e += 256) {
            const int k = e / 64, j = e % 64, i = i0 + k;
            const double g = (i < n && sg[i] > cut) ? 1.0 / (sg[i] * sg[i]) : 0.0;
            As[k][j] = (i < n && r0 + j < n) ? V[(size_t)i * n + r0 + j] * g : 0.0;
            Bs[k][j] = (i < n && c0 + j < n) ? W[(size_t)i * n + c0 + j] : 0.0;
        }
        __syncthreads();
        for (int k = 0; k < 16; k++) {
            double a[4], b[4];
            for (int i = 0; i < 4; i++) a[i] = As[k][ty * 4 + i];
            for (int j = 0; j < 4; j++) b[j] = Bs[k][tx * 4 + j];
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) acc[i][j] += a[i] * b[j];
        }
        __syncthreads();
    }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            const int r = r0 + ty * 4 + i, c = c0 + tx * 4 + j;
            if (r < n && c < n) P[(size_t)r * n + c] = acc[i][j];
        }
}

// Rows per block of the block Jacobi for an n x n matrix: the largest of 32, 16, 8,
// 4 whose LDS (2b rows of W, J, the staged V columns) fits 64 KiB with room for the
// kernel's static LDS; 0 when none does.
static int bjac_block(int n) {
    for (int b : {32, 16, 8, 4}) {
        const size_t bytes = ((size_t)2 * b * n + 4 * (size_t)b * b + (size_t)2 * b * 64) * sizeof(double);
        if (bytes + 256 <= 65536 && (n + b - 1) / b >= 2) return b;
    }
    return 0;
}

template <int B>
static void bjac_launch(int n, int nb, int round, int nmat, double* W, double* V, const double* tiny, int* flags,
                        hipStream_t s) {
    const size_t shm = ((size_t)2 * B * n + 4 * (size_t)B * B + (size_t)2 * B * 64) * sizeof(double);
    hipLaunchKernelGGL(k_bjac_round<B>, dim3(nb / 2, nmat), dim3(kBjThreads), shm, s, n, nb, round, W, V, tiny,
                       flags);
    TBEM_LAUNCH_CHECK();
}

// W, V (allocated by the caller, nmat n x n each) <- the SVD factors of the matrices;
// false when the block Jacobi does not serve this size (bjac_block(n) == 0).
static bool block_jacobi_svd(const double* d_mats, int nmat, int n, double* W, double* V, hipStream_t s) {
    const int b = bjac_block(n);
    if (b == 0) return false;
    int nb = (n + b - 1) / b;
    nb += nb & 1;
    DevBuf<double> tiny(nmat);
    DevBuf<int> flags(nmat);
    hipLaunchKernelGGL(k_svd_init, dim3(std::min(256, div_up((size_t)n * n, 256)), nmat), dim3(256), 0, s, n,
                       d_mats, W, V, tiny.p);
    TBEM_LAUNCH_CHECK();
    std::vector<int> hf(nmat);
    int sweep = 0;
    for (; sweep < 60; sweep++) {
        TBEM_HIP(hipMemsetAsync(flags.p, 0, nmat * sizeof(int), s));
        for (int r = 0; r < nb - 1; r++) {
            switch (b) {
            case 32: bjac_launch<32>(n, nb, r, nmat, W, V, tiny.p, flags.p, s); break;
            case 16: bjac_launch<16>(n, nb, r, nmat, W, V, tiny.p, flags.p, s); break;
            case 8: bjac_launch<8>(n, nb, r, nmat, W, V, tiny.p, flags.p, s); break;
            default: bjac_launch<4>(n, nb, r, nmat, W, V, tiny.p, flags.p, s); break;
            }
        }
        TBEM_HIP(hipMemcpyAsync(hf.data(), flags.p, nmat * sizeof(int), hipMemcpyDeviceToHost, s));
        TBEM_HIP(hipStreamSynchronize(s));
        bool any = false;
        for (int f : hf) any = any || f != 0;
        if (!any) break;
    }
    TBEM_REQUIRE(sweep < 60, TBEM_E_NUMERIC, "Jacobi SVD did not converge in 60 sweeps");
    return true;
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
    if (block_jacobi_svd(d_mats, nmat, n, W.p, V.p, s)) {
        DevBuf<double> sig((size_t)nmat * (n + 1));
        hipLaunchKernelGGL(k_svd_sigma, dim3(nmat), dim3(256), 0, s, n, W.p, sig.p);
        TBEM_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_pinv_product, dim3(div_up(n, 64), div_up(n, 64), nmat), dim3(256), 0, s, n, threshold,
                           W.p, V.p, sig.p, d_out);
        TBEM_LAUNCH_CHECK();
        TBEM_HIP(hipStreamSynchronize(s));  // sig is freed on return
    } else {
        DevBuf<int> st(nmat);
        hipLaunchKernelGGL(k_jacobi_pinv, dim3(nmat), dim3(1024), n * sizeof(double), s, n, threshold,
                           d_mats, W.p, V.p, d_out, st.p);
        TBEM_LAUNCH_CHECK();
        auto hs = st.to_host(s);
        for (int v: hs) TBEM_REQUIRE(v == 0, TBEM_E_NUMERIC, "Jacobi SVD did not converge in 60 sweeps");
    }
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
    if (block_jacobi_svd(d_mats, nmat, n, W.p, V.p, s)) return;
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
