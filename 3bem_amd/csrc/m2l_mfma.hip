// M2L of the mixed-precision path as grouped FP32 MFMA GEMMs (gfx950).
//
// Plan: pairs -> classes (level_t, level_s, integer offset). The C2E matrices are
// factored by Jacobi SVD, C2E = sum_i w_i v_i^T (w_i = s_i a_i in equivalent space,
// v_i in check space, dropped modes have s_i ~ 0), which gives per level
//   up:   Bu = [v_i^T] (rows),        Ru = [w_i] (columns)      UC2E = Ru Bu
//   down: Ld = [s_i v_i^T] (rows),    Vd = [w_i / s_i] (columns) DC2E = Vd Ld
// and per class, in FP64 on the GPU: K_c assembled from the kernel's reference
// expression (tensor_ref, as the C2E systems), then M_c = Ld * (K_c * Ru) by a tiled
// FP64 GEMM, rounded to FP32 in MFMA A-operand order. The FP32 data are the
// up-check potential in modal form, Bu * check (computed in FP64 by the P2M/M2M
// stage), and M_c's output z; the target adds Vd * sum(z) in FP64. In this
// basis each entry and the input it multiplies scale together mode by mode, so the
// FP32 sums carry no cancellation: measured in a single-pair model at orders
// 20-150, <= 5e-7 relative, against up to 3e-4 for the operator DC2E K_c UC2E
// applied in FP32 in the physical basis. Work units are (class, up to 64 pairs); units of
// one class go to one XCD (workgroup i runs on XCD i % 8), so each S_c is fetched
// from HBM once and then served by that XCD's L2.
//
// Apply: one workgroup (4 or 8 wave64s) per unit stages its 64 up-check potentials in
// LDS as FP32 (k x 64; column n of row k stored at n ^ 16*(k%4), which keeps both
// the staging writes and the B-operand reads free of bank conflicts), then each
// wave runs row tiles of S_c:
// v_mfma_f32_16x16x4f32 with A from global (coalesced 256 B per wave and k-step,
// L2-resident) and B from LDS, FP32 accumulation over k = blk. Results go to a
// per-pair buffer Y (no atomics); k_m2l_pair_sum sums each target's pairs in FP64,
// and a batched per-level GEMM adds Vd times the sums to the down-equivalent densities.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <type_traits>

#include "c2e.h"
#include "m2l_mfma.h"
#include "stages.h"

namespace tbem {

namespace {

constexpr int kUnitCols = 64;   // pairs per work unit (= LDS row length of the staged columns)
constexpr int kUnitCols64 = 32; // the same for the FP64 operators
// The FP64 kernel's shape, as for the FP32 one by how many workgroups the staged
// columns leave per CU: eight waves with A 8 k-steps ahead where at most two fit
// (order 100 elastic, blk 297), four waves 4 k-steps ahead otherwise (order 75
// Laplace, blk 74: M2L stage 52.5 -> 42.3 ms; 4 waves with 8 ahead 46.5, 2 waves 43.4).
constexpr int kPre64 = 8;
constexpr int kWaves64 = 8;
constexpr int kPre64Narrow = 4;
constexpr int kWaves64Narrow = 4;
constexpr int kMaxBlk = 512;    // staged columns: up to 512 x 64 x 4 B = 128 KiB of LDS
// Below this operator size the FP64 M2L stage is the cheaper one: measured on the
// 1M-point elastic workload at order 6 (blk 18), the MFMA path with its per-pair
// buffer and modal products took longer than the 0.4 ms FP64 stage; at blk 60
// (order 20) it is already faster (2M points: 3.0 vs 5.2 ms).
constexpr int kMinBlk = 48;
constexpr int kXcds = 8;
constexpr size_t kLdsPerCu = 160 * 1024;  // gfx950
// k-steps of A loaded ahead (see m2l_row_tile): the wide kernel (eight waves, two
// workgroups per CU) hides more latency per wave with 16; the narrow one, with many
// workgroups per CU, does better with 8 (order 75 Laplace: 15.2 -> 14.0 ms).
constexpr int kPreWide = 16;
constexpr int kPreNarrow = 8;
constexpr int kMaxChunk = 16384;  // classes per plan-time batch (grid y/z limit)

typedef float f32x4 __attribute__((ext_vector_type(4)));

// K_c for a chunk of classes: rows (d1, i) = component d1 at target check point i
// (tsurf, scaled by the target cell's radius), columns (d2, j) = component d2 of
// source equivalent point j (ssurf, scaled by the source cell's radius).
__global__ void k_m2l_assemble(int kid, KConst kc, int n_e, int R, const double* __restrict__ tc,
                               const double* __restrict__ sc, const double* __restrict__ tr,
                               const double* __restrict__ sr, const double* __restrict__ tsurf,
                               const double* __restrict__ ssurf, const double* __restrict__ unit,
                               double* __restrict__ mats) {
#pragma clang fp contract(off)
    const int c = blockIdx.y;
    const int n = R * n_e;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n_e * n_e; p += gridDim.x * blockDim.x) {
        const int i = p / n_e, j = p % n_e;
        double ck[3], eq[3];
        for (int d = 0; d < 3; d++) {
            ck[d] = tr[c] * tsurf[3 * i + d] + tc[3 * c + d];
            eq[d] = sr[c] * ssurf[3 * j + d] + sc[3 * c + d];
        }
        double kv[9];
        tensor_ref_dispatch(kid, ck, eq, unit + 3 * i, unit + 3 * j, kc, kv);
        double* M = mats + (size_t)c * n * n;
        for (int d1 = 0; d1 < R; d1++)
            for (int d2 = 0; d2 < R; d2++) M[(size_t)(d1 * n_e + i) * n + d2 * n_e + j] = kv[d1 * R + d2];
    }
}

// C[c] = A[ai[c]] * B[bi[c]], n x n row-major FP64; 64 x 64 tile per workgroup.
__global__ __launch_bounds__(256) void k_gemm64(int n, const double* __restrict__ A0, const int* __restrict__ ai,
                                                const double* __restrict__ B0, const int* __restrict__ bi,
                                                double* __restrict__ C0) {
    const int c = blockIdx.z;
    const double* A = A0 + (size_t)ai[c] * n * n;
    const double* B = B0 + (size_t)bi[c] * n * n;
    double* C = C0 + (size_t)c * n * n;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    __shared__ double As[16][65], Bs[16][65];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    double acc[4][4] = {};
    for (int k0 = 0; k0 < n; k0 += 16) {
        for (int e = threadIdx.x; e < 1024; e += 256) {
            const int r = e / 16, k = e % 16;
            As[k][r] = (m0 + r < n && k0 + k < n) ? A[(size_t)(m0 + r) * n + k0 + k] : 0.0;
            const int kb = e / 64, cb = e % 64;
            Bs[kb][cb] = (k0 + kb < n && n0 + cb < n) ? B[(size_t)(k0 + kb) * n + n0 + cb] : 0.0;
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
            const int r = m0 + ty * 4 + i, cc = n0 + tx * 4 + j;
            if (r < n && cc < n) C[(size_t)r * n + cc] = acc[i][j];
        }
}

// FP64 row-major S_c (or its transpose) -> A-operand order of v_mfma_f32_16x16x4f32
// (T = float) or v_mfma_f64_16x16x4f64 (T = double), which share it; zero padded to
// mt*16 x kt*4.
template <class T>
__global__ void k_m2l_swizzle(int n, int mt, int kt, int trans, const double* __restrict__ S64,
                              T* __restrict__ out) {
    const int c = blockIdx.y;
    const size_t per = (size_t)mt * kt * 64;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < per; e += (size_t)gridDim.x * blockDim.x) {
        const int rt = (int)(e / ((size_t)kt * 64)), ks = (int)((e / 64) % kt), l = (int)(e % 64);
        const int row = rt * 16 + l % 16, col = ks * 4 + l / 16;
        const size_t src = trans ? (size_t)col * n + row : (size_t)row * n + col;
        out[c * per + e] = (row < n && col < n) ? (T)S64[(size_t)c * n * n + src] : T(0);
    }
}

// One row tile (16 rows of M_c) times the NCT staged column tiles. A comes from
// global memory (L2-resident), loaded kPre k-steps ahead of its use so that the
// loads' latency overlaps kPre * NCT MFMAs. The last block of a row tile already
// loads the first kPre k-steps of the wave's next row tile (An; null for none), so
// the pipeline does not drain between row tiles: cur holds A's first kPre k-steps
// on entry and An's on exit.
template <int NCT, int kPre>
__device__ __forceinline__ void m2l_row_tile(const float* __restrict__ A, const float* __restrict__ An,
                                             const float* __restrict__ b, int kt, int s0, int s1, int s2, int s3,
                                             f32x4* acc, float* cur) {
    float nxt[kPre];
    for (int k0 = 0; k0 < kt; k0 += kPre) {
        if (k0 + kPre >= kt) {
            for (int i = 0; i < kPre; i++) nxt[i] = (An && i < kt) ? An[(size_t)i * 64] : 0.f;
        } else {
            for (int i = 0; i < kPre; i++) nxt[i] = k0 + kPre + i < kt ? A[(size_t)(k0 + kPre + i) * 64] : 0.f;
        }
#pragma unroll
        for (int i = 0; i < kPre; i++) {
            if (k0 + i >= kt) break;
            const float* bk = b + (k0 + i) * 4 * 64;
            acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(cur[i], bk[s0], acc[0], 0, 0, 0);
            if (NCT > 1) acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(cur[i], bk[s1], acc[1], 0, 0, 0);
            if (NCT > 2) acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(cur[i], bk[s2], acc[2], 0, 0, 0);
            if (NCT > 3) acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(cur[i], bk[s3], acc[3], 0, 0, 0);
        }
        for (int i = 0; i < kPre; i++) cur[i] = nxt[i];
    }
}

// NW waves per workgroup share the unit's staged columns; wave w takes row tiles
// w, w + NW, ... With large operators the staged columns (up to 128 KiB) leave room
// for two workgroups per CU, and eight waves per workgroup then keep more MFMA
// work in flight (order 100 elastic, blk 297: 5.65 -> 4.11 ms); with small operators
// many workgroups fit anyway and four waves balance the row tiles better (order 75
// Laplace, blk 74: 15.2 ms at four waves against 16.5 at eight).
template <int NW, int kPre>
__global__ __launch_bounds__(NW * 64) void k_m2l_mfma(int n, int mt, int kt, int ystride, const float* __restrict__ S,
                                                  const int* __restrict__ ucls, const int* __restrict__ uoff,
                                                  const int* __restrict__ ucnt, const int* __restrict__ colpair,
                                                  const int* __restrict__ pair_src, int ybase,
                                                  const float* __restrict__ chk, float* __restrict__ Y) {
    extern __shared__ float Bs[];  // [kt*4][64], swizzled
    const int u = blockIdx.x;
    const int cls = uni(ucls[u]), off = uni(uoff[u]), cnt = uni(ucnt[u]);
    if (cnt == 0) return;  // padding unit: uniform exit before any barrier
    const int lane = threadIdx.x & 63, wave = uni(threadIdx.x >> 6);
    const int kpad = kt * 4;
    // stage: lane = column (pair), waves stride over k
    const float* col = lane < cnt ? chk + (size_t)pair_src[colpair[off + lane]] * n : nullptr;
    for (int kk = wave; kk < kpad; kk += NW)
        Bs[kk * 64 + (lane ^ ((kk & 3) << 4))] = (col && kk < n) ? col[kk] : 0.f;
    __syncthreads();
    // B operand of column tile ct at k-step ks: row ks*4 + lq, column ct*16 + lr,
    // stored at ((ct ^ lq) * 16 + lr)
    const int lr = lane % 16, lq = lane / 16;
    const float* b = Bs + lq * 64 + lr;
    const int s0 = (0 ^ lq) * 16, s1 = (1 ^ lq) * 16, s2 = (2 ^ lq) * 16, s3 = (3 ^ lq) * 16;
    const int nct = (cnt + 15) / 16;  // column tiles holding pairs (uniform)
    const float* Acls = S + (size_t)cls * mt * kt * 64 + lane;
    float cur[kPre];
    for (int i = 0; i < kPre; i++) cur[i] = (wave < mt && i < kt) ? Acls[((size_t)wave * kt + i) * 64] : 0.f;
    for (int rt = wave; rt < mt; rt += NW) {
        f32x4 acc[4] = {};
        const float* A = Acls + (size_t)rt * kt * 64;
        const float* An = rt + NW < mt ? A + (size_t)NW * kt * 64 : nullptr;
        switch (nct) {
        case 1: m2l_row_tile<1, kPre>(A, An, b, kt, s0, s1, s2, s3, acc, cur); break;
        case 2: m2l_row_tile<2, kPre>(A, An, b, kt, s0, s1, s2, s3, acc, cur); break;
        case 3: m2l_row_tile<3, kPre>(A, An, b, kt, s0, s1, s2, s3, acc, cur); break;
        default: m2l_row_tile<4, kPre>(A, An, b, kt, s0, s1, s2, s3, acc, cur); break;
        }
        // D[4*lq + r][lr] of column tile ct -> Y[pair][rt*16 + 4*lq + r]
        for (int ct = 0; ct < nct; ct++) {
            const int nn = ct * 16 + lr;
            if (nn < cnt) *(f32x4*)(Y + (size_t)(colpair[off + nn] - ybase) * ystride + rt * 16 + 4 * lq) = acc[ct];
        }
    }
}

// The pairs of item k whose outputs are in Y (those of one slice of the pairs, Y[0]
// holding pair ybase): the range ptr[k] .. ptr[k+1] (forward: the slice's items), or,
// with pairs given (the transpose's source-side lists, each ascending), the entries
// of pairs[ptr[k] .. ptr[k+1]] within [pb, pe).
__device__ __forceinline__ void slice_pairs(const int* __restrict__ ptr, const int* __restrict__ pairs, int k, int pb,
                                            int pe, int& j0, int& j1) {
    j0 = uni(ptr[k]);
    j1 = uni(ptr[k + 1]);
    if (!pairs) return;
    auto lower = [&](int lo, int hi, int v) {
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (pairs[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    const int a = lower(j0, j1, pb);
    j1 = uni(lower(a, j1, pe));
    j0 = uni(a);
}

// z[t] (= or += with add) sum_p Y[p] over the pairs p of target t in this slice, in
// FP64 (blk entries per cell); the Vd product follows as a batched per-level GEMM.
// One 256-thread workgroup per target. The target's pairs are contiguous in Y, so the
// block is read as consecutive float4s: thread i takes float4 column i % q (q =
// ystride / 4) of every npar-th pair starting at i / q, and the npar partial sums are
// combined through LDS.
__global__ __launch_bounds__(256) void k_m2l_pair_sum(const int* __restrict__ items, const int* __restrict__ ptr,
                                                      const int* __restrict__ pairs, const float* __restrict__ Y,
                                                      int ystride, int blk, int ybase, int pb, int pe, int add,
                                                      double* __restrict__ z) {
    extern __shared__ double part[];  // npar * ystride partial sums
    const int k = blockIdx.x;
    const int t = uni(items[k]);
    int p0, p1;
    slice_pairs(ptr, pairs, k, pb, pe, p0, p1);
    const int q = ystride / 4, npar = 256 / q;
    const int c = threadIdx.x % q, ph = threadIdx.x / q;
    if (ph < npar) {
        double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        const f32x4* y4 = (const f32x4*)Y;
        for (int j = p0 + ph; j < p1; j += npar) {
            const size_t p = (size_t)((pairs ? pairs[j] : j) - ybase);
            const f32x4 v = y4[p * q + c];
            s0 += (double)v[0];
            s1 += (double)v[1];
            s2 += (double)v[2];
            s3 += (double)v[3];
        }
        double* o = part + (size_t)ph * ystride + 4 * c;
        o[0] = s0; o[1] = s1; o[2] = s2; o[3] = s3;
    }
    __syncthreads();
    for (int row = threadIdx.x; row < blk; row += blockDim.x) {
        double s = 0;
        for (int h = 0; h < npar; h++) s += part[(size_t)h * ystride + row];
        double* o = z + (size_t)t * blk + row;
        *o = add ? *o + s : s;
    }
}

// ------------------------------------------------------------- FP64 operators
// The FP64 path's M2L on the same classes: Y[p] = K_c * mult[src(p)] with K_c the
// kernel between the target's down-check and the source's up-equivalent surface,
// in FP64 (v_mfma_f64_16x16x4f64, FP64 accumulation), then the target's check
// potential is the sum of its pairs' Y in pair order and the FP64 DC2E follows as in
// the reference. Units are (class, up to 32 pairs): the staged columns (k x 32
// doubles, 77 KB at blk 297) leave room for two workgroups per CU.
typedef double f64x4 __attribute__((ext_vector_type(4)));

// One row tile (16 rows of K_c) times NCT column tiles; A from global memory
// (L2-resident) loaded kPre k-steps ahead, continuing into the wave's next row tile
// (An) as m2l_row_tile does.
template <int NCT, int kPre>
__device__ __forceinline__ void m2l_row_tile64(const double* __restrict__ A, const double* __restrict__ An,
                                               const double* __restrict__ b, int kt, f64x4* acc, double* cur) {
    double nxt[kPre];
    for (int k0 = 0; k0 < kt; k0 += kPre) {
        if (k0 + kPre >= kt) {
            for (int i = 0; i < kPre; i++) nxt[i] = (An && i < kt) ? An[(size_t)i * 64] : 0.0;
        } else {
            for (int i = 0; i < kPre; i++) nxt[i] = k0 + kPre + i < kt ? A[(size_t)(k0 + kPre + i) * 64] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < kPre; i++) {
            if (k0 + i >= kt) break;
            const double* bk = b + (k0 + i) * 4 * kUnitCols64;
#pragma unroll
            for (int ct = 0; ct < NCT; ct++)
                acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(cur[i], bk[16 * ct], acc[ct], 0, 0, 0);
        }
        for (int i = 0; i < kPre; i++) cur[i] = nxt[i];
    }
}

// One workgroup of NW waves per unit: the unit's source coefficients staged in LDS
// (row k, column = pair, kpad x 32 doubles), then wave w takes row tiles w, w + NW, ...
// The FP64 MFMA leaves D[4*r + lq][lr] in register r of lane (lq, lr) (not the f32
// 16x16x4 order), so Y[pair][rt*16 + 4*r + lq] = D[4*r + lq][lr] of column tile ct
// (pair ct*16 + lr): the four lanes lq of a pair write 32 contiguous bytes per r.
template <int NW, int kPre>
__global__ __launch_bounds__(NW * 64) void k_m2l_mfma64(int n, int mt, int kt, int ystride,
                                                     const double* __restrict__ S, const int* __restrict__ ucls,
                                                     const int* __restrict__ uoff, const int* __restrict__ ucnt,
                                                     const int* __restrict__ colpair, const int* __restrict__ pair_src,
                                                     int ybase, const double* __restrict__ x, double* __restrict__ Y) {
    extern __shared__ double Bd[];  // [kt*4][kUnitCols64]
    const int u = blockIdx.x;
    const int cls = uni(ucls[u]), off = uni(uoff[u]), cnt = uni(ucnt[u]);
    if (cnt == 0) return;  // padding unit: uniform exit before any barrier
    const int lane = threadIdx.x & 63, wave = uni(threadIdx.x >> 6);
    const int kpad = kt * 4;
    // stage: lane % kUnitCols64 = column (pair); the rest of the lane index and the
    // waves stride over k
    constexpr int RPW = 64 / kUnitCols64;  // k rows per wave and pass
    const int c = lane % kUnitCols64, h = lane / kUnitCols64;
    const double* col = c < cnt ? x + (size_t)pair_src[colpair[off + c]] * n : nullptr;
    for (int kk = RPW * wave + h; kk < kpad; kk += RPW * NW) Bd[kk * kUnitCols64 + c] = (col && kk < n) ? col[kk] : 0.0;
    __syncthreads();
    // B operand of column tile ct at k-step ks: row ks*4 + lq, column ct*16 + lr
    const int lr = lane % 16, lq = lane / 16;
    const double* b = Bd + lq * kUnitCols64 + lr;
    constexpr int MAXCT = kUnitCols64 / 16;
    const int nct = (cnt + 15) / 16;  // 1 .. MAXCT (uniform)
    const double* Acls = S + (size_t)cls * mt * kt * 64 + lane;
    double cur[kPre];
    for (int i = 0; i < kPre; i++) cur[i] = (wave < mt && i < kt) ? Acls[((size_t)wave * kt + i) * 64] : 0.0;
    for (int rt = wave; rt < mt; rt += NW) {
        f64x4 acc[MAXCT] = {};
        const double* A = Acls + (size_t)rt * kt * 64;
        const double* An = rt + NW < mt ? A + (size_t)NW * kt * 64 : nullptr;
        if (nct == 1) m2l_row_tile64<1, kPre>(A, An, b, kt, acc, cur);
        else if (MAXCT == 2 || nct == 2) m2l_row_tile64<(MAXCT < 2 ? MAXCT : 2), kPre>(A, An, b, kt, acc, cur);
        else if (nct == 3) m2l_row_tile64<(MAXCT < 3 ? MAXCT : 3), kPre>(A, An, b, kt, acc, cur);
        else m2l_row_tile64<MAXCT, kPre>(A, An, b, kt, acc, cur);
        for (int ct = 0; ct < nct; ct++) {
            const int nn = ct * 16 + lr;
            if (nn < cnt) {
                double* y = Y + (size_t)(colpair[off + nn] - ybase) * ystride + rt * 16 + lq;
                for (int r = 0; r < 4; r++) y[4 * r] = acc[ct][r];
            }
        }
    }
}

// out[t] (= or += with add) sum over the pairs of item k in this slice (slice_pairs)
// of Y[p][0 .. blk), FP64. As k_m2l_pair_sum: thread i takes double2 column i % q
// (q = ystride / 2) of every npar-th pair from i / q on, and the npar partial sums
// (each in pair order) are added in order through LDS. With small operators most of
// the workgroup would otherwise idle while blk threads each walk all the target's
// pairs (4M-point Laplace, order 75: blk 74 of 256 threads, 2.1 ms per 6 GB slice).
// Above 256 rows (q > 128) npar is 1 and each thread takes two columns in turn.
__global__ __launch_bounds__(256) void k_m2l_pair_sum64(const int* __restrict__ items, const int* __restrict__ ptr,
                                                        const int* __restrict__ pairs, const double* __restrict__ Y,
                                                        int ystride, int blk, int ybase, int pb, int pe, int add,
                                                        double* __restrict__ out) {
    extern __shared__ double part64[];  // npar * ystride partial sums (npar > 1 only)
    const int k = blockIdx.x;
    const int t = uni(items[k]);
    int p0, p1;
    slice_pairs(ptr, pairs, k, pb, pe, p0, p1);
    const int q = ystride / 2, npar = max(1, 256 / q);
    typedef double f64x2 __attribute__((ext_vector_type(2)));
    const f64x2* y2 = (const f64x2*)Y;
    if (npar == 1) {
        for (int c = threadIdx.x; c < q; c += blockDim.x) {
            double s0 = 0, s1 = 0;
            for (int j = p0; j < p1; j++) {
                const size_t p = (size_t)((pairs ? pairs[j] : j) - ybase);
                const f64x2 v = y2[p * q + c];
                s0 += v[0];
                s1 += v[1];
            }
            for (int h = 0; h < 2; h++) {
                const int row = 2 * c + h;
                if (row >= blk) break;
                double* o = out + (size_t)t * blk + row;
                const double s = h ? s1 : s0;
                *o = add ? *o + s : s;
            }
        }
        return;
    }
    const int c = threadIdx.x % q, ph = threadIdx.x / q;
    if (ph < npar) {
        double s0 = 0, s1 = 0;
        for (int j = p0 + ph; j < p1; j += npar) {
            const size_t p = (size_t)((pairs ? pairs[j] : j) - ybase);
            const f64x2 v = y2[p * q + c];
            s0 += v[0];
            s1 += v[1];
        }
        part64[(size_t)ph * ystride + 2 * c] = s0;
        part64[(size_t)ph * ystride + 2 * c + 1] = s1;
    }
    __syncthreads();
    for (int row = threadIdx.x; row < blk; row += blockDim.x) {
        double s = 0;
        for (int h = 0; h < npar; h++) s += part64[(size_t)h * ystride + row];
        double* o = out + (size_t)t * blk + row;
        *o = add ? *o + s : s;
    }
}

// Per-level modal factors of C2E = sum_i W_i V_i^T (rows of W and V from the
// Jacobi SVD). Modes with |W_i| <= 1e-13 max_j |W_j| (the ones the pseudo-inverse
// dropped) are zeroed. up: Bu rows V_i, Ru columns W_i; down: Ld rows s_i V_i,
// Vd columns W_i / s_i.
__global__ __launch_bounds__(256) void k_modal_factors(int n, int up, const double* __restrict__ W,
                                                       const double* __restrict__ V, double* __restrict__ Lrows,
                                                       double* __restrict__ Rcols) {
    extern __shared__ double sv[];  // n norms + 1
    const int L = blockIdx.x;
    const double* w = W + (size_t)L * n * n;
    const double* v = V + (size_t)L * n * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double s = 0;
        for (int j = 0; j < n; j++) s += w[(size_t)i * n + j] * w[(size_t)i * n + j];
        sv[i] = sqrt(s);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double mx = 0;
        for (int i = 0; i < n; i++) mx = fmax(mx, sv[i]);
        sv[n] = mx;
    }
    __syncthreads();
    const double cut = 1e-13 * sv[n];
    double* Lr = Lrows + (size_t)L * n * n;
    double* Rc = Rcols + (size_t)L * n * n;
    for (size_t e = threadIdx.x; e < (size_t)n * n; e += blockDim.x) {
        const int i = (int)(e / n), j = (int)(e % n);
        const bool keep = sv[i] > cut;
        const double vij = v[(size_t)i * n + j], wij = w[(size_t)i * n + j];
        Lr[(size_t)i * n + j] = keep ? (up ? vij : sv[i] * vij) : 0.0;     // row i
        Rc[(size_t)j * n + i] = keep ? (up ? wij : wij / sv[i]) : 0.0;    // column i
    }
}

// target of every pair: pair p of item k (p in [ptr[k], ptr[k+1])) has target items[k]
__global__ void k_pair_targets(const int* __restrict__ items, const int* __restrict__ ptr, int* __restrict__ tgt) {
    const int k = blockIdx.x, t = items[k];
    for (int p = ptr[k] + threadIdx.x; p < ptr[k + 1]; p += blockDim.x) tgt[p] = t;
}

// Translation class of pair p: (level t, level s, integer offset of the centres in
// units of the finer half width per axis, each biased by 2^17 into 18 bits). bad is
// set when an offset is not within tol of an integer (cells not on one grid) or
// out of range. idx[p] = p (the values the sort carries).
__global__ void k_class_keys(int n, const int* __restrict__ tgt, const int* __restrict__ src, TreeView ot,
                             const double* __restrict__ ohw, TreeView st, const double* __restrict__ shw, double tol,
                             unsigned long long* __restrict__ key, int* __restrict__ idx, int* __restrict__ bad) {
#pragma clang fp contract(off)
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int t = tgt[p], c = src[p];
    const int Lt = ot.level[t], Ls = st.level[c];
    const double ct[3] = {ot.cx[t], ot.cy[t], ot.cz[t]}, cs[3] = {st.cx[c], st.cy[c], st.cz[c]};
    unsigned long long k = ((unsigned long long)Lt << 59) | ((unsigned long long)Ls << 54);
    bool ok = true;
    for (int a = 0; a < 3; a++) {
        const double hf = fmin(ohw[3 * Lt + a], shw[3 * Ls + a]);
        const double q = (cs[a] - ct[a]) / hf;
        const double qi = rint(q);
        ok = ok && fabs(q - qi) < tol && fabs(qi) < (double)(1 << 17);
        k |= (unsigned long long)((long long)qi + (1 << 17)) << (36 - 18 * a);
    }
    key[p] = k;
    idx[p] = p;
    if (!ok) atomicOr(bad, 1);
}

// flag[i] = 1 where sorted key i starts a run of equal keys
template <class K>
__global__ void k_run_starts(int n, const K* __restrict__ key, int* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

// start[pos[i]] = i for every run start i; start[runs] = n
__global__ void k_run_compact(int n, const int* __restrict__ flag, const int* __restrict__ pos, int runs,
                              int* __restrict__ start) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) start[pos[i]] = i;
    if (i == 0) start[runs] = n;
}

// the representative pair of each class: its first pair (in pair order)
__global__ void k_class_reps(int nc, const int* __restrict__ cstart, const int* __restrict__ sorted_pairs,
                             const int* __restrict__ tgt, const int* __restrict__ src, int* __restrict__ rep_t,
                             int* __restrict__ rep_s) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int p = sorted_pairs[cstart[c]];
    rep_t[c] = tgt[p];
    rep_s[c] = src[p];
}

// split[c * (S+1) + q] = first position in class c's run of sorted pairs
// (colpair[cstart[c] .. cstart[c+1]), ascending) holding a pair >= slice_pairs[q]
__global__ void k_class_splits(int nc, int S, const int* __restrict__ cstart, const int* __restrict__ colpair,
                               const int* __restrict__ slice_pairs, int* __restrict__ split) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int a = cstart[c], b = cstart[c + 1];
    for (int q = 0; q <= S; q++) {
        const int v = slice_pairs[q];
        int lo = a, hi = b;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (colpair[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        split[(size_t)c * (S + 1) + q] = lo;
    }
}

// key of a pair for the transpose's grouping: its source cell
__global__ void k_source_keys(int n, const int* __restrict__ src, unsigned int* __restrict__ key,
                              int* __restrict__ idx) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    key[p] = (unsigned int)src[p];
    idx[p] = p;
}

__global__ void k_gather_int(int n, const int* __restrict__ idx, const unsigned int* __restrict__ v,
                             int* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (int)v[idx[i]];
}

}  // namespace

static size_t staged_bytes(int n, bool fp64) {
    return (size_t)((n + 3) / 4) * 4 * (fp64 ? kUnitCols64 * sizeof(double) : kUnitCols * sizeof(float));
}

static bool size_ok(int n, bool fp64) {
    if (n > kMaxBlk || n < kMinBlk) return false;
    // the staged columns must fit one workgroup's LDS on this device
    int dev = 0, max_shm = 0;
    TBEM_HIP(hipGetDevice(&dev));
    TBEM_HIP(hipDeviceGetAttribute(&max_shm, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    return staged_bytes(n, fp64) <= (size_t)max_shm;
}

bool m2l_mfma_size_ok(int n) { return size_ok(n, false); }
bool m2l_fp64_size_ok(int n) { return size_ok(n, true); }

void build_m2l_mfma(M2lMfma& M, int kid, const KConst& kc, const HostTree& ot, const HostTree& st,
                    const DevTree& dot, const DevTree& dst, bool same_grid, bool fp64, size_t y_budget,
                    const HostList& m2l, const DevList& dm2l, const Surfaces& surf, const double* d_uc2e,
                    const double* d_dc2e, hipStream_t s, const DevBuf<double>* up_fac, const DevBuf<double>* down_fac,
                    const std::vector<double>* t_surf, const std::vector<double>* s_surf) {
    auto t0 = std::chrono::steady_clock::now();
    M = M2lMfma{};
    const int R = kernel_dim(kid), n_e = surf.n_e, n = R * n_e;
    M.blk = n;
    M.fp64 = fp64;
    if (!same_grid || !size_ok(n, fp64) || m2l.cell.empty()) return;
    const size_t shm = staged_bytes(n, fp64);
    if (shm > 65536) {
        if (fp64) {
            TBEM_HIP(hipFuncSetAttribute((const void*)k_m2l_mfma64<kWaves64, kPre64>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
            TBEM_HIP(hipFuncSetAttribute((const void*)k_m2l_mfma64<kWaves64Narrow, kPre64Narrow>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        } else {
            TBEM_HIP(hipFuncSetAttribute((const void*)k_m2l_mfma<4, kPreNarrow>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
            TBEM_HIP(hipFuncSetAttribute((const void*)k_m2l_mfma<8, kPreWide>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        }
    }
    const int n_pairs = (int)m2l.cell.size();

    // classes: (level_t, level_s, q) with q = (c_s - c_t) / finer half-width, integral;
    // on the GPU: a key per pair, a stable sort of the pairs by key, the key runs as
    // classes (numbered in key order; each represented by its first pair)
    const int gp = div_up(std::max(n_pairs, 1), 256);
    DevBuf<int> tgt(n_pairs), colpair(n_pairs), itmp(n_pairs), flag(n_pairs), pos(n_pairs), bad(1);
    DevBuf<unsigned long long> key(n_pairs), ktmp(n_pairs);
    DevBuf<double> ohw, shw;
    ohw.upload(ot.level_hw, s);
    shw.upload(st.level_hw, s);
    TBEM_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_pair_targets, dim3(dm2l.n_items), dim3(64), 0, s, dm2l.items.p, dm2l.ptr.p, tgt.p);
    TBEM_LAUNCH_CHECK();
    // FP64 operators serve every pair of a class from its first pair's geometry, so
    // their classes must be exact to rounding: offsets within 1e-10 of the integers
    // (a relative change of the kernel below 1e-10 even for touching cells; in
    // practice the offsets are exact to ~1e-13). The FP32 operators allow 1e-6.
    hipLaunchKernelGGL(k_class_keys, dim3(gp), dim3(256), 0, s, n_pairs, tgt.p, dm2l.cell.p, view(dot), ohw.p,
                       view(dst), shw.p, fp64 ? 1e-10 : 1e-6, key.p, colpair.p, bad.p);
    TBEM_LAUNCH_CHECK();
    int h_bad = 0;
    TBEM_HIP(hipMemcpyAsync(&h_bad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    TBEM_HIP(hipStreamSynchronize(s));
    if (h_bad) return;  // not on one grid
    radix_sort_pairs_u64(key.p, colpair.p, ktmp.p, itmp.p, n_pairs, 64, s);
    hipLaunchKernelGGL(k_run_starts<unsigned long long>, dim3(gp), dim3(256), 0, s, n_pairs, key.p, flag.p);
    TBEM_LAUNCH_CHECK();
    const int nc = device_exclusive_scan(flag.p, n_pairs, pos.p, s);
    DevBuf<int> cstart(nc + 1), drep_t(nc), drep_s(nc);
    hipLaunchKernelGGL(k_run_compact, dim3(gp), dim3(256), 0, s, n_pairs, flag.p, pos.p, nc, cstart.p);
    TBEM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_class_reps, dim3(div_up(nc, 256)), dim3(256), 0, s, nc, cstart.p, colpair.p, tgt.p,
                       dm2l.cell.p, drep_t.p, drep_s.p);
    TBEM_LAUNCH_CHECK();
    std::vector<int> rep_t(nc), rep_s(nc), hstart(nc + 1);
    drep_t.download(rep_t.data(), nc, s);
    drep_s.download(rep_s.data(), nc, s);
    cstart.download(hstart.data(), nc + 1, s);
    TBEM_HIP(hipStreamSynchronize(s));
    const int mt = (n + 15) / 16, kt = (n + 3) / 4;
    M.n_classes = nc;
    M.mt = mt;
    M.kt = kt;
    M.ystride = mt * 16;
    M.n_pairs = n_pairs;
    if (fp64) {
        // the FP64 operators and their transposes must fit an eighth of the device
        size_t free_b = 0, total_b = 0;
        TBEM_HIP(hipMemGetInfo(&free_b, &total_b));
        if ((size_t)nc * mt * kt * 64 * 2 * sizeof(double) > total_b / 8) return;
    }

    // modal factors of UC2E (source tree levels) and DC2E (observation tree levels)
    const size_t nn = (size_t)n * n, per = (size_t)mt * kt * 64;
    DevBuf<double> Ru, Ld;
    if (!fp64) {
        DevBuf<double> W, V;
        const bool have_up = up_fac && up_fac[0].n && up_fac[1].n, have_down = down_fac && down_fac[0].n && down_fac[1].n;
        if (!have_up) gpu_svd_batched(d_uc2e, st.n_levels, n, W, V, s);
        M.Bu.alloc((size_t)st.n_levels * nn);
        Ru.alloc((size_t)st.n_levels * nn);
        hipLaunchKernelGGL(k_modal_factors, dim3(st.n_levels), dim3(256), (n + 1) * sizeof(double), s, n, 1,
                           have_up ? up_fac[0].p : W.p, have_up ? up_fac[1].p : V.p, M.Bu.p, Ru.p);
        TBEM_LAUNCH_CHECK();
        if (!have_down) gpu_svd_batched(d_dc2e, ot.n_levels, n, W, V, s);
        Ld.alloc((size_t)ot.n_levels * nn);
        M.Vd.alloc((size_t)ot.n_levels * nn);
        hipLaunchKernelGGL(k_modal_factors, dim3(ot.n_levels), dim3(256), (n + 1) * sizeof(double), s, n, 0,
                           have_down ? down_fac[0].p : W.p, have_down ? down_fac[1].p : V.p, Ld.p, M.Vd.p);
        TBEM_LAUNCH_CHECK();
        TBEM_HIP(hipStreamSynchronize(s));
    }

    // operators M_c = Ld * K_c * Ru (FP64 path: K_c itself), in chunks of classes
    // (FP64 temporaries <= ~0.5 GB each)
    if (fp64) {
        M.S64.alloc((size_t)nc * per);
        M.ST64.alloc((size_t)nc * per);
    } else {
        M.S.alloc((size_t)nc * per);
        M.ST.alloc((size_t)nc * per);
    }
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>({(size_t)nc, (size_t)kMaxChunk, (size_t)(5e8 / (nn * 8))}));
    DevBuf<double> K((size_t)chunk * nn), T1(fp64 ? 0 : (size_t)chunk * nn), S64(fp64 ? 0 : (size_t)chunk * nn);
    DevBuf<double> d_tc, d_sc, d_tr, d_sr, d_tsurf, d_ssurf, d_unit;
    DevBuf<int> d_iota, d_ls, d_lt;
    d_tsurf.upload(t_surf ? *t_surf : surf.inner, s);
    d_ssurf.upload(s_surf ? *s_surf : surf.inner, s);
    d_unit.upload(surf.unit, s);
    std::vector<int> iota(chunk);
    for (int i = 0; i < chunk; i++) iota[i] = i;
    d_iota.upload(iota, s);
    for (int c0 = 0; c0 < nc; c0 += chunk) {
        const int m = std::min(chunk, nc - c0);
        std::vector<double> tc(3 * m), scn(3 * m), tr(m), sr(m);
        std::vector<int> lt(m), ls(m);
        for (int i = 0; i < m; i++) {
            const int t = rep_t[c0 + i], sc = rep_s[c0 + i];
            tc[3 * i] = ot.cx[t]; tc[3 * i + 1] = ot.cy[t]; tc[3 * i + 2] = ot.cz[t];
            scn[3 * i] = st.cx[sc]; scn[3 * i + 1] = st.cy[sc]; scn[3 * i + 2] = st.cz[sc];
            lt[i] = ot.level[t];
            ls[i] = st.level[sc];
            tr[i] = ot.level_r[lt[i]];
            sr[i] = st.level_r[ls[i]];
        }
        d_tc.upload(tc, s); d_sc.upload(scn, s); d_tr.upload(tr, s); d_sr.upload(sr, s);
        d_lt.upload(lt, s); d_ls.upload(ls, s);
        hipLaunchKernelGGL(k_m2l_assemble, dim3(std::max(1, std::min(64, div_up((size_t)n_e * n_e, 256))), m),
                           dim3(256), 0, s, kid, kc, n_e, R, d_tc.p, d_sc.p, d_tr.p, d_sr.p, d_tsurf.p, d_ssurf.p,
                           d_unit.p,
                           K.p);
        TBEM_LAUNCH_CHECK();
        const dim3 gs(std::max(1, std::min(256, div_up(per, 256))), m);
        if (fp64) {
            hipLaunchKernelGGL(k_m2l_swizzle<double>, gs, dim3(256), 0, s, n, mt, kt, 0, K.p,
                               M.S64.p + (size_t)c0 * per);
            TBEM_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_m2l_swizzle<double>, gs, dim3(256), 0, s, n, mt, kt, 1, K.p,
                               M.ST64.p + (size_t)c0 * per);
            TBEM_LAUNCH_CHECK();
            TBEM_HIP(hipStreamSynchronize(s));  // the host vectors above are reused next chunk
            continue;
        }
        const dim3 gg(div_up(n, 64), div_up(n, 64), m);
        hipLaunchKernelGGL(k_gemm64, gg, dim3(256), 0, s, n, K.p, d_iota.p, Ru.p, d_ls.p, T1.p);
        TBEM_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_gemm64, gg, dim3(256), 0, s, n, Ld.p, d_lt.p, T1.p, d_iota.p, S64.p);
        TBEM_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_m2l_swizzle<float>, gs, dim3(256), 0, s, n, mt, kt, 0, S64.p, M.S.p + (size_t)c0 * per);
        TBEM_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_m2l_swizzle<float>, gs, dim3(256), 0, s, n, mt, kt, 1, S64.p, M.ST.p + (size_t)c0 * per);
        TBEM_LAUNCH_CHECK();
        TBEM_HIP(hipStreamSynchronize(s));  // the host vectors above are reused next chunk
    }

    // slices of the pairs (runs of whole target items) whose per-pair outputs fit
    // y_budget bytes; one slice when all fit
    {
        const size_t ybytes = (size_t)M.ystride * (fp64 ? sizeof(double) : sizeof(float));
        const int n_items = (int)m2l.items.size();
        size_t max_item = 0;
        for (int k = 0; k < n_items; k++) max_item = std::max<size_t>(max_item, m2l.ptr[k + 1] - m2l.ptr[k]);
        const size_t cap = std::max<size_t>(max_item, y_budget / ybytes);
        M.slice_items.assign(1, 0);
        M.slice_pairs.assign(1, 0);
        size_t acc = 0;
        for (int k = 0; k < n_items; k++) {
            const size_t c = m2l.ptr[k + 1] - m2l.ptr[k];
            if (acc + c > cap) {
                M.slice_items.push_back(k);
                M.slice_pairs.push_back(m2l.ptr[k]);
                acc = 0;
            }
            acc += c;
        }
        M.slice_items.push_back(n_items);
        M.slice_pairs.push_back(n_pairs);
        M.n_slices = (int)M.slice_items.size() - 1;
        M.y_pairs = 0;
        for (int q = 0; q < M.n_slices; q++) M.y_pairs = std::max(M.y_pairs, M.slice_pairs[q + 1] - M.slice_pairs[q]);
    }
    // split[c * (S+1) + q]: where slice q's pairs start in class c's run of the sorted
    // pairs (each run ascends in pair order)
    const int S = M.n_slices;
    std::vector<int> split((size_t)nc * (S + 1));
    if (S == 1) {
        for (int c = 0; c < nc; c++) {
            split[2 * c] = hstart[c];
            split[2 * c + 1] = hstart[c + 1];
        }
    } else {
        DevBuf<int> d_sp, d_split((size_t)nc * (S + 1));
        d_sp.upload(M.slice_pairs, s);
        hipLaunchKernelGGL(k_class_splits, dim3(div_up(nc, 256)), dim3(256), 0, s, nc, S, cstart.p, colpair.p, d_sp.p,
                           d_split.p);
        TBEM_LAUNCH_CHECK();
        d_split.download(split.data(), split.size(), s);
        TBEM_HIP(hipStreamSynchronize(s));
    }

    // work units, slice by slice: classes binned to XCDs by pair count, units
    // interleaved so that workgroup u lands on XCD u % 8 with the units of its bin; a
    // unit's pairs are colpair[off .. off + cnt) (pairs of one class and one slice, in
    // pair order, in the sorted array)
    const int ucols = fp64 ? kUnitCols64 : kUnitCols;
    std::vector<int> ucls, uoff, ucnt;
    M.slice_units.assign(1, 0);
    std::vector<int> order(nc);
    for (int q = 0; q < S; q++) {
        auto lo = [&](int c) { return split[(size_t)c * (S + 1) + q]; };
        auto csize = [&](int c) { return split[(size_t)c * (S + 1) + q + 1] - lo(c); };
        for (int c = 0; c < nc; c++) order[c] = c;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return csize(a) > csize(b); });
        std::vector<size_t> load(kXcds, 0);
        std::vector<std::vector<std::array<int, 3>>> bins(kXcds);  // (class, off, cnt)
        for (int c: order) {
            const int sz = csize(c);
            if (sz == 0) break;  // sorted by size: the rest are empty too
            const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            for (int j = 0; j < sz; j += ucols) bins[x].push_back({c, lo(c) + j, std::min(ucols, sz - j)});
            load[x] += (sz + ucols - 1) / ucols;
        }
        size_t rounds = 0;
        for (auto& b: bins) rounds = std::max(rounds, b.size());
        const size_t u0 = ucls.size();
        ucls.resize(u0 + rounds * kXcds, 0);
        uoff.resize(u0 + rounds * kXcds, 0);
        ucnt.resize(u0 + rounds * kXcds, 0);
        for (size_t r = 0; r < rounds; r++)
            for (int x = 0; x < kXcds; x++)
                if (r < bins[x].size()) {
                    ucls[u0 + r * kXcds + x] = bins[x][r][0];
                    uoff[u0 + r * kXcds + x] = bins[x][r][1];
                    ucnt[u0 + r * kXcds + x] = bins[x][r][2];
                }
        M.slice_units.push_back((int)ucls.size());
    }
    M.n_units = (int)ucls.size();
    M.unit_class.upload(ucls, s);
    M.unit_off.upload(uoff, s);
    M.unit_cnt.upload(ucnt, s);
    M.colpair = std::move(colpair);

    // transpose support: target of every pair; per source cell, its pairs (a stable
    // sort of the pairs by source cell keeps each source's pairs contiguous, in pair
    // order, the sources ascending)
    M.pair_tgt = std::move(tgt);
    {
        DevBuf<unsigned int> skey(n_pairs), sktmp(n_pairs);
        DevBuf<int> spairs(n_pairs);
        hipLaunchKernelGGL(k_source_keys, dim3(gp), dim3(256), 0, s, n_pairs, dm2l.cell.p, skey.p, spairs.p);
        TBEM_LAUNCH_CHECK();
        int bits = 1;
        while (bits < 31 && (1ll << bits) < (long long)st.n_cells) bits++;
        radix_sort_pairs_u32(skey.p, spairs.p, sktmp.p, itmp.p, n_pairs, bits, s);
        hipLaunchKernelGGL(k_run_starts<unsigned int>, dim3(gp), dim3(256), 0, s, n_pairs, skey.p, flag.p);
        TBEM_LAUNCH_CHECK();
        const int na = device_exclusive_scan(flag.p, n_pairs, pos.p, s);
        M.adj_ptr.alloc(na + 1);
        M.adj_items.alloc(na);
        hipLaunchKernelGGL(k_run_compact, dim3(gp), dim3(256), 0, s, n_pairs, flag.p, pos.p, na, M.adj_ptr.p);
        TBEM_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_gather_int, dim3(div_up(std::max(na, 1), 256)), dim3(256), 0, s, na, M.adj_ptr.p,
                           skey.p, M.adj_items.p);
        TBEM_LAUNCH_CHECK();
        std::vector<int> adj_items(na);
        M.adj_items.download(adj_items.data(), na, s);
        TBEM_HIP(hipStreamSynchronize(s));
        M.n_adj = na;
        M.adj_tiles = make_cell_tiles(adj_items, st, s);
        M.adj_pairs = std::move(spairs);
    }
    TBEM_HIP(hipStreamSynchronize(s));
    M.enabled = true;
    M.plan_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// LDS of k_m2l_pair_sum64: npar partial rows of ystride doubles when npar > 1
static size_t pair_sum64_shm(int ystride) {
    const int q = ystride / 2, npar = std::max(1, 256 / q);
    return npar > 1 ? (size_t)npar * ystride * sizeof(double) : 0;
}

// The slice's units: grid and offset unit arrays
struct SliceUnits {
    int n;
    const int *ucls, *uoff, *ucnt;
};
static SliceUnits slice_units(const M2lMfma& M, int q) {
    const int u0 = M.slice_units[q];
    return {M.slice_units[q + 1] - u0, M.unit_class.p + u0, M.unit_off.p + u0, M.unit_cnt.p + u0};
}

// Y (slice q) = M_c * chk_up[src(p)] (transpose: M_c^T * chk_up[tgt(p)]), FP32 MFMA
static void mfma32_slice(const M2lMfma& M, int q, const int* pair_src, const float* chk_up, float* Y,
                         bool transpose, hipStream_t s) {
    const SliceUnits U = slice_units(M, q);
    if (U.n == 0) return;
    const size_t shm = staged_bytes(M.blk, false);
    // eight waves when the staged columns allow at most two workgroups per CU (160 KiB of LDS)
    const bool wide = 3 * shm > kLdsPerCu;
    const float* S = transpose ? M.ST.p : M.S.p;
    const int* src = transpose ? M.pair_tgt.p : pair_src;
    const int yb = M.slice_pairs[q];
    if (wide)
        hipLaunchKernelGGL((k_m2l_mfma<8, kPreWide>), dim3(U.n), dim3(512), shm, s, M.blk, M.mt, M.kt, M.ystride, S,
                           U.ucls, U.uoff, U.ucnt, M.colpair.p, src, yb, chk_up, Y);
    else
        hipLaunchKernelGGL((k_m2l_mfma<4, kPreNarrow>), dim3(U.n), dim3(256), shm, s, M.blk, M.mt, M.kt, M.ystride,
                           S, U.ucls, U.uoff, U.ucnt, M.colpair.p, src, yb, chk_up, Y);
    TBEM_LAUNCH_CHECK();
}

// Y (slice q) = K_c * x[src(p)] (transpose: K_c^T * x[tgt(p)]), FP64 MFMA
static void mfma64_slice(const M2lMfma& M, int q, const int* pair_src, const double* x, double* Y, bool transpose,
                         hipStream_t s) {
    const SliceUnits U = slice_units(M, q);
    if (U.n == 0) return;
    const size_t shm = staged_bytes(M.blk, true);
    const double* S = transpose ? M.ST64.p : M.S64.p;
    const int* src = transpose ? M.pair_tgt.p : pair_src;
    if (3 * shm > kLdsPerCu)
        hipLaunchKernelGGL((k_m2l_mfma64<kWaves64, kPre64>), dim3(U.n), dim3(kWaves64 * 64), shm, s, M.blk, M.mt,
                           M.kt, M.ystride, S, U.ucls, U.uoff, U.ucnt, M.colpair.p, src, M.slice_pairs[q], x, Y);
    else
        hipLaunchKernelGGL((k_m2l_mfma64<kWaves64Narrow, kPre64Narrow>), dim3(U.n), dim3(kWaves64Narrow * 64), shm,
                           s, M.blk, M.mt, M.kt, M.ystride, S, U.ucls, U.uoff, U.ucnt, M.colpair.p, src,
                           M.slice_pairs[q], x, Y);
    TBEM_LAUNCH_CHECK();
}

// the forward sums of slice q: its items' pairs, out[t] = sum (FP32 Y into FP64 z, or FP64 Y)
template <class T>
static void forward_sum_slice(const M2lMfma& M, int q, const int* items, const int* ptr, const T* Y, double* out,
                              hipStream_t s) {
    const int k0 = M.slice_items[q], nk = M.slice_items[q + 1] - k0;
    if (nk == 0) return;
    if constexpr (std::is_same<T, float>::value) {
        const int npar = 256 / (M.ystride / 4);
        hipLaunchKernelGGL(k_m2l_pair_sum, dim3(nk), dim3(256), (size_t)npar * M.ystride * sizeof(double), s,
                           items + k0, ptr + k0, nullptr, Y, M.ystride, M.blk, M.slice_pairs[q], 0, 0, 0, out);
    } else {
        hipLaunchKernelGGL(k_m2l_pair_sum64, dim3(nk), dim3(256), pair_sum64_shm(M.ystride), s, items + k0,
                           ptr + k0, nullptr, Y, M.ystride, M.blk, M.slice_pairs[q], 0, 0, 0, out);
    }
    TBEM_LAUNCH_CHECK();
}

// the transpose's sums of slice q: per source cell, its pairs within the slice;
// out[s] = (q == 0 and !add_first ? 0 : out[s]) + sum
template <class T>
static void adjoint_sum_slice(const M2lMfma& M, int q, const T* Y, double* out, bool add_first, hipStream_t s) {
    if (M.n_adj == 0) return;
    const int add = (q > 0 || add_first) ? 1 : 0;
    const int pb = M.slice_pairs[q], pe = M.slice_pairs[q + 1];
    if constexpr (std::is_same<T, float>::value) {
        const int npar = 256 / (M.ystride / 4);
        hipLaunchKernelGGL(k_m2l_pair_sum, dim3(M.n_adj), dim3(256), (size_t)npar * M.ystride * sizeof(double), s,
                           M.adj_items.p, M.adj_ptr.p, M.adj_pairs.p, Y, M.ystride, M.blk, pb, pb, pe, add, out);
    } else {
        hipLaunchKernelGGL(k_m2l_pair_sum64, dim3(M.n_adj), dim3(256), pair_sum64_shm(M.ystride), s,
                           M.adj_items.p, M.adj_ptr.p, M.adj_pairs.p, Y, M.ystride, M.blk, pb, pb, pe, add, out);
    }
    TBEM_LAUNCH_CHECK();
}

void m2l_forward32(const M2lMfma& M, const int* items, const int* ptr, const int* pair_src, const CellTiles& targets,
                   const float* chk_up, float* Y, double* z, double* locals, hipStream_t s, hipEvent_t kernel_end) {
    if (!M.enabled || M.fp64) return;
    for (int q = 0; q < M.n_slices; q++) {
        mfma32_slice(M, q, pair_src, chk_up, Y, false, s);
        if (kernel_end && q + 1 == M.n_slices) TBEM_HIP(hipEventRecord(kernel_end, s));
        forward_sum_slice<float>(M, q, items, ptr, Y, z, s);
    }
    launch_level_gemm(targets, M.blk, M.Vd.p, false, z, nullptr, locals, nullptr, true, s);
}

void m2l_forward64(const M2lMfma& M, const int* items, const int* ptr, const int* pair_src, const double* mult,
                   double* Y, double* fld, hipStream_t s, hipEvent_t kernel_end) {
    if (!M.enabled || !M.fp64) return;
    for (int q = 0; q < M.n_slices; q++) {
        mfma64_slice(M, q, pair_src, mult, Y, false, s);
        if (kernel_end && q + 1 == M.n_slices) TBEM_HIP(hipEventRecord(kernel_end, s));
        forward_sum_slice<double>(M, q, items, ptr, Y, fld, s);
    }
}

void m2l_adjoint32(const M2lMfma& M, const CellTiles& targets, const double* lbar, float* w, float* Y, double* z,
                   double* e2, hipStream_t s) {
    if (!M.enabled || M.fp64) return;
    launch_level_gemm(targets, M.blk, M.Vd.p, true, lbar, nullptr, nullptr, w, false, s);
    for (int q = 0; q < M.n_slices; q++) {
        mfma32_slice(M, q, nullptr, w, Y, true, s);
        adjoint_sum_slice<float>(M, q, Y, z, false, s);
    }
    if (M.n_adj > 0) launch_level_gemm(M.adj_tiles, M.blk, M.Bu.p, true, z, nullptr, e2, nullptr, false, s);
}

void m2l_adjoint64(const M2lMfma& M, const double* cbar, double* Y, double* mbar, hipStream_t s) {
    if (!M.enabled || !M.fp64) return;
    for (int q = 0; q < M.n_slices; q++) {
        mfma64_slice(M, q, nullptr, cbar, Y, true, s);
        adjoint_sum_slice<double>(M, q, Y, mbar, true, s);
    }
}

}  // namespace tbem
