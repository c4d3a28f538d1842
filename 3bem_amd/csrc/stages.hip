// FMM stage kernels for gfx950.
//
// Every stage is "accumulate the field of a stream of sources at a set of targets":
//   * surface targets (P2M, M2M, M2L, P2L, L2L and their transposes): the n_e
//     points of a cell's check/equivalent surface; one kSurfBlock-thread
//     workgroup per target cell; the FP64 C2E products follow as batched per-level
//     GEMMs (k_level_gemm);
//   * point targets (P2P + M2P + L2P fused, or P2P^T + P2L^T + P2M^T): the
//     points of one leaf; one kLeafBlock-thread workgroup per leaf.
// Sources are staged through LDS in NB-entry SoA tiles (NB = threads per
// workgroup). Point sources come as contiguous ranges of the sorted SoA arrays
// (neighbouring leaves of a list are merged at plan time; Morton order makes most
// of them adjacent). Surface sources are generated from the cell center and the
// scaled unit sphere. The next tile is loaded into registers while the current
// one is computed.
// Lanes split into S = floor(NB/ntp) source slices that share the tile (ntp =
// target slots; the last NB - S*ntp lanes sit out), summed through LDS at the
// end by the whole workgroup (slice_sum: the FP64 elastic leaf stage went 4.41 ->
// 4.08 ms when the slice-0 lanes alone added up to 42 partials each); in the leaf
// stage each lane carries two targets, so one LDS read feeds two pairs.
// Workgroup width is the main performance lever of the leaf stage: wider
// workgroups give each leaf's targets more source slices, hence more waves in
// flight per leaf and fewer tile barriers per pair. Measured on the 1M-point
// elastic workload (order 6, leaf 64), FP64 leaf stage: 6.4 ms at 128 threads,
// 4.7 ms at 256, 4.4 ms at 512 (8.6 ms at 1024) with tiles packed across range
// boundaries. The surface stages have few sources per target and prefer 128.
//
// Precision. Surface-target stages are FP64 only: each check potential they form
// feeds a check-to-equivalent pseudo-inverse (condition number up to ~1e10 at the
// 1e-10 SVD threshold), which turns FP32 rounding into O(1) errors at order >= 50.
// The leaf stage runs its point sources (a) in T: FP64, or FP32 for the mixed
// path. Tiles always hold FP64 coordinates and every displacement is formed in
// FP64 and then rounded, so near-coincident pairs (1/r up to 1e6 near the r^2 <
// 1e-12 cutoff) keep full relative accuracy; FP32 partial sums are folded per
// tile into Neumaier-compensated accumulators. Surface sources (b), (c) evaluate
// equivalent densities, whose entries grow like (1.73/0.52)^degree because the
// equivalent sphere lies inside the cell: they stay FP64 (a few percent of the
// pairs of a leaf).
#include <algorithm>
#include <type_traits>

#include "stages.h"

namespace tbem {

// SoA source tile: FP64 coordinates, normals and densities in VT.
template <class VT, int R, bool NRM, int NB>
struct Tile {
    double x[NB], y[NB], z[NB];
    VT nx[NRM ? NB : 1], ny[NRM ? NB : 1], nz[NRM ? NB : 1];
    VT d[R][NB];
};

// LDS of a stage kernel: the point tile (values in PT), the surface tile (FP64),
// and the cross-slice reduction (RED doubles per thread) share bytes, since they
// are live one at a time; the range cache serves the packed point stream.
template <class PT, int R, bool NRM, int NB, int RED>
struct StageSmem {
    union {
        Tile<PT, R, NRM, NB> pt;
        Tile<double, R, NRM, NB> st;
        double red[RED * NB];
    };
    int rpre[NB + 1], rbeg[NB];
    int wflag[NB / 64];  // stream_ranges_packed: per wave, a close point in the tile
    double part[NB];  // group sums of slice_sum
};

// Sums the S slice partials red[v*NB + k*ntp + tg] (k < S) of nv value rows and
// ntp target slots with the whole block, and hands each total to emit(v, tg, s).
// When G = min(S, NB / (nv*ntp)) > 1 groups fit, group g first adds slices
// g, g+G, ... into part, then one thread per (v, tg) adds the G group sums:
// chains of about S/G + G additions instead of S. The caller has written red and
// synchronized; every thread of the block must call this.
template <int NB, class F>
__device__ __forceinline__ void slice_sum(const double* red, double* part, int nv, int ntp, int S, F&& emit) {
    const int tid = threadIdx.x;
    const int nout = nv * ntp;
    const int G = min(S, NB / nout);
    if (G <= 1) {
        for (int o = tid; o < nout; o += NB) {
            const int v = o / ntp, tg = o - v * ntp;
            double s = 0;
            for (int k = 0; k < S; k++) s += red[v * NB + k * ntp + tg];
            emit(v, tg, s);
        }
        return;
    }
    if (tid < nout * G) {
        const int g = tid / nout, o = tid - g * nout;
        const int v = o / ntp, tg = o - v * ntp;
        double s = 0;
        for (int k = g; k < S; k += G) s += red[v * NB + k * ntp + tg];
        part[tid] = s;
    }
    __syncthreads();
    if (tid < nout) {
        double s = 0;
        for (int g = 0; g < G; g++) s += part[g * nout + tid];
        const int v = tid / ntp;
        emit(v, tid - v * ntp, s);
    }
}

// one staged source held in registers (FP64, absolute) until it is written to LDS
template <int R, bool NRM>
struct SrcReg {
    double x, y, z, nx, ny, nz, d[R];
};

// Wait for every outstanding memory access of this wave. Issued after the target
// loads so the waitcnt pass knows they are complete; otherwise it inserts
// conservative vmcnt(0) waits inside the pair loop, which also wait on the next
// tile's prefetch and serialise load latency with compute.
__device__ __forceinline__ void drain_loads() { __builtin_amdgcn_s_waitcnt(0); }

// moved surface point r*s + c, rounded as TranslationSurface::move (cpp/fmm.h:19-31)
__device__ inline void moved(double r, const double* s, double cx, double cy, double cz, double* p) {
#pragma clang fp contract(off)
    p[0] = r * s[0] + cx;
    p[1] = r * s[1] + cy;
    p[2] = r * s[2] + cz;
}

template <class T, int R>
struct Acc;

template <int R>
struct Acc<double, R> {
    double t[R];
    __device__ void init() {
        for (int q = 0; q < R; q++) t[q] = 0;
    }
    __device__ double* tile() { return t; }
    __device__ void flush() {}
    __device__ double value(int q) const { return t[q]; }
};

// Per-tile FP32 partials folded into Neumaier-compensated FP32 sums.
template <int R>
struct Acc<float, R> {
    float t[R], s[R], c[R];
    __device__ void init() {
        for (int q = 0; q < R; q++) t[q] = s[q] = c[q] = 0.f;
    }
    __device__ float* tile() { return t; }
    __device__ void flush() {
        for (int q = 0; q < R; q++) {
            const float tt = s[q] + t[q];
            if (fabsf(s[q]) >= fabsf(t[q])) c[q] += (s[q] - tt) + t[q];
            else c[q] += (t[q] - tt) + s[q];
            s[q] = tt;
            t[q] = 0.f;
        }
    }
    __device__ double value(int q) const { return (double)s[q] + (double)c[q]; }
};

// ----------------------------------------------------------------- source loaders
template <int R, bool NRM>
__device__ inline void load_point(SrcReg<R, NRM>& r, int k, const PtsView& sp, const double* __restrict__ dens,
                                  int n_src) {
    r.x = sp.x[k]; r.y = sp.y[k]; r.z = sp.z[k];
    if (NRM) { r.nx = sp.nx[k]; r.ny = sp.ny[k]; r.nz = sp.nz[k]; }
    for (int q = 0; q < R; q++) r.d[q] = dens[(size_t)q * n_src + k];
}

template <int R, bool NRM>
__device__ inline void load_surf(SrcReg<R, NRM>& r, int c, int j, const TreeView& st, const double* __restrict__ surf,
                                 const double* __restrict__ unit, const double* __restrict__ coeff, int n_e, int blk) {
    double p[3];
    moved(st.level_r[st.level[c]], surf + 3 * j, st.cx[c], st.cy[c], st.cz[c], p);
    r.x = p[0]; r.y = p[1]; r.z = p[2];
    if (NRM) { r.nx = unit[3 * j]; r.ny = unit[3 * j + 1]; r.nz = unit[3 * j + 2]; }
    for (int q = 0; q < R; q++) r.d[q] = coeff[(size_t)c * blk + q * n_e + j];
}

template <class VT, int R, bool NRM, int NB>
__device__ inline void store_tile(Tile<VT, R, NRM, NB>& tl, int slot, const SrcReg<R, NRM>& r) {
    tl.x[slot] = r.x;
    tl.y[slot] = r.y;
    tl.z[slot] = r.z;
    if (NRM) { tl.nx[slot] = VT(r.nx); tl.ny[slot] = VT(r.ny); tl.nz[slot] = VT(r.nz); }
    for (int q = 0; q < R; q++) tl.d[q][slot] = VT(r.d[q]);
}

// Streams the points of ranges e0..e1 (sorted points [rbeg[e], rbeg[e]+rcnt[e]))
// through the tile one range at a time: a tile never spans two ranges, so each
// lane's source index is rbeg[e] + off + tid. Used by the surface stages, whose
// point lists (P2M, P2L) are short.
template <class VT, int R, bool NRM, int NB, class F>
__device__ inline void stream_ranges(Tile<VT, R, NRM, NB>& tl, int e0, int e1,
                                     const int* __restrict__ rbeg, const int* __restrict__ rcnt,
                                     const PtsView& sp, const double* __restrict__ dens, int n_src,
                                     F&& compute) {
    const int tid = threadIdx.x;
    int e = e0, off = 0;
    if (e >= e1) return;
    SrcReg<R, NRM> r;
    int len = uni(rcnt[e]), beg = uni(rbeg[e]);
    int cnt = min(NB, len);
    if (tid < cnt) load_point(r, beg + tid, sp, dens, n_src);
    while (true) {
        __syncthreads();
        if (tid < cnt) store_tile(tl, tid, r);
        __syncthreads();
        const int cur = cnt;
        off += NB;
        if (off >= len) {
            e++;
            off = 0;
            if (e < e1) {
                len = uni(rcnt[e]);
                beg = uni(rbeg[e]);
            }
        }
        const bool more = e < e1;
        if (more) {
            cnt = min(NB, len - off);
            if (tid < cnt) load_point(r, beg + off + tid, sp, dens, n_src);
        }
        compute(cur);
        if (!uni(more)) break;
    }
}

// Streams the points of ranges e0..e1 through the tile in full NB-point tiles cut
// from the list's virtual concatenation (range e starts at offset rpre[e] of it),
// so a tile may span range boundaries and only the last one is partial. Up to NB
// ranges are cached in LDS; each lane finds its range by a short forward scan
// from the range holding the tile's first point. Ranges must be nonempty: then a
// cache that starts at the range holding point `base` covers at least
// [base, base + NB), which is what one tile needs.
//
// rclose (optional): per range, 1 when some of its points may lie within the
// kernels' cutoff radius of the targets. With tile_close given, *tile_close is set
// before each compute() to whether any point of the tile comes from such a range
// (1 for every tile when rclose is null), so that compute() can leave the cutoff
// test out of the other tiles: a ballot per wave into s_wflag (NB/64 ints of LDS)
// before the tile's second barrier, read back after it. The flag travels in bit 30
// of the cached range starts (point indices are below 2^30).
constexpr int kCloseBit = 1 << 30;

template <class VT, int R, bool NRM, int NB, class F>
__device__ inline void stream_ranges_packed(Tile<VT, R, NRM, NB>& tl, int* s_pre, int* s_beg, int e0, int e1,
                                            const int* __restrict__ rbeg, const int* __restrict__ rcnt,
                                            const int* __restrict__ rpre, const PtsView& sp,
                                            const double* __restrict__ dens, int n_src, F&& compute,
                                            const int* __restrict__ rclose = nullptr, int* tile_close = nullptr,
                                            int* s_wflag = nullptr) {
    const int tid = threadIdx.x;
    if (e0 >= e1) return;
    const int total = uni(rpre[e1 - 1] + rcnt[e1 - 1]);
    SrcReg<R, NRM> r;
    int mclose = 0;  // the lane's next point comes from a close range
    int lbase = e0;  // list index of cache entry 0
    auto cache = [&](int from) {
        __syncthreads();
        lbase = from;
        const int e = from + tid;
        if (e < e1) {
            s_pre[tid] = rpre[e];
            s_beg[tid] = rbeg[e] | ((!rclose || rclose[e]) ? kCloseBit : 0);
        }
        if (tid == 0) s_pre[NB] = (from + NB < e1) ? rpre[from + NB] : total;
        __syncthreads();
    };
    cache(e0);
    int ecur = 0;  // cache entry of the range holding the current tile's first point (uniform)
    auto fetch = [&](int base) {
        const int v = base + tid;
        mclose = 0;
        if (v < total) {
            const int nl = min(NB, e1 - lbase);
            int f = ecur;
            while (f + 1 < nl && s_pre[f + 1] <= v) f++;
            const int bf = s_beg[f];
            mclose = bf & kCloseBit;
            load_point(r, (bf & ~kCloseBit) + (v - s_pre[f]), sp, dens, n_src);
        }
    };
    int base = 0;
    fetch(0);
    while (true) {
        const int cnt = min(NB, total - base);
        __syncthreads();
        if (tid < cnt) store_tile(tl, tid, r);
        if (tile_close) {
            const unsigned long long bal = __ballot(tid < cnt && mclose);
            if ((tid & 63) == 0) s_wflag[tid >> 6] = bal != 0ull;
        }
        __syncthreads();
        if (tile_close) {
            int any = 0;
            for (int w = 0; w < NB / 64; w++) any |= s_wflag[w];
            *tile_close = any;
        }
        base += NB;
        const bool more = base < total;
        if (more) {
            const int nl = min(NB, e1 - lbase);
            while (uni(ecur + 1 < nl && s_pre[ecur + 1] <= base)) ecur++;
            const int cache_end = uni(s_pre[NB]);
            if (uni(lbase + NB < e1 && cache_end < base + NB)) {  // the next tile reaches past the cache
                // the previous tile was covered, so cache_end >= base; equality means
                // the range holding base is the first one past the cache
                cache(cache_end <= base ? lbase + NB : lbase + ecur);
                ecur = 0;
            }
            fetch(base);
        }
        compute(cnt);
        if (!uni(more)) break;
    }
}

// Streams the equivalent/check surfaces of cells cell[e0..e1) (n_e points each).
template <int R, bool NRM, int NB, class F>
__device__ inline void stream_surfaces(Tile<double, R, NRM, NB>& tl, int e0, int e1,
                                       const int* __restrict__ cell, const TreeView& st,
                                       const double* __restrict__ surf, const double* __restrict__ unit,
                                       const double* __restrict__ coeff, int n_e, int blk, F&& compute) {
    const int tid = threadIdx.x;
    const int total = (e1 - e0) * n_e;
    if (total <= 0) return;
    SrcReg<R, NRM> r;
    int base = 0;
    auto fetch = [&](int b) {
        const int v = b + tid;
        if (v < total) {
            const int k = v / n_e;
            load_surf(r, cell[e0 + k], v - k * n_e, st, surf, unit, coeff, n_e, blk);
        }
    };
    fetch(0);
    while (true) {
        const int cnt = min(NB, total - base);
        __syncthreads();
        if (tid < cnt) store_tile(tl, tid, r);
        __syncthreads();
        base += NB;
        const bool more = base < total;
        if (more) fetch(base);
        compute(cnt);
        if (!uni(more)) break;
    }
}

// ---------------------------------------------------------------------- pair loop
// One source j of the tile on one target (FP64 coordinates tx, ty, tz): the
// displacement is formed in FP64, then the kernel runs in T.
template <int KID, bool ADJ, bool FAST = false, bool CUT = true, class T, class VT, int R, bool NRM, int NB>
__device__ __forceinline__ void pair(const Tile<VT, R, NRM, NB>& tl, int j, double tx, double ty, double tz,
                                     const T* tn, T* acc, const KC<T>& kc) {
    const T sn[3] = {NRM ? T(tl.nx[j]) : T(0), NRM ? T(tl.ny[j]) : T(0), NRM ? T(tl.nz[j]) : T(0)};
    T v[R];
    for (int q = 0; q < R; q++) v[q] = T(tl.d[q][j]);
    if (!ADJ) {
        const T dx = T(tl.x[j] - tx), dy = T(tl.y[j] - ty), dz = T(tl.z[j] - tz);
        const T r2 = dx * dx + dy * dy + dz * dz;
        kop<KID, false, T, FAST, CUT>(dx, dy, dz, r2, tn, sn, v, acc, kc);
    } else {
        const T dx = T(tx - tl.x[j]), dy = T(ty - tl.y[j]), dz = T(tz - tl.z[j]);
        const T r2 = dx * dx + dy * dy + dz * dz;
        kop<KID, true, T, FAST, CUT>(dx, dy, dz, r2, sn, tn, v, acc, kc);
    }
}

// ------------------------------------------------------------------ surface targets
// CUT = false (surface sources only): the r2 < 1e-12 cutoff test is left out, when
// the plan has shown every pair of the stage to be farther apart (SurfStage.no_cutoff).
template <int KID, bool SRC_SURF, bool ADJ, bool DUAL, bool CUT = true>
__global__ __launch_bounds__(kSurfBlock) void k_surf_stage(SurfStage a, KC<double> kc, double scale, int add) {
    constexpr int NB = kSurfBlock;
    constexpr int R = kernel_dim(KID);
    constexpr bool NRM = kernel_uses_nsrc(KID) || kernel_uses_nobs(KID);
    __shared__ StageSmem<double, R, NRM, NB, (DUAL ? 2 : 1) * R> sm;
    extern __shared__ double field[];  // R * n_e
    const int item = blockIdx.x;
    const int cell = uni(a.items[item]);
    const int lvl = uni(a.tt.level[cell]);
    const double rc = a.tt.level_r[lvl];
    const double cx = a.tt.cx[cell], cy = a.tt.cy[cell], cz = a.tt.cz[cell];
    const int e0 = uni(a.ptr[item]), e1 = uni(a.ptr[item + 1]);
    const int tid = threadIdx.x;

    // DUAL (launched for n_e > NB/2): two targets per lane (slots tg and tg + ntp;
    // a lane without a second target evaluates the first twice and drops the copy),
    // so one LDS read of a source feeds two pairs. Otherwise one target per lane.
    // The sources are split over S slices, summed at the end.
    constexpr bool dual = DUAL;
    const int per_pass = dual ? 2 * NB : NB;
    for (int t0 = 0; t0 < a.n_e; t0 += per_pass) {
        const int nt = min(per_pass, a.n_e - t0);
        const int ntp = dual ? (nt + 1) / 2 : nt, S = NB / ntp;
        const int tg = tid % ntp, slice = tid / ntp;
        const bool active = slice < S;
        const bool two = dual && tg + ntp < nt;
        const int i0 = t0 + tg, i1 = two ? i0 + ntp : i0;
        double p0[3], p1[3] = {0, 0, 0};
        moved(rc, a.t_surf + 3 * i0, cx, cy, cz, p0);
        if (dual) moved(rc, a.t_surf + 3 * i1, cx, cy, cz, p1);
        const double tn0[3] = {a.unit[3 * i0], a.unit[3 * i0 + 1], a.unit[3 * i0 + 2]};
        double tn1[3] = {0, 0, 0};
        if (dual) for (int m = 0; m < 3; m++) tn1[m] = a.unit[3 * i1 + m];
        drain_loads();
        double acc0[R], acc1[R];
        for (int q = 0; q < R; q++) acc0[q] = acc1[q] = 0.0;
        auto compute = [&](int cnt) {
            if (!active) return;
            if constexpr (dual) {
                for (int j = slice; j < cnt; j += S) {
                    pair<KID, ADJ, false, CUT>(sm.st, j, p0[0], p0[1], p0[2], tn0, acc0, kc);
                    pair<KID, ADJ, false, CUT>(sm.st, j, p1[0], p1[1], p1[2], tn1, acc1, kc);
                }
            } else {
                for (int j = slice; j < cnt; j += S)
                    pair<KID, ADJ, false, CUT>(sm.st, j, p0[0], p0[1], p0[2], tn0, acc0, kc);
            }
        };
        if (SRC_SURF)
            stream_surfaces(sm.st, e0, e1, a.cell, a.st, a.s_surf, a.unit, a.s_coeff, a.n_e, a.blk, compute);
        else
            stream_ranges(sm.st, e0, e1, a.rbeg, a.rcnt, a.sp, a.dens, a.n_src_pts, compute);
        __syncthreads();
        for (int q = 0; q < R; q++) {
            sm.red[q * NB + tid] = active ? acc0[q] : 0.0;
            if constexpr (dual) sm.red[(R + q) * NB + tid] = active ? acc1[q] : 0.0;
        }
        __syncthreads();
        slice_sum<NB>(sm.red, sm.part, (dual ? 2 : 1) * R, ntp, S, [&](int v, int g, double s) {
            const int h = v / R, q = v - h * R;
            if (h == 1 && ntp + g >= nt) return;
            field[q * a.n_e + t0 + h * ntp + g] = s * scale;
        });
        __syncthreads();
    }
    double* o = a.out + (size_t)cell * a.blk;
    for (int row = tid; row < a.blk; row += NB) o[row] = add ? o[row] + field[row] : field[row];
}

// -------------------------------------------------------------------- point targets
// Two targets per lane: slot tg holds targets t0+tg and t0+nslot+tg. Point
// sources accumulate in T (acA, acB); surface sources in FP64 (sfA, sfB, which
// are acA, acB themselves when T is double). Register budget: at 512 threads per
// workgroup, <= 128 VGPRs keeps two workgroups per CU; one more register halves
// the resident waves (the elastic-U FP32 instance at 139 VGPRs ran 1.5x slower),
// so leaf_waves() makes the compiler hold that budget (leaf_block() for the rest).
// CLOSE: tiles without close points (LeafStage.rp_close) skip the cutoff test; the
// instances without it have one pair loop and fewer registers, which pays where
// leaves are small and per-leaf work dominates.
template <int KID, class T, bool ADJ, bool FAST, bool CLOSE = false>
__global__ __launch_bounds__((leaf_block<KID, T>()), (leaf_waves<KID, T>())) void k_leaf_stage(LeafStage a, KC<T> kc, KC<double> kc64,
                                                                  double scale) {
    constexpr int NB = leaf_block<KID, T>();
    constexpr int R = kernel_dim(KID);
    constexpr bool NRM = kernel_uses_nsrc(KID) || kernel_uses_nobs(KID);
    constexpr bool SPLIT = !std::is_same<T, double>::value;
    __shared__ StageSmem<T, R, NRM, NB, 2 * R> sm;
    const int pos = a.order ? uni(a.order[blockIdx.x]) : (int)blockIdx.x;
    if (pos < 0) return;  // padding of the layout (the whole workgroup)
    const int leaf = uni(a.leaves[pos]);
    const int b = uni(a.tt.begin[leaf]), m = uni(a.tt.count[leaf]);
    const int pe0 = a.rp_ptr ? uni(a.rp_ptr[pos]) : 0, pe1 = a.rp_ptr ? uni(a.rp_ptr[pos + 1]) : 0;
    const int se0 = a.sl_ptr ? uni(a.sl_ptr[pos]) : 0, se1 = a.sl_ptr ? uni(a.sl_ptr[pos + 1]) : 0;
    const int tid = threadIdx.x;

    for (int t0 = 0; t0 < m; t0 += 2 * NB) {
        const int nt = min(2 * NB, m - t0);
        const int nslot = (nt + 1) >> 1;
        const int ntp = nslot, S = NB / ntp;  // lanes >= ntp*S sit out
        const int tg = tid % ntp, slice = tid / ntp;
        const bool va = slice < S, vb = va && nslot + tg < nt;
        const int ia = b + t0 + tg;
        const int ib = b + t0 + (vb ? nslot + tg : tg);
        const double ax = a.tp.x[ia], ay = a.tp.y[ia], az = a.tp.z[ia];
        const double bx = a.tp.x[ib], by = a.tp.y[ib], bz = a.tp.z[ib];
        const double an64[3] = {NRM ? a.tp.nx[ia] : 0.0, NRM ? a.tp.ny[ia] : 0.0, NRM ? a.tp.nz[ia] : 0.0};
        const double bn64[3] = {NRM ? a.tp.nx[ib] : 0.0, NRM ? a.tp.ny[ib] : 0.0, NRM ? a.tp.nz[ib] : 0.0};
        const T an[3] = {T(an64[0]), T(an64[1]), T(an64[2])};
        const T bn[3] = {T(bn64[0]), T(bn64[1]), T(bn64[2])};
        drain_loads();
        Acc<T, R> acA, acB;
        acA.init();
        acB.init();
        double sfA[SPLIT ? R : 1], sfB[SPLIT ? R : 1];
        for (int q = 0; q < (SPLIT ? R : 1); q++) sfA[q] = sfB[q] = 0.0;
        int tclose = 1;  // set per tile by stream_ranges_packed
        auto points = [&](int cnt) {
            if (va) {
                if (uni(tclose)) {
                    for (int j = slice; j < cnt; j += S) {
                        pair<KID, ADJ, FAST, true>(sm.pt, j, ax, ay, az, an, acA.tile(), kc);
                        pair<KID, ADJ, FAST, true>(sm.pt, j, bx, by, bz, bn, acB.tile(), kc);
                    }
                } else {  // no point of the tile within the cutoff of the leaf: same result without the test
                    for (int j = slice; j < cnt; j += S) {
                        pair<KID, ADJ, FAST, false>(sm.pt, j, ax, ay, az, an, acA.tile(), kc);
                        pair<KID, ADJ, FAST, false>(sm.pt, j, bx, by, bz, bn, acB.tile(), kc);
                    }
                }
            }
            acA.flush();
            acB.flush();
        };
        auto surfaces = [&](int cnt) {
            if (va) {
                double *sa, *sb;
                if constexpr (SPLIT) {
                    sa = sfA;
                    sb = sfB;
                } else {
                    sa = acA.tile();
                    sb = acB.tile();
                }
                for (int j = slice; j < cnt; j += S) {
                    pair<KID, ADJ>(sm.st, j, ax, ay, az, an64, sa, kc64);
                    pair<KID, ADJ>(sm.st, j, bx, by, bz, bn64, sb, kc64);
                }
            }
        };
        stream_ranges_packed(sm.pt, sm.rpre, sm.rbeg, pe0, pe1, a.rp_beg, a.rp_cnt, a.rp_pre, a.ps, a.pdens,
                             a.n_ps, points, CLOSE ? a.rp_close : nullptr, CLOSE ? &tclose : nullptr,
                             sm.wflag);                                                              // (a)
        stream_surfaces(sm.st, se0, se1, a.sl_cell, a.slt, a.sl_surf, a.unit, a.sl_coeff, a.n_e, a.blk,
                        surfaces);                                                                // (b)
        if (a.self_coeff)
            stream_surfaces(sm.st, 0, 1, &leaf, a.tt, a.self_surf, a.unit, a.self_coeff, a.n_e, a.blk,
                            surfaces);                                                            // (c)
        __syncthreads();
        for (int q = 0; q < R; q++) {
            const double ea = SPLIT ? sfA[q] : 0.0, eb = SPLIT ? sfB[q] : 0.0;
            sm.red[q * NB + tid] = va ? acA.value(q) + ea : 0.0;
            sm.red[(R + q) * NB + tid] = vb ? acB.value(q) + eb : 0.0;
        }
        __syncthreads();
        slice_sum<NB>(sm.red, sm.part, 2 * R, ntp, S, [&](int v, int g, double s) {
            const int h = v / R, q = v - h * R;
            if (h == 1 && nslot + g >= nt) return;
            const int i = b + t0 + h * nslot + g;
            const double w = a.t_weight ? a.t_weight[i] : 1.0;
            double* o = a.out_sorted + (size_t)q * a.n_tgt + i;
            *o = a.accumulate ? *o + w * (s * scale) : w * (s * scale);
        });
        __syncthreads();
    }
}

// ----------------------------------------------------- point targets, small leaves
// Leaves of at most kWaveLeafMax points (every leaf at the reference's leaf sizes;
// the plan sends larger ones to k_leaf_stage): one wave64 per work item, a whole
// leaf or, for a leaf of more than kWavePairsMax pairs, one chunk of its sources
// (WaveItem), laid out per XCD in runs of neighbouring leaves, the most work first
// (padding items have pos < 0); kWaveBlock / 64 items per workgroup, each wave streaming its own sources
// through its own 64-entry LDS tiles. There is no workgroup barrier and no
// workgroup-wide slice sum, and a CU holds 16 leaves in flight instead of the 512-
// thread kernel's 2, which is what hides a small leaf's chain of dependent loads
// (targets, lists, tiles, surfaces, output): on the 1M-point leaf-64 workload the
// 512-thread kernel spends more than half its time outside the pair loop. The
// arithmetic is the 512-thread kernel's: two targets per lane, S = 64 / ceil(m/2)
// source slices, pair() on FP64 coordinates, FP32 tiles folded into compensated
// sums per tile, FP64 surfaces; slices are summed in order at the end. Cutting the
// largest leaves into chunks (k_leaf_parts adds them, in chunk order) keeps one leaf
// from setting the launch's length: on the 1M-point leaf-64 workload a 3-point leaf
// has 674k point sources, and 20 leaves have 1-2.4M pairs each.

template <class PT, int R, bool NRM>
struct WaveSmem {
    union {
        Tile<PT, R, NRM, 64> pt;
        Tile<double, R, NRM, 64> st;
        double red[2 * R * 64];
    };
};

// Orders one wave's LDS writes before its later reads (and reads before later
// writes): the lanes of a wave run in lockstep, so only the compiler's ordering is
// at stake.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int KID, class T, bool ADJ, bool FAST>
__global__ __launch_bounds__(kWaveBlock) void k_leaf_wave(LeafStage a, KC<T> kc, KC<double> kc64, double scale) {
    constexpr int R = kernel_dim(KID);
    constexpr bool NRM = kernel_uses_nsrc(KID) || kernel_uses_nobs(KID);
    constexpr bool SPLIT = !std::is_same<T, double>::value;
    __shared__ WaveSmem<T, R, NRM> smw[kWaveBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int idx = blockIdx.x * (kWaveBlock / 64) + w;
    if (idx >= a.n_wave) return;  // the whole wave: nothing below synchronizes the block
    const WaveItem* it = a.wave_items + idx;
    const int pos = uni(it->pos);
    if (pos < 0) return;  // padding of the layout
    WaveSmem<T, R, NRM>& sm = smw[w];
    const int leaf = uni(a.leaves[pos]);
    const int b = uni(a.tt.begin[leaf]), m = uni(a.tt.count[leaf]);  // m <= kWaveLeafMax
    const int nslot = (m + 1) >> 1, S = 64 / nslot;
    const int tg = lane % nslot, slice = lane / nslot;
    const bool va = slice < S, vb = va && nslot + tg < m;
    const int ia = b + tg, ib = b + (vb ? nslot + tg : tg);
    const double ax = a.tp.x[ia], ay = a.tp.y[ia], az = a.tp.z[ia];
    const double bx = a.tp.x[ib], by = a.tp.y[ib], bz = a.tp.z[ib];
    const double an64[3] = {NRM ? a.tp.nx[ia] : 0.0, NRM ? a.tp.ny[ia] : 0.0, NRM ? a.tp.nz[ia] : 0.0};
    const double bn64[3] = {NRM ? a.tp.nx[ib] : 0.0, NRM ? a.tp.ny[ib] : 0.0, NRM ? a.tp.nz[ib] : 0.0};
    const T an[3] = {T(an64[0]), T(an64[1]), T(an64[2])};
    const T bn[3] = {T(bn64[0]), T(bn64[1]), T(bn64[2])};
    const int f0 = uni(it->f0), f1 = uni(it->f1), v0 = uni(it->v0), v1 = uni(it->v1);
    const int se0 = uni(it->c0), se1 = uni(it->c1), self = uni(it->self), part = uni(it->part);
    drain_loads();
    Acc<T, R> acA, acB;
    acA.init();
    acB.init();
    double sfA[SPLIT ? R : 1], sfB[SPLIT ? R : 1];
    for (int q = 0; q < (SPLIT ? R : 1); q++) sfA[q] = sfB[q] = 0.0;

    // (a) point sources: [v0, v1) of the leaf's packed ranges, 64 per tile, the next
    // tile in registers while the current one is computed; each lane walks the
    // ranges f0 .. f1-1 forward (a null rp_ptr, sl_ptr or self_coeff leaves out (a),
    // (b) or (c) as in k_leaf_stage; a chunk then sums what is left of it)
    if (a.rp_ptr && v0 < v1) {
        const int total = v1 - v0;
        SrcReg<R, NRM> r;
        int f = f0;
        int mclose = 0;  // the lane's fetched point comes from a range flagged close (rp_close)
        auto fetch = [&](int base) {
            const int v = v0 + base + lane;
            mclose = 0;
            if (v < v1) {
                while (f + 1 < f1 && a.rp_pre[f + 1] <= v) f++;
                load_point(r, a.rp_beg[f] + (v - a.rp_pre[f]), a.ps, a.pdens, a.n_ps);
                mclose = a.rp_close ? a.rp_close[f] : 1;
            }
        };
        fetch(0);
        for (int base = 0; base < total; base += 64) {
            const int cnt = min(64, total - base);
            wave_sync();
            if (lane < cnt) store_tile(sm.pt, lane, r);
            // a tile none of whose points is flagged close runs without the cutoff test
            // (the same result: every pair is farther apart than the cutoff radius)
            const bool tclose = __ballot(mclose) != 0ull;
            wave_sync();
            if (base + 64 < total) fetch(base + 64);
            if (va) {
                if (tclose) {
                    for (int j = slice; j < cnt; j += S) {
                        pair<KID, ADJ, FAST, true>(sm.pt, j, ax, ay, az, an, acA.tile(), kc);
                        pair<KID, ADJ, FAST, true>(sm.pt, j, bx, by, bz, bn, acB.tile(), kc);
                    }
                } else {
                    for (int j = slice; j < cnt; j += S) {
                        pair<KID, ADJ, FAST, false>(sm.pt, j, ax, ay, az, an, acA.tile(), kc);
                        pair<KID, ADJ, FAST, false>(sm.pt, j, bx, by, bz, bn, acB.tile(), kc);
                    }
                }
            }
            acA.flush();
            acB.flush();
        }
    }
    // (b), (c) surface sources, FP64: n_e points per cell of the list, then the own surface
    auto surfaces = [&](const int* cells, int c0, int c1, const TreeView& st, const double* surf,
                        const double* coeff) {
        const int total = (c1 - c0) * a.n_e;
        if (total <= 0) return;
        SrcReg<R, NRM> r;
        auto fetch = [&](int base) {
            const int v = base + lane;
            if (v < total) {
                const int k = v / a.n_e;
                load_surf(r, cells[c0 + k], v - k * a.n_e, st, surf, a.unit, coeff, a.n_e, a.blk);
            }
        };
        double *sa, *sb;
        if constexpr (SPLIT) {
            sa = sfA;
            sb = sfB;
        } else {
            sa = acA.tile();
            sb = acB.tile();
        }
        fetch(0);
        for (int base = 0; base < total; base += 64) {
            const int cnt = min(64, total - base);
            wave_sync();
            if (lane < cnt) store_tile(sm.st, lane, r);
            wave_sync();
            if (base + 64 < total) fetch(base + 64);
            if (va) {
                for (int j = slice; j < cnt; j += S) {
                    pair<KID, ADJ>(sm.st, j, ax, ay, az, an64, sa, kc64);
                    pair<KID, ADJ>(sm.st, j, bx, by, bz, bn64, sb, kc64);
                }
            }
        }
    };
    if (a.sl_ptr && se0 < se1) surfaces(a.sl_cell, se0, se1, a.slt, a.sl_surf, a.sl_coeff);
    if (self && a.self_coeff) surfaces(&leaf, 0, 1, a.tt, a.self_surf, a.self_coeff);

    // the S slices of each target, summed in slice order: the output, or the chunk's
    // partial sums
    wave_sync();
    for (int q = 0; q < R; q++) {
        const double ea = SPLIT ? sfA[q] : 0.0, eb = SPLIT ? sfB[q] : 0.0;
        sm.red[q * 64 + lane] = va ? acA.value(q) + ea : 0.0;
        sm.red[(R + q) * 64 + lane] = vb ? acB.value(q) + eb : 0.0;
    }
    wave_sync();
    if (lane < nslot) {
        for (int h = 0; h < 2; h++) {
            if (h == 1 && nslot + lane >= m) break;
            const int t = h * nslot + lane, i = b + t;
            const double wgt = (part < 0 && a.t_weight) ? a.t_weight[i] : 1.0;
            for (int q = 0; q < R; q++) {
                double s = 0;
                for (int k = 0; k < S; k++) s += sm.red[(h * R + q) * 64 + k * nslot + lane];
                if (part >= 0) {
                    a.part_buf[(size_t)part + q * m + t] = s;
                    continue;
                }
                double* o = a.out_sorted + (size_t)q * a.n_tgt + i;
                *o = a.accumulate ? *o + wgt * (s * scale) : wgt * (s * scale);
            }
        }
    }
}

// The leaves k_leaf_wave ran in chunks: one workgroup per leaf, one target per thread
// (m <= kWaveLeafMax), the chunks' partial sums added in chunk order.
__global__ __launch_bounds__(kWaveLeafMax) void k_leaf_parts(LeafStage a, int R, double scale) {
    const WaveLeafParts lp = a.wave_parts[blockIdx.x];
    const int leaf = a.leaves[lp.pos];
    const int b = a.tt.begin[leaf], m = a.tt.count[leaf];
    const int t = threadIdx.x;
    if (t >= m) return;
    const int i = b + t;
    const double wgt = a.t_weight ? a.t_weight[i] : 1.0;
    for (int q = 0; q < R; q++) {
        const double* p = a.part_buf + (size_t)lp.part + q * m + t;
        double s = 0;
        for (int c = 0; c < lp.n; c++) s += p[(size_t)c * kPartRows * m];
        double* o = a.out_sorted + (size_t)q * a.n_tgt + i;
        *o = a.accumulate ? *o + wgt * (s * scale) : wgt * (s * scale);
    }
}

// ------------------------------------------------------------ symmetric point pairs
// Kelvin U is symmetric, U(x <- y) = U(y <- x) (a symmetric 3x3 tensor, even in
// y - x), and so is the Laplace single layer 1/r, so with one point set on both sides
// a pair (i, j) of neighbouring leaves serves both u_i += U v_j and u_j += U v_i. The
// workgroup of a leaf (its m points, kSymLeafMin <= m <= kSymLeafMax, the targets)
// evaluates each pair with the points of the leaves it owns (the sources; every
// unordered leaf pair has one owner) once. A wave takes 64 x kSymSrc sources at a
// time, kSymSrc per lane, and visits the m targets in the rotated order
// t = (r + k) mod m, r = lane mod m: the lanes of one group (lane / m; one group when
// m >= 64) update distinct targets at every step, and each group has its own copy of
// the wave's target sums in LDS (one LDS add per component and step; a slot's adds
// arrive in instruction order, the same on every run); the source sums stay in
// registers over all m targets, then go to the reverse buffer at the source's slot;
// k_sym_gather adds those slots in list order.
// The transpose runs the same pairs (U^T = U) with each sum times t_weight, the
// source weight of the point it goes to.
struct alignas(16) SymTarget {
    double x, y, z, v0, v1, v2;
};
// (T: the pair arithmetic; positions stay FP64, so the displacement is exact before
// it is rounded to T, as in the other point-pair kernels)
template <class T>
struct SymSource {
    double x, y, z;
    T v0, v1, v2, r0, r1, r2;
};

// U v = C1 / r (C2 v + d (d.v) / r^2), C1 applied by the caller: f (the target's
// sum, set when SET) gets the source's density, s.r the target's
// (K_LS: u = v / r, 1/(4 pi) applied by the caller)
template <int KID, class T, bool CUT, bool SET>
__device__ __forceinline__ void sym_pair(SymSource<T>& s, double tx, double ty, double tz, T tv0, T tv1, T tv2,
                                         T u2, T* f) {
    const T dx = T(s.x - tx), dy = T(s.y - ty), dz = T(s.z - tz);
    const T r2 = dx * dx + dy * dy + dz * dz;
    T ri = rsqrt_fast(r2);
    if (CUT) ri = r2 < cutoff_r2<T>() ? T(0) : ri;
    if constexpr (KID == K_LS) {
        if (SET) f[0] = ri * s.v0;
        else f[0] = fma(ri, s.v0, f[0]);
        s.r0 = fma(ri, tv0, s.r0);
        return;
    }
    const T ri3 = ri * ri * ri, au = u2 * ri;
    const T bs = ri3 * (dx * s.v0 + dy * s.v1 + dz * s.v2);
    const T bt = ri3 * (dx * tv0 + dy * tv1 + dz * tv2);
    if (SET) {
        f[0] = fma(au, s.v0, bs * dx);
        f[1] = fma(au, s.v1, bs * dy);
        f[2] = fma(au, s.v2, bs * dz);
    } else {
        f[0] = fma(bs, dx, fma(au, s.v0, f[0]));
        f[1] = fma(bs, dy, fma(au, s.v1, f[1]));
        f[2] = fma(bs, dz, fma(au, s.v2, f[2]));
    }
    s.r0 = fma(bt, dx, fma(au, tv0, s.r0));
    s.r1 = fma(bt, dy, fma(au, tv1, s.r1));
    s.r2 = fma(bt, dz, fma(au, tv2, s.r2));
}

// Stream position p of the leaf's owned ranges [e0, e1) into s (beyond n: a point far
// away with zero density, whose pairs add exact zeros); close |= the point lies within
// kCloseMargin of the box of the leaf's points (box: lo x, y, z, hi x, y, z), where a
// pair might fall inside the r2 < 1e-12 cutoff. (Straight-line: a position past n
// reads point 0 and then takes the far point.)
template <int R, class T>
__device__ inline void sym_load(SymSource<T>& s, int p, int n, int e0, int e1, const LeafStage& a,
                                const double* box, int& close) {
    const bool ok = p < n;
    int lo = e0, hi = e1 - 1;  // the last range starting at or before p
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.sy_pre[mid] <= p) lo = mid;
        else hi = mid - 1;
    }
    const int k = ok ? a.sy_beg[lo] + (p - a.sy_pre[lo]) : 0;
    const size_t ns = (size_t)a.n_ps;
    const double x = a.ps.x[k], y = a.ps.y[k], z = a.ps.z[k];
    const double gx = fmax(fmax(box[0] - x, x - box[3]), 0.0), gy = fmax(fmax(box[1] - y, y - box[4]), 0.0),
                 gz = fmax(fmax(box[2] - z, z - box[5]), 0.0);
    if (ok && gx * gx + gy * gy + gz * gz <= kCloseMargin * kCloseMargin) close = 1;
    const double v0 = a.pdens[k], v1 = R > 1 ? a.pdens[ns + k] : 0.0, v2 = R > 1 ? a.pdens[2 * ns + k] : 0.0;
    s.x = ok ? x : 1e30;
    s.y = ok ? y : 1e30;
    s.z = ok ? z : 1e30;
    s.v0 = ok ? T(v0) : T(0);
    s.v1 = ok ? T(v1) : T(0);
    s.v2 = ok ? T(v2) : T(0);
    s.r0 = T(0);
    s.r1 = T(0);
    s.r2 = T(0);
}

// Step k takes slot r + k of the records and of the group's sums: records repeat the
// targets past slot m (slot s holds target s mod m), and the sums of slots m.. are
// folded onto their targets at the end. The target sums are FP64 in LDS whatever T is.
template <int KID, class T, bool CUT>
__device__ __forceinline__ void sym_step(SymSource<T>* sv, double2 A, double2 B, double2 C, double* acc, T u2) {
    constexpr int R = kernel_dim(KID);
    const T tv0 = T(B.y), tv1 = R > 1 ? T(C.x) : T(0), tv2 = R > 1 ? T(C.y) : T(0);
    T f[3];
    sym_pair<KID, T, CUT, true>(sv[0], A.x, A.y, B.x, tv0, tv1, tv2, u2, f);
    for (int u = 1; u < kSymSrc; u++) sym_pair<KID, T, CUT, false>(sv[u], A.x, A.y, B.x, tv0, tv1, tv2, u2, f);
    for (int q = 0; q < R; q++)
        __hip_atomic_fetch_add(acc + q * kSymSlots, (double)f[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Two steps per iteration, both records read first: LDS operations complete in
// order, so a read issued after a step's adds waits for them; this waits once per two.
template <int KID, class T, bool CUT>
__device__ __forceinline__ void sym_block(SymSource<T>* sv, const SymTarget* tg, double* fa, int m, int r, T u2) {
    const double2* rec = reinterpret_cast<const double2*>(tg + r);
    double* acc = fa + r;
    int k = 0;
    for (; k + 2 <= m; k += 2) {
        const double2 A0 = rec[0], B0 = rec[1], C0 = rec[2], A1 = rec[3], B1 = rec[4], C1 = rec[5];
        sym_step<KID, T, CUT>(sv, A0, B0, C0, acc, u2);
        sym_step<KID, T, CUT>(sv, A1, B1, C1, acc + 1, u2);
        rec += 6;
        acc += 2;
    }
    if (k < m) sym_step<KID, T, CUT>(sv, rec[0], rec[1], rec[2], acc, u2);
}

template <int KID, class T>
__global__ __launch_bounds__(kSymBlock) void k_leaf_sym(LeafStage a, KC<T> kc, double scale) {
    constexpr int NW = kSymBlock / 64;
    constexpr int R = kernel_dim(KID);
    __shared__ SymTarget tg[kSymSlots];
    __shared__ double facc[NW][3 * kSymSlots];
    const int pos = uni(a.sy_order[blockIdx.x]);
    if (pos < 0) return;  // padding of the layout (the whole workgroup)
    const int leaf = uni(a.leaves[pos]);
    const int b = uni(a.tt.begin[leaf]), m = uni(a.tt.count[leaf]);
    const int e0 = uni(a.sy_ptr[pos]), e1 = uni(a.sy_ptr[pos + 1]);
    const int n = uni(a.sy_pre[e1 - 1] + a.sy_cnt[e1 - 1]);  // listed positions own at least one range
    const size_t base = (size_t)uni(a.sy_base[pos]);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    // slots per group copy (r + k < S), the copies, and this lane's group and offset:
    // C * S <= 128 + 2m <= kSymSlots below 64 targets, S = m + 64 <= kSymSlots above
    const int S = m < 64 ? 2 * m : m + 64, C = m < 64 ? (64 + m - 1) / m : 1;
    const int gc = lane / m, gr = lane - gc * m;
    const size_t ns = (size_t)a.n_ps;
    for (int i = tid; i < S; i += kSymBlock) {
        const int j = b + (i < m ? i : i - m);
        SymTarget t;
        t.x = a.tp.x[j]; t.y = a.tp.y[j]; t.z = a.tp.z[j];
        t.v0 = a.pdens[j];
        t.v1 = R > 1 ? a.pdens[ns + j] : 0.0;
        t.v2 = R > 1 ? a.pdens[2 * ns + j] : 0.0;
        tg[i] = t;
    }
    for (int i = tid; i < NW * 3 * kSymSlots; i += kSymBlock) (&facc[0][0])[i] = 0.0;
    // the box of the leaf's points (one per thread: m <= kSymBlock)
    static_assert(kSymLeafMax <= kSymBlock, "one target per thread");
    __shared__ double wbox[NW][6];
    double bx[6];
    {
        const int i = tid < m ? b + tid : b;
        bx[0] = bx[3] = a.tp.x[i];
        bx[1] = bx[4] = a.tp.y[i];
        bx[2] = bx[5] = a.tp.z[i];
        for (int off = 32; off > 0; off >>= 1)
            for (int c = 0; c < 3; c++) {
                bx[c] = fmin(bx[c], __shfl_xor(bx[c], off));
                bx[3 + c] = fmax(bx[3 + c], __shfl_xor(bx[3 + c], off));
            }
        if (lane == 0)
            for (int c = 0; c < 6; c++) wbox[w][c] = bx[c];
    }
    __syncthreads();
    for (int c = 0; c < 3; c++) {
        bx[c] = wbox[0][c];
        bx[3 + c] = wbox[0][3 + c];
        for (int v = 1; v < NW; v++) {
            bx[c] = fmin(bx[c], wbox[v][c]);
            bx[3 + c] = fmax(bx[3 + c], wbox[v][3 + c]);
        }
    }
    double* fa = facc[w] + gc * S;
    constexpr int PB = 64 * kSymSrc;  // stream positions per wave block
    for (int p0 = w * PB; p0 < n; p0 += NW * PB) {
        SymSource<T> sv[kSymSrc];
        int close = 0;
        for (int u = 0; u < kSymSrc; u++) sym_load<R>(sv[u], p0 + 64 * u + lane, n, e0, e1, a, bx, close);
        if (__ballot(close) != 0ull) sym_block<KID, T, true>(sv, tg, fa, m, gr, kc.u2);
        else sym_block<KID, T, false>(sv, tg, fa, m, gr, kc.u2);
        for (int u = 0; u < kSymSrc; u++) {
            const int p = p0 + 64 * u + lane;
            if (p >= n) continue;
            double* r = a.rev + base + p;
            r[0] = (double)sv[u].r0;
            if (R > 1) {
                r[a.n_rev] = (double)sv[u].r1;
                r[2 * a.n_rev] = (double)sv[u].r2;
            }
        }
    }
    __syncthreads();
    for (int o = tid; o < R * m; o += kSymBlock) {
        const int q = o / m, i = o - q * m;
        double s = 0;
        for (int v = 0; v < NW; v++)
            for (int c = 0; c < C; c++) {
                const double* f = &facc[v][q * kSymSlots + c * S];
                s += f[i];
                if (i + m < S) s += f[i + m];
            }
        const double wgt = a.t_weight ? a.t_weight[b + i] : 1.0;
        double* out = a.out_sorted + (size_t)q * a.n_tgt + b + i;
        *out += wgt * (s * scale);
    }
}

// u of the leaves at sg_pos += the reverse sums addressed to them, slot by slot in
// list order
__global__ __launch_bounds__(256) void k_sym_gather(LeafStage a, int R, double scale) {
    const int pos = a.sg_pos[blockIdx.x];
    const int leaf = a.leaves[pos];
    const int b = a.tt.begin[leaf], m = a.tt.count[leaf];
    const int e0 = a.sg_ptr[blockIdx.x], e1 = a.sg_ptr[blockIdx.x + 1];
    for (int o = threadIdx.x; o < R * m; o += 256) {
        const int q = o / m, i = o - q * m;
        const double* r = a.rev + (size_t)q * a.n_rev + i;
        double s = 0;
        int e = e0;
        for (; e + 8 <= e1; e += 8) {  // eight loads in flight, added in list order
            double v[8];
            for (int u = 0; u < 8; u++) v[u] = r[a.sg_off[e + u]];
            for (int u = 0; u < 8; u++) s += v[u];
        }
        for (; e < e1; e++) s += r[a.sg_off[e]];
        const double wgt = a.t_weight ? a.t_weight[b + i] : 1.0;
        double* out = a.out_sorted + (size_t)q * a.n_tgt + b + i;
        *out += wgt * (s * scale);
    }
}

// ------------------------------------------------- observation-coordinate gradient
// Forward leaf stage (P2P + M2P + L2P) differentiated with respect to the target
// coordinates; writes grad[3*i + m] = sum_q g_q(i) d u_q(i) / d x_i[m]. This
// kernel serves the kernels without a closed-form gradient (adjoint traction,
// hypersingular, identities), in FP64 dual numbers, one target per lane;
// k_leaf_grad2 below serves the others.
template <int KID>
__global__ __launch_bounds__(kGradBlock) void k_leaf_grad(LeafStage a, KC<D3> kc, double scale) {
    static_assert(!has_analytic_grad(KID), "closed-form kernels use k_leaf_grad2");
    constexpr int NB = kGradBlock;
    constexpr int R = kernel_dim(KID);
    constexpr bool NRM = kernel_uses_nsrc(KID) || kernel_uses_nobs(KID);
    __shared__ StageSmem<double, R, NRM, NB, 3> sm;
    const int pos = blockIdx.x;
    const int leaf = uni(a.leaves[pos]);
    const int b = uni(a.tt.begin[leaf]), m = uni(a.tt.count[leaf]);
    const int pe0 = a.rp_ptr ? uni(a.rp_ptr[pos]) : 0, pe1 = a.rp_ptr ? uni(a.rp_ptr[pos + 1]) : 0;
    const int se0 = a.sl_ptr ? uni(a.sl_ptr[pos]) : 0, se1 = a.sl_ptr ? uni(a.sl_ptr[pos + 1]) : 0;
    const int tid = threadIdx.x;
    for (int t0 = 0; t0 < m; t0 += NB) {
        const int nt = min(NB, m - t0);
        const int ntp = nt, S = NB / ntp;
        const int tg = tid % ntp, slice = tid / ntp;
        const bool active = slice < S;
        const int i = b + t0 + tg;
        const D3 tx(a.tp.x[i], 1, 0, 0), ty(a.tp.y[i], 0, 1, 0), tz(a.tp.z[i], 0, 0, 1);
        const D3 tn[3] = {D3(NRM ? a.tp.nx[i] : 0.0), D3(NRM ? a.tp.ny[i] : 0.0), D3(NRM ? a.tp.nz[i] : 0.0)};
        double gt[R];
        for (int q = 0; q < R; q++) gt[q] = a.grad_g[(size_t)q * a.n_tgt + i];
        drain_loads();
        D3 acc[R];
        // one FP64 source j of tile tl, differentiated by the dual parts of the target
        auto pair_d = [&](const Tile<double, R, NRM, NB>& tl, int j) {
            const D3 sn[3] = {D3(NRM ? tl.nx[j] : 0.0), D3(NRM ? tl.ny[j] : 0.0), D3(NRM ? tl.nz[j] : 0.0)};
            D3 v[R];
            for (int q = 0; q < R; q++) v[q] = D3(tl.d[q][j]);
            const D3 dx = D3(tl.x[j]) - tx, dy = D3(tl.y[j]) - ty, dz = D3(tl.z[j]) - tz;
            const D3 r2 = dx * dx + dy * dy + dz * dz;
            kop<KID, false, D3>(dx, dy, dz, r2, tn, sn, v, acc, kc);
        };
        auto points = [&](int cnt) {
            if (!active) return;
            for (int j = slice; j < cnt; j += S) pair_d(sm.pt, j);
        };
        auto surfaces = [&](int cnt) {
            if (!active) return;
            for (int j = slice; j < cnt; j += S) pair_d(sm.st, j);
        };
        stream_ranges_packed(sm.pt, sm.rpre, sm.rbeg, pe0, pe1, a.rp_beg, a.rp_cnt, a.rp_pre, a.ps, a.pdens,
                             a.n_ps, points);
        stream_surfaces(sm.st, se0, se1, a.sl_cell, a.slt, a.sl_surf, a.unit, a.sl_coeff, a.n_e, a.blk, surfaces);
        if (a.self_coeff)
            stream_surfaces(sm.st, 0, 1, &leaf, a.tt, a.self_surf, a.unit, a.self_coeff, a.n_e, a.blk, surfaces);
        double gr[3] = {0, 0, 0};
        if (active) {
            for (int q = 0; q < R; q++) {
                gr[0] += gt[q] * acc[q].d0;
                gr[1] += gt[q] * acc[q].d1;
                gr[2] += gt[q] * acc[q].d2;
            }
        }
        __syncthreads();
        for (int mm = 0; mm < 3; mm++) sm.red[mm * NB + tid] = gr[mm];
        __syncthreads();
        slice_sum<NB>(sm.red, sm.part, 3, ntp, S, [&](int mm, int g, double s) {
            a.grad_out[3 * (size_t)(b + t0 + g) + mm] = s * scale;
        });
        __syncthreads();
    }
}

// The same gradient for the kernels with closed forms, two targets per lane (as
// k_leaf_stage: slots tg and nslot + tg, a lane without a second target repeats
// the first and drops it), so one LDS read of a source feeds two pairs. Point
// sources in T with Neumaier-compensated per-tile sums when T is float; surface
// sources FP64.
// Four waves per SIMD (<= 128 VGPRs) except the FP32 elastic-U instance, which
// spills under that cap (measured 2M points, order 100: 30.1 ms uncapped against
// 33.5 capped; the FP64 and traction instances are 1-5% faster capped).
// A lane's gradient for one target: the FP64 sums G (and, when SPLIT, the FP32 tile
// sums P) of grad_pair, with the FP64 scalar slot of K_EU applied (G[m] - g_m G[3]).
template <int KID, bool SPLIT, class P>
__device__ __forceinline__ void grad_total(const double* G, const P& p, const double* g, double* out) {
    for (int m = 0; m < 3; m++) {
        out[m] = G[m] + (SPLIT ? p.value(m) : 0.0);
        if constexpr (grad_slots<KID, double>() == 4) out[m] -= g[m] * G[3];
    }
}

template <int KID, class T>
constexpr int grad_waves() {
    return (KID == K_EU && std::is_same<T, float>::value) ? 1 : 4;
}

template <int KID, class T>
__global__ __launch_bounds__(kGradBlock, (grad_waves<KID, T>())) void k_leaf_grad2(LeafStage a, KC<double> kc64,
                                                                                 KC<T> kct, double scale) {
    static_assert(has_analytic_grad(KID), "closed-form gradient kernels only");
    constexpr int NB = kGradBlock;
    constexpr int R = kernel_dim(KID);
    constexpr bool NRM = kernel_uses_nsrc(KID) || kernel_uses_nobs(KID);
    constexpr bool SPLIT = !std::is_same<T, double>::value;
    __shared__ StageSmem<T, R, NRM, NB, 6> sm;
    const int pos = a.order ? uni(a.order[blockIdx.x]) : (int)blockIdx.x;
    if (pos < 0) return;  // padding of the layout (the whole workgroup)
    const int leaf = uni(a.leaves[pos]);
    const int b = uni(a.tt.begin[leaf]), m = uni(a.tt.count[leaf]);
    const int pe0 = a.rp_ptr ? uni(a.rp_ptr[pos]) : 0, pe1 = a.rp_ptr ? uni(a.rp_ptr[pos + 1]) : 0;
    const int se0 = a.sl_ptr ? uni(a.sl_ptr[pos]) : 0, se1 = a.sl_ptr ? uni(a.sl_ptr[pos + 1]) : 0;
    const int tid = threadIdx.x;
    for (int t0 = 0; t0 < m; t0 += 2 * NB) {
        const int nt = min(2 * NB, m - t0);
        const int nslot = (nt + 1) >> 1;
        const int ntp = nslot, S = NB / ntp;  // lanes >= ntp*S sit out
        const int tg = tid % ntp, slice = tid / ntp;
        const bool va = slice < S, vb = va && nslot + tg < nt;
        const int ia = b + t0 + tg;
        const int ib = b + t0 + (vb ? nslot + tg : tg);
        const double ax = a.tp.x[ia], ay = a.tp.y[ia], az = a.tp.z[ia];
        const double bx = a.tp.x[ib], by = a.tp.y[ib], bz = a.tp.z[ib];
        constexpr int NG = grad_slots<KID, double>();
    double gA[R], gB[R], GA[NG] = {}, GB[NG] = {};
        T gAT[R], gBT[R];
        for (int q = 0; q < R; q++) {
            gA[q] = a.grad_g[(size_t)q * a.n_tgt + ia];
            gB[q] = a.grad_g[(size_t)q * a.n_tgt + ib];
            gAT[q] = T(gA[q]);
            gBT[q] = T(gB[q]);
        }
        drain_loads();
        Acc<T, grad_slots<KID, T>()> PA, PB;
        PA.init();
        PB.init();
        int tclose = 1;  // set per tile by stream_ranges_packed when the plan has close flags
        auto point_loop = [&](int cnt, auto cut) {
            constexpr bool C = decltype(cut)::value;
            const Tile<T, R, NRM, NB>& tl = sm.pt;
            for (int j = slice; j < cnt; j += S) {
                const T sn[3] = {NRM ? tl.nx[j] : T(0), NRM ? tl.ny[j] : T(0), NRM ? tl.nz[j] : T(0)};
                T v[R];
                for (int q = 0; q < R; q++) v[q] = tl.d[q][j];
                const double sx = tl.x[j], sy = tl.y[j], sz = tl.z[j];
                const T dxa = T(sx - ax), dya = T(sy - ay), dza = T(sz - az);
                const T dxb = T(sx - bx), dyb = T(sy - by), dzb = T(sz - bz);
                if constexpr (SPLIT) {
                    grad_pair<KID, T, true, C>(dxa, dya, dza, dxa * dxa + dya * dya + dza * dza, sn, v, gAT, PA.tile(), kct);
                    grad_pair<KID, T, true, C>(dxb, dyb, dzb, dxb * dxb + dyb * dyb + dzb * dzb, sn, v, gBT, PB.tile(), kct);
                } else {
                    grad_pair<KID, double, true, C>(dxa, dya, dza, dxa * dxa + dya * dya + dza * dza, sn, v, gA, GA, kc64);
                    grad_pair<KID, double, true, C>(dxb, dyb, dzb, dxb * dxb + dyb * dyb + dzb * dzb, sn, v, gB, GB, kc64);
                }
            }
        };
        auto points = [&](int cnt) {
            if (va) {
                if (uni(tclose)) point_loop(cnt, std::true_type{});
                else point_loop(cnt, std::false_type{});  // no point of the tile near the leaf
            }
            PA.flush();
            PB.flush();
        };
        auto surfaces = [&](int cnt) {
            if (!va) return;
            const Tile<double, R, NRM, NB>& tl = sm.st;
            for (int j = slice; j < cnt; j += S) {
                const double sn[3] = {NRM ? tl.nx[j] : 0.0, NRM ? tl.ny[j] : 0.0, NRM ? tl.nz[j] : 0.0};
                double v[R];
                for (int q = 0; q < R; q++) v[q] = tl.d[q][j];
                const double dxa = tl.x[j] - ax, dya = tl.y[j] - ay, dza = tl.z[j] - az;
                const double dxb = tl.x[j] - bx, dyb = tl.y[j] - by, dzb = tl.z[j] - bz;
                grad_pair<KID, double>(dxa, dya, dza, dxa * dxa + dya * dya + dza * dza, sn, v, gA, GA, kc64);
                grad_pair<KID, double>(dxb, dyb, dzb, dxb * dxb + dyb * dyb + dzb * dzb, sn, v, gB, GB, kc64);
            }
        };
        stream_ranges_packed(sm.pt, sm.rpre, sm.rbeg, pe0, pe1, a.rp_beg, a.rp_cnt, a.rp_pre, a.ps, a.pdens,
                             a.n_ps, points, a.rp_close, a.rp_close ? &tclose : nullptr, sm.wflag);
        stream_surfaces(sm.st, se0, se1, a.sl_cell, a.slt, a.sl_surf, a.unit, a.sl_coeff, a.n_e, a.blk, surfaces);
        if (a.self_coeff)
            stream_surfaces(sm.st, 0, 1, &leaf, a.tt, a.self_surf, a.unit, a.self_coeff, a.n_e, a.blk, surfaces);
        __syncthreads();
        double fa[3], fb[3];
        grad_total<KID, SPLIT>(GA, PA, gA, fa);
        grad_total<KID, SPLIT>(GB, PB, gB, fb);
        for (int mm = 0; mm < 3; mm++) {
            sm.red[mm * NB + tid] = va ? fa[mm] : 0.0;
            sm.red[(3 + mm) * NB + tid] = vb ? fb[mm] : 0.0;
        }
        __syncthreads();
        slice_sum<NB>(sm.red, sm.part, 6, ntp, S, [&](int v, int g, double s) {
            const int h = v / 3, mm = v - h * 3;
            if (h == 1 && nslot + g >= nt) return;
            a.grad_out[3 * (size_t)(b + t0 + h * nslot + g) + mm] = s * scale;
        });
        __syncthreads();
    }
}

// k_leaf_grad2 one wave per work item (a small leaf or a chunk of a large one), as
// k_leaf_wave is to k_leaf_stage: the same items and layout, two targets per lane,
// S = 64 / ceil(m/2) source slices, each wave on its own 64-entry LDS tiles with no
// workgroup barrier. A chunk writes its 3 x m sums to part_buf and k_leaf_grad_parts
// adds a leaf's chunks in chunk order.
template <class PT, int R, bool NRM>
struct GradWaveSmem {
    union {
        Tile<PT, R, NRM, 64> pt;
        Tile<double, R, NRM, 64> st;
        double red[6 * 64];
    };
};

template <int KID, class T>
__global__ __launch_bounds__(kWaveBlock) void k_leaf_grad2_wave(LeafStage a, KC<double> kc64, KC<T> kct,
                                                                double scale) {
    static_assert(has_analytic_grad(KID), "closed-form gradient kernels only");
    constexpr int R = kernel_dim(KID);
    constexpr bool NRM = kernel_uses_nsrc(KID) || kernel_uses_nobs(KID);
    constexpr bool SPLIT = !std::is_same<T, double>::value;
    __shared__ GradWaveSmem<T, R, NRM> smw[kWaveBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int idx = blockIdx.x * (kWaveBlock / 64) + w;
    if (idx >= a.n_wave) return;  // the whole wave: nothing below synchronizes the block
    const WaveItem* it = a.wave_items + idx;
    const int pos = uni(it->pos);
    if (pos < 0) return;  // padding of the layout
    GradWaveSmem<T, R, NRM>& sm = smw[w];
    const int leaf = uni(a.leaves[pos]);
    const int b = uni(a.tt.begin[leaf]), m = uni(a.tt.count[leaf]);  // m <= kWaveLeafMax
    const int nslot = (m + 1) >> 1, S = 64 / nslot;
    const int tg = lane % nslot, slice = lane / nslot;
    const bool va = slice < S, vb = va && nslot + tg < m;
    const int ia = b + tg, ib = b + (vb ? nslot + tg : tg);
    const double ax = a.tp.x[ia], ay = a.tp.y[ia], az = a.tp.z[ia];
    const double bx = a.tp.x[ib], by = a.tp.y[ib], bz = a.tp.z[ib];
    constexpr int NG = grad_slots<KID, double>();
    double gA[R], gB[R], GA[NG] = {}, GB[NG] = {};
    T gAT[R], gBT[R];
    for (int q = 0; q < R; q++) {
        gA[q] = a.grad_g[(size_t)q * a.n_tgt + ia];
        gB[q] = a.grad_g[(size_t)q * a.n_tgt + ib];
        gAT[q] = T(gA[q]);
        gBT[q] = T(gB[q]);
    }
    const int f0 = uni(it->f0), f1 = uni(it->f1), v0 = uni(it->v0), v1 = uni(it->v1);
    const int se0 = uni(it->c0), se1 = uni(it->c1), self = uni(it->self), part = uni(it->part);
    drain_loads();
    Acc<T, grad_slots<KID, T>()> PA, PB;
    PA.init();
    PB.init();

    // (a) point sources [v0, v1) of the leaf's packed ranges, as in k_leaf_wave
    if (a.rp_ptr && v0 < v1) {
        const int total = v1 - v0;
        SrcReg<R, NRM> r;
        int f = f0;
        int mclose = 0;  // as in k_leaf_wave
        auto fetch = [&](int base) {
            const int v = v0 + base + lane;
            mclose = 0;
            if (v < v1) {
                while (f + 1 < f1 && a.rp_pre[f + 1] <= v) f++;
                load_point(r, a.rp_beg[f] + (v - a.rp_pre[f]), a.ps, a.pdens, a.n_ps);
                mclose = a.rp_close ? a.rp_close[f] : 1;
            }
        };
        auto point_loop = [&](int cnt, auto cut) {
            constexpr bool C = decltype(cut)::value;
            const Tile<T, R, NRM, 64>& tl = sm.pt;
            for (int j = slice; j < cnt; j += S) {
                const T sn[3] = {NRM ? tl.nx[j] : T(0), NRM ? tl.ny[j] : T(0), NRM ? tl.nz[j] : T(0)};
                T v[R];
                for (int q = 0; q < R; q++) v[q] = tl.d[q][j];
                const double sx = tl.x[j], sy = tl.y[j], sz = tl.z[j];
                const T dxa = T(sx - ax), dya = T(sy - ay), dza = T(sz - az);
                const T dxb = T(sx - bx), dyb = T(sy - by), dzb = T(sz - bz);
                if constexpr (SPLIT) {
                    grad_pair<KID, T, true, C>(dxa, dya, dza, dxa * dxa + dya * dya + dza * dza, sn, v, gAT, PA.tile(), kct);
                    grad_pair<KID, T, true, C>(dxb, dyb, dzb, dxb * dxb + dyb * dyb + dzb * dzb, sn, v, gBT, PB.tile(), kct);
                } else {
                    grad_pair<KID, double, true, C>(dxa, dya, dza, dxa * dxa + dya * dya + dza * dza, sn, v, gA, GA, kc64);
                    grad_pair<KID, double, true, C>(dxb, dyb, dzb, dxb * dxb + dyb * dyb + dzb * dzb, sn, v, gB, GB, kc64);
                }
            }
        };
        fetch(0);
        for (int base = 0; base < total; base += 64) {
            const int cnt = min(64, total - base);
            wave_sync();
            if (lane < cnt) store_tile(sm.pt, lane, r);
            const bool tclose = __ballot(mclose) != 0ull;
            wave_sync();
            if (base + 64 < total) fetch(base + 64);
            if (va) {
                if (tclose) point_loop(cnt, std::true_type{});
                else point_loop(cnt, std::false_type{});
            }
            PA.flush();
            PB.flush();
        }
    }
    // (b), (c) surface sources, FP64
    auto surfaces = [&](const int* cells, int c0, int c1, const TreeView& st, const double* surf,
                        const double* coeff) {
        const int total = (c1 - c0) * a.n_e;
        if (total <= 0) return;
        SrcReg<R, NRM> r;
        auto fetch = [&](int base) {
            const int v = base + lane;
            if (v < total) {
                const int k = v / a.n_e;
                load_surf(r, cells[c0 + k], v - k * a.n_e, st, surf, a.unit, coeff, a.n_e, a.blk);
            }
        };
        fetch(0);
        for (int base = 0; base < total; base += 64) {
            const int cnt = min(64, total - base);
            wave_sync();
            if (lane < cnt) store_tile(sm.st, lane, r);
            wave_sync();
            if (base + 64 < total) fetch(base + 64);
            if (va) {
                const Tile<double, R, NRM, 64>& tl = sm.st;
                for (int j = slice; j < cnt; j += S) {
                    const double sn[3] = {NRM ? tl.nx[j] : 0.0, NRM ? tl.ny[j] : 0.0, NRM ? tl.nz[j] : 0.0};
                    double v[R];
                    for (int q = 0; q < R; q++) v[q] = tl.d[q][j];
                    const double dxa = tl.x[j] - ax, dya = tl.y[j] - ay, dza = tl.z[j] - az;
                    const double dxb = tl.x[j] - bx, dyb = tl.y[j] - by, dzb = tl.z[j] - bz;
                    grad_pair<KID, double>(dxa, dya, dza, dxa * dxa + dya * dya + dza * dza, sn, v, gA, GA, kc64);
                    grad_pair<KID, double>(dxb, dyb, dzb, dxb * dxb + dyb * dyb + dzb * dzb, sn, v, gB, GB, kc64);
                }
            }
        }
    };
    if (a.sl_ptr && se0 < se1) surfaces(a.sl_cell, se0, se1, a.slt, a.sl_surf, a.sl_coeff);
    if (self && a.self_coeff) surfaces(&leaf, 0, 1, a.tt, a.self_surf, a.self_coeff);

    // the S slices of each target, summed in slice order
    wave_sync();
    double fa[3], fb[3];
    grad_total<KID, SPLIT>(GA, PA, gA, fa);
    grad_total<KID, SPLIT>(GB, PB, gB, fb);
    for (int mm = 0; mm < 3; mm++) {
        sm.red[mm * 64 + lane] = va ? fa[mm] : 0.0;
        sm.red[(3 + mm) * 64 + lane] = vb ? fb[mm] : 0.0;
    }
    wave_sync();
    if (lane < nslot) {
        for (int h = 0; h < 2; h++) {
            if (h == 1 && nslot + lane >= m) break;
            const int t = h * nslot + lane;
            for (int mm = 0; mm < 3; mm++) {
                double s = 0;
                for (int k = 0; k < S; k++) s += sm.red[(h * 3 + mm) * 64 + k * nslot + lane];
                if (part >= 0) a.part_buf[(size_t)part + mm * m + t] = s;
                else a.grad_out[3 * (size_t)(b + t) + mm] = s * scale;
            }
        }
    }
}

// The gradient's chunked leaves: grad_out[3 i + mm] = the chunks' sums, in chunk order.
__global__ __launch_bounds__(kWaveLeafMax) void k_leaf_grad_parts(LeafStage a, double scale) {
    const WaveLeafParts lp = a.wave_parts[blockIdx.x];
    const int leaf = a.leaves[lp.pos];
    const int b = a.tt.begin[leaf], m = a.tt.count[leaf];
    const int t = threadIdx.x;
    if (t >= m) return;
    for (int mm = 0; mm < 3; mm++) {
        const double* p = a.part_buf + (size_t)lp.part + mm * m + t;
        double s = 0;
        for (int c = 0; c < lp.n; c++) s += p[(size_t)c * kPartRows * m];
        a.grad_out[3 * (size_t)(b + t) + mm] = s * scale;
    }
}

// ------------------------------------------------------------------------ helpers
// out[c] (= or +=) op(M[level c]) in[c] (+ addin[c]) for the cells of each tile:
// a 64 (rows) x 64 (cells) output tile per workgroup, 4 x 4 per thread (rows ty*4+i,
// cells tx+16j), k in chunks of kGemmK staged in LDS; the next chunk is loaded into
// registers while the current one is multiplied, so its latency overlaps the FMAs.
constexpr int kGemmK = 32;

__global__ __launch_bounds__(256) void k_level_gemm(int n, const int* __restrict__ cells,
                                                    const int* __restrict__ tiles, const double* __restrict__ M,
                                                    int trans, const double* __restrict__ in,
                                                    const double* __restrict__ addin, double* __restrict__ out,
                                                    float* __restrict__ out32, int add) {
    constexpr int KC = kGemmK, PER = 64 * KC / 256;  // elements of each operand per thread and chunk
    const int t = blockIdx.y;
    const int first = uni(tiles[3 * t]), cnt = uni(tiles[3 * t + 1]), lvl = uni(tiles[3 * t + 2]);
    const double* A = M + (size_t)lvl * n * n;
    const int m0 = blockIdx.x * 64;
    __shared__ double As[KC][65], Bs[KC][65];
    __shared__ int cl[64];
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    if (tid < 64) cl[tid] = tid < cnt ? cells[first + tid] : -1;
    __syncthreads();
    double pa[PER], pb[PER];
    // A chunk: As[k][r] = op(A)[m0 + r][k0 + k]; B chunk: Bs[k][c] = in[cell c][k0 + k]
    auto fetch = [&](int k0) {
        for (int q = 0; q < PER; q++) {
            const int e = tid + 256 * q;
            const int r = trans ? e % 64 : e / KC, k = trans ? e / 64 : e % KC;
            const int row = m0 + r, kk = k0 + k;
            pa[q] = (row < n && kk < n) ? (trans ? A[(size_t)kk * n + row] : A[(size_t)row * n + kk]) : 0.0;
            const int c = e / KC, kb = e % KC, cell = cl[c];
            pb[q] = (cell >= 0 && k0 + kb < n) ? in[(size_t)cell * n + k0 + kb] : 0.0;
        }
    };
    auto store = [&]() {
        for (int q = 0; q < PER; q++) {
            const int e = tid + 256 * q;
            const int r = trans ? e % 64 : e / KC, k = trans ? e / 64 : e % KC;
            As[k][r] = pa[q];
            Bs[e % KC][e / KC] = pb[q];
        }
    };
    double acc[4][4] = {};
    fetch(0);
    for (int k0 = 0; k0 < n; k0 += KC) {
        store();
        __syncthreads();
        if (k0 + KC < n) fetch(k0 + KC);
        for (int k = 0; k < KC; k++) {
            double av[4], bv[4];
            for (int i = 0; i < 4; i++) av[i] = As[k][ty * 4 + i];
            for (int j = 0; j < 4; j++) bv[j] = Bs[k][tx + 16 * j];
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) acc[i][j] += av[i] * bv[j];
        }
        __syncthreads();
    }
    for (int j = 0; j < 4; j++) {
        const int cell = cl[tx + 16 * j];
        if (cell < 0) continue;
        for (int i = 0; i < 4; i++) {
            const int row = m0 + ty * 4 + i;
            if (row >= n) continue;
            const size_t o = (size_t)cell * n + row;
            double v = acc[i][j];
            if (addin) v += addin[o];
            if (out32) out32[o] = (float)v;
            else out[o] = add ? out[o] + v : v;
        }
    }
}

__global__ void k_gather_dens(const double* __restrict__ x, const int* __restrict__ perm,
                              const double* __restrict__ w, int n, int R, double* __restrict__ dens) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int g = perm[i];
    const double wi = w ? w[i] : 1.0;
    for (int q = 0; q < R; q++) dens[(size_t)q * n + i] = x[(size_t)q * n + g] * wi;
}

__global__ void k_scatter_out(const double* __restrict__ out_sorted, const int* __restrict__ perm, int n,
                              int R, double* __restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int g = perm[i];
    for (int q = 0; q < R; q++) y[(size_t)q * n + g] = out_sorted[(size_t)q * n + i];
}

__global__ void k_fill(double* __restrict__ p, size_t n, double v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

CellTiles make_cell_tiles(const std::vector<int>& cells_in, const HostTree& t, hipStream_t s) {
    CellTiles T;
    std::vector<int> cells(cells_in);
    std::stable_sort(cells.begin(), cells.end(), [&](int a, int b) { return t.level[a] < t.level[b]; });
    std::vector<int> tiles;
    for (size_t i = 0; i < cells.size();) {
        const int lvl = t.level[cells[i]];
        size_t j = i;
        while (j < cells.size() && j - i < 64 && t.level[cells[j]] == lvl) j++;
        tiles.push_back((int)i);
        tiles.push_back((int)(j - i));
        tiles.push_back(lvl);
        i = j;
    }
    T.n = (int)(tiles.size() / 3);
    T.cells.upload(cells, s);
    T.tiles.upload(tiles, s);
    TBEM_HIP(hipStreamSynchronize(s));  // the host vectors go out of scope
    return T;
}

CellTiles make_level_tiles(const HostTree& t, int level, hipStream_t s) {
    std::vector<int> cells;
    for (int c = t.level_begin[level]; c < t.level_begin[level + 1]; c++) cells.push_back(c);
    return make_cell_tiles(cells, t, s);
}

void launch_level_gemm(const CellTiles& T, int blk, const double* M, bool trans, const double* in,
                       const double* addin, double* out, float* out32, bool add, hipStream_t s) {
    if (T.n == 0) return;
    hipLaunchKernelGGL(k_level_gemm, dim3(div_up(blk, 64), T.n), dim3(256), 0, s, blk, T.cells.p, T.tiles.p, M,
                       trans ? 1 : 0, in, addin, out, out32, add ? 1 : 0);
    TBEM_LAUNCH_CHECK();
}

// One workgroup per item: the tight box of the item's target points (block min/max),
// then per range whether any of its points lies within margin of that box.
__global__ __launch_bounds__(256) void k_range_close(const int* __restrict__ leaves, TreeView tt, PtsView tp,
                                                     const int* __restrict__ ptr, const int* __restrict__ rbeg,
                                                     const int* __restrict__ rcnt, PtsView sp, double margin,
                                                     int* __restrict__ close) {
    __shared__ double red[6][256];
    __shared__ double box[6];
    const int k = blockIdx.x, tid = threadIdx.x;
    const int leaf = uni(leaves[k]);
    const int b = uni(tt.begin[leaf]), m = uni(tt.count[leaf]);
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = tid; i < m; i += 256) {
        const double p[3] = {tp.x[b + i], tp.y[b + i], tp.z[b + i]};
        for (int a = 0; a < 3; a++) {
            lo[a] = fmin(lo[a], p[a]);
            hi[a] = fmax(hi[a], p[a]);
        }
    }
    for (int a = 0; a < 3; a++) {
        red[a][tid] = lo[a];
        red[3 + a][tid] = hi[a];
    }
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h)
            for (int a = 0; a < 3; a++) {
                red[a][tid] = fmin(red[a][tid], red[a][tid + h]);
                red[3 + a][tid] = fmax(red[3 + a][tid], red[3 + a][tid + h]);
            }
        __syncthreads();
    }
    if (tid < 6) box[tid] = red[tid][0];
    __syncthreads();
    const double m2 = margin * margin;
    const int e0 = uni(ptr[k]), e1 = uni(ptr[k + 1]);
    for (int e = e0; e < e1; e++) {
        const int rb = uni(rbeg[e]), rc = uni(rcnt[e]);
        int near = 0;
        for (int j = tid; j < rc; j += 256) {
            const double p[3] = {sp.x[rb + j], sp.y[rb + j], sp.z[rb + j]};
            double d2 = 0;
            for (int a = 0; a < 3; a++) {
                const double g = fmax(fmax(box[a] - p[a], p[a] - box[3 + a]), 0.0);
                d2 += g * g;
            }
            near |= (d2 <= m2) ? 1 : 0;
        }
        near = __syncthreads_or(near);
        if (tid == 0) close[e] = near ? 1 : 0;
    }
}

void launch_range_close(int n_items, const int* leaves, const TreeView& tt, const PtsView& tp, const int* ptr,
                        const int* rbeg, const int* rcnt, const PtsView& sp, double margin, int* close,
                        hipStream_t s) {
    if (n_items == 0) return;
    hipLaunchKernelGGL(k_range_close, dim3(n_items), dim3(256), 0, s, leaves, tt, tp, ptr, rbeg, rcnt, sp, margin,
                       close);
    TBEM_LAUNCH_CHECK();
}

void launch_gather_dens(const double* x, const int* perm, const double* w, int n, int R, double* dens,
                        hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_gather_dens, dim3(div_up(n, 256)), dim3(256), 0, s, x, perm, w, n, R, dens);
    TBEM_LAUNCH_CHECK();
}

void launch_scatter_out(const double* out_sorted, const int* perm, int n, int R, double* y, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_scatter_out, dim3(div_up(n, 256)), dim3(256), 0, s, out_sorted, perm, n, R, y);
    TBEM_LAUNCH_CHECK();
}

void launch_fill(double* p, size_t n, double v, hipStream_t s) {
    if (n == 0) return;
    int blocks = (int)std::min<size_t>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, s, p, n, v);
    TBEM_LAUNCH_CHECK();
}

// ------------------------------------------------------------------------ dispatch
template <class T>
static KC<T> kc_of(const KConst& c);
template <>
KC<double> kc_of<double>(const KConst& c) { return make_kc64(c); }
template <>
KC<float> kc_of<float>(const KConst& c) { return make_kc32(c); }

template <int KID, class T>
static void surf_launch(bool src_surf, bool add, bool adj, const SurfStage& a, const KConst& c, hipStream_t s) {
    if (a.n_items == 0) return;
    const KC<double> kc = make_kc64(c);
    const double scale = (double)kop_scale<KID, double>(kc);
    const size_t shm = (size_t)kernel_dim(KID) * a.n_e * sizeof(double);
    dim3 g(a.n_items), b(kSurfBlock);
    const int v = (src_surf && a.no_cutoff ? 8 : src_surf ? 4 : 0) + (adj ? 2 : 0) + (a.n_e > kSurfBlock / 2 ? 1 : 0);
    const int ad = add ? 1 : 0;
    switch (v) {
    case 8: hipLaunchKernelGGL((k_surf_stage<KID, true, false, false, false>), g, b, shm, s, a, kc, scale, ad); break;
    case 9: hipLaunchKernelGGL((k_surf_stage<KID, true, false, true, false>), g, b, shm, s, a, kc, scale, ad); break;
    case 10: hipLaunchKernelGGL((k_surf_stage<KID, true, true, false, false>), g, b, shm, s, a, kc, scale, ad); break;
    case 11: hipLaunchKernelGGL((k_surf_stage<KID, true, true, true, false>), g, b, shm, s, a, kc, scale, ad); break;
    case 0: hipLaunchKernelGGL((k_surf_stage<KID, false, false, false>), g, b, shm, s, a, kc, scale, ad); break;
    case 1: hipLaunchKernelGGL((k_surf_stage<KID, false, false, true>), g, b, shm, s, a, kc, scale, ad); break;
    case 2: hipLaunchKernelGGL((k_surf_stage<KID, false, true, false>), g, b, shm, s, a, kc, scale, ad); break;
    case 3: hipLaunchKernelGGL((k_surf_stage<KID, false, true, true>), g, b, shm, s, a, kc, scale, ad); break;
    case 4: hipLaunchKernelGGL((k_surf_stage<KID, true, false, false>), g, b, shm, s, a, kc, scale, ad); break;
    case 5: hipLaunchKernelGGL((k_surf_stage<KID, true, false, true>), g, b, shm, s, a, kc, scale, ad); break;
    case 6: hipLaunchKernelGGL((k_surf_stage<KID, true, true, false>), g, b, shm, s, a, kc, scale, ad); break;
    default: hipLaunchKernelGGL((k_surf_stage<KID, true, true, true>), g, b, shm, s, a, kc, scale, ad); break;
    }
    TBEM_LAUNCH_CHECK();
}

template <int KID, class T>
static void leaf_launch(bool adj, bool fast, const LeafStage& a, const KConst& c, hipStream_t s) {
    if (a.n_leaves == 0) return;
    const KC<T> kc = kc_of<T>(c);
    const KC<double> kc64 = make_kc64(c);
    const double scale = (double)kop_scale<KID, double>(kc64);
    if (a.wave_items || a.block_pos) {  // small leaves one wave per item, the rest one workgroup each
        if (a.n_wave > 0) {
            const dim3 g((a.n_wave + kWaveBlock / 64 - 1) / (kWaveBlock / 64)), b(kWaveBlock);
            const bool f64 = fast && std::is_same<T, double>::value;
            if (adj) {
                if (f64) hipLaunchKernelGGL((k_leaf_wave<KID, T, true, true>), g, b, 0, s, a, kc, kc64, scale);
                else hipLaunchKernelGGL((k_leaf_wave<KID, T, true, false>), g, b, 0, s, a, kc, kc64, scale);
            } else {
                if (f64) hipLaunchKernelGGL((k_leaf_wave<KID, T, false, true>), g, b, 0, s, a, kc, kc64, scale);
                else hipLaunchKernelGGL((k_leaf_wave<KID, T, false, false>), g, b, 0, s, a, kc, kc64, scale);
            }
            TBEM_LAUNCH_CHECK();
            if (a.n_parts > 0) {
                hipLaunchKernelGGL(k_leaf_parts, dim3(a.n_parts), dim3(kWaveLeafMax), 0, s, a,
                                   kernel_dim(KID), scale);
                TBEM_LAUNCH_CHECK();
            }
        }
        if (a.n_block > 0) {
            LeafStage bl = a;
            bl.order = a.block_pos;
            bl.n_order = a.n_block;
            bl.wave_items = nullptr;
            bl.block_pos = nullptr;
            bl.n_wave = bl.n_parts = 0;
            bl.sy_order = nullptr;
            leaf_launch<KID, T>(adj, fast, bl, c, s);
        }
        if (a.sy_order && a.rp_ptr) {  // the pairs left out of the rp_* lists, then their reverse sums
            TBEM_REQUIRE((KID == K_EU || KID == K_LS) && fast, TBEM_E_INVALID,
                         "symmetric pairs: Kelvin U and the Laplace single layer only");
            if (a.n_sy_order > 0) {
                if constexpr (KID == K_EU || KID == K_LS)
                    hipLaunchKernelGGL((k_leaf_sym<KID, T>), dim3(a.n_sy_order), dim3(kSymBlock), 0, s, a, kc,
                                       scale);
                TBEM_LAUNCH_CHECK();
            }
            if (a.n_sg > 0) {
                hipLaunchKernelGGL(k_sym_gather, dim3(a.n_sg), dim3(256), 0, s, a, kernel_dim(KID), scale);
                TBEM_LAUNCH_CHECK();
            }
        }
        return;
    }
    dim3 g(a.order ? a.n_order : a.n_leaves), b(leaf_block<KID, T>());
    // the one-step rsqrt only differs from the full one in FP64
    if (fast && std::is_same<T, double>::value && a.rp_close && a.rp_ptr) {
        if (adj) hipLaunchKernelGGL((k_leaf_stage<KID, T, true, true, true>), g, b, 0, s, a, kc, kc64, scale);
        else hipLaunchKernelGGL((k_leaf_stage<KID, T, false, true, true>), g, b, 0, s, a, kc, kc64, scale);
    } else if (std::is_same<T, float>::value && a.rp_close && a.rp_ptr) {
        if (adj) hipLaunchKernelGGL((k_leaf_stage<KID, T, true, false, true>), g, b, 0, s, a, kc, kc64, scale);
        else hipLaunchKernelGGL((k_leaf_stage<KID, T, false, false, true>), g, b, 0, s, a, kc, kc64, scale);
    } else if (fast && std::is_same<T, double>::value) {
        if (adj) hipLaunchKernelGGL((k_leaf_stage<KID, T, true, true>), g, b, 0, s, a, kc, kc64, scale);
        else hipLaunchKernelGGL((k_leaf_stage<KID, T, false, true>), g, b, 0, s, a, kc, kc64, scale);
    } else {
        if (adj) hipLaunchKernelGGL((k_leaf_stage<KID, T, true, false>), g, b, 0, s, a, kc, kc64, scale);
        else hipLaunchKernelGGL((k_leaf_stage<KID, T, false, false>), g, b, 0, s, a, kc, kc64, scale);
    }
    TBEM_LAUNCH_CHECK();
}

template <int KID, class T>
static void grad_launch(bool fp32_points, const LeafStage& a, const KConst& c, hipStream_t s) {
    if (a.n_leaves == 0) return;
    const double scale = (double)kop_scale<KID, double>(make_kc64(c));
    if constexpr (has_analytic_grad(KID)) {
        if (a.wave_items || a.block_pos) {  // the forward's split: small leaves and chunks on waves
            if (a.n_wave > 0) {
                const dim3 g((a.n_wave + kWaveBlock / 64 - 1) / (kWaveBlock / 64)), b(kWaveBlock);
                if (fp32_points)
                    hipLaunchKernelGGL((k_leaf_grad2_wave<KID, float>), g, b, 0, s, a, make_kc64(c), make_kc32(c),
                                       scale);
                else
                    hipLaunchKernelGGL((k_leaf_grad2_wave<KID, double>), g, b, 0, s, a, make_kc64(c), make_kc64(c),
                                       scale);
                TBEM_LAUNCH_CHECK();
                if (a.n_parts > 0) {
                    hipLaunchKernelGGL(k_leaf_grad_parts, dim3(a.n_parts), dim3(kWaveLeafMax), 0, s, a, scale);
                    TBEM_LAUNCH_CHECK();
                }
            }
            if (a.n_block == 0) return;
            LeafStage bl = a;
            bl.order = a.block_pos;
            bl.n_order = a.n_block;
            bl.wave_items = nullptr;
            bl.block_pos = nullptr;
            bl.n_wave = bl.n_parts = 0;
            grad_launch<KID, T>(fp32_points, bl, c, s);
            return;
        }
        const dim3 g(a.order ? a.n_order : a.n_leaves);
        if (fp32_points)
            hipLaunchKernelGGL((k_leaf_grad2<KID, float>), g, dim3(kGradBlock), 0, s, a, make_kc64(c),
                               make_kc32(c), scale);
        else
            hipLaunchKernelGGL((k_leaf_grad2<KID, double>), g, dim3(kGradBlock), 0, s, a, make_kc64(c),
                               make_kc64(c), scale);
    } else {
        hipLaunchKernelGGL((k_leaf_grad<KID>), dim3(a.n_leaves), dim3(kGradBlock), 0, s, a, make_kcd(c), scale);
    }
    TBEM_LAUNCH_CHECK();
}

#define TBEM_KID_SWITCH(kid, F, ...)                                                  \
    switch (kid) {                                                                    \
    case K_LS: F<K_LS, T>(__VA_ARGS__); break;                                        \
    case K_LD: F<K_LD, T>(__VA_ARGS__); break;                                        \
    case K_LH: F<K_LH, T>(__VA_ARGS__); break;                                        \
    case K_EU: F<K_EU, T>(__VA_ARGS__); break;                                        \
    case K_ET: F<K_ET, T>(__VA_ARGS__); break;                                        \
    case K_EA: F<K_EA, T>(__VA_ARGS__); break;                                        \
    case K_EH: F<K_EH, T>(__VA_ARGS__); break;                                        \
    case K_ID1: F<K_ID1, T>(__VA_ARGS__); break;                                      \
    case K_ID3: F<K_ID3, T>(__VA_ARGS__); break;                                      \
    default: throw Error(TBEM_E_INVALID, "unknown kernel id");                        \
    }

void launch_surf_stage(int kid, bool src_surf, bool add, bool adj, const SurfStage& a, const KConst& kc,
                       hipStream_t s) {
    using T = double;
    TBEM_KID_SWITCH(kid, surf_launch, src_surf, add, adj, a, kc, s);
}

template <class T>
void launch_leaf_stage(int kid, bool adj, const LeafStage& a, const KConst& kc, hipStream_t s, bool fast) {
    TBEM_KID_SWITCH(kid, leaf_launch, adj, fast, a, kc, s);
}

void launch_leaf_grad(int kid, const LeafStage& a, const KConst& kc, hipStream_t s, bool fp32_points) {
    using T = double;
    TBEM_KID_SWITCH(kid, grad_launch, fp32_points, a, kc, s);
}

template void launch_leaf_stage<double>(int, bool, const LeafStage&, const KConst&, hipStream_t, bool);
template void launch_leaf_stage<float>(int, bool, const LeafStage&, const KConst&, hipStream_t, bool);

}  // namespace tbem
