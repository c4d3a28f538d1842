// Green's function kernels of 3bem on gfx950.
//
// Two forms of every kernel:
//  * tensor_ref<KID>(): the full R x R block in FP64 with the reference's
//    association order (cpp/laplace_kernels.h:18-64, cpp/elastic_kernels.h:191-365,
//    cpp/identity_kernels.h:9-39, cutoff cpp/kernel.h:19-33). Callers compile it
//    with FP contraction off, so each entry is the reference's value bit for bit
//    when sqrt and division are correctly rounded (they are, in HIP FP64). Used
//    where the matrix itself matters: the check-to-equivalent systems, the
//    per-offset M2L/M2M/L2L operators and the kernel parity entry point.
//  * kop<KID,ADJ,T>(): the block applied to a density vector,
//    acc += K v or acc += K^T v, contracted by hand so that the hot loops
//    (P2P and the point<->surface stages) do ~25 flops per pair instead of
//    forming 9 entries. T is double (FP64 path) or float (mixed path).
//    The r^2 < 1e-12 cutoff is applied by zeroing 1/r, which every term of
//    these kernels carries as a positive power.
#pragma once
#include <type_traits>

#include <hip/hip_runtime.h>
#include <cmath>

namespace tbem {

enum KernelId {
    K_LS = 0, K_LD = 1, K_LH = 2, K_EU = 3, K_ET = 4, K_EA = 5, K_EH = 6, K_ID1 = 7, K_ID3 = 8
};

constexpr double kPi = 3.14159265358979323846;

struct KConst {
    double mu, nu;
    double disp_C1, disp_C2, trac_C1, trac_C2, hyp_C1, hyp_C2, hyp_C3;
};

// Constants exactly as the reference constructors form them
// (cpp/elastic_kernels.h:199-204, 238-243, 280-285, 322-329).
inline KConst make_kconst(double mu, double nu) {
    KConst c;
    c.mu = mu;
    c.nu = nu;
    c.disp_C1 = 1.0 / (16 * kPi * mu * (1 - nu));
    c.disp_C2 = 3 - 4 * nu;
    c.trac_C1 = 1.0 / (8 * kPi * (1 - nu));
    c.trac_C2 = 1 - 2 * nu;
    c.hyp_C1 = mu / (4 * kPi * (1 - nu));
    c.hyp_C2 = -1 + 4 * nu;
    c.hyp_C3 = 3 * nu;
    return c;
}

__host__ __device__ constexpr int kernel_dim(int kid) {
    return (kid == K_LS || kid == K_LD || kid == K_LH || kid == K_ID1) ? 1 : 3;
}
__host__ __device__ constexpr bool kernel_uses_nsrc(int kid) {
    return kid == K_LD || kid == K_ET || kid == K_EH;
}
__host__ __device__ constexpr bool kernel_uses_nobs(int kid) {
    return kid == K_EA || kid == K_EH;
}
inline bool kernel_valid(int kid) { return kid >= 0 && kid <= 8; }

// ------------------------------------------------------------------ tensor_ref
// d = src - obs; r2 = (d0*d0 + d1*d1) + d2*d2 exactly as dot_product/sum
// (cpp/geometry.h:51-54, cpp/vec_ops.h:248-250).
template <int KID>
__host__ __device__ inline void tensor_ref(const double* obs, const double* src, const double* nobs,
                                           const double* nsrc, const KConst& c, double* out) {
#pragma clang fp contract(off)
    const double d[3] = {src[0] - obs[0], src[1] - obs[1], src[2] - obs[2]};
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    constexpr int R = kernel_dim(KID);
    if (r2 < 1e-12) {
        for (int i = 0; i < R * R; i++) out[i] = 0.0;
        return;
    }
    if (KID == K_LS) {
        out[0] = 1.0 / (4.0 * kPi * sqrt(r2));
    } else if (KID == K_LD) {
        const double dn = nsrc[0] * d[0] + nsrc[1] * d[1] + nsrc[2] * d[2];
        out[0] = dn / (4 * kPi * r2 * sqrt(r2));
    } else if (KID == K_LH) {
        out[0] = 0.0;
    } else if (KID == K_ID1) {
        out[0] = 1.0;
    } else if (KID == K_ID3) {
        for (int k = 0; k < 3; k++)
            for (int j = 0; j < 3; j++) out[3 * k + j] = (k == j) ? 1.0 : 0.0;
    } else if (KID == K_EU) {
        const double r = sqrt(r2);
        for (int k = 0; k < 3; k++)
            for (int j = 0; j < 3; j++)
                out[3 * k + j] = (c.disp_C1 / r) * (c.disp_C2 * (k == j ? 1.0 : 0.0) + d[k] * d[j] / r2);
    } else if (KID == K_ET) {
        const double r = sqrt(r2);
        const double drdn = (d[0] * nsrc[0] + d[1] * nsrc[1] + d[2] * nsrc[2]) / r;
        for (int k = 0; k < 3; k++)
            for (int j = 0; j < 3; j++) {
                const double t1 = c.trac_C2 * (k == j ? 1.0 : 0.0) + 3 * d[k] * d[j] / r2;
                const double t2 = c.trac_C2 * (nsrc[j] * d[k] - nsrc[k] * d[j]) / r;
                out[3 * k + j] = -(c.trac_C1 / r2) * (t1 * drdn - t2);
            }
    } else if (KID == K_EA) {
        const double r = sqrt(r2);
        const double drdm = (d[0] * nobs[0] + d[1] * nobs[1] + d[2] * nobs[2]) / r;
        for (int k = 0; k < 3; k++)
            for (int j = 0; j < 3; j++) {
                const double t1 = c.trac_C2 * (k == j ? 1.0 : 0.0) + 3 * d[k] * d[j] / r2;
                const double t2 = c.trac_C2 * (nobs[j] * d[k] - nobs[k] * d[j]) / r;
                out[3 * k + j] = (c.trac_C1 / r2) * (t1 * drdm + t2);
            }
    } else if (KID == K_EH) {
        const double r = sqrt(r2);
        const double inv = 1 / r;  // Vec / s multiplies by 1/s (cpp/vec_ops.h:150-153)
        const double dr[3] = {d[0] * inv, d[1] * inv, d[2] * inv};
        const double drdm = (d[0] * nobs[0] + d[1] * nobs[1] + d[2] * nobs[2]) / r;
        const double drdn = dr[0] * nsrc[0] + dr[1] * nsrc[1] + dr[2] * nsrc[2];
        const double nn = nsrc[0] * nobs[0] + nsrc[1] * nobs[1] + nsrc[2] * nobs[2];
        for (int k = 0; k < 3; k++)
            for (int j = 0; j < 3; j++) {
                const double kr = (k == j) ? 1.0 : 0.0;
                const double l1 = 3 * drdn * (c.trac_C2 * nobs[k] * dr[j] +
                                  c.nu * (nobs[j] * dr[k] + kr * drdm) -
                                  5 * dr[k] * dr[j] * drdm);
                const double l2 = c.trac_C2 * (3 * nsrc[j] * dr[k] * drdm + kr * nn + nsrc[k] * nobs[j]);
                const double l3 = c.hyp_C3 * (nsrc[k] * dr[j] * drdm + nn * dr[k] * dr[j]);
                const double l4 = c.hyp_C2 * nsrc[j] * nobs[k];
                const double C = c.hyp_C1 / (r2 * r);
                out[3 * k + j] = C * (l1 + l2 + l3 + l4);
            }
    }
}

// runtime dispatch of tensor_ref (host + device)
__host__ __device__ inline void tensor_ref_dispatch(int kid, const double* obs, const double* src,
                                                    const double* nobs, const double* nsrc,
                                                    const KConst& c, double* out) {
    switch (kid) {
    case K_LS: tensor_ref<K_LS>(obs, src, nobs, nsrc, c, out); break;
    case K_LD: tensor_ref<K_LD>(obs, src, nobs, nsrc, c, out); break;
    case K_LH: tensor_ref<K_LH>(obs, src, nobs, nsrc, c, out); break;
    case K_EU: tensor_ref<K_EU>(obs, src, nobs, nsrc, c, out); break;
    case K_ET: tensor_ref<K_ET>(obs, src, nobs, nsrc, c, out); break;
    case K_EA: tensor_ref<K_EA>(obs, src, nobs, nsrc, c, out); break;
    case K_EH: tensor_ref<K_EH>(obs, src, nobs, nsrc, c, out); break;
    case K_ID1: tensor_ref<K_ID1>(obs, src, nobs, nsrc, c, out); break;
    default: tensor_ref<K_ID3>(obs, src, nobs, nsrc, c, out); break;
    }
}

// --------------------------------------------------------------- fast contraction
template <class T>
struct KC {
    T ls, ld;                   // 1/(4 pi)
    T u1, u2;                   // disp C1, C2
    T t1, t2;                   // trac C1, C2
    T h1, h2, h3, nu;           // hyp C1, C2, C3, nu
};

inline KC<double> make_kc64(const KConst& c) {
    KC<double> k;
    k.ls = 1.0 / (4.0 * kPi);
    k.ld = 1.0 / (4.0 * kPi);
    k.u1 = c.disp_C1; k.u2 = c.disp_C2;
    k.t1 = c.trac_C1; k.t2 = c.trac_C2;
    k.h1 = c.hyp_C1; k.h2 = c.hyp_C2; k.h3 = c.hyp_C3; k.nu = c.nu;
    return k;
}
inline KC<float> make_kc32(const KConst& c) {
    KC<double> d = make_kc64(c);
    KC<float> k;
    k.ls = (float)d.ls; k.ld = (float)d.ld;
    k.u1 = (float)d.u1; k.u2 = (float)d.u2;
    k.t1 = (float)d.t1; k.t2 = (float)d.t2;
    k.h1 = (float)d.h1; k.h2 = (float)d.h2; k.h3 = (float)d.h3; k.nu = (float)d.nu;
    return k;
}

__device__ inline double rsqrt_t(double x) {
    // v_rsq_f64 (~2^-22 relative) refined by two Newton-Raphson steps
    double y = __builtin_amdgcn_rsq(x);
    double h = x * y;
    double e = fma(-h, y, 1.0);
    y = fma(0.5 * y, e, y);
    h = x * y;
    e = fma(-h, y, 1.0);
    y = fma(0.5 * y, e, y);
    return y;
}
__device__ inline float rsqrt_t(float x) { return __builtin_amdgcn_rsqf(x); }

// One Newton step: relative error <= 1.5 * (2^-22)^2 ~ 9e-14, one-sided. Used for
// point-to-point pairs only, whose densities are of natural size; check potentials
// (amplified by the C2E) and equivalent-density evaluations keep rsqrt_t.
__device__ inline double rsqrt_fast(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = x * y;
    const double e = fma(-h, y, 1.0);
    return fma(0.5 * y, e, y);
}
__device__ inline float rsqrt_fast(float x) { return __builtin_amdgcn_rsqf(x); }

template <class T> __device__ inline T cutoff_r2();
template <> __device__ inline double cutoff_r2<double>() { return 1e-12; }
template <> __device__ inline float cutoff_r2<float>() { return 1e-12f; }


// ------------------------------------------------------- forward-mode dual numbers
// Value plus its gradient with respect to the three target coordinates; running
// kop<> on D3 gives d(K v)/d(obs) for the observation-coordinate gradient.
struct D3 {
    double v, d0, d1, d2;
    __host__ __device__ D3() : v(0), d0(0), d1(0), d2(0) {}
    __host__ __device__ D3(double x) : v(x), d0(0), d1(0), d2(0) {}
    __host__ __device__ D3(double x, double a, double b, double c) : v(x), d0(a), d1(b), d2(c) {}
};
__device__ inline D3 operator+(const D3& a, const D3& b) { return D3(a.v + b.v, a.d0 + b.d0, a.d1 + b.d1, a.d2 + b.d2); }
__device__ inline D3 operator-(const D3& a, const D3& b) { return D3(a.v - b.v, a.d0 - b.d0, a.d1 - b.d1, a.d2 - b.d2); }
__device__ inline D3 operator-(const D3& a) { return D3(-a.v, -a.d0, -a.d1, -a.d2); }
__device__ inline D3 operator*(const D3& a, const D3& b) {
    return D3(a.v * b.v, a.d0 * b.v + a.v * b.d0, a.d1 * b.v + a.v * b.d1, a.d2 * b.v + a.v * b.d2);
}
__device__ inline D3& operator+=(D3& a, const D3& b) { a = a + b; return a; }
__device__ inline bool operator<(const D3& a, const D3& b) { return a.v < b.v; }
__device__ inline D3 rsqrt_t(const D3& x) {
    const double y = rsqrt_t(x.v);
    const double dy = -0.5 * y * y * y;
    return D3(y, dy * x.d0, dy * x.d1, dy * x.d2);
}
template <> __device__ inline D3 cutoff_r2<D3>() { return D3(1e-12); }

__host__ __device__ inline KC<double> make_kc64_dev(const KC<D3>& d) {
    KC<double> k;
    k.ls = d.ls.v; k.ld = d.ld.v; k.u1 = d.u1.v; k.u2 = d.u2.v; k.t1 = d.t1.v; k.t2 = d.t2.v;
    k.h1 = d.h1.v; k.h2 = d.h2.v; k.h3 = d.h3.v; k.nu = d.nu.v;
    return k;
}

inline KC<D3> make_kcd(const KConst& c) {
    KC<double> d = make_kc64(c);
    KC<D3> k;
    k.ls = d.ls; k.ld = d.ld; k.u1 = d.u1; k.u2 = d.u2; k.t1 = d.t1; k.t2 = d.t2;
    k.h1 = d.h1; k.h2 = d.h2; k.h3 = d.h3; k.nu = d.nu;
    return k;
}

// acc[R] += K(obs <- src) v   (ADJ = false)
// acc[R] += K(obs <- src)^T v (ADJ = true)
// d = src - obs, r2 = |d|^2; no / ns are the obs and src normals.
// CUT = false leaves out the r2 < 1e-12 test, for callers that have shown that no
// pair they pass is that close (the result is then the same bit for bit).
template <int KID, bool ADJ, class T, bool FAST = false, bool CUT = true>
__device__ __forceinline__ void kop(T dx, T dy, T dz, T r2, const T* no, const T* ns, const T* v,
                                    T* acc, const KC<T>& k) {
    if (KID == K_LH) return;
    if (KID == K_ID1 || KID == K_ID3) {
        const T m = (CUT && r2 < cutoff_r2<T>()) ? T(0) : T(1);
        acc[0] += m * v[0];
        if (KID == K_ID3) {
            acc[1] += m * v[1];
            acc[2] += m * v[2];
        }
        return;
    }
    T ri;
    if constexpr (FAST) ri = rsqrt_fast(r2);
    else ri = rsqrt_t(r2);
    if constexpr (CUT) ri = (r2 < cutoff_r2<T>()) ? T(0) : ri;
    if (KID == K_LS) {
        acc[0] += ri * v[0];             // times 1/(4 pi) at the end
        return;
    }
    if (KID == K_LD) {
        // (ns . d) / r^3 is the same for K and K^T (scalar)
        const T dn = ns[0] * dx + ns[1] * dy + ns[2] * dz;
        acc[0] += dn * (ri * ri * ri) * v[0];
        return;
    }
    const T ri2 = ri * ri;
    const T dv = dx * v[0] + dy * v[1] + dz * v[2];
    if (KID == K_EU) {  // symmetric: fwd == adj
        const T t = dv * ri2;
        acc[0] += ri * (k.u2 * v[0] + dx * t);
        acc[1] += ri * (k.u2 * v[1] + dy * t);
        acc[2] += ri * (k.u2 * v[2] + dz * t);
        return;
    }
    const T ri3 = ri2 * ri;
    if (KID == K_ET || KID == K_EA) {
        const T* n = (KID == K_ET) ? ns : no;
        const T dn = dx * n[0] + dy * n[1] + dz * n[2];
        const T nv = n[0] * v[0] + n[1] * v[1] + n[2] * v[2];
        // K_ET:  -C1 ri3 [ (C2 v_k + 3 d_k dv ri2) dn - C2 (d_k nv - n_k dv) ]        (fwd)
        //        -C1 ri3 [ (C2 v_j + 3 d_j dv ri2) dn - C2 (n_j dv - d_j nv) ]        (adj)
        // K_EA:   C1 ri3 [ (C2 v_k + 3 d_k dv ri2) dn + C2 (d_k nv - n_k dv) ]        (fwd)
        //         C1 ri3 [ (C2 v_j + 3 d_j dv ri2) dn + C2 (n_j dv - d_j nv) ]        (adj)
        // collected per vector: acc += cv v + cd d + cn n (C1 applied at the end)
        const T s = (KID == K_ET) ? -ri3 : ri3;
        constexpr int sgsa = ((KID == K_ET) ? -1 : 1) * (ADJ ? -1 : 1);
        const T st2 = s * k.t2;
        const T cv = st2 * dn;
        const T cd = s * (T(3) * dv * ri2 * dn) + T(sgsa) * (st2 * nv);
        const T cn = T(-sgsa) * (st2 * dv);
        acc[0] += cv * v[0] + cd * dx + cn * n[0];
        acc[1] += cv * v[1] + cd * dy + cn * n[1];
        acc[2] += cv * v[2] + cd * dz + cn * n[2];
        return;
    }
    if (KID == K_EH) {
        // r-hat = d ri; all terms are C1 ri3 times a combination of r-hat, the
        // observation normal m, the density v and the source normal n, collected
        // per vector: acc += ri3 (ca ri d + cm m + cv v + cn n) (C1 at the end)
        const T drdm = (dx * no[0] + dy * no[1] + dz * no[2]) * ri;
        const T drdn = (dx * ns[0] + dy * ns[1] + dz * ns[2]) * ri;
        const T nn = ns[0] * no[0] + ns[1] * no[1] + ns[2] * no[2];
        const T drv = dv * ri;
        const T nov = no[0] * v[0] + no[1] * v[1] + no[2] * v[2];
        const T nsv = ns[0] * v[0] + ns[1] * v[1] + ns[2] * v[2];
        const T d3 = T(3) * drdn;
        const T pd = drv * drdm;
        T ca, cm, cv, cn;
        if (!ADJ) {
            // d3 (t2 m drv + nu (r nov + v drdm) - 5 r drv drdm)
            //   + t2 (3 nsv r drdm + v nn + n nov) + h3 (n drv drdm + nn r drv) + h2 nsv m
            ca = d3 * (k.nu * nov - T(5) * pd) + T(3) * k.t2 * nsv * drdm + k.h3 * nn * drv;
            cm = d3 * k.t2 * drv + k.h2 * nsv;
            cv = d3 * k.nu * drdm + k.t2 * nn;
            cn = k.t2 * nov + k.h3 * pd;
        } else {
            // d3 (t2 nov r + nu (m drv + v drdm) - 5 drv r drdm)
            //   + t2 (3 n drv drdm + v nn + nsv m) + h3 (nsv r drdm + nn drv r) + h2 n nov
            ca = d3 * (k.t2 * nov - T(5) * pd) + k.h3 * (nsv * drdm + nn * drv);
            cm = d3 * k.nu * drv + k.t2 * nsv;
            cv = d3 * k.nu * drdm + k.t2 * nn;
            cn = T(3) * k.t2 * pd + k.h2 * nov;
        }
        const T ca4 = ri3 * ri * ca, cm4 = ri3 * cm, cv4 = ri3 * cv, cn4 = ri3 * cn;
        acc[0] += ca4 * dx + cm4 * no[0] + cv4 * v[0] + cn4 * ns[0];
        acc[1] += ca4 * dy + cm4 * no[1] + cv4 * v[1] + cn4 * ns[1];
        acc[2] += ca4 * dz + cm4 * no[2] + cv4 * v[2] + cn4 * ns[2];
        return;
    }
}


// ------------------------------------------------- analytic observation gradients
// G[m] += sum_k g_k d(K v)_k / d obs_m for kernels with a closed form
// (d = src - obs, so d d / d obs = -I and d (1/r) / d obs_m = d_m / r^3).
// The constant factor kop_scale<KID>() is applied by the caller. G has
// grad_slots<KID, T>() entries: for K_EU in FP64 the term -g_m (d.v) / r^3 is summed
// as the scalar G[3] and applied once per target by the caller (G[m] -= g_m G[3]),
// which saves two FMAs per pair (in FP32 the extra compensated sum costs more than
// that). FAST: the one-Newton-step FP64 rsqrt (as the forward point pairs of the FMM).
__host__ __device__ constexpr bool has_analytic_grad(int kid) {
    return kid == K_LS || kid == K_LD || kid == K_EU || kid == K_ET;
}
template <int KID, class T>
__host__ __device__ constexpr int grad_slots() {
    return (KID == K_EU && std::is_same<T, double>::value) ? 4 : 3;
}

template <int KID, class T, bool FAST = false, bool CUT = true>
__device__ __forceinline__ void grad_pair(T dx, T dy, T dz, T r2, const T* ns, const T* v, const T* g, T* G,
                                          const KC<T>& k) {
    T ri;
    if constexpr (FAST) ri = rsqrt_fast(r2);
    else ri = rsqrt_t(r2);
    if constexpr (CUT) ri = (r2 < cutoff_r2<T>()) ? T(0) : ri;
    const T ri2 = ri * ri, ri3 = ri2 * ri;
    const T d[3] = {dx, dy, dz};
    if (KID == K_LS) {
        // u = v / r
        const T s = g[0] * v[0] * ri3;
        for (int m = 0; m < 3; m++) G[m] += s * d[m];
    } else if (KID == K_LD) {
        // u = (n.d) v / r^3
        const T b = ns[0] * dx + ns[1] * dy + ns[2] * dz;
        const T s = g[0] * v[0] * ri3;
        for (int m = 0; m < 3; m++) G[m] += s * (T(3) * b * d[m] * ri2 - ns[m]);
    } else if (KID == K_EU) {
        // u_k = C2 v_k / r + d_k (d.v) / r^3
        const T a = dx * v[0] + dy * v[1] + dz * v[2];
        const T gd = g[0] * dx + g[1] * dy + g[2] * dz;
        const T gv = g[0] * v[0] + g[1] * v[1] + g[2] * v[2];
        const T c1 = k.u2 * gv + T(3) * gd * a * ri2;
        if constexpr (grad_slots<KID, T>() == 4) {
            const T p = ri3 * c1, w = ri3 * gd;
            for (int m = 0; m < 3; m++) G[m] += p * d[m] - w * v[m];
            G[3] += ri3 * a;  // times -g_m, per target
        } else {
            for (int m = 0; m < 3; m++) G[m] += ri3 * (c1 * d[m] - g[m] * a - gd * v[m]);
        }
    } else if (KID == K_ET) {
        // u_k = -[C2 v_k b / r^3 + 3 d_k a b / r^5 - C2 d_k c / r^3 + C2 n_k a / r^3]
        const T a = dx * v[0] + dy * v[1] + dz * v[2];
        const T b = dx * ns[0] + dy * ns[1] + dz * ns[2];
        const T c = ns[0] * v[0] + ns[1] * v[1] + ns[2] * v[2];
        const T gv = g[0] * v[0] + g[1] * v[1] + g[2] * v[2];
        const T gn = g[0] * ns[0] + g[1] * ns[1] + g[2] * ns[2];
        const T gd = g[0] * dx + g[1] * dy + g[2] * dz;
        const T ri5 = ri3 * ri2, ri7 = ri5 * ri2;
        const T cd = T(3) * k.t2 * gv * b * ri5 + T(15) * gd * a * b * ri7 - T(3) * k.t2 * c * gd * ri5 +
                     T(3) * k.t2 * gn * a * ri5;
        const T cn = -k.t2 * gv * ri3 - T(3) * gd * a * ri5;
        const T cg = T(-3) * a * b * ri5 + k.t2 * c * ri3;
        const T cv = T(-3) * gd * b * ri5 - k.t2 * gn * ri3;
        for (int m = 0; m < 3; m++) G[m] -= cd * d[m] + cn * ns[m] + cg * g[m] + cv * v[m];
    }
}

// constant factor applied once after accumulation
template <int KID, class T>
__host__ __device__ inline T kop_scale(const KC<T>& k) {
    switch (KID) {
    case K_LS: return k.ls;
    case K_LD: return k.ld;
    case K_EU: return k.u1;
    case K_ET: return k.t1;
    case K_EA: return k.t1;
    case K_EH: return k.h1;
    default: return T(1);
    }
}

}  // namespace tbem
