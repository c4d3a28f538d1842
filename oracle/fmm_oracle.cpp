// ORACLE TEST INFRASTRUCTURE — not product code.
//
// A from-scratch FP64 CPU restatement of 3bem's kernel-independent FMM matvec,
// written for checking the HIP path. Only tests/, tests/golden/, the smoke check
// and bench.py's cpu_baseline leg may load it; the product never links it.
//
// What it restates (reference file:line):
//   kernels + r^2 < 1e-12 cutoff ... cpp/kernel.h:19-33, cpp/laplace_kernels.h:18-64,
//                                     cpp/elastic_kernels.h:191-365, cpp/identity_kernels.h:9-39
//   bounding box / subcell / child id cpp/geometry.cpp:12-95, cpp/octree.cpp:66-76
//   octree build .................... cpp/octree.cpp:95-180
//   translation surfaces ............ cpp/fmm.h:19-66, cpp/fmm.cpp:22-51
//   check-to-equivalent pinv ........ cpp/fmm.cpp:112-160, cpp/blas_wrapper.cpp:67-127
//   stages P2M..L2P, mat_vec ........ cpp/fmm.cpp:57-69, 162-441
//   execute order / apply ........... cpp/fmm.cpp:444-558
//   traversals (MAC, splitting) ..... cpp/fmm.cpp:560-648
//   direct sum ...................... cpp/nbody_operator.h:117-141
// Floating-point expressions keep the reference's association order and the
// file is compiled with -ffp-contract=off, so with one OpenMP thread the result
// matches the compiled reference to rounding of identical operation sequences.
// LAPACK: dgesvd/dgemm from OpenBLAS 0.3.28 as bundled by scipy 1.15.3.

#include <chrono>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>
#include <array>
#include <memory>
#include <algorithm>

extern "C" void scipy_dgesvd_(char* JOBU, char* JOBVT, int* M, int* N, double* A, int* LDA,
                              double* S, double* U, int* LDU, double* VT, int* LDVT,
                              double* WORK, int* LWORK, int* INFO);
extern "C" void scipy_dgemm_(char* TA, char* TB, int* M, int* N, int* K, double* alpha,
                             double* A, int* LDA, double* B, int* LDB, double* beta,
                             double* C, int* LDC);

namespace orc {

typedef std::array<double, 3> V3;

enum { K_LS = 0, K_LD = 1, K_LH = 2, K_EU = 3, K_ET = 4, K_EA = 5, K_EH = 6, K_ID1 = 7, K_ID3 = 8 };

static inline double dot3(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline double norm3(const V3& a) { return std::sqrt(dot3(a, a)); }

// ---------------------------------------------------------------- kernels
struct Kern {
    int id;
    int R;
    double mu, nu;
    // derived constants, same formulas as the reference constructors
    double disp_C1, disp_C2, trac_C1, trac_C2, hyp_C1, hyp_C2, hyp_C3;

    Kern(int id_, double mu_, double nu_): id(id_), mu(mu_), nu(nu_) {
        R = (id == K_LS || id == K_LD || id == K_LH || id == K_ID1) ? 1 : 3;
        disp_C1 = 1.0 / (16 * M_PI * mu * (1 - nu));
        disp_C2 = 3 - 4 * nu;
        trac_C1 = 1.0 / (8 * M_PI * (1 - nu));
        trac_C2 = 1 - 2 * nu;
        hyp_C1 = mu / (4 * M_PI * (1 - nu));
        hyp_C2 = -1 + 4 * nu;
        hyp_C3 = 3 * nu;
    }

    // K(obs <- src) as an R x R row-major block (cpp/kernel.h:19-33)
    void eval(const V3& obs, const V3& src, const V3& nobs, const V3& nsrc, double* out) const {
        V3 d = {src[0] - obs[0], src[1] - obs[1], src[2] - obs[2]};
        double r2 = dot3(d, d);
        int n = R * R;
        if (r2 < 1e-12) {
            for (int i = 0; i < n; i++) out[i] = 0.0;
            return;
        }
        static const double kr[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        switch (id) {
        case K_LS:
            out[0] = 1.0 / (4.0 * M_PI * std::sqrt(r2));
            return;
        case K_LD:
            out[0] = dot3(nsrc, d) / (4 * M_PI * r2 * std::sqrt(r2));
            return;
        case K_LH:
            out[0] = 0.0;
            return;
        case K_ID1:
            out[0] = 1.0;
            return;
        case K_ID3:
            for (int k = 0; k < 3; k++)
                for (int j = 0; j < 3; j++) out[3 * k + j] = kr[k][j];
            return;
        case K_EU: {
            double r = std::sqrt(r2);
            for (int k = 0; k < 3; k++)
                for (int j = 0; j < 3; j++)
                    out[3 * k + j] = (disp_C1 / r) * (disp_C2 * kr[k][j] + d[k] * d[j] / r2);
            return;
        }
        case K_ET: {
            double r = std::sqrt(r2);
            double drdn = dot3(d, nsrc) / r;
            for (int k = 0; k < 3; k++)
                for (int j = 0; j < 3; j++) {
                    double t1 = trac_C2 * kr[k][j] + 3 * d[k] * d[j] / r2;
                    double t2 = trac_C2 * (nsrc[j] * d[k] - nsrc[k] * d[j]) / r;
                    out[3 * k + j] = -(trac_C1 / r2) * (t1 * drdn - t2);
                }
            return;
        }
        case K_EA: {
            double r = std::sqrt(r2);
            double drdm = dot3(d, nobs) / r;
            for (int k = 0; k < 3; k++)
                for (int j = 0; j < 3; j++) {
                    double t1 = trac_C2 * kr[k][j] + 3 * d[k] * d[j] / r2;
                    double t2 = trac_C2 * (nobs[j] * d[k] - nobs[k] * d[j]) / r;
                    out[3 * k + j] = (trac_C1 / r2) * (t1 * drdm + t2);
                }
            return;
        }
        case K_EH: {
            double r = std::sqrt(r2);
            double inv = 1 / r;  // the reference divides a Vec by multiplying with 1/s
            V3 dr = {d[0] * inv, d[1] * inv, d[2] * inv};
            double drdm = dot3(d, nobs) / r;
            double drdn = dot3(dr, nsrc);
            double nn = dot3(nsrc, nobs);
            for (int k = 0; k < 3; k++)
                for (int j = 0; j < 3; j++) {
                    double l1 = 3 * drdn * (trac_C2 * nobs[k] * dr[j] +
                                nu * (nobs[j] * dr[k] + kr[k][j] * drdm) -
                                5 * dr[k] * dr[j] * drdm);
                    double l2 = trac_C2 * (3 * nsrc[j] * dr[k] * drdm + kr[k][j] * nn +
                                nsrc[k] * nobs[j]);
                    double l3 = hyp_C3 * (nsrc[k] * dr[j] * drdm + nn * dr[k] * dr[j]);
                    double l4 = hyp_C2 * nsrc[j] * nobs[k];
                    double C = hyp_C1 / (r2 * r);
                    out[3 * k + j] = C * (l1 + l2 + l3 + l4);
                }
            return;
        }
        }
    }
};

// --------------------------------------------------- point sets and direct sum
struct Points {
    std::vector<V3> locs, normals;
    std::vector<double> weights;  // only for sources
};

// out[d1*n_obs + i] += sum_j (w_j K_ij)[d1][d2] x[d2*n_src + j]  (cpp/nbody_operator.h:117-141)
static std::vector<double> nbody_eval(const Kern& K, const std::vector<V3>& ol, const std::vector<V3>& on,
                                      const std::vector<V3>& sl, const std::vector<V3>& sn,
                                      const std::vector<double>& sw, const double* x) {
    int R = K.R;
    size_t no = ol.size(), ns = sl.size();
    std::vector<double> out(R * no, 0.0);
    double kv[9];
    for (size_t i = 0; i < no; i++) {
        for (size_t j = 0; j < ns; j++) {
            K.eval(ol[i], sl[j], on[i], sn[j], kv);
            for (int q = 0; q < R * R; q++) kv[q] = sw[j] * kv[q];
            for (int d1 = 0; d1 < R; d1++) {
                size_t row = d1 * no + i;
                for (int d2 = 0; d2 < R; d2++) {
                    out[row] += kv[d1 * R + d2] * x[d2 * ns + j];
                }
            }
        }
    }
    return out;
}

// square row-major matvec with a sequential inner sum (cpp/fmm.cpp:57-69)
static std::vector<double> mat_vec(const std::vector<double>& A, const std::vector<double>& x) {
    size_t n = x.size();
    std::vector<double> out(n, 0.0);
    for (size_t i = 0; i < n; i++)
        for (size_t j = 0; j < n; j++) out[i] += A[i * n + j] * x[j];
    return out;
}

// ----------------------------------------------------------------- octree
struct Box {
    V3 c, hw;
};

struct Cell {
    Box b;
    int level;
    int child[8];
    std::vector<size_t> idx;  // ascending original indices
    bool leaf() const {
        for (int i = 0; i < 8; i++)
            if (child[i] >= 0) return false;
        return true;
    }
};

struct Tree {
    std::vector<Cell> cells;  // cells[0] is the root; children referenced by position
    int root = 0;
};

static Box bounding_box(const std::vector<V3>& p, const std::vector<size_t>& idx) {
    if (idx.empty()) return Box{{0, 0, 0}, {0, 0, 0}};
    V3 lo = p[idx[0]], hi = p[idx[0]];
    for (size_t k = 1; k < idx.size(); k++) {
        const V3& q = p[idx[k]];
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], q[d]);
            hi[d] = std::max(hi[d], q[d]);
        }
    }
    Box b;
    for (int d = 0; d < 3; d++) {
        b.c[d] = (hi[d] + lo[d]) * 0.5;   // Vec / 2.0 is a multiply by 1/2.0
        b.hw[d] = (hi[d] - lo[d]) * 0.5;
    }
    return b;
}

static int containing_child(const Box& b, const V3& p) {
    return ((p[0] > b.c[0]) << 2) | ((p[1] > b.c[1]) << 1) | (p[2] > b.c[2]);
}

static Box subcell(const Box& b, int child) {
    Box out;
    for (int d = 0; d < 3; d++) {
        int bit = (child >> (2 - d)) & 1;
        out.hw[d] = b.hw[d] * 0.5;
        out.c[d] = b.c[d] + ((double)bit * 2 - 1) * out.hw[d];
    }
    return out;
}

static int add_cell(Tree& t, const Box& b, int level, std::vector<size_t>&& idx) {
    Cell c;
    c.b = b;
    c.level = level;
    for (int i = 0; i < 8; i++) c.child[i] = -1;
    c.idx = std::move(idx);
    t.cells.push_back(std::move(c));
    return (int)t.cells.size() - 1;
}

static void split(Tree& t, int cell, const Box& split_box, const std::vector<V3>& p, size_t leaf_size) {
    if (t.cells[cell].idx.size() <= leaf_size) return;
    std::vector<size_t> parts[8];
    for (size_t i: t.cells[cell].idx) parts[containing_child(split_box, p[i])].push_back(i);
    int level = t.cells[cell].level + 1;
    for (int ch = 0; ch < 8; ch++) {
        if (parts[ch].empty()) continue;
        Box cb = subcell(split_box, ch);
        Box tight = bounding_box(p, parts[ch]);
        bool zero_extent = tight.hw[0] == 0 && tight.hw[1] == 0 && tight.hw[2] == 0;
        int id = add_cell(t, cb, level, std::move(parts[ch]));
        t.cells[cell].child[ch] = id;
        if (!zero_extent) split(t, id, cb, p, leaf_size);
    }
}

// root bounds = tight box expanded by 1e-5 * max half width; children split the tight box
static Tree build_tree(const std::vector<V3>& p, size_t leaf_size) {
    Tree t;
    std::vector<size_t> all(p.size());
    for (size_t i = 0; i < p.size(); i++) all[i] = i;
    Box tight = bounding_box(p, all);
    double m = std::max(tight.hw[0], std::max(tight.hw[1], tight.hw[2]));
    Box root{tight.c, {tight.hw[0] + 1e-5 * m, tight.hw[1] + 1e-5 * m, tight.hw[2] + 1e-5 * m}};
    add_cell(t, root, 0, std::move(all));
    split(t, 0, tight, p, leaf_size);
    return t;
}

// ---------------------------------------------------------------- surfaces
static std::vector<V3> sphere(size_t order) {
    std::vector<V3> pts;
    double a = 4 * M_PI / order;
    double d = std::sqrt(a);
    size_t Mt = (size_t)std::round(M_PI / d);
    double dt = M_PI / Mt;
    double dp = a / dt;
    for (size_t m = 0; m < Mt; m++) {
        double th = M_PI * (m + 0.5) / Mt;
        size_t Mp = (size_t)std::round(2 * M_PI * std::sin(th) / dp);
        for (size_t n = 0; n < Mp; n++) {
            double ph = 2 * M_PI * n / Mp;
            pts.push_back({std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th)});
        }
    }
    return pts;
}

struct Surface {
    std::vector<V3> pts;      // unit sphere scaled by r_ref
    std::vector<V3> normals;  // unit sphere points
    std::vector<V3> move(const Box& b) const {
        double r = norm3(b.hw);
        std::vector<V3> out(pts.size());
        for (size_t i = 0; i < pts.size(); i++)
            for (int d = 0; d < 3; d++) out[i][d] = r * pts[i][d] + b.c[d];
        return out;
    }
};

static Surface make_surface(size_t order, double r_ref) {
    Surface s;
    s.normals = sphere(order);
    s.pts.resize(s.normals.size());
    for (size_t i = 0; i < s.pts.size(); i++)
        for (int d = 0; d < 3; d++) s.pts[i][d] = r_ref * s.normals[i][d];
    return s;
}

// ---------------------------------------------------- thresholded pseudo-inverse
// LAPACK sees the row-major matrix as its transpose; the pinv assembly below
// follows cpp/blas_wrapper.cpp:67-127 so the row-major result is pinv(A).
static std::vector<double> pinv(const std::vector<double>& M, int n, double threshold) {
    std::vector<double> A = M, s(n), U((size_t)n * n), VT((size_t)n * n);
    char job = 'A';
    int lwork = 5 * n * n, info = 0;
    std::vector<double> work(std::max(lwork, 1));
    scipy_dgesvd_(&job, &job, &n, &n, A.data(), &n, s.data(), U.data(), &n, VT.data(), &n,
                  work.data(), &lwork, &info);
    if (info != 0) {
        std::fprintf(stderr, "oracle: dgesvd info=%d\n", info);
        std::abort();
    }
    double cutoff = threshold * s[0];
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) {
            double& u = U[(size_t)i * n + j];
            u = s[i] > cutoff ? u / s[i] : 0.0;
        }
    }
    std::vector<double> out((size_t)n * n);
    char tt = 'T';
    double one = 1, zero = 0;
    scipy_dgemm_(&tt, &tt, &n, &n, &n, &one, VT.data(), &n, U.data(), &n, &zero, out.data(), &n);
    return out;
}

// ------------------------------------------------------------------- FMM
struct Pair { int obs, src; };

struct FMM {
    Kern K;
    Points obs, src;
    double mac, d;
    size_t order, leaf;
    bool small_cells;
    Surface up_equiv, up_check, down_check, down_equiv;
    Tree st, ot;
    std::vector<std::vector<double>> uc2e, dc2e;
    std::vector<int> p2m, m2m, l2l, l2p;           // cell positions
    std::vector<Pair> p2p, m2p, p2l, m2l;
    std::vector<int> src_slot, obs_slot;           // cell position -> coefficient slot

    size_t ne() const { return up_equiv.pts.size(); }

    FMM(const Kern& k, Points o, Points s, double mac_, size_t order_, size_t leaf_, double d_, bool sc)
        : K(k), obs(std::move(o)), src(std::move(s)), mac(mac_), d(d_), order(order_), leaf(leaf_),
          small_cells(sc) {
        double r_far = 4.0 - std::sqrt(2) - 2 * d;
        up_equiv = make_surface(order, 0.3);
        up_check = make_surface(order, r_far);
        down_check = make_surface(order, 0.3);
        down_equiv = make_surface(order, r_far);
        st = build_tree(src.locs, leaf);
        ot = build_tree(obs.locs, leaf);
        build_c2e(st, up_equiv, up_check, uc2e);
        build_c2e(ot, down_equiv, down_check, dc2e);
        up(0);
        dual(0, 0);
        down(0);
        src_slot = slots(st);
        obs_slot = slots(ot);
    }

    static std::vector<int> slots(const Tree& t) {
        std::vector<int> s(t.cells.size());
        for (size_t i = 0; i < s.size(); i++) s[i] = (int)i;
        return s;
    }

    std::vector<double> c2e_matrix(const Box& b, const Surface& eq, const Surface& ck) const {
        auto cp = ck.move(b), ep = eq.move(b);
        int n = (int)(K.R * eq.pts.size());
        std::vector<double> M((size_t)n * n);
        size_t m = eq.pts.size();
        double kv[9];
        for (size_t i = 0; i < m; i++)
            for (size_t j = 0; j < m; j++) {
                K.eval(cp[i], ep[j], ck.normals[i], eq.normals[j], kv);
                for (int d1 = 0; d1 < K.R; d1++)
                    for (int d2 = 0; d2 < K.R; d2++)
                        M[(d1 * m + i) * n + d2 * m + j] = 1.0 * kv[d1 * K.R + d2];
            }
        return pinv(M, n, 1e-10);
    }

    // one matrix per level, from the first depth-first cell reaching that level
    void build_c2e(const Tree& t, const Surface& eq, const Surface& ck, std::vector<std::vector<double>>& out) {
        std::vector<int> stack{0};
        while (!stack.empty()) {
            int c = stack.back();
            stack.pop_back();
            const Cell& cell = t.cells[c];
            if ((size_t)cell.level == out.size()) out.push_back(c2e_matrix(cell.b, eq, ck));
            for (int ch = 7; ch >= 0; ch--)
                if (cell.child[ch] >= 0) stack.push_back(cell.child[ch]);
        }
    }

    void up(int c) {
        const Cell& cell = st.cells[c];
        if (cell.leaf()) {
            p2m.push_back(c);
            return;
        }
        for (int ch = 0; ch < 8; ch++)
            if (cell.child[ch] >= 0) up(cell.child[ch]);
        m2m.push_back(c);
    }

    void down(int c) {
        const Cell& cell = ot.cells[c];
        if (cell.leaf()) {
            l2p.push_back(c);
            return;
        }
        l2l.push_back(c);
        for (int ch = 0; ch < 8; ch++)
            if (cell.child[ch] >= 0) down(cell.child[ch]);
    }

    void dual(int o, int s) {
        const Cell& oc = ot.cells[o];
        const Cell& sc = st.cells[s];
        double rs = norm3(sc.b.hw), ro = norm3(oc.b.hw);
        double rmax = std::max(rs, ro), rmin = std::min(rs, ro);
        V3 dc = {oc.b.c[0] - sc.b.c[0], oc.b.c[1] - sc.b.c[1], oc.b.c[2] - sc.b.c[2]};
        double sep = norm3(dc);
        if (rmax + mac * rmin <= mac * sep) {
            bool small_src = sc.idx.size() <= down_equiv.pts.size();
            bool small_obs = oc.idx.size() <= up_equiv.pts.size();
            if (small_cells && small_obs && small_src) p2p.push_back({o, s});
            else if (small_cells && small_obs) m2p.push_back({o, s});
            else if (small_cells && small_src) p2l.push_back({o, s});
            else m2l.push_back({o, s});
            return;
        }
        if (oc.leaf() && sc.leaf()) {
            p2p.push_back({o, s});
            return;
        }
        bool split_src = (oc.level > sc.level && !sc.leaf()) || oc.leaf();
        if (split_src) {
            for (int ch = 0; ch < 8; ch++)
                if (sc.child[ch] >= 0) dual(o, sc.child[ch]);
        } else {
            for (int ch = 0; ch < 8; ch++)
                if (oc.child[ch] >= 0) dual(oc.child[ch], s);
        }
    }

    // gather the source points of a cell with their densities (component-major)
    void gather_src(const Cell& c, const double* x, std::vector<V3>& l, std::vector<V3>& n,
                    std::vector<double>& w, std::vector<double>& str) const {
        size_t m = c.idx.size(), N = src.locs.size();
        l.resize(m); n.resize(m); w.resize(m); str.resize(K.R * m);
        for (size_t i = 0; i < m; i++) {
            size_t g = c.idx[i];
            l[i] = src.locs[g]; n[i] = src.normals[g]; w[i] = src.weights[g];
            for (int q = 0; q < K.R; q++) str[q * m + i] = x[q * N + g];
        }
    }

    void gather_obs(const Cell& c, std::vector<V3>& l, std::vector<V3>& n) const {
        size_t m = c.idx.size();
        l.resize(m); n.resize(m);
        for (size_t i = 0; i < m; i++) {
            l[i] = obs.locs[c.idx[i]];
            n[i] = obs.normals[c.idx[i]];
        }
    }

    void scatter_obs(const Cell& c, const std::vector<double>& res, double* out) const {
        size_t m = c.idx.size(), N = obs.locs.size();
        for (size_t i = 0; i < m; i++)
            for (int q = 0; q < K.R; q++) {
                #pragma omp atomic
                out[q * N + c.idx[i]] += res[q * m + i];
            }
    }

    std::vector<double> ones(size_t m) const { return std::vector<double>(m, 1.0); }

    void apply(const double* x, double* out) const {
        size_t n_e = ne(), R = K.R, blk = n_e * R;
        std::vector<double> mult(st.cells.size() * blk, 0.0), loc(ot.cells.size() * blk, 0.0);
        size_t NO = obs.locs.size();
        for (size_t i = 0; i < R * NO; i++) out[i] = 0.0;

        #pragma omp parallel for schedule(dynamic)
        for (size_t t = 0; t < p2m.size(); t++) {  // P2M: cpp/fmm.cpp:162-202
            const Cell& c = st.cells[p2m[t]];
            std::vector<V3> l, n; std::vector<double> w, str;
            gather_src(c, x, l, n, w, str);
            auto chk = nbody_eval(K, up_check.move(c.b), up_check.normals, l, n, w, str.data());
            auto eq = mat_vec(uc2e[c.level], chk);
            std::copy(eq.begin(), eq.end(), mult.begin() + src_slot[p2m[t]] * blk);
        }
        for (size_t t = 0; t < m2m.size(); t++) {  // M2M: cpp/fmm.cpp:204-251
            const Cell& c = st.cells[m2m[t]];
            std::vector<V3> l, n; std::vector<double> str;
            int nch = 0;
            for (int ch = 0; ch < 8; ch++) nch += c.child[ch] >= 0;
            size_t m = nch * n_e;
            l.resize(m); n.resize(m); str.resize(R * m);
            int k = 0;
            for (int ch = 0; ch < 8; ch++) {
                if (c.child[ch] < 0) continue;
                auto ep = up_equiv.move(st.cells[c.child[ch]].b);
                const double* cm = &mult[src_slot[c.child[ch]] * blk];
                for (size_t i = 0; i < n_e; i++) {
                    size_t idx = k * n_e + i;
                    l[idx] = ep[i]; n[idx] = up_equiv.normals[i];
                    for (size_t q = 0; q < R; q++) str[q * m + idx] = cm[q * n_e + i];
                }
                k++;
            }
            auto chk = nbody_eval(K, up_check.move(c.b), up_check.normals, l, n, ones(m), str.data());
            auto eq = mat_vec(uc2e[c.level], chk);
            std::copy(eq.begin(), eq.end(), mult.begin() + src_slot[m2m[t]] * blk);
        }
        #pragma omp parallel for schedule(dynamic)
        for (size_t t = 0; t < p2p.size(); t++) {  // P2P: cpp/fmm.cpp:254-294
            const Cell& o = ot.cells[p2p[t].obs];
            const Cell& s = st.cells[p2p[t].src];
            std::vector<V3> sl, sn, ol, on; std::vector<double> w, str;
            gather_src(s, x, sl, sn, w, str);
            gather_obs(o, ol, on);
            scatter_obs(o, nbody_eval(K, ol, on, sl, sn, w, str.data()), out);
        }
        #pragma omp parallel for schedule(dynamic)
        for (size_t t = 0; t < m2p.size(); t++) {  // M2P: cpp/fmm.cpp:332-360
            const Cell& o = ot.cells[m2p[t].obs];
            const Cell& s = st.cells[m2p[t].src];
            std::vector<V3> ol, on;
            gather_obs(o, ol, on);
            auto res = nbody_eval(K, ol, on, up_equiv.move(s.b), up_equiv.normals, ones(n_e),
                                  &mult[src_slot[m2p[t].src] * blk]);
            scatter_obs(o, res, out);
        }
        #pragma omp parallel for schedule(dynamic)
        for (size_t t = 0; t < p2l.size(); t++) {  // P2L: cpp/fmm.cpp:296-330
            const Cell& o = ot.cells[p2l[t].obs];
            const Cell& s = st.cells[p2l[t].src];
            std::vector<V3> l, n; std::vector<double> w, str;
            gather_src(s, x, l, n, w, str);
            auto chk = nbody_eval(K, down_check.move(o.b), down_check.normals, l, n, w, str.data());
            auto eq = mat_vec(dc2e[o.level], chk);
            double* dst = &loc[obs_slot[p2l[t].obs] * blk];
            for (size_t i = 0; i < eq.size(); i++) {
                #pragma omp atomic
                dst[i] += eq[i];
            }
        }
        #pragma omp parallel for schedule(dynamic)
        for (size_t t = 0; t < m2l.size(); t++) {  // M2L: cpp/fmm.cpp:362-382
            const Cell& o = ot.cells[m2l[t].obs];
            const Cell& s = st.cells[m2l[t].src];
            auto chk = nbody_eval(K, down_check.move(o.b), down_check.normals, up_equiv.move(s.b),
                                  up_equiv.normals, ones(n_e), &mult[src_slot[m2l[t].src] * blk]);
            auto eq = mat_vec(dc2e[o.level], chk);
            double* dst = &loc[obs_slot[m2l[t].obs] * blk];
            for (size_t i = 0; i < eq.size(); i++) {
                #pragma omp atomic
                dst[i] += eq[i];
            }
        }
        for (size_t t = 0; t < l2l.size(); t++) {  // L2L: cpp/fmm.cpp:384-413
            const Cell& c = ot.cells[l2l[t]];
            auto ep = down_equiv.move(c.b);
            const double* pl = &loc[obs_slot[l2l[t]] * blk];
            for (int ch = 0; ch < 8; ch++) {
                if (c.child[ch] < 0) continue;
                const Cell& cc = ot.cells[c.child[ch]];
                auto chk = nbody_eval(K, down_check.move(cc.b), down_check.normals, ep, down_equiv.normals,
                                      ones(n_e), pl);
                auto eq = mat_vec(dc2e[c.level + 1], chk);
                double* dst = &loc[obs_slot[c.child[ch]] * blk];
                for (size_t i = 0; i < eq.size(); i++) dst[i] += eq[i];
            }
        }
        #pragma omp parallel for schedule(dynamic)
        for (size_t t = 0; t < l2p.size(); t++) {  // L2P: cpp/fmm.cpp:415-441
            const Cell& c = ot.cells[l2p[t]];
            std::vector<V3> ol, on;
            gather_obs(c, ol, on);
            auto res = nbody_eval(K, ol, on, down_equiv.move(c.b), down_equiv.normals, ones(n_e),
                                  &loc[obs_slot[l2p[t]] * blk]);
            scatter_obs(c, res, out);
        }
    }
};

static std::vector<V3> to_v3(const double* p, size_t n) {
    std::vector<V3> v(n);
    for (size_t i = 0; i < n; i++) v[i] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    return v;
}

static Points make_points(const double* l, const double* nrm, const double* w, size_t n) {
    Points p;
    p.locs = to_v3(l, n);
    p.normals = to_v3(nrm, n);
    if (w) p.weights.assign(w, w + n);
    return p;
}

}  // namespace orc

using namespace orc;

extern "C" {

void* orc_fmm_create(int kernel, double mu, double nu, size_t n_obs, const double* obs_locs,
                     const double* obs_normals, size_t n_src, const double* src_locs,
                     const double* src_normals, const double* src_weights, double mac, size_t order,
                     size_t leaf, double d, int small_cells) {
    return new FMM(Kern(kernel, mu, nu), make_points(obs_locs, obs_normals, nullptr, n_obs),
                   make_points(src_locs, src_normals, src_weights, n_src), mac, order, leaf, d,
                   small_cells != 0);
}

void orc_fmm_apply(void* h, const double* x, double* out) { ((FMM*)h)->apply(x, out); }

void orc_fmm_counts(void* h, long* c) {
    FMM* f = (FMM*)h;
    c[0] = (long)f->ne();
    c[1] = (long)f->st.cells.size();
    c[2] = (long)f->ot.cells.size();
    c[3] = (long)f->uc2e.size();
    c[4] = (long)f->dc2e.size();
    c[5] = (long)f->p2m.size();
    c[6] = (long)f->m2m.size();
    c[7] = (long)f->p2p.size();
    c[8] = (long)f->m2p.size();
    c[9] = (long)f->p2l.size();
    c[10] = (long)f->m2l.size();
    c[11] = (long)f->l2l.size();
    c[12] = (long)f->l2p.size();
}

void orc_fmm_c2e(void* h, int down, size_t level, double* out) {
    FMM* f = (FMM*)h;
    auto& v = down ? f->dc2e[level] : f->uc2e[level];
    std::memcpy(out, v.data(), v.size() * sizeof(double));
}

void orc_fmm_pairs(void* h, int which, double* out) {
    FMM* f = (FMM*)h;
    auto& v = which == 0 ? f->p2p : which == 1 ? f->m2l : which == 2 ? f->m2p : f->p2l;
    for (size_t i = 0; i < v.size(); i++) {
        const Cell* cs[2] = {&f->ot.cells[v[i].obs], &f->st.cells[v[i].src]};
        for (int k = 0; k < 2; k++) {
            double* o = out + 14 * i + 7 * k;
            o[0] = cs[k]->level;
            for (int d = 0; d < 3; d++) {
                o[1 + d] = cs[k]->b.c[d];
                o[4 + d] = cs[k]->b.hw[d];
            }
        }
    }
}

void orc_fmm_destroy(void* h) { delete (FMM*)h; }

void orc_direct_eval(int kernel, double mu, double nu, size_t n_obs, const double* obs_locs,
                     const double* obs_normals, size_t n_src, const double* src_locs,
                     const double* src_normals, const double* src_weights, const double* x, double* out) {
    Kern K(kernel, mu, nu);
    auto o = make_points(obs_locs, obs_normals, nullptr, n_obs);
    auto s = make_points(src_locs, src_normals, src_weights, n_src);
    auto y = nbody_eval(K, o.locs, o.normals, s.locs, s.normals, s.weights, x);
    std::memcpy(out, y.data(), y.size() * sizeof(double));
}

// Parallel direct sum over observation points (same per-row summation order as
// nbody_eval); used for FP64 error references at sizes where the serial one is slow.
void orc_direct_eval_par(int kernel, double mu, double nu, size_t n_obs, const double* obs_locs,
                         const double* obs_normals, size_t n_src, const double* src_locs,
                         const double* src_normals, const double* src_weights, const double* x,
                         double* out) {
    Kern K(kernel, mu, nu);
    auto o = make_points(obs_locs, obs_normals, nullptr, n_obs);
    auto s = make_points(src_locs, src_normals, src_weights, n_src);
    int R = K.R;
    #pragma omp parallel for schedule(dynamic, 16)
    for (size_t i = 0; i < n_obs; i++) {
        double acc[3] = {0, 0, 0}, kv[9];
        for (size_t j = 0; j < n_src; j++) {
            K.eval(o.locs[i], s.locs[j], o.normals[i], s.normals[j], kv);
            for (int q = 0; q < R * R; q++) kv[q] = s.weights[j] * kv[q];
            for (int d1 = 0; d1 < R; d1++)
                for (int d2 = 0; d2 < R; d2++) acc[d1] += kv[d1 * R + d2] * x[d2 * n_src + j];
        }
        for (int d1 = 0; d1 < R; d1++) out[d1 * n_obs + i] = acc[d1];
    }
}

// Transposed direct sum: out[d2*n_src + j] = w_j sum_i K(x_i <- y_j)[d1][d2] g[d1*n_obs + i]
void orc_direct_adjoint(int kernel, double mu, double nu, size_t n_obs, const double* obs_locs,
                        const double* obs_normals, size_t n_src, const double* src_locs,
                        const double* src_normals, const double* src_weights, const double* g,
                        double* out) {
    Kern K(kernel, mu, nu);
    auto o = make_points(obs_locs, obs_normals, nullptr, n_obs);
    auto s = make_points(src_locs, src_normals, src_weights, n_src);
    int R = K.R;
    #pragma omp parallel for schedule(dynamic, 16)
    for (size_t j = 0; j < n_src; j++) {
        double acc[3] = {0, 0, 0}, kv[9];
        for (size_t i = 0; i < n_obs; i++) {
            K.eval(o.locs[i], s.locs[j], o.normals[i], s.normals[j], kv);
            for (int d1 = 0; d1 < R; d1++)
                for (int d2 = 0; d2 < R; d2++) acc[d2] += kv[d1 * R + d2] * g[d1 * n_obs + i];
        }
        for (int d2 = 0; d2 < R; d2++) out[d2 * n_src + j] = s.weights[j] * acc[d2];
    }
}

void orc_kernel_eval(int kernel, double mu, double nu, const double* obs, const double* src,
                     const double* nobs, const double* nsrc, double* out) {
    Kern K(kernel, mu, nu);
    K.eval({obs[0], obs[1], obs[2]}, {src[0], src[1], src[2]}, {nobs[0], nobs[1], nobs[2]},
           {nsrc[0], nsrc[1], nsrc[2]}, out);
}

// Octree cells in depth-first pre-order, same 10-double layout as the reference driver.
void orc_octree(const double* pts, size_t n, size_t leaf, long* n_cells, long* n_idx,
                double* cells_out, long* idx_out) {
    auto p = to_v3(pts, n);
    Tree t = build_tree(p, leaf);
    std::vector<double> cells;
    std::vector<long> idx;
    std::vector<int> stack{0};
    while (!stack.empty()) {
        int c = stack.back();
        stack.pop_back();
        const Cell& cell = t.cells[c];
        cells.push_back(cell.level);
        for (int d = 0; d < 3; d++) cells.push_back(cell.b.c[d]);
        for (int d = 0; d < 3; d++) cells.push_back(cell.b.hw[d]);
        cells.push_back(cell.leaf() ? 1.0 : 0.0);
        cells.push_back((double)cell.idx.size());
        cells.push_back((double)idx.size());
        for (auto i: cell.idx) idx.push_back((long)i);
        for (int ch = 7; ch >= 0; ch--)
            if (cell.child[ch] >= 0) stack.push_back(cell.child[ch]);
    }
    *n_cells = (long)(cells.size() / 10);
    *n_idx = (long)idx.size();
    if (cells_out) std::memcpy(cells_out, cells.data(), cells.size() * sizeof(double));
    if (idx_out) std::memcpy(idx_out, idx.data(), idx.size() * sizeof(long));
}

long orc_surface(size_t order, double* out) {
    auto s = sphere(order);
    if (out)
        for (size_t i = 0; i < s.size(); i++)
            for (int d = 0; d < 3; d++) out[3 * i + d] = s[i][d];
    return (long)s.size();
}

// CSR from COO triplets as cpp/sparse_operator.cpp:140-172 builds it: rows
// counted, row pointers as prefix sums, entries placed in input order within their
// row (duplicates kept). With x, the product as cpp/sparse_operator.cpp:19-28
// computes it: y_i = sum over the row's entries in stored order of value * x[col].
void orc_sparse_coo(size_t n_rows, size_t n_cols, size_t n_entries, const long* rows, const long* cols,
                    const double* vals, long* row_ptrs, long* column_indices, double* values, const double* x,
                    double* y) {
    std::vector<long> count(n_rows, 0), fill(n_rows, 0);
    for (size_t i = 0; i < n_entries; i++) count[rows[i]]++;
    row_ptrs[0] = 0;
    for (size_t r = 0; r < n_rows; r++) row_ptrs[r + 1] = row_ptrs[r] + count[r];
    for (size_t i = 0; i < n_entries; i++) {
        const long slot = row_ptrs[rows[i]] + fill[rows[i]]++;
        column_indices[slot] = cols[i];
        values[slot] = vals[i];
    }
    if (!x) return;
    (void)n_cols;
    for (size_t r = 0; r < n_rows; r++) {
        double s = 0.0;
        for (long k = row_ptrs[r]; k < row_ptrs[r + 1]; k++) s += values[k] * x[column_indices[k]];
        y[r] = s;
    }
}

// distribute_vector as cpp/constraint_matrix.cpp:54-82 computes it, for solved
// constraints x_{cdof[i]} = trhs[i] + sum_k tw[k] x_{tdof[k]} (cdof ascending): the
// free dofs take in[] in order, then each constrained dof, ascending, sums its terms
// in order starting from its rhs.
void orc_distribute(long n_c, const long* cdof, const long* tptr, const long* tdof, const double* tw,
                    const double* trhs, const double* in, size_t n_total, double* out) {
    std::vector<long> which(n_total, -1);
    for (long i = 0; i < n_c; i++) which[cdof[i]] = i;
    size_t next = 0;
    for (size_t d = 0; d < n_total; d++)
        if (which[d] < 0) out[d] = in[next++];
    for (size_t d = 0; d < n_total; d++) {
        const long i = which[d];
        if (i < 0) continue;
        double val = trhs[i];
        for (long k = tptr[i]; k < tptr[i + 1]; k++) val += tw[k] * out[tdof[k]];
        out[d] = val;
    }
}

// condense_vector as cpp/constraint_matrix.cpp:99-125 computes it: dofs from the
// last down; a dof's entry is all[d] plus what it has accumulated; a constrained dof
// passes weight * entry to each of its terms' dofs, a free dof keeps 0.0 + entry.
void orc_condense(long n_c, const long* cdof, const long* tptr, const long* tdof, const double* tw,
                  const double* trhs, const double* all, size_t n_total, double* out) {
    (void)trhs;
    std::vector<long> which(n_total, -1);
    for (long i = 0; i < n_c; i++) which[cdof[i]] = i;
    std::vector<double> acc(n_total, 0.0);
    for (size_t d = n_total; d-- > 0;) {
        const double entry = all[d] + acc[d];
        acc[d] = 0.0;
        const long i = which[d];
        if (i < 0) {
            acc[d] += entry;
            continue;
        }
        for (long k = tptr[i]; k < tptr[i + 1]; k++) acc[tdof[k]] += tw[k] * entry;
    }
    size_t next = 0;
    for (size_t d = 0; d < n_total; d++)
        if (which[d] < 0) out[next++] = acc[d];
}

// The CSR product of orc_sparse_coo on ready CSR arrays, reps times; returns the
// mean seconds per apply and leaves y.
double orc_sparse_csr_apply(size_t n_rows, size_t n_cols, const long* rp, const long* ci, const double* v,
                            const double* x, double* y, int reps) {
    (void)n_cols;
    auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < reps; r++)
        for (size_t i = 0; i < n_rows; i++) {
            double s = 0.0;
            for (long k = rp[i]; k < rp[i + 1]; k++) s += v[k] * x[ci[k]];
            y[i] = s;
        }
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count() / reps;
}

// Sparse product A B as cpp/sparse_operator.cpp:114-138 forms it: a dictionary
// keyed by row * n_cols(B) + col, each value 0.0 plus the products b_jc * a_ij in
// traversal order (A's entries in stored order, then B's row j in stored order),
// emitted in key order through csr_from_coo. Writes at most cap entries; returns nnz.
long orc_sparse_right_multiply(size_t a_rows, size_t a_cols, const long* a_rp, const long* a_ci, const double* a_v,
                               size_t b_cols, const long* b_rp, const long* b_ci, const double* b_v,
                               long cap, long* rp, long* ci, double* v) {
    (void)a_cols;
    std::map<size_t, double> dok;
    for (size_t i = 0; i < a_rows; i++)
        for (long ka = a_rp[i]; ka < a_rp[i + 1]; ka++) {
            const long j = a_ci[ka];
            for (long kb = b_rp[j]; kb < b_rp[j + 1]; kb++) dok[i * b_cols + b_ci[kb]] += b_v[kb] * a_v[ka];
        }
    const long nnz = (long)dok.size();
    if (nnz > cap) return nnz;
    std::vector<long> rows, cols;
    std::vector<double> vals;
    for (const auto& e : dok) {
        rows.push_back((long)(e.first / b_cols));
        cols.push_back((long)(e.first % b_cols));
        vals.push_back(e.second);
    }
    orc_sparse_coo(a_rows, b_cols, (size_t)nnz, rows.data(), cols.data(), vals.data(), rp, ci, v, nullptr, nullptr);
    return nnz;
}

void orc_pinv(const double* A, int n, double threshold, double* out) {
    std::vector<double> m(A, A + (size_t)n * n);
    auto p = pinv(m, n, threshold > 0 ? threshold : 1e-16);
    std::memcpy(out, p.data(), p.size() * sizeof(double));
}

}  // extern "C"
