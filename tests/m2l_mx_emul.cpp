// Lane-by-lane CPU emulation of the matrix-core M2L engine (fastmultipole_amd/csrc/m2l_mx.hpp).
// Test infrastructure: the engine's source is compiled here with a backend whose per-lane values are
// arrays of 64 doubles, whose cross-lane moves are permutations of those arrays, and whose MFMA is
// v_mfma_f64_16x16x4f64's product with its lane layout (lane l supplies A(l % 16, l / 16) and
// B(l / 16, l % 16) and holds D(l / 16 + 4 r, l % 16) in element r). tests/test_m2l_mx_emul.py compares
// the result with the oracle's multipole_to_local.
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../fastmultipole_amd/csrc/m2l_mx.hpp"
#include "../fastmultipole_amd/csrc/tables_host.hpp"

namespace {

constexpr int W = 64;

struct CpuB {
    struct M {
        std::array<bool, W> v;
        M() { v.fill(false); }
        M(bool b) { v.fill(b); }
    };
    struct I {
        std::array<int64_t, W> v;
        I() { v.fill(0); }
        I(int64_t x) { v.fill(x); }
    };
    struct V {
        std::array<double, W> v;
        V() { v.fill(0.0); }
        V(double x) { v.fill(x); }
    };
    struct V4 {
        V r[4];
    };
    struct PV {
        std::array<const double*, W> v;
    };

    static I lane() {
        I r;
        for (int i = 0; i < W; ++i) r.v[i] = i;
        return r;
    }
    static V sel(const M& m, const V& a, const V& b) {
        V r;
        for (int i = 0; i < W; ++i) r.v[i] = m.v[i] ? a.v[i] : b.v[i];
        return r;
    }
    static I seli(const M& m, const I& a, const I& b) {
        I r;
        for (int i = 0; i < W; ++i) r.v[i] = m.v[i] ? a.v[i] : b.v[i];
        return r;
    }
    static I opaque(const I& a) { return a; }
    static void sched_fence() {}
    static I mini(const I& a, int b) {
        I r;
        for (int i = 0; i < W; ++i) r.v[i] = a.v[i] < b ? a.v[i] : b;
        return r;
    }
    static V fabs(const V& a) {
        V r;
        for (int i = 0; i < W; ++i) r.v[i] = std::fabs(a.v[i]);
        return r;
    }
    static V sqrt(const V& a) {
        V r;
        for (int i = 0; i < W; ++i) r.v[i] = std::sqrt(a.v[i]);
        return r;
    }
    static V fma(const V& a, const V& b, const V& c) {
        V r;
        for (int i = 0; i < W; ++i) r.v[i] = std::fma(a.v[i], b.v[i], c.v[i]);
        return r;
    }
    static V tod(const I& a) {
        V r;
        for (int i = 0; i < W; ++i) r.v[i] = (double)a.v[i];
        return r;
    }
    static V from(const V& a, int (*src)(int)) {
        V r;
        for (int i = 0; i < W; ++i) r.v[i] = a.v[src(i)];
        return r;
    }
    static V perm1(const V& a) { return from(a, [](int i) { return i ^ 1; }); }
    static V perm2(const V& a) { return from(a, [](int i) { return i ^ 2; }); }
    static V perm3(const V& a) { return from(a, [](int i) { return i ^ 3; }); }
    static V bperm(const V& a, const I& src) {
        V r;
        for (int i = 0; i < W; ++i) r.v[i] = a.v[src.v[i] & 63];
        return r;
    }
    static V quadsum(const V& a) {
        V r;
        for (int i = 0; i < W; ++i) r.v[i] = (a.v[i] + a.v[i ^ 16]) + (a.v[i ^ 32] + a.v[i ^ 48]);
        return r;
    }
    template <int NC>
    static V slotsum(const V& a) {
        V r;
        for (int i = 0; i < W; ++i) {
            const int row = i & ~15, c = i & 15;
            double s = 0.0;
            for (int j = c % NC; j < 16; j += NC) s += a.v[row + j];
            r.v[i] = s;
        }
        return r;
    }
    template <int NC, int j>
    static V qbcast(const V& a) {
        V r;
        for (int i = 0; i < W; ++i) r.v[i] = a.v[(i & ~(NC - 1)) + j];
        return r;
    }
    static V4 zero4() { return V4{}; }
    static V get4(const V4& a, int r) { return a.r[r]; }
    static void set4(V4& a, int r, const V& x) { a.r[r] = x; }
    static void mfma(const V& a, const V& b, V4& d) {
        for (int r = 0; r < 4; ++r)
            for (int l = 0; l < W; ++l) {
                const int i = l / 16 + 4 * r, j = l % 16;
                double s = d.r[r].v[l];
                for (int kk = 0; kk < 4; ++kk) s = std::fma(a.v[16 * kk + i], b.v[16 * kk + j], s);
                d.r[r].v[l] = s;
            }
    }
    static bool any(const M& m) {
        for (int i = 0; i < W; ++i)
            if (m.v[i]) return true;
        return false;
    }
    static bool all(const M& m) {
        for (int i = 0; i < W; ++i)
            if (!m.v[i]) return false;
        return true;
    }
    static int maxi(const I& a, const M& m) {
        int64_t r = INT32_MIN;
        for (int i = 0; i < W; ++i)
            if (m.v[i] && a.v[i] > r) r = a.v[i];
        return (int)r;
    }
    static V gld(const double* p, const I& idx, const M& ok) {
        V r;
        for (int i = 0; i < W; ++i) r.v[i] = ok.v[i] ? p[idx.v[i]] : 0.0;
        return r;
    }
    static I gldi(const int32_t* p, const I& idx, const M& ok) {
        I r;
        for (int i = 0; i < W; ++i) r.v[i] = ok.v[i] ? p[idx.v[i]] : 0;
        return r;
    }
    static PV ptr(const double* base, const I& js, int64_t stride) {
        PV r;
        for (int i = 0; i < W; ++i) r.v[i] = base + js.v[i] * stride;
        return r;
    }
    static V gldp(const PV& p, const I& idx, const M& ok) {
        V r;
        for (int i = 0; i < W; ++i) r.v[i] = ok.v[i] ? p.v[i][idx.v[i]] : 0.0;
        return r;
    }
    static V lds(const double* p, const I& idx) {
        V r;
        for (int i = 0; i < W; ++i) r.v[i] = p[idx.v[i]];
        return r;
    }
    static void lds_st(double* p, const I& idx, const V& x, const M& ok) {
        for (int i = 0; i < W; ++i)
            if (ok.v[i]) p[idx.v[i]] = x.v[i];
    }
    static void lds_add(double* p, const I& idx, const V& x, const M& ok) {
        for (int i = 0; i < W; ++i)
            if (ok.v[i]) p[idx.v[i]] += x.v[i];
    }
    static void sti(int32_t* p, const I& idx, const I& x, const M& ok) {
        for (int i = 0; i < W; ++i)
            if (ok.v[i]) p[idx.v[i]] = (int32_t)x.v[i];
    }
    static void wsync() {}
};

// ---- element-wise operators of the emulated lane types
#define MX_BIN(T, R, op)                                                                     \
    inline R operator op(const T& a, const T& b) {                                           \
        R r;                                                                                 \
        for (int i = 0; i < W; ++i) r.v[i] = a.v[i] op b.v[i];                               \
        return r;                                                                            \
    }
#define MX_BIN_S(T, S, R, op)                                                                \
    inline R operator op(const T& a, S b) {                                                  \
        R r;                                                                                 \
        for (int i = 0; i < W; ++i) r.v[i] = a.v[i] op b;                                    \
        return r;                                                                            \
    }                                                                                        \
    inline R operator op(S a, const T& b) {                                                  \
        R r;                                                                                 \
        for (int i = 0; i < W; ++i) r.v[i] = a op b.v[i];                                    \
        return r;                                                                            \
    }
using V = CpuB::V;
using I = CpuB::I;
using M = CpuB::M;
MX_BIN(V, V, +) MX_BIN(V, V, -) MX_BIN(V, V, *) MX_BIN(V, V, /)
MX_BIN_S(V, double, V, +) MX_BIN_S(V, double, V, -) MX_BIN_S(V, double, V, *) MX_BIN_S(V, double, V, /)
MX_BIN(V, M, <) MX_BIN(V, M, <=) MX_BIN(V, M, >) MX_BIN(V, M, >=)
MX_BIN_S(V, double, M, <) MX_BIN_S(V, double, M, <=) MX_BIN_S(V, double, M, >) MX_BIN_S(V, double, M, >=)
MX_BIN(I, I, +) MX_BIN(I, I, -) MX_BIN(I, I, *) MX_BIN(I, I, /) MX_BIN(I, I, %) MX_BIN(I, I, &)
MX_BIN_S(I, int, I, +) MX_BIN_S(I, int, I, -) MX_BIN_S(I, int, I, *) MX_BIN_S(I, int, I, /) MX_BIN_S(I, int, I, %)
MX_BIN_S(I, int, I, &) MX_BIN_S(I, int, I, >>)
MX_BIN(I, M, <) MX_BIN(I, M, <=) MX_BIN(I, M, >) MX_BIN(I, M, >=) MX_BIN(I, M, ==) MX_BIN(I, M, !=)
MX_BIN_S(I, int, M, <) MX_BIN_S(I, int, M, <=) MX_BIN_S(I, int, M, >) MX_BIN_S(I, int, M, >=) MX_BIN_S(I, int, M, ==)
MX_BIN_S(I, int, M, !=)
MX_BIN(M, M, &&) MX_BIN(M, M, ||) MX_BIN(M, M, !=) MX_BIN(M, M, ==)
inline V operator-(const V& a) {
    V r;
    for (int i = 0; i < W; ++i) r.v[i] = -a.v[i];
    return r;
}
inline M operator!(const M& a) {
    M r;
    for (int i = 0; i < W; ++i) r.v[i] = !a.v[i];
    return r;
}
inline M operator&&(const M& a, bool b) { return a && M(b); }
inline M operator||(const M& a, bool b) { return a || M(b); }

struct HostTabs {
    std::vector<double> A, zin, zoh, ein, eoh, Mt, Lt, tf, i0;
    explicit HostTabs(int cap) {
        using namespace fmmhip::tabh;
        std::vector<double> Hs, z, e;
        host_Hs(Hs, cap);
        host_mags(z, e, cap);
        host_power_norms(Mt, Lt, cap);
        A = host_dense_A(Hs, cap);
        tf = host_test_factorials(cap);
        const int H = nh(cap);
        zin.resize(H); zoh.resize(H); ein.resize(H); eoh.resize(H);
        i0.resize(2 * (cap + 1));
        for (int n = 0; n <= cap; ++n) {
            i0[2 * n] = 1.0 / Mt[nh(n - 1)];
            i0[2 * n + 1] = 1.0 / Lt[nh(n - 1)];
        }
        for (int n = 0, i = 0; n <= cap; ++n)
            for (int m = 0; m <= n; ++m, ++i) {
                const int b = len_Ts(n - 1);
                zin[i] = z[b + m];
                zoh[i] = 0.5 * z[b + m * (n + 1)];
                ein[i] = e[b + m];
                eoh[i] = 0.5 * e[b + m * (n + 1)];
            }
    }
    fmmhip::mx::Tabs view() const {
        return {A.data(), zin.data(), zoh.data(), ein.data(), eoh.data(), Mt.data(), Lt.data(), tf.data(), i0.data()};
    }
};

template <int CAP, bool LH>
int run(int P, int kind, int relative, double eps, int BE, const double* tg, int nsrc, const double* sg,
        const double* sex, int64_t he4, double* out, int32_t* orders, int* pmax, int* ok) {
    using E = fmmhip::mx::Engine<CpuB, CAP, LH>;
    static const HostTabs tabs(CAP);
    const fmmhip::mx::Tabs T = tabs.view();
    std::vector<double> eim(E::BP * E::NS * 2), cs(E::BP * E::NS * 2), acc(E::NC * fmmhip::mx::mnh(CAP), 0.0),
        F(E::BP * E::NF);
    fmmhip::mx::Wave Wv{eim.data(), cs.data(), acc.data(), F.data()};
    fmmhip::mx::Box X;
    for (int d = 0; d < 3; ++d) { X.tc[d] = tg[3 + d]; X.tbox[d] = tg[11 + d]; }
    X.r_l = tg[7];
    const double e2 = relative ? tg[14] * eps : eps;
    X.limit = e2 * 4 * 3.141592653589793;
    X.P = P; X.kind = kind; X.BE = BE;
    // source centres as the branch array (stride 15, source centre at offset 0)
    std::vector<int32_t> ms(nsrc);
    for (int i = 0; i < nsrc; ++i) ms[i] = i;
    int pm = 0, okf = 1;
    for (int b0 = 0; b0 < nsrc; b0 += E::BP)
        E::batch(T, Wv, X, fmmhip::mx::Batch{ms.data(), nullptr, b0, std::min(E::BP, nsrc - b0), orders}, sg, 15, 0, sex, he4,
                 pm, okf);
    const int H = fmmhip::mx::mnh(P);
    for (int i = 0; i < H; ++i)
        for (int c = 0; c < 4; ++c) out[4 * i + c] = c < E::NC ? acc[i * E::NC + c] : 0.0;
    *pmax = pm;
    *ok = okf;
    return 0;
}

}  // namespace

extern "C" int mx_emul_m2l(int LH, int cap, int P, int kind, int relative, double eps, int BE, const double* tgeom,
                           int nsrc, const double* sgeoms, const double* sexp, int64_t he4, double* out, int32_t* orders,
                           int* pmax, int* ok) {
#define MX_CASE(C)                                                                                                   \
    if (cap == C)                                                                                                    \
        return LH ? run<C, true>(P, kind, relative, eps, BE, tgeom, nsrc, sgeoms, sexp, he4, out, orders, pmax, ok)  \
                  : run<C, false>(P, kind, relative, eps, BE, tgeom, nsrc, sgeoms, sexp, he4, out, orders, pmax, ok);
    MX_CASE(8) MX_CASE(20)
#undef MX_CASE
    return -1;
}
