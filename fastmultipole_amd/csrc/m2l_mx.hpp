// Register-resident M2L on the matrix cores (v_mfma_f64_16x16x4f64, gfx950).
//
// multipole_to_local! (src/translate.jl:519-550 fixed order, :555-1412 error-controlled) as rotate,
// translate along z, Lamb-Helmholtz transform, rotate back, for BP interactions of one target box at once
// per wavefront. The wavefront's 16 matrix columns are the real components of the BP interactions
// ("slots"): NC = 4 with the Lamb-Helmholtz potential (phi re, phi im, chi re, chi im), NC = 2 without
// (phi re, phi im), so BP = 16 / NC.
//
// Register layout. A degree-n coefficient vector (rows m = 0..n, one column per component) is split by
// the parity of m: the even part E holds m = 8e + 2q in register e of the lanes of quad q = lane / 16, the
// odd part O holds m = 8o + 2q + 1, column = lane % 16. This is the layout in which the MFMA returns rows
// i = 4r + q of a 16-row result when its rows are m = 2i (E) or m = 2i + 1 (O), and the layout in which
// the next MFMA takes its B operand (lane l supplies B(l / 16, l % 16), so a result register r is the
// K-block of rows 4r..4r+3). Products therefore chain with no data movement, and since a given m sits in
// the same lane and register index for every degree, the z-translation L_j^m += F(j + k) M_k^m is a
// lane-local multiply-add.
//
// Rotations. The y-rotation of degree n is the factorised form of the reference's Wigner matrices
// (update_Ts!, src/rotate.jl:260-441, with the zeta / eta magnitudes of :625-1004):
//   U = g_in(m') X,  S_E(v) = sum_{m' even} A(m', v) U_m',  S_O(v) = sum_{m' odd} A(m', v) U_m'
//   P+ = (S_E - A(0, v) U_0) - S_O + s_v conj(S_E + S_O),  P- = (S_E - A(0, v) U_0) + S_O + s_v conj(S_E - S_O)
//   Q+- = w_v (cos v b +- i sigma sin v b) P+-,  R_e = Q+ + Q-,  R_o = Q+ - Q-
//   out_m = g_out(m) / 2 sum_v A(m, v) R_(parity of m)(v)
// with A the degree-n pi/2 table (symmetric), s_v = (-1)^(n+v), w_0 = 1, w_v = 2. The two sums over m'
// and the two sums over v are MFMA products whose A operand is the shared table (read from LDS with the
// rows or columns of the other parity masked to zero); the combination is per element on the vector ALU
// (the re / im partner of a component is the neighbouring lane, reached by DPP).
//
// z-translation, streamed by degree. The reference forms L_n^m = sum_{k=m}^{P} F(n + k) M_k^m with
// F(j) = j! / t^(j+1) after the whole rotated multipole exists. Here degree k of the multipole is
// rotated, added to L_k (which then holds the degree-k part translate_multipole_to_local_z_n! and
// translate_multipole_to_local_z_m01_n! compute for the error tests), the order tests of degree k run,
// and only if the interaction's order is not settled below k is M_k added to the other degrees L_j.
// F(j) follows the reference's product chain F(j) = F(j - 1) ((j) / t) from F(0) = 1 / t, so every
// factor is the same double the reference uses. The terms of each L_n^m are added in the reference's
// order (ascending k); multiply-adds are fused.
//
// The tests of the error-controlled methods (RotatedCoefficients*, PowerPotential / PowerGradient,
// absolute or relative) use the reference's formulas and running products; the sums over m of the
// power tests are taken as lane reductions, so a test value can differ from the reference's in the last
// bits and an order decision sitting on the tolerance can flip.
//
// The engine is written against a backend B (per-lane value types, cross-lane moves, the MFMA) so the
// same source runs on the GPU (backend in expansions.hip) and lane by lane on the CPU (the emulation
// used by the tests). It is plain C++ apart from the backend.
#pragma once
#include <cstdint>
#include <type_traits>

#if defined(__HIPCC__)
#define MX_FN __device__ __forceinline__
#else
#define MX_FN inline
#endif

namespace fmmhip {
namespace mx {

constexpr int mnh(int p) { return p < 0 ? 0 : ((p + 1) * (p + 2)) >> 1; }
constexpr int KE(int n) { return (n / 2 + 4) / 4; }          // registers of the even part of degree n
constexpr int KO(int n) { return ((n + 1) / 2 + 3) / 4; }    // registers of the odd part
constexpr int KB(int n) { return (n + 4) / 4; }              // 4-row blocks of rows 0..n
constexpr int TE(int n) { return (n + 16) / 16; }            // 16-row tiles of rows 0..n
constexpr int LOFF(int n) { return n <= 0 ? 0 : LOFF(n - 1) + KE(n - 1) + KO(n - 1); }   // degree n in L
constexpr int NLREG(int cap) { return LOFF(cap + 1); }
template <int n>
struct LO {   // LOFF(n) as a constant (register arrays need compile-time indices)
    static constexpr int v = LOFF(n);
};
constexpr int AST(int n) { return (n + 2) & ~1; }            // row length of the dense pi/2 table of degree n
constexpr int ABASE(int n) { return (n * (n + 1) * (2 * n + 1)) / 6 + ((n + 1) / 2) * ((n + 1) / 2); }
constexpr int ALEN(int cap) { return ABASE(cap + 1); }
constexpr int imin(int a, int b) { return a < b ? a : b; }
constexpr int imax(int a, int b) { return a > b ? a : b; }

template <int I0, int I1, class F>
MX_FN void sfor(F&& f) {
    if constexpr (I0 < I1) {
        f(std::integral_constant<int, I0>());
        sfor<I0 + 1, I1>(f);
    }
}

// Shared tables (LDS on the GPU), for degrees 0..CAP; per (n, m) tables are indexed mnh(n - 1) + m.
struct Tabs {
    const double* A;     // dense pi/2 tables: degree n at ABASE(n), row-major, rows of AST(n)
    const double* zin;   // zeta input scale (multipole rotation)
    const double* zoh;   // half the zeta output scale
    const double* ein;   // eta input scale (local rotation)
    const double* eoh;   // half the eta output scale
    const double* Mt;    // multipole power normalisation
    const double* Lt;    // local power normalisation
    const double* tf;    // per degree n: 1/(n-1)!, (n+1)!
    const double* i0;    // per degree n: 1 / Mt(n, 0), 1 / Lt(n, 0)
};

// per-wavefront scratch
struct Wave {
    double* eim;   // BP x (CAP + 1) x 2: e^{i m alpha} of each slot
    double* cs;    // BP x (CAP + 1) x 2: cos / sin (v beta) of each slot
    double* acc;   // mnh(CAP) x NC: the target box's local expansion, summed over the wave's interactions
    double* F;     // BP x (2 CAP + 1): z-translation factors j! / t^(j+1) of each slot
};

// BP consecutive entries b0 .. b0 + nb - 1 of the target box's part of the M2L list, taken in the
// order perm gives (list indices; null: list order)
struct Batch {
    const int32_t* m2l_s;   // source branch per list entry
    const int32_t* perm;    // processing order of the list entries, or null
    int b0, nb;
    int32_t* orders;        // expansion order per list entry (written)
};

// the target box and the method (uniform over the wavefront)
struct Box {
    double tc[3], tbox[3];
    double r_l;     // target radius
    double limit;   // 4 pi eps (eps scaled by max_influence for the relative methods)
    int P;          // expansion order (fixed) or order cap (error-controlled), <= CAP
    int kind;       // 0 fixed order, 1 rotated coefficients (gradient), 2 power potential, 3 power gradient
    int BE;         // the order settles at the passing degree (1) or one below (0)
};

template <class B, int CAP, bool LH, bool FULL = false>
struct Engine {
    using V = typename B::V;
    using I = typename B::I;
    using M = typename B::M;
    using V4 = typename B::V4;
    using PV = typename B::PV;
    static constexpr int NC = LH ? 4 : 2;
    static constexpr int BP = 16 / NC;
    static constexpr int NL = NLREG(CAP);
    static constexpr int NS = CAP + 1;   // slot-table entries
    static constexpr int NF = 2 * CAP + 1;   // translation factors per slot
    static constexpr int NX = KE(CAP) + KO(CAP);   // registers of one degree
    static constexpr double SQ3 = 1.7320508075688772;

    // src/error.jl:5-10
    MX_FN static V md1(V dx, double rm, double rp) {
        const V right = dx + rp, left = dx + rm;
        const M same = right * left >= 0.0;
        const V mn = B::sel(B::fabs(right) <= B::fabs(left), right, left);
        return B::sel(same, mn, V(0.0));
    }

    // list index of slot `slot` of the batch
    MX_FN static I entry(const Batch& Bt, I slot, M ok) {
        const I k = B::seli(ok, slot, I(0)) + Bt.b0;
        return Bt.perm ? B::gldi(Bt.perm, k, ok) : k;
    }

    // translation vector target centre - source centre of source branch js (benign values for empty slots)
    MX_FN static void geometry(const Box& X, I js, M ok, const double* sbr, int br_stride, int sc_off, V& dx, V& dy, V& dz) {
        const I base = js * br_stride + sc_off;
        const V sx = B::gld(sbr, base, ok), sy = B::gld(sbr, base + 1, ok), sz = B::gld(sbr, base + 2, ok);
        dx = B::sel(ok, X.tc[0] - sx, V(0.0));
        dy = B::sel(ok, X.tc[1] - sy, V(0.0));
        dz = B::sel(ok, X.tc[2] - sz, V(1.0));
    }

    // cos / sin of the azimuth and polar angle of (x, y, z) from the Cartesian components
    // (cartesian_to_spherical, src/rotate.jl:5-21, up to rounding); returns r
    MX_FN static V angles(V x, V y, V z, V& cp, V& sp, V& cb, V& sb) {
        const V x2y2 = x * x + y * y;
        const V r = B::sqrt(x2y2 + z * z);
        const M has = x2y2 > 0.0;
        const V rho = B::sqrt(x2y2);
        const V rs = B::sel(has, rho, V(1.0)), rr = B::sel(r > 0.0, r, V(1.0));
        cp = B::sel(has, x / rs, V(1.0));
        sp = B::sel(has, y / rs, V(0.0));
        cb = B::sel(has, z / rr, B::sel(z < 0.0, V(-1.0), V(1.0)));
        sb = B::sel(has, rho / rr, V(0.0));
        return r;
    }

    // Per-slot tables in the wave's LDS: e^{i m alpha} and (cos, sin)(m beta), m = 0..CAP, and the
    // z-translation factors F(j) = j! / t^(j+1), j = 0..2 CAP. Lane l serves slot l / (64 / BP) and
    // entries l % (64 / BP) + h (64 / BP): it forms its first entry by repeated multiplication and each
    // later one by continuing the same recurrence (update_eimphis!, src/rotate.jl:25-36; the
    // angle-addition recurrence of update_Ts!, :380-435; the factor chain of
    // translate_multipole_to_local_z!, src/translate.jl:181-232), so every entry is the double the
    // reference's chain produces.
    MX_FN static void slot_tables(const Box& X, const Wave& W, const Batch& Bt, const double* sbr, int br_stride, int sc_off) {
        constexpr int LPS = 64 / BP;
        B::wsync();   // the previous batch's reads of the tables are done
        const I l = B::lane();
        const I sT = l / LPS, kT = l % LPS;
        const M vT = sT < Bt.nb;
        V dx, dy, dz;
        geometry(X, B::gldi(Bt.m2l_s, entry(Bt, sT, vT), vT), vT, sbr, br_stride, sc_off, dx, dy, dz);
        V cp, sp, cb, sb;
        const V r = angles(dx, dy, dz, cp, sp, cb, sb);
        const V ot = 1.0 / r;
        V er = 1.0, ei = 0.0, c = 1.0, s = 0.0, f = ot;
        auto step = [&](M st, int x) {
            const V t = er;
            const V ner = er * cp - ei * sp;
            const V nei = t * sp + ei * cp;
            const V cn = cb * c - sb * s;
            const V sn = sb * c + cb * s;
            er = B::sel(st, ner, er);
            ei = B::sel(st, nei, ei);
            c = B::sel(st, cn, c);
            s = B::sel(st, sn, s);
            (void)x;
        };
        // F's own chain: F(x) = F(x - 1) (x / t)
        for (int j = 0; j < LPS - 1; ++j) {
            const M st = kT > j;
            step(st, j);
            f = B::sel(st, f * ((double)(j + 1) * ot), f);
        }
        sfor<0, (2 * CAP + LPS) / LPS>([&](auto HH) {
            constexpr int h = decltype(HH)::value;
            if constexpr (h > 0) {
                for (int j = 0; j < LPS; ++j) {
                    if (h * LPS - LPS + j < CAP) step(M(true), j);
                    f = f * (B::tod(kT + (h * LPS - LPS + j + 1)) * ot);
                }
            }
            const I k = kT + h * LPS;
            const M okk = k <= CAP, okf = k <= 2 * CAP;
            const I idx = B::seli(okk, (sT * NS + k) * 2, I(0));
            B::lds_st(W.eim, idx, er, okk);
            B::lds_st(W.eim, idx + 1, ei, okk);
            B::lds_st(W.cs, idx, c, okk);
            B::lds_st(W.cs, idx + 1, s, okk);
            B::lds_st(W.F, B::seli(okf, sT * NF + k, I(0)), f, okf);
        });
        B::wsync();
    }

    // Lane constants of the rotation's combination step
    struct RotLane {
        I c, q, s;
        V sq;    // (-1)^q, negated on imaginary lanes: (-1)^(n+v) conj() for rows v = 16 t + 4 r + q
        V nj;    // +1 on imaginary lanes, -1 on real ones: the i of a complex product
        V w0;    // w_v of row v = q (1 for v = 0, else 2)
    };

    // One rotation of degree n (compile-time) of the coefficient vector u (E / O registers, input scale
    // applied, rows m > n zero): out = g_out / 2 * A . R(A^T . u). sigma = +1 multipole (zeta), -1 local
    // (eta). The A operands are read unmasked: entries outside the degree's table multiply zero rows of u
    // (first products) or rows of R that are set to zero (second products), or land in output rows m > n,
    // which are discarded; every address stays inside the shared tables.
    template <int n>
    MX_FN static void rotate(const Tabs& T, const Wave& W, const double* gout, double sigma, const RotLane& RL,
                             const V (&uE)[imax(KE(n), 1)], const V (&uO)[imax(KO(n), 1)], V (&oE)[imax(KE(n), 1)],
                             V (&oO)[imax(KO(n), 1)]) {
        constexpr int ke = KE(n), ko = KO(n), te = TE(n), kb = KB(n), st = AST(n), ab = ABASE(n);
        const I c = RL.c, q = RL.q, s = RL.s;
        V4 SE[te], SO[te];
        sfor<0, te>([&](auto TT) {
            constexpr int t = decltype(TT)::value;
            SE[t] = B::zero4();
            SO[t] = B::zero4();
        });
        // first products: rows v = 16 t + (lane % 16), K = the even (odd) rows m' of u
        const I lo1 = q * (2 * st) + c;
        sfor<0, te>([&](auto TT) {
            constexpr int t = decltype(TT)::value;
            sfor<0, ke>([&](auto EE) {
                constexpr int e = decltype(EE)::value;
                B::mfma(B::lds(T.A, lo1 + (ab + 8 * e * st + 16 * t)), uE[e], SE[t]);
            });
            sfor<0, ko>([&](auto OO) {
                constexpr int o = decltype(OO)::value;
                B::mfma(B::lds(T.A, lo1 + (ab + (8 * o + 1) * st + 16 * t)), uO[o], SO[t]);
            });
        });
        // combination, per element (rows v of the first products)
        const V U0 = B::bperm(uE[0], c);
        const I sb = s * (NS * 2);
        sfor<0, te>([&](auto TT) {
            constexpr int t = decltype(TT)::value;
            sfor<0, 4>([&](auto RR) {
                constexpr int r = decltype(RR)::value;
                constexpr int v0 = 16 * t + 4 * r;
                if constexpr (v0 <= n) {
                    const I v = q + v0;
                    const V se = B::get4(SE[t], r), so = B::get4(SO[t], r);
                    const V a00 = B::lds(T.A, q + (ab + v0));
                    const I ci = sb + B::mini(v, n) * 2;
                    const V cv = B::lds(W.cs, ci), sv = B::lds(W.cs, ci + 1);
                    V cc, sn;
                    if constexpr (v0 == 0) {
                        cc = RL.w0 * cv;
                        sn = RL.w0 * sigma * sv;
                    } else {
                        cc = 2.0 * cv;
                        sn = (2.0 * sigma) * sv;
                    }
                    const V sgc = (n & 1) ? -RL.sq : RL.sq;
                    const V D = se - a00 * U0;
                    const V Pp = (D - so) + sgc * (se + so);
                    const V Pm = (D + so) + sgc * (se - so);
                    const V snj = sn * RL.nj;
                    const V Qp = cc * Pp + snj * B::perm1(Pp);
                    const V Qm = cc * Pm - snj * B::perm1(Pm);
                    if constexpr (v0 + 3 <= n) {
                        B::set4(SE[t], r, Qp + Qm);   // R_e
                        B::set4(SO[t], r, Qp - Qm);   // R_o
                    } else {
                        const M okv = v <= n;
                        B::set4(SE[t], r, B::sel(okv, Qp + Qm, V(0.0)));
                        B::set4(SO[t], r, B::sel(okv, Qp - Qm, V(0.0)));
                    }
                } else {
                    B::set4(SE[t], r, V(0.0));
                    B::set4(SO[t], r, V(0.0));
                }
            });
        });
        // second products: rows m = 2 (lane % 16) (+1), K = rows v
        V4 PE = B::zero4(), PO = B::zero4();
        const I lo2 = q * st + c * 2;
        sfor<0, kb>([&](auto BB) {
            constexpr int b = decltype(BB)::value;
            B::mfma(B::lds(T.A, lo2 + (ab + 4 * b * st)), B::get4(SE[b / 4], b % 4), PE);
            B::mfma(B::lds(T.A, lo2 + (ab + 4 * b * st + 1)), B::get4(SO[b / 4], b % 4), PO);
        });
        auto out = [&](auto par, auto RI, const V4& Pv, V& o) {
            constexpr int pr = decltype(par)::value, ri = decltype(RI)::value;
            constexpr int m0 = 8 * ri + pr;
            const I m = q * 2 + m0;
            if constexpr (m0 + 6 <= n) {
                o = B::lds(gout, m + mnh(n - 1)) * B::get4(Pv, ri);
            } else {
                const M ok = m <= n;
                o = B::sel(ok, B::lds(gout, B::mini(m, n) + mnh(n - 1)) * B::get4(Pv, ri), V(0.0));
            }
        };
        sfor<0, ke>([&](auto EE) { out(std::integral_constant<int, 0>(), EE, PE, oE[decltype(EE)::value]); });
        sfor<0, ko>([&](auto OO) { out(std::integral_constant<int, 1>(), OO, PO, oO[decltype(OO)::value]); });
    }

    // the source coefficients of degree n for this lane's rows (E registers, then O), zero for rows m > n
    // and for empty slots
    template <int n>
    MX_FN static void fetch(const PV& src, I q, I comp, M valid, V (&xn)[NX]) {
        auto one = [&](auto par, auto RI, V& x) {
            constexpr int pr = decltype(par)::value, ri = decltype(RI)::value;
            constexpr int m0 = 8 * ri + pr;
            const I m = q * 2 + m0;
            const M ok = (m0 + 6 <= n) ? valid : (valid && (m <= n));
            const I mi = (m0 + 6 <= n) ? m : B::mini(m, n);
            x = B::gldp(src, (mi + mnh(n - 1)) * 4 + comp, ok);
        };
        sfor<0, KE(n)>([&](auto EE) { one(std::integral_constant<int, 0>(), EE, xn[decltype(EE)::value]); });
        sfor<0, KO(n)>([&](auto OO) { one(std::integral_constant<int, 1>(), OO, xn[KE(n) + decltype(OO)::value]); });
    }

    // value of component j of this lane's slot, from the same lane quad
    template <int j>
    MX_FN static V comp_of(V x) { return B::template qbcast<NC, j>(x); }

    // One batch of interactions of one target box: their local expansions are added to W.acc and their
    // orders written to Bt.orders. pmax / ok collect the largest order and whether every tolerance was met.
    MX_FN static void batch(const Tabs& T, const Wave& W, const Box& X, const Batch& Bt, const double* sbr, int br_stride,
                            int sc_off, const double* sex, int64_t HE4, int& pmax, int& okflag) {
        slot_tables(X, W, Bt, sbr, br_stride, sc_off);
        const I l0 = B::lane();
        const I c0 = l0 & 15, q0 = l0 >> 4;
        const I c = c0, q = q0, s = c / NC, comp = c % NC, pot = comp >> 1;
        const M isIm = (comp & 1) != 0;
        const M valid = s < Bt.nb;
        const int P = FULL ? CAP : X.P;   // FULL: the launch's order is the engine's cap
        const bool err = X.kind != 0;

        const I li = entry(Bt, s, valid);
        const I js = B::gldi(Bt.m2l_s, li, valid);
        V dx, dy, dz;
        geometry(X, js, valid, sbr, br_stride, sc_off, dx, dy, dz);
        const V r = B::sqrt(dx * dx + dy * dy + dz * dz);
        const V d0 = md1(dx, -X.tbox[0], X.tbox[0]), d1 = md1(dy, -X.tbox[1], X.tbox[1]), d2 = md1(dz, -X.tbox[2], X.tbox[2]);
        const V r_mp = B::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        const V rmi = 1.0 / r_mp;
        const PV src = B::ptr(sex, js, HE4);
        const V cj = B::sel(isIm, V(-1.0), V(1.0));
        const V phi_sign = B::sel(comp == 0, V(-1.0), B::sel(comp == 1, V(1.0), V(0.0)));   // Lamb-Helmholtz phi term
        const M chi_lane = pot == 1;
        RotLane RL0;
        RL0.c = c;
        RL0.q = q;
        RL0.s = s;
        RL0.sq = B::sel((q & 1) != 0, -cj, cj);
        RL0.nj = -cj;
        RL0.w0 = B::sel(q == 0, V(1.0), V(2.0));

        V L[NL];
        sfor<0, NL>([&](auto II) { L[decltype(II)::value] = V(0.0); });
        // degree k's source coefficients are loaded during degree k - 1 (their HBM latency hidden)
        V xn[NX];
        fetch<0>(src, q, comp, valid, xn);
        M active = valid;
        M succ = valid && !err;
        I Pl = I(P);
        V rmi_np2 = rmi * rmi * rmi;  // r_mp^-(n+2) at the test of degree n
        double rl_nm1 = 1.0;          // r_l^(n-1)
        V mp_pow = V(0.0);            // multipole power (phi) of the previous degree
        bool running = true;
        int kdone = -1;

        sfor<0, CAP + 1>([&](auto KK) {
            constexpr int k = decltype(KK)::value;
            if (!running || k > P) return;
            B::sched_fence();   // keep each degree's instructions together (limits register pressure)
            kdone = k;
            constexpr int ke = KE(k), ko = KO(k), ke1 = imax(ke, 1), ko1 = imax(ko, 1);
            // lane indices made opaque per degree: their derived addresses and masks are recomputed here
            // instead of being hoisted (one register each for the whole kernel)
            const I c = B::opaque(c0), q = B::opaque(q0), s = c / NC, comp = c % NC, pot = comp >> 1;
            const I fb = s * NF, sb2 = s * (NS * 2);
            RotLane RL = RL0;
            RL.c = c;
            RL.q = q;
            RL.s = s;
            // ---- degree k of the source multipoles: power terms, z-rotation, input scale
            V uE[ke1], uO[ko1], ME[ke1], MO[ko1];
            V xc[NX];
            sfor<0, ke + ko>([&](auto II) { xc[decltype(II)::value] = xn[decltype(II)::value]; });
            if constexpr (k + 1 <= CAP) {
                if (FULL || k + 1 <= P) fetch<k + 1>(src, q, comp, valid, xn);
            }
            V pw_t = V(0.0);
            const bool power = X.kind >= 2;
            auto load = [&](auto par, auto RI, V& u) {
                constexpr int pr = decltype(par)::value, ri = decltype(RI)::value;
                constexpr int m0 = 8 * ri + pr;
                const I m = q * 2 + m0;
                const I mi = (m0 + 6 <= k) ? m : B::mini(m, k);
                const V x = xc[(pr ? ke : 0) + ri];
                if (power) {
                    const V nt = B::lds(T.Mt, mi + mnh(k - 1));
                    const V x2 = x * x;
                    pw_t = pw_t + (x2 + B::perm1(x2)) * nt * nt;
                }
                const I ei = sb2 + mi * 2;
                const V er = B::lds(W.eim, ei), eimg = B::lds(W.eim, ei + 1);
                const V y = x * er + RL.nj * B::perm1(x) * eimg;
                u = B::lds(T.zin, mi + mnh(k - 1)) * y;
            };
            sfor<0, ke>([&](auto EE) { load(std::integral_constant<int, 0>(), EE, uE[decltype(EE)::value]); });
            sfor<0, ko>([&](auto OO) { load(std::integral_constant<int, 1>(), OO, uO[decltype(OO)::value]); });
            if constexpr (ko == 0) uO[0] = V(0.0);
            V pw_phi = V(0.0), pw_chi = V(0.0);
            if (power) {
                const V t = B::quadsum(pw_t);
                if (LH) {
                    const V o = B::perm2(t);
                    pw_phi = B::sqrt(B::sel(pot == 0, t, o));
                    pw_chi = B::sqrt(B::sel(pot == 1, t, o));
                } else {
                    pw_phi = B::sqrt(t);
                }
            }
            // ---- y-rotation of degree k (rotate_multipole_y!, zeta magnitudes)
            rotate<k>(T, W, T.zoh, 1.0, RL, uE, uO, ME, MO);
            // ---- L_k += F(2k) M_k
            {
                const V fac = B::sel(active, B::lds(W.F, fb + 2 * k), V(0.0));
                sfor<0, ke>([&](auto EE) {
                    constexpr int e = decltype(EE)::value;
                    L[LO<k>::v + e] = B::fma(fac, ME[e], L[LO<k>::v + e]);
                });
                sfor<0, ko>([&](auto OO) {
                    constexpr int o = decltype(OO)::value;
                    L[LO<k>::v + ke + o] = B::fma(fac, MO[o], L[LO<k>::v + ke + o]);
                });
            }
            // ---- order tests of degree k. Uniform factors are formed once per degree from reciprocal
            // tables (reordered products: test values agree with the reference's to a few ulps), and the
            // local part is skipped when no active interaction's multipole part is within the tolerance
            // (the reference evaluates it only then).
            M pass = M(false);
            if constexpr (k >= 1) {
                if (err) {
                    constexpr double ik = 1.0 / k, ik1 = 1.0 / (k + 1);
                    const double nm1_f_inv = T.tf[2 * k], np1_f = T.tf[2 * k + 1];
                    V eps_mp, eps_l = V(0.0);
                    if (X.kind == 1) {
                        // rotated coefficients (src/translate.jl:555-826): m = 0, 1 of the rotated multipole and
                        // of the degree-k local (translate_multipole_to_local_z_m01_n!, :302-362)
                        const V e0 = B::bperm(ME[0], c), e1 = B::bperm(MO[0], c);
                        const V p0 = comp_of<0>(e0), p1r = comp_of<0>(e1), p1i = comp_of<1>(e1);
                        eps_mp = (B::sqrt(p1r * p1r + p1i * p1i) + B::fabs(p0)) * np1_f * rmi_np2;
                        if constexpr (LH) {
                            const V c1r = comp_of<2>(e1), c1i = comp_of<3>(e1);
                            eps_mp = eps_mp + B::sqrt(c1r * c1r + c1i * c1i) * np1_f * rmi_np2 * r_mp;
                        }
                        if (B::any(active && (eps_mp <= X.limit))) {
                            const V l0 = B::bperm(L[LO<k>::v], c), l1 = B::bperm(L[LO<k>::v + ke], c);
                            V q0 = comp_of<0>(l0), q1r = comp_of<0>(l1), q1i = comp_of<1>(l1);
                            const double cl = rl_nm1 * nm1_f_inv;
                            if constexpr (LH) {
                                const V d1r = comp_of<2>(l1), d1i = comp_of<3>(l1);
                                const V ton = r * ik;
                                q1r = q1r - ton * d1i;
                                q1i = q1i + ton * d1r;
                                eps_l = (B::fabs(q1r) + B::fabs(q1i) + B::fabs(q0)) * cl + B::sqrt(d1r * d1r + d1i * d1i) * (cl * X.r_l);
                            } else {
                                eps_l = (B::fabs(q1r) + B::fabs(q1i) + B::fabs(q0)) * cl;
                            }
                        }
                    } else {
                        // powers (src/translate.jl:831-1412): multipole power of degree k (or k - 1), local
                        // power of the degree-k local (local_power, :455-502)
                        const double iM0 = T.i0[2 * (k - 1)], iM1 = T.i0[2 * k], iN = T.i0[2 * k + 1];
                        if (X.kind == 3) {
                            eps_mp = mp_pow * (SQ3 * np1_f * iM0);
                            if (LH) eps_mp = eps_mp + pw_chi * (SQ3 * (double)k * np1_f * iM1);
                            eps_mp = eps_mp * rmi_np2;
                        } else {
                            eps_mp = pw_phi * (SQ3 * np1_f * iM0 * ik1);
                            if (LH) eps_mp = eps_mp + pw_chi * (SQ3 * (double)k * np1_f * iM1 * ik1);
                            eps_mp = eps_mp * rmi_np2 * r_mp;
                        }
                        if (B::any(active && (eps_mp <= X.limit))) {
                            const V r_n = r * ik, r_np1 = r * ik1;
                            V lsum = V(0.0);
                            auto term = [&](auto par, auto RI) {
                                constexpr int pr = decltype(par)::value, ri = decltype(RI)::value;
                                constexpr int m0 = 8 * ri + pr;
                                const I m = q * 2 + m0;
                                const V x = pr ? L[LO<k>::v + ke + ri] : L[LO<k>::v + ri];
                                V val = x;
                                if constexpr (LH) {
                                    const V y = B::perm3(x);
                                    const V rm = r_n * B::tod(m);
                                    // chi of (k, m + 1): the odd register of the same lane for even m; for odd m
                                    // the even part's next entry, one quad up (or register ri + 1 of quad 0)
                                    V nxt;
                                    if constexpr (pr == 0) {
                                        nxt = ri < ko ? L[LO<k>::v + ke + imin(ri, ko1 - 1)] : V(0.0);
                                    } else {
                                        const V hi = ri + 1 < ke ? L[LO<k>::v + imin(ri + 1, ke - 1)] : V(0.0);
                                        const V tmp = B::sel(q == 0, hi, L[LO<k>::v + ri]);
                                        nxt = B::bperm(tmp, ((q + 1) & 3) * 16 + c);
                                    }
                                    if constexpr (m0 + 7 > k) nxt = B::sel(m + 1 <= k, nxt, V(0.0));
                                    val = B::sel(comp == 0, x - y * rm, B::sel(comp == 1, x + y * rm, x - nxt * r_np1));
                                }
                                const V t2 = val * val;
                                if constexpr (m0 + 6 <= k) {
                                    const V nt = B::lds(T.Lt, m + mnh(k - 1));
                                    lsum = lsum + (t2 + B::perm1(t2)) * nt * nt;
                                } else {
                                    const M ok = m <= k;
                                    const V nt = B::lds(T.Lt, B::mini(m, k) + mnh(k - 1));
                                    lsum = lsum + B::sel(ok, (t2 + B::perm1(t2)) * nt * nt, V(0.0));
                                }
                            };
                            sfor<0, ke>([&](auto EE) { term(std::integral_constant<int, 0>(), EE); });
                            sfor<0, ko>([&](auto OO) { term(std::integral_constant<int, 1>(), OO); });
                            const V tl = B::quadsum(lsum);
                            V lp, lc = V(0.0);
                            if constexpr (LH) {
                                const V o = B::perm2(tl);
                                lp = B::sqrt(B::sel(pot == 0, tl, o));
                                lc = B::sqrt(B::sel(pot == 1, tl, o));
                            } else {
                                lp = B::sqrt(tl);
                            }
                            const double cl = SQ3 * rl_nm1 * nm1_f_inv * iN;
                            eps_l = lp * (X.kind == 3 ? cl : cl * X.r_l * ik);
                            if (LH) eps_l = eps_l + lc * (cl * X.r_l);
                        }
                    }
                    pass = (eps_mp <= X.limit) && (eps_mp + eps_l * 2.0 <= X.limit);
                }
            }
            const M newly = active && pass;
            Pl = B::seli(newly, I(k - 1 + X.BE), Pl);
            succ = succ || newly;
            const M addk = active && (!pass || X.BE != 0);
            active = active && !pass;
            // ---- M_k into the other degrees: L_j += F(j + k) M_k
            sfor<0, ke>([&](auto EE) { ME[decltype(EE)::value] = B::sel(addk, ME[decltype(EE)::value], V(0.0)); });
            sfor<0, ko>([&](auto OO) { MO[decltype(OO)::value] = B::sel(addk, MO[decltype(OO)::value], V(0.0)); });
            sfor<0, CAP + 1>([&](auto JJ) {
                constexpr int j = decltype(JJ)::value;
                if constexpr (j != k) {
                    const V fac = B::lds(W.F, fb + (k + j));
                    sfor<0, imin(KE(j), ke)>([&](auto EE) {
                        constexpr int e = decltype(EE)::value;
                        L[LO<j>::v + e] = B::fma(fac, ME[e], L[LO<j>::v + e]);
                    });
                    sfor<0, imin(KO(j), ko)>([&](auto OO) {
                        constexpr int o = decltype(OO)::value;
                        L[LO<j>::v + KE(j) + o] = B::fma(fac, MO[o], L[LO<j>::v + KE(j) + o]);
                    });
                }
            });
            // ---- running products for the next degree
            if constexpr (k >= 1) {
                rmi_np2 = rmi_np2 * rmi;
                rl_nm1 = rl_nm1 * X.r_l;
            }
            mp_pow = pw_phi;
            running = !err || B::any(active);
        });

        // ---- Lamb-Helmholtz transform of the translated locals (transform_lh_local!, src/translate.jl:98-142)
        if constexpr (LH) {
            sfor<0, CAP + 1>([&](auto NN) {
                constexpr int n = decltype(NN)::value;
                if (n > kdone) return;
                constexpr int ke = KE(n), ko = KO(n);
                const I q = B::opaque(q0);
                // phi += (r m / n) (-/+) chi on the phi lanes; chi -= r / (n + 1) chi_(n+1) on the chi lanes below
                // the interaction's order. Per-lane coefficients, zero where a term does not apply, so each
                // register is two multiply-adds with the reference's rounding (x + 0 y = x).
                const V r_np1 = r * (1.0 / (n + 1));
                const V r_n = r * (1.0 / (n > 0 ? n : 1));
                const V bco = B::sel(chi_lane && (Pl > n), -r_np1, V(0.0));
                auto lh = [&](auto par, auto RI) {
                    constexpr int pr = decltype(par)::value, ri = decltype(RI)::value;
                    constexpr int at = LOFF(n) + (pr ? ke : 0) + ri;
                    constexpr int nn = imin(n + 1, CAP);
                    constexpr int at1 = LOFF(nn) + (pr ? KE(nn) : 0) + ri;
                    const I m = q * 2 + (8 * ri + pr);
                    const V x = L[at];
                    V nv = x;
                    if constexpr (n > 0) nv = x + (phi_sign * (r_n * B::tod(m))) * B::perm3(x);
                    if constexpr (n < CAP) nv = nv + bco * L[at1];
                    L[at] = nv;
                };
                sfor<0, ke>([&](auto EE) { lh(std::integral_constant<int, 0>(), EE); });
                sfor<0, ko>([&](auto OO) { lh(std::integral_constant<int, 1>(), OO); });
            });
        }

        // ---- back rotation (rotate_local_y!, eta magnitudes), back z-rotation, sum over slots
        sfor<0, CAP + 1>([&](auto NN) {
            constexpr int n = decltype(NN)::value;
            if (n > kdone) return;
            B::sched_fence();
            constexpr int ke = KE(n), ko = KO(n), ke1 = imax(ke, 1), ko1 = imax(ko, 1);
            const I c = B::opaque(c0), q = B::opaque(q0), s = c / NC, comp = c % NC;
            const I sb2 = s * (NS * 2);
            RotLane RL = RL0;
            RL.c = c;
            RL.q = q;
            RL.s = s;
            const M on = valid && (Pl >= n);
            V uE[ke1], uO[ko1], oE[ke1], oO[ko1];
            auto scale = [&](auto par, auto RI, V& u) {
                constexpr int pr = decltype(par)::value, ri = decltype(RI)::value;
                const I m = q * 2 + (8 * ri + pr);
                const M ok = on && (m <= n);
                const V g = B::lds(T.ein, B::seli(ok, m, I(0)) + mnh(n - 1));
                u = B::sel(ok, g * L[LO<n>::v + (pr ? ke : 0) + ri], V(0.0));
            };
            sfor<0, ke>([&](auto EE) { scale(std::integral_constant<int, 0>(), EE, uE[decltype(EE)::value]); });
            sfor<0, ko>([&](auto OO) { scale(std::integral_constant<int, 1>(), OO, uO[decltype(OO)::value]); });
            if constexpr (ko == 0) uO[0] = V(0.0);
            rotate<n>(T, W, T.eoh, -1.0, RL, uE, uO, oE, oO);
            auto store = [&](auto par, auto RI, const V& o) {
                constexpr int pr = decltype(par)::value, ri = decltype(RI)::value;
                const I m = q * 2 + (8 * ri + pr);
                const M ok = m <= n;
                const I ei = sb2 + B::mini(m, n) * 2;
                const V er = B::lds(W.eim, ei), eimg = B::lds(W.eim, ei + 1);
                const V y = o * er - (RL.nj * B::perm1(o)) * eimg;   // back_rotate_z! (src/rotate.jl:584-616)
                const V z = B::template slotsum<NC>(y);
                const M w = ok && (s == 0);
                B::lds_add(W.acc, B::seli(w, (m + mnh(n - 1)) * NC + comp, I(0)), z, w);
            };
            sfor<0, ke>([&](auto EE) { store(std::integral_constant<int, 0>(), EE, oE[decltype(EE)::value]); });
            sfor<0, ko>([&](auto OO) { store(std::integral_constant<int, 1>(), OO, oO[decltype(OO)::value]); });
        });
        B::wsync();

        // ---- orders and flags
        const M wr = valid && (q == 0) && (comp == 0);
        B::sti(Bt.orders, B::seli(wr, li, I(0)), Pl, wr);
        const int mx = B::maxi(Pl, valid);
        if (mx > pmax) pmax = mx;
        if (!B::all(succ || !valid)) okflag = 0;
    }
};

}  // namespace mx
}  // namespace fmmhip
