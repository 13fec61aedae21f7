// Far-field operators on the GPU: P2M, M2M, M2L (fixed order and the
// error-controlled variants), L2L, L2P for FastMultipole's solid-harmonic
// expansions (src/bodytomultipole.jl, src/rotate.jl, src/translate.jl,
// src/evaluate_expansions.jl). Expansions keep the reference layout per branch:
// [H][phi/chi][re/im] doubles.
//
// Translations use the reference's rotation scheme: z-rotation, y-rotation by
// a Wigner matrix assembled from the pi/2 tables, a translation along z, the
// Lamb-Helmholtz correction, then the inverse rotations. One wavefront runs one
// translation at a time with all intermediates in LDS: lanes split the (n, m)
// coefficients for rotations and translations and the (n, m, m') entries for
// the Wigner matrix. A wavefront owns one destination expansion (target box for
// M2L, parent for M2M, child for L2L) and accumulates in LDS, so results are
// written once and summation order is the reference's list order.
//
// P2M: one thread per output coefficient of a leaf, bodies staged in LDS;
// L2P: per-leaf coefficient sets for potential, field and field gradient, then
// one thread per target evaluating the regular solid harmonics by their
// Cartesian recurrences (no trigonometry per target).
//
// Compiled with -ffp-contract=off so that the error estimates that choose the
// expansion order follow the reference's arithmetic.
#include "fmmhip_internal.hpp"
#include <algorithm>
#include <cmath>

namespace fmmhip {

static constexpr int PMAX = 30;                 // largest supported expansion order
static constexpr double PI_D = 3.141592653589793;
static constexpr double ONE_OVER_4PI = 1.0 / (4.0 * 3.141592653589793);
static constexpr double SQRT3_D = 1.7320508075688772;
static constexpr double LOCAL_ERROR_SAFETY = 2.0;

__host__ __device__ inline int nh(int p) { return ((p + 1) * (p + 2)) >> 1; }
__host__ __device__ inline int len_Hs(int p) { int n = p + 1; return (n * (n + 1) * (n + 2)) / 6; }
__host__ __device__ inline int len_Ts(int p) { return p < 0 ? 0 : ((p + 1) * (p + 2) * (2 * p + 3)) / 6; }

// idx -> (n, m) for the packed (n, m>=0) ordering
__device__ inline void idx_nm(int idx, int& n, int& m) {
    int k = (int)((sqrtf(8.0f * (float)idx + 1.0f) - 1.0f) * 0.5f);
    while (((k + 1) * (k + 2)) / 2 <= idx) ++k;
    while ((k * (k + 1)) / 2 > idx) --k;
    n = k;
    m = idx - (k * (k + 1)) / 2;
}

// ============================================================== tables
struct Tables {
    int p = -1;
    DBuf<double> Hs, zeta, eta, Mt, Lt, inv_fact;
    DBuf<uint32_t> tent;   // Wigner entries (n | m<<8 | mp<<16), ordered by n
};

// Host computation of the pi/2 Wigner tables, zeta/eta magnitudes and power
// normalisations (src/rotate.jl:95-223, 625-1004; src/translate.jl:366-448),
// in the reference's operation order.
static std::vector<double> host_pnms(int p, double s, double c) {
    std::vector<double> P(nh(p), 0.0);
    double Pn0 = 1.0, fact = 1.0;
    for (int m = 0; m <= p; ++m) {
        double Pnm = Pn0;
        P[(m * (m + 1)) / 2 + m] = Pnm;
        double t1 = Pnm;
        Pnm = c * (double)(2 * m + 1) * t1;
        for (int n = m + 1; n <= p; ++n) {
            P[(n * (n + 1)) / 2 + m] = Pnm;
            double t2 = t1;
            t1 = Pnm;
            Pnm = (c * (double)(2 * n + 1) * t1 - (double)(n + m) * t2) / (double)(n - m + 1);
        }
        Pn0 = -Pn0 * fact * s;
        fact += 2;
    }
    return P;
}

static void host_Hs(std::vector<double>& Hs, int pmax) {
    const double beta = PI_D / 2;
    const double sb = std::sin(beta), cb = std::cos(beta);
    std::vector<double> P = host_pnms(pmax + 1, sb, cb);
    auto Pn = [&](int n, int m) { return P[(n * (n + 1)) / 2 + m]; };
    std::vector<double> Hn1(pmax + 3), b(2 * (pmax + 1) + 2), d(pmax + 2);
    Hs.assign(len_Hs(pmax), 0.0);
    Hs[0] = 1.0;
    for (int n = 1; n <= pmax; ++n) {
        double* H = &Hs[len_Hs(n - 1)];
        auto Hat = [&](int mp, int m) -> double& { return H[(m * (m + 1)) / 2 + mp]; };
        double s = 1.0, f = 1.0;
        for (int m = 0; m <= n; ++m) { Hat(0, m) = s * f * Pn(n, m); s = -s; f /= std::sqrt((double)((n + m + 1) * (n - m))); }
        s = 1.0; f = 1.0;
        for (int m = 0; m <= n + 1; ++m) { Hn1[m] = s * f * Pn(n + 1, m); s = -s; f /= std::sqrt((double)((n + 1 - m) * (n + m + 2))); }
        const int bn = n + 1;
        for (int m = -bn; m <= -1; ++m)
            b[bn + m] = -std::sqrt((double)(bn - m - 1) / (double)(2 * bn - 1) * (double)(bn - m) / (double)(2 * bn + 1));
        for (int m = 0; m <= bn; ++m)
            b[bn + m] = std::sqrt((double)(bn - m - 1) / (double)(2 * bn - 1) * (double)(bn - m) / (double)(2 * bn + 1));
        const double omc = (1.0 - cb) / 2, opc = (1.0 + cb) / 2;
        for (int m = 1; m <= n; ++m) {   // b at 1-based index k is b[k-1]
            double bmm1 = b[bn - m - 1], bm1 = b[bn + m - 1], b0 = b[bn];
            double anm = std::sqrt((double)(n + 1 + m) / (double)(2 * n + 1) * (double)(n + 1 - m) / (double)(2 * n + 3));
            Hat(1, m) = (bmm1 * omc * Hn1[m + 1] - bm1 * opc * Hn1[m - 1] - anm * sb * Hn1[m]) / b0;
        }
        for (int m = 0; m <= n; ++m) d[m] = std::sqrt((double)((n - m) * (n + m + 1))) / 2;
        auto D = [&](int k) { return d[k - 1]; };
        for (int mp = 1; mp <= n - 1; ++mp) {
            for (int m = mp + 1; m <= n - 1; ++m)
                Hat(mp + 1, m) = (D(mp) * Hat(mp - 1, m) - D(m) * Hat(mp, m - 1) + D(m + 1) * Hat(mp, m + 1)) / D(mp + 1);
            int m = n;
            Hat(mp + 1, m) = (D(mp) * Hat(mp - 1, m) - D(m) * Hat(mp, m - 1)) / D(mp + 1);
        }
    }
}

static void host_mags(std::vector<double>& z, std::vector<double>& e, int pmax) {
    z.assign(len_Ts(pmax), 0.0);
    e.assign(len_Ts(pmax), 0.0);
    z[0] = e[0] = 1.0;
    for (int n = 1; n <= pmax; ++n) {
        double* zz = &z[len_Ts(n - 1)];
        double* ee = &e[len_Ts(n - 1)];
        double zm = 1.0, em = 1.0;
        int i = 0;
        for (int m = 0; m <= n; ++m) {
            double zmp = zm, emp = em;
            for (int mp = 0; mp <= n; ++mp) {
                zz[i] = zmp;
                ee[i] = emp;
                int a = (mp + 1 < n) ? n - mp : 1;
                zmp *= std::sqrt((double)(n + mp + 1) / (double)a);
                emp *= std::sqrt((double)a / (double)(n + mp + 1));
                ++i;
            }
            int a = (m < n - 1) ? n - m : 1;
            zm *= std::sqrt((double)a / (double)(n + m + 1));
            em *= std::sqrt((double)(n + m + 1) / (double)a);
        }
    }
}

// k! correctly rounded to double (exact integer product, one rounding)
static double host_factorial(int k) {
    std::vector<uint32_t> limbs(1, 1u);
    for (int f = 2; f <= k; ++f) {
        uint64_t carry = 0;
        for (auto& l : limbs) { uint64_t v = (uint64_t)l * (uint64_t)f + carry; l = (uint32_t)v; carry = v >> 32; }
        if (carry) limbs.push_back((uint32_t)carry);
    }
    int top = (int)limbs.size() - 1;
    int nbits = 32 * top + (32 - __builtin_clz(limbs[top]));
    if (nbits <= 53) { double r = 0; for (int i = top; i >= 0; --i) r = r * 4294967296.0 + limbs[i]; return r; }
    auto bit = [&](int b) { return (int)((limbs[b / 32] >> (b % 32)) & 1u); };
    uint64_t lead = 0;
    for (int b2 = nbits - 1; b2 >= nbits - 64; --b2) lead = (lead << 1) | (uint64_t)bit(b2);
    bool sticky = false;
    for (int b2 = nbits - 65; b2 >= 0 && !sticky; --b2) sticky = bit(b2);
    uint64_t mant = lead >> 11, rem = lead & 0x7FF;
    if (rem > 0x400 || (rem == 0x400 && (sticky || (mant & 1)))) mant += 1;
    return std::ldexp((double)mant, nbits - 53);
}

void ensure_tables(Context& C, int p) {
    if (p > PMAX) throw Error(FMMHIP_EINVAL, "expansion order above 30 is not supported");
    if (!C.tables) C.tables = new Tables();
    Tables& T = *C.tables;
    if (T.p >= PMAX) return;
    const int pm = PMAX;
    std::vector<double> Hs, z, e, Mt(nh(pm)), Lt(nh(pm)), invf(2 * pm + 4);
    host_Hs(Hs, pm);
    host_mags(z, e, pm);
    Mt[0] = Lt[0] = 1.0;
    for (int n = 1, i = 1; n <= pm; ++n)
        for (int m = 0; m <= n; ++m, ++i) {
            double a = host_factorial(n + m), b = host_factorial(n - m);
            Mt[i] = std::sqrt(a * b / (double)(2 * n + 1));
            Lt[i] = std::sqrt(1.0 / (a * b * (double)(2 * n + 1)));
        }
    for (int k = 0; k < (int)invf.size(); ++k) invf[k] = 1.0 / host_factorial(k);
    std::vector<uint32_t> tent;
    for (int n = 1; n <= pm; ++n)
        for (int m = 0; m <= n; ++m)
            for (int mp = 0; mp <= m; ++mp) tent.push_back((uint32_t)n | ((uint32_t)m << 8) | ((uint32_t)mp << 16));
    auto up = [&](DBuf<double>& d, const std::vector<double>& h) {
        d.resize(h.size());
        HIP_CHECK(hipMemcpy(d.data(), h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    };
    up(T.Hs, Hs); up(T.zeta, z); up(T.eta, e); up(T.Mt, Mt); up(T.Lt, Lt); up(T.inv_fact, invf);
    T.tent.resize(tent.size());
    HIP_CHECK(hipMemcpy(T.tent.data(), tent.data(), tent.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    T.p = pm;
}

void free_tables(Tables* t) { delete t; }

struct TablePtrs {
    const double *Hs, *zeta, *eta, *Mt, *Lt, *inv_fact;
    const uint32_t* tent;
};
static TablePtrs ptrs(Context& C) {
    Tables& T = *C.tables;
    return {T.Hs.data(), T.zeta.data(), T.eta.data(), T.Mt.data(), T.Lt.data(), T.inv_fact.data(), T.tent.data()};
}

// ======================================================= wavefront engine
// LDS expansion arrays use the per-branch global layout: [idx][pot][reim]
#define EX(a, i, pot, ri) (a)[4 * (i) + 2 * (pot) + (ri)]

struct Lds {
    double* acc;   // 4H
    double* w1;    // 4H
    double* w2;    // 4H
    double* T;     // len_Ts(p)
    double* eim;   // 2(p+1)
    double* misc;  // 16 + 4(p+1)
};

__device__ inline Lds carve(double* base, int p) {
    const int H4 = 4 * nh(p);
    Lds L;
    L.acc = base;
    L.w1 = L.acc + H4;
    L.w2 = L.w1 + H4;
    L.T = L.w2 + H4;
    L.eim = L.T + len_Ts(p);
    L.misc = L.eim + 2 * (p + 1);
    return L;
}
static size_t lds_bytes(int p) { return sizeof(double) * (3 * 4 * nh(p) + len_Ts(p) + 2 * (p + 1) + 16 + 4 * (p + 1)); }

__device__ inline void c2s(double x, double y, double z, double& r, double& th, double& ph) {   // src/rotate.jl:9-21
    double x2y2 = x * x + y * y;
    double r2 = x2y2 + z * z;
    r = (r2 == 0.0) ? r2 : sqrt(r2);
    double z_r = z / r;
    if (r > 0) th = x2y2 > 0 ? acos(z_r) : (z < 0 ? PI_D : 0.0);
    else th = 0.0;
    ph = (x2y2 == 0.0) ? 0.0 : atan2(y, x);
}

// e^{i m phi}, m = 0..P by repeated multiplication (src/rotate.jl:25-36); lane 0 only
__device__ inline void eim_fill(double* eim, double phi, int P) {
    if (threadIdx.x == 0) {
        double ei_i = sin(phi), ei_r = cos(phi);
        double er = 1.0, e_i = 0.0;
        for (int m = 0; m <= P; ++m) {
            eim[2 * m] = er;
            eim[2 * m + 1] = e_i;
            double t = er;
            er = er * ei_r - e_i * ei_i;
            e_i = t * ei_i + e_i * ei_r;
        }
    }
}

__device__ inline double get_scalar(int m_mp) {
    if (m_mp & 1) m_mp += 1;
    return ((m_mp >> 1) & 1) ? -1.0 : 1.0;
}

// one Wigner entry (n, m, mp): the (m, +-mp) pair of T for angle beta (src/rotate.jl:380-435)
__device__ inline void wigner_entry(double* T, const double* Hs, double sb, double cb, int n, int m, int mp) {
    const double* Hp = Hs + len_Hs(n - 1);
    auto Hat = [&](int a, int b) { return Hp[(b * (b + 1)) / 2 + a]; };
    const bool even = ((m + mp) & 1) == 0;
    double scalar = get_scalar(m + mp);
    double vpos = 0.0, vneg = 0.0, s_prev = 0.0, c_prev = 1.0;
    double sgn = ((n + mp + 1) & 1) ? -1.0 : 1.0;           // (-1)^(n+mp+nu) at nu = 1
    const double one_mp_odd = (mp & 1) ? -1.0 : 1.0;
    for (int nu = 1; nu <= n; ++nu) {
        double c_nb = cb * c_prev - sb * s_prev;
        double s_nb = sb * c_prev + cb * s_prev;
        double sc = even ? scalar * c_nb : scalar * s_nb;
        double val = Hat(min(mp, nu), max(mp, nu)) * Hat(min(m, nu), max(m, nu)) * sc;
        vpos += val;
        vneg += val * sgn * one_mp_odd;
        c_prev = c_nb;
        s_prev = s_nb;
        sgn = -sgn;
    }
    vpos *= 2.0;
    vneg *= 2.0;
    scalar = even ? scalar : 0.0;
    double val = Hat(0, m) * Hat(0, mp) * scalar;
    const double one_n_mp = ((n + mp) & 1) ? -1.0 : 1.0;
    vpos += val;
    vneg += val * one_n_mp * one_mp_odd;
    double* Tn = T + len_Ts(n - 1);
    Tn[m * m + mp + m] = vpos;
    Tn[m * m - mp + m] = vneg;
}

// all Wigner entries for degrees 1..P (lane-parallel); T[0] = 1
__device__ inline void wigner_all(double* T, const TablePtrs& tb, double beta, int P) {
    const double sb = sin(beta), cb = cos(beta);
    if (threadIdx.x == 0) T[0] = 1.0;
    const int ne = len_Hs(P) - 1;
    for (int k = threadIdx.x; k < ne; k += 64) {
        uint32_t e = tb.tent[k];
        wigner_entry(T, tb.Hs, sb, cb, (int)(e & 0xff), (int)((e >> 8) & 0xff), (int)(e >> 16));
    }
}

// y-rotation of coefficient (n, m) using magnitudes `mags` (zeta: multipole, eta: local)
template <bool ETA, bool LH>
__device__ inline void rotate_y_one(const double* in, const double* T, const double* mags, int n, int m, double out[4]) {
    const double* Tn = T + len_Ts(n - 1);
    const double* mg = mags + len_Ts(n - 1) + m * (n + 1);
    double v1r = 0, v1i = 0, v2r = 0, v2i = 0;
    auto sign4 = [&](int mp, double mag, double& zr, double& zi) {
        int mod = ETA ? (m - abs(mp)) % 4 : (abs(mp) - m) % 4;
        if (mod < 0) mod += 4;
        zr = (mod == 0) ? mag : (mod == 2) ? -mag : 0.0;
        zi = (mod == 1) ? mag : (mod == 3) ? -mag : 0.0;
    };
    const int base = (n * (n + 1)) / 2;
    for (int mp = -n; mp <= 0; ++mp) {
        int i1 = min(-mp, m), i2 = max(-mp, m);
        double Tv = (n == 0) ? T[0] : Tn[i2 * i2 - i1 + i2];
        double zr, zi;
        sign4(mp, mg[-mp] * Tv, zr, zi);
        int iw = base - mp;
        double Mr = EX(in, iw, 0, 0), Mi = -EX(in, iw, 0, 1);
        if (mp & 1) { Mr = -Mr; Mi = -Mi; }
        v1r += Mr * zr - Mi * zi;
        v1i += Mr * zi + Mi * zr;
        if (LH) {
            Mr = EX(in, iw, 1, 0); Mi = -EX(in, iw, 1, 1);
            if (mp & 1) { Mr = -Mr; Mi = -Mi; }
            v2r += Mr * zr - Mi * zi;
            v2i += Mr * zi + Mi * zr;
        }
    }
    for (int mp = 1; mp <= n; ++mp) {
        int i1 = min(mp, m), i2 = max(mp, m);
        double Tv = Tn[i2 * i2 + i1 + i2];
        double zr, zi;
        sign4(mp, mg[mp] * Tv, zr, zi);
        int iw = base + mp;
        double Mr = EX(in, iw, 0, 0), Mi = EX(in, iw, 0, 1);
        v1r += Mr * zr - Mi * zi;
        v1i += Mr * zi + Mi * zr;
        if (LH) {
            Mr = EX(in, iw, 1, 0); Mi = EX(in, iw, 1, 1);
            v2r += Mr * zr - Mi * zi;
            v2i += Mr * zi + Mi * zr;
        }
    }
    out[0] = v1r; out[1] = v1i; out[2] = v2r; out[3] = v2i;
}

enum { OP_M2M = 0, OP_M2L = 1, OP_L2L = 2 };

// z-translation output (n, m) (src/translate.jl:6-50, 181-232, 1851-1891)
template <int OP, bool LH>
__device__ inline void translate_one(const double* in, double t, int P, int n, int m, double out[4]) {
    double v1r = 0, v1i = 0, v2r = 0, v2i = 0;
    if (OP == OP_M2M) {
        const double _t = -t;
        double n_np_f = 1.0, t_n_np = 1.0;
        int n_np = 1;
        for (int np = n; np >= m; --np) {
            double tmp = t_n_np / n_np_f;
            int iw = (np * (np + 1)) / 2 + m;
            v1r += tmp * EX(in, iw, 0, 0); v1i += tmp * EX(in, iw, 0, 1);
            if (LH) { v2r += tmp * EX(in, iw, 1, 0); v2i += tmp * EX(in, iw, 1, 1); }
            t_n_np *= _t;
            n_np_f *= (double)n_np;
            n_np += 1;
        }
    } else if (OP == OP_M2L) {
        const double one_over_t = 1.0 / t;
        double nf = one_over_t;   // n!/t^(n+1)
        for (int k = 0; k < n; ++k) nf *= (double)(k + 1) * one_over_t;
        double nmf = nf;          // (n+m)!/t^(n+m+1)
        for (int k = 0; k < m; ++k) nmf *= (double)(n + k + 1) * one_over_t;
        double f = nmf;
        int n_np = n + m;
        for (int np = m; np <= P; ++np) {
            int iw = (np * (np + 1)) / 2 + m;
            v1r += f * EX(in, iw, 0, 0); v1i += f * EX(in, iw, 0, 1);
            if (LH) { v2r += f * EX(in, iw, 1, 0); v2i += f * EX(in, iw, 1, 1); }
            n_np += 1;
            f *= (double)n_np * one_over_t;
        }
    } else {
        const double _t = -t;
        double t_n_nf = 1.0;
        int n_np = 1;
        for (int np = n; np <= P; ++np) {
            int iw = (np * (np + 1)) / 2 + m;
            v1r += t_n_nf * EX(in, iw, 0, 0); v1i += t_n_nf * EX(in, iw, 0, 1);
            if (LH) { v2r += t_n_nf * EX(in, iw, 1, 0); v2i += t_n_nf * EX(in, iw, 1, 1); }
            t_n_nf *= _t / (double)n_np;
            n_np += 1;
        }
    }
    out[0] = v1r; out[1] = v1i; out[2] = v2r; out[3] = v2i;
}

// Lamb-Helmholtz corrections, elementwise from the untransformed coefficients
// (multipole: src/translate.jl:52-96; local: :98-142)
template <bool LOCAL>
__device__ inline void lh_one(const double* in, double r, int P, int n, int m, double out[4]) {
    const int i = (n * (n + 1)) / 2 + m;
    double pr = EX(in, i, 0, 0), pi = EX(in, i, 0, 1), cr = EX(in, i, 1, 0), ci = EX(in, i, 1, 1);
    if (!LOCAL) {
        if (n >= 1) {
            double r_m_np1 = r * (double)m / (double)(n + 1);
            out[0] = pr + r_m_np1 * ci;
            out[1] = pi - r_m_np1 * cr;
            double c1r = 0, c1i = 0;
            if (m < n) { c1r = EX(in, i - n, 1, 0); c1i = EX(in, i - n, 1, 1); }
            double r_n = r / (double)n;
            out[2] = cr + r_n * c1r;
            out[3] = ci + r_n * c1i;
        } else { out[0] = pr; out[1] = pi; out[2] = cr; out[3] = ci; }
    } else {
        if (n > 0) {
            double r_m_n = r * (double)m / (double)n;
            out[0] = pr - r_m_n * ci;
            out[1] = pi + r_m_n * cr;
        } else { out[0] = pr; out[1] = pi; }
        if (n < P) {
            double c1r = EX(in, i + n + 1, 1, 0), c1i = EX(in, i + n + 1, 1, 1);
            double r_np1 = r / (double)(n + 1);
            out[2] = cr - r_np1 * c1r;
            out[3] = ci - r_np1 * c1i;
        } else { out[2] = cr; out[3] = ci; }
    }
}

// full rotate-translate-rotate of one expansion into acc (+=), all lanes
template <int OP, bool LH>
__device__ void translate_full(const double* __restrict__ src, double dx, double dy, double dz, const Lds& L, int P,
                               const TablePtrs& tb) {
    const int H = nh(P);
    double r, th, ph;
    c2s(dx, dy, dz, r, th, ph);
    eim_fill(L.eim, ph, P);
    wigner_all(L.T, tb, th, P);
    __syncthreads();
    for (int i = threadIdx.x; i < H; i += 64) {   // rotate_z (src/rotate.jl:448-483)
        int n, m;
        idx_nm(i, n, m);
        double er = L.eim[2 * m], ei = L.eim[2 * m + 1];
        double xr = src[4 * i], xi = src[4 * i + 1];
        EX(L.w1, i, 0, 0) = xr * er - xi * ei;
        EX(L.w1, i, 0, 1) = xr * ei + xi * er;
        if (LH) {
            xr = src[4 * i + 2]; xi = src[4 * i + 3];
            EX(L.w1, i, 1, 0) = xr * er - xi * ei;
            EX(L.w1, i, 1, 1) = xr * ei + xi * er;
        }
    }
    __syncthreads();
    const double* mag_fwd = (OP == OP_L2L) ? tb.eta : tb.zeta;
    for (int i = threadIdx.x; i < H; i += 64) {
        int n, m;
        idx_nm(i, n, m);
        double o[4];
        if (OP == OP_L2L) rotate_y_one<true, LH>(L.w1, L.T, mag_fwd, n, m, o);
        else rotate_y_one<false, LH>(L.w1, L.T, mag_fwd, n, m, o);
        EX(L.w2, i, 0, 0) = o[0]; EX(L.w2, i, 0, 1) = o[1];
        if (LH) { EX(L.w2, i, 1, 0) = o[2]; EX(L.w2, i, 1, 1) = o[3]; }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < H; i += 64) {
        int n, m;
        idx_nm(i, n, m);
        double o[4];
        translate_one<OP, LH>(L.w2, r, P, n, m, o);
        EX(L.w1, i, 0, 0) = o[0]; EX(L.w1, i, 0, 1) = o[1];
        if (LH) { EX(L.w1, i, 1, 0) = o[2]; EX(L.w1, i, 1, 1) = o[3]; }
    }
    __syncthreads();
    const double* back_in = L.w1;
    if (LH) {
        for (int i = threadIdx.x; i < H; i += 64) {
            int n, m;
            idx_nm(i, n, m);
            double o[4];
            if (OP == OP_M2M) lh_one<false>(L.w1, r, P, n, m, o);
            else lh_one<true>(L.w1, r, P, n, m, o);
            EX(L.w2, i, 0, 0) = o[0]; EX(L.w2, i, 0, 1) = o[1]; EX(L.w2, i, 1, 0) = o[2]; EX(L.w2, i, 1, 1) = o[3];
        }
        __syncthreads();
        back_in = L.w2;
    }
    const double* mag_back = (OP == OP_M2M) ? tb.zeta : tb.eta;
    for (int i = threadIdx.x; i < H; i += 64) {   // back y-rotation, back z-rotation, accumulate
        int n, m;
        idx_nm(i, n, m);
        double o[4];
        if (OP == OP_M2M) rotate_y_one<false, LH>(back_in, L.T, mag_back, n, m, o);
        else rotate_y_one<true, LH>(back_in, L.T, mag_back, n, m, o);
        double er = L.eim[2 * m], ei = -L.eim[2 * m + 1];
        EX(L.acc, i, 0, 0) += o[0] * er - o[1] * ei;
        EX(L.acc, i, 0, 1) += o[0] * ei + o[1] * er;
        if (LH) {
            EX(L.acc, i, 1, 0) += o[2] * er - o[3] * ei;
            EX(L.acc, i, 1, 1) += o[2] * ei + o[3] * er;
        }
    }
    __syncthreads();
}

__device__ inline void acc_zero(const Lds& L, int P) {
    for (int i = threadIdx.x; i < 4 * nh(P); i += 64) L.acc[i] = 0.0;
    __syncthreads();
}

// ================================================================= M2M
__global__ __launch_bounds__(64) void k_m2m_level(Branch* __restrict__ br, double* __restrict__ ex, int64_t b0, int64_t b1,
                                                  int P, int LHi, TablePtrs tb) {
    extern __shared__ double lds[];
    const int64_t ib = b0 + blockIdx.x;
    if (ib > b1) return;
    const Branch B = br[ib];
    if (B.n_branches == 0) return;
    Lds L = carve(lds, P);
    acc_zero(L, P);
    const int H4 = 4 * nh(P);
    for (int64_t c = B.child_first; c < B.child_first + B.n_branches; ++c) {
        const Branch& Ch = br[c];
        double dx = B.source_center[0] - Ch.source_center[0], dy = B.source_center[1] - Ch.source_center[1],
               dz = B.source_center[2] - Ch.source_center[2];
        if (LHi) translate_full<OP_M2M, true>(ex + c * H4, dx, dy, dz, L, P, tb);
        else translate_full<OP_M2M, false>(ex + c * H4, dx, dy, dz, L, P, tb);
    }
    for (int i = threadIdx.x; i < H4; i += 64) ex[ib * H4 + i] += L.acc[i];
}

void run_m2m(Context& C, Tree& st, int p, bool LH) {
    ensure_tables(C, p);
    TablePtrs tb = ptrs(C);
    const size_t sh = lds_bytes(p);
    for (int64_t l = (int64_t)st.level_first.size() - 2; l >= 0; --l) {
        int64_t b0 = st.level_first[l], b1 = st.level_last[l];
        k_m2m_level<<<(unsigned)(b1 - b0 + 1), 64, sh, C.stream>>>(st.branches.data(), st.expansions.data(), b0, b1, p, LH ? 1 : 0, tb);
        LAUNCH_CHECK();
    }
}

// ================================================================= L2L
__global__ __launch_bounds__(64) void k_l2l_level(Branch* __restrict__ br, double* __restrict__ ex, int64_t b0, int64_t b1,
                                                  int P, int LHi, TablePtrs tb) {
    extern __shared__ double lds[];
    const int64_t ib = b0 + blockIdx.x;
    if (ib > b1) return;
    const Branch B = br[ib];
    const Branch& Pa = br[B.parent];
    Lds L = carve(lds, P);
    acc_zero(L, P);
    const int H4 = 4 * nh(P);
    double dx = B.target_center[0] - Pa.target_center[0], dy = B.target_center[1] - Pa.target_center[1],
           dz = B.target_center[2] - Pa.target_center[2];
    if (LHi) translate_full<OP_L2L, true>(ex + B.parent * H4, dx, dy, dz, L, P, tb);
    else translate_full<OP_L2L, false>(ex + B.parent * H4, dx, dy, dz, L, P, tb);
    for (int i = threadIdx.x; i < H4; i += 64) ex[ib * H4 + i] += L.acc[i];
}

void run_l2l(Context& C, Tree& tt, int p, bool LH) {
    ensure_tables(C, p);
    TablePtrs tb = ptrs(C);
    const size_t sh = lds_bytes(p);
    for (size_t l = 1; l < tt.level_first.size(); ++l) {
        int64_t b0 = tt.level_first[l], b1 = tt.level_last[l];
        k_l2l_level<<<(unsigned)(b1 - b0 + 1), 64, sh, C.stream>>>(tt.branches.data(), tt.expansions.data(), b0, b1, p, LH ? 1 : 0, tb);
        LAUNCH_CHECK();
    }
}

// ================================================================= M2L
struct ErrParams {
    int kind;        // 1 rotated coefficients (gradient), 2 power potential, 3 power gradient
    int relative;
    double eps;
    int BE;
};

__device__ inline double md1(double dx, double rm, double rp) {   // src/error.jl:3-10
    double right = dx + rp, left = dx + rm;
    bool same = right * left >= 0;
    double mn = (fabs(right) <= fabs(left)) ? right : left;
    return same ? mn : 0.0;
}

// error-controlled M2L of one pair into acc; returns the expansion order used (src/translate.jl:555-1412)
template <bool LH>
__device__ int m2l_error_pair(const double* __restrict__ src, const Branch& tb_, const Branch& sb_, const Lds& L, int P,
                              const TablePtrs& tb, const ErrParams& E, int& success) {
    const int H = nh(P);
    double* misc = L.misc;             // [0] decision flag, [1..] scratch
    double* tmp = L.misc + 16;         // 4(P+1) translated degree-n coefficients
    const double eps = E.relative ? tb_.max_influence * E.eps : E.eps;
    double dx = tb_.target_center[0] - sb_.source_center[0], dy = tb_.target_center[1] - sb_.source_center[1],
           dz = tb_.target_center[2] - sb_.source_center[2];
    double r, th, ph;
    c2s(dx, dy, dz, r, th, ph);
    double d0 = md1(tb_.target_center[0] - sb_.source_center[0], -tb_.target_box[0], tb_.target_box[0]);
    double d1 = md1(tb_.target_center[1] - sb_.source_center[1], -tb_.target_box[1], tb_.target_box[1]);
    double d2 = md1(tb_.target_center[2] - sb_.source_center[2], -tb_.target_box[2], tb_.target_box[2]);
    const double r_mp = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    const double r_l = tb_.target_radius;
    const double rinv = 1.0 / r, r_mp_inv = 1.0 / r_mp;
    double r_mp_inv_np2 = r_mp_inv * r_mp_inv * r_mp_inv;
    double r_l_nm1 = 1.0;
    const double sbeta = sin(th), cbeta = cos(th);
    eim_fill(L.eim, ph, P);
    if (threadIdx.x == 0) {
        L.T[0] = 1.0;
        EX(L.w1, 0, 0, 0) = src[0]; EX(L.w1, 0, 0, 1) = src[1];
        if (LH) { EX(L.w1, 0, 1, 0) = src[2]; EX(L.w1, 0, 1, 1) = src[3]; }
    }
    __syncthreads();
    if (threadIdx.x == 0) {   // degree 0 rotation (T[0] = 1)
        double o[4];
        rotate_y_one<false, LH>(L.w1, L.T, tb.zeta, 0, 0, o);
        EX(L.w2, 0, 0, 0) = o[0]; EX(L.w2, 0, 0, 1) = o[1]; EX(L.w2, 0, 1, 0) = o[2]; EX(L.w2, 0, 1, 1) = o[3];
    }
    double n_f_t_np1 = rinv * rinv, nm1_f_inv = 1.0, np1_f = 2.0;
    const bool power = E.kind >= 2;
    double Mt_n0 = tb.Mt[0], mp_pow = 0.0;
    if (power) { double a = src[0], b = src[1]; mp_pow = sqrt((a * a + b * b) * Mt_n0 * Mt_n0); }
    int Pused = P;
    success = 0;
    const double limit = eps * 4 * PI_D;
    for (int n = 1; n <= P; ++n) {
        const int in0 = (n * (n + 1)) / 2;
        // rotate_z degree n (src/rotate.jl:503-579)
        for (int m = threadIdx.x; m <= n; m += 64) {
            double er = L.eim[2 * m], ei = L.eim[2 * m + 1];
            int i = in0 + m;
            double xr = src[4 * i], xi = src[4 * i + 1];
            EX(L.w1, i, 0, 0) = xr * er - xi * ei;
            EX(L.w1, i, 0, 1) = xr * ei + xi * er;
            if (LH) {
                xr = src[4 * i + 2]; xi = src[4 * i + 3];
                EX(L.w1, i, 1, 0) = xr * er - xi * ei;
                EX(L.w1, i, 1, 1) = xr * ei + xi * er;
            }
        }
        // Wigner entries of degree n
        for (int k = threadIdx.x; k < nh(n); k += 64) {
            int m, mp;
            idx_nm(k, m, mp);
            wigner_entry(L.T, tb.Hs, sbeta, cbeta, n, m, mp);
        }
        __syncthreads();
        for (int m = threadIdx.x; m <= n; m += 64) {
            double o[4];
            rotate_y_one<false, LH>(L.w1, L.T, tb.zeta, n, m, o);
            int i = in0 + m;
            EX(L.w2, i, 0, 0) = o[0]; EX(L.w2, i, 0, 1) = o[1];
            EX(L.w2, i, 1, 0) = LH ? o[2] : 0.0; EX(L.w2, i, 1, 1) = LH ? o[3] : 0.0;
        }
        __syncthreads();
        // multipole error (uniform across lanes: every lane reads the same LDS values)
        double Mt_next = tb.Mt[in0], pow_next = 0.0, pow_c_next = 0.0;
        if (power) {
            for (int m = 0; m <= n; ++m) {
                int i = in0 + m;
                double Nt = tb.Mt[i];
                double xr = src[4 * i], xi = src[4 * i + 1];
                pow_next += (xr * xr + xi * xi) * Nt * Nt;
                if (LH) { xr = src[4 * i + 2]; xi = src[4 * i + 3]; pow_c_next += (xr * xr + xi * xi) * Nt * Nt; }
            }
            pow_next = sqrt(pow_next);
            pow_c_next = sqrt(pow_c_next);
        }
        double eps_mp;
        if (E.kind == 1) {
            double p0 = EX(L.w2, in0, 0, 0), p1r = EX(L.w2, in0 + 1, 0, 0), p1i = EX(L.w2, in0 + 1, 0, 1);
            eps_mp = (sqrt(p1r * p1r + p1i * p1i) + fabs(p0)) * np1_f * r_mp_inv_np2;
            if (LH) {
                double c1r = EX(L.w2, in0 + 1, 1, 0), c1i = EX(L.w2, in0 + 1, 1, 1);
                eps_mp += sqrt(c1r * c1r + c1i * c1i) * np1_f * r_mp_inv_np2 * r_mp;
            }
        } else if (E.kind == 3) {
            eps_mp = SQRT3_D * mp_pow * r_mp_inv_np2 * np1_f / Mt_n0;
            if (LH) eps_mp += SQRT3_D * (double)n * pow_c_next * r_mp_inv_np2 * np1_f / Mt_next;
        } else {
            eps_mp = SQRT3_D * pow_next * r_mp_inv_np2 * r_mp * np1_f / (Mt_n0 * (double)(n + 1));
            if (LH) eps_mp += SQRT3_D * (double)n * pow_c_next * r_mp_inv_np2 * r_mp * np1_f / (Mt_next * (double)(n + 1));
        }
        bool done = false;
        if (eps_mp <= limit) {
            double eps_l;
            if (E.kind == 1) {   // translate_multipole_to_local_z_m01_n (:302-362), all lanes redundantly
                const double one_over_t = rinv;
                double n_m_f = n_f_t_np1;
                int n_m = n;
                double p0r = 0.0, n_np_f = n_m_f;
                int n_np = n_m;
                for (int np = 0; np <= n; ++np) {
                    p0r += n_np_f * EX(L.w2, (np * (np + 1)) / 2, 0, 0);
                    n_np += 1;
                    n_np_f *= (double)n_np * one_over_t;
                }
                n_m += 1;
                n_m_f *= (double)n_m * one_over_t;
                double p1r = 0, p1i = 0, c1r = 0, c1i = 0;
                n_np_f = n_m_f;
                n_np = n_m;
                for (int np = 1; np <= n; ++np) {
                    int iw = (np * (np + 1)) / 2 + 1;
                    p1r += n_np_f * EX(L.w2, iw, 0, 0);
                    p1i += n_np_f * EX(L.w2, iw, 0, 1);
                    if (LH) { c1r += n_np_f * EX(L.w2, iw, 1, 0); c1i += n_np_f * EX(L.w2, iw, 1, 1); }
                    n_np += 1;
                    n_np_f *= (double)n_np * one_over_t;
                }
                if (LH) {
                    double t_over_n = r / (double)n;
                    p1r -= t_over_n * c1i;
                    p1i += t_over_n * c1r;
                }
                eps_l = (fabs(p1r) + fabs(p1i) + fabs(p0r)) * r_l_nm1 * nm1_f_inv;
                if (LH) eps_l += sqrt(c1r * c1r + c1i * c1i) * r_l_nm1 * r_l * nm1_f_inv;
            } else {
                // translate_multipole_to_local_z_n! (:238-295) for degree n, lanes over m
                for (int m = threadIdx.x; m <= n; m += 64) {
                    double n_m_f = n_f_t_np1;
                    for (int k = 0; k < m; ++k) n_m_f *= (double)(n + k + 1) * rinv;
                    double v1r = 0, v1i = 0, v2r = 0, v2i = 0, f = n_m_f;
                    int n_np = n + m;
                    for (int np = m; np <= n; ++np) {
                        int iw = (np * (np + 1)) / 2 + m;
                        v1r += f * EX(L.w2, iw, 0, 0); v1i += f * EX(L.w2, iw, 0, 1);
                        if (LH) { v2r += f * EX(L.w2, iw, 1, 0); v2i += f * EX(L.w2, iw, 1, 1); }
                        n_np += 1;
                        f *= (double)n_np * rinv;
                    }
                    tmp[4 * m] = v1r; tmp[4 * m + 1] = v1i; tmp[4 * m + 2] = v2r; tmp[4 * m + 3] = v2i;
                }
                __syncthreads();
                // local_power (:455-502)
                double lp = 0, lc = 0, Nt = 0, cnr = 0, cni = 0;
                for (int m = n; m >= 0; --m) {
                    Nt = tb.Lt[in0 + m];
                    double pr = tmp[4 * m], pi = tmp[4 * m + 1];
                    if (LH) {
                        double cor = tmp[4 * m + 2], coi = tmp[4 * m + 3];
                        pr = pr - coi * r * (double)m / (double)n;
                        pi = pi + cor * r * (double)m / (double)n;
                        double cr = cor - cnr * r / (double)(n + 1);
                        double ci = coi - cni * r / (double)(n + 1);
                        cnr = cor; cni = coi;
                        lc += (cr * cr + ci * ci) * Nt * Nt;
                    }
                    lp += (pr * pr + pi * pi) * Nt * Nt;
                }
                lp = sqrt(lp);
                if (LH) lc = sqrt(lc);
                if (E.kind == 3) {
                    eps_l = SQRT3_D * lp * r_l_nm1 * nm1_f_inv / Nt;
                    if (LH) eps_l += SQRT3_D * (double)n * lc * r_l_nm1 * r_l * nm1_f_inv / (Nt * (double)n);
                } else {
                    eps_l = SQRT3_D * lp * r_l_nm1 * r_l * nm1_f_inv / (Nt * (double)n);
                    if (LH) eps_l += SQRT3_D * (double)n * lc * r_l_nm1 * r_l * nm1_f_inv / (Nt * (double)n);
                }
                __syncthreads();
            }
            if (eps_mp + eps_l * LOCAL_ERROR_SAFETY <= limit) {
                Pused = n - 1 + (E.BE ? 1 : 0);
                success = 1;
                done = true;
            }
        }
        if (done) break;
        if (E.kind != 1 || n < P) {
            n_f_t_np1 *= (double)(n + 1) * rinv;
            r_l_nm1 *= r_l;
            r_mp_inv_np2 *= r_mp_inv;
            nm1_f_inv /= (double)n;
            np1_f *= (double)(n + 1);
        }
        if (power) { mp_pow = pow_next; Mt_n0 = Mt_next; }
    }
    __syncthreads();
    // translate, LH correction, back rotations into acc at order Pused
    const int Hu = nh(Pused);
    for (int i = threadIdx.x; i < Hu; i += 64) {
        int n, m;
        idx_nm(i, n, m);
        double o[4];
        translate_one<OP_M2L, LH>(L.w2, r, Pused, n, m, o);
        EX(L.w1, i, 0, 0) = o[0]; EX(L.w1, i, 0, 1) = o[1];
        if (LH) { EX(L.w1, i, 1, 0) = o[2]; EX(L.w1, i, 1, 1) = o[3]; }
    }
    __syncthreads();
    const double* back_in = L.w1;
    if (LH) {
        for (int i = threadIdx.x; i < Hu; i += 64) {
            int n, m;
            idx_nm(i, n, m);
            double o[4];
            lh_one<true>(L.w1, r, Pused, n, m, o);
            EX(L.w2, i, 0, 0) = o[0]; EX(L.w2, i, 0, 1) = o[1]; EX(L.w2, i, 1, 0) = o[2]; EX(L.w2, i, 1, 1) = o[3];
        }
        __syncthreads();
        back_in = L.w2;
    }
    for (int i = threadIdx.x; i < Hu; i += 64) {
        int n, m;
        idx_nm(i, n, m);
        double o[4];
        rotate_y_one<true, LH>(back_in, L.T, tb.eta, n, m, o);
        double er = L.eim[2 * m], ei = -L.eim[2 * m + 1];
        EX(L.acc, i, 0, 0) += o[0] * er - o[1] * ei;
        EX(L.acc, i, 0, 1) += o[0] * ei + o[1] * er;
        if (LH) {
            EX(L.acc, i, 1, 0) += o[2] * er - o[3] * ei;
            EX(L.acc, i, 1, 1) += o[2] * ei + o[3] * er;
        }
    }
    __syncthreads();
    (void)H;
    (void)misc;
    return Pused;
}

__global__ __launch_bounds__(64) void k_m2l(const int32_t* __restrict__ m2l_targets, const int64_t* __restrict__ m2l_offsets,
                                            const int32_t* __restrict__ m2l_s, const Branch* __restrict__ tbr,
                                            const Branch* __restrict__ sbr, const double* __restrict__ sex,
                                            double* __restrict__ tex, int P, int LHi, ErrParams E, int32_t* __restrict__ orders,
                                            int32_t* __restrict__ pmax, int32_t* __restrict__ all_ok, TablePtrs tb) {
    extern __shared__ double lds[];
    const int32_t it = m2l_targets[blockIdx.x];
    const int64_t l0 = m2l_offsets[blockIdx.x], l1 = m2l_offsets[blockIdx.x + 1];
    Lds L = carve(lds, P);
    acc_zero(L, P);
    const int H4 = 4 * nh(P);
    const Branch T = tbr[it];
    int local_max = 0, local_ok = 1;
    for (int64_t li = l0; li < l1; ++li) {
        const int32_t js = m2l_s[li];
        const Branch S = sbr[js];
        const double* src = sex + (int64_t)js * H4;
        int Pu = P, ok = 1;
        if (E.kind == 0) {
            double dx = T.target_center[0] - S.source_center[0], dy = T.target_center[1] - S.source_center[1],
                   dz = T.target_center[2] - S.source_center[2];
            if (LHi) translate_full<OP_M2L, true>(src, dx, dy, dz, L, P, tb);
            else translate_full<OP_M2L, false>(src, dx, dy, dz, L, P, tb);
        } else {
            if (LHi) Pu = m2l_error_pair<true>(src, T, S, L, P, tb, E, ok);
            else Pu = m2l_error_pair<false>(src, T, S, L, P, tb, E, ok);
        }
        if (threadIdx.x == 0) orders[li] = Pu;
        local_max = max(local_max, Pu);
        local_ok = local_ok && ok;
    }
    for (int i = threadIdx.x; i < H4; i += 64) tex[(int64_t)it * H4 + i] += L.acc[i];
    if (threadIdx.x == 0) {
        atomicMax(pmax, local_max);
        if (!local_ok) atomicAnd(all_ok, 0);
    }
}

__global__ void k_init_flags(int32_t* pmax, int32_t* ok) { *pmax = 0; *ok = 1; }

void run_m2l(Context& C, const Tree& tt, Tree& st, Tree& tt_dst, Lists& Ls, int p, bool LH, int err_kind, bool err_rel,
             double eps, bool BE, int precision, int32_t* pmax, int32_t* success) {
    (void)precision;
    (void)tt;
    ensure_tables(C, p);
    TablePtrs tb = ptrs(C);
    static DBuf<int32_t>& flags = *new DBuf<int32_t>();   // process lifetime
    flags.resize(2);
    k_init_flags<<<1, 1, 0, C.stream>>>(flags.data(), flags.data() + 1);
    LAUNCH_CHECK();
    if (Ls.n_m2l_targets > 0) {
        ErrParams E{err_kind, err_rel ? 1 : 0, eps, BE ? 1 : 0};
        k_m2l<<<(unsigned)Ls.n_m2l_targets, 64, lds_bytes(p), C.stream>>>(
            Ls.m2l_targets.data(), Ls.m2l_offsets.data(), Ls.m2l_s.data(), tt_dst.branches.data(), st.branches.data(),
            st.expansions.data(), tt_dst.expansions.data(), p, LH ? 1 : 0, E, Ls.m2l_order.data(), flags.data(),
            flags.data() + 1, tb);
        LAUNCH_CHECK();
    }
    int32_t h[2];
    HIP_CHECK(hipMemcpyAsync(h, flags.data(), sizeof(h), hipMemcpyDeviceToHost, C.stream));
    HIP_CHECK(hipStreamSynchronize(C.stream));
    *pmax = h[0];
    *success = h[1];
}

// ================================================================= P2M
// Regular solid harmonic R_j^k(v) (Gumerov normalisation with the
// Condon-Shortley phase) from the Cartesian recurrence of
// Q_j^k = rho^j P_j^k(cos theta) e^{i k phi}:
//   Q_k^k = -(2k-1)(x+iy) Q_{k-1}^{k-1},  Q_{k+1}^k = (2k+1) z Q_k^k,
//   Q_j^k = ((2j-1) z Q_{j-1}^k - (j+k-1) rho^2 Q_{j-2}^k) / (j-k),
//   R_j^k = (-1)^j i^k Q_j^k / (j+k)!.
// Returns R_n^k and R_{n-1}^k (k >= 0; zero where undefined).
__device__ inline void solid_pair(double x, double y, double z, int n, int k, const double* inv_fact, double R[4]) {
    R[0] = R[1] = R[2] = R[3] = 0.0;
    if (k < 0 || k > n) return;
    double qr = 1.0, qi = 0.0;
    for (int j = 1; j <= k; ++j) {   // Q_j^j
        double f = -(double)(2 * j - 1);
        double nr = f * (x * qr - y * qi), ni = f * (x * qi + y * qr);
        qr = nr; qi = ni;
    }
    const double rho2 = x * x + y * y + z * z;
    double pr = 0.0, pi = 0.0;   // Q_{j-1}
    double cr = qr, ci = qi;     // Q_j at j = k
    for (int j = k + 1; j <= n; ++j) {
        double a = (double)(2 * j - 1) * z, b = (double)(j + k - 1) * rho2, d = 1.0 / (double)(j - k);
        double nr = (a * cr - b * pr) * d, ni = (a * ci - b * pi) * d;
        pr = cr; pi = ci; cr = nr; ci = ni;
    }
    // i^k (-1)^j / (j+k)!
    auto finish = [&](int j, double q_r, double q_i, double& o_r, double& o_i) {
        double s = ((j & 1) ? -1.0 : 1.0) * inv_fact[j + k];
        double ar = q_r * s, ai = q_i * s;
        switch (k & 3) {
            case 0: o_r = ar; o_i = ai; break;
            case 1: o_r = -ai; o_i = ar; break;
            case 2: o_r = -ar; o_i = -ai; break;
            default: o_r = ai; o_i = -ar; break;
        }
    };
    finish(n, cr, ci, R[0], R[1]);
    if (n - 1 >= k) finish(n - 1, pr, pi, R[2], R[3]);
}

template <int KERNEL>
__global__ __launch_bounds__(64) void k_p2m(const Branch* __restrict__ br, const int64_t* __restrict__ leaf_index,
                                            const double* __restrict__ x, const double* __restrict__ y,
                                            const double* __restrict__ z, const double* __restrict__ q, int64_t ns,
                                            double* __restrict__ ex, int P, const double* __restrict__ inv_fact) {
    __shared__ double lb[6][64];
    const int64_t ib = leaf_index[blockIdx.x];
    const Branch B = br[ib];
    const int H = nh(P);
    const double cx = B.source_center[0], cy = B.source_center[1], cz = B.source_center[2];
    for (int i0 = 0; i0 < H; i0 += 64) {
        const int i = i0 + threadIdx.x;
        int n = 0, m = 0;
        if (i < H) idx_nm(i, n, m);
        double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
        for (int64_t b0 = 0; b0 < B.n_bodies; b0 += 64) {
            const int cnt = (int)min<int64_t>(64, B.n_bodies - b0);
            __syncthreads();
            if ((int)threadIdx.x < cnt) {
                const int64_t j = B.first + b0 + threadIdx.x;
                lb[0][threadIdx.x] = x[j] - cx; lb[1][threadIdx.x] = y[j] - cy; lb[2][threadIdx.x] = z[j] - cz;
                lb[3][threadIdx.x] = q[j];
                if (KERNEL == FMMHIP_POINT_VORTEX) { lb[4][threadIdx.x] = q[ns + j]; lb[5][threadIdx.x] = q[2 * ns + j]; }
            }
            __syncthreads();
            if (i >= H) continue;
            for (int s = 0; s < cnt; ++s) {
                if (KERNEL == FMMHIP_POINT_SOURCE) {
                    // M_n^m += (-1)^(n+m) (-q) conj(R_n^m(dx))   (src/bodytomultipole.jl:403-429)
                    double R[4];
                    solid_pair(lb[0][s], lb[1][s], lb[2][s], n, m, inv_fact, R);
                    double f = (((n + m) & 1) ? 1.0 : -1.0) * lb[3][s];
                    a0 += f * R[0];
                    a1 -= f * R[1];
                } else {
                    // mirrored_source_to_vortex! with q = R(-dx) (src/bodytomultipole.jl:283-334, 465-489)
                    const double vx = lb[3][s], vy = lb[4][s], vz = lb[5][s];
                    const double ux = -lb[0][s], uy = -lb[1][s], uz = -lb[2][s];
                    double Rm1[4], R0[4], Rp1[4];
                    solid_pair(ux, uy, uz, n, m - 1 >= 0 ? m - 1 : 1, inv_fact, Rm1);
                    solid_pair(ux, uy, uz, n, m, inv_fact, R0);
                    solid_pair(ux, uy, uz, n, m + 1, inv_fact, Rp1);
                    const double om = (m & 1) ? -1.0 : 1.0;
                    // phi part: q_n^{m-1}, q_n^m, q_n^{m+1}
                    double qm1r, qm1i;
                    if (m > 0) { qm1r = Rm1[0]; qm1i = Rm1[1]; }
                    else if (n == 0) { qm1r = 0.0; qm1i = 0.0; }
                    else { qm1r = -om * Rm1[0]; qm1i = om * Rm1[1]; }   // q_n^{-1} = -conj(q_n^1)
                    double qmi = R0[1], qmr = R0[0];
                    double qp1r = (m < n) ? Rp1[0] : 0.0, qp1i = (m < n) ? Rp1[1] : 0.0;
                    const double A = (double)(n - m + 1) * 0.5, Bc = (double)(n + m + 1) * 0.5, inp1 = 1.0 / (double)(n + 1);
                    a0 += om * ((-vx * qm1r + vy * qm1i) * A + (vx * qp1r + vy * qp1i) * Bc - vz * (double)m * qmi) * inp1;
                    a1 += om * ((vx * qm1i + vy * qm1r) * A + (-vx * qp1i + vy * qp1r) * Bc - vz * (double)m * qmr) * inp1;
                    if (n >= 1) {
                        // chi part: q_{n-1}^{m-1}, q_{n-1}^m, q_{n-1}^{m+1} (get_nm1 rules)
                        double pm1r, pm1i;
                        if (m > 0) { pm1r = Rm1[2]; pm1i = Rm1[3]; }
                        else if (n > 1) { pm1r = -om * Rm1[2]; pm1i = om * Rm1[3]; }
                        else { pm1r = 0.0; pm1i = 0.0; }
                        double p0r = (m < n) ? R0[2] : 0.0, p0i = (m < n) ? R0[3] : 0.0;
                        double pp1r = (m + 1 < n) ? Rp1[2] : 0.0, pp1i = (m + 1 < n) ? Rp1[3] : 0.0;
                        const double ion = 1.0 / (double)n;
                        a2 -= om * ion * (0.5 * (-vy * pm1r - vx * pm1i + vy * pp1r - vx * pp1i) - vz * p0r);
                        a3 -= om * ion * (0.5 * (vy * pm1i - vx * pm1r - vy * pp1i - vx * pp1r) + vz * p0i);
                    }
                }
            }
        }
        if (i < H) {
            double* e = ex + ib * 4 * H + 4 * i;
            e[0] += a0; e[1] += a1;
            if (KERNEL == FMMHIP_POINT_VORTEX) { e[2] += a2; e[3] += a3; }
        }
    }
}

void run_p2m(Context& C, int kernel, Tree& st, int p) {
    ensure_tables(C, p);
    if (st.n_leaves == 0) return;
    TablePtrs tb = ptrs(C);
    if (kernel == FMMHIP_POINT_SOURCE)
        k_p2m<FMMHIP_POINT_SOURCE><<<(unsigned)st.n_leaves, 64, 0, C.stream>>>(st.branches.data(), st.leaf_index.data(), st.x.data(), st.y.data(),
                                                                              st.z.data(), st.strength.data(), st.n, st.expansions.data(), p, tb.inv_fact);
    else
        k_p2m<FMMHIP_POINT_VORTEX><<<(unsigned)st.n_leaves, 64, 0, C.stream>>>(st.branches.data(), st.leaf_index.data(), st.x.data(), st.y.data(),
                                                                              st.z.data(), st.strength.data(), st.n, st.expansions.data(), p, tb.inv_fact);
    LAUNCH_CHECK();
}

// ================================================================= L2P
// field-expansion coefficients of the local expansion (src/evaluate_expansions.jl:86-300)
template <bool LH>
__device__ inline void vcoef(const double* Lx, int P, int n, int m, double V[6]) {
    for (int k = 0; k < 6; ++k) V[k] = 0.0;   // x(re,im), y(re,im), z(re,im)
    auto Lre = [&](int nn, int mm) { return Lx[4 * ((nn * (nn + 1)) / 2 + mm)]; };
    auto Lim = [&](int nn, int mm) { return Lx[4 * ((nn * (nn + 1)) / 2 + mm) + 1]; };
    auto Cre = [&](int nn, int mm) { return Lx[4 * ((nn * (nn + 1)) / 2 + mm) + 2]; };
    auto Cim = [&](int nn, int mm) { return Lx[4 * ((nn * (nn + 1)) / 2 + mm) + 3]; };
    if (n == 0) {
        if (P >= 1) {
            V[0] -= Lim(1, 1); V[2] -= Lre(1, 1); V[4] -= Lre(1, 0); V[5] -= Lim(1, 0);
        }
        return;
    }
    if (m == 0) {
        if (n < P) { V[0] -= Lim(n + 1, 1); V[2] -= Lre(n + 1, 1); V[4] -= Lre(n + 1, 0); V[5] -= Lim(n + 1, 0); }
        if (LH) { V[0] += (double)n * Cre(n, 1); V[2] -= (double)n * Cim(n, 1); }
        return;
    }
    if (n < P) {
        double ar = Lre(n + 1, m - 1), ai = Lim(n + 1, m - 1), br_ = Lre(n + 1, m), bi = Lim(n + 1, m);
        double cr = Lre(n + 1, m + 1), ci = Lim(n + 1, m + 1);
        V[0] -= (ai + ci) * 0.5; V[1] += (ar + cr) * 0.5;
        V[2] += (ar - cr) * 0.5; V[3] += (ai - ci) * 0.5;
        V[4] -= br_; V[5] -= bi;
    }
    if (LH) {
        double cmr = Cre(n, m - 1), cmi = Cim(n, m - 1);
        V[0] -= (double)(n + m) * cmr * 0.5; V[1] -= (double)(n + m) * cmi * 0.5;
        V[2] -= (double)(n + m) * cmi * 0.5; V[3] += (double)(n + m) * cmr * 0.5;
        if (m < n) {
            double cpr = Cre(n, m + 1), cpi = Cim(n, m + 1);
            V[0] += (double)(n - m) * cpr * 0.5; V[1] += (double)(n - m) * cpi * 0.5;
            V[2] -= (double)(n - m) * cpi * 0.5; V[3] += (double)(n - m) * cpr * 0.5;
        }
        double c0r = Cre(n, m), c0i = Cim(n, m);
        V[4] += (double)m * c0i; V[5] -= (double)m * c0r;
    }
}

// per-leaf evaluation coefficients: for each (n, m) 13 complex values C' such that
// the contribution to output k is Re(Q_n^m(dx) * C'_k), with Q as in solid_pair
template <bool LH>
__global__ __launch_bounds__(64) void k_l2p_coef(const int64_t* __restrict__ leaf_index, const double* __restrict__ ex, int P,
                                                 int PS, int GS, int HS, const double* __restrict__ inv_fact,
                                                 double* __restrict__ coef) {
    const int64_t ib = leaf_index[blockIdx.x];
    const int H = nh(P);
    const double* Lx = ex + ib * 4 * H;
    double* out = coef + (int64_t)blockIdx.x * 26 * H;
    for (int i = threadIdx.x; i < H; i += 64) {
        int n, m;
        idx_nm(i, n, m);
        double c[26];
        for (int k = 0; k < 26; ++k) c[k] = 0.0;
        const double w = (m == 0) ? 1.0 : 2.0;
        if (PS && (n == 0 || !LH)) { c[0] = w * Lx[4 * i]; c[1] = w * Lx[4 * i + 1]; }
        if (GS) {
            double V[6];
            vcoef<LH>(Lx, P, n, m, V);
            for (int k = 0; k < 6; ++k) c[2 + k] = w * V[k];
        }
        if (HS && n <= P - 1) {
            double A[6], B[6], Cc[6];
            vcoef<LH>(Lx, P, n + 1, m > 0 ? m - 1 : 0, A);
            vcoef<LH>(Lx, P, n + 1, m, B);
            vcoef<LH>(Lx, P, n + 1, m + 1, Cc);
            for (int comp = 0; comp < 3; ++comp) {   // field component x, y, z
                double ar = A[2 * comp], ai = A[2 * comp + 1], br_ = B[2 * comp], bi = B[2 * comp + 1];
                double cr = Cc[2 * comp], ci = Cc[2 * comp + 1];
                double Xr, Xi, Yr, Yi, Zr, Zi;
                if (m == 0) {
                    Xr = -ci; Xi = 0.0; Yr = -cr; Yi = 0.0; Zr = -br_; Zi = -bi;
                } else {
                    Xr = -(ai + ci) * 0.5; Xi = (ar + cr) * 0.5;
                    Yr = (ar - cr) * 0.5;  Yi = (ai - ci) * 0.5;
                    Zr = -br_; Zi = -bi;
                }
                // output index 3*d + comp: d = derivative direction x, y, z
                c[8 + 2 * (0 + comp)] = w * Xr; c[8 + 2 * (0 + comp) + 1] = w * Xi;
                c[8 + 2 * (3 + comp)] = w * Yr; c[8 + 2 * (3 + comp) + 1] = w * Yi;
                c[8 + 2 * (6 + comp)] = w * Zr; c[8 + 2 * (6 + comp) + 1] = w * Zi;
            }
        }
        // fold (-1)^n i^m / (n+m)!: Re(R C) = Re(Q * s i^m C)
        const double s = ((n & 1) ? -1.0 : 1.0) * inv_fact[n + m];
        for (int k = 0; k < 13; ++k) {
            double cr = c[2 * k] * s, ci = c[2 * k + 1] * s, orr, oi;
            switch (m & 3) {
                case 0: orr = cr; oi = ci; break;
                case 1: orr = -ci; oi = cr; break;
                case 2: orr = -cr; oi = -ci; break;
                default: orr = ci; oi = -cr; break;
            }
            out[26 * i + 2 * k] = orr;
            out[26 * i + 2 * k + 1] = oi;
        }
    }
}

__global__ __launch_bounds__(128) void k_l2p_eval(const Branch* __restrict__ br, const int64_t* __restrict__ leaf_index,
                                                  const double* __restrict__ x, const double* __restrict__ y,
                                                  const double* __restrict__ z, int P, const double* __restrict__ coef,
                                                  double* __restrict__ tout, int64_t nt, int PS, int GS, int HS) {
    extern __shared__ double lc[];
    const int64_t ib = leaf_index[blockIdx.x];
    const Branch B = br[ib];
    const int H = nh(P);
    const double* cg = coef + (int64_t)blockIdx.x * 26 * H;
    for (int k = threadIdx.x; k < 26 * H; k += blockDim.x) lc[k] = cg[k];
    __syncthreads();
    for (int64_t t = threadIdx.x; t < B.n_bodies; t += blockDim.x) {
        const int64_t it = B.first + t;
        const double dx = x[it] - B.target_center[0], dy = y[it] - B.target_center[1], dz = z[it] - B.target_center[2];
        const double rho2 = dx * dx + dy * dy + dz * dz;
        double acc[13];
        for (int k = 0; k < 13; ++k) acc[k] = 0.0;
        double mr = 1.0, mi = 0.0;   // Q_m^m
        for (int m = 0; m <= P; ++m) {
            if (m > 0) {
                double f = -(double)(2 * m - 1);
                double nr = f * (dx * mr - dy * mi), ni = f * (dx * mi + dy * mr);
                mr = nr; mi = ni;
            }
            double pr = 0.0, pi = 0.0, cr = mr, ci = mi;
            for (int n = m; n <= P; ++n) {
                if (n > m) {
                    double a = (double)(2 * n - 1) * dz, b = (double)(n + m - 1) * rho2, d = 1.0 / (double)(n - m);
                    double nr = (a * cr - b * pr) * d, ni = (a * ci - b * pi) * d;
                    pr = cr; pi = ci; cr = nr; ci = ni;
                }
                const double* c = lc + 26 * ((n * (n + 1)) / 2 + m);
                for (int k = 0; k < 13; ++k) acc[k] += cr * c[2 * k] - ci * c[2 * k + 1];
            }
        }
        if (PS) tout[it] += -acc[0] * ONE_OVER_4PI;
        if (GS) for (int k = 1; k < 4; ++k) tout[(int64_t)k * nt + it] += acc[k] * ONE_OVER_4PI;
        if (HS) for (int k = 4; k < 13; ++k) tout[(int64_t)k * nt + it] += acc[k] * ONE_OVER_4PI;
    }
}

void run_l2p(Context& C, Tree& tt, double* tout, int64_t nt, int p, bool LH, bool PS, bool GS, bool HS) {
    ensure_tables(C, p);
    if (tt.n_leaves == 0) return;
    TablePtrs tb = ptrs(C);
    static DBuf<double>& coef = *new DBuf<double>();      // process lifetime
    coef.resize((size_t)tt.n_leaves * 26 * nh(p));
    if (LH)
        k_l2p_coef<true><<<(unsigned)tt.n_leaves, 64, 0, C.stream>>>(tt.leaf_index.data(), tt.expansions.data(), p, PS, GS, HS, tb.inv_fact, coef.data());
    else
        k_l2p_coef<false><<<(unsigned)tt.n_leaves, 64, 0, C.stream>>>(tt.leaf_index.data(), tt.expansions.data(), p, PS, GS, HS, tb.inv_fact, coef.data());
    LAUNCH_CHECK();
    k_l2p_eval<<<(unsigned)tt.n_leaves, 128, sizeof(double) * 26 * nh(p), C.stream>>>(
        tt.branches.data(), tt.leaf_index.data(), tt.x.data(), tt.y.data(), tt.z.data(), p, coef.data(), tout, nt, PS, GS, HS);
    LAUNCH_CHECK();
}

}  // namespace fmmhip
