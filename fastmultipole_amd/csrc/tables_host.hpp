// Host-side construction of the rotation and normalisation tables shared by the
// translation kernels (expansions.hip) and their lane-by-lane CPU emulation.
// Plain C++: no HIP types.
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

namespace fmmhip {
namespace tabh {

constexpr double PI_T = 3.141592653589793;
constexpr int nh(int p) { return ((p + 1) * (p + 2)) >> 1; }
constexpr int len_Hs(int p) { return ((p + 1) * (p + 2) * (p + 3)) / 6; }
constexpr int len_Ts(int p) { return p < 0 ? 0 : ((p + 1) * (p + 2) * (2 * p + 3)) / 6; }

// Host computation of the pi/2 Wigner tables, zeta/eta magnitudes and power
// normalisations (src/rotate.jl:95-223, 625-1004; src/translate.jl:366-448),
// in the reference's operation order.
inline std::vector<double> host_pnms(int p, double s, double c) {
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

inline void host_Hs(std::vector<double>& Hs, int pmax) {
    const double beta = PI_T / 2;
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

inline void host_mags(std::vector<double>& z, std::vector<double>& e, int pmax) {
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
inline double host_factorial(int k) {
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


// 1/(n-1)! and (n+1)! as the error tests' running loops form them (same operations, so the
// same doubles): start 1 and 2, then divide by k and multiply by k + 1 for k = 1..n-1
inline std::vector<double> host_test_factorials(int pm) {
    std::vector<double> tf(2 * (pm + 1), 0.0);
    for (int n = 1; n <= pm; ++n) {
        double a = 1.0, b = 2.0;
        for (int k = 1; k < n; ++k) {
            a /= (double)k;
            b *= (double)(k + 1);
        }
        tf[2 * n] = a;
        tf[2 * n + 1] = b;
    }
    return tf;
}

// dense symmetric copy of the pi/2 tables: A_n(a, b) = Hs_n(min(a, b), max(a, b)), n + 1 rows per
// degree, rows padded to an even length
inline std::vector<double> host_dense_A(const std::vector<double>& Hs, int pm) {
    std::vector<double> Ad;
    for (int n = 0; n <= pm; ++n) {
        const int s = (n + 2) & ~1;
        const double* Hn = &Hs[((n) * (n + 1) * (n + 2)) / 6];
        for (int a = 0; a <= n; ++a)
            for (int b = 0; b < s; ++b) {
                const int lo = a < b ? a : b, hi = a < b ? b : a;
                Ad.push_back(b <= n ? Hn[(hi * (hi + 1)) / 2 + lo] : 0.0);
            }
    }
    return Ad;
}

// power normalisations Mt (multipole) and Lt (local) per (n, m)
inline void host_power_norms(std::vector<double>& Mt, std::vector<double>& Lt, int pm) {
    Mt.assign(nh(pm), 0.0);
    Lt.assign(nh(pm), 0.0);
    Mt[0] = Lt[0] = 1.0;
    for (int n = 1, i = 1; n <= pm; ++n)
        for (int m = 0; m <= n; ++m, ++i) {
            double a = host_factorial(n + m), b = host_factorial(n - m);
            Mt[i] = std::sqrt(a * b / (double)(2 * n + 1));
            Lt[i] = std::sqrt(1.0 / (a * b * (double)(2 * n + 1)));
        }
}

}  // namespace tabh
}  // namespace fmmhip
