// Oracle (test infrastructure only): statement-by-statement FP64 restatement
// of FastMultipole's expansion arithmetic. Each function cites the reference
// lines it follows. Built with -O2 -ffp-contract=off (no FMA contraction, no
// fast-math), matching Julia's evaluation of a+b+c as (a+b)+c.
#include "oracle_math.hpp"
#include <cstring>

namespace orc {

// ------------------------------------------------------------ factorials
// Float64(factorial(big(k))) (src/translate.jl:399,442): exact big-integer
// product, then a single round-to-nearest-even conversion.
double factorial_to_double(int k) {
    std::vector<uint32_t> limbs(1, 1u);
    for (int f = 2; f <= k; ++f) {
        uint64_t carry = 0;
        for (auto& l : limbs) {
            uint64_t v = (uint64_t)l * (uint64_t)f + carry;
            l = (uint32_t)v;
            carry = v >> 32;
        }
        if (carry) limbs.push_back((uint32_t)carry);
    }
    int top = (int)limbs.size() - 1;
    int nbits = 32 * top + (32 - __builtin_clz(limbs[top]));
    if (nbits <= 53) {
        double r = 0.0;
        for (int i = top; i >= 0; --i) r = r * 4294967296.0 + (double)limbs[i];
        return r;
    }
    // extract the leading 64 bits and a sticky bit
    auto bit = [&](int b) -> int { return (limbs[b / 32] >> (b % 32)) & 1u; };
    uint64_t lead = 0;
    for (int b = nbits - 1; b >= nbits - 64; --b) lead = (lead << 1) | (uint64_t)bit(b);
    bool sticky = false;
    for (int b = nbits - 65; b >= 0; --b) if (bit(b)) { sticky = true; break; }
    uint64_t mant = lead >> 11;
    uint64_t rem = lead & 0x7FFull;
    bool round_up = rem > 0x400 || (rem == 0x400 && (sticky || (mant & 1)));
    if (round_up) mant += 1;
    return std::ldexp((double)mant, nbits - 53);
}

// ------------------------------------------------------------- harmonics
void regular_harmonics(Exp h, double rho, double theta, double phi, int P) {   // src/harmonics.jl:13-58
    double y = std::sin(theta), x = std::cos(theta);
    double fact = 1.0, pn = 1.0, rhom = 1.0;
    double i_ei_imag = std::sin(phi + PI / 2), i_ei_real = std::cos(phi + PI / 2);
    double i_eim_real = 1.0, i_eim_imag = 0.0;
    for (int m = 0; m <= P; ++m) {
        double p = pn;
        int i = harmonic_index(m, m);
        double rhom_p = rhom * p;
        h(1, 1, i) = rhom_p * i_eim_real;
        h(2, 1, i) = rhom_p * i_eim_imag;
        double p1 = p;
        p = x * (double)(2 * m + 1) * p1;
        rhom *= rho;
        double rhon = rhom;
        for (int n = m + 1; n <= P; ++n) {
            i = harmonic_index(n, m);
            rhon /= (double)(-(n + m));
            double rhon_p = rhon * p;
            h(1, 1, i) = rhon_p * i_eim_real;
            h(2, 1, i) = rhon_p * i_eim_imag;
            double p2 = p1;
            p1 = p;
            p = (x * (double)(2 * n + 1) * p1 - (double)(n + m) * p2) / (double)(n - m + 1);
            rhon *= rho;
        }
        rhom /= (double)(-(2 * m + 2) * (2 * m + 1));
        pn = -pn * fact * y;
        fact += 2;
        double tr = i_eim_real, ti = i_eim_imag;
        i_eim_real = tr * i_ei_real - ti * i_ei_imag;
        i_eim_imag = tr * i_ei_imag + ti * i_ei_real;
    }
}

void irregular_harmonics(Exp h, double rho, double theta, double phi, int P) {   // src/harmonics.jl:64-104
    double y = std::sin(theta), x = std::cos(theta);
    double fact = 1.0, pn = 1.0;
    double one_over_rho = 1.0 / rho;
    double rhom = one_over_rho;
    double i_ei_imag = std::sin(phi + PI / 2), i_ei_real = std::cos(phi + PI / 2);
    double i_eim_real = 1.0, i_eim_imag = 0.0;
    for (int m = 0; m <= P; ++m) {
        double p = pn;
        int i = harmonic_index(m, m);
        double rhom_p = rhom * p;
        h(1, 1, i) = rhom_p * i_eim_real;
        h(2, 1, i) = rhom_p * i_eim_imag;
        double p1 = p;
        p = x * (double)(2 * m + 1) * p1;
        rhom *= -one_over_rho;
        double rhon = -rhom;
        for (int n = m + 1; n <= P; ++n) {
            i = harmonic_index(n, m);
            double rhon_p = rhon * p;
            h(1, 1, i) = rhon_p * i_eim_real;
            h(2, 1, i) = rhon_p * i_eim_imag;
            double p2 = p1;
            p1 = p;
            p = (x * (double)(2 * n + 1) * p1 - (double)(n + m) * p2) / (double)(n - m + 1);
            rhon *= one_over_rho * (double)(n - m + 1);
        }
        pn *= -fact * y;
        fact += 2;
        double tr = i_eim_real, ti = i_eim_imag;
        i_eim_real = tr * i_ei_real - ti * i_ei_imag;
        i_eim_imag = tr * i_ei_imag + ti * i_ei_real;
    }
}

void cartesian_to_spherical(double x, double y, double z, double& r, double& theta, double& phi) {  // src/rotate.jl:9-21
    double x2y2 = x * x + y * y;
    double r2 = x2y2 + z * z;
    r = (r2 == 0.0) ? r2 : std::sqrt(r2);
    double z_r = z / r;
    if (r > 0) theta = x2y2 > 0 ? std::acos(z_r) : (z < 0 ? PI : 0.0);
    else theta = 0.0;
    phi = (x2y2 == 0.0) ? 0.0 : std::atan2(y, x);
}

// ------------------------------------------------------- rotation tables
int length_Hs(int p) { int n = p + 1; return (n * (n + 1) * (n + 2)) / 6; }       // src/rotate.jl:67-71
static int length_H(int n) { return ((n + 1) * (n + 2)) >> 1; }                    // :73-75
int length_Ts(int p) { return ((p + 1) * (p + 2) * (2 * p + 3)) / 6; }            // :227-230

std::vector<double> compute_Pnms(int p, double s, double c) {                     // :95-117
    std::vector<double> P(n_harmonics(p), 0.0);
    double Pn0 = 1.0, fact = 1.0;
    for (int m = 0; m <= p; ++m) {
        double Pnm = Pn0;
        P[harmonic_index(m, m) - 1] = Pnm;
        double tmp1 = Pnm;
        Pnm = c * (double)(2 * m + 1) * tmp1;
        for (int n = m + 1; n <= p; ++n) {
            P[harmonic_index(n, m) - 1] = Pnm;
            double tmp2 = tmp1;
            tmp1 = Pnm;
            Pnm = (c * (double)(2 * n + 1) * tmp1 - (double)(n + m) * tmp2) / (double)(n - m + 1);
        }
        Pn0 = -Pn0 * fact * s;
        fact += 2;
    }
    return P;
}

void compute_Hs(std::vector<double>& Hs, double beta, int n_start, int n_end) {   // :119-203
    double sb = std::sin(beta), cb = std::cos(beta);
    std::vector<double> Pnms = compute_Pnms(n_end + 1, sb, cb);
    auto Pn = [&](int n, int m) { return Pnms[harmonic_index(n, m) - 1]; };
    std::vector<double> H_np1(n_end + 2, 0.0), bnms(2 * (n_end + 1) + 1, 0.0), dnms(n_end + 1, 0.0);
    int i_start = length_Hs(n_start - 1) + 1;
    int i_end = i_start + length_H(n_start) - 1;
    for (int n = n_start; n <= n_end; ++n) {
        double* H = &Hs[i_start - 1];   // H[k] 1-based -> H[k-1]
        auto Hat = [&](int mp, int m) -> double& { return H[H_index(mp, m) - 1]; };
        {
            double one_m = 1.0, fact = 1.0;
            for (int m = 0; m <= n; ++m) {
                Hat(0, m) = one_m * fact * Pn(n, m);
                one_m *= -1;
                fact /= std::sqrt((double)((n + m + 1) * (n - m)));
            }
        }
        {
            double one_m = 1.0, fact = 1.0;
            for (int m = 0; m <= n + 1; ++m) {
                H_np1[m] = one_m * fact * Pn(n + 1, m);
                one_m *= -1;
                fact /= std::sqrt((double)((n + 1 - m) * (n + m + 2)));
            }
        }
        int b_n = n + 1;
        for (int m = -b_n; m <= -1; ++m)
            bnms[b_n + m] = -std::sqrt((double)(b_n - m - 1) / (double)(2 * b_n - 1) * (double)(b_n - m) / (double)(2 * b_n + 1));
        for (int m = 0; m <= b_n; ++m)
            bnms[b_n + m] = std::sqrt((double)(b_n - m - 1) / (double)(2 * b_n - 1) * (double)(b_n - m) / (double)(2 * b_n + 1));
        double one_m_cb_2 = (1.0 - cb) / 2, one_p_cb_2 = (1.0 + cb) / 2;
        auto B = [&](int k) { return bnms[k - 1]; };          // bnms[k] 1-based
        auto Hn1 = [&](int k) { return H_np1[k - 1]; };      // H_np1[k] 1-based
        for (int m = 1; m <= n; ++m) {
            double b__mm1 = B(b_n - m), b_mm1 = B(b_n + m), b_0 = B(b_n + 1);
            double a_n_m = std::sqrt((double)(n + 1 + m) / (double)(2 * n + 1) * (double)(n + 1 - m) / (double)(2 * n + 3));
            Hat(1, m) = (b__mm1 * one_m_cb_2 * Hn1(m + 2) - b_mm1 * one_p_cb_2 * Hn1(m) - a_n_m * sb * Hn1(m + 1)) / b_0;
        }
        for (int m = 0; m <= n; ++m) dnms[m] = std::sqrt((double)((n - m) * (n + m + 1))) / 2;
        auto D = [&](int k) { return dnms[k - 1]; };
        for (int mp = 1; mp <= n - 1; ++mp) {
            for (int m = mp + 1; m <= n - 1; ++m) {
                double val = D(mp) * Hat(mp - 1, m) - D(m) * Hat(mp, m - 1) + D(m + 1) * Hat(mp, m + 1);
                Hat(mp + 1, m) = val / D(mp + 1);
            }
            int m = n;
            double val = D(mp) * Hat(mp - 1, m) - D(m) * Hat(mp, m - 1);
            Hat(mp + 1, m) = val / D(mp + 1);
        }
        i_start = i_end + 1;
        i_end = i_start + length_H(n + 1) - 1;
    }
}

static double get_scalar(int m_mp) {   // src/rotate.jl:251-258
    if (m_mp & 1) m_mp += 1;
    return ((m_mp >> 1) % 2 == 0) ? 1.0 : -1.0;
}

void update_eimphis(double* e, double phi, int P) {   // src/rotate.jl:25-36 (e is 2 x (P+1))
    double ei_imag = std::sin(phi), ei_real = std::cos(phi);
    double er = 1.0, ei = 0.0;
    for (int m = 0; m <= P; ++m) {
        e[2 * m] = er;
        e[2 * m + 1] = ei;
        double t = er;
        er = er * ei_real - ei * ei_imag;
        ei = t * ei_imag + ei * ei_real;
    }
}

// one degree of the Wigner d-matrix assembly (shared by update_Ts and update_Ts_n)
static void Ts_degree(double* Ts, const double* Hs, double sb, double cb, double one_n, int n) {
    const double* Hp = Hs + length_Hs(n - 1);        // view(Hs_pi2, i_H_start:i_H_end), 0-based start
    double* T = Ts + length_Ts(n - 1);
    auto Hat = [&](int a, int b) { return Hp[H_index(a, b) - 1]; };
    for (int m = 0; m <= n; ++m) {
        double H_n_m_0 = Hat(0, m);
        double one_mp_odd = 1.0;
        int m_mp = m;
        double one_n_mp = one_n;
        for (int mp = 0; mp <= m; ++mp) {
            double H_n_mp_0 = Hat(0, mp);
            double vpos = 0.0, vneg = 0.0, s_prev = 0.0, c_prev = 1.0;
            bool even = (m_mp % 2) == 0;
            double scalar = get_scalar(m_mp);
            double one_n_mp_nu = -one_n_mp;
            for (int nu = 1; nu <= n; ++nu) {
                double c_nb = cb * c_prev - sb * s_prev;
                double s_nb = sb * c_prev + cb * s_prev;
                double sc = even ? scalar * c_nb : scalar * s_nb;
                double H_mp_nu = Hat(std::min(mp, nu), std::max(mp, nu));
                double H_m_nu = Hat(std::min(m, nu), std::max(m, nu));
                double val = H_mp_nu * H_m_nu * sc;
                vpos += val;
                vneg += val * one_n_mp_nu * one_mp_odd;
                c_prev = c_nb;
                s_prev = s_nb;
                one_n_mp_nu = -one_n_mp_nu;
            }
            vpos *= 2.0;
            vneg *= 2.0;
            scalar = even ? scalar : 0.0;
            double val = H_n_m_0 * H_n_mp_0 * scalar;
            vpos += val;
            vneg += val * one_n_mp * one_mp_odd;
            T[T_index(mp, m) - 1] = vpos;
            T[T_index(-mp, m) - 1] = vneg;
            one_n_mp = -one_n_mp;
            one_mp_odd = -one_mp_odd;
            m_mp += 1;
        }
    }
}

void update_Ts(double* Ts, const double* Hs, double beta, int p) {   // src/rotate.jl:260-346
    double sb = std::sin(beta), cb = std::cos(beta);
    double one_n = -1.0;
    Ts[0] = 1.0;
    for (int n = 1; n <= p; ++n) {
        Ts_degree(Ts, Hs, sb, cb, one_n, n);
        one_n = -one_n;
    }
}

void update_Ts_0(double* Ts, double beta, double& sb, double& cb, double& one_n) {   // :351-361
    Ts[0] = 1.0;
    sb = std::sin(beta);
    cb = std::cos(beta);
    one_n = -1.0;
}

void update_Ts_n(double* Ts, const double* Hs, double sb, double cb, double& one_n, int n) {   // :363-441
    Ts_degree(Ts, Hs, sb, cb, one_n, n);
    one_n = -one_n;
}

void rotate_z(Exp out, Exp in, double* e, double phi, int P, bool LH) {   // :448-483
    update_eimphis(e, phi, P);
    int iw = 1;
    out(1, 1, iw) = in(1, 1, iw);
    out(2, 1, iw) = in(2, 1, iw);
    if (LH) { out(1, 2, iw) = in(1, 2, iw); out(2, 2, iw) = in(2, 2, iw); }
    iw += 1;
    for (int n = 1; n <= P; ++n)
        for (int m = 0; m <= n; ++m) {
            double er = e[2 * m], ei = e[2 * m + 1];
            double xr = in(1, 1, iw), xi = in(2, 1, iw);
            out(1, 1, iw) = xr * er - xi * ei;
            out(2, 1, iw) = xr * ei + xi * er;
            if (LH) {
                xr = in(1, 2, iw); xi = in(2, 2, iw);
                out(1, 2, iw) = xr * er - xi * ei;
                out(2, 2, iw) = xr * ei + xi * er;
            }
            iw += 1;
        }
}

void back_rotate_z(Exp out, Exp in, const double* e, int P, bool LH) {   // :584-616
    int iw = 1;
    out(1, 1, iw) += in(1, 1, iw);
    out(2, 1, iw) += in(2, 1, iw);
    if (LH) { out(1, 2, iw) += in(1, 2, iw); out(2, 2, iw) += in(2, 2, iw); }
    iw += 1;
    for (int n = 1; n <= P; ++n)
        for (int m = 0; m <= n; ++m) {
            double er = e[2 * m], ei = -e[2 * m + 1];
            double xr = in(1, 1, iw), xi = in(2, 1, iw);
            out(1, 1, iw) += xr * er - xi * ei;
            out(2, 1, iw) += xr * ei + xi * er;
            if (LH) {
                xr = in(1, 2, iw); xi = in(2, 2, iw);
                out(1, 2, iw) += xr * er - xi * ei;
                out(2, 2, iw) += xr * ei + xi * er;
            }
            iw += 1;
        }
}

void zeta_sign(double mag, int mp, int m, double& re, double& im) {   // :644-656
    int mod = (std::abs(mp) - std::abs(m)) % 4;
    if (mod < 0) mod += 4;
    switch (mod) {
        case 0: re = mag; im = 0.0; break;
        case 1: re = 0.0; im = mag; break;
        case 2: re = -mag; im = 0.0; break;
        default: re = 0.0; im = -mag; break;
    }
}

void eta_sign(double mag, int mp, int m, double& re, double& im) {   // :959-971
    int mod = (std::abs(m) - std::abs(mp)) % 4;
    if (mod < 0) mod += 4;
    switch (mod) {
        case 0: re = mag; im = 0.0; break;
        case 1: re = 0.0; im = mag; break;
        case 2: re = -mag; im = 0.0; break;
        default: re = 0.0; im = -mag; break;
    }
}

void update_zetas_mag(std::vector<double>& z, int n_start, int n_end) {   // :658-689
    if (n_start == 0) { z[0] = 1.0; n_start = 1; }
    for (int n = n_start; n <= n_end; ++n) {
        double* zeta = &z[length_Ts(n - 1)];
        int i = 0;
        double mag_m = 1.0;
        for (int m = 0; m <= n; ++m) {
            double mag_mp = mag_m;
            for (int mp = 0; mp <= n; ++mp) {
                zeta[i] = mag_mp;
                int num = n + mp + 1;
                int denom = (mp + 1 < n) ? n - mp : 1;
                mag_mp *= std::sqrt((double)num / (double)denom);
                i += 1;
            }
            int num = (m < n - 1) ? n - m : 1;
            int denom = n + m + 1;
            mag_m *= std::sqrt((double)num / (double)denom);
        }
    }
}

void update_etas_mag(std::vector<double>& e, int n_start, int n_end) {   // :973-1004
    if (n_start == 0) { e[0] = 1.0; n_start = 1; }
    for (int n = n_start; n <= n_end; ++n) {
        double* eta = &e[length_Ts(n - 1)];
        int i = 0;
        double mag_m = 1.0;
        for (int m = 0; m <= n; ++m) {
            double mag_mp = mag_m;
            for (int mp = 0; mp <= n; ++mp) {
                eta[i] = mag_mp;
                int num = (mp + 1 < n) ? n - mp : 1;
                int denom = n + mp + 1;
                mag_mp *= std::sqrt((double)num / (double)denom);
                i += 1;
            }
            int num = n + m + 1;
            int denom = (m < n - 1) ? n - m : 1;
            mag_m *= std::sqrt((double)num / (double)denom);
        }
    }
}

// one degree of the y-axis rotation; `use_eta` selects the local normalisation
static void rotate_y_degree(Exp out, Exp in, const double* Ts, const double* mags, bool LH, int& i_z, int i_T,
                            int n, bool use_eta) {
    for (int m = 0; m <= n; ++m) {
        double v1r = 0, v1i = 0, v2r = 0, v2i = 0;
        i_z += n + 1;
        int i_z_mp = i_z;
        for (int mp = -n; mp <= 0; ++mp) {
            int i1 = std::min(-mp, m), i2 = std::max(-mp, m);
            double T = Ts[i_T + T_index(-i1, i2) - 1];
            double zr, zi;
            if (use_eta) eta_sign(mags[i_z_mp - 1] * T, mp, m, zr, zi);
            else zeta_sign(mags[i_z_mp - 1] * T, mp, m, zr, zi);
            int iw = harmonic_index(n, -mp);
            double Mr = in(1, 1, iw), Mi = -in(2, 1, iw);
            if (mp & 1) { Mr = -Mr; Mi = -Mi; }
            v1r += Mr * zr - Mi * zi;
            v1i += Mr * zi + Mi * zr;
            if (LH) {
                Mr = in(1, 2, iw); Mi = -in(2, 2, iw);
                if (mp & 1) { Mr = -Mr; Mi = -Mi; }
                v2r += Mr * zr - Mi * zi;
                v2i += Mr * zi + Mi * zr;
            }
            i_z_mp -= 1;
        }
        i_z_mp += 2;
        for (int mp = 1; mp <= n; ++mp) {
            int i1 = std::min(mp, m), i2 = std::max(mp, m);
            double T = Ts[i_T + T_index(i1, i2) - 1];
            double zr, zi;
            if (use_eta) eta_sign(mags[i_z_mp - 1] * T, mp, m, zr, zi);
            else zeta_sign(mags[i_z_mp - 1] * T, mp, m, zr, zi);
            int iw = harmonic_index(n, mp);
            double Mr = in(1, 1, iw), Mi = in(2, 1, iw);
            v1r += Mr * zr - Mi * zi;
            v1i += Mr * zi + Mi * zr;
            if (LH) {
                Mr = in(1, 2, iw); Mi = in(2, 2, iw);
                v2r += Mr * zr - Mi * zi;
                v2i += Mr * zi + Mi * zr;
            }
            i_z_mp += 1;
        }
        int i = harmonic_index(n, m);
        out(1, 1, i) += v1r;
        out(2, 1, i) += v1i;
        if (LH) { out(1, 2, i) += v2r; out(2, 2, i) += v2i; }
    }
}

void rotate_multipole_y_full(Exp out, Exp in, const double* Ts, const double* zetas, int P, bool LH) {   // :699-803
    zero_exp(out, P);
    int i_z = 0, i_T = 0;
    for (int n = 0; n <= P; ++n) {
        rotate_y_degree(out, in, Ts, zetas, LH, i_z, i_T, n, false);
        i_T += (n + 1) * (n + 1);
    }
}

void rotate_multipole_y_n(Exp out, Exp in, const double* Ts, const double* zetas, bool LH, int& i_z, int& i_T, int n) {  // :808-910
    int in0 = ((n * (n + 1)) >> 1) + 1;
    for (int k = in0; k <= in0 + n; ++k) { out(1, 1, k) = 0; out(2, 1, k) = 0; out(1, 2, k) = 0; out(2, 2, k) = 0; }
    rotate_y_degree(out, in, Ts, zetas, LH, i_z, i_T, n, false);
    i_T += (n + 1) * (n + 1);
}

void rotate_local_y_full(Exp out, Exp in, const double* Ts, const double* etas, int P, bool LH) {   // :1014-1113
    zero_exp(out, P);
    int i_z = 0, i_T = 0;
    for (int n = 0; n <= P; ++n) {
        rotate_y_degree(out, in, Ts, etas, LH, i_z, i_T, n, true);
        i_T += (n + 1) * (n + 1);
    }
}

// ---------------------------------------------------------- translations
void translate_multipole_z(Exp out, Exp in, double t, int p, bool LH) {   // src/translate.jl:6-50
    double _t = -t;
    int i = 1;
    for (int n = 0; n <= p; ++n)
        for (int m = 0; m <= n; ++m) {
            double v1r = 0, v1i = 0, v2r = 0, v2i = 0;
            double n_np_f = 1.0;
            int n_np = 1;
            double t_n_np = 1.0;
            for (int np = n; np >= m; --np) {
                double tmp = t_n_np / n_np_f;
                int iw = harmonic_index(np, m);
                v1r += tmp * in(1, 1, iw);
                v1i += tmp * in(2, 1, iw);
                if (LH) { v2r += tmp * in(1, 2, iw); v2i += tmp * in(2, 2, iw); }
                t_n_np *= _t;
                n_np_f *= (double)n_np;
                n_np += 1;
            }
            out(1, 1, i) = v1r; out(2, 1, i) = v1i;
            if (LH) { out(1, 2, i) = v2r; out(2, 2, i) = v2i; }
            i += 1;
        }
}

void transform_lh_multipole(Exp e, double r, int p) {   // :52-96
    int i_P_m = harmonic_index(p, 0);
    for (int m = 0; m <= p; ++m) {
        double c_nm1_r = e(1, 2, i_P_m), c_nm1_i = e(2, 2, i_P_m);
        int i_n_m = i_P_m;
        for (int n = p; n >= std::max(m, 1); --n) {
            double pr = e(1, 1, i_n_m), pi = e(2, 1, i_n_m);
            double cr = c_nm1_r, ci = c_nm1_i;
            double r_m_np1 = r * (double)m / (double)(n + 1);
            e(1, 1, i_n_m) = pr + r_m_np1 * ci;
            e(2, 1, i_n_m) = pi - r_m_np1 * cr;
            if (m < n) { c_nm1_r = e(1, 2, i_n_m - n); c_nm1_i = e(2, 2, i_n_m - n); }
            else { c_nm1_r = 0.0; c_nm1_i = 0.0; }
            double r_n = r / (double)n;
            e(1, 2, i_n_m) = cr + r_n * c_nm1_r;
            e(2, 2, i_n_m) = ci + r_n * c_nm1_i;
            i_n_m -= n;
        }
        i_P_m += 1;
    }
}

void transform_lh_local(Exp e, double r, int p) {   // :98-142
    int i_m_m = 1;
    for (int m = 0; m <= p; ++m) {
        double c_np1_r = e(1, 2, i_m_m), c_np1_i = e(2, 2, i_m_m);
        int i_n_m = i_m_m;
        for (int n = m; n <= p; ++n) {
            double cr = c_np1_r, ci = c_np1_i;
            if (n > 0) {
                double pr = e(1, 1, i_n_m), pi = e(2, 1, i_n_m);
                double r_m_n = r * (double)m / (double)n;
                e(1, 1, i_n_m) = pr - r_m_n * ci;
                e(2, 1, i_n_m) = pi + r_m_n * cr;
            }
            if (n < p) {
                c_np1_r = e(1, 2, i_n_m + n + 1);
                c_np1_i = e(2, 2, i_n_m + n + 1);
                double r_np1 = r / (double)(n + 1);
                e(1, 2, i_n_m) = cr - r_np1 * c_np1_r;
                e(2, 2, i_n_m) = ci - r_np1 * c_np1_i;
            }
            i_n_m += n + 1;
        }
        i_m_m += m + 2;
    }
}

void translate_multipole_to_local_z(Exp out, Exp in, double t, int p, bool LH) {   // :181-232
    double one_over_t = 1.0 / t;
    double n_f_t_np1 = one_over_t;
    int i = 1;
    for (int n = 0; n <= p; ++n) {
        double n_m_f = n_f_t_np1;
        int n_m = n;
        for (int m = 0; m <= n; ++m) {
            double v1r = 0, v1i = 0, v2r = 0, v2i = 0;
            double n_np_f = n_m_f;
            int n_np = n_m;
            for (int np = m; np <= p; ++np) {
                int iw = harmonic_index(np, m);
                v1r += n_np_f * in(1, 1, iw);
                v1i += n_np_f * in(2, 1, iw);
                if (LH) { v2r += n_np_f * in(1, 2, iw); v2i += n_np_f * in(2, 2, iw); }
                n_np += 1;
                n_np_f *= (double)n_np * one_over_t;
            }
            out(1, 1, i) = v1r; out(2, 1, i) = v1i;
            if (LH) { out(1, 2, i) = v2r; out(2, 2, i) = v2i; }
            i += 1;
            n_m += 1;
            n_m_f *= (double)n_m * one_over_t;
        }
        n_f_t_np1 *= (double)(n + 1) * one_over_t;
    }
}

// :238-295
static void translate_multipole_to_local_z_n(Exp out, Exp in, double one_over_t, double n_f_t_np1, int n, bool LH) {
    int i = ((n * (n + 1)) >> 1) + 1;
    double n_m_f = n_f_t_np1;
    int n_m = n;
    for (int m = 0; m <= n; ++m) {
        double v1r = 0, v1i = 0, v2r = 0, v2i = 0;
        double n_np_f = n_m_f;
        int n_np = n_m;
        for (int np = m; np <= n; ++np) {
            int iw = harmonic_index(np, m);
            v1r += n_np_f * in(1, 1, iw);
            v1i += n_np_f * in(2, 1, iw);
            if (LH) { v2r += n_np_f * in(1, 2, iw); v2i += n_np_f * in(2, 2, iw); }
            n_np += 1;
            n_np_f *= (double)n_np * one_over_t;
        }
        out(1, 1, i) = v1r; out(2, 1, i) = v1i;
        if (LH) { out(1, 2, i) = v2r; out(2, 2, i) = v2i; }
        i += 1;
        n_m += 1;
        n_m_f *= (double)n_m * one_over_t;
    }
}

// :302-362
static void translate_multipole_to_local_z_m01_n(Exp sw, double t, double one_over_t, bool LH, double n_f_t_np1, int n,
                                                 double& p0r, double& p1r, double& p1i, double& c1r, double& c1i) {
    double n_m_f = n_f_t_np1;
    int n_m = n;
    p0r = 0.0;
    double n_np_f = n_m_f;
    int n_np = n_m;
    for (int np = 0; np <= n; ++np) {
        int iw = harmonic_index(np, 0);
        p0r += n_np_f * sw(1, 1, iw);
        n_np += 1;
        n_np_f *= (double)n_np * one_over_t;
    }
    n_m += 1;
    n_m_f *= (double)n_m * one_over_t;
    p1r = 0; p1i = 0; c1r = 0; c1i = 0;
    n_np_f = n_m_f;
    n_np = n_m;
    for (int np = 1; np <= n; ++np) {
        int iw = harmonic_index(np, 1);
        p1r += n_np_f * sw(1, 1, iw);
        p1i += n_np_f * sw(2, 1, iw);
        if (LH) { c1r += n_np_f * sw(1, 2, iw); c1i += n_np_f * sw(2, 2, iw); }
        n_np += 1;
        n_np_f *= (double)n_np * one_over_t;
    }
    if (LH) {
        double t_over_n = t / (double)n;
        p1r -= t_over_n * c1i;
        p1i += t_over_n * c1r;
    }
}

// :455-502
static void local_power(Exp w, double r, int n, const std::vector<double>& Lt, bool LH, double& lp, double& lc, double& Nt) {
    lp = 0; lc = 0; Nt = 0;
    double c_np1_r = 0, c_np1_i = 0;
    int i = ((n * (n + 1)) >> 1) + 1;
    for (int m = n; m >= 0; --m) {
        Nt = Lt[i + m - 1];
        double pr = w(1, 1, i + m), pi = w(2, 1, i + m);
        if (LH) {
            double cor = w(1, 2, i + m), coi = w(2, 2, i + m);
            pr = pr - coi * r * (double)m / (double)n;
            pi = pi + cor * r * (double)m / (double)n;
            double cr = cor - c_np1_r * r / (double)(n + 1);
            double ci = coi - c_np1_i * r / (double)(n + 1);
            c_np1_r = cor; c_np1_i = coi;
            // (1 + m>0) parses as (1+m)>0 == true, i.e. a factor of one
            lc += (cr * cr + ci * ci) * Nt * Nt;
        }
        lp += (pr * pr + pi * pi) * Nt * Nt;
    }
    lp = std::sqrt(lp);
    if (LH) lc = std::sqrt(lc);
}

void Scratch::ensure(int p) {
    size_t H = n_harmonics(p);
    if (w1.size() < 4 * H) { w1.assign(4 * H, 0.0); w2.assign(4 * H, 0.0); w3.assign(4 * H, 0.0); }
    if (Ts.size() < (size_t)length_Ts(p)) Ts.assign(length_Ts(p), 0.0);
    if (eimphis.size() < (size_t)(2 * (p + 2))) eimphis.assign(2 * (p + 2), 0.0);
}

void multipole_to_multipole(Exp target, const double* tc, Exp source, const double* sc, Scratch& s, int p, bool LH) {  // :144-174
    s.ensure(p);
    Tables& T = tables();
    T.ensure(p);
    double r, th, ph;
    cartesian_to_spherical(tc[0] - sc[0], tc[1] - sc[1], tc[2] - sc[2], r, th, ph);
    Exp w1{s.w1.data()}, w2{s.w2.data()};
    rotate_z(w1, source, s.eimphis.data(), ph, p, LH);
    update_Ts(s.Ts.data(), T.Hs_pi2.data(), th, p);
    rotate_multipole_y_full(w2, w1, s.Ts.data(), T.zetas.data(), p, LH);
    translate_multipole_z(w1, w2, r, p, LH);
    if (LH) transform_lh_multipole(w1, r, p);
    rotate_multipole_y_full(w2, w1, s.Ts.data(), T.zetas.data(), p, LH);
    back_rotate_z(target, w2, s.eimphis.data(), p, LH);
}

void translate_local_z(Exp out, Exp in, double t, int p, bool LH) {   // :1851-1891
    double _t = -t;
    int i = 1;
    for (int n = 0; n <= p; ++n)
        for (int m = 0; m <= n; ++m) {
            double v1r = 0, v1i = 0, v2r = 0, v2i = 0;
            int n_np = 1;
            double t_n_nf = 1.0;
            for (int np = n; np <= p; ++np) {
                int iw = harmonic_index(np, m);
                v1r += t_n_nf * in(1, 1, iw);
                v1i += t_n_nf * in(2, 1, iw);
                if (LH) { v2r += t_n_nf * in(1, 2, iw); v2i += t_n_nf * in(2, 2, iw); }
                t_n_nf *= _t / (double)n_np;
                n_np += 1;
            }
            out(1, 1, i) = v1r; out(2, 1, i) = v1i;
            if (LH) { out(1, 2, i) = v2r; out(2, 2, i) = v2i; }
            i += 1;
        }
}

void local_to_local(Exp target, const double* tc, Exp source, const double* sc, Scratch& s, int p, bool LH) {   // :1896-1926
    s.ensure(p);
    Tables& T = tables();
    T.ensure(p);
    double r, th, ph;
    cartesian_to_spherical(tc[0] - sc[0], tc[1] - sc[1], tc[2] - sc[2], r, th, ph);
    Exp w1{s.w1.data()}, w2{s.w2.data()};
    rotate_z(w1, source, s.eimphis.data(), ph, p, LH);
    update_Ts(s.Ts.data(), T.Hs_pi2.data(), th, p);
    rotate_local_y_full(w2, w1, s.Ts.data(), T.etas.data(), p, LH);
    translate_local_z(w1, w2, r, p, LH);
    if (LH) transform_lh_local(w1, r, p);
    rotate_local_y_full(w2, w1, s.Ts.data(), T.etas.data(), p, LH);
    back_rotate_z(target, w2, s.eimphis.data(), p, LH);
}

static double minabs(double x, double y) { return (std::fabs(x) <= std::fabs(y)) ? x : y; }   // src/error.jl:3
static double md1(double dx, double rm, double rp) {   // src/error.jl:5-10
    double right = dx + rp, left = dx + rm;
    bool same_sign = right * left >= 0;
    return same_sign ? minabs(right, left) : 0.0;
}
void minimum_distance(const double* c1, const double* c2, const double* box2, double* out) {   // src/error.jl:12-20
    out[0] = md1(c2[0] - c1[0], -box2[0], box2[0]);
    out[1] = md1(c2[1] - c1[1], -box2[1], box2[1]);
    out[2] = md1(c2[2] - c1[2], -box2[2], box2[2]);
}

thread_local double last_m2l_margin = INFINITY;

static void note_margin(double lhs, double limit) {
    double m = std::fabs(lhs / limit - 1.0);
    if (!(m >= last_m2l_margin)) last_m2l_margin = m;   // NaN ratios count as zero margin
}

int multipole_to_local(Exp target, const BranchGeom& tb, Exp source, const BranchGeom& sb, Scratch& s, int p, bool LH,
                       const ErrorMethod& em, bool& error_success) {
    last_m2l_margin = INFINITY;
    s.ensure(p);
    Tables& TB = tables();
    TB.ensure(p);
    Exp w1{s.w1.data()}, w2{s.w2.data()}, w3{s.w3.data()};
    double r, th, ph;
    cartesian_to_spherical(tb.target_center[0] - sb.source_center[0], tb.target_center[1] - sb.source_center[1],
                           tb.target_center[2] - sb.source_center[2], r, th, ph);
    if (em.kind == 0) {   // fixed order, src/translate.jl:519-550
        rotate_z(w1, source, s.eimphis.data(), ph, p, LH);
        update_Ts(s.Ts.data(), TB.Hs_pi2.data(), th, p);
        rotate_multipole_y_full(w2, w1, s.Ts.data(), TB.zetas.data(), p, LH);
        translate_multipole_to_local_z(w1, w2, r, p, LH);
        if (LH) transform_lh_local(w1, r, p);
        rotate_local_y_full(w2, w1, s.Ts.data(), TB.etas.data(), p, LH);
        back_rotate_z(target, w2, s.eimphis.data(), p, LH);
        error_success = true;
        return p;
    }
    // error-controlled variants: RotatedCoefficients* :555-826, PowerPotential :831-1119, PowerGradient :1124-1412
    double eps = em.relative ? tb.max_influence * em.eps : em.eps;
    double d[3];
    minimum_distance(sb.source_center, tb.target_center, tb.target_box, d);
    double r_mp = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double r_l = tb.target_radius;
    double rinv = 1.0 / r, r_mp_inv = 1.0 / r_mp;
    double r_mp_inv_np2 = r_mp_inv * r_mp_inv * r_mp_inv;
    double r_l_nm1 = 1.0;
    // rotate_z_0!
    double* E = s.eimphis.data();
    double ei_imag = std::sin(ph), ei_real = std::cos(ph);
    double em_r = 1.0, em_i = 0.0;
    E[0] = em_r; E[1] = em_i;
    w1(1, 1, 1) = source(1, 1, 1); w1(2, 1, 1) = source(2, 1, 1);
    if (LH) { w1(1, 2, 1) = source(1, 2, 1); w1(2, 2, 1) = source(2, 2, 1); }
    double sbeta, cbeta, one_n;
    update_Ts_0(s.Ts.data(), th, sbeta, cbeta, one_n);
    int i_z = 0, i_T = 0;
    rotate_multipole_y_n(w2, w1, s.Ts.data(), TB.zetas.data(), LH, i_z, i_T, 0);
    double n_f_t_np1 = rinv * rinv, nm1_f_inv = 1.0, np1_f = 2.0;
    bool power = em.kind >= 2;
    double mp_pow = 0, mp_pow_c = 0, Mt_n0 = TB.Mt[0];
    (void)mp_pow_c;   // the reference reads only the next-degree chi power
    if (power) {
        double a = source(1, 1, 1), b = source(2, 1, 1);
        mp_pow = std::sqrt((a * a + b * b) * Mt_n0 * Mt_n0);
        if (LH) { a = source(1, 2, 1); b = source(2, 2, 1); mp_pow_c = std::sqrt((a * a + b * b) * Mt_n0 * Mt_n0); }
    }
    double eps_mp = 0, eps_l = 0;
    error_success = false;
    int P = p;
    const double limit = eps * 4 * PI;
    for (int n = 1; n <= p; ++n) {
        // rotate_z_n! / rotate_z_n_power!
        double tr = em_r;
        em_r = em_r * ei_real - em_i * ei_imag;
        em_i = tr * ei_imag + em_i * ei_real;
        E[2 * n] = em_r; E[2 * n + 1] = em_i;
        int iw = ((n * (n + 1)) >> 1) + 1;
        double Mt_next = TB.Mt[iw - 1], pow_next = 0, pow_c_next = 0;
        for (int m = 0; m <= n; ++m) {
            double er = E[2 * m], ei = E[2 * m + 1];
            double xr = source(1, 1, iw), xi = source(2, 1, iw);
            if (power) { double Nt = TB.Mt[iw - 1]; pow_next += (xr * xr + xi * xi) * Nt * Nt; }
            w1(1, 1, iw) = xr * er - xi * ei;
            w1(2, 1, iw) = xr * ei + xi * er;
            if (LH) {
                xr = source(1, 2, iw); xi = source(2, 2, iw);
                if (power) { double Nt = TB.Mt[iw - 1]; pow_c_next += (xr * xr + xi * xi) * Nt * Nt; }
                w1(1, 2, iw) = xr * er - xi * ei;
                w1(2, 2, iw) = xr * ei + xi * er;
            }
            iw += 1;
        }
        pow_next = std::sqrt(pow_next);
        pow_c_next = std::sqrt(pow_c_next);
        update_Ts_n(s.Ts.data(), TB.Hs_pi2.data(), sbeta, cbeta, one_n, n);
        rotate_multipole_y_n(w2, w1, s.Ts.data(), TB.zetas.data(), LH, i_z, i_T, n);
        if (em.kind == 1) {
            int in0 = ((n * (n + 1)) >> 1) + 1;
            double p0 = w2(1, 1, in0), p1r = w2(1, 1, in0 + 1), p1i = w2(2, 1, in0 + 1);
            eps_mp = (std::sqrt(p1r * p1r + p1i * p1i) + std::fabs(p0)) * np1_f * r_mp_inv_np2;
            if (LH) {
                double c1r = w2(1, 2, in0 + 1), c1i = w2(2, 2, in0 + 1);
                eps_mp += std::sqrt(c1r * c1r + c1i * c1i) * np1_f * r_mp_inv_np2 * r_mp;
            }
        } else if (em.kind == 3) {
            eps_mp = SQRT3 * mp_pow * r_mp_inv_np2 * np1_f / Mt_n0;
            if (LH) eps_mp += SQRT3 * (double)n * pow_c_next * r_mp_inv_np2 * np1_f / Mt_next;
        } else {
            eps_mp = SQRT3 * pow_next * r_mp_inv_np2 * r_mp * np1_f / (Mt_n0 * (double)(n + 1));
            if (LH) eps_mp += SQRT3 * (double)n * pow_c_next * r_mp_inv_np2 * r_mp * np1_f / (Mt_next * (double)(n + 1));
        }
        note_margin(eps_mp, limit);
        if (eps_mp <= limit) {
            if (em.kind == 1) {
                double q0, q1r, q1i, c1r, c1i;
                translate_multipole_to_local_z_m01_n(w2, r, rinv, LH, n_f_t_np1, n, q0, q1r, q1i, c1r, c1i);
                eps_l = (std::fabs(q1r) + std::fabs(q1i) + std::fabs(q0)) * r_l_nm1 * nm1_f_inv;
                if (LH) eps_l += std::sqrt(c1r * c1r + c1i * c1i) * r_l_nm1 * r_l * nm1_f_inv;
            } else {
                translate_multipole_to_local_z_n(w3, w2, rinv, n_f_t_np1, n, LH);
                double lp, lc, Lt_n0;
                local_power(w3, r, n, TB.Lt, LH, lp, lc, Lt_n0);
                if (em.kind == 3) {
                    eps_l = SQRT3 * lp * r_l_nm1 * nm1_f_inv / Lt_n0;
                    if (LH) eps_l += SQRT3 * (double)n * lc * r_l_nm1 * r_l * nm1_f_inv / (Lt_n0 * (double)n);
                } else {
                    eps_l = SQRT3 * lp * r_l_nm1 * r_l * nm1_f_inv / (Lt_n0 * (double)n);
                    if (LH) eps_l += SQRT3 * (double)n * lc * r_l_nm1 * r_l * nm1_f_inv / (Lt_n0 * (double)n);
                }
            }
            note_margin(eps_mp + eps_l * LOCAL_ERROR_SAFETY, limit);
            if (eps_mp + eps_l * LOCAL_ERROR_SAFETY <= limit) {
                P = n - 1 + (em.BE ? 1 : 0);
                error_success = true;
                break;
            }
        }
        if (em.kind != 1 || n < p) {
            n_f_t_np1 *= (double)(n + 1) * rinv;
            r_l_nm1 *= r_l;
            r_mp_inv_np2 *= r_mp_inv;
            nm1_f_inv /= (double)n;
            np1_f *= (double)(n + 1);
        }
        if (power) { mp_pow = pow_next; mp_pow_c = pow_c_next; Mt_n0 = Mt_next; }
    }
    translate_multipole_to_local_z(w1, w2, r, P, LH);
    if (LH) transform_lh_local(w1, r, P);
    rotate_local_y_full(w2, w1, s.Ts.data(), TB.etas.data(), P, LH);
    back_rotate_z(target, w2, s.eimphis.data(), P, LH);
    return P;
}

// ------------------------------------------------------------------- L2P
void evaluate_local(const double* dxv, Exp harmonics, std::vector<double>& gnm, Exp L, int p, bool LH,
                    bool PS, bool GS, bool HS, double& phi_out, double* v_out, double* h_out) {   // src/evaluate_expansions.jl:45-399
    double r, th, ph;
    cartesian_to_spherical(dxv[0], dxv[1], dxv[2], r, th, ph);
    regular_harmonics(harmonics, r, th, ph, p);
    int Hn = n_harmonics(p);
    if (gnm.size() < (size_t)(6 * Hn)) gnm.resize(6 * Hn);
    auto G = [&](int a, int b, int i) -> double& { return gnm[(a - 1) + 2 * ((b - 1) + 3 * (size_t)(i - 1))]; };
    double u = 0, vx = 0, vy = 0, vz = 0;
    if (HS) std::fill(gnm.begin(), gnm.begin() + 6 * Hn, 0.0);
    double vxx = 0, vxy = 0, vxz = 0, vyx = 0, vyy = 0, vyz = 0, vzx = 0, vzy = 0, vzz = 0;
    int i_n_m = 0;
    {   // n = 0, m = 0
        int n = 0;
        i_n_m += 1;
        double Rr = harmonics(1, 1, i_n_m), Ri = harmonics(2, 1, i_n_m);
        if (PS) u += Rr * L(1, 1, i_n_m) - Ri * L(2, 1, i_n_m);
        if (GS || HS) {
            double vxr = 0, vyr = 0, vzr = 0, vzi = 0;
            double p_np1_m_r = L(1, 1, i_n_m + n + 1), p_np1_m_i = L(2, 1, i_n_m + n + 1);
            double p_np1_mp1_r = L(1, 1, i_n_m + n + 2), p_np1_mp1_i = L(2, 1, i_n_m + n + 2);
            vxr -= p_np1_mp1_i;
            vyr -= p_np1_mp1_r;
            vzr -= p_np1_m_r;
            vzi -= p_np1_m_i;
            if (GS) { vx += vxr * Rr; vy += vyr * Rr; vz += vzr * Rr - vzi * Ri; }
            if (HS) { G(1, 1, i_n_m) = vxr; G(1, 2, i_n_m) = vyr; G(1, 3, i_n_m) = vzr; G(2, 3, i_n_m) = vzi; }
        }
    }
    for (int n = 1; n <= p; ++n) {
        i_n_m += 1;   // m = 0
        double Rr = harmonics(1, 1, i_n_m), Ri = harmonics(2, 1, i_n_m);
        if (PS && !LH) u += Rr * L(1, 1, i_n_m) - Ri * L(2, 1, i_n_m);
        if (GS || HS) {
            double vxr = 0, vyr = 0, vzr = 0, vzi = 0;
            if (n < p) {
                double p_np1_m_r = L(1, 1, i_n_m + n + 1), p_np1_m_i = L(2, 1, i_n_m + n + 1);
                double p_np1_mp1_r = L(1, 1, i_n_m + n + 2), p_np1_mp1_i = L(2, 1, i_n_m + n + 2);
                vxr -= p_np1_mp1_i;
                vyr -= p_np1_mp1_r;
                vzr -= p_np1_m_r;
                vzi -= p_np1_m_i;
            }
            if (LH) {
                double c_r = L(1, 2, i_n_m + 1), c_i = L(2, 2, i_n_m + 1);
                vxr += (double)n * c_r;
                vyr -= (double)n * c_i;
            }
            if (GS) { vx += vxr * Rr; vy += vyr * Rr; vz += vzr * Rr - vzi * Ri; }
            if (HS) { G(1, 1, i_n_m) = vxr; G(1, 2, i_n_m) = vyr; G(1, 3, i_n_m) = vzr; G(2, 3, i_n_m) = vzi; }
        }
        for (int m = 1; m <= n; ++m) {
            i_n_m += 1;
            Rr = harmonics(1, 1, i_n_m); Ri = harmonics(2, 1, i_n_m);
            if (PS && !LH) u += 2 * (Rr * L(1, 1, i_n_m) - Ri * L(2, 1, i_n_m));
            if (GS || HS) {
                double vxr = 0, vxi = 0, vyr = 0, vyi = 0, vzr = 0, vzi = 0;
                if (n < p) {
                    double a_r = L(1, 1, i_n_m + n), a_i = L(2, 1, i_n_m + n);
                    double b_r = L(1, 1, i_n_m + n + 1), b_i = L(2, 1, i_n_m + n + 1);
                    double c_r = L(1, 1, i_n_m + n + 2), c_i = L(2, 1, i_n_m + n + 2);
                    vxr -= (a_i + c_i) * 0.5;
                    vxi += (a_r + c_r) * 0.5;
                    vyr += (a_r - c_r) * 0.5;
                    vyi += (a_i - c_i) * 0.5;
                    vzr -= b_r;
                    vzi -= b_i;
                }
                if (LH) {
                    double cm_r = L(1, 2, i_n_m - 1), cm_i = L(2, 2, i_n_m - 1);
                    vxr -= (double)(n + m) * cm_r * 0.5;
                    vxi -= (double)(n + m) * cm_i * 0.5;
                    vyr -= (double)(n + m) * cm_i * 0.5;
                    vyi += (double)(n + m) * cm_r * 0.5;
                    if (m < n) {
                        double cp_r = L(1, 2, i_n_m + 1), cp_i = L(2, 2, i_n_m + 1);
                        vxr += (double)(n - m) * cp_r * 0.5;
                        vxi += (double)(n - m) * cp_i * 0.5;
                        vyr -= (double)(n - m) * cp_i * 0.5;
                        vyi += (double)(n - m) * cp_r * 0.5;
                    }
                    double c_r = L(1, 2, i_n_m), c_i = L(2, 2, i_n_m);
                    vzr += (double)m * c_i;
                    vzi -= (double)m * c_r;
                }
                vx += 2 * (vxr * Rr - vxi * Ri);
                vy += 2 * (vyr * Rr - vyi * Ri);
                vz += 2 * (vzr * Rr - vzi * Ri);
                if (HS) {
                    G(1, 1, i_n_m) = vxr; G(2, 1, i_n_m) = vxi;
                    G(1, 2, i_n_m) = vyr; G(2, 2, i_n_m) = vyi;
                    G(1, 3, i_n_m) = vzr; G(2, 3, i_n_m) = vzi;
                }
            }
        }
    }
    if (HS) {
        i_n_m = 0;
        for (int n = 0; n <= p - 1; ++n) {
            i_n_m += 1;
            double Rr = harmonics(1, 1, i_n_m), Ri = harmonics(2, 1, i_n_m);
            vxx += -G(2, 1, i_n_m + n + 2) * Rr;
            vyx += -G(1, 1, i_n_m + n + 2) * Rr;
            { double a = -G(1, 1, i_n_m + n + 1), b = -G(2, 1, i_n_m + n + 1); vzx += a * Rr - b * Ri; }
            vxy += -G(2, 2, i_n_m + n + 2) * Rr;
            vyy += -G(1, 2, i_n_m + n + 2) * Rr;
            { double a = -G(1, 2, i_n_m + n + 1), b = -G(2, 2, i_n_m + n + 1); vzy += a * Rr - b * Ri; }
            vxz += -G(2, 3, i_n_m + n + 2) * Rr;
            vyz += -G(1, 3, i_n_m + n + 2) * Rr;
            { double a = -G(1, 3, i_n_m + n + 1), b = -G(2, 3, i_n_m + n + 1); vzz += a * Rr - b * Ri; }
            for (int m = 1; m <= n; ++m) {
                i_n_m += 1;
                Rr = harmonics(1, 1, i_n_m); Ri = harmonics(2, 1, i_n_m);
                for (int c = 1; c <= 3; ++c) {
                    double xr = -(G(2, c, i_n_m + n) + G(2, c, i_n_m + n + 2)) * 0.5;
                    double xi = (G(1, c, i_n_m + n) + G(1, c, i_n_m + n + 2)) * 0.5;
                    double yr = (G(1, c, i_n_m + n) - G(1, c, i_n_m + n + 2)) * 0.5;
                    double yi = (G(2, c, i_n_m + n) - G(2, c, i_n_m + n + 2)) * 0.5;
                    double zr = -G(1, c, i_n_m + n + 1), zi = -G(2, c, i_n_m + n + 1);
                    double ax = 2 * (xr * Rr - xi * Ri), ay = 2 * (yr * Rr - yi * Ri), az = 2 * (zr * Rr - zi * Ri);
                    if (c == 1) { vxx += ax; vyx += ay; vzx += az; }
                    else if (c == 2) { vxy += ax; vyy += ay; vzy += az; }
                    else { vxz += ax; vyz += ay; vzz += az; }
                }
            }
        }
    }
    phi_out = -u * ONE_OVER_4PI;
    v_out[0] = vx * ONE_OVER_4PI; v_out[1] = vy * ONE_OVER_4PI; v_out[2] = vz * ONE_OVER_4PI;
    // SMatrix{3,3}(vxx, vxy, vxz, vyx, vyy, vyz, vzx, vzy, vzz): linear order as written
    double hh[9] = {vxx, vxy, vxz, vyx, vyy, vyz, vzx, vzy, vzz};
    for (int k = 0; k < 9; ++k) h_out[k] = hh[k] * ONE_OVER_4PI;
}

// ------------------------------------------------------------------- P2M
static void get_nm1(Exp h, int index, int n, int m, int i_nm1_m, double one_m, double* q) {   // src/bodytomultipole.jl:77-110
    // q = {mm1_r, mm1_i, m_r, m_i, mp1_r, mp1_i}
    if (m > 0) { q[0] = h(1, index, i_nm1_m - 1); q[1] = h(2, index, i_nm1_m - 1); }
    else if (n > 1) { q[0] = -one_m * h(1, index, i_nm1_m + 1); q[1] = one_m * h(2, index, i_nm1_m + 1); }
    else { q[0] = 0.0; q[1] = 0.0; }
    if (m < n) { q[2] = h(1, index, i_nm1_m); q[3] = h(2, index, i_nm1_m); } else { q[2] = 0.0; q[3] = 0.0; }
    if (m + 1 < n) { q[4] = h(1, index, i_nm1_m + 1); q[5] = h(2, index, i_nm1_m + 1); } else { q[4] = 0.0; q[5] = 0.0; }
}

static void get_n(Exp h, int index, int n, int m, int i, double one_m, double* q) {   // :48-74
    if (m > 0) { q[0] = h(1, index, i - 1); q[1] = h(2, index, i - 1); }
    else if (n == 0) { q[0] = 0.0; q[1] = 0.0; }
    else { q[0] = -one_m * h(1, index, i + 1); q[1] = one_m * h(2, index, i + 1); }
    q[2] = h(1, index, i); q[3] = h(2, index, i);
    if (m < n) { q[4] = h(1, index, i + 1); q[5] = h(2, index, i + 1); } else { q[4] = 0.0; q[5] = 0.0; }
}

static void calculate_q(Exp h, double xr, double xi, double er, double ei, double z, int p) {   // :112-139
    int i = 1;
    h(1, 1, i) = 1.0;
    h(2, 1, i) = 0.0;
    i += 1;
    for (int n = 1; n <= p; ++n) {
        int i_nm1_m = i - n;
        double one_m = 1.0;
        for (int m = 0; m <= n; ++m) {
            double q[6];
            get_nm1(h, 1, n, m, i_nm1_m, one_m, q);
            h(1, 1, i) = (-(xr * q[1] + xi * q[0]) - (er * q[5] + ei * q[4]) - z * q[2]) / (double)n;
            h(2, 1, i) = ((xr * q[0] - xi * q[1]) + (er * q[4] - ei * q[5]) - z * q[3]) / (double)n;
            i += 1;
            i_nm1_m += 1;
            one_m = -one_m;
        }
    }
}

static void source_to_dipole(Exp h, int i_dipole, int i_source, double qx, double qy, double qz, int p) {   // :251-278
    h(1, i_dipole, 1) = 0.0;
    h(2, i_dipole, 1) = 0.0;
    int i = 2;
    for (int n = 1; n <= p; ++n) {
        int i_nm1_m = i - n;
        double one_m = 1.0;
        for (int m = 0; m <= n; ++m) {
            double q[6];
            get_nm1(h, i_source, n, m, i_nm1_m, one_m, q);
            h(1, i_dipole, i) = -qx * 0.5 * (q[5] + q[1]) + qy * 0.5 * (q[4] - q[0]) - qz * q[2];
            h(2, i_dipole, i) = qx * 0.5 * (q[4] + q[0]) + qy * 0.5 * (q[5] - q[1]) - qz * q[3];
            i += 1;
            i_nm1_m += 1;
            one_m = -one_m;
        }
    }
}

// p_n^m / j_n^m of a triangle (src/bodytomultipole.jl:146-176): index 2, from q (index 1)
static void calculate_pj(Exp h, double xr, double xi, double er, double ei, double z, int p) {
    int i = 1;
    h(1, 2, i) = 1.0;
    h(2, 2, i) = 0.0;
    i += 1;
    for (int n = 1; n <= p; ++n) {
        int i_nm1_m = i - n;
        double one_m = 1.0;
        for (int m = 0; m <= n; ++m) {
            double q[6];
            get_nm1(h, 2, n, m, i_nm1_m, one_m, q);
            const double qr = h(1, 1, i), qi = h(2, 1, i);
            h(1, 2, i) = (-(xr * q[1] + xi * q[0]) - (er * q[5] + ei * q[4]) - z * q[2] + qr) / (double)(n + 1);
            h(2, 2, i) = ((xr * q[0] - xi * q[1]) + (er * q[4] - ei * q[5]) - z * q[3] + qi) / (double)(n + 1);
            i += 1;
            i_nm1_m += 1;
            one_m = -one_m;
        }
    }
}

// i_n^m of a triangle (:183-213): index 1, from j (index 2)
static void calculate_ib(Exp h, double xr, double xi, double er, double ei, double z, int p) {
    int i = 1;
    h(1, 1, i) = 0.5;
    h(2, 1, i) = 0.0;
    i += 1;
    for (int n = 1; n <= p; ++n) {
        int i_nm1_m = i - n;
        double one_m = 1.0;
        for (int m = 0; m <= n; ++m) {
            double q[6];
            get_nm1(h, 1, n, m, i_nm1_m, one_m, q);
            const double jr = h(1, 2, i), ji = h(2, 2, i);
            h(1, 1, i) = (-(xr * q[1] + xi * q[0]) - (er * q[5] + ei * q[4]) - z * q[2] + jr) / (double)(n + 2);
            h(2, 1, i) = ((xr * q[0] - xi * q[1]) + (er * q[4] - ei * q[5]) - z * q[3] + ji) / (double)(n + 2);
            i += 1;
            i_nm1_m += 1;
            one_m = -one_m;
        }
    }
}

// the three recurrences of a triangle x0 (relative to the expansion centre), x0+xu, x0+xv
static void panel_harmonics(Exp h, const double* x0, const double* xu, const double* xv, int p) {
    // xyz_to_xi_eta_z (:39-46) of each point, added componentwise as the reference does
    const double x0r = x0[0] * 0.5, x0i = x0[1] * 0.5, e0r = x0r, e0i = -x0i;
    const double xur = xu[0] * 0.5, xui = xu[1] * 0.5, eur = xur, eui = -xui;
    const double xvr = xv[0] * 0.5, xvi = xv[1] * 0.5, evr = xvr, evi = -xvi;
    calculate_q(h, x0r + xur, x0i + xui, e0r + eur, e0i + eui, x0[2] + xu[2], p);
    calculate_pj(h, x0r + xvr, x0i + xvi, e0r + evr, e0i + evi, x0[2] + xv[2], p);
    calculate_ib(h, x0r, x0i, e0r, e0i, x0[2], p);
}

static double cross_norm(const double* u, const double* v) {
    const double c[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    return std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
}

void p2m_panel_source(Exp M, Exp h, const double* x0, const double* xu, const double* xv, double strength, int p) {   // :722-774
    panel_harmonics(h, x0, xu, xv, p);
    const double Jq = cross_norm(xu, xv) * (-strength);
    int i = 1;
    double one_n = 1.0;
    for (int n = 0; n <= p; ++n) {
        double one_n_m = one_n;
        for (int m = 0; m <= n; ++m) {
            M(1, 1, i) += Jq * one_n_m * h(1, 1, i);
            M(2, 1, i) -= Jq * one_n_m * h(2, 1, i);
            i += 1;
            one_n_m = -one_n_m;
        }
        one_n = -one_n;
    }
}

void p2m_panel_dipole(Exp M, Exp h, const double* x0, const double* xu, const double* xv, const double* normal,
                      double strength, int p) {   // :776-831
    panel_harmonics(h, x0, xu, xv, p);
    source_to_dipole(h, 2, 1, -normal[0] * strength, -normal[1] * strength, -normal[2] * strength, p);
    const double J = cross_norm(xu, xv);
    int i = 1;
    double one_n = 1.0;
    for (int n = 0; n <= p; ++n) {
        double one_n_m = one_n;
        for (int m = 0; m <= n; ++m) {
            M(1, 1, i) += J * one_n_m * h(1, 2, i);
            M(2, 1, i) -= J * one_n_m * h(2, 2, i);
            i += 1;
            one_n_m = -one_n_m;
        }
        one_n = -one_n;
    }
}

void p2m_point_source(Exp M, Exp h, const double* dx, double q, int p) {   // :403-429
    double xr = dx[0] * 0.5, xi = dx[1] * 0.5, er = xr, ei = -xi, z = dx[2];
    double s = -q;
    calculate_q(h, xr, xi, er, ei, z, p);
    double one_n = 1.0;
    int i = 1;
    for (int n = 0; n <= p; ++n) {
        double one_n_m = one_n;
        for (int m = 0; m <= n; ++m) {
            M(1, 1, i) += h(1, 1, i) * one_n_m * s;
            M(2, 1, i) -= h(2, 1, i) * one_n_m * s;
            one_n_m = -one_n_m;
            i += 1;
        }
        one_n = -one_n;
    }
}

void p2m_point_dipole(Exp M, Exp h, const double* dx, const double* qv, int p) {   // :431-463
    double xr = dx[0] * 0.5, xi = dx[1] * 0.5, er = xr, ei = -xi, z = dx[2];
    calculate_q(h, xr, xi, er, ei, z, p);
    source_to_dipole(h, 2, 1, -qv[0], -qv[1], -qv[2], p);
    double one_n = 1.0;
    int i = 1;
    for (int n = 0; n <= p; ++n) {
        double one_n_m = one_n;
        for (int m = 0; m <= n; ++m) {
            M(1, 1, i) += h(1, 2, i) * one_n_m;
            M(2, 1, i) -= h(2, 2, i) * one_n_m;
            one_n_m = -one_n_m;
            i += 1;
        }
        one_n = -one_n;
    }
}

void p2m_point_vortex(Exp M, Exp h, const double* dx, const double* g, int p) {   // :465-489, :283-334
    double mx = -dx[0], my = -dx[1], mz = -dx[2];
    double xr = mx * 0.5, xi = my * 0.5, er = xr, ei = -xi, z = mz;
    calculate_q(h, xr, xi, er, ei, z, p);
    const double vx = g[0], vy = g[1], vz = g[2], mult = 1.0;
    int i = 1;
    for (int n = 0; n <= p; ++n) {
        double one_m = 1.0, one_np1 = 1.0 / (double)(n + 1);
        for (int m = 0; m <= n; ++m) {
            double q[6];
            get_n(h, 1, n, m, i, one_m, q);
            double a = (double)(n - m + 1) * 0.5, b = (double)(n + m + 1) * 0.5;
            M(1, 1, i) += mult * one_m * ((-vx * q[0] + vy * q[1]) * a + (vx * q[4] + vy * q[5]) * b - vz * (double)m * q[3]) * one_np1;
            M(2, 1, i) += mult * one_m * ((vx * q[1] + vy * q[0]) * a + (-vx * q[5] + vy * q[4]) * b - vz * (double)m * q[2]) * one_np1;
            i += 1;
            one_m = -one_m;
        }
    }
    i = 2;
    for (int n = 1; n <= p; ++n) {
        int i_nm1_m = i - n;
        double one_m = 1.0, one_over_n = 1.0 / (double)n;
        for (int m = 0; m <= n; ++m) {
            double q[6];
            get_nm1(h, 1, n, m, i_nm1_m, one_m, q);
            M(1, 2, i) -= mult * one_m * one_over_n * (0.5 * (-vy * q[0] - vx * q[1] + vy * q[4] - vx * q[5]) - vz * q[2]);
            M(2, 2, i) -= mult * one_m * one_over_n * (0.5 * (vy * q[1] - vx * q[0] - vy * q[5] - vx * q[4]) + vz * q[3]);
            i += 1;
            i_nm1_m += 1;
            one_m = -one_m;
        }
    }
}

// ---------------------------------------------------------------- tables
Tables& tables() {
    static Tables t;
    return t;
}

void Tables::ensure(int p_needed) {   // src/FastMultipole.jl:137-148 and src/fmm.jl:983-991
    int pn = std::max(p_needed, 21);
    if (pn <= p) return;
    Hs_pi2.assign(length_Hs(pn), 0.0);
    Hs_pi2[0] = 1.0;
    compute_Hs(Hs_pi2, PI / 2, 1, pn);
    zetas.assign(length_Ts(pn), 0.0);
    update_zetas_mag(zetas, 0, pn);
    etas.assign(length_Ts(pn), 0.0);
    update_etas_mag(etas, 0, pn);
    Mt.assign(n_harmonics(pn), 0.0);
    Lt.assign(n_harmonics(pn), 0.0);
    Mt[0] = 1.0;
    Lt[0] = 1.0;
    int i = 2;
    for (int n = 1; n <= pn; ++n)
        for (int m = 0; m <= n; ++m) {
            double a = factorial_to_double(n + m), b = factorial_to_double(n - m);
            Mt[i - 1] = std::sqrt(a * b / (double)(2 * n + 1));
            Lt[i - 1] = std::sqrt(1.0 / (a * b * (double)(2 * n + 1)));
            i += 1;
        }
    p = pn;
}

}  // namespace orc
