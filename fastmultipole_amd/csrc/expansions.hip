// Far-field operators on the GPU: P2M, M2M, M2L (fixed order and the
// error-controlled variants), L2L, L2P for FastMultipole's solid-harmonic
// expansions (src/bodytomultipole.jl, src/rotate.jl, src/translate.jl,
// src/evaluate_expansions.jl). Expansions keep the reference layout per branch:
// [H][phi/chi][re/im] doubles.
//
// Translations use the reference's rotation scheme: z-rotation, y-rotation, a
// translation along z, the Lamb-Helmholtz correction, then the inverse
// rotations. The y-rotations apply the reference's Wigner matrices in factorised
// form (two products with the fixed pi/2 tables around a diagonal phase), so no
// per-translation matrix is assembled: O(p^3) work per translation. One
// wavefront runs one translation at a time with its intermediates in LDS and
// its destination expansion in registers; workgroups of four wavefronts share
// the pi/2 tables in LDS. M2L: one workgroup per target box, its wavefronts
// take the box's sources in turn and the partial locals are summed once.
//
// P2M: one thread per output coefficient of a leaf, bodies staged in LDS;
// L2P: per-leaf coefficient sets for potential, field and field gradient, then
// one thread per target evaluating the regular solid harmonics by their
// Cartesian recurrences (no trigonometry per target).
//
// Compiled with -ffp-contract=off so that the error estimates that choose the
// expansion order follow the reference's arithmetic; the rotation and
// translation products use explicit fma.
#include "expansions_common.hpp"
#include "tables_host.hpp"
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <cmath>

namespace fmmhip {

static constexpr int PMAX = 30;                 // largest supported expansion order
static constexpr double ONE_OVER_4PI = 1.0 / (4.0 * 3.141592653589793);
static constexpr double SQRT3_D = 1.7320508075688772;
static constexpr double LOCAL_ERROR_SAFETY = 2.0;



// ============================================================== tables
struct Tables {
    int p = -1;
    DBuf<double> Hs, zeta, eta, Mt, Lt, inv_fact, fact;
    DBuf<double> Ad;     // dense, even-padded pi/2 tables per degree (rotation kernels)
    DBuf<double> tfac;   // per degree n: 1/(n-1)! and (n+1)! by the error tests' running loops
};

void ensure_tables(Context& C, int p) {
    using namespace tabh;
    if (p > PMAX) throw Error(FMMHIP_EINVAL, "expansion order above 30 is not supported");
    if (!C.tables) C.tables = new Tables();
    Tables& T = *C.tables;
    if (T.p >= PMAX) return;
    const int pm = PMAX;
    std::vector<double> Hs, z, e, Mt, Lt, invf(2 * pm + 4), fct(2 * pm + 4);
    host_Hs(Hs, pm);
    host_mags(z, e, pm);
    host_power_norms(Mt, Lt, pm);
    for (int k = 0; k < (int)invf.size(); ++k) {
        fct[k] = host_factorial(k);
        invf[k] = 1.0 / fct[k];
    }
    auto up = [&](DBuf<double>& d, const std::vector<double>& h) {
        d.resize(h.size());
        HIP_CHECK(hipMemcpy(d.data(), h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    };
    const std::vector<double> Ad = host_dense_A(Hs, pm), tf = host_test_factorials(pm);
    up(T.Hs, Hs); up(T.zeta, z); up(T.eta, e); up(T.Mt, Mt); up(T.Lt, Lt); up(T.inv_fact, invf); up(T.fact, fct); up(T.Ad, Ad);
    up(T.tfac, tf);
    T.p = pm;
}

void free_tables(Tables* t) { delete t; }

static TablePtrs ptrs(Context& C) {
    Tables& T = *C.tables;
    return {T.Hs.data(), T.zeta.data(), T.eta.data(), T.Mt.data(), T.Lt.data(), T.inv_fact.data(), T.fact.data(), T.Ad.data(),
            T.tfac.data()};
}

// ====================================================== rotation engine
// y-rotations use the reference's matrices T (update_Ts!, src/rotate.jl:260-441)
// with the zeta / eta magnitudes and the i^k signs of _rotate_multipole_y! /
// _rotate_local_y! (:699-1128), evaluated in factorised form instead of
// assembling T for each translation. With A_n(m, v) the degree-n pi/2 table
// (Hs_pi2, read as a symmetric matrix), w_0 = 1, w_v = 2 (v >= 1) and
// s_v = (-1)^(n+v), the entries of T are
//   T(m, +-m') = sum_v w_v A(m, v) A(m', v) (i^(m+m') e^{i v b} + (-i)^(m+m') e^{-i v b}) / 2 * (1 | s_v),
// and the magnitudes factor as zeta(m, m') = g(m') / g(m) = 1 / eta(m, m'),
// g(k) = sqrt((n+k)! (n-k)!) / n!. Substituting both into the rotation sum,
// the i^k signs cancel, and one rotation of degree n becomes
//   U_m'   = g(m')^(+-1) X_m'                                   (input scale)
//   SE_v   = sum_{m' even} A(m', v) U_m',   SO_v = sum_{m' odd} A(m', v) U_m'
//   P+_v   = (SE_v - A(0, v) U_0) - SO_v + s_v conj(SE_v + SO_v)
//   P-_v   = (SE_v - A(0, v) U_0) + SO_v + s_v conj(SE_v - SO_v)
//   Q+-_v  = w_v e^{+-i sigma v b} P+-_v,  R_v^(e,o) = Q+_v +- Q-_v
//   out_m  = g(m)^(-+1) / 2 * sum_v A(m, v) R_v^(parity of m)
// with sigma = +1 and g^+1 in / g^-1 out for zeta (multipole) rotations,
// sigma = -1 and g^-1 in / g^+1 out for eta (local) rotations. The back
// rotations of the reference reuse the same T (rotate.jl:927-930, 1115-1128), so
// the same algorithm serves both directions. Work per degree: two real
// (n+1) x (n+1) matrix-vector products per complex potential, no T assembly.
//
// One wavefront runs one translation. Lanes own work items of a "chunk" of
// consecutive degrees: item (n, j) covers the column pair v = 2j, 2j+1 in the
// first product and the output pair m = 2j, 2j+1 in the second. A chunk holds at
// most 64 items, so each product is a single pass of the wavefront.

static constexpr int RBUF = 1024;      // doubles of per-wavefront item buffer (64 items x 16)
static constexpr int MAXCH = 2;        // back-rotation chunks held in registers: orders up to 20
static constexpr int FWD_DEGREES = 64; // degrees per forward chunk of the error-controlled M2L (64 items bound it)
// error-controlled M2L passes: orders up to the first cap in one pass with more wavefronts per
// workgroup, the interactions needing more at the full order in a second
enum { M2L_PASS_ALL = 0, M2L_PASS_FIRST = 1, M2L_PASS_DEFERRED = 2 };
static constexpr int M2L_FIRST_CAP = 18;
static constexpr int SMALL_P = 4;      // fixed orders up to this run one M2L per lane (k_m2l_lane)
static constexpr size_t LDS_LIMIT = 160 * 1024;



// items of degrees < n (ceil((k+1)/2) per degree k)
__host__ __device__ inline int items_below(int n) { return ((n + 1) * (n + 1)) >> 2; }

// end degree of the chunk starting at n0: at most 64 items and at most `maxdeg` degrees
__host__ __device__ inline int chunk_end(int n0, int P, int maxdeg) {
    int n1 = n0;
    while (n1 + 1 <= P && n1 + 1 - n0 < maxdeg && items_below(n1 + 2) - items_below(n0) <= 64) ++n1;
    return n1;
}

// global item index g -> (n, j)
__device__ inline void item_nj(int g, int& n, int& j) {
    int k = (int)sqrtf(4.0f * (float)g + 3.0f) - 1;
    while (items_below(k + 1) <= g) ++k;
    while (k > 0 && items_below(k) > g) --k;
    n = k;
    j = g - items_below(k);
}

// Dense pi/2 tables: degree n is an (n+1) x s_n row-major matrix A_n(a, b), with
// the row length s_n = n+1 rounded up to even (zero padding), so a column pair
// (2j, 2j+1) is one aligned 16-byte read.
__host__ __device__ constexpr int a_stride(int n) { return (n + 2) & ~1; }
__host__ __device__ constexpr int a_base(int n) { return (n * (n + 1) * (2 * n + 1)) / 6 + ((n + 1) / 2) * ((n + 1) / 2); }
__host__ __device__ constexpr int len_A(int P) { return a_base(P + 1); }

struct WaveLds {
    double* X;     // 4H: expansion workspace [idx][phi/chi][re/im]
    double* R;     // max(RBUF, 4H): item buffer / translated expansion
    double* eim;   // 2(P+1): e^{i m phi}
    double* cs;    // 2(P+1): cos(v beta), sin(v beta)
    double* fac;   // 6(P+1): running factors of the error tests per degree
    double* pw;    // 2(P+1): multipole power per degree (phi, chi)
    double* F;     // 2(P+1): k!/t^(k+1), k = 0..2P (M2L z-translation factors)
    const double* Mt;   // shared: multipole power normalisations, nh(P)
    const double* Lt;   // shared: local power normalisations, nh(P)
    const double* tf;   // shared: 1/(n-1)!, (n+1)! per degree, 2(P+1)
};

__host__ __device__ inline int wave_lds_doubles(int P) {
    const int H4 = 4 * (((P + 1) * (P + 2)) >> 1);
    return H4 + (H4 > RBUF ? H4 : RBUF) + 14 * (P + 1);
}
// shared part: the pi/2 tables, then Mt, Lt and the test factors
__host__ __device__ inline int tables_len(int P) { return len_A(P) + 2 * nh(P) + 2 * (P + 1); }
static size_t translate_lds_bytes(int P, int nw) { return sizeof(double) * (size_t)(tables_len(P) + nw * wave_lds_doubles(P)); }

// wavefronts per workgroup: as many (8, 6 or 4) as the LDS holds
static int translate_waves(int P, int max_nw = 8) {
    for (int nw : {12, 10, 8, 6, 4})
        if (nw <= max_nw && translate_lds_bytes(P, nw) <= LDS_LIMIT) return nw;
    throw Error(FMMHIP_EINVAL, "expansion order too large for the translation kernels");
}

// wavefronts per workgroup of the wave M2L: as many wavefronts per CU as the LDS holds (at most 12,
// the register limit), where two or more resident workgroups count 1.3 times, since a workgroup that
// waits for its slowest wavefront then leaves the CU to another (measured on the 1M ring at an order
// cap of 14: 2 x 4 wavefronts 55.0 ms, 1 x 10 wavefronts 58.9 ms; at 20: 1 x 6 beats 2 x 2).
// Even wavefront counts only.
static int m2l_waves(int P) {
    int best = 0;
    double best_eff = 0.0;
    for (int nw : {12, 10, 8, 6, 4, 2}) {
        const size_t bytes = translate_lds_bytes(P, nw);
        if (bytes > LDS_LIMIT) continue;
        // resident workgroups, with a 4 KB margin (two 5-wavefront workgroups of 79 KB at an order
        // cap of 14 measured as one)
        const int wg = std::max(1, (int)((LDS_LIMIT - 4096) / bytes));
        const double eff = std::min(12, nw * wg) * (wg >= 2 ? 1.3 : 1.0);
        if (eff > best_eff) {
            best_eff = eff;
            best = nw;
        }
    }
    if (!best) throw Error(FMMHIP_EINVAL, "expansion order too large for the translation kernels");
    return best;
}

// launch set-up shared by the translation kernels: order check and the LDS limit
template <class K>
static size_t translate_setup(K kernel, int P, int nw) {
    int chunks = 0;
    for (int n0 = 0; n0 <= P; n0 = chunk_end(n0, P, 64) + 1) ++chunks;
    if (chunks > MAXCH) throw Error(FMMHIP_EINVAL, "expansion order too large for the translation kernels");
    const size_t sh = translate_lds_bytes(P, nw);
    HIP_CHECK(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    return sh;
}

__device__ inline WaveLds carve_wave(double* lds, int P) {
    WaveLds W;
    W.X = lds + tables_len(P) + (size_t)wave_id() * wave_lds_doubles(P);
    const int H4 = 4 * nh(P);
    W.R = W.X + H4;
    W.eim = W.R + (H4 > RBUF ? H4 : RBUF);
    W.cs = W.eim + 2 * (P + 1);
    W.fac = W.cs + 2 * (P + 1);
    W.pw = W.fac + 6 * (P + 1);
    W.F = W.pw + 2 * (P + 1);
    W.Mt = lds + len_A(P);
    W.Lt = W.Mt + nh(P);
    W.tf = W.Lt + nh(P);
    return W;
}

// copies the dense pi/2 tables, Mt, Lt and the test factors into LDS (whole
// workgroup, ends with a barrier)
__device__ inline const double* load_tables(double* lds, const TablePtrs& tb, int P) {
    const int n = len_A(P), h = nh(P);
    for (int i = threadIdx.x; i < n; i += blockDim.x) lds[i] = tb.Ad[i];
    for (int i = threadIdx.x; i < h; i += blockDim.x) {
        lds[n + i] = tb.Mt[i];
        lds[n + h + i] = tb.Lt[i];
    }
    for (int i = threadIdx.x; i < 2 * (P + 1); i += blockDim.x) lds[n + 2 * h + i] = tb.tfac[i];
    __syncthreads();
    return lds;
}

// Spherical angles of the translation vector (cartesian_to_spherical,
// src/rotate.jl:5-21) enter only through e^{i phi} and (cos b, sin b); they are
// taken from the Cartesian components directly (e^{i phi} = (x + i y)/rho,
// cos b = z/r, sin b = rho/r), equal to the reference's acos/atan2 route up to
// rounding. Lane k <= P then forms e^{i k phi} by repeated multiplication
// (update_eimphis!, :25-36) and cos/sin(k b) by the angle-addition recurrence of
// update_Ts! (:380-435). Returns r.
__device__ inline double fill_angles(const WaveLds& W, double x, double y, double z, int P) {
    const double x2y2 = x * x + y * y;
    const double r = sqrt(x2y2 + z * z);
    double cp = 1.0, sp = 0.0, cb = z < 0.0 ? -1.0 : 1.0, sb = 0.0;
    if (x2y2 > 0.0) {
        const double rho = sqrt(x2y2);
        cp = x / rho; sp = y / rho;
        cb = z / r; sb = rho / r;
    }
    const int l = lane_id();
    if (l <= P) {
        double er = 1.0, ei = 0.0, c = 1.0, s = 0.0;
        for (int k = 0; k < l; ++k) {
            const double t = er;
            er = er * cp - ei * sp;
            ei = t * sp + ei * cp;
            const double cn = cb * c - sb * s;
            const double sn = sb * c + cb * s;
            c = cn;
            s = sn;
        }
        W.eim[2 * l] = er;
        W.eim[2 * l + 1] = ei;
        W.cs[2 * l] = c;
        W.cs[2 * l + 1] = s;
    }
    return r;
}

// F[k] = k!/t^(k+1) for k = 0..2P by the product chain of
// translate_multipole_to_local_z! (src/translate.jl:181-232); lane k
__device__ inline void fill_m2l_factors(const WaveLds& W, double t, int P) {
    const int k = lane_id();
    if (k <= 2 * P) {
        const double ot = 1.0 / t;
        double f = ot, dk = 1.0;
        for (int j = 0; j < k; ++j) {
            f *= dk * ot;
            dk += 1.0;
        }
        W.F[k] = f;
    }
}

// X[n][m] = g_n(m)^(+-1) * (src[n][m] e^{i m phi}) for degrees n0..n1 (rotate_z!, src/rotate.jl:448-483)
// mags row m = 0 of zeta (g) or eta (1/g). With POW, also the power terms
// |src_n^m|^2 Mt_n^m^2 (phi, chi) of the unrotated coefficients into R.
template <bool LH, bool POW>
__device__ inline void zrot_in(const double* __restrict__ src, const WaveLds& W, const double* mags, const double* Mt,
                               int n0, int n1) {
    const int i0 = nh(n0 - 1), i1 = nh(n1);
    for (int i = i0 + lane_id(); i < i1; i += 64) {
        int n, m;
        idx_nm(i, n, m);
        const double g = mags[len_Ts(n - 1) + m];
        const double er = W.eim[2 * m], ei = W.eim[2 * m + 1];
        const double2 s01 = *(const double2*)(src + 4 * i);
        double2 o01;
        o01.x = g * (s01.x * er - s01.y * ei);
        o01.y = g * (s01.x * ei + s01.y * er);
        *(double2*)(W.X + 4 * i) = o01;
        double2 s23 = make_double2(0.0, 0.0);
        if (LH) {
            s23 = *(const double2*)(src + 4 * i + 2);
            double2 o23;
            o23.x = g * (s23.x * er - s23.y * ei);
            o23.y = g * (s23.x * ei + s23.y * er);
            *(double2*)(W.X + 4 * i + 2) = o23;
        }
        if (POW) {
            const double Nt = Mt[i];
            W.R[2 * (i - i0)] = (s01.x * s01.x + s01.y * s01.y) * Nt * Nt;
            W.R[2 * (i - i0) + 1] = LH ? (s23.x * s23.x + s23.y * s23.y) * Nt * Nt : 0.0;
        }
    }
}

// Item buffer of a rotation chunk in R: slot-major, 8 double2 slots per item (v half h,
// potential, m parity: slot 4h + 2 pot + parity) of the chunk's (at most 64) items, so the
// lanes of a wavefront write consecutive 16-byte words.
__device__ inline double2* rslot(const WaveLds& W, int slot, int item) { return (double2*)W.R + slot * 64 + item; }

// first product of the chunk [n0, n1] (degrees <= nlim): X -> R
template <bool LH>
__device__ inline void rot_step1(const WaveLds& W, const double* A, double sigma, int n0, int n1, int nlim) {
    const int k = lane_id();
    const int base = items_below(n0);
    if (k >= items_below(n1 + 1) - base) return;
    int n, j;
    item_nj(base + k, n, j);
    if (n > nlim) return;
    const int s = a_stride(n);
    const double* An = A + a_base(n) + 2 * j;   // row m' at An + m' s: (A(m', 2j), A(m', 2j+1))
    const double* Xn = W.X + 4 * nh(n - 1);
    double se0[4] = {0.0, 0.0, 0.0, 0.0}, so0[4] = {0.0, 0.0, 0.0, 0.0};
    double se1[4] = {0.0, 0.0, 0.0, 0.0}, so1[4] = {0.0, 0.0, 0.0, 0.0};
    int mp = 0;
    for (; mp + 1 <= n; mp += 2) {
        const double2 ae = *(const double2*)(An + mp * s);
        const double2 ao = *(const double2*)(An + (mp + 1) * s);
        const double2 xe0 = *(const double2*)(Xn + 4 * mp), xo0 = *(const double2*)(Xn + 4 * mp + 4);
        se0[0] = fma(ae.x, xe0.x, se0[0]); se0[1] = fma(ae.x, xe0.y, se0[1]);
        se1[0] = fma(ae.y, xe0.x, se1[0]); se1[1] = fma(ae.y, xe0.y, se1[1]);
        so0[0] = fma(ao.x, xo0.x, so0[0]); so0[1] = fma(ao.x, xo0.y, so0[1]);
        so1[0] = fma(ao.y, xo0.x, so1[0]); so1[1] = fma(ao.y, xo0.y, so1[1]);
        if (LH) {
            const double2 xe1 = *(const double2*)(Xn + 4 * mp + 2), xo1 = *(const double2*)(Xn + 4 * mp + 6);
            se0[2] = fma(ae.x, xe1.x, se0[2]); se0[3] = fma(ae.x, xe1.y, se0[3]);
            se1[2] = fma(ae.y, xe1.x, se1[2]); se1[3] = fma(ae.y, xe1.y, se1[3]);
            so0[2] = fma(ao.x, xo1.x, so0[2]); so0[3] = fma(ao.x, xo1.y, so0[3]);
            so1[2] = fma(ao.y, xo1.x, so1[2]); so1[3] = fma(ao.y, xo1.y, so1[3]);
        }
    }
    if (mp <= n) {   // last (even) m'
        const double2 ae = *(const double2*)(An + mp * s);
        const double2 xe0 = *(const double2*)(Xn + 4 * mp);
        se0[0] = fma(ae.x, xe0.x, se0[0]); se0[1] = fma(ae.x, xe0.y, se0[1]);
        se1[0] = fma(ae.y, xe0.x, se1[0]); se1[1] = fma(ae.y, xe0.y, se1[1]);
        if (LH) {
            const double2 xe1 = *(const double2*)(Xn + 4 * mp + 2);
            se0[2] = fma(ae.x, xe1.x, se0[2]); se0[3] = fma(ae.x, xe1.y, se0[3]);
            se1[2] = fma(ae.y, xe1.x, se1[2]); se1[3] = fma(ae.y, xe1.y, se1[3]);
        }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int v = 2 * j + h;
        if (v > n) break;
        const double* se = h ? se1 : se0;
        const double* so = h ? so1 : so0;
        const double a00 = An[h];   // A(0, v)
        const double sg = ((n + v) & 1) ? -1.0 : 1.0;
        const double w = v ? 2.0 : 1.0;
        const double c = w * W.cs[2 * v], sn = w * sigma * W.cs[2 * v + 1];
#pragma unroll
        for (int pot = 0; pot < (LH ? 2 : 1); ++pot) {
            const double ser = se[2 * pot], sei = se[2 * pot + 1], sor = so[2 * pot], soi = so[2 * pot + 1];
            const double dr = ser - a00 * Xn[2 * pot], di = sei - a00 * Xn[2 * pot + 1];
            // P+ = (SE - S0) - SO + s conj(SE + SO),  P- = (SE - S0) + SO + s conj(SE - SO)
            const double ppr = dr - sor + sg * (ser + sor), ppi = di - soi - sg * (sei + soi);
            const double pmr = dr + sor + sg * (ser - sor), pmi = di + soi - sg * (sei - soi);
            // Q+ = (c + i sn) P+,  Q- = (c - i sn) P-
            const double qpr = c * ppr - sn * ppi, qpi = c * ppi + sn * ppr;
            const double qmr = c * pmr + sn * pmi, qmi = c * pmi - sn * pmr;
            *rslot(W, 4 * h + 2 * pot, k) = make_double2(qpr + qmr, qpi + qmi);       // even m
            *rslot(W, 4 * h + 2 * pot + 1, k) = make_double2(qpr - qmr, qpi - qmi);   // odd m
        }
    }
}

// second product for item (n, j): outputs m0 = 2j, m1 = 2j+1 (scaled), o[0..3] and o[4..7]
template <bool LH>
__device__ inline void rot_step2(const WaveLds& W, const double* A, const double* mags, int n0, int n, int j, double o[8]) {
    const int s = a_stride(n);
    const double* An = A + a_base(n) + 2 * j;   // row v: (A(v, 2j), A(v, 2j+1)) = (A(2j, v), A(2j+1, v))
    const int kn = items_below(n) - items_below(n0);   // the degree's first item in the chunk
    double e[4] = {0.0, 0.0, 0.0, 0.0}, d[4] = {0.0, 0.0, 0.0, 0.0};
    for (int v = 0; v <= n; ++v) {
        const double2 a = *(const double2*)(An + v * s);
        const int kk = kn + (v >> 1), h4 = 4 * (v & 1);   // item v/2, half v&1
        const double2 re0 = *rslot(W, h4, kk), ro0 = *rslot(W, h4 + 1, kk);
        e[0] = fma(a.x, re0.x, e[0]); e[1] = fma(a.x, re0.y, e[1]);
        d[0] = fma(a.y, ro0.x, d[0]); d[1] = fma(a.y, ro0.y, d[1]);
        if (LH) {
            const double2 re1 = *rslot(W, h4 + 2, kk), ro1 = *rslot(W, h4 + 3, kk);
            e[2] = fma(a.x, re1.x, e[2]); e[3] = fma(a.x, re1.y, e[3]);
            d[2] = fma(a.y, ro1.x, d[2]); d[3] = fma(a.y, ro1.y, d[3]);
        }
    }
    const double* mg = mags + len_Ts(n - 1);
    const int m0 = 2 * j, m1 = 2 * j + 1;
    const double g0 = 0.5 * mg[m0 * (n + 1)], g1 = m1 <= n ? 0.5 * mg[m1 * (n + 1)] : 0.0;
    o[0] = g0 * e[0]; o[1] = g0 * e[1]; o[2] = g0 * e[2]; o[3] = g0 * e[3];
    o[4] = g1 * d[0]; o[5] = g1 * d[1]; o[6] = g1 * d[2]; o[7] = g1 * d[3];
}

// rotation of the chunk [n0, n1]: X (scaled input) -> X (rotated)
template <bool LH>
__device__ inline void rot_chunk_store(const WaveLds& W, const double* A, const double* mags, double sigma, int n0, int n1) {
    rot_step1<LH>(W, A, sigma, n0, n1, n1);
    wave_sync();
    const int k = lane_id();
    const int base = items_below(n0);
    if (k < items_below(n1 + 1) - base) {
        int n, j;
        item_nj(base + k, n, j);
        double o[8];
        rot_step2<LH>(W, A, mags, n0, n, j, o);
        double* x = W.X + 4 * (nh(n - 1) + 2 * j);
        *(double2*)x = make_double2(o[0], o[1]);
        if (LH) *(double2*)(x + 2) = make_double2(o[2], o[3]);
        if (2 * j + 1 <= n) {
            *(double2*)(x + 4) = make_double2(o[4], o[5]);
            if (LH) *(double2*)(x + 6) = make_double2(o[6], o[7]);
        }
    }
    wave_sync();
}

// back rotation (degrees <= nlim) + back z-rotation, accumulated into registers:
// acc[c] holds item (chunk c, lane) = outputs (n, 2j) and (n, 2j+1)
template <bool LH>
__device__ inline void rot_back_acc(const WaveLds& W, const double* A, const double* mags, double sigma, int P, int nlim,
                                    double (&acc)[MAXCH][8]) {
    int n0 = 0;
#pragma unroll
    for (int c = 0; c < MAXCH; ++c) {
        if (n0 <= nlim) {
            const int n1 = chunk_end(n0, P, 64);
            rot_step1<LH>(W, A, sigma, n0, n1, nlim);
            wave_sync();
            const int k = lane_id();
            const int base = items_below(n0);
            if (k < items_below(n1 + 1) - base) {
                int n, j;
                item_nj(base + k, n, j);
                if (n <= nlim) {
                    double o[8];
                    rot_step2<LH>(W, A, mags, n0, n, j, o);
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const int m = 2 * j + h;
                        if (m > n) break;
                        const double er = W.eim[2 * m], ei = -W.eim[2 * m + 1];   // back_rotate_z! (src/rotate.jl:584-616)
                        acc[c][4 * h] += o[4 * h] * er - o[4 * h + 1] * ei;
                        acc[c][4 * h + 1] += o[4 * h] * ei + o[4 * h + 1] * er;
                        if (LH) {
                            acc[c][4 * h + 2] += o[4 * h + 2] * er - o[4 * h + 3] * ei;
                            acc[c][4 * h + 3] += o[4 * h + 2] * ei + o[4 * h + 3] * er;
                        }
                    }
                }
            }
            wave_sync();
            n0 = n1 + 1;
        }
    }
}

// register accumulators -> expansion layout (every (n, m) <= P written exactly once)
template <bool ADD>
__device__ inline void acc_store(double* dst, int P, const double (&acc)[MAXCH][8]) {
    int n0 = 0;
#pragma unroll
    for (int c = 0; c < MAXCH; ++c) {
        if (n0 <= P) {
            const int n1 = chunk_end(n0, P, 64);
            const int k = lane_id();
            const int base = items_below(n0);
            if (k < items_below(n1 + 1) - base) {
                int n, j;
                item_nj(base + k, n, j);
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int m = 2 * j + h;
                    if (m > n) break;
                    double* d = dst + 4 * (nh(n - 1) + m);
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        if (ADD) d[q] += acc[c][4 * h + q];
                        else d[q] = acc[c][4 * h + q];
                    }
                }
            }
            n0 = n1 + 1;
        }
    }
}

enum { OP_M2M = 0, OP_M2L = 1, OP_L2L = 2 };

// z-translation output (n, m) (src/translate.jl:6-50, 181-232, 1851-1891)
template <int OP, bool LH>
__device__ inline void translate_one(const double* in, const double* F, double t, int P, int n, int m, double out[4]) {
    double v1r = 0, v1i = 0, v2r = 0, v2i = 0;
    if (OP == OP_M2M) {
        const double _t = -t;
        double n_np_f = 1.0, t_n_np = 1.0, d_np = 1.0;
        int iw = nh(n - 1) + m;
        for (int np = n; np >= m; --np) {
            const double tmp = t_n_np / n_np_f;
            const double2 a = *(const double2*)(in + 4 * iw);
            v1r = fma(tmp, a.x, v1r); v1i = fma(tmp, a.y, v1i);
            if (LH) { const double2 b = *(const double2*)(in + 4 * iw + 2); v2r = fma(tmp, b.x, v2r); v2i = fma(tmp, b.y, v2i); }
            t_n_np *= _t;
            n_np_f *= d_np;
            d_np += 1.0;
            iw -= np;
        }
    } else if (OP == OP_M2L) {
        int iw = nh(m - 1) + m;
        for (int np = m; np <= P; ++np) {   // L_n^m += (n+np)!/t^(n+np+1) M_np^m
            const double f = F[n + np];
            const double2 a = *(const double2*)(in + 4 * iw);
            v1r = fma(f, a.x, v1r); v1i = fma(f, a.y, v1i);
            if (LH) { const double2 b = *(const double2*)(in + 4 * iw + 2); v2r = fma(f, b.x, v2r); v2i = fma(f, b.y, v2i); }
            iw += np + 1;
        }
    } else {
        const double _t = -t;
        double t_n_nf = 1.0, d_np = 1.0;
        int iw = nh(n - 1) + m;
        for (int np = n; np <= P; ++np) {
            const double2 a = *(const double2*)(in + 4 * iw);
            v1r = fma(t_n_nf, a.x, v1r); v1i = fma(t_n_nf, a.y, v1i);
            if (LH) { const double2 b = *(const double2*)(in + 4 * iw + 2); v2r = fma(t_n_nf, b.x, v2r); v2i = fma(t_n_nf, b.y, v2i); }
            t_n_nf *= _t / d_np;
            d_np += 1.0;
            iw += np + 1;
        }
    }
    out[0] = v1r; out[1] = v1i; out[2] = v2r; out[3] = v2i;
}

// Lamb-Helmholtz corrections, elementwise from the untransformed coefficients
// (multipole: src/translate.jl:52-96; local: :98-142)
template <bool LOCAL>
__device__ inline void lh_one(const double* in, double r, int P, int n, int m, double out[4]) {
    const int i = (n * (n + 1)) / 2 + m;
    double pr = in[4 * i], pi = in[4 * i + 1], cr = in[4 * i + 2], ci = in[4 * i + 3];
    if (!LOCAL) {
        if (n >= 1) {
            double r_m_np1 = r * (double)m / (double)(n + 1);
            out[0] = pr + r_m_np1 * ci;
            out[1] = pi - r_m_np1 * cr;
            double c1r = 0, c1i = 0;
            if (m < n) { c1r = in[4 * (i - n) + 2]; c1i = in[4 * (i - n) + 3]; }
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
            double c1r = in[4 * (i + n + 1) + 2], c1i = in[4 * (i + n + 1) + 3];
            double r_np1 = r / (double)(n + 1);
            out[2] = cr - r_np1 * c1r;
            out[3] = ci - r_np1 * c1i;
        } else { out[2] = cr; out[3] = ci; }
    }
}

// z-translation X -> R, then Lamb-Helmholtz transform and input scale of the back
// rotation R -> X, for orders <= Pu
template <int OP, bool LH>
__device__ inline void translate_and_transform(const WaveLds& W, double r, int Pu, const double* back_mags) {
    const int H = nh(Pu);
    for (int i = lane_id(); i < H; i += 64) {
        int n, m;
        idx_nm(i, n, m);
        double o[4];
        translate_one<OP, LH>(W.X, W.F, r, Pu, n, m, o);
        *(double2*)(W.R + 4 * i) = make_double2(o[0], o[1]);
        if (LH) *(double2*)(W.R + 4 * i + 2) = make_double2(o[2], o[3]);
    }
    wave_sync();
    for (int i = lane_id(); i < H; i += 64) {
        int n, m;
        idx_nm(i, n, m);
        const double g = back_mags[len_Ts(n - 1) + m];
        if (LH) {
            double o[4];
            lh_one<OP != OP_M2M>(W.R, r, Pu, n, m, o);
            *(double2*)(W.X + 4 * i) = make_double2(g * o[0], g * o[1]);
            *(double2*)(W.X + 4 * i + 2) = make_double2(g * o[2], g * o[3]);
        } else {
            const double2 a = *(const double2*)(W.R + 4 * i);
            *(double2*)(W.X + 4 * i) = make_double2(g * a.x, g * a.y);
        }
    }
    wave_sync();
}

// complete fixed-order translation of one expansion, accumulated into acc
template <int OP, bool LH>
__device__ void translate_fixed(const double* __restrict__ src, double dx, double dy, double dz, const WaveLds& W,
                                const double* A, int P, const TablePtrs& tb, double (&acc)[MAXCH][8]) {
    const double r = fill_angles(W, dx, dy, dz, P);
    if (OP == OP_M2L) fill_m2l_factors(W, r, P);
    wave_sync();
    const double* fwd_mags = (OP == OP_L2L) ? tb.eta : tb.zeta;
    const double* back_mags = (OP == OP_M2M) ? tb.zeta : tb.eta;
    const double sf = (OP == OP_L2L) ? -1.0 : 1.0, sb = (OP == OP_M2M) ? 1.0 : -1.0;
    zrot_in<LH, false>(src, W, fwd_mags, tb.Mt, 0, P);
    wave_sync();
    for (int n0 = 0; n0 <= P;) {
        const int n1 = chunk_end(n0, P, 64);
        rot_chunk_store<LH>(W, A, fwd_mags, sf, n0, n1);
        n0 = n1 + 1;
    }
    translate_and_transform<OP, LH>(W, r, P, back_mags);
    rot_back_acc<LH>(W, A, back_mags, sb, P, P, acc);
}

// ================================================================= M2M
// Per level, one wavefront per child translates the child's multipole to its parent's centre into a
// per-child slot; a second launch adds each parent's children's slots, in child order, into the
// parent's expansion. A wavefront per parent translating its up to 8 children in turn left the top
// levels (1, 8, 64 ... parents) to a handful of wavefronts.
template <bool LH, int NW>
__global__ __launch_bounds__(64 * NW) void k_m2m_child(const Branch* __restrict__ br, const double* __restrict__ ex,
                                                       int64_t c0, int64_t c1, int P, TablePtrs tb, double* __restrict__ slot) {
    extern __shared__ double lds[];
    const double* A = load_tables(lds, tb, P);
    const WaveLds W = carve_wave(lds, P);
    const int64_t ic = c0 + (int64_t)blockIdx.x * NW + wave_id();
    if (ic > c1) return;
    const Branch& Ch = br[ic];
    const Branch& B = br[Ch.parent];
    const int H4 = 4 * nh(P);
    double acc[MAXCH][8] = {};
    const double dx = B.source_center[0] - Ch.source_center[0], dy = B.source_center[1] - Ch.source_center[1],
                 dz = B.source_center[2] - Ch.source_center[2];
    translate_fixed<OP_M2M, LH>(ex + ic * H4, dx, dy, dz, W, A, P, tb, acc);
    acc_store<false>(slot + (ic - c0) * H4, P, acc);
}

// every coefficient of every parent of a level: the sum of its children's slots, in child order
__global__ void k_m2m_sum(const Branch* __restrict__ br, double* __restrict__ ex, int64_t b0, int64_t b1, int64_t c0,
                          int H4, const double* __restrict__ slot) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (b1 - b0 + 1) * H4) return;
    const int64_t ib = b0 + t / H4;
    const int q = (int)(t % H4);
    const Branch& B = br[ib];
    if (B.n_branches == 0) return;
    double s = 0.0;
    for (int64_t c = B.child_first; c < B.child_first + B.n_branches; ++c) s += slot[(c - c0) * H4 + q];
    ex[ib * H4 + q] += s;
}

template <bool LH, int NW>
static void launch_m2m(Context& C, Tree& st, int p, const TablePtrs& tb) {
    const size_t sh = translate_setup(k_m2m_child<LH, NW>, p, NW);
    const int H4 = 4 * nh(p);
    int64_t widest = 1;
    for (size_t l = 1; l < st.level_first.size(); ++l) widest = std::max<int64_t>(widest, st.level_last[l] - st.level_first[l] + 1);
    C.m2m_slot.resize((size_t)widest * H4);
    for (int64_t l = (int64_t)st.level_first.size() - 2; l >= 0; --l) {
        const int64_t b0 = st.level_first[l], b1 = st.level_last[l];
        const int64_t c0 = st.level_first[l + 1], c1 = st.level_last[l + 1];
        k_m2m_child<LH, NW><<<(unsigned)((c1 - c0 + NW) / NW), 64 * NW, sh, C.stream>>>(st.branches.data(), st.expansions.data(),
                                                                                       c0, c1, p, tb, C.m2m_slot.data());
        LAUNCH_CHECK();
        const int64_t items = (b1 - b0 + 1) * H4;
        k_m2m_sum<<<(unsigned)((items + 255) / 256), 256, 0, C.stream>>>(st.branches.data(), st.expansions.data(), b0, b1, c0,
                                                                        H4, C.m2m_slot.data());
        LAUNCH_CHECK();
    }
}

void run_m2m(Context& C, Tree& st, int p, bool LH) {
    ensure_tables(C, p);
    TablePtrs tb = ptrs(C);
    const int nw = translate_waves(p);
    if (LH) { if (nw == 8) launch_m2m<true, 8>(C, st, p, tb); else if (nw == 6) launch_m2m<true, 6>(C, st, p, tb); else launch_m2m<true, 4>(C, st, p, tb); }
    else { if (nw == 8) launch_m2m<false, 8>(C, st, p, tb); else if (nw == 6) launch_m2m<false, 6>(C, st, p, tb); else launch_m2m<false, 4>(C, st, p, tb); }
}

// ================================================================= L2L
template <bool LH, int NW>
__global__ __launch_bounds__(64 * NW) void k_l2l_level(Branch* __restrict__ br, double* __restrict__ ex, int64_t b0,
                                                       int64_t b1, int P, TablePtrs tb) {
    extern __shared__ double lds[];
    const double* A = load_tables(lds, tb, P);
    const WaveLds W = carve_wave(lds, P);
    const int64_t ib = b0 + (int64_t)blockIdx.x * NW + wave_id();
    if (ib > b1) return;
    const Branch B = br[ib];
    const Branch& Pa = br[B.parent];
    const int H4 = 4 * nh(P);
    double acc[MAXCH][8] = {};
    double dx = B.target_center[0] - Pa.target_center[0], dy = B.target_center[1] - Pa.target_center[1],
           dz = B.target_center[2] - Pa.target_center[2];
    translate_fixed<OP_L2L, LH>(ex + B.parent * H4, dx, dy, dz, W, A, P, tb, acc);
    acc_store<true>(ex + ib * H4, P, acc);
}

template <bool LH, int NW>
static void launch_l2l(Context& C, Tree& tt, int p, const TablePtrs& tb) {
    const size_t sh = translate_setup(k_l2l_level<LH, NW>, p, NW);
    for (size_t l = 1; l < tt.level_first.size(); ++l) {
        int64_t b0 = tt.level_first[l], b1 = tt.level_last[l];
        const unsigned grid = (unsigned)((b1 - b0 + NW) / NW);
        k_l2l_level<LH, NW><<<grid, 64 * NW, sh, C.stream>>>(tt.branches.data(), tt.expansions.data(), b0, b1, p, tb);
        LAUNCH_CHECK();
    }
}

void run_l2l(Context& C, Tree& tt, int p, bool LH) {
    ensure_tables(C, p);
    TablePtrs tb = ptrs(C);
    const int nw = translate_waves(p);
    if (LH) { if (nw == 8) launch_l2l<true, 8>(C, tt, p, tb); else if (nw == 6) launch_l2l<true, 6>(C, tt, p, tb); else launch_l2l<true, 4>(C, tt, p, tb); }
    else { if (nw == 8) launch_l2l<false, 8>(C, tt, p, tb); else if (nw == 6) launch_l2l<false, 6>(C, tt, p, tb); else launch_l2l<false, 4>(C, tt, p, tb); }
}

// ================================================================= M2L

__device__ inline double md1(double dx, double rm, double rp) {   // src/error.jl:3-10
    double right = dx + rp, left = dx + rm;
    bool same = right * left >= 0;
    double mn = (fabs(right) <= fabs(left)) ? right : left;
    return same ? mn : 0.0;
}

// Geometry of the error-controlled multipole_to_local! methods
// (src/translate.jl:555-1412). The reference walks the degrees in order and
// updates running products after each degree; here lane n replays the same
// product sequence for its degree once per interaction (fac), and each degree's
// test is evaluated by its own lane.
struct ErrGeo {
    double limit, rinv, r, r_mp, r_mp_inv, r_l;
};

// running factors at the start of iteration n: n!/t^(n+1)-type chain, r_l^(n-1),
// r_mp^-(n+2), 1/(n-1)!, (n+1)!  -> fac[6n .. 6n+4]
__device__ inline void fill_factors(const WaveLds& W, const ErrGeo& G, int P) {
    const int n = lane_id();
    if (n < 1 || n > P) return;
    double n_f_t_np1 = G.rinv * G.rinv, r_l_nm1 = 1.0, r_mp_inv_np2 = G.r_mp_inv * G.r_mp_inv * G.r_mp_inv;
    for (int k = 1; k < n; ++k) {
        n_f_t_np1 *= (double)(k + 1) * G.rinv;
        r_l_nm1 *= G.r_l;
        r_mp_inv_np2 *= G.r_mp_inv;
    }
    double* f = W.fac + 6 * n;
    f[0] = n_f_t_np1; f[1] = r_l_nm1; f[2] = r_mp_inv_np2; f[3] = W.tf[2 * n]; f[4] = W.tf[2 * n + 1];
}

// degree powers of the chunk from the terms zrot_in left in R (summed in m order)
template <bool LH>
__device__ inline void chunk_powers(const WaveLds& W, int n0, int n1) {
    const int n = n0 + lane_id();
    if (n > n1) return;
    const double* t = W.R + 2 * (nh(n - 1) - nh(n0 - 1));
    double a = 0.0, b = 0.0;
    for (int m = 0; m <= n; ++m) {
        a += t[2 * m];
        if (LH) b += t[2 * m + 1];
    }
    W.pw[2 * n] = sqrt(a);
    W.pw[2 * n + 1] = sqrt(b);
}

// Tests of degrees nA..nB (rotated multipole of degrees <= nB in W.X). Returns
// the first degree that meets the tolerance, or -1.
template <bool LH>
__device__ int m2l_test_chunk(const WaveLds& W, const ErrGeo& G, const ErrParams& E, const TablePtrs& tb, int nA, int nB) {
    if (E.kind >= 2) {
        // degree-n part of the z-translation, all m (translate_multipole_to_local_z_n!, :238-295) -> R.
        // The weight of X[np][m] in degree n is the running product n!/t^(n+1) * prod_{k=n+1}^{n+np} (k/t),
        // the same chain whatever m is, so each degree's chain is formed once (lane n) and kept in R
        // after the results.
        const int i0 = nh(nA - 1), cnt = nh(nB) - i0;
        double* chain = W.R + 4 * cnt;
        {
            const int n = nA + lane_id();
            if (n <= nB) {
                double f = W.fac[6 * n];
                double* cf = chain + (nh(n - 1) - i0);
                cf[0] = f;
                for (int np = 1; np <= n; ++np) {
                    f *= (double)(n + np) * G.rinv;
                    cf[np] = f;
                }
            }
        }
        wave_sync();
        for (int k = lane_id(); k < cnt; k += 64) {
            int n, m;
            idx_nm(i0 + k, n, m);
            const double* cf = chain + (nh(n - 1) - i0);
            double v1r = 0, v1i = 0, v2r = 0, v2i = 0;
            int iw = nh(m - 1) + m;
            for (int np = m; np <= n; ++np) {
                const double f = cf[np];
                v1r += f * W.X[4 * iw]; v1i += f * W.X[4 * iw + 1];
                if (LH) { v2r += f * W.X[4 * iw + 2]; v2i += f * W.X[4 * iw + 3]; }
                iw += np + 1;
            }
            double* t = W.R + 4 * k;
            t[0] = v1r; t[1] = v1i; t[2] = v2r; t[3] = v2i;
        }
        wave_sync();
    }
    bool pass = false;
    const int n = nA + lane_id();
    if (n <= nB) {
        const double* F = W.fac + 6 * n;
        const double n_f_t_np1 = F[0], r_l_nm1 = F[1], r_mp_inv_np2 = F[2], nm1_f_inv = F[3], np1_f = F[4];
        const int in0 = nh(n - 1);
        double eps_mp;
        if (E.kind == 1) {
            const double p0 = W.X[4 * in0], p1r = W.X[4 * (in0 + 1)], p1i = W.X[4 * (in0 + 1) + 1];
            eps_mp = (sqrt(p1r * p1r + p1i * p1i) + fabs(p0)) * np1_f * r_mp_inv_np2;
            if (LH) {
                const double c1r = W.X[4 * (in0 + 1) + 2], c1i = W.X[4 * (in0 + 1) + 3];
                eps_mp += sqrt(c1r * c1r + c1i * c1i) * np1_f * r_mp_inv_np2 * G.r_mp;
            }
        } else {
            const double pw_prev = W.pw[2 * (n - 1)], pw_n = W.pw[2 * n], pc_n = W.pw[2 * n + 1];
            const double Mt_n0 = W.Mt[nh(n - 2)], Mt_next = W.Mt[in0];
            if (E.kind == 3) {
                eps_mp = SQRT3_D * pw_prev * r_mp_inv_np2 * np1_f / Mt_n0;
                if (LH) eps_mp += SQRT3_D * (double)n * pc_n * r_mp_inv_np2 * np1_f / Mt_next;
            } else {
                eps_mp = SQRT3_D * pw_n * r_mp_inv_np2 * G.r_mp * np1_f / (Mt_n0 * (double)(n + 1));
                if (LH) eps_mp += SQRT3_D * (double)n * pc_n * r_mp_inv_np2 * G.r_mp * np1_f / (Mt_next * (double)(n + 1));
            }
        }
        if (eps_mp <= G.limit) {
            double eps_l;
            if (E.kind == 1) {   // translate_multipole_to_local_z_m01_n (:302-362)
                double n_m_f = n_f_t_np1;
                int n_m = n;
                double p0r = 0.0, n_np_f = n_m_f;
                int n_np = n_m;
                for (int np = 0; np <= n; ++np) {
                    p0r += n_np_f * W.X[4 * ((np * (np + 1)) / 2)];
                    n_np += 1;
                    n_np_f *= (double)n_np * G.rinv;
                }
                n_m += 1;
                n_m_f *= (double)n_m * G.rinv;
                double p1r = 0, p1i = 0, c1r = 0, c1i = 0;
                n_np_f = n_m_f;
                n_np = n_m;
                for (int np = 1; np <= n; ++np) {
                    const int iw = (np * (np + 1)) / 2 + 1;
                    p1r += n_np_f * W.X[4 * iw];
                    p1i += n_np_f * W.X[4 * iw + 1];
                    if (LH) { c1r += n_np_f * W.X[4 * iw + 2]; c1i += n_np_f * W.X[4 * iw + 3]; }
                    n_np += 1;
                    n_np_f *= (double)n_np * G.rinv;
                }
                if (LH) {
                    const double t_over_n = G.r / (double)n;
                    p1r -= t_over_n * c1i;
                    p1i += t_over_n * c1r;
                }
                eps_l = (fabs(p1r) + fabs(p1i) + fabs(p0r)) * r_l_nm1 * nm1_f_inv;
                if (LH) eps_l += sqrt(c1r * c1r + c1i * c1i) * r_l_nm1 * G.r_l * nm1_f_inv;
            } else {   // local_power (:455-502) of the degree-n translation
                const double* tmp = W.R + 4 * (in0 - nh(nA - 1));
                double lp = 0, lc = 0, Nt = 0, cnr = 0, cni = 0;
                // the Lamb-Helmholtz corrections multiply by r/n and r/(n+1) formed once
                // (the reference divides per term: rounding-level differences)
                const double r_n = G.r / (double)n, r_np1 = G.r / (double)(n + 1);
                for (int m = n; m >= 0; --m) {
                    Nt = W.Lt[in0 + m];
                    double pr = tmp[4 * m], pi = tmp[4 * m + 1];
                    if (LH) {
                        const double cor = tmp[4 * m + 2], coi = tmp[4 * m + 3];
                        const double rm = r_n * (double)m;
                        pr = pr - coi * rm;
                        pi = pi + cor * rm;
                        const double cr = cor - cnr * r_np1;
                        const double ci = coi - cni * r_np1;
                        cnr = cor; cni = coi;
                        lc += (cr * cr + ci * ci) * Nt * Nt;
                    }
                    lp += (pr * pr + pi * pi) * Nt * Nt;
                }
                lp = sqrt(lp);
                if (LH) lc = sqrt(lc);
                if (E.kind == 3) {
                    eps_l = SQRT3_D * lp * r_l_nm1 * nm1_f_inv / Nt;
                    if (LH) eps_l += SQRT3_D * (double)n * lc * r_l_nm1 * G.r_l * nm1_f_inv / (Nt * (double)n);
                } else {
                    eps_l = SQRT3_D * lp * r_l_nm1 * G.r_l * nm1_f_inv / (Nt * (double)n);
                    if (LH) eps_l += SQRT3_D * (double)n * lc * r_l_nm1 * G.r_l * nm1_f_inv / (Nt * (double)n);
                }
            }
            pass = (eps_mp + eps_l * LOCAL_ERROR_SAFETY <= G.limit);
        }
    }
    const unsigned long long mask = __ballot(pass);
    wave_sync();
    return mask ? nA + __ffsll((long long)mask) - 1 : -1;
}

// one M2L interaction (fixed order when E.kind == 0) accumulated into acc; returns the order used
template <bool LH>
__device__ int m2l_pair(const double* __restrict__ src, const Branch& T, const double (&sc)[3], const WaveLds& W,
                        const double* A, int P, const TablePtrs& tb, const ErrParams& E, int& success,
                        double (&acc)[MAXCH][8], bool defer, int n_test0) {
    const double dx = T.target_center[0] - sc[0], dy = T.target_center[1] - sc[1], dz = T.target_center[2] - sc[2];
    const double r = fill_angles(W, dx, dy, dz, P);
    fill_m2l_factors(W, r, P);
    int Pu = P;
    success = 1;
    if (E.kind == 0) {
        wave_sync();
        zrot_in<LH, false>(src, W, tb.zeta, tb.Mt, 0, P);
        wave_sync();
        for (int n0 = 0; n0 <= P;) {
            const int n1 = chunk_end(n0, P, 64);
            rot_chunk_store<LH>(W, A, tb.zeta, 1.0, n0, n1);
            n0 = n1 + 1;
        }
    } else {
        ErrGeo G;
        const double eps = E.relative ? T.max_influence * E.eps : E.eps;
        G.limit = eps * 4 * PI_D;
        const double d0 = md1(dx, -T.target_box[0], T.target_box[0]);
        const double d1 = md1(dy, -T.target_box[1], T.target_box[1]);
        const double d2 = md1(dz, -T.target_box[2], T.target_box[2]);
        G.r = r;
        G.r_mp = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        G.r_l = T.target_radius;
        G.rinv = 1.0 / r;
        G.r_mp_inv = 1.0 / G.r_mp;
        fill_factors(W, G, P);
        wave_sync();
        success = 0;
        const bool power = E.kind >= 2;
        for (int n0 = 0; n0 <= P;) {
            const int n1 = chunk_end(n0, P, FWD_DEGREES);
            if (power) {
                zrot_in<LH, true>(src, W, tb.zeta, W.Mt, n0, n1);
                wave_sync();
                chunk_powers<LH>(W, n0, n1);
            } else {
                zrot_in<LH, false>(src, W, tb.zeta, tb.Mt, n0, n1);
            }
            wave_sync();
            rot_chunk_store<LH>(W, A, tb.zeta, 1.0, n0, n1);
            // degrees below n_test0 are known to fail (tested by the first pass)
            const int nA = max(n0 < 1 ? 1 : n0, n_test0);
            if (nA <= n1) {
                const int hit = m2l_test_chunk<LH>(W, G, E, tb, nA, n1);
                if (hit >= 0) {
                    Pu = hit - 1 + (E.BE ? 1 : 0);
                    success = 1;
                    break;
                }
            }
            n0 = n1 + 1;
        }
        if (!success && defer) return -1;   // left for the pass at the full order
    }
    translate_and_transform<OP_M2L, LH>(W, r, Pu, tb.eta);
    rot_back_acc<LH>(W, A, tb.eta, -1.0, P, Pu, acc);
    return Pu;
}

// One workgroup per target box: its wavefronts take the box's M2L sources in
// turn, accumulate in registers, and the partial locals are summed through LDS
// and added to the box's expansion once.
template <bool LH, int NW>
__global__ __launch_bounds__(64 * NW) void k_m2l(const int32_t* __restrict__ m2l_targets, const int64_t* __restrict__ m2l_offsets,
                                                 const int32_t* __restrict__ m2l_s, const Branch* __restrict__ tbr,
                                                 const Branch* __restrict__ sbr, const double* __restrict__ sex,
                                                 double* __restrict__ tex, int P, int PE, int mode, ErrParams E,
                                                 int32_t* __restrict__ orders, int32_t* __restrict__ pmax,
                                                 int32_t* __restrict__ all_ok, int32_t* __restrict__ box_ok, TablePtrs tb) {
    extern __shared__ double lds[];
    const int64_t l0 = m2l_offsets[blockIdx.x], l1 = m2l_offsets[blockIdx.x + 1];
    if (mode == M2L_PASS_DEFERRED) {   // only the interactions the first pass left (order -1)
        int pending = 0;
        for (int64_t li = l0 + threadIdx.x; li < l1; li += 64 * NW) pending |= orders[li] < 0;
        if (!__syncthreads_or(pending)) return;
    }
    const double* A = load_tables(lds, tb, P);
    const WaveLds W = carve_wave(lds, P);
    const int32_t it = m2l_targets[blockIdx.x];
    const int H4 = 4 * nh(P), HE4 = 4 * nh(PE);   // coefficients worked on / stored per expansion
    const Branch T = tbr[it];
    double acc[MAXCH][8] = {};
    int local_max = 0, local_ok = 1;
    // the next interaction's source index and centre are loaded one iteration ahead
    int64_t li = l0 + wave_id();
    int32_t js_next = 0;
    double sc_next[3] = {0.0, 0.0, 0.0};
    if (li < l1) {
        js_next = m2l_s[li];
        for (int d = 0; d < 3; ++d) sc_next[d] = sbr[js_next].source_center[d];
    }
    for (; li < l1; li += NW) {
        const int32_t js = js_next;
        const double sc[3] = {sc_next[0], sc_next[1], sc_next[2]};
        if (li + NW < l1) {
            js_next = m2l_s[li + NW];
            for (int d = 0; d < 3; ++d) sc_next[d] = sbr[js_next].source_center[d];
        }
        if (mode == M2L_PASS_DEFERRED && orders[li] >= 0) continue;
        int ok = 1;
        // the deferred pass redoes interactions whose tests of degrees 1 .. cap all failed in the first
        // pass: each degree's test is independent, so it resumes testing at cap + 1
        const int Pu = m2l_pair<LH>(sex + (int64_t)js * HE4, T, sc, W, A, P, tb, E, ok, acc, mode == M2L_PASS_FIRST,
                                    mode == M2L_PASS_DEFERRED ? M2L_FIRST_CAP + 1 : 1);
        if (lane_id() == 0) orders[li] = Pu;
        if (Pu < 0) continue;
        local_max = max(local_max, Pu);
        local_ok = local_ok && ok;
    }
    acc_store<false>(W.X, P, acc);
    __syncthreads();
    const double* base = lds + tables_len(P);
    const int stride = wave_lds_doubles(P);
    for (int i = threadIdx.x; i < H4; i += 64 * NW) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) s += base[(size_t)w * stride + i];
        tex[(int64_t)it * HE4 + i] += s;
    }
    if (lane_id() == 0) {
        atomicMax(pmax, local_max);
        if (!local_ok) {
            atomicAnd(all_ok, 0);
            if (box_ok) box_ok[blockIdx.x] = 0;
        }
    }
}

template <bool LH, int NW>
static void launch_m2l(Context& C, int p, int pe, int mode, const Lists& Ls, const Tree& st, Tree& tt_dst, const ErrParams& E,
                       int32_t* flags, const TablePtrs& tb) {
    const size_t sh = translate_setup(k_m2l<LH, NW>, p, NW);
    k_m2l<LH, NW><<<(unsigned)Ls.n_m2l_targets, 64 * NW, sh, C.stream>>>(
        Ls.m2l_targets.data(), Ls.m2l_offsets.data(), Ls.m2l_s.data(), tt_dst.branches.data(), st.branches.data(),
        st.expansions.data(), tt_dst.expansions.data(), p, pe, mode, E, Ls.m2l_order.data(), flags, flags + 1,
        Ls.track_box_ok ? Ls.m2l_box_ok.data() : nullptr, tb);
    LAUNCH_CHECK();
}

// the wave M2L at order cap p over arrays of order pe, in the given pass mode
template <bool LH>
static void run_m2l_wave(Context& C, int p, int pe, int mode, const Lists& Ls, const Tree& st, Tree& tt_dst, const ErrParams& E,
                         int32_t* flags, const TablePtrs& tb) {
    const int nw = m2l_waves(p);
    switch (nw) {
        case 2: launch_m2l<LH, 2>(C, p, pe, mode, Ls, st, tt_dst, E, flags, tb); break;
        case 3: launch_m2l<LH, 3>(C, p, pe, mode, Ls, st, tt_dst, E, flags, tb); break;
        case 12: launch_m2l<LH, 12>(C, p, pe, mode, Ls, st, tt_dst, E, flags, tb); break;
        case 10: launch_m2l<LH, 10>(C, p, pe, mode, Ls, st, tt_dst, E, flags, tb); break;
        case 5: launch_m2l<LH, 5>(C, p, pe, mode, Ls, st, tt_dst, E, flags, tb); break;
        case 4: launch_m2l<LH, 4>(C, p, pe, mode, Ls, st, tt_dst, E, flags, tb); break;
        case 8: launch_m2l<LH, 8>(C, p, pe, mode, Ls, st, tt_dst, E, flags, tb); break;
        case 6: launch_m2l<LH, 6>(C, p, pe, mode, Ls, st, tt_dst, E, flags, tb); break;
        default: launch_m2l<LH, 4>(C, p, pe, mode, Ls, st, tt_dst, E, flags, tb); break;
    }
}

// ============================================ M2L on matrix cores (FP64 MFMA)
// The factorised rotation above applies the same pi/2 table A_n to every
// interaction, so a batch of interactions turns each of its products into a
// dense matrix product with a shared left operand: the batch's coefficients
// form the columns of X (16 columns = BP interactions x NC real components:
// phi re/im and, with Lamb-Helmholtz terms, chi re/im), and v_mfma_f64_16x16x4f64
// multiplies 16-row tiles of A_n into them. The z-translation is a shared
// product too: F(n+k) = (n+k)! t^-(n+1) t^-k, so its Hankel matrix (n+k)! is
// common and the t powers scale the columns before and the rows after it.
// Everything else (z-rotations, the P+- / Q+- combination, Lamb-Helmholtz
// terms, the error tests) is elementwise per column on the vector ALU.
//
// Lane layout of v_mfma_f64_16x16x4f64 (D = A B + C, 16 x 4 times 4 x 16): lane l
// supplies A(l % 16, l / 16) and B(l / 16, l % 16) and holds D(l / 16 + 4 r, l % 16)
// in element r of its result.
//
// One workgroup per target box: it takes the box's sources in batches of BP
// consecutive list entries, so every column of a batch accumulates into the same
// local expansion, and its wavefronts share each batch (one degree or one column
// m at a time each); the tables live in LDS. The error-controlled order is found in groups of
// degrees: the group is rotated, then one lane per (interaction, degree) runs
// the reference's tests for that degree (the tests of m2l_test_chunk), and each
// interaction takes its first passing degree. Every interaction is then
// translated and rotated back at its own order: its columns are masked above
// that order, so each result is the one of an M2L truncated there.

typedef double mf4 __attribute__((ext_vector_type(4)));

__device__ inline mf4 mfma_f64(double a, double b, mf4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

enum : int { MF_EMPTY = -1, MF_SEARCHING = -2 };
static constexpr int MF_GROUP = 8;   // degrees rotated between two rounds of error tests

// LDS of one workgroup: the tables every product reads, the batch and the
// target's accumulator, and one product buffer per wavefront
struct MfLds {
    double* Hs;    // len_Hs(P): pi/2 tables (packed lower triangles, read as symmetric A_n)
    double* zin;   // H: zeta input scale g(m)      (zeta(0, m))
    double* zout;  // H: zeta output scale 1/g(m)   (zeta(m, 0))
    double* ein;   // H: eta input scale            (eta(0, m))
    double* eout;  // H: eta output scale           (eta(m, 0))
    double* fact;  // 2P+2: k!
    double* X;     // 16 H: coefficients of the batch, row = (n, m), column = (interaction, component)
    double* acc;   // 4 H: partial local expansion of the target box
    double* eim;   // BP x (P+1) x 2: e^{i m phi}
    double* cs;    // BP x (P+1) x 2: cos / sin (v beta)
    double* fac;   // BP x (P+1) x 6: running factors of the error tests
    double* pw;    // BP x (P+1) x 2: multipole power per degree
    double* rp;    // BP x (P+2): t^-k
    double* geo;   // BP x 8: t, 1/t, r_mp, 1/r_mp, dx, dy, dz
    int* ord;      // BP: order of the interaction (MF_EMPTY, MF_SEARCHING)
    int* js;       // BP: source branch
    double* T;     // 16 (P+1) for this wavefront: z-translation rows
    double* T0;    // all wavefronts' buffers (NW x 16 (P+1)): error-test scratch between rotations
};

__host__ __device__ inline int mf_lds_doubles(int P, int BP, int NW) {
    const int H = nh(P);
    return len_Hs(P) + 4 * H + 2 * P + 2 + 16 * H + 4 * H + BP * (P + 1) * 12 + BP * (P + 2) + 8 * BP + BP + NW * 16 * (P + 1);
}
static size_t mf_lds_bytes(int P, int BP, int NW) { return sizeof(double) * (size_t)mf_lds_doubles(P, BP, NW); }

__device__ inline MfLds mf_carve(double* b, int P, int BP) {
    MfLds W;
    const int H = nh(P);
    W.Hs = b; b += len_Hs(P);
    W.zin = b; b += H;
    W.zout = b; b += H;
    W.ein = b; b += H;
    W.eout = b; b += H;
    W.fact = b; b += 2 * P + 2;
    W.X = b; b += 16 * H;
    W.acc = b; b += 4 * H;
    W.eim = b; b += BP * (P + 1) * 2;
    W.cs = b; b += BP * (P + 1) * 2;
    W.fac = b; b += BP * (P + 1) * 6;
    W.pw = b; b += BP * (P + 1) * 2;
    W.rp = b; b += BP * (P + 2);
    W.geo = b; b += 8 * BP;
    W.ord = (int*)b; b += BP;
    W.js = W.ord + BP;
    W.T0 = b;
    W.T = b + (size_t)wave_id() * 16 * (P + 1);
    return W;
}

// A_n(a, b) from the packed table of degree n
__device__ inline double a_sym(const double* Hn, int a, int b) {
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    return Hn[((hi * (hi + 1)) >> 1) + lo];
}

// per-interaction set-up of a batch: geometry, e^{i k phi}, (cos, sin)(k beta),
// t^-k and, for the error-controlled forms, the test factors (fill_factors)
template <int BP>
__device__ void mf_setup(const MfLds& W, const int32_t* __restrict__ m2l_s, int64_t li0, int nb, const Branch& T,
                         const Branch* __restrict__ sbr, int P, const ErrParams& E) {
    const int l = threadIdx.x;
    if (l < BP) {
        const bool valid = l < nb;
        const int32_t js = valid ? m2l_s[li0 + l] : 0;
        W.js[l] = js;
        W.ord[l] = valid ? (E.kind == 0 ? P : MF_SEARCHING) : MF_EMPTY;
        double* g = W.geo + 8 * l;
        if (valid) {
            const Branch& S = sbr[js];
            const double dx = T.target_center[0] - S.source_center[0], dy = T.target_center[1] - S.source_center[1],
                         dz = T.target_center[2] - S.source_center[2];
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            const double d0 = md1(dx, -T.target_box[0], T.target_box[0]);
            const double d1 = md1(dy, -T.target_box[1], T.target_box[1]);
            const double d2 = md1(dz, -T.target_box[2], T.target_box[2]);
            const double r_mp = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            g[0] = r; g[1] = 1.0 / r; g[2] = r_mp; g[3] = 1.0 / r_mp; g[4] = dx; g[5] = dy; g[6] = dz;
        } else {
            g[0] = 1.0; g[1] = 1.0; g[2] = 1.0; g[3] = 1.0; g[4] = 0.0; g[5] = 0.0; g[6] = 1.0;
        }
    }
    __syncthreads();
    for (int it = l; it < BP * (P + 2); it += blockDim.x) {
        const int q = it / (P + 2), k = it - q * (P + 2);
        const double* g = W.geo + 8 * q;
        double f = 1.0;
        for (int j = 0; j < k; ++j) f *= g[1];
        W.rp[it] = f;
        if (k > P) continue;
        const double x = g[4], y = g[5], z = g[6];
        const double x2y2 = x * x + y * y;
        double cp = 1.0, sp = 0.0, cb = z < 0.0 ? -1.0 : 1.0, sb = 0.0;
        if (x2y2 > 0.0) {
            const double rho = sqrt(x2y2);
            cp = x / rho; sp = y / rho;
            cb = z / g[0]; sb = rho / g[0];
        }
        double er = 1.0, ei = 0.0, c = 1.0, s = 0.0;
        for (int j = 0; j < k; ++j) {
            const double t = er;
            er = er * cp - ei * sp;
            ei = t * sp + ei * cp;
            const double cn = cb * c - sb * s;
            const double sn = sb * c + cb * s;
            c = cn;
            s = sn;
        }
        const int o = 2 * (q * (P + 1) + k);
        W.eim[o] = er; W.eim[o + 1] = ei;
        W.cs[o] = c; W.cs[o + 1] = s;
        if (E.kind != 0 && k >= 1) {
            const double rinv = g[1], r_l = T.target_radius, r_mp_inv = g[3];
            double n_f_t_np1 = rinv * rinv, r_l_nm1 = 1.0, r_mp_inv_np2 = r_mp_inv * r_mp_inv * r_mp_inv;
            double nm1_f_inv = 1.0, np1_f = 2.0;
            for (int j = 1; j < k; ++j) {
                n_f_t_np1 *= (double)(j + 1) * rinv;
                r_l_nm1 *= r_l;
                r_mp_inv_np2 *= r_mp_inv;
                nm1_f_inv /= (double)j;
                np1_f *= (double)(j + 1);
            }
            double* f6 = W.fac + 6 * (q * (P + 1) + k);
            f6[0] = n_f_t_np1; f6[1] = r_l_nm1; f6[2] = r_mp_inv_np2; f6[3] = nm1_f_inv; f6[4] = np1_f;
        }
    }
    __syncthreads();
}

// the batch's sources, all degrees: z-rotation and input scale into X (zrot_in);
// with POW the multipole power per degree (chunk_powers, summed in m order).
// One thread per (interaction, degree).
template <bool LH, bool POW>
__device__ void mf_load(const MfLds& W, const double* __restrict__ sex, int H4, const double* __restrict__ Mt, int P) {
    constexpr int NC = LH ? 4 : 2, BP = 16 / NC;
    for (int it = threadIdx.x; it < BP * (P + 1); it += blockDim.x) {
        const int q = it / (P + 1), n = it - q * (P + 1);
        const int i0 = nh(n - 1);
        double* x = W.X + 16 * i0 + q * NC;
        if (W.ord[q] == MF_EMPTY) {
            for (int m = 0; m <= n; ++m)
#pragma unroll
                for (int c = 0; c < NC; ++c) x[16 * m + c] = 0.0;
            continue;
        }
        const double* src = sex + (int64_t)W.js[q] * H4 + 4 * i0;
        const double* eq = W.eim + 2 * q * (P + 1);
        double a = 0.0, b = 0.0;
#pragma unroll 4
        for (int m = 0; m <= n; ++m) {
            const double gm = W.zin[i0 + m];
            const double er = eq[2 * m], ei = eq[2 * m + 1];
            const double2 s01 = *(const double2*)(src + 4 * m);
            x[16 * m] = gm * (s01.x * er - s01.y * ei);
            x[16 * m + 1] = gm * (s01.x * ei + s01.y * er);
            double2 s23 = make_double2(0.0, 0.0);
            if (LH) {
                s23 = *(const double2*)(src + 4 * m + 2);
                x[16 * m + 2] = gm * (s23.x * er - s23.y * ei);
                x[16 * m + 3] = gm * (s23.x * ei + s23.y * er);
            }
            if (POW) {
                const double Nt = Mt[i0 + m];
                a += (s01.x * s01.x + s01.y * s01.y) * Nt * Nt;
                if (LH) b += (s23.x * s23.x + s23.y * s23.y) * Nt * Nt;
            }
        }
        if (POW) {
            W.pw[2 * (q * (P + 1) + n)] = sqrt(a);
            W.pw[2 * (q * (P + 1) + n) + 1] = sqrt(b);
        }
    }
    __syncthreads();
}

// One y-rotation of degree n for the whole batch, by one wavefront (the
// factorised form above): step 1 SE / SO = A(v, even / odd m') X(m') on the matrix
// cores, combination into R^e / R^o, step 2 out(m) = A(m, v) R^(parity m)(v),
// scaled by mout(m) / 2. The result layout of step 1 (row v = 16 t + l/16 + 4 r in
// element r) is the operand layout step 2 needs for the chunk of rows 4 cc + l/16
// with cc = 4 t + r, so R^e / R^o pass from step 1 to step 2 in registers.
// Forward: out -> X (degree n rows). BACK: out -> back z-rotation -> sum over the
// interactions of the batch -> acc. The operand loops run to fixed bounds
// (degree <= 30) so that their loads issue ahead of the products.
template <bool LH, bool BACK>
__device__ void mf_rotate_degree(const MfLds& W, const double* mout, double sigma, int P, int n) {
    constexpr int NC = LH ? 4 : 2;
    const int lane = lane_id(), lr = lane & 15, lg = lane >> 4;
    const int q = lr / NC, c = lr % NC, isim = c & 1, pot = c >> 1;
    const double* Hn = W.Hs + len_Hs(n - 1);
    double* Xn = W.X + 16 * nh(n - 1);
    const double* mo = mout + nh(n - 1);
    const int ne = (n >> 1) + 1, no = (n + 1) >> 1;
    const int ce = (ne + 3) >> 2, co = (no + 3) >> 2;
    const double* csq = W.cs + 2 * q * (P + 1);
    const double x0r = Xn[q * NC + 2 * pot], x0i = Xn[q * NC + 2 * pot + 1];
    double re[2][4], ro[2][4];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
#pragma unroll
        for (int r = 0; r < 4; ++r) { re[t][r] = 0.0; ro[t][r] = 0.0; }
        if (16 * t > n) continue;
        mf4 se = {0.0, 0.0, 0.0, 0.0}, so = {0.0, 0.0, 0.0, 0.0};
        const int va = 16 * t + lr;
        double ae[4], be[4], ao[4], bo[4];
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) {
            const int me = 2 * (4 * cc + lg), mo2 = me + 1;
            ae[cc] = (cc < ce && va <= n && me <= n) ? a_sym(Hn, va, me) : 0.0;
            be[cc] = (cc < ce && me <= n) ? Xn[16 * me + lr] : 0.0;
            ao[cc] = (cc < co && va <= n && mo2 <= n) ? a_sym(Hn, va, mo2) : 0.0;
            bo[cc] = (cc < co && mo2 <= n) ? Xn[16 * mo2 + lr] : 0.0;
        }
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) {
            if (cc < ce) se = mfma_f64(ae[cc], be[cc], se);
            if (cc < co) so = mfma_f64(ao[cc], bo[cc], so);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int v = 16 * t + lg + 4 * r;
            const double sx = se[r], ox = so[r];
            const double sp = __shfl_xor(sx, 1, 64), op = __shfl_xor(ox, 1, 64);
            if (v > n) continue;
            const double ser = isim ? sp : sx, sei = isim ? sx : sp, sor = isim ? op : ox, soi = isim ? ox : op;
            const double a00 = Hn[(v * (v + 1)) >> 1];   // A(0, v)
            const double dr = ser - a00 * x0r, di = sei - a00 * x0i;
            const double sg = ((n + v) & 1) ? -1.0 : 1.0;
            const double w = v ? 2.0 : 1.0;
            const double cb = w * csq[2 * v], sb = w * sigma * csq[2 * v + 1];
            const double ppr = dr - sor + sg * (ser + sor), ppi = di - soi - sg * (sei + soi);
            const double pmr = dr + sor + sg * (ser - sor), pmi = di + soi - sg * (sei - soi);
            const double qpr = cb * ppr - sb * ppi, qpi = cb * ppi + sb * ppr;
            const double qmr = cb * pmr + sb * pmi, qmi = cb * pmi - sb * pmr;
            re[t][r] = isim ? (qpi + qmi) : (qpr + qmr);
            ro[t][r] = isim ? (qpi - qmi) : (qpr - qmr);
        }
    }
    const int kc = (n + 4) >> 2;
    const double* eq = W.eim + 2 * q * (P + 1);
    mf4 o[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
#pragma unroll
    for (int par = 0; par < 2; ++par) {
        const int ma = 2 * lr + par;
        double a2[8];
#pragma unroll
        for (int cc = 0; cc < 8; ++cc) {
            const int v = 4 * cc + lg;
            a2[cc] = (cc < kc && ma <= n && v <= n) ? a_sym(Hn, ma, v) : 0.0;
        }
#pragma unroll
        for (int cc = 0; cc < 8; ++cc)
            if (cc < kc) o[par] = mfma_f64(a2[cc], par ? ro[cc >> 2][cc & 3] : re[cc >> 2][cc & 3], o[par]);
    }
#pragma unroll
    for (int par = 0; par < 2; ++par) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int m = 2 * (lg + 4 * r) + par;
            const bool in = m <= n;
            const double val = in ? 0.5 * mo[m] * o[par][r] : 0.0;
            if (!BACK) {
                if (in) Xn[16 * m + lr] = val;
            } else {
                const double vp = __shfl_xor(val, 1, 64);
                const double er = in ? eq[2 * m] : 1.0, ei = in ? -eq[2 * m + 1] : 0.0;   // back_rotate_z!
                const double re2 = isim ? vp : val, im2 = isim ? val : vp;
                double y = isim ? (re2 * ei + im2 * er) : (re2 * er - im2 * ei);
                for (int k = NC; k < 16; k <<= 1) y += __shfl_xor(y, k, 64);
                if (q == 0 && in) W.acc[4 * (nh(n - 1) + m) + c] += y;
            }
        }
    }
    wave_sync();
}

// Error tests of the degrees n0..n1 (n0 >= 1, at most mf_group(NW) of them): the
// tests of m2l_test_chunk. For the power methods the whole workgroup first forms
// the degree-n part of the z-translation (translate_multipole_to_local_z_n!,
// src/translate.jl:238-295) for every (interaction, degree, m) into the
// wavefronts' product buffers; then wavefront 0 runs one lane per (interaction,
// degree) and consumes it in local_power's m order. Each searching interaction
// takes its first passing degree d: order d - 1 + BE. Called by every thread.
template <bool LH>
__device__ void mf_test_group(const MfLds& W, const ErrParams& E, const TablePtrs& tb, double limit, double r_l, int P, int n0,
                              int n1) {
    constexpr int NC = LH ? 4 : 2, BP = 16 / NC;
    const int G = n1 - n0 + 1;
    if (E.kind >= 2) {
        for (int it = threadIdx.x; it < BP * G * (P + 1); it += blockDim.x) {
            const int q = it / (G * (P + 1)), rem = it - q * G * (P + 1);
            const int gi = rem / (P + 1), m = rem - gi * (P + 1), n = n0 + gi;
            if (m > n || W.ord[q] != MF_SEARCHING) continue;
            const double rinv = W.geo[8 * q + 1];
            const double* Xq = W.X + q * NC;
            double n_m_f = W.fac[6 * (q * (P + 1) + n)];
            for (int k = 0; k < m; ++k) n_m_f *= (double)(n + k + 1) * rinv;
            double v1r = 0, v1i = 0, v2r = 0, v2i = 0, f = n_m_f;
            int n_np = n + m;
            int iw = nh(m - 1) + m;
            #pragma unroll 4
            for (int np = m; np <= n; ++np) {
                v1r += f * Xq[16 * iw]; v1i += f * Xq[16 * iw + 1];
                if (LH) { v2r += f * Xq[16 * iw + 2]; v2i += f * Xq[16 * iw + 3]; }
                n_np += 1;
                f *= (double)n_np * rinv;
                iw += np + 1;
            }
            double* t = W.T0 + (size_t)NC * ((q * G + gi) * (P + 1) + m);
            t[0] = v1r; t[1] = v1i;
            if (LH) { t[2] = v2r; t[3] = v2i; }
        }
        __syncthreads();
    }
    if (wave_id() == 0) {
    const int l = lane_id();
    const int q = l / MF_GROUP, n = n0 + (l - q * MF_GROUP);
    bool pass = false;
    if (q < BP && n <= n1 && W.ord[q] == MF_SEARCHING) {
        const double* g = W.geo + 8 * q;
        const double r = g[0], rinv = g[1], r_mp = g[2];
        const double* Xq = W.X + q * NC;
        const double* F = W.fac + 6 * (q * (P + 1) + n);
        const double n_f_t_np1 = F[0], r_l_nm1 = F[1], r_mp_inv_np2 = F[2], nm1_f_inv = F[3], np1_f = F[4];
        const double* pwq = W.pw + 2 * q * (P + 1);
        const int in0 = nh(n - 1);
        double eps_mp;
        if (E.kind == 1) {
            const double p0 = Xq[16 * in0], p1r = Xq[16 * (in0 + 1)], p1i = Xq[16 * (in0 + 1) + 1];
            eps_mp = (sqrt(p1r * p1r + p1i * p1i) + fabs(p0)) * np1_f * r_mp_inv_np2;
            if (LH) {
                const double c1r = Xq[16 * (in0 + 1) + 2], c1i = Xq[16 * (in0 + 1) + 3];
                eps_mp += sqrt(c1r * c1r + c1i * c1i) * np1_f * r_mp_inv_np2 * r_mp;
            }
        } else {
            const double pw_prev = pwq[2 * (n - 1)], pw_n = pwq[2 * n], pc_n = pwq[2 * n + 1];
            const double Mt_n0 = tb.Mt[nh(n - 2)], Mt_next = tb.Mt[in0];
            if (E.kind == 3) {
                eps_mp = SQRT3_D * pw_prev * r_mp_inv_np2 * np1_f / Mt_n0;
                if (LH) eps_mp += SQRT3_D * (double)n * pc_n * r_mp_inv_np2 * np1_f / Mt_next;
            } else {
                eps_mp = SQRT3_D * pw_n * r_mp_inv_np2 * r_mp * np1_f / (Mt_n0 * (double)(n + 1));
                if (LH) eps_mp += SQRT3_D * (double)n * pc_n * r_mp_inv_np2 * r_mp * np1_f / (Mt_next * (double)(n + 1));
            }
        }
        if (eps_mp <= limit) {
            double eps_l;
            if (E.kind == 1) {   // translate_multipole_to_local_z_m01_n (:302-362)
                double n_m_f = n_f_t_np1;
                int n_m = n;
                double p0r = 0.0, n_np_f = n_m_f;
                int n_np = n_m;
                for (int np = 0; np <= n; ++np) {
                    p0r += n_np_f * Xq[16 * ((np * (np + 1)) / 2)];
                    n_np += 1;
                    n_np_f *= (double)n_np * rinv;
                }
                n_m += 1;
                n_m_f *= (double)n_m * rinv;
                double p1r = 0, p1i = 0, c1r = 0, c1i = 0;
                n_np_f = n_m_f;
                n_np = n_m;
                for (int np = 1; np <= n; ++np) {
                    const int iw = (np * (np + 1)) / 2 + 1;
                    p1r += n_np_f * Xq[16 * iw];
                    p1i += n_np_f * Xq[16 * iw + 1];
                    if (LH) { c1r += n_np_f * Xq[16 * iw + 2]; c1i += n_np_f * Xq[16 * iw + 3]; }
                    n_np += 1;
                    n_np_f *= (double)n_np * rinv;
                }
                if (LH) {
                    const double t_over_n = r / (double)n;
                    p1r -= t_over_n * c1i;
                    p1i += t_over_n * c1r;
                }
                eps_l = (fabs(p1r) + fabs(p1i) + fabs(p0r)) * r_l_nm1 * nm1_f_inv;
                if (LH) eps_l += sqrt(c1r * c1r + c1i * c1i) * r_l_nm1 * r_l * nm1_f_inv;
            } else {
                // local_power (:455-502) of the degree-n translation, in its m order
                const double* tq = W.T0 + (size_t)NC * ((q * G + (n - n0)) * (P + 1));
                double lp = 0, lc = 0, Nt = 0, cnr = 0, cni = 0;
                #pragma unroll 4
                for (int m = n; m >= 0; --m) {
                    Nt = tb.Lt[in0 + m];
                    double pr = tq[NC * m], pi = tq[NC * m + 1];
                    if (LH) {
                        const double cor = tq[NC * m + 2], coi = tq[NC * m + 3];
                        pr = pr - coi * r * (double)m / (double)n;
                        pi = pi + cor * r * (double)m / (double)n;
                        const double cr = cor - cnr * r / (double)(n + 1);
                        const double ci = coi - cni * r / (double)(n + 1);
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
            }
            pass = (eps_mp + eps_l * LOCAL_ERROR_SAFETY <= limit);
        }
    }
    const unsigned long long mask = __ballot(pass);
    wave_sync();
    if (l < BP && W.ord[l] == MF_SEARCHING) {
        const unsigned long long mine = (mask >> (l * MF_GROUP)) & ((1ull << G) - 1ull);
        if (mine) W.ord[l] = n0 + __ffsll((long long)mine) - 1 - 1 + (E.BE ? 1 : 0);
    }
    }
    __syncthreads();
}

// z-translation of column m of the batch at each interaction's own order (rows
// above it masked), Lamb-Helmholtz transform (lh_one<true>) and the input scale
// of the back rotation; in place in X, by one wavefront
template <bool LH>
__device__ void mf_translate_m(const MfLds& W, int P, int Pb, int m) {
    constexpr int NC = LH ? 4 : 2;
    const int lane = lane_id(), lr = lane & 15, lg = lane >> 4, q = lr / NC;
    const int Pq = W.ord[q];
    const double* rpq = W.rp + q * (P + 2);
    const int sz = Pb - m + 1;
    const int kc = (sz + 3) >> 2;
    for (int t = 0; 16 * t < sz; ++t) {
        mf4 o = {0.0, 0.0, 0.0, 0.0};
        const int na = m + 16 * t + lr;
        double a[8], b[8];
#pragma unroll
        for (int cc = 0; cc < 8; ++cc) {
            const int k = m + 4 * cc + lg;
            a[cc] = (cc < kc && na <= Pb && k <= Pb) ? W.fact[na + k] : 0.0;
            b[cc] = (cc < kc && k <= Pq) ? W.X[16 * (nh(k - 1) + m) + lr] * rpq[k] : 0.0;
        }
#pragma unroll
        for (int cc = 0; cc < 8; ++cc)
            if (cc < kc) o = mfma_f64(a[cc], b[cc], o);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int n = m + 16 * t + lg + 4 * r;
            if (n <= Pb) W.T[16 * (n - m) + lr] = (n <= Pq) ? o[r] * rpq[n + 1] : 0.0;
        }
    }
    wave_sync();
    for (int it = lane; it < 16 * sz; it += 64) {
        const int n = m + (it >> 4), col = it & 15, qq = col / NC, cq = col % NC;
        const double* Tn = W.T + 16 * (n - m) + qq * NC;
        double out = Tn[cq];
        if (LH) {
            const double rq = W.geo[8 * qq];
            if (cq < 2) {
                if (n > 0) {
                    const double r_m_n = rq * (double)m / (double)n;
                    out = (cq == 0) ? Tn[0] - r_m_n * Tn[3] : Tn[1] + r_m_n * Tn[2];
                }
            } else if (n < W.ord[qq]) {
                out = Tn[cq] - rq / (double)(n + 1) * Tn[16 + cq];
            }
        }
        W.X[16 * (nh(n - 1) + m) + col] = W.ein[nh(n - 1) + m] * out;
    }
    wave_sync();
}

// One workgroup per target box; its wavefronts share each batch of BP sources:
// degrees (rotations) and columns m (z-translation) are dealt out to them.
template <bool LH, int NW>
__global__ __launch_bounds__(64 * NW, 2) void k_m2l_mfma(const int32_t* __restrict__ m2l_targets, const int64_t* __restrict__ m2l_offsets,
                                                      const int32_t* __restrict__ m2l_s, const Branch* __restrict__ tbr,
                                                      const Branch* __restrict__ sbr, const double* __restrict__ sex,
                                                      double* __restrict__ tex, int P, ErrParams E, int32_t* __restrict__ orders,
                                                      int32_t* __restrict__ pmax, int32_t* __restrict__ all_ok,
                                                      int32_t* __restrict__ box_ok, TablePtrs tb) {
    constexpr int NC = LH ? 4 : 2, BP = 16 / NC;
    extern __shared__ double lds[];
    const MfLds W = mf_carve(lds, P, BP);
    const int32_t it = m2l_targets[blockIdx.x];
    const int64_t l0 = m2l_offsets[blockIdx.x], l1 = m2l_offsets[blockIdx.x + 1];
    const int H = nh(P), H4 = 4 * H;
    const int tid = threadIdx.x, wv = wave_id();
    for (int i = tid; i < len_Hs(P); i += 64 * NW) W.Hs[i] = tb.Hs[i];
    for (int i = tid; i < H; i += 64 * NW) {
        int n, m;
        idx_nm(i, n, m);
        const int b = len_Ts(n - 1);
        W.zin[i] = tb.zeta[b + m]; W.zout[i] = tb.zeta[b + m * (n + 1)];
        W.ein[i] = tb.eta[b + m]; W.eout[i] = tb.eta[b + m * (n + 1)];
    }
    for (int i = tid; i < 2 * P + 2; i += 64 * NW) W.fact[i] = tb.fact[i];
    for (int i = tid; i < H4; i += 64 * NW) W.acc[i] = 0.0;
    const Branch T = tbr[it];
    const double limit = (E.relative ? T.max_influence * E.eps : E.eps) * 4 * PI_D;
    int local_max = 0, local_ok = 1;
    __syncthreads();
    for (int64_t li0 = l0; li0 < l1; li0 += BP) {
        const int nb = (int)min((int64_t)BP, l1 - li0);
        mf_setup<BP>(W, m2l_s, li0, nb, T, sbr, P, E);
        if (E.kind >= 2) mf_load<LH, true>(W, sex, H4, tb.Mt, P);
        else mf_load<LH, false>(W, sex, H4, tb.Mt, P);
        if (E.kind == 0) {
            for (int n = wv; n <= P; n += NW) mf_rotate_degree<LH, false>(W, W.zout, 1.0, P, n);
            __syncthreads();
        } else {
            // the tests' scratch (BP x group x (P+1) x NC) lives in the NW product buffers
            constexpr int GMAX = (NW < MF_GROUP) ? NW : MF_GROUP;
            for (int n0 = 0; n0 <= P; n0 += GMAX) {
                const int n1 = min(P, n0 + GMAX - 1);
                for (int n = n0 + wv; n <= n1; n += NW) mf_rotate_degree<LH, false>(W, W.zout, 1.0, P, n);
                __syncthreads();
                mf_test_group<LH>(W, E, tb, limit, T.target_radius, P, max(n0, 1), n1);
                bool searching = false;
#pragma unroll
                for (int q = 0; q < BP; ++q) searching |= (W.ord[q] == MF_SEARCHING);
                if (!searching) break;
            }
        }
        int Pb = 0;
        bool ok = true;
#pragma unroll
        for (int q = 0; q < BP; ++q) {
            int o = W.ord[q];
            if (o == MF_SEARCHING) { o = P; ok = false; }
            Pb = max(Pb, o);
        }
        __syncthreads();
        if (tid < BP && W.ord[tid] == MF_SEARCHING) W.ord[tid] = P;
        __syncthreads();
        for (int m = wv; m <= Pb; m += NW) mf_translate_m<LH>(W, P, Pb, m);
        __syncthreads();
        for (int n = wv; n <= Pb; n += NW) mf_rotate_degree<LH, true>(W, W.eout, -1.0, P, n);
        if (tid < nb) orders[li0 + tid] = W.ord[tid];
        local_max = max(local_max, Pb);
        local_ok = local_ok && ok;
        __syncthreads();
    }
    for (int i = tid; i < H4; i += 64 * NW) tex[(int64_t)it * H4 + i] += W.acc[i];
    if (tid == 0) {
        atomicMax(pmax, local_max);
        if (!local_ok) {
            atomicAnd(all_ok, 0);
            if (box_ok) box_ok[blockIdx.x] = 0;
        }
    }
}

template <bool LH>
static int mf_waves(int P) {
    constexpr int BP = LH ? 4 : 8;
    // two 4-wavefront workgroups per CU keep two batches in flight when their LDS fits
    if (2 * mf_lds_bytes(P, BP, 4) <= LDS_LIMIT) return 4;
    for (int nw : {8, 4, 2, 1})
        if (mf_lds_bytes(P, BP, nw) <= LDS_LIMIT) return nw;
    throw Error(FMMHIP_EINVAL, "expansion order too large for the M2L kernel");
}

template <bool LH, int NW>
static void launch_m2l_mfma(Context& C, int p, const Lists& Ls, const Tree& st, Tree& tt_dst, const ErrParams& E, int32_t* flags,
                            const TablePtrs& tb) {
    constexpr int BP = LH ? 4 : 8;
    const size_t sh = mf_lds_bytes(p, BP, NW);
    HIP_CHECK(hipFuncSetAttribute((const void*)k_m2l_mfma<LH, NW>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    k_m2l_mfma<LH, NW><<<(unsigned)Ls.n_m2l_targets, 64 * NW, sh, C.stream>>>(
        Ls.m2l_targets.data(), Ls.m2l_offsets.data(), Ls.m2l_s.data(), tt_dst.branches.data(), st.branches.data(),
        st.expansions.data(), tt_dst.expansions.data(), p, E, Ls.m2l_order.data(), flags, flags + 1,
        Ls.track_box_ok ? Ls.m2l_box_ok.data() : nullptr, tb);
    LAUNCH_CHECK();
}

template <bool LH>
static void run_m2l_mfma(Context& C, int p, const Lists& Ls, const Tree& st, Tree& tt_dst, const ErrParams& E, int32_t* flags,
                         const TablePtrs& tb) {
    switch (mf_waves<LH>(p)) {
        case 8: launch_m2l_mfma<LH, 8>(C, p, Ls, st, tt_dst, E, flags, tb); break;
        case 4: launch_m2l_mfma<LH, 4>(C, p, Ls, st, tt_dst, E, flags, tb); break;
        case 2: launch_m2l_mfma<LH, 2>(C, p, Ls, st, tt_dst, E, flags, tb); break;
        default: launch_m2l_mfma<LH, 1>(C, p, Ls, st, tt_dst, E, flags, tb); break;
    }
}

// ================================================= small orders: lane per M2L
// For p <= 4 an expansion (at most 15 coefficients x 2 potentials) fits in
// registers, so one lane runs one M2L interaction end to end with the same
// factorised rotations, and the table reads are wavefront-uniform (scalar
// loads). A wavefront owns one target box and takes 64 of its sources at a
// time; the partial locals are summed across lanes at the end. This is the
// order used by estimate_influence (p = 3, src/fmm.jl:1193-1276). Angles come
// from the Cartesian components (cos b = z/r, sin b = rho/r, e^{i phi} = (x + iy)/rho),
// which equal the reference's acos/atan2 route up to rounding.

__host__ __device__ constexpr int hidx(int a, int b) { return a <= b ? ((b * (b + 1)) >> 1) + a : ((a * (a + 1)) >> 1) + b; }

template <int P, bool LH>
__device__ inline void lane_rotate(double (&x)[nh(P)][4], const double* __restrict__ Hs, const double* __restrict__ mags,
                                   double sigma, const double (&cv)[P + 1], const double (&sv)[P + 1]) {
#pragma clang fp contract(fast)   // this file is exact elsewhere; the lane M2L's outputs are compared at 1e-11
    constexpr int NQ = LH ? 4 : 2;
#pragma unroll
    for (int n = 0; n <= P; ++n) {
        const int i0 = nh(n - 1);
        const double* Hn = Hs + len_Hs(n - 1);
        const double* mg = mags + len_Ts(n - 1);
        // inputs scaled in place by g(m'); outputs accumulated over v per m
#pragma unroll
        for (int mp = 0; mp <= n; ++mp) {
            const double g = mg[mp];
#pragma unroll
            for (int q = 0; q < NQ; ++q) x[i0 + mp][q] *= g;
        }
        double a4[P + 1][4];
#pragma unroll
        for (int m = 0; m <= n; ++m)
#pragma unroll
            for (int q = 0; q < 4; ++q) a4[m][q] = 0.0;
#pragma unroll
        for (int v = 0; v <= n; ++v) {
            double se[4] = {0.0, 0.0, 0.0, 0.0}, so[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int mp = 0; mp <= n; ++mp) {
                const double a = Hn[hidx(mp, v)];
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    if (mp & 1) so[q] = fma(a, x[i0 + mp][q], so[q]);
                    else se[q] = fma(a, x[i0 + mp][q], se[q]);
                }
            }
            const double a00 = Hn[hidx(0, v)];
            const double sg = ((n + v) & 1) ? -1.0 : 1.0;
            const double w = v ? 2.0 : 1.0;
            const double c = w * cv[v], sn = w * sigma * sv[v];
            double re[4], ro[4];
#pragma unroll
            for (int pot = 0; pot < NQ / 2; ++pot) {
                const double ser = se[2 * pot], sei = se[2 * pot + 1], sor = so[2 * pot], soi = so[2 * pot + 1];
                const double dr = ser - a00 * x[i0][2 * pot], di = sei - a00 * x[i0][2 * pot + 1];
                const double ppr = dr - sor + sg * (ser + sor), ppi = di - soi - sg * (sei + soi);
                const double pmr = dr + sor + sg * (ser - sor), pmi = di + soi - sg * (sei - soi);
                const double qpr = c * ppr - sn * ppi, qpi = c * ppi + sn * ppr;
                const double qmr = c * pmr + sn * pmi, qmi = c * pmi - sn * pmr;
                re[2 * pot] = qpr + qmr; re[2 * pot + 1] = qpi + qmi;
                ro[2 * pot] = qpr - qmr; ro[2 * pot + 1] = qpi - qmi;
            }
#pragma unroll
            for (int m = 0; m <= n; ++m) {
                const double a = Hn[hidx(m, v)];
#pragma unroll
                for (int q = 0; q < NQ; ++q) a4[m][q] = fma(a, (m & 1) ? ro[q] : re[q], a4[m][q]);
            }
        }
#pragma unroll
        for (int m = 0; m <= n; ++m) {
            const double g = 0.5 * mg[m * (n + 1)];
#pragma unroll
            for (int q = 0; q < NQ; ++q) x[i0 + m][q] = g * a4[m][q];
        }
    }
}

template <int P, bool LH>
__device__ inline void lane_m2l(double tx, double ty, double tz, const Branch& S, const double* __restrict__ sex, int32_t js,
                                const TablePtrs& tb, double (&x)[nh(P)][4]) {
#pragma clang fp contract(fast)
    constexpr int H = nh(P);
    constexpr int NQ = LH ? 4 : 2;
    const double dx = tx - S.source_center[0], dy = ty - S.source_center[1], dz = tz - S.source_center[2];
    const double x2y2 = dx * dx + dy * dy;
    const double r = sqrt(x2y2 + dz * dz);
    double cb, sb, ce, se;
    if (x2y2 > 0.0) {
        const double rho = sqrt(x2y2);
        cb = dz / r; sb = rho / r; ce = dx / rho; se = dy / rho;
    } else {
        cb = dz < 0.0 ? -1.0 : 1.0; sb = 0.0; ce = 1.0; se = 0.0;
    }
    double cv[P + 1], sv[P + 1], er[P + 1], ei[P + 1];
    cv[0] = 1.0; sv[0] = 0.0; er[0] = 1.0; ei[0] = 0.0;
#pragma unroll
    for (int k = 1; k <= P; ++k) {
        cv[k] = cb * cv[k - 1] - sb * sv[k - 1];
        sv[k] = sb * cv[k - 1] + cb * sv[k - 1];
        er[k] = er[k - 1] * ce - ei[k - 1] * se;
        ei[k] = er[k - 1] * se + ei[k - 1] * ce;
    }
    const double* src = sex + (int64_t)js * 4 * H;
#pragma unroll
    for (int n = 0; n <= P; ++n)
#pragma unroll
        for (int m = 0; m <= n; ++m) {   // rotate_z! (src/rotate.jl:448-483)
            const int i = nh(n - 1) + m;
#pragma unroll
            for (int pot = 0; pot < NQ / 2; ++pot) {
                const double a = src[4 * i + 2 * pot], b = src[4 * i + 2 * pot + 1];
                x[i][2 * pot] = a * er[m] - b * ei[m];
                x[i][2 * pot + 1] = a * ei[m] + b * er[m];
            }
        }
    lane_rotate<P, LH>(x, tb.Hs, tb.zeta, 1.0, cv, sv);
    // translate_multipole_to_local_z! (src/translate.jl:181-232): F[k] = k!/t^(k+1),
    // one order m at a time (the translation couples the degrees of one m only), then the
    // Lamb-Helmholtz local transform (:98-142) of that column, written back in place
    double F[2 * P + 1];
    const double ot = 1.0 / r;
    F[0] = ot;
#pragma unroll
    for (int k = 1; k <= 2 * P; ++k) F[k] = F[k - 1] * ((double)k * ot);
#pragma unroll
    for (int m = 0; m <= P; ++m) {
        double y[P + 1][4];
#pragma unroll
        for (int n = m; n <= P; ++n) {
            double v4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int np = m; np <= P; ++np)
#pragma unroll
                for (int q = 0; q < NQ; ++q) v4[q] = fma(F[n + np], x[nh(np - 1) + m][q], v4[q]);
#pragma unroll
            for (int q = 0; q < 4; ++q) y[n][q] = v4[q];
        }
#pragma unroll
        for (int n = m; n <= P; ++n) {
            const int i = nh(n - 1) + m;
            double o0 = y[n][0], o1 = y[n][1], o2 = 0.0, o3 = 0.0;
            if (LH) {
                const double cr = y[n][2], ci = y[n][3];
                if (n > 0) {
                    const double r_m_n = r * (double)m / (double)n;
                    o0 = y[n][0] - r_m_n * ci;
                    o1 = y[n][1] + r_m_n * cr;
                }
                o2 = cr; o3 = ci;
                if (n < P) {
                    const double r_np1 = r / (double)(n + 1);
                    o2 = cr - r_np1 * y[n + 1][2];
                    o3 = ci - r_np1 * y[n + 1][3];
                }
            }
            x[i][0] = o0; x[i][1] = o1;
            if (LH) { x[i][2] = o2; x[i][3] = o3; }
        }
    }
    lane_rotate<P, LH>(x, tb.Hs, tb.eta, -1.0, cv, sv);
#pragma unroll
    for (int n = 0; n <= P; ++n)
#pragma unroll
        for (int m = 0; m <= n; ++m) {   // back_rotate_z! (src/rotate.jl:584-616)
            const int i = nh(n - 1) + m;
#pragma unroll
            for (int pot = 0; pot < NQ / 2; ++pot) {
                const double a = x[i][2 * pot], b = x[i][2 * pot + 1];
                x[i][2 * pot] = a * er[m] + b * ei[m];
                x[i][2 * pot + 1] = b * er[m] - a * ei[m];
            }
        }
}

// Sum over the 64 lanes of v[k] (k < 64), delivered to lane k: recursive halving, each
// step keeping the half of the array selected by one lane bit and adding the partner's
// other half (63 shuffles for 64 values).
template <int L>
__device__ inline void rs_step(double (&v)[64], int lane) {
    const bool up = (lane & L) != 0;
#pragma unroll
    for (int j = 0; j < L; ++j) {
        const double mine = up ? v[j + L] : v[j];
        const double other = up ? v[j] : v[j + L];
        v[j] = mine + __shfl_xor(other, L, 64);
    }
}

template <int P, bool LH>
__global__ __launch_bounds__(256) void k_m2l_lane(const int32_t* __restrict__ m2l_targets, const int64_t* __restrict__ m2l_offsets,
                                                  const int32_t* __restrict__ m2l_s, const Branch* __restrict__ tbr,
                                                  const Branch* __restrict__ sbr, const double* __restrict__ sex,
                                                  double* __restrict__ tex, int64_t n_targets, int32_t* __restrict__ orders,
                                                  int32_t* __restrict__ pmax, TablePtrs tb) {
    constexpr int H = nh(P);
    constexpr int NQ = LH ? 4 : 2;
    static_assert(H * NQ <= 64, "one coefficient per lane");
    const int64_t wt = (int64_t)blockIdx.x * 4 + wave_id();
    if (wt >= n_targets) return;   // uniform per wavefront
    const int lane = lane_id();
    const int32_t it = m2l_targets[wt];
    const int64_t l0 = m2l_offsets[wt], l1 = m2l_offsets[wt + 1];
    const double tx = tbr[it].target_center[0], ty = tbr[it].target_center[1], tz = tbr[it].target_center[2];
    double acc = 0.0;   // lane k: coefficient k = NQ i + q of the target's local expansion
    for (int64_t base = l0; base < l1; base += 64) {
        const int64_t li = base + lane;
        double x[H][4];
#pragma unroll
        for (int i = 0; i < H; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) x[i][q] = 0.0;
        if (li < l1) {
            const int32_t js = m2l_s[li];
            lane_m2l<P, LH>(tx, ty, tz, sbr[js], sex, js, tb, x);
            orders[li] = P;
        }
        double v[64];
#pragma unroll
        for (int k = 0; k < 64; ++k) v[k] = k < H * NQ ? x[k / NQ][k % NQ] : 0.0;
        rs_step<32>(v, lane);
        rs_step<16>(v, lane);
        rs_step<8>(v, lane);
        rs_step<4>(v, lane);
        rs_step<2>(v, lane);
        rs_step<1>(v, lane);
        acc += v[0];
    }
    if (lane < H * NQ) tex[(int64_t)it * 4 * H + 4 * (lane / NQ) + lane % NQ] += acc;
    if (lane == 0 && l1 > l0) atomicMax(pmax, P);
}

template <bool LH>
static void launch_m2l_lane(Context& C, int p, const Lists& Ls, const Tree& st, Tree& tt_dst, int32_t* pmax, const TablePtrs& tb) {
    const unsigned grid = (unsigned)((Ls.n_m2l_targets + 3) / 4);
#define FMMHIP_M2L_LANE(PP)                                                                                          \
    k_m2l_lane<PP, LH><<<grid, 256, 0, C.stream>>>(Ls.m2l_targets.data(), Ls.m2l_offsets.data(), Ls.m2l_s.data(),    \
                                                   tt_dst.branches.data(), st.branches.data(), st.expansions.data(), \
                                                   tt_dst.expansions.data(), Ls.n_m2l_targets, Ls.m2l_order.data(),  \
                                                   pmax, tb)
    switch (p) {
        case 1: FMMHIP_M2L_LANE(1); break;
        case 2: FMMHIP_M2L_LANE(2); break;
        case 3: FMMHIP_M2L_LANE(3); break;
        default: FMMHIP_M2L_LANE(4); break;
    }
#undef FMMHIP_M2L_LANE
    LAUNCH_CHECK();
}


__global__ void k_init_flags(int32_t* pmax, int32_t* ok) { *pmax = 0; *ok = 1; ok[1] = 0; }

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// M2L kernel for orders above SMALL_P: the wavefront-per-interaction engine
// (k_m2l) or the matrix-core batches (k_m2l_mfma). Both compute the same
// interactions at the same orders; FMMHIP_M2L_KERNEL=mfma selects the second.
enum M2LEngine { M2L_WAVE = 0, M2L_MFMA = 1, M2L_MX = 2 };
static M2LEngine m2l_engine() {
    const char* e = std::getenv("FMMHIP_M2L_KERNEL");
    if (e && std::strcmp(e, "mfma") == 0) return M2L_MFMA;
    if (e && std::strcmp(e, "wave") == 0) return M2L_WAVE;
    return M2L_MX;
}

void run_m2l(Context& C, const Tree& tt, Tree& st, Tree& tt_dst, Lists& Ls, int p, bool LH, int err_kind, bool err_rel,
             double eps, bool BE, int precision, int32_t* pmax, int32_t* success) {
    (void)precision;
    (void)tt;
    ensure_tables(C, p);
    TablePtrs tb = ptrs(C);
    DBuf<int32_t>& flags = C.m2l_flags;
    flags.resize(3);
    k_init_flags<<<1, 1, 0, C.stream>>>(flags.data(), flags.data() + 1);
    LAUNCH_CHECK();
    if (Ls.track_box_ok) {
        Ls.m2l_box_ok.resize(std::max<int64_t>(1, Ls.n_m2l_targets));
        k_fill_i32<<<(unsigned)std::min<int64_t>(1024, (Ls.n_m2l_targets + 255) / 256 + 1), 256, 0, C.stream>>>(
            Ls.m2l_box_ok.data(), Ls.n_m2l_targets, 1);
        LAUNCH_CHECK();
    }
    if (Ls.n_m2l_targets > 0 && err_kind == FMMHIP_ERR_NONE && p <= SMALL_P) {
        if (LH) launch_m2l_lane<true>(C, p, Ls, st, tt_dst, flags.data(), tb);
        else launch_m2l_lane<false>(C, p, Ls, st, tt_dst, flags.data(), tb);
    } else if (Ls.n_m2l_targets > 0) {
        ErrParams E{err_kind, err_rel ? 1 : 0, eps, BE ? 1 : 0};
        const M2LEngine eng = m2l_engine();
        if (eng == M2L_MX && p <= 20) {
            run_m2l_mx(C, p, LH, Ls, st, tt_dst, E, flags.data(), tb);
        } else if (eng == M2L_MFMA) {
            if (LH) run_m2l_mfma<true>(C, p, Ls, st, tt_dst, E, flags.data(), tb);
            else run_m2l_mfma<false>(C, p, Ls, st, tt_dst, E, flags.data(), tb);
        } else if (err_kind != FMMHIP_ERR_NONE && p > M2L_FIRST_CAP) {
            // error-controlled orders above the cap: a first pass at the cap (more wavefronts fit
            // in LDS) settles every interaction whose tolerance is met by then; the rest run at p
            auto run = LH ? run_m2l_wave<true> : run_m2l_wave<false>;
            run(C, M2L_FIRST_CAP, p, M2L_PASS_FIRST, Ls, st, tt_dst, E, flags.data(), tb);
            run(C, p, p, M2L_PASS_DEFERRED, Ls, st, tt_dst, E, flags.data(), tb);
        } else {
            auto run = LH ? run_m2l_wave<true> : run_m2l_wave<false>;
            run(C, p, p, M2L_PASS_ALL, Ls, st, tt_dst, E, flags.data(), tb);
        }
    }
    if (!pmax) return;   // no read-back: the launch sequence stays free of host synchronisation (graph capture)
    int32_t h[2];
    HIP_CHECK(hipMemcpyAsync(h, flags.data(), sizeof(h), hipMemcpyDeviceToHost, C.stream));
    HIP_CHECK(hipStreamSynchronize(C.stream));
    *pmax = h[0];
    *success = h[1];
}

__global__ void k_fill_zero(double* __restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

void fill_zero(Context& C, double* p, size_t n) {
    if (!n) return;
    const unsigned grid = (unsigned)std::min<size_t>(4096, (n + 255) / 256);
    k_fill_zero<<<grid, 256, 0, C.stream>>>(p, (int64_t)n);
    LAUNCH_CHECK();
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

// Column k of the regular solid harmonics R_j^k, j = k..P, of two positions by the Cartesian
// recurrence above (no trigonometry), into tab0 / tab1[(j(j+1)/2 + k)][re/im]; the two recursions are
// independent, so their latencies overlap; the second is skipped when has1 is false
__device__ inline void harmonic_column2(double x0, double y0, double z0, double x1, double y1, double z1, bool has1, int k,
                                        int P, const double* inv_fact, double* tab0, double* tab1) {
    double qr0 = 1.0, qi0 = 0.0, qr1 = 1.0, qi1 = 0.0;
    for (int j = 1; j <= k; ++j) {
        const double f = -(double)(2 * j - 1);
        const double nr0 = f * (x0 * qr0 - y0 * qi0), ni0 = f * (x0 * qi0 + y0 * qr0);
        const double nr1 = f * (x1 * qr1 - y1 * qi1), ni1 = f * (x1 * qi1 + y1 * qr1);
        qr0 = nr0; qi0 = ni0; qr1 = nr1; qi1 = ni1;
    }
    const double rho20 = x0 * x0 + y0 * y0 + z0 * z0, rho21 = x1 * x1 + y1 * y1 + z1 * z1;
    double pr0 = 0.0, pi0 = 0.0, cr0 = qr0, ci0 = qi0;
    double pr1 = 0.0, pi1 = 0.0, cr1 = qr1, ci1 = qi1;
    for (int j = k; j <= P; ++j) {
        if (j > k) {
            const double d = 1.0 / (double)(j - k);
            const double a0 = (double)(2 * j - 1) * z0, b0 = (double)(j + k - 1) * rho20;
            const double nr0 = (a0 * cr0 - b0 * pr0) * d, ni0 = (a0 * ci0 - b0 * pi0) * d;
            pr0 = cr0; pi0 = ci0; cr0 = nr0; ci0 = ni0;
            const double a1 = (double)(2 * j - 1) * z1, b1 = (double)(j + k - 1) * rho21;
            const double nr1 = (a1 * cr1 - b1 * pr1) * d, ni1 = (a1 * ci1 - b1 * pi1) * d;
            pr1 = cr1; pi1 = ci1; cr1 = nr1; ci1 = ni1;
        }
        const double s = ((j & 1) ? -1.0 : 1.0) * inv_fact[j + k];
        const int i = ((j * (j + 1)) >> 1) + k;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (h == 1 && !has1) break;
            const double ar = (h ? cr1 : cr0) * s, ai = (h ? ci1 : ci0) * s;
            double o0, o1;
            switch (k & 3) {
                case 0: o0 = ar; o1 = ai; break;
                case 1: o0 = -ai; o1 = ar; break;
                case 2: o0 = -ar; o1 = -ai; break;
                default: o0 = ai; o1 = -ar; break;
            }
            double* tab = h ? tab1 : tab0;
            tab[2 * i] = o0;
            tab[2 * i + 1] = o1;
        }
    }
}

// R_j^k from a table (zero outside 0 <= k <= j)
__device__ inline void tab_get(const double* tab, int j, int k, double& r, double& i) {
    if (j < 0 || k < 0 || k > j) { r = 0.0; i = 0.0; return; }
    const int idx = ((j * (j + 1)) >> 1) + k;
    r = tab[2 * idx];
    i = tab[2 * idx + 1];
}

// P2M (body_to_multipole!, src/bodytomultipole.jl:387-496): one wavefront per leaf. The
// bodies go through in groups of 2 x 64/(P+1): lane l builds column l % (P+1) of the harmonic
// tables of bodies l / (P+1) and G + l / (P+1) (no barrier inside; 1/k! from LDS), then every
// lane adds the group's contributions, body by body, to the coefficients it owns
// (i = lane + 64 r), reading neighbours from the tables. Point{Vortex} uses the mirrored
// position (mirrored_source_to_vortex!).
template <int KERNEL>
__global__ __launch_bounds__(64) void k_p2m(const Branch* __restrict__ br, const int64_t* __restrict__ leaf_index,
                                            const double* __restrict__ x, const double* __restrict__ y,
                                            const double* __restrict__ z, const double* __restrict__ q, int64_t ns,
                                            double* __restrict__ ex, int P, const double* __restrict__ inv_fact) {
    extern __shared__ double lds[];
    constexpr int NR = 4;   // coefficients per lane: H(20) = 231 <= 4 x 64
    const int H = nh(P);
    const int G = 64 / (P + 1);                       // bodies per column pass; two passes per group
    double* tabs = lds;                               // 2G tables of 2H doubles
    double* qs = lds + (size_t)2 * G * 2 * H;         // 2G x 3 strengths
    double* fs = qs + (size_t)2 * G * 3;              // 1/k!, k = 0 .. 2P
    for (int i = threadIdx.x; i <= 2 * P; i += 64) fs[i] = inv_fact[i];
    __syncthreads();
    const int64_t ib = leaf_index[blockIdx.x];
    const Branch B = br[ib];
    const double cx = B.source_center[0], cy = B.source_center[1], cz = B.source_center[2];
    const int tid = threadIdx.x;
    int rn[NR], rm[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int i = tid + 64 * r;
        if (i < H) idx_nm(i, rn[r], rm[r]);
        else { rn[r] = -1; rm[r] = 0; }
    }
    double acc[NR][4];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r][0] = acc[r][1] = acc[r][2] = acc[r][3] = 0.0;
    const bool mirror = KERNEL == FMMHIP_POINT_VORTEX;
    for (int64_t b0 = 0; b0 < B.n_bodies; b0 += 2 * G) {
        const int cnt = (int)min<int64_t>(2 * G, B.n_bodies - b0);
        const int g = tid / (P + 1), k = tid % (P + 1);
        if (g < G && g < cnt) {
            const bool has1 = G + g < cnt;
            const int64_t j0 = B.first + b0 + g, j1 = has1 ? j0 + G : j0;
            double dx0 = x[j0] - cx, dy0 = y[j0] - cy, dz0 = z[j0] - cz;
            double dx1 = x[j1] - cx, dy1 = y[j1] - cy, dz1 = z[j1] - cz;
            if (mirror) { dx0 = -dx0; dy0 = -dy0; dz0 = -dz0; dx1 = -dx1; dy1 = -dy1; dz1 = -dz1; }
            harmonic_column2(dx0, dy0, dz0, dx1, dy1, dz1, has1, k, P, fs, tabs + (size_t)g * 2 * H,
                             tabs + (size_t)(G + g) * 2 * H);
            if (k == 0) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if (h == 1 && !has1) break;
                    const int gg = g + h * G;
                    const int64_t j = h ? j1 : j0;
                    qs[3 * gg] = q[j];
                    if (KERNEL != FMMHIP_POINT_SOURCE) { qs[3 * gg + 1] = q[ns + j]; qs[3 * gg + 2] = q[2 * ns + j]; }
                }
            }
        }
        __syncthreads();
        for (int gb = 0; gb < cnt; ++gb) {
            const double* T = tabs + (size_t)gb * 2 * H;
            const double s0 = qs[3 * gb], s1 = qs[3 * gb + 1], s2 = qs[3 * gb + 2];
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const int n = rn[r], m = rm[r];
                if (n < 0) continue;
                const double om = (m & 1) ? -1.0 : 1.0;
                if (KERNEL == FMMHIP_POINT_SOURCE) {
                    // M_n^m += (-1)^(n+m) (-q) conj(R_n^m(dx))   (src/bodytomultipole.jl:403-429)
                    double Rr, Ri;
                    tab_get(T, n, m, Rr, Ri);
                    const double f = (((n + m) & 1) ? 1.0 : -1.0) * s0;
                    acc[r][0] += f * Rr;
                    acc[r][1] -= f * Ri;
                } else if (KERNEL == FMMHIP_POINT_DIPOLE) {
                    // source_to_dipole! of -q on q_{n-1}^{m-1,m,m+1} (get_nm1 rules, :77-110, 251-278, 431-463)
                    if (n >= 1) {
                        double q0, q1, q2, q3, q4, q5;
                        tab_get(T, n - 1, m > 0 ? m - 1 : 1, q0, q1);
                        if (m == 0) q0 = -q0;   // q_{n-1}^{-1} = -conj(q_{n-1}^1)
                        tab_get(T, n - 1, m, q2, q3);
                        tab_get(T, n - 1, m + 1, q4, q5);
                        const double sxq = -s0, syq = -s1, szq = -s2;
                        const double hr = -sxq * 0.5 * (q5 + q1) + syq * 0.5 * (q4 - q0) - szq * q2;
                        const double hi = sxq * 0.5 * (q4 + q0) + syq * 0.5 * (q5 - q1) - szq * q3;
                        const double f = ((n + m) & 1) ? -1.0 : 1.0;
                        acc[r][0] += f * hr;
                        acc[r][1] -= f * hi;
                    }
                } else {
                    // mirrored_source_to_vortex! with q = R(-dx) (src/bodytomultipole.jl:283-334, 465-489)
                    const double vx = s0, vy = s1, vz = s2;
                    double m1r, m1i, m1pr, m1pi, r0r, r0i, r0pr, r0pi, p1r, p1i, p1pr, p1pi;
                    const int km1 = m - 1 >= 0 ? m - 1 : 1;
                    tab_get(T, n, km1, m1r, m1i);
                    tab_get(T, n - 1, km1, m1pr, m1pi);
                    tab_get(T, n, m, r0r, r0i);
                    tab_get(T, n - 1, m, r0pr, r0pi);
                    tab_get(T, n, m + 1, p1r, p1i);
                    tab_get(T, n - 1, m + 1, p1pr, p1pi);
                    // phi part: q_n^{m-1}, q_n^m, q_n^{m+1}
                    double qm1r, qm1i;
                    if (m > 0) { qm1r = m1r; qm1i = m1i; }
                    else if (n == 0) { qm1r = 0.0; qm1i = 0.0; }
                    else { qm1r = -om * m1r; qm1i = om * m1i; }   // q_n^{-1} = -conj(q_n^1)
                    const double qmi = r0i, qmr = r0r;
                    const double qp1r = (m < n) ? p1r : 0.0, qp1i = (m < n) ? p1i : 0.0;
                    const double A = (double)(n - m + 1) * 0.5, Bc = (double)(n + m + 1) * 0.5, inp1 = 1.0 / (double)(n + 1);
                    acc[r][0] += om * ((-vx * qm1r + vy * qm1i) * A + (vx * qp1r + vy * qp1i) * Bc - vz * (double)m * qmi) * inp1;
                    acc[r][1] += om * ((vx * qm1i + vy * qm1r) * A + (-vx * qp1i + vy * qp1r) * Bc - vz * (double)m * qmr) * inp1;
                    if (n >= 1) {
                        // chi part: q_{n-1}^{m-1}, q_{n-1}^m, q_{n-1}^{m+1} (get_nm1 rules)
                        double pm1r, pm1i;
                        if (m > 0) { pm1r = m1pr; pm1i = m1pi; }
                        else if (n > 1) { pm1r = -om * m1pr; pm1i = om * m1pi; }
                        else { pm1r = 0.0; pm1i = 0.0; }
                        const double p0r = (m < n) ? r0pr : 0.0, p0i = (m < n) ? r0pi : 0.0;
                        const double pp1r = (m + 1 < n) ? p1pr : 0.0, pp1i = (m + 1 < n) ? p1pi : 0.0;
                        const double ion = 1.0 / (double)n;
                        acc[r][2] -= om * ion * (0.5 * (-vy * pm1r - vx * pm1i + vy * pp1r - vx * pp1i) - vz * p0r);
                        acc[r][3] -= om * ion * (0.5 * (vy * pm1i - vx * pm1r - vy * pp1i - vx * pp1r) + vz * p0i);
                    }
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int i = tid + 64 * r;
        if (i >= H) continue;
        double* e = ex + ib * 4 * H + 4 * i;
        e[0] += acc[r][0]; e[1] += acc[r][1];
        if (KERNEL == FMMHIP_POINT_VORTEX) { e[2] += acc[r][2]; e[3] += acc[r][3]; }   // Lamb-Helmholtz chi part
    }
}

// --------------------------------------------------------- panel P2M
// body_to_multipole! for constant-strength triangles (src/bodytomultipole.jl:643-831):
// the regular-harmonic recurrences of calculate_q! at x0+xu, calculate_pj! at
// x0+xv and calculate_ib! at x0 (:112-213), each degree from the previous one
// (get_nm1 rules, :77-110), then M += |xu x xv| (-1)^(n+m) conj(i_n^m) * (-q)
// (source) or the dipole conversion source_to_dipole! of i with moment -q n
// (:251-278). One wavefront per leaf; lane m carries coefficient m of the
// current degree for all three recurrences at once (one barrier per degree).
// Quadrilaterals are the triangles (v1, v2, v3) and (v1, v3, v4) with the unit
// normal of each triangle.

// get_nm1 for lane (n, m) from the degree n-1 values of A ([index][re/im] in LDS)
__device__ inline void nm1_vals(const double* A, int n, int m, double q[6]) {
    const int b = nh(n - 2);   // start of degree n-1
    if (m > 0) { q[0] = A[2 * (b + m - 1)]; q[1] = A[2 * (b + m - 1) + 1]; }
    else if (n > 1) { q[0] = -A[2 * (b + 1)]; q[1] = A[2 * (b + 1) + 1]; }
    else { q[0] = 0.0; q[1] = 0.0; }
    if (m < n) { q[2] = A[2 * (b + m)]; q[3] = A[2 * (b + m) + 1]; } else { q[2] = 0.0; q[3] = 0.0; }
    if (m + 1 < n) { q[4] = A[2 * (b + m + 1)]; q[5] = A[2 * (b + m + 1) + 1]; } else { q[4] = 0.0; q[5] = 0.0; }
}

template <bool DIPOLE>
__device__ void p2m_triangle(const double* v1, const double* v2, const double* v3, const double* center, double strength, int P,
                             double* Q, double* J, double* I, double* Macc) {
    const int m = (int)threadIdx.x;
    const double x0[3] = {v1[0] - center[0], v1[1] - center[1], v1[2] - center[2]};
    const double xu[3] = {v2[0] - v1[0], v2[1] - v1[1], v2[2] - v1[2]};
    const double xv[3] = {v3[0] - v1[0], v3[1] - v1[1], v3[2] - v1[2]};
    // xyz_to_xi_eta_z (:39-46) of each vector, added componentwise
    const double x0r = x0[0] * 0.5, x0i = x0[1] * 0.5, xur = xu[0] * 0.5, xui = xu[1] * 0.5, xvr = xv[0] * 0.5, xvi = xv[1] * 0.5;
    const double qxr = x0r + xur, qxi = x0i + xui, qer = x0r + xur, qei = -x0i + -xui, qz = x0[2] + xu[2];
    const double jxr = x0r + xvr, jxi = x0i + xvi, jer = x0r + xvr, jei = -x0i + -xvi, jz = x0[2] + xv[2];
    const double ixr = x0r, ixi = x0i, ier = x0r, iei = -x0i, iz = x0[2];
    const double c0 = xu[1] * xv[2] - xu[2] * xv[1], c1 = xu[2] * xv[0] - xu[0] * xv[2], c2 = xu[0] * xv[1] - xu[1] * xv[0];
    const double area2 = sqrt(c0 * c0 + c1 * c1 + c2 * c2);   // |xu x xv|
    double dqx = 0.0, dqy = 0.0, dqz = 0.0;
    if (DIPOLE) {   // moment -q n, n the unit normal of this triangle
        dqx = -c0 / area2 * strength; dqy = -c1 / area2 * strength; dqz = -c2 / area2 * strength;
    }
    const double w = DIPOLE ? area2 : area2 * (-strength);
    for (int n = 0; n <= P; ++n) {
        const bool act = m <= n;
        const int i = nh(n - 1) + m;
        double qr = 0.0, qi = 0.0, jr = 0.0, ji = 0.0, ir = 0.0, ii = 0.0, dr = 0.0, di = 0.0;
        if (act) {
            if (n == 0) {
                qr = 1.0; jr = 1.0; ir = 0.5;
            } else {
                double a[6];
                nm1_vals(Q, n, m, a);   // calculate_q! (:112-139)
                qr = (-(qxr * a[1] + qxi * a[0]) - (qer * a[5] + qei * a[4]) - qz * a[2]) / (double)n;
                qi = ((qxr * a[0] - qxi * a[1]) + (qer * a[4] - qei * a[5]) - qz * a[3]) / (double)n;
                nm1_vals(J, n, m, a);   // calculate_pj! (:146-176)
                jr = (-(jxr * a[1] + jxi * a[0]) - (jer * a[5] + jei * a[4]) - jz * a[2] + qr) / (double)(n + 1);
                ji = ((jxr * a[0] - jxi * a[1]) + (jer * a[4] - jei * a[5]) - jz * a[3] + qi) / (double)(n + 1);
                nm1_vals(I, n, m, a);   // calculate_ib! (:183-213)
                ir = (-(ixr * a[1] + ixi * a[0]) - (ier * a[5] + iei * a[4]) - iz * a[2] + jr) / (double)(n + 2);
                ii = ((ixr * a[0] - ixi * a[1]) + (ier * a[4] - iei * a[5]) - iz * a[3] + ji) / (double)(n + 2);
                if (DIPOLE) {   // source_to_dipole! (:251-278) from i of degree n-1
                    dr = -dqx * 0.5 * (a[5] + a[1]) + dqy * 0.5 * (a[4] - a[0]) - dqz * a[2];
                    di = dqx * 0.5 * (a[4] + a[0]) + dqy * 0.5 * (a[5] - a[1]) - dqz * a[3];
                }
            }
            const double sg = ((n + m) & 1) ? -1.0 : 1.0;
            if (DIPOLE) { Macc[2 * i] += w * sg * dr; Macc[2 * i + 1] -= w * sg * di; }
            else { Macc[2 * i] += w * sg * ir; Macc[2 * i + 1] -= w * sg * ii; }
        }
        __syncthreads();   // every lane has read degree n-1 of the three arrays
        if (act) {
            Q[2 * i] = qr; Q[2 * i + 1] = qi;
            J[2 * i] = jr; J[2 * i + 1] = ji;
            I[2 * i] = ir; I[2 * i + 1] = ii;
        }
        __syncthreads();   // degree n published
    }
}

template <int KERNEL>
__global__ __launch_bounds__(64) void k_p2m_panel(const Branch* __restrict__ br, const int64_t* __restrict__ leaf_index,
                                                  const double* __restrict__ rows, int64_t ns, double* __restrict__ ex, int P,
                                                  int srow, int voff) {
    extern __shared__ double lds[];
    constexpr int NV = panel_vertices(KERNEL);
    constexpr bool DIPOLE = panel_dipole(KERNEL);
    const int H = nh(P);
    double* Q = lds;
    double* J = Q + 2 * H;
    double* I = J + 2 * H;
    double* Macc = I + 2 * H;
    const int64_t ib = leaf_index[blockIdx.x];
    const Branch B = br[ib];
    for (int k = threadIdx.x; k < 2 * H; k += 64) Macc[k] = 0.0;
    __syncthreads();
    for (int64_t j = B.first; j < B.first + B.n_bodies; ++j) {
        double v[NV][3];
#pragma unroll
        for (int k = 0; k < NV; ++k)
#pragma unroll
            for (int a = 0; a < 3; ++a) v[k][a] = rows[(int64_t)(voff + 3 * k + a) * ns + j];
        const double q = rows[(int64_t)srow * ns + j];   // strength row srow, vertices from row voff
#pragma unroll
        for (int t = 0; t < NV - 2; ++t)
            p2m_triangle<DIPOLE>(v[0], v[t + 1], v[t + 2], B.source_center, q, P, Q, J, I, Macc);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < H; k += 64) {
        ex[ib * 4 * H + 4 * k] += Macc[2 * k];
        ex[ib * 4 * H + 4 * k + 1] += Macc[2 * k + 1];
    }
}

void run_p2m(Context& C, int kernel, Tree& st, int p, const Branch* sbr) {
    if (!sbr) sbr = st.branches.data();
    ensure_tables(C, p);
    if (st.n_leaves == 0) return;
    TablePtrs tb = ptrs(C);
    if (is_panel(kernel)) {
        const size_t sh = sizeof(double) * 8 * nh(p);
#define FMMHIP_P2M_PANEL(K, SROW, VOFF)                                                                                \
    k_p2m_panel<K><<<(unsigned)st.n_leaves, 64, sh, C.stream>>>(sbr, st.leaf_index.data(), st.strength.data(),             \
                                                               st.n, st.expansions.data(), p, SROW, VOFF)
        switch (kernel) {
            case FMMHIP_PANEL_SOURCE_TRI: FMMHIP_P2M_PANEL(FMMHIP_PANEL_SOURCE_TRI, 0, 1); break;
            case FMMHIP_PANEL_DIPOLE_TRI: FMMHIP_P2M_PANEL(FMMHIP_PANEL_DIPOLE_TRI, 0, 1); break;
            case FMMHIP_PANEL_SOURCE_QUAD: FMMHIP_P2M_PANEL(FMMHIP_PANEL_SOURCE_QUAD, 0, 1); break;
            case FMMHIP_PANEL_SOURCE_DIPOLE_TRI:   // the source's expansion, then the dipole's (src/bodytomultipole.jl:717-720)
                FMMHIP_P2M_PANEL(FMMHIP_PANEL_SOURCE_TRI, 0, 2);
                LAUNCH_CHECK();
                FMMHIP_P2M_PANEL(FMMHIP_PANEL_DIPOLE_TRI, 1, 2);
                break;
            case FMMHIP_PANEL_SOURCE_DIPOLE_QUAD:
                FMMHIP_P2M_PANEL(FMMHIP_PANEL_SOURCE_QUAD, 0, 2);
                LAUNCH_CHECK();
                FMMHIP_P2M_PANEL(FMMHIP_PANEL_DIPOLE_QUAD, 1, 2);
                break;
            default: FMMHIP_P2M_PANEL(FMMHIP_PANEL_DIPOLE_QUAD, 0, 1); break;
        }
#undef FMMHIP_P2M_PANEL
        LAUNCH_CHECK();
        return;
    }
    const size_t p2m_sh = sizeof(double) * ((size_t)2 * (64 / (p + 1)) * (2 * nh(p) + 3) + 2 * p + 1);
#define FMMHIP_P2M(K, Q)                                                                                               \
    k_p2m<K><<<(unsigned)st.n_leaves, 64, p2m_sh, C.stream>>>(sbr, st.leaf_index.data(), st.x.data(), st.y.data(),             \
                                                        st.z.data(), Q, st.n, st.expansions.data(), p, tb.inv_fact)
    const double* q = st.strength.data();
    if (kernel == FMMHIP_POINT_SOURCE) FMMHIP_P2M(FMMHIP_POINT_SOURCE, q);
    else if (kernel == FMMHIP_POINT_DIPOLE) FMMHIP_P2M(FMMHIP_POINT_DIPOLE, q);
    else if (kernel == FMMHIP_POINT_SOURCE_VORTEX) {   // the source's expansion (row 0), then the vortex's (rows 1-3)
        FMMHIP_P2M(FMMHIP_POINT_SOURCE, q);
        LAUNCH_CHECK();
        FMMHIP_P2M(FMMHIP_POINT_VORTEX, q + st.n);
    } else FMMHIP_P2M(FMMHIP_POINT_VORTEX, q);   // the Gaussian core changes only the nearfield
#undef FMMHIP_P2M
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

// One wavefront per leaf. Every lane reads the same coefficients at the same time, so they come from
// memory through the scalar cache (wave-uniform addresses) rather than being staged in LDS, whose 48 KB
// per leaf at order 20 held a CU to three leaves.
__global__ __launch_bounds__(64) void k_l2p_eval(const Branch* __restrict__ br, const int64_t* __restrict__ leaf_index,
                                                 const double* __restrict__ x, const double* __restrict__ y,
                                                 const double* __restrict__ z, int P, const double* __restrict__ coef,
                                                 double* __restrict__ tout, int64_t nt, int PS, int GS, int HS) {
    const int64_t ib = leaf_index[blockIdx.x];
    const Branch B = br[ib];
    const int H = nh(P);
    const double* __restrict__ lc = coef + (int64_t)blockIdx.x * 26 * H;
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
    DBuf<double>& coef = C.l2p_coef;
    coef.resize((size_t)tt.n_leaves * 26 * nh(p));
    if (LH)
        k_l2p_coef<true><<<(unsigned)tt.n_leaves, 64, 0, C.stream>>>(tt.leaf_index.data(), tt.expansions.data(), p, PS, GS, HS, tb.inv_fact, coef.data());
    else
        k_l2p_coef<false><<<(unsigned)tt.n_leaves, 64, 0, C.stream>>>(tt.leaf_index.data(), tt.expansions.data(), p, PS, GS, HS, tb.inv_fact, coef.data());
    LAUNCH_CHECK();
    k_l2p_eval<<<(unsigned)tt.n_leaves, 64, 0, C.stream>>>(
        tt.branches.data(), tt.leaf_index.data(), tt.x.data(), tt.y.data(), tt.z.data(), p, coef.data(), tout, nt, PS, GS, HS);
    LAUNCH_CHECK();
}

}  // namespace fmmhip
