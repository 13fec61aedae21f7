// Oracle (test infrastructure only): CPU FP64 restatement of FastMultipole's
// expansion arithmetic. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may use this code, and only as the checker.
//
// Indexing convention: every array below is 1-based through the accessor
// helpers so that each statement can be read against the reference line it
// restates. Expansions use the reference layout expansions[re/im, phi/chi, H]
// (src/tree.jl:1334-1339), column-major.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstddef>
#include <vector>
#include <algorithm>

namespace orc {

static const double PI = 3.141592653589793;
static const double ONE_OVER_4PI = 1.0 / (4.0 * 3.141592653589793);   // src/FastMultipole.jl:12
static const double SQRT3 = 1.7320508075688772;                      // src/FastMultipole.jl:16 (sqrt(3.0))
static const double LOCAL_ERROR_SAFETY = 2.0;                         // src/FastMultipole.jl:17

// ---------------------------------------------------------------- indexing
inline int harmonic_index(int n, int m) { return ((n * (n + 1)) >> 1) + m + 1; }   // src/harmonics.jl:5-7
inline int n_harmonics(int p) { return ((p + 1) * (p + 2)) >> 1; }

// expansion accessor: a in {1,2} re/im, b in {1,2} phi/chi, i 1-based
struct Exp {
    double* d;
    inline double& operator()(int a, int b, int i) const { return d[(a - 1) + 2 * ((b - 1) + 2 * (size_t)(i - 1))]; }
};
inline void zero_exp(Exp e, int p) { for (int k = 0; k < 4 * n_harmonics(p); ++k) e.d[k] = 0.0; }

// --------------------------------------------------------- tables (global)
struct Tables {
    int p = -1;
    std::vector<double> Hs_pi2, zetas, etas, Mt, Lt;  // 1-based via at()
    void ensure(int p_needed);
};
Tables& tables();

// ------------------------------------------------------------- harmonics
void regular_harmonics(Exp h, double rho, double theta, double phi, int p);     // src/harmonics.jl:13-58
void irregular_harmonics(Exp h, double rho, double theta, double phi, int p);   // src/harmonics.jl:64-104
void cartesian_to_spherical(double x, double y, double z, double& r, double& theta, double& phi);  // src/rotate.jl:9-21

// ------------------------------------------------------------- rotations
int length_Hs(int p);
int length_Ts(int p);
inline int H_index(int mp, int m) { return ((m * (m + 1)) >> 1) + mp + 1; }
inline int T_index(int mp, int m) { return m * m + mp + m + 1; }
std::vector<double> compute_Pnms(int p, double s, double c);                     // src/rotate.jl:95-117
void compute_Hs(std::vector<double>& Hs, double beta, int n_start, int n_end);  // src/rotate.jl:119-203
void update_eimphis(double* eimphis, double phi, int P);                        // src/rotate.jl:25-36
void update_Ts(double* Ts, const double* Hs, double beta, int p);               // src/rotate.jl:260-346
void update_Ts_0(double* Ts, double beta, double& sb, double& cb, double& one_n); // :351-361
void update_Ts_n(double* Ts, const double* Hs, double sb, double cb, double& one_n, int n); // :363-441
void rotate_z(Exp out, Exp in, double* eimphis, double phi, int P, bool LH);    // :448-483
void back_rotate_z(Exp out, Exp in, const double* eimphis, int P, bool LH);     // :584-616
void rotate_multipole_y_full(Exp out, Exp in, const double* Ts, const double* zetas, int P, bool LH);  // :699-803
void rotate_multipole_y_n(Exp out, Exp in, const double* Ts, const double* zetas, bool LH, int& i_z, int& i_T, int n); // :808-910
void rotate_local_y_full(Exp out, Exp in, const double* Ts, const double* etas, int P, bool LH);      // :1014-1113
void update_zetas_mag(std::vector<double>& z, int n_start, int n_end);          // :658-689
void update_etas_mag(std::vector<double>& e, int n_start, int n_end);           // :973-1004
void zeta_sign(double mag, int mp, int m, double& re, double& im);              // :644-656
void eta_sign(double mag, int mp, int m, double& re, double& im);               // :959-971

// ----------------------------------------------------------- translations
void translate_multipole_z(Exp out, Exp in, double t, int p, bool LH);          // src/translate.jl:6-50
void transform_lh_multipole(Exp e, double r, int p);                            // :52-96
void transform_lh_local(Exp e, double r, int p);                                // :98-142
void translate_multipole_to_local_z(Exp out, Exp in, double t, int p, bool LH); // :181-232
void translate_local_z(Exp out, Exp in, double t, int p, bool LH);              // :1851-1891

struct Scratch {
    std::vector<double> w1, w2, w3, Ts, eimphis;
    void ensure(int p);
};

void multipole_to_multipole(Exp target, const double* target_center, Exp source, const double* source_center,
                            Scratch& s, int p, bool LH);                        // :144-174
void local_to_local(Exp target, const double* target_center, Exp source, const double* source_center,
                    Scratch& s, int p, bool LH);                                // :1896-1926

// error method: 0 none, 1 RotatedCoefficients(Abs/Rel)Gradient, 2 Power(Abs/Rel)Potential,
// 3 Power(Abs/Rel)Gradient; `relative` scales epsilon by the target max_influence.
struct ErrorMethod { int kind = 0; bool relative = false; double eps = 0.0; bool BE = true; };

struct BranchGeom {
    double source_center[3], target_center[3];
    double source_radius, target_radius;
    double source_box[3], target_box[3];
    double max_influence;
};

// returns expansion order used; error_success via out parameter (src/translate.jl:519-1412)
int multipole_to_local(Exp target, const BranchGeom& tb, Exp source, const BranchGeom& sb,
                       Scratch& s, int p, bool LH, const ErrorMethod& em, bool& error_success);
// closeness of the last error-controlled multipole_to_local's order decision to its tolerance:
// min over every comparison it made (eps_mp <= limit, eps_mp + 2 eps_l <= limit) of |lhs/limit - 1|
// (+inf for a fixed-order call). Test infrastructure: an interaction whose margin is well above
// rounding cannot change order under last-bit differences in the rotated coefficients.
extern thread_local double last_m2l_margin;

void minimum_distance(const double* c1, const double* c2, const double* box2, double* out); // src/error.jl:5-20

// ------------------------------------------------------------------- L2P
// src/evaluate_expansions.jl:45-399; returns -u/4pi, v/4pi, hessian/4pi (9, column-major)
void evaluate_local(const double* dx, Exp harmonics, std::vector<double>& grad_nm, Exp local, int p, bool LH,
                    bool PS, bool GS, bool HS, double& phi_out, double* v_out, double* h_out);

// ------------------------------------------------------------------- P2M
// body_to_multipole_point! for Point{Source}/Point{Dipole}/Point{Vortex}; src/bodytomultipole.jl:403-496
void p2m_point_source(Exp M, Exp harm, const double* dx, double q, int p);
void p2m_point_dipole(Exp M, Exp harm, const double* dx, const double* q, int p);
// body_to_multipole_panel! for Panel{Source} / Panel{Dipole} triangles (src/bodytomultipole.jl:722-831);
// x0 = first vertex relative to the expansion centre, xu = v2 - v1, xv = v3 - v1; harm needs 4 H(p) doubles
void p2m_panel_source(Exp M, Exp harm, const double* x0, const double* xu, const double* xv, double strength, int p);
void p2m_panel_dipole(Exp M, Exp harm, const double* x0, const double* xu, const double* xv, const double* normal,
                      double strength, int p);
void p2m_point_vortex(Exp M, Exp harm, const double* dx, const double* g, int p);

double factorial_to_double(int k);   // Float64(factorial(big(k))), correctly rounded

}  // namespace orc
