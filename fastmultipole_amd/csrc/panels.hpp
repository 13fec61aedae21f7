// Device functions of the constant-strength panel kernels, shared by the
// nearfield (p2p.hip) and the Gauss-Seidel influence matrices (fgs.hip).
#pragma once
#include "fmmhip_internal.hpp"

namespace fmmhip {

static constexpr double PANEL_INV4PI = 1.0 / (4.0 * 3.141592653589793);

// ================================================================ panels
// Constant-strength planar panels, nearfield: induced() of test/panels.jl
// (rotate_to_panel :78-93, source_dipole_preliminaries :95-133, Panel{Source}
// :139-216, Panel{Dipole} :224-315) for every (target, panel) pair, FP64 in both
// precision forms (log/atan edge terms). The per-panel frame and edge geometry
// is computed once when a panel is staged into LDS.

static constexpr int PANEL_TILE = 64;   // panels per LDS tile

template <int NV> struct PanelStage {
    double c[3], ex[3], ey[3], ez[3];   // centroid, panel frame (R' rows)
    double vx[NV + 1], vy[NV + 1];      // vertices in the panel frame (closed polygon)
    double dx[NV], dy[NV], ds[NV], ms[NV];
    double strength;
    double strength2;   // a source-dipole panel's dipole strength
};

// panel j of the sorted arrays (rows after radius: NSTR strengths, v1, v2, ...) into a stage
template <int NV, int NSTR = 1>
__device__ inline void stage_panel(const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
                                   const double* __restrict__ rows, int64_t n, int64_t j, PanelStage<NV>& P) {
    double v[NV][3];
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
        for (int a = 0; a < 3; ++a) v[k][a] = rows[(int64_t)(NSTR + 3 * k + a) * n + j];
    P.c[0] = px[j]; P.c[1] = py[j]; P.c[2] = pz[j];
    P.strength = rows[j];
    P.strength2 = NSTR > 1 ? rows[n + j] : 0.0;
    // unit normal from the first three vertices (get_normal, src/compatibility.jl:65-72)
    const double a0 = v[1][0] - v[0][0], a1 = v[1][1] - v[0][1], a2 = v[1][2] - v[0][2];
    const double b0 = v[2][0] - v[0][0], b1 = v[2][1] - v[0][1], b2 = v[2][2] - v[0][2];
    double nz[3] = {a1 * b2 - a2 * b1, a2 * b0 - a0 * b2, a0 * b1 - a1 * b0};
    const double nzn = sqrt(nz[0] * nz[0] + nz[1] * nz[1] + nz[2] * nz[2]);
    double nx[3] = {b0, b1, b2};
    const double nxn = sqrt(b0 * b0 + b1 * b1 + b2 * b2);
#pragma unroll
    for (int a = 0; a < 3; ++a) { nz[a] /= nzn; nx[a] /= nxn; }
    const double ny[3] = {nz[1] * nx[2] - nz[2] * nx[1], nz[2] * nx[0] - nz[0] * nx[2], nz[0] * nx[1] - nz[1] * nx[0]};
#pragma unroll
    for (int a = 0; a < 3; ++a) { P.ex[a] = nx[a]; P.ey[a] = ny[a]; P.ez[a] = nz[a]; }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double d0 = v[k][0] - P.c[0], d1 = v[k][1] - P.c[1], d2 = v[k][2] - P.c[2];
        P.vx[k] = nx[0] * d0 + nx[1] * d1 + nx[2] * d2;
        P.vy[k] = ny[0] * d0 + ny[1] * d1 + ny[2] * d2;
    }
    P.vx[NV] = P.vx[0];
    P.vy[NV] = P.vy[0];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        P.dx[k] = P.vx[k + 1] - P.vx[k];
        P.dy[k] = P.vy[k + 1] - P.vy[k];
        P.ds[k] = sqrt(P.dx[k] * P.dx[k] + P.dy[k] * P.dy[k]);
        P.ms[k] = P.dy[k] / P.dx[k];
    }
}

// one (target, panel) pair of strength st; accumulates [phi, velocity slot (3), Hessian slot (9)] into a
template <int NV, bool DIPOLE, bool PS, bool GS, bool HS>
__device__ inline void panel_pair(const PanelStage<NV>& P, double tx, double ty, double tz, double* a, double st) {
    const double EPS = 2.220446049250313e-16;   // the reference's eps() perturbation
    const double d0 = tx - P.c[0], d1 = ty - P.c[1], d2 = tz - P.c[2];
    const double tRx = P.ex[0] * d0 + P.ex[1] * d1 + P.ex[2] * d2;
    const double tRy = P.ey[0] * d0 + P.ey[1] * d1 + P.ey[2] * d2;
    double dz = P.ez[0] * d0 + P.ez[1] * d1 + P.ez[2] * d2;
    if (dz == 0.0) dz = EPS;
    const double dz2 = dz * dz;
    double rxs[NV + 1], rys[NV + 1], es[NV + 1], hs[NV + 1], rs[NV + 1];
#pragma unroll
    for (int i = 0; i <= NV; ++i) {
        rxs[i] = tRx - P.vx[i];
        rys[i] = tRy - P.vy[i];
        es[i] = rxs[i] * rxs[i] + dz2;
        hs[i] = rxs[i] * rys[i];
        const double dy = P.vy[i] - tRy;
        rs[i] = sqrt(es[i] + dy * dy);
    }
    double pot = 0.0, u0 = 0.0, u1 = 0.0, u2 = 0.0;
    double h[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};   // xx, xy, xz, yy, yz, zz
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double ri = rs[i], rip1 = rs[i + 1];
        const double r_plus = ri + rip1;
        double num = r_plus - P.ds[i];
        if (num == 0.0) num += EPS;
        const double log_term = (!DIPOLE && (PS || GS)) ? log(num / (r_plus + P.ds[i])) : 0.0;
        if (ri == 0.0) ri += EPS;
        if (rip1 == 0.0) rip1 += EPS;
        const double tan_term = (PS || !DIPOLE) ? atan((P.ms[i] * es[i] - hs[i]) / dz / ri) -
                                                    atan((P.ms[i] * es[i + 1] - hs[i + 1]) / dz / rip1)
                                              : 0.0;
        const double dx = P.dx[i], dy = P.dy[i];
        const double r_times = rs[i] * rs[i + 1];
        const double rho = r_times + rxs[i] * rxs[i + 1] + rys[i] * rys[i + 1] + dz2;
        const double lambda = rxs[i] * rys[i + 1] - rxs[i + 1] * rys[i];
        if (!DIPOLE) {
            if (PS) { pot += (rxs[i] * dy - rys[i] * dx) / P.ds[i] * log_term; pot += dz * tan_term; }
            if (GS) { u0 += dy / P.ds[i] * log_term; u1 += -dx / P.ds[i] * log_term; u2 += tan_term; }
            if (HS) {
                const double val1 = r_plus * r_plus - P.ds[i] * P.ds[i];
                const double val2 = rxs[i] / rs[i] + rxs[i + 1] / rs[i + 1];
                const double val3 = rys[i] / rs[i] + rys[i + 1] / rs[i + 1];
                const double val4 = r_plus / (r_times * rho);
                h[0] += 2 * dy / val1 * val2;
                h[1] += -2 * dx / val1 * val2;
                h[2] += dz * dy * val4;
                h[3] += -2 * dx / val1 * val3;
                h[4] += -dz * dx * val4;
                h[5] += lambda * val4;
            }
        } else {
            if (PS) pot -= tan_term;
            if (GS) {
                const double val4 = r_plus / (r_times * rho);
                u0 += dz * dy * val4;
                u1 += -dz * dx * val4;
                u2 += lambda * val4;
            }
            if (HS) {
                const double rp2 = r_plus * r_plus;
                const double val1 = (r_times * rp2 + rho * rs[i + 1] * rs[i + 1]) / (rho * rs[i] * r_plus);
                const double val2 = (r_times * rp2 + rho * rs[i] * rs[i]) / (rho * rs[i + 1] * r_plus);
                const double val3 = r_plus / (rho * r_times * r_times);
                h[0] += dz * dy * val3 * (rxs[i] * val1 + rxs[i + 1] * val2);
                h[1] += dz * dy * val3 * (rys[i] * val1 + rys[i + 1] * val2);
                h[3] += -dz * dx * val3 * (rys[i] * val1 + rys[i + 1] * val2);
                const double val6 = dz * (rp2 / rho + (rs[i] * rs[i] - r_times + rs[i + 1] * rs[i + 1]) / r_times);
                h[5] += lambda * val3 * val6;
                const double val8 = val3 * (r_times - dz * val6);
                h[2] += -dy * val8;
                h[4] += dx * val8;
            }
        }
    }
    const double s4 = st * PANEL_INV4PI;
    if (PS) a[0] += pot * s4;
    if (GS) {   // velocity = (+-) s/(4 pi) R u, R = [ex ey ez] as columns
        const double sv = DIPOLE ? s4 : -s4;
#pragma unroll
        for (int c = 0; c < 3; ++c) a[1 + c] += sv * (P.ex[c] * u0 + P.ey[c] * u1 + P.ez[c] * u2);
    }
    if (HS) {   // -s/(4 pi) R h R'
        const double H[3][3] = {{h[0], h[1], h[2]}, {h[1], h[3], h[4]}, {h[2], h[4], h[5]}};
        const double* E[3] = {P.ex, P.ey, P.ez};
        double RH[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int k = 0; k < 3; ++k) RH[r][k] = E[0][r] * H[0][k] + E[1][r] * H[1][k] + E[2][r] * H[2][k];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) a[4 + 3 * r + c] += -s4 * (RH[r][0] * E[0][c] + RH[r][1] * E[1][c] + RH[r][2] * E[2][c]);
    }
}

// one (target, panel) pair of a panel kind: MODE 0 a source panel, 1 a dipole panel, 2 a source-dipole panel
// (the source's closed form with the first strength, then the dipole's with the second, test/panels.jl:323-429)
template <int NV, int MODE, bool PS, bool GS, bool HS>
__device__ inline void panel_pair_kind(const PanelStage<NV>& P, double tx, double ty, double tz, double* a) {
    if (MODE == 1) panel_pair<NV, true, PS, GS, HS>(P, tx, ty, tz, a, P.strength);
    else panel_pair<NV, false, PS, GS, HS>(P, tx, ty, tz, a, P.strength);
    if (MODE == 2) panel_pair<NV, true, PS, GS, HS>(P, tx, ty, tz, a, P.strength2);
}

}  // namespace fmmhip
