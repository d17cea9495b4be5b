// cmp_device.hpp - pointwise five-equation physics, scalar reconstructions and the face flux of
// the gfx950 kernels.
//
// FP64 throughout. The formulas are the reference's (file:line on each function), arranged for
// the GPU: primitives are computed once per staged cell (one division), WENO5 weights use the
// product form d_k prod_{j!=k} IS_j^2 / sum (one division instead of nine, with the reference's
// ratio form kept for the cases the product form cannot represent), the Riemann solver works on
// the face primitives it is given instead of re-deriving them, and the HLLC flux evaluates only
// the upwind star state. Results agree with the reference's operation order to rounding.
#pragma once

#include <hip/hip_runtime.h>
#include "../../include/compas_amd.h"

namespace cmp {

// Kernel-side parameter block (passed by value as a kernel argument).
struct KParams {
    double A[CMP_MAX_PHASES], B[CMP_MAX_PHASES], C[CMP_MAX_PHASES], D[CMP_MAX_PHASES];
    int source_term;
    int nasg;          // any C_k > 0
    int thinc;
    int thinc_use_normal;
    double thinc_eps, thinc_beta, thinc_thr;
    double dx[3];
    double inv_dx[3];
    double weno5b_weight;   // quadrature weight of the WENO5B bound limiter (1/12 for ID_Poly = 4)
    double pf_pmin[CMP_MAX_PHASES];    // phasic energy floors of the phasic Phase-Field formulation
};

__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }  // std::min
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }  // std::max
__device__ __forceinline__ double dsign(double x) { return (x > 0.0) ? 1.0 : ((x < 0.0) ? -1.0 : 0.0); }

// c = sqrt(c^2); c^2 <= 0 (p + p_inf <= 0 in the mixture or a phase) sets bad, where the reference's
// sound speed would be NaN or zero.
__device__ __forceinline__ double checked_sqrt(double c2, int& bad)
{
    if (c2 <= 0.0) bad = CMP_ERR_PRESSURE;
    return sqrt(c2);
}

// NASG_Phi (include/EquationOfState_NASG.H:34-41); the covolume violation sets bad.
__device__ __forceinline__ double nasg_phi(double alpha, double alpha_rho, double b, int& bad)
{
    if (b > 0.0) {
        if (alpha > 0.0 && !(1.0 - b * (alpha_rho / alpha) > 0.0)) bad = CMP_ERR_COVOLUME;
        return alpha - b * alpha_rho;
    }
    return alpha;
}

// Mixture sums of EquationOfState_Mixture
// (include/Physics_Compressible5Eq/EquationOfState_Mixture_Compressible5Eq.H:15-54).
__device__ __forceinline__ void mix_AB(const KParams& kp, double a1, double a1r1, double a2r2,
                                       double& AP, double& BP, double& rho_D, int& bad)
{
    const double a2 = 1.0 - a1;
    double phi1 = a1, phi2 = a2;
    if (kp.nasg) {
        phi1 = nasg_phi(a1, a1r1, kp.C[0], bad);
        phi2 = nasg_phi(a2, a2r2, kp.C[1], bad);
    }
    rho_D = a1r1 * kp.D[0] + a2r2 * kp.D[1];
    AP = phi1 * kp.A[0] + phi2 * kp.A[1];
    BP = phi1 * kp.B[0] + phi2 * kp.B[1];
}

// Conservative2Primitive (include/Physics_Compressible5Eq/Conservative2Primitive_Compressible5Eq.H:15-61),
// with 1/rho and 1/AP from one division.
template <int DIM>
__device__ __forceinline__ void cons2prim(const KParams& kp, const double* U, double* W, int& bad)
{
    const double a1r1 = U[0], a2r2 = U[1];
    const double rho = a1r1 + a2r2;
    if (rho <= 0.0) bad = CMP_ERR_DENSITY;
    double AP, BP, rho_D;
    mix_AB(kp, U[3 + DIM], a1r1, a2r2, AP, BP, rho_D, bad);
    const double inv = 1.0 / (rho * AP);
    const double ir = AP * inv;
    double ke = 0.0;
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
        const double v = U[2 + m] * ir;
        W[2 + m] = v;
        ke = (m == 0) ? v * v : ke + v * v;
    }
    const double rho_e = U[2 + DIM] - (0.5 * rho) * ke;
    W[0] = a1r1;
    W[1] = a2r2;
    W[2 + DIM] = (rho_e - rho_D - BP) * (rho * inv);
    W[3 + DIM] = U[3 + DIM];
}

// Primitive2Conservative (include/Physics_Compressible5Eq/Primitive2Conservative_Compressible5Eq.H:15-62);
// also returns AP, BP and rho for the sound speed.
template <int DIM>
__device__ __forceinline__ void prim2cons(const KParams& kp, const double* W, double* U, double& AP,
                                          double& BP, double& rho, int& bad)
{
    const double a1r1 = W[0], a2r2 = W[1];
    rho = a1r1 + a2r2;
    double ke = 0.0;
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
        U[2 + m] = rho * W[2 + m];
        ke = (m == 0) ? W[2 + m] * W[2 + m] : ke + W[2 + m] * W[2 + m];
    }
    double rho_D;
    mix_AB(kp, W[3 + DIM], a1r1, a2r2, AP, BP, rho_D, bad);
    const double rho_e = AP * W[2 + DIM] + BP + rho_D;
    U[0] = a1r1;
    U[1] = a2r2;
    U[2 + DIM] = rho_e + (0.5 * rho) * ke;
    U[3 + DIM] = W[3 + DIM];
}

// SoundSpeed with rho_c2_Phase / rho_c2_Mixture
// (include/Physics_Compressible5Eq/SoundSpeed_Compressible5Eq.H:15-54, rho_c2_Phase_…H:17-23,
//  rho_c2_Mixture_…H:15-49). K = 0 without covolume: sum_k alpha_k A_k rho_k c_k^2 = (AP+1) p + BP,
// so c^2 = ((AP+1) p + BP)/(AP rho); K = 1 (Wood) and NASG use the general sums.
__device__ __forceinline__ double sound_speed(const KParams& kp, double p, double a1, double a1r1,
                                              double a2r2, double rho, double AP, double BP, int& bad)
{
    if (kp.source_term == 0 && !kp.nasg) return checked_sqrt(((AP + 1.0) * p + BP) / (AP * rho), bad);
    const double g1 = (kp.A[0] + 1.0) * p + kp.B[0];
    const double g2 = (kp.A[1] + 1.0) * p + kp.B[1];
    const double a2 = 1.0 - a1;
    double phi1 = a1, phi2 = a2;
    if (kp.nasg) {
        phi1 = nasg_phi(a1, a1r1, kp.C[0], bad);
        phi2 = nasg_phi(a2, a2r2, kp.C[1], bad);
    }
    if (kp.source_term == 0) return checked_sqrt((a1 * g1 + a2 * g2) / (AP * rho), bad);
    // Wood: rho c^2 = 1/(phi1 A1/g1 + phi2 A2/g2)
    return checked_sqrt(g1 * g2 / ((phi1 * kp.A[0] * g2 + phi2 * kp.A[1] * g1) * rho), bad);
}

// Kapila coefficient phi1 (1 + K1) of the alpha1 equation's div(u) term
// (FVM_SurfaceIntegral_NC, include/Operators_Compressible5Eq/FVM_SurfaceIntegralNC_Compressible5Eq.H:29-44,
//  NonConservative include/Physics_Compressible5Eq/NonConservative_Compressible5Eq.H:15-58).
template <int DIM>
__device__ __forceinline__ double nc_coefficient(const KParams& kp, const double* U, int& bad)
{
    const double a1 = U[3 + DIM];
    if (kp.source_term == 0) return a1;
    double W[4 + DIM];
    cons2prim<DIM>(kp, U, W, bad);
    const double p = W[2 + DIM];
    const double g1 = (kp.A[0] + 1.0) * p + kp.B[0];   // A1 rho1 c1^2
    const double g2 = (kp.A[1] + 1.0) * p + kp.B[1];
    double phi1 = a1, phi2 = 1.0 - a1;
    if (kp.nasg) {
        phi1 = nasg_phi(a1, U[0], kp.C[0], bad);
        phi2 = nasg_phi(1.0 - a1, U[1], kp.C[1], bad);
    }
    // phi1 (1 + K1) = phi1 rho c^2 / (rho1 c1^2) with Wood's rho c^2
    return phi1 * (kp.A[0] * g2) / (phi1 * kp.A[0] * g2 + phi2 * kp.A[1] * g1);
}

// ---- WENO5 (include/Operators_FV/WENO5.H:12-75, Reconstruction2Face_WENO5_D.H:12-37) ------------

// Candidates and smoothness indicators of one side. FACE = +1: high face of cell c.
template <int FACE>
__device__ __forceinline__ void weno5_parts(double f_mm, double f_m, double f_c, double f_p, double f_pp,
                                            double* f, double* IS)
{
    if (FACE == 1) {
        f[0] = f_c + (2.0 / 3.0) * (f_p - f_c) - (1.0 / 6.0) * (f_pp - f_p);
        f[1] = f_c + (1.0 / 6.0) * (f_c - f_m) + (1.0 / 3.0) * (f_p - f_c);
        f[2] = f_c - (1.0 / 3.0) * (f_m - f_mm) + (5.0 / 6.0) * (f_c - f_m);
    } else {
        f[0] = f_c - (5.0 / 6.0) * (f_p - f_c) + (1.0 / 3.0) * (f_pp - f_p);
        f[1] = f_c - (1.0 / 3.0) * (f_c - f_m) - (1.0 / 6.0) * (f_p - f_c);
        f[2] = f_c + (1.0 / 6.0) * (f_m - f_mm) - (2.0 / 3.0) * (f_c - f_m);
    }
    const double a = f_p - f_c, b = f_pp - f_p, c = f_c - f_m, e = f_m - f_mm;
    IS[0] = (10.0 / 3.0) * a * a - (11.0 / 3.0) * a * b + (4.0 / 3.0) * b * b;
    IS[1] = -(5.0 / 3.0) * c * a + (4.0 / 3.0) * c * c + (4.0 / 3.0) * a * a;
    IS[2] = (10.0 / 3.0) * c * c - (11.0 / 3.0) * e * c + (4.0 / 3.0) * e * e;
}

// Nonlinear weights. The reference's ratio form d_k / sum_j (IS_k/IS_j)^2 d_j with 0/0 -> 1 equals
// d_k b_k / sum_j d_j b_j with b_k = prod_{j != k} IS_j^2; the IS are first scaled by a power of two
// (exact) so the products neither overflow nor underflow. When that sum vanishes (two or three IS
// equal to zero, or ratios beyond 1e-150) the ratio form itself is evaluated.
template <int FACE>
__device__ __forceinline__ void weno5_weights(const double* IS, double* w)
{
    constexpr double d0 = (FACE == 1) ? 3.0 / 10.0 : 1.0 / 10.0;
    constexpr double d2 = (FACE == 1) ? 1.0 / 10.0 : 3.0 / 10.0;
    constexpr double d1 = 1.0 - d0 - d2;
    const double M = smax(smax(IS[0], IS[1]), IS[2]);
    if (M == 0.0) {
        // locally constant data: every ratio is 0/0 -> 1, i.e. the linear weights d_k / sum d
        constexpr double den = (d0 + d1) + d2;
        w[0] = d0 / den;
        w[1] = d1 / den;
        w[2] = d2 / den;
        return;
    }
    int e = 0;
    (void)frexp(M, &e);
    const double r0 = ldexp(IS[0], -e), r1 = ldexp(IS[1], -e), r2 = ldexp(IS[2], -e);
    const double q01 = r0 * r1, q02 = r0 * r2, q12 = r1 * r2;
    const double b0 = q12 * q12, b1 = q02 * q02, b2 = q01 * q01;
    const double S = d0 * b0 + d1 * b1 + d2 * b2;
    if (S > 1e-290) {
        const double inv = 1.0 / S;
        w[0] = d0 * b0 * inv;
        w[1] = d1 * b1 * inv;
        w[2] = d2 * b2 * inv;
        return;
    }
    double IS01 = IS[0] / IS[1];
    if (__builtin_isnan(IS01)) IS01 = 1.0;
    double IS02 = IS[0] / IS[2];
    if (__builtin_isnan(IS02)) IS02 = 1.0;
    double IS12 = IS[1] / IS[2];
    if (__builtin_isnan(IS12)) IS12 = 1.0;
    const double IS10 = 1.0 / IS01, IS20 = 1.0 / IS02, IS21 = 1.0 / IS12;
    w[0] = d0 / (d0 + (IS01 * IS01) * d1 + (IS02 * IS02) * d2);
    w[1] = d1 / ((IS10 * IS10) * d0 + d1 + (IS12 * IS12) * d2);
    w[2] = d2 / ((IS20 * IS20) * d0 + (IS21 * IS21) * d1 + d2);
}

// Face value sum_k w_k f_k. Fast path: with t_k = d_k prod_{j != k} IS_j^2 the value is
// (sum_k t_k f_k) / sum_k t_k, taken when that sum lies well inside the double range (no scaling,
// one division); constant data (all IS zero), extreme IS ratios, NaN and inf go through
// weno5_weights.
template <int FACE>
__device__ __forceinline__ double weno5(double f_mm, double f_m, double f_c, double f_p, double f_pp)
{
    constexpr double d0 = (FACE == 1) ? 3.0 / 10.0 : 1.0 / 10.0;
    constexpr double d2 = (FACE == 1) ? 1.0 / 10.0 : 3.0 / 10.0;
    constexpr double d1 = 1.0 - d0 - d2;
    double f[3], IS[3];
    weno5_parts<FACE>(f_mm, f_m, f_c, f_p, f_pp, f, IS);
    const double q01 = IS[0] * IS[1], q02 = IS[0] * IS[2], q12 = IS[1] * IS[2];
    const double t0 = d0 * (q12 * q12), t1 = d1 * (q02 * q02), t2 = d2 * (q01 * q01);
    const double S = (t0 + t1) + t2;
    if (S > 1e-290 && S < 1e290) return (t0 * f[0] + t1 * f[1] + t2 * f[2]) / S;
    double w[3];
    weno5_weights<FACE>(IS, w);
    return w[0] * f[0] + w[1] * f[1] + w[2] * f[2];
}

// Slope_TVD + Reconstruction2Face_MUSCL_D + Reconstruction2Face_Linear
// (include/Operators_FV/Slope_TVD.H:12-43, Reconstruction2Face_MUSCL_D.H:12-36, Limiter_*.H).
__device__ __forceinline__ double muscl(double f_m, double f_c, double f_p, double dx, double idx,
                                        int face, int limiter)
{
    const double sm = (f_c - f_m) * idx;
    const double sp = (f_p - f_c) * idx;
    double s = 0.0;
    if (limiter == 1) {
        s = smin(smax(sm, 0.0), smax(sp, 0.0)) + smax(smin(sm, 0.0), smin(sp, 0.0));
    } else if (limiter == 2) {
        s = (sm * sp > 0.0) ? 2.0 * sm * sp / (sm + sp) : 0.0;
    } else if (limiter == 3) {
        const double am = fabs(sm), ap = fabs(sp);
        s = (dsign(sm) + dsign(sp)) * smin(smin(am, ap), 0.5 * smax(am, ap));
    } else if (limiter == 4) {
        const double sc = (f_p - f_m) * (0.5 * idx);
        const double hi = smin(smin(smax(sc, 0.0), smax(2.0 * sm, 0.0)), smax(2.0 * sp, 0.0));
        const double lo = smax(smax(smin(sc, 0.0), smin(2.0 * sm, 0.0)), smin(2.0 * sp, 0.0));
        s = hi + lo;
    }
    return f_c + double(face) * 0.5 * s * dx;
}

// ---- WENO5B (include/Operators_Compressible5Eq/FVM_Reconstruction2Face_Conservative_WENO5B_D_Compressible5Eq.H:15-320)
// fC_QuadratureRule + Limiter_BoundMin with f_min = 0 (include/Operators_FV/fC_QuadratureRule.H,
// Limiter_BoundMin.H): theta = max(0, f_c)/max(0, f_c - min(fC, fL, fR)), at most 1.
__device__ __forceinline__ double theta_min0(double f_c, double fL, double fR, double W, double iw)
{
    const double fC = (f_c - W * fL - W * fR) * iw;
    const double mf = smin(smin(fC, fL), fR);
    const double th = smax(0.0, f_c) / smax(0.0, f_c - mf);
    return smin(1.0, th);
}

// Bounded face values of the phase masses and alpha1 of the middle cell of a 5-cell stencil of
// primitives (s(t, n), t = 0..4). FACE selects which face is returned; both faces enter the limiter.
template <int DIM, int FACE, typename S>
__device__ __forceinline__ void weno5b_bounded(const KParams& kp, const S& s, double& a1r1F,
                                               double& a2r2F, double& a1F)
{
    constexpr int IA = 3 + DIM;
    const double W = kp.weno5b_weight, iw = 1.0 / (1.0 - 2.0 * W);
    double arL[2], alL[2], arR[2], alR[2];
    double sumL = 0.0, sumR = 0.0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        double v[5], va[5];
#pragma unroll
        for (int t = 0; t < 5; ++t) {
            v[t] = s(t, k);
            va[t] = (k == 0) ? s(t, IA) : 1.0 - s(t, IA);
        }
        double f[3], IS[3], w[3], fa[3], ISa[3];
        weno5_parts<1>(v[0], v[1], v[2], v[3], v[4], f, IS);
        weno5_weights<1>(IS, w);
        arL[k] = w[0] * f[0] + w[1] * f[1] + w[2] * f[2];
        weno5_parts<1>(va[0], va[1], va[2], va[3], va[4], fa, ISa);
        alL[k] = w[0] * fa[0] + w[1] * fa[1] + w[2] * fa[2];
        weno5_parts<-1>(v[0], v[1], v[2], v[3], v[4], f, IS);
        weno5_weights<-1>(IS, w);
        arR[k] = w[0] * f[0] + w[1] * f[1] + w[2] * f[2];
        weno5_parts<-1>(va[0], va[1], va[2], va[3], va[4], fa, ISa);
        alR[k] = w[0] * fa[0] + w[1] * fa[1] + w[2] * fa[2];
        const double th = smin(theta_min0(v[2], arL[k], arR[k], W, iw),
                               theta_min0(va[2], alL[k], alR[k], W, iw));
        arL[k] = v[2] + th * (arL[k] - v[2]);
        alL[k] = va[2] + th * (alL[k] - va[2]);
        arR[k] = v[2] + th * (arR[k] - v[2]);
        alR[k] = va[2] + th * (alR[k] - va[2]);
        sumL += alL[k];
        sumR += alR[k];
    }
    const double isL = 1.0 / sumL, isR = 1.0 / sumR;
    double theta = 1.0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double c_ar = s(2, k);
        const double c_al = (k == 0) ? s(2, IA) : 1.0 - s(2, IA);
        arL[k] *= isL; alL[k] *= isL;
        arR[k] *= isR; alR[k] *= isR;
        theta = smin(smin(theta, theta_min0(c_ar, arL[k], arR[k], W, iw)),
                     theta_min0(c_al, alL[k], alR[k], W, iw));
    }
    const double c1 = s(2, 0), c2 = s(2, 1), ca = s(2, IA);
    const double r1 = (FACE == 1) ? arL[0] : arR[0];
    const double r2 = (FACE == 1) ? arL[1] : arR[1];
    const double ra = (FACE == 1) ? alL[0] : alR[0];
    a1r1F = c1 + theta * (r1 - c1);
    a2r2F = c2 + theta * (r2 - c2);
    a1F = ca + theta * (ra - ca);
}

// ---- THINC (include/Operators_Compressible5Eq/FVM_Reconstruction2Face_Conservative_THINC_K_Compressible5Eq.H:70-192,
//      include/Operators_FV/Reconstruction2Face_THINC_D.H:13-35) --------------------------------------
__device__ __forceinline__ double thinc_face(double beta, double sigma, double abar, int face_hi)
{
    const double a = face_hi ? 0.0 : 1.0;
    const double b = face_hi ? 1.0 : -1.0;
    const double e2a = exp(2.0 * beta * abar);
    const double arg_ln = (e2a - 1.0) / (exp(2.0 * beta) - e2a);
    const double c = 0.5 / beta * log(arg_ln);
    const double af = 0.5 * (1.0 + tanh(beta * (a + 0.5 * b * (sigma + 1.0) + c)));
    const double lo = (0.0 < af) ? af : 0.0;
    return (lo < 1.0) ? lo : 1.0;
}

// Face state of a THINC side (Reconstruct_State_From_VolumeFraction, :16-64): the phase densities,
// velocity and pressure of the base face state W, alpha1 replaced by a1n (primitive form).
template <int DIM>
__device__ __forceinline__ void thinc_state(double* W, double a1n)
{
    const double a1 = W[3 + DIM];
    const double rho1 = W[0] / a1;
    const double rho2 = W[1] / (1.0 - a1);
    W[0] = a1n * rho1;
    W[1] = (1.0 - a1n) * rho2;
    W[3 + DIM] = a1n;
}

// ---- Riemann solvers from face primitives ---------------------------------------------------------
// HLLC: Riemann_HLLC (include/Physics_Compressible5Eq/Riemann_HLLC_Compressible5Eq.H:16-194), only
// the star state of the side sign(S*) selects (both halves when S* == 0). HLL: Riemann_HLL
// (Riemann_HLL_Compressible5Eq.H:16-133). Conservative2Flux (Conservative2Flux_Compressible5Eq.H:15-80)
// gives the physical fluxes. Returns u*_n; F gets N fluxes; cmax the face wave speed (:164).
template <int DIM>
struct FaceSide {
    double W[4 + DIM];
    double U[4 + DIM];
    double rho, AP, BP, c;
};

template <int DIM>
__device__ __forceinline__ double hllc_half(const FaceSide<DIM>& s, int dir, double S, double Sb,
                                            double S_Star, double* F)
{
    const double V = s.W[2 + dir];
    const double p = s.W[2 + DIM];
    const double inv = 1.0 / (S - S_Star);
    const double q = (S - V) * inv;
    F[0] = s.W[0] * V + Sb * (q * s.U[0] - s.U[0]);
    F[1] = s.W[1] * V + Sb * (q * s.U[1] - s.U[1]);
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
        double Fm = s.U[2 + m] * V;
        double Cs = q * s.U[2 + m];
        if (m == dir) {
            Fm += p;
            Cs += q * s.rho * (S_Star - V);
        }
        F[2 + m] = Fm + Sb * (Cs - s.U[2 + m]);
    }
    const double CsE = q * s.U[2 + DIM] + (S_Star - V) * (q * s.rho * S_Star + p * inv);
    F[2 + DIM] = (s.U[2 + DIM] + p) * V + Sb * (CsE - s.U[2 + DIM]);
    F[3 + DIM] = s.W[3 + DIM] * V + Sb * (q * s.U[3 + DIM] - s.U[3 + DIM]);
    return V + Sb * (q - 1.0);
}

template <int DIM, int RIEMANN>
__device__ __forceinline__ double riemann(const FaceSide<DIM>& L, const FaceSide<DIM>& R, int dir,
                                          double* F, double& cmax)
{
    constexpr int N = 4 + DIM;
    const double VL = L.W[2 + dir], VR = R.W[2 + dir];
    double qL = 0.0, qR = 0.0;
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
        qL = (m == 0) ? L.W[2] * L.W[2] : qL + L.W[2 + m] * L.W[2 + m];
        qR = (m == 0) ? R.W[2] * R.W[2] : qR + R.W[2 + m] * R.W[2 + m];
    }
    cmax = smax(L.c + sqrt(qL), R.c + sqrt(qR));
    if (RIEMANN == 2) {
        const double pL = L.W[2 + DIM], pR = R.W[2 + DIM];
        const double V_avg = 0.5 * (VL + VR);
        const double c_avg = 0.5 * (L.c + R.c);
        const double SL = smin(V_avg - c_avg, VL - L.c);
        const double SR = smax(V_avg + c_avg, VR + R.c);
        const double S_Star = (pR - pL + L.rho * VL * (SL - VL) - R.rho * VR * (SR - VR))
                            / (L.rho * (SL - VL) - R.rho * (SR - VR));
        if (S_Star > 0.0) return hllc_half<DIM>(L, dir, SL, smin(0.0, SL), S_Star, F);
        if (S_Star < 0.0) return hllc_half<DIM>(R, dir, SR, smax(0.0, SR), S_Star, F);
        double FR_[N];
        const double uL = hllc_half<DIM>(L, dir, SL, smin(0.0, SL), S_Star, F);
        const double uR = hllc_half<DIM>(R, dir, SR, smax(0.0, SR), S_Star, FR_);
#pragma unroll
        for (int n = 0; n < N; ++n) F[n] = 0.5 * F[n] + 0.5 * FR_[n];
        return 0.5 * uL + 0.5 * uR;
    } else {
        const double Sp = smax(smax(0.0, VL + L.c), VR + R.c);
        const double Sm = smin(smin(0.0, VL - L.c), VR - R.c);
        const double inv = 1.0 / (Sp - Sm);
        const double pL = L.W[2 + DIM], pR = R.W[2 + DIM];
#pragma unroll
        for (int n = 0; n < N; ++n) {
            double FL, FR;
            if (n < 2 || n == 3 + DIM) { FL = L.W[n] * VL; FR = R.W[n] * VR; }
            else if (n == 2 + DIM) { FL = (L.U[n] + pL) * VL; FR = (R.U[n] + pR) * VR; }
            else {
                FL = L.U[n] * VL + ((n == 2 + dir) ? pL : 0.0);
                FR = R.U[n] * VR + ((n == 2 + dir) ? pR : 0.0);
            }
            F[n] = (Sp * FL - Sm * FR + Sp * Sm * (R.U[n] - L.U[n])) * inv;
        }
        return (Sp * VL - Sm * VR) * inv;
    }
}

// Primitive face state -> the side record the Riemann solver reads.
template <int DIM>
__device__ __forceinline__ void make_side(const KParams& kp, FaceSide<DIM>& s, int& bad)
{
    prim2cons<DIM>(kp, s.W, s.U, s.AP, s.BP, s.rho, bad);
    s.c = sound_speed(kp, s.W[2 + DIM], s.W[3 + DIM], s.W[0], s.W[1], s.rho, s.AP, s.BP, bad);
}

// ---- one face --------------------------------------------------------------------------------------
// FVM_Conservative2FluxHyperbolic_K (include/Operators_Compressible5Eq/FVM_Conservative2FluxHyperbolic_K_Compressible5Eq.H:15-170)
// for the face between stencil cells G-1 and G: s(t, n) returns primitive n of stencil cell t
// (t = 0..2G-1, cell G-1 is the low side). SCHEME: 0 Godunov, 1 MUSCL, 5 WENO5, 6 WENO5B. With
// THINC, normal() returns the face normal's component along dir (from the volume fraction); it is
// evaluated only when the volume-fraction gates already admit THINC on a side.
// THINC decision of one face (FVM_Reconstruction2Face_Conservative_THINC_K_Compressible5Eq.H:143-170):
// the interface gates of each side and, for a side that passes them, its THINC volume fraction.
// aL / aR stay negative for a side THINC leaves to the base scheme. G is the stencil half-width of
// the accessor s (cells mm, m, c, p of the face are G-2 .. G+1).
template <int DIM, int G, typename S, typename NRM>
__device__ __forceinline__ void thinc_decide(const KParams& kp, const S& s, const NRM& normal,
                                             double& aL, double& aR)
{
    constexpr int IA = 3 + DIM;
    const double amm = s(G - 2, IA), am = s(G - 1, IA), ac = s(G, IA), ap = s(G + 1, IA);
    const bool gL = (ac - am) * (am - amm) > 0.0;
    const bool gR = (ap - ac) * (ac - am) > 0.0;
    const bool sym = (amm == ap) && (am == ac) && (am != amm);
    const double eps = kp.thinc_eps;
    const bool cL = (am > eps && am < 1.0 - eps) && (gL || sym);
    const bool cR = (ac > eps && ac < 1.0 - eps) && (gR || sym);
    aL = -1.0;
    aR = -1.0;
    if (cL || cR) {
        const double nd = normal();
        const bool gate = (!kp.thinc_use_normal) || (fabs(nd) > kp.thinc_thr);
        if (gate) {
            const double beta = kp.thinc_use_normal ? (kp.thinc_beta * fabs(nd) + 0.01) : kp.thinc_beta;
            if (cL) aL = thinc_face(beta, (ac - amm > 0.0) ? 1.0 : -1.0, am, 1);
            if (cR) aR = thinc_face(beta, (ap - am > 0.0) ? 1.0 : -1.0, ac, 0);
        }
    }
}

// Reconstruction part: the primitive face states WL (low side) and WR (high side). With THINC,
// aL / aR are thinc_decide's results for this face: where either side uses THINC, both sides take
// the MUSCL-minmod base state and a THINC side replaces its volume fraction.
template <int DIM, int SCHEME, bool THINC, typename S>
__device__ __forceinline__ void recon_face_decided(const KParams& kp, int limiter, int dir, const S& s,
                                                   double aL, double aR, double* LW, double* RW)
{
    constexpr int N = 4 + DIM;
    constexpr int G = (SCHEME >= 5) ? 3 : (SCHEME == 1 ? 2 : 1);
    constexpr int IA = 3 + DIM;
    const double dx = kp.dx[dir], idx = kp.inv_dx[dir];
    struct { double* W; } L{LW}, R{RW};
    bool done = false;
    if (THINC && (aL >= 0.0 || aR >= 0.0)) {
#pragma unroll
        for (int n = 0; n < N; ++n) {
            L.W[n] = muscl(s(G - 2, n), s(G - 1, n), s(G, n), dx, idx, 1, 1);
            R.W[n] = muscl(s(G - 1, n), s(G, n), s(G + 1, n), dx, idx, -1, 1);
        }
        if (aL >= 0.0) thinc_state<DIM>(L.W, aL);
        if (aR >= 0.0) thinc_state<DIM>(R.W, aR);
        done = true;
    }
    (void)IA;
    if (!done) {
        if (SCHEME == 5 || SCHEME == 6) {
            constexpr int n0 = (SCHEME == 6) ? 2 : 0;
            constexpr int n1 = (SCHEME == 6) ? 2 + DIM : N - 1;
#pragma unroll
            for (int n = n0; n <= n1; ++n) {
                const double v0 = s(0, n), v1 = s(1, n), v2 = s(2, n), v3 = s(3, n), v4 = s(4, n), v5 = s(5, n);
                L.W[n] = weno5<1>(v0, v1, v2, v3, v4);
                R.W[n] = weno5<-1>(v1, v2, v3, v4, v5);
            }
            if (SCHEME == 6) {
                auto sl = [&](int t, int n) { return s(t, n); };
                auto sr = [&](int t, int n) { return s(t + 1, n); };
                weno5b_bounded<DIM, 1>(kp, sl, L.W[0], L.W[1], L.W[IA]);
                weno5b_bounded<DIM, -1>(kp, sr, R.W[0], R.W[1], R.W[IA]);
            }
        } else if (SCHEME == 1) {
#pragma unroll
            for (int n = 0; n < N; ++n) {
                L.W[n] = muscl(s(0, n), s(1, n), s(2, n), dx, idx, 1, limiter);
                R.W[n] = muscl(s(1, n), s(2, n), s(3, n), dx, idx, -1, limiter);
            }
        } else {
#pragma unroll
            for (int n = 0; n < N; ++n) {
                L.W[n] = s(0, n);
                R.W[n] = s(1, n);
            }
        }
    }
}

// Reconstruction of one face, THINC decision included.
template <int DIM, int SCHEME, bool THINC, typename S, typename NRM>
__device__ __forceinline__ void recon_face(const KParams& kp, int limiter, int dir, const S& s,
                                           const NRM& normal, double* LW, double* RW)
{
    constexpr int G = (SCHEME >= 5) ? 3 : (SCHEME == 1 ? 2 : 1);
    double aL = -1.0, aR = -1.0;
    if (THINC) thinc_decide<DIM, G>(kp, s, normal, aL, aR);
    recon_face_decided<DIM, SCHEME, THINC>(kp, limiter, dir, s, aL, aR, LW, RW);
}

// Flow flux of one face with the THINC decision already taken (aL, aR of thinc_decide).
template <int DIM, int SCHEME, int RIEMANN, bool THINC, typename S>
__device__ __forceinline__ double face_flux_decided(const KParams& kp, int limiter, int dir, const S& s,
                                                    double aL, double aR, double* F, double& cmax,
                                                    int& bad)
{
    FaceSide<DIM> L, R;
    recon_face_decided<DIM, SCHEME, THINC>(kp, limiter, dir, s, aL, aR, L.W, R.W);
    make_side<DIM>(kp, L, bad);
    make_side<DIM>(kp, R, bad);
    return riemann<DIM, RIEMANN>(L, R, dir, F, cmax);
}

template <int DIM, int SCHEME, int RIEMANN, bool THINC, typename S, typename NRM>
__device__ __forceinline__ double face_flux(const KParams& kp, int limiter, int dir, const S& s,
                                            const NRM& normal, double* F, double& cmax, int& bad)
{
    constexpr int G = (SCHEME >= 5) ? 3 : (SCHEME == 1 ? 2 : 1);
    double aL = -1.0, aR = -1.0;
    if (THINC) thinc_decide<DIM, G>(kp, s, normal, aL, aR);
    return face_flux_decided<DIM, SCHEME, RIEMANN, THINC>(kp, limiter, dir, s, aL, aR, F, cmax, bad);
}

// ---- the N-phase model (include/Physics_Compressible5Eq_NPhase/, Operators_Compressible5Eq_NPhase/) ---
// State alpha_k rho_k (k < NP), rho u_d, rho E, alpha_k (k < NP); primitives alpha_k rho_k, u_d, p,
// alpha_k (Variables_Compressible5Eq_NPhase.H:21-29). Every phase's own alpha_k enters the mixture.
template <int DIM, int NP>
struct NpIdx {
    static constexpr int N = 2 * NP + DIM + 1;
    static constexpr int MOM = NP;            // first momentum component
    static constexpr int E = NP + DIM;        // energy (conservative) / pressure (primitive)
    static constexpr int A = NP + DIM + 1;    // first volume fraction
};

// Sums of EquationOfState_Mixture (…EquationOfState_Mixture_Compressible5Eq_NPhase.H:15-67).
template <int NP>
__device__ __forceinline__ void np_mix(const KParams& kp, const double* alpha, const double* arho,
                                       double& AP, double& BP, double& rho_D, int& bad)
{
    rho_D = arho[0] * kp.D[0];
#pragma unroll
    for (int k = 1; k < NP; ++k) rho_D += arho[k] * kp.D[k];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const double phi = kp.nasg ? nasg_phi(alpha[k], arho[k], kp.C[k], bad) : alpha[k];
        AP = (k == 0) ? phi * kp.A[k] : AP + phi * kp.A[k];
        BP = (k == 0) ? phi * kp.B[k] : BP + phi * kp.B[k];
    }
}

// Conservative2Primitive (…Conservative2Primitive_Compressible5Eq_NPhase.H:16-60), 1/rho and 1/AP
// from one division.
template <int DIM, int NP>
__device__ __forceinline__ void np_cons2prim(const KParams& kp, const double* U, double* W, int& bad)
{
    using I = NpIdx<DIM, NP>;
    double rho = U[0];
#pragma unroll
    for (int k = 1; k < NP; ++k) rho += U[k];
    if (rho <= 0.0) bad = CMP_ERR_DENSITY;
    double AP, BP, rho_D;
    np_mix<NP>(kp, U + I::A, U, AP, BP, rho_D, bad);
    const double inv = 1.0 / (rho * AP);
    const double ir = AP * inv;
    double ke = 0.0;
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
        const double v = U[I::MOM + m] * ir;
        W[I::MOM + m] = v;
        ke = (m == 0) ? v * v : ke + v * v;
    }
    const double rho_e = U[I::E] - (0.5 * rho) * ke;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        W[k] = U[k];
        W[I::A + k] = U[I::A + k];
    }
    W[I::E] = (rho_e - rho_D - BP) * (rho * inv);
}

// Primitive2Conservative (…Primitive2Conservative_Compressible5Eq_NPhase.H:15-62); AP, BP, rho out.
template <int DIM, int NP>
__device__ __forceinline__ void np_prim2cons(const KParams& kp, const double* W, double* U, double& AP,
                                             double& BP, double& rho, int& bad)
{
    using I = NpIdx<DIM, NP>;
    rho = W[0];
#pragma unroll
    for (int k = 1; k < NP; ++k) rho += W[k];
    double ke = 0.0;
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
        U[I::MOM + m] = rho * W[I::MOM + m];
        ke = (m == 0) ? W[I::MOM] * W[I::MOM] : ke + W[I::MOM + m] * W[I::MOM + m];
    }
    double rho_D;
    np_mix<NP>(kp, W + I::A, W, AP, BP, rho_D, bad);
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        U[k] = W[k];
        U[I::A + k] = W[I::A + k];
    }
    U[I::E] = (AP * W[I::E] + BP + rho_D) + (0.5 * rho) * ke;
}

// SoundSpeed with rho_c2_Phase / rho_c2_Mixture (…SoundSpeed_Compressible5Eq_NPhase.H:15-52,
// …rho_c2_Mixture_Compressible5Eq_NPhase.H:15-52): K = 0 rho c^2 = sum_k alpha_k A_k rho_k c_k^2 / AP
// with A_k rho_k c_k^2 = (A_k + 1) p + B_k; K = 1 Wood's 1/sum_k phi_k/(rho_k c_k^2).
template <int NP>
__device__ __forceinline__ double np_sound_speed(const KParams& kp, double p, const double* alpha,
                                                 const double* arho, double rho, double AP, int& bad)
{
    if (kp.source_term == 0) {
        double s = alpha[0] * ((kp.A[0] + 1.0) * p + kp.B[0]);
#pragma unroll
        for (int k = 1; k < NP; ++k) s += alpha[k] * ((kp.A[k] + 1.0) * p + kp.B[k]);
        return checked_sqrt(s / (AP * rho), bad);
    }
    double inv = 0.0;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const double phi = kp.nasg ? nasg_phi(alpha[k], arho[k], kp.C[k], bad) : alpha[k];
        inv += phi * kp.A[k] / ((kp.A[k] + 1.0) * p + kp.B[k]);
    }
    return checked_sqrt(1.0 / (inv * rho), bad);
}

// Kapila coefficients phi_k (1 + K_k) of the alpha_k equations (FVM_SurfaceIntegral_NC,
// …FVM_SurfaceIntegralNC_Compressible5Eq_NPhase.H:16-56; NonConservative …NPhase.H:15-58).
template <int DIM, int NP>
__device__ __forceinline__ void np_nc_coefficients(const KParams& kp, const double* U, double* f, int& bad)
{
    using I = NpIdx<DIM, NP>;
    if (kp.source_term == 0) {
#pragma unroll
        for (int k = 0; k < NP; ++k) f[k] = U[I::A + k];
        return;
    }
    double W[I::N];
    np_cons2prim<DIM, NP>(kp, U, W, bad);
    const double p = W[I::E];
    double g[NP], inv = 0.0, phi[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        g[k] = ((kp.A[k] + 1.0) * p + kp.B[k]) / kp.A[k];   // rho_k c_k^2
        phi[k] = kp.nasg ? nasg_phi(U[I::A + k], U[k], kp.C[k], bad) : U[I::A + k];
        inv += phi[k] / g[k];
    }
    const double rc2 = 1.0 / inv;
#pragma unroll
    for (int k = 0; k < NP; ++k) f[k] = phi[k] * (1.0 + (rc2 / g[k] - 1.0));
}

// Limiter_Consistent (include/Operators_FV/Limiter_Consistent.H:12-39).
template <int NP>
__device__ __forceinline__ void np_limiter_consistent(const double* s, double* th)
{
    double S = s[0];
    int imax = 0;
#pragma unroll
    for (int k = 1; k < NP; ++k) {
        S += s[k];
        if (fabs(s[imax]) < fabs(s[k])) imax = k;
    }
    double smx = s[0];
#pragma unroll
    for (int k = 1; k < NP; ++k)
        if (k == imax) smx = s[k];
    if (fabs(smx) < 1.0e-10 || fabs(S) < 1.0e-10) {
#pragma unroll
        for (int k = 0; k < NP; ++k) th[k] = 1.0;
        return;
    }
    const double smS = smx - S;
    const double to = smax(smin(smx / smS, 1.0), 0.0);
    const double tm = smax(smin(smS / smx, 1.0), 0.0);
#pragma unroll
    for (int k = 0; k < NP; ++k) th[k] = (k == imax) ? tm : to;
}

// FVM_Reconstruction2Face_Conservative_LimiterConsistent (…LimiterConsistent_Compressible5Eq_NPhase.H:15-73)
// on a primitive face state W of the side whose cell primitives are c(n): prim->cons->prim as the
// reference takes it, the alpha_k rho_k and alpha_k slopes to the face, theta from the alpha slopes,
// the linear face values with theta times the slope.
template <int DIM, int NP, typename C>
__device__ __forceinline__ void np_limit_face(const KParams& kp, double* W, const C& c, double dx, int face,
                                              int& bad)
{
    using I = NpIdx<DIM, NP>;
    double U[I::N], AP, BP, rho;
    np_prim2cons<DIM, NP>(kp, W, U, AP, BP, rho, bad);
    np_cons2prim<DIM, NP>(kp, U, W, bad);
    const double h = double(face) * 0.5 * dx;
    double sr[NP], sa[NP], th[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        sr[k] = (W[k] - c(k)) / h;
        sa[k] = (W[I::A + k] - c(I::A + k)) / h;
    }
    np_limiter_consistent<NP>(sa, th);
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        W[k] = c(k) + double(face) * 0.5 * (th[k] * sr[k]) * dx;
        W[I::A + k] = c(I::A + k) + double(face) * 0.5 * (th[k] * sa[k]) * dx;
    }
}

template <int DIM, int NP>
struct FaceSideN {
    double W[NpIdx<DIM, NP>::N];
    double U[NpIdx<DIM, NP>::N];
    double rho, AP, BP, c;
};

template <int DIM, int NP>
__device__ __forceinline__ void np_make_side(const KParams& kp, FaceSideN<DIM, NP>& s, int& bad)
{
    using I = NpIdx<DIM, NP>;
    np_prim2cons<DIM, NP>(kp, s.W, s.U, s.AP, s.BP, s.rho, bad);
    s.c = np_sound_speed<NP>(kp, s.W[I::E], s.W + I::A, s.W, s.rho, s.AP, bad);
}

// One side's HLLC flux (Riemann_HLLC …NPhase.H:16-162) with the star state of the side S* picks.
template <int DIM, int NP>
__device__ __forceinline__ double np_hllc_half(const FaceSideN<DIM, NP>& s, int dir, double S, double Sb,
                                               double S_Star, double* F)
{
    using I = NpIdx<DIM, NP>;
    const double V = s.W[I::MOM + dir];
    const double p = s.W[I::E];
    const double inv = 1.0 / (S - S_Star);
    const double q = (S - V) * inv;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        F[k] = s.W[k] * V + Sb * (q * s.U[k] - s.U[k]);
        F[I::A + k] = s.W[I::A + k] * V + Sb * (q * s.U[I::A + k] - s.U[I::A + k]);
    }
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
        double Fm = s.U[I::MOM + m] * V;
        double Cs = q * s.U[I::MOM + m];
        if (m == dir) {
            Fm += p;
            Cs += q * s.rho * (S_Star - V);
        }
        F[I::MOM + m] = Fm + Sb * (Cs - s.U[I::MOM + m]);
    }
    const double CsE = q * s.U[I::E] + (S_Star - V) * (q * s.rho * S_Star + p * inv);
    F[I::E] = (s.U[I::E] + p) * V + Sb * (CsE - s.U[I::E]);
    return V + Sb * (q - 1.0);
}

// HLLC (RIEMANN 2) or HLL (1, Riemann_HLL …NPhase.H) of the N-phase model; returns u*_n.
template <int DIM, int RIEMANN, int NP>
__device__ __forceinline__ double np_riemann(const FaceSideN<DIM, NP>& L, const FaceSideN<DIM, NP>& R, int dir,
                                             double* F, double& cmax)
{
    using I = NpIdx<DIM, NP>;
    constexpr int N = I::N;
    const double VL = L.W[I::MOM + dir], VR = R.W[I::MOM + dir];
    double qL = 0.0, qR = 0.0;
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
        qL = (m == 0) ? L.W[I::MOM] * L.W[I::MOM] : qL + L.W[I::MOM + m] * L.W[I::MOM + m];
        qR = (m == 0) ? R.W[I::MOM] * R.W[I::MOM] : qR + R.W[I::MOM + m] * R.W[I::MOM + m];
    }
    cmax = smax(L.c + sqrt(qL), R.c + sqrt(qR));
    const double pL = L.W[I::E], pR = R.W[I::E];
    if (RIEMANN == 2) {
        const double V_avg = 0.5 * (VL + VR);
        const double c_avg = 0.5 * (L.c + R.c);
        const double SL = smin(V_avg - c_avg, VL - L.c);
        const double SR = smax(V_avg + c_avg, VR + R.c);
        const double S_Star = (pR - pL + L.rho * VL * (SL - VL) - R.rho * VR * (SR - VR))
                            / (L.rho * (SL - VL) - R.rho * (SR - VR));
        if (S_Star > 0.0) return np_hllc_half<DIM, NP>(L, dir, SL, smin(0.0, SL), S_Star, F);
        if (S_Star < 0.0) return np_hllc_half<DIM, NP>(R, dir, SR, smax(0.0, SR), S_Star, F);
        double FR_[N];
        const double uL = np_hllc_half<DIM, NP>(L, dir, SL, smin(0.0, SL), S_Star, F);
        const double uR = np_hllc_half<DIM, NP>(R, dir, SR, smax(0.0, SR), S_Star, FR_);
#pragma unroll
        for (int n = 0; n < N; ++n) F[n] = 0.5 * F[n] + 0.5 * FR_[n];
        return 0.5 * uL + 0.5 * uR;
    } else {
        const double Sp = smax(smax(0.0, VL + L.c), VR + R.c);
        const double Sm = smin(smin(0.0, VL - L.c), VR - R.c);
        const double inv = 1.0 / (Sp - Sm);
#pragma unroll
        for (int n = 0; n < N; ++n) {
            double FL, FR;
            if (n < NP || n >= I::A) { FL = L.W[n] * VL; FR = R.W[n] * VR; }
            else if (n == I::E) { FL = (L.U[n] + pL) * VL; FR = (R.U[n] + pR) * VR; }
            else {
                FL = L.U[n] * VL + ((n == I::MOM + dir) ? pL : 0.0);
                FR = R.U[n] * VR + ((n == I::MOM + dir) ? pR : 0.0);
            }
            F[n] = (Sp * FL - Sm * FR + Sp * Sm * (R.U[n] - L.U[n])) * inv;
        }
        return (Sp * VL - Sm * VR) * inv;
    }
}

// Primitive face states of the N-phase model (FVM_Reconstruction2FaceLR_Conservative_K …NPhase):
// WENO5 or MUSCL per primitive, then the consistent limiter; Godunov without it. s(t, n) as in
// face_flux.
template <int DIM, int SCHEME, int NP, typename S>
__device__ __forceinline__ void np_recon_face(const KParams& kp, int limiter, int dir, const S& s, double* LW,
                                              double* RW, int& bad)
{
    constexpr int N = NpIdx<DIM, NP>::N;
    const double dx = kp.dx[dir], idx = kp.inv_dx[dir];
    if (SCHEME == 5) {
#pragma unroll
        for (int n = 0; n < N; ++n) {
            const double v0 = s(0, n), v1 = s(1, n), v2 = s(2, n), v3 = s(3, n), v4 = s(4, n), v5 = s(5, n);
            LW[n] = weno5<1>(v0, v1, v2, v3, v4);
            RW[n] = weno5<-1>(v1, v2, v3, v4, v5);
        }
        np_limit_face<DIM, NP>(kp, LW, [&](int n) { return s(2, n); }, dx, 1, bad);
        np_limit_face<DIM, NP>(kp, RW, [&](int n) { return s(3, n); }, dx, -1, bad);
    } else if (SCHEME == 1) {
#pragma unroll
        for (int n = 0; n < N; ++n) {
            LW[n] = muscl(s(0, n), s(1, n), s(2, n), dx, idx, 1, limiter);
            RW[n] = muscl(s(1, n), s(2, n), s(3, n), dx, idx, -1, limiter);
        }
        np_limit_face<DIM, NP>(kp, LW, [&](int n) { return s(1, n); }, dx, 1, bad);
        np_limit_face<DIM, NP>(kp, RW, [&](int n) { return s(2, n); }, dx, -1, bad);
    } else {
#pragma unroll
        for (int n = 0; n < N; ++n) {
            LW[n] = s(0, n);
            RW[n] = s(1, n);
        }
    }
}

// One face of the N-phase model (FVM_Conservative2FluxHyperbolic_K …NPhase.H:16-150): the face
// states, then the Riemann solver; returns u*_n.
template <int DIM, int SCHEME, int RIEMANN, int NP, typename S>
__device__ __forceinline__ double np_face_flux(const KParams& kp, int limiter, int dir, const S& s, double* F,
                                               double& cmax, int& bad)
{
    FaceSideN<DIM, NP> L, R;
    np_recon_face<DIM, SCHEME, NP>(kp, limiter, dir, s, L.W, R.W, bad);
    np_make_side<DIM, NP>(kp, L, bad);
    np_make_side<DIM, NP>(kp, R, bad);
    return np_riemann<DIM, RIEMANN, NP>(L, R, dir, F, cmax);
}

// ---- phase field ------------------------------------------------------------------------------------
// Mapping_Backward / Mapping_Forward (include/Physics_PhaseField/Mapping_Backward.H:12-29,
// Mapping_Forward.H:12-26).
__device__ __forceinline__ double pf_map_backward(double phi, double eta)
{
    const double ep = 1.0e-12;
    double t = phi;
    if (t < ep) t = ep;
    else if (t > 1.0 - ep) t = 1.0 - ep;
    return 1.4142135623730951 * eta * atanh(2.0 * t - 1.0);
}

__device__ __forceinline__ double pf_map_forward(double psi, double eta)
{
    return 0.5 + 0.5 * tanh(psi / (1.4142135623730951 * eta));
}

__device__ __forceinline__ double pf_map_forward_d(double psi, double eta)
{
    const double x = psi / (1.4142135623730951 * eta);
    const double ch = cosh(x);
    return 0.5 / (1.4142135623730951 * eta) / ch / ch;
}

// ACDI face flux J = M (grad phi - F'(psi_f) n_f) (FDM_PhaseFieldFlux_ACDI_D,
// include/Operators_PhaseField/FDM_PhaseFieldFlux_ACDI_D.H:12-45; PhaseFieldFlux_ACDI.H:12-26).
// psi_m, psi_c are Mapping_Backward of pf_m, pf_c.
__device__ __forceinline__ double pf_acdi_flux(double pf_m, double pf_c, double psi_m, double psi_c,
                                               double n_m, double n_c, double mob, double eta, double dx)
{
    const double grad = (pf_c - pf_m) / dx;
    const double psi = 0.5 * psi_m + 0.5 * psi_c;
    const double nrm = 0.5 * n_m + 0.5 * n_c;
    return mob * (grad - pf_map_forward_d(psi, eta) * nrm);
}

// CAC-Adv mechanism of one cell, mech = -M mu, mu = Lambda (g'(phi)/eta^2 - lap phi - |grad phi| kappa),
// kappa = -div n with the normals averaged to the faces (FDM_PhaseFieldMechanism_ACAdv_D.H,
// FDM_ChemicalPotential_Adv_D.H:12-126, ChemicalPotential_Adv.H, PhaseFieldMechanism_AC.H:12-17,
// PotentialFunction_DoubleWell.H:22-23). fm/fp: phi at the neighbours along each direction;
// nc/nm/np: component d of the normal at the cell and its neighbours along d.
template <int DIM>
__device__ __forceinline__ double pf_cac_mech(const double* dx, double f, const double* fm,
                                              const double* fp, const double* nc, const double* nm,
                                              const double* np, double mob, double lambda, double eta)
{
    double lap = 0.0, divn = 0.0, absg = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const double gp = (fp[d] - f) / dx[d];
        const double gm = (f - fm[d]) / dx[d];
        const double t = (gp - gm) / dx[d];
        lap = (d == 0) ? t : lap + t;
        const double nfp = 0.5 * nc[d] + 0.5 * np[d];
        const double nfm = 0.5 * nm[d] + 0.5 * nc[d];
        const double tn = (nfp - nfm) / dx[d];
        divn = (d == 0) ? tn : divn + tn;
        const double g = 0.5 * gm + 0.5 * gp;
        absg += g * g;
    }
    absg = sqrt(absg);
    const double g1 = 2.0 * f * (1.0 - f) * (1.0 - 2.0 * f);
    const double mu = lambda * ((1.0 / eta / eta) * g1 - lap - absg * (-divn));
    return -mob * mu;
}

// Explicit Grad_Q of one face (source/Class_Compressible_PhaseField.cpp:2358-2391,
// include/Physics_PhaseField/PhaseFieldFlux2GradQPF.H:12-28) from the re-sharpened cell values
// m = F(B(alpha_m, eta), eta_factor eta), c likewise.
__device__ __forceinline__ double pf_gradq_sharp(double m, double c, double jhat)
{
    const double epsilon = 1.0e-3;
    const double pf = 0.5 * m + 0.5 * c;
    const double flux = (pf < epsilon || pf > 1.0 - epsilon) ? 0.0 : jhat;
    return flux / (pf * (1.0 - pf) + epsilon);
}

// f_Upwind / f_Downwind (include/Operators_FV/f_Upwind.H:12-27, f_Downwind.H:12-27)
__device__ __forceinline__ double f_up(double fL, double fR, double V)
{
    return (V > 0.0) ? fL : ((V < 0.0) ? fR : 0.5 * (fL + fR));
}
__device__ __forceinline__ double f_down(double fL, double fR, double V)
{
    return (V > 0.0) ? fR : ((V < 0.0) ? fL : 0.5 * (fL + fR));
}

// Phasic internal energy alpha_k rho_k e_k from p, floored at p_min
// (EquationOfState_Phase, Conservative2ConservativePhase_Compressible5Eq.H:34-91).
template <int DIM>
__device__ __forceinline__ void pf_phase_energies(const KParams& kp, const double* W, double& e1,
                                                  double& e2, int& bad)
{
    const double a1 = W[3 + DIM], a2 = 1.0 - a1, p = W[2 + DIM];
    const double phi1 = nasg_phi(a1, W[0], kp.C[0], bad);
    const double phi2 = nasg_phi(a2, W[1], kp.C[1], bad);
    e1 = phi1 * kp.A[0] * p + phi1 * kp.B[0] + W[0] * kp.D[0];
    const double e1m = phi1 * kp.A[0] * kp.pf_pmin[0] + phi1 * kp.B[0] + W[0] * kp.D[0];
    e2 = phi2 * kp.A[1] * p + phi2 * kp.B[1] + W[1] * kp.D[1];
    const double e2m = phi2 * kp.A[1] * kp.pf_pmin[1] + phi2 * kp.B[1] + W[1] * kp.D[1];
    const double rho_e = e1 + e2;
    if (e1 < e1m) { e1 = e1m; e2 = rho_e - e1; }
    else if (e2 < e2m) { e2 = e2m; e1 = rho_e - e2; }
}

// Phase-Field flux fhat of one face from the primitive face states: FluxPhaseField
// (include/Physics_Compressible5Eq_PhaseField/FluxPhaseField_Compressible5Eq_PhaseField.H:12-65) with
// the UD phasic flux (…/FluxPhasePhaseField_UD_Compressible_TwoPhase_PhaseField.H:12-194) and the
// mixture (FORM 0, FluxPhaseField_Mixture_Compressible5Eq_PhaseField.H:12-116) or phasic
// (FORM 1, VectorPhase2Vector_Compressible5Eq.H:15-54) formulation. Grad_Q1 = g, Grad_Q2 = -g.
template <int DIM, int FORM>
__device__ __forceinline__ void pf_flux(const KParams& kp, const double* WL, const double* WR, double g,
                                        double* F, int& bad)
{
    constexpr int IA = 3 + DIM;
    const double g1 = g, g2 = -g;
    const double a1L = WL[IA], a1R = WR[IA];
    const double a2L = 1.0 - a1L, a2R = 1.0 - a1R;
    const double a2u = f_up(a2L, a2R, g1);
    const double a1u = f_up(a1L, a1R, g2);
    const double J1r1 = f_down(WL[0], WR[0], g1) * a2u * g1;
    const double J2r2 = f_down(WL[1], WR[1], g2) * a1u * g2;
    const double J1 = f_down(a1L, a1R, g1) * a2u * g1;
    F[0] = J1r1;
    F[1] = J2r2;
    F[IA] = J1;
    if (FORM == 0) {
        const double J2 = f_down(a2L, a2R, g2) * a1u * g2;
        const double M = J1r1 + J2r2;
        double ke = 0.0;
#pragma unroll
        for (int m = 0; m < DIM; ++m) {
            const double u = f_down(WL[2 + m], WR[2 + m], M);
            F[2 + m] = M * u;
            ke = (m == 0) ? u * u : ke + u * u;
        }
        const double MeK = M * (0.5 * ke);
        const double p1 = f_down(WL[2 + DIM], WR[2 + DIM], J1);
        const double J1phi = (kp.C[0] > 0.0) ? J1 - kp.C[0] * J1r1 : J1;
        const double J1e1 = kp.A[0] * J1phi * p1 + kp.B[0] * J1phi + kp.D[0] * J1r1;
        const double p2 = f_down(WL[2 + DIM], WR[2 + DIM], J2);
        const double J2phi = (kp.C[1] > 0.0) ? J2 - kp.C[1] * J2r2 : J2;
        const double J2e2 = kp.A[1] * J2phi * p2 + kp.B[1] * J2phi + kp.D[1] * J2r2;
        F[2 + DIM] = MeK + J1e1 + J2e2;
    } else {
        double keL = 0.0, keR = 0.0;
#pragma unroll
        for (int m = 0; m < DIM; ++m) {
            const double uL = WL[2 + m], uR = WR[2 + m];
            keL = (m == 0) ? uL * uL : keL + uL * uL;
            keR = (m == 0) ? uR * uR : keR + uR * uR;
            const double m1 = f_down(WL[0] * uL, WR[0] * uR, g1) * a2u * g1;
            const double m2 = f_down(WL[1] * uL, WR[1] * uR, g2) * a1u * g2;
            F[2 + m] = m1 + m2;
        }
        double e1L, e2L, e1R, e2R;
        pf_phase_energies<DIM>(kp, WL, e1L, e2L, bad);
        pf_phase_energies<DIM>(kp, WR, e1R, e2R, bad);
        const double E1L = e1L + 0.5 * WL[0] * keL, E1R = e1R + 0.5 * WR[0] * keR;
        const double E2L = e2L + 0.5 * WL[1] * keL, E2R = e2R + 0.5 * WR[1] * keR;
        const double JE1 = f_down(E1L, E1R, g1) * a2u * g1;
        const double JE2 = f_down(E2L, E2R, g2) * a1u * g2;
        F[2 + DIM] = JE1 + JE2;
    }
}

// FVM_Conservative2FluxPhaseField_K (ID_Bound 0): base reconstruction (no THINC on this path),
// then the Phase-Field flux.
template <int DIM, int SCHEME, int FORM, typename S>
__device__ __forceinline__ void pf_face_flux(const KParams& kp, int limiter, int dir, const S& s, double g,
                                             double* F, int& bad)
{
    constexpr int N = 4 + DIM;
    double LW[N], RW[N];
    auto no_normal = []() -> double { return 0.0; };
    recon_face<DIM, SCHEME, false>(kp, limiter, dir, s, no_normal, LW, RW);
    pf_flux<DIM, FORM>(kp, LW, RW, g, F, bad);
}

// ---- the N-phase Phase-Field flux ---------------------------------------------------------------
// Phasic internal energies alpha_k rho_k e_k of one face state at its pressure
// (Conservative2ConservativePhase_Compressible5Eq_NPhase.H:43-120): a phase below its p_min energy
// is set to it and the energy it gained is taken from the other phases in proportion to their
// volume fractions; repeated until none is below (each pass sets at least one more phase, so at
// most NP passes).
template <int DIM, int NP>
__device__ __forceinline__ void np_pf_phase_energies(const KParams& kp, const double* W, double* e, int& bad)
{
    using I = NpIdx<DIM, NP>;
    const double p = W[I::E];
    double em[NP];
    bool set[NP];
    double rho_e = 0.0;
    bool low = false;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const double phi = nasg_phi(W[I::A + k], W[k], kp.C[k], bad);
        e[k] = phi * kp.A[k] * p + phi * kp.B[k] + W[k] * kp.D[k];
        em[k] = phi * kp.A[k] * kp.pf_pmin[k] + phi * kp.B[k] + W[k] * kp.D[k];
        rho_e += e[k];
        set[k] = false;
        low = low || (e[k] < em[k]);
    }
    for (int pass = 0; low && pass < NP; ++pass) {
#pragma unroll
        for (int k = 0; k < NP; ++k)
            if (e[k] < em[k]) { e[k] = em[k]; set[k] = true; }
        double se = 0.0, sa = 0.0;
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            se += e[k];
            if (!set[k]) sa += W[I::A + k];
        }
        low = false;
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            if (!set[k]) e[k] += W[I::A + k] / sa * (rho_e - se);
        }
#pragma unroll
        for (int k = 0; k < NP; ++k) low = low || (e[k] < em[k]);
    }
}

// FluxPhaseField of the N-phase model (…FluxPhaseField_Compressible5Eq_NPhase_PhaseField.H:12-79)
// from primitive face states, with the efficient UD phase flux
// (FluxPhasePhaseField_UD_Efficient_Compressible_NPhase_PhaseField.H:12-232): for each phase i the
// phase-i quantities downwind of Grad_Q_i times (1 - alpha_i) upwind of it times Grad_Q_i; then,
// with Sum_J the sum of the volume-fraction fluxes, every flux less the quantity upwind of Sum_J
// times Sum_J. FORM 0: mixture momentum and energy (FluxPhaseField_Mixture …NPhase_PhaseField.H:12-104);
// FORM 1: phasic sums (VectorPhase2Vector).
template <int DIM, int FORM, int NP>
__device__ __forceinline__ void np_pf_flux(const KParams& kp, const double* WL, const double* WR, const double* g,
                                           double* F, int& bad)
{
    using I = NpIdx<DIM, NP>;
    double aju[NP], Jr[NP], Ja[NP];
    double SumJ = 0.0;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        aju[i] = f_up(1.0 - WL[I::A + i], 1.0 - WR[I::A + i], g[i]);
        Jr[i] = f_down(WL[i], WR[i], g[i]) * aju[i] * g[i];
        Ja[i] = f_down(WL[I::A + i], WR[I::A + i], g[i]) * aju[i] * g[i];
        SumJ += Ja[i];
    }
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        Jr[i] = Jr[i] - f_up(WL[i], WR[i], SumJ) * SumJ;
        Ja[i] = Ja[i] - f_up(WL[I::A + i], WR[I::A + i], SumJ) * SumJ;
        F[i] = Jr[i];
        F[I::A + i] = Ja[i];
    }
    if (FORM == 0) {
        double M = 0.0;
#pragma unroll
        for (int i = 0; i < NP; ++i) M += Jr[i];
        double ke = 0.0;
#pragma unroll
        for (int m = 0; m < DIM; ++m) {
            const double u = f_down(WL[I::MOM + m], WR[I::MOM + m], M);
            F[I::MOM + m] = M * u;
            ke = (m == 0) ? u * u : ke + u * u;
        }
        const double MeK = M * (0.5 * ke);
        double Je = 0.0;
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            const double pk = f_down(WL[I::E], WR[I::E], Ja[k]);
            const double Jphi = (kp.C[k] > 0.0) ? Ja[k] - kp.C[k] * Jr[k] : Ja[k];
            Je += kp.A[k] * Jphi * pk + kp.B[k] * Jphi + kp.D[k] * Jr[k];
        }
        F[I::E] = MeK + Je;
    } else {
        double keL = 0.0, keR = 0.0;
#pragma unroll
        for (int m = 0; m < DIM; ++m) {
            const double uL = WL[I::MOM + m], uR = WR[I::MOM + m];
            keL = (m == 0) ? uL * uL : keL + uL * uL;
            keR = (m == 0) ? uR * uR : keR + uR * uR;
            double t = 0.0;
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                double Ji = f_down(WL[i] * uL, WR[i] * uR, g[i]) * aju[i] * g[i];
                Ji = Ji - f_up(WL[i] * uL, WR[i] * uR, SumJ) * SumJ;
                t += Ji;
            }
            F[I::MOM + m] = t;
        }
        double eL[NP], eR[NP];
        np_pf_phase_energies<DIM, NP>(kp, WL, eL, bad);
        np_pf_phase_energies<DIM, NP>(kp, WR, eR, bad);
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const double EL = 0.5 * WL[i] * keL + eL[i], ER = 0.5 * WR[i] * keR + eR[i];
            double Ji = f_down(EL, ER, g[i]) * aju[i] * g[i];
            Ji = Ji - f_up(EL, ER, SumJ) * SumJ;
            t += Ji;
        }
        F[I::E] = t;
    }
}

// FVM_Conservative2FluxPhaseField_K of the N-phase model (ID_Bound 0,
// …FVM_Conservative2FluxPhaseField_K_Compressible5Eq_NPhase_PhaseField.H:15-79): the model's face
// states, then its Phase-Field flux; g holds Grad_Q of every phase on this face.
template <int DIM, int SCHEME, int FORM, int NP, typename S>
__device__ __forceinline__ void np_pf_face_flux(const KParams& kp, int limiter, int dir, const S& s,
                                                const double* g, double* F, int& bad)
{
    constexpr int N = NpIdx<DIM, NP>::N;
    double LW[N], RW[N];
    np_recon_face<DIM, SCHEME, NP>(kp, limiter, dir, s, LW, RW, bad);
    np_pf_flux<DIM, FORM, NP>(kp, LW, RW, g, F, bad);
}

// Normal component along dir of the volume-fraction gradient at a face
// (FDM_Normal_{X,Y,Z}FaceFromCenter_D, include/Operators_FD/FDM_Normal_XFaceFromCenter_D.H:12-56,
//  with FDM_Gradient*_*FaceFromCenter_D and Interpolation2Edge_Linear_D). a(o_d, o_t1, o_t2) returns
// alpha1 at the cell offset by o_d along dir (0: high cell, -1: low cell) and o_t along the
// transverse directions (ascending axis order).
template <int DIM, typename A>
__device__ __forceinline__ double face_normal_component(const KParams& kp, int dir, const A& a)
{
    const double fc = a(0, 0, 0), fm = a(-1, 0, 0);
    const double gd = (fc - fm) * kp.inv_dx[dir];
    double sum = gd * gd;
    int tdim[2] = {0, 0};
    int nt = 0;
#pragma unroll
    for (int t = 0; t < DIM; ++t)
        if (t != dir) tdim[nt++] = t;
#pragma unroll
    for (int q = 0; q < DIM - 1; ++q) {
        double edge[2];
#pragma unroll
        for (int sgn = 0; sgn < 2; ++sgn) {
            const int o = sgn ? 1 : -1;
            const double C = (q == 0) ? a(0, o, 0) : a(0, 0, o);
            const double D = (q == 0) ? a(-1, o, 0) : a(-1, 0, o);
            edge[sgn] = (dir < tdim[q]) ? 0.25 * fc + 0.25 * fm + 0.25 * C + 0.25 * D
                                        : 0.25 * fc + 0.25 * C + 0.25 * fm + 0.25 * D;
        }
        const double g = (edge[1] - edge[0]) * kp.inv_dx[tdim[q]];
        sum += g * g;
    }
    return (sum == 0.0) ? 0.0 : gd / sqrt(sum);
}

}  // namespace cmp
