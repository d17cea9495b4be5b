// cmp_device.hpp - pointwise five-equation physics and scalar reconstructions for gfx950 kernels.
//
// FP64 throughout. Formulas follow the reference's operators (file:line on each function);
// the kernels call these on primitives that are computed once per cell and staged in LDS, and
// the HLLC flux evaluates only the upwind star state. FMA contraction is left to the compiler.
#pragma once

#include <hip/hip_runtime.h>
#include "../../include/compas_amd.h"

namespace cmp {

// Kernel-side parameter block (passed by value as a kernel argument).
struct KParams {
    double A[2], B[2], C[2], D[2];
    int source_term;
    int nasg;          // any C_k > 0
    int thinc;
    int thinc_use_normal;
    double thinc_eps, thinc_beta, thinc_thr;
    double dx[3];
    double inv_dx[3];
};

__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }  // std::min
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }  // std::max
__device__ __forceinline__ double dsign(double x) { return (x > 0.0) ? 1.0 : ((x < 0.0) ? -1.0 : 0.0); }

// NASG_Phi (include/EquationOfState_NASG.H:34-41); the covolume violation sets *bad.
__device__ __forceinline__ double nasg_phi(double alpha, double alpha_rho, double b, int& bad)
{
    if (b > 0.0) {
        if (alpha > 0.0 && !(1.0 - b * (alpha_rho / alpha) > 0.0)) bad = CMP_ERR_COVOLUME;
        return alpha - b * alpha_rho;
    }
    return alpha;
}

// Mixture A and B sums of EquationOfState_Mixture
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

// Conservative2Primitive (include/Physics_Compressible5Eq/Conservative2Primitive_Compressible5Eq.H:15-61).
template <int DIM>
__device__ __forceinline__ void cons2prim(const KParams& kp, const double* U, double* W, int& bad)
{
    const double a1r1 = U[0], a2r2 = U[1];
    const double rho = a1r1 + a2r2;
    double ke = 0.0;
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
        const double v = U[2 + m] / rho;
        W[2 + m] = v;
        ke = (m == 0) ? v * v : ke + v * v;
    }
    const double rho_e = U[2 + DIM] - (0.5 * rho) * ke;
    double AP, BP, rho_D;
    mix_AB(kp, U[3 + DIM], a1r1, a2r2, AP, BP, rho_D, bad);
    W[0] = a1r1;
    W[1] = a2r2;
    W[2 + DIM] = (rho_e - rho_D - BP) / AP;
    W[3 + DIM] = U[3 + DIM];
}

// Primitive2Conservative (include/Physics_Compressible5Eq/Primitive2Conservative_Compressible5Eq.H:15-62).
template <int DIM>
__device__ __forceinline__ void prim2cons(const KParams& kp, const double* W, double* U, int& bad)
{
    const double a1r1 = W[0], a2r2 = W[1];
    const double rho = a1r1 + a2r2;
    double ke = 0.0;
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
        U[2 + m] = rho * W[2 + m];
        ke = (m == 0) ? W[2 + m] * W[2 + m] : ke + W[2 + m] * W[2 + m];
    }
    double AP, BP, rho_D;
    mix_AB(kp, W[3 + DIM], a1r1, a2r2, AP, BP, rho_D, bad);
    const double rho_e = AP * W[2 + DIM] + BP + rho_D;
    U[0] = a1r1;
    U[1] = a2r2;
    U[2 + DIM] = rho_e + (0.5 * rho) * ke;
    U[3 + DIM] = W[3 + DIM];
}

// SoundSpeed with rho_c2_Phase / rho_c2_Mixture
// (include/Physics_Compressible5Eq/SoundSpeed_Compressible5Eq.H:15-54, rho_c2_Phase_…H:17-23,
//  rho_c2_Mixture_…H:15-49). K = 0: A-weighted frozen mixture; K = 1: Wood.
__device__ __forceinline__ double sound_speed(const KParams& kp, double p, double a1, double a1r1,
                                              double a2r2, int& bad)
{
    const double r1c1 = ((kp.A[0] + 1.0) * p + kp.B[0]) / kp.A[0];
    const double r2c2 = ((kp.A[1] + 1.0) * p + kp.B[1]) / kp.A[1];
    const double a2 = 1.0 - a1;
    double phi1 = a1, phi2 = a2;
    if (kp.nasg) {
        phi1 = nasg_phi(a1, a1r1, kp.C[0], bad);
        phi2 = nasg_phi(a2, a2r2, kp.C[1], bad);
    }
    double rho_c2;
    if (kp.source_term == 0) {
        const double num = a1 * kp.A[0] * r1c1 + a2 * kp.A[1] * r2c2;
        const double AP = phi1 * kp.A[0] + phi2 * kp.A[1];
        rho_c2 = num / AP;
    } else {
        rho_c2 = 1.0 / (phi1 / r1c1 + phi2 / r2c2);
    }
    return sqrt(rho_c2 / (a1r1 + a2r2));
}

// Kapila coefficient phi1 (1 + K1) of the alpha1 equation's div(u) term
// (FVM_SurfaceIntegral_NC, include/Operators_Compressible5Eq/FVM_SurfaceIntegralNC_Compressible5Eq.H:29-44,
//  NonConservative include/Physics_Compressible5Eq/NonConservative_Compressible5Eq.H:15-58).
template <int DIM>
__device__ __forceinline__ double nc_coefficient(const KParams& kp, const double* U, int& bad)
{
    const double a1 = U[3 + DIM];
    if (kp.source_term == 0) return a1 * (1.0 + 0.0);
    double W[4 + DIM];
    cons2prim<DIM>(kp, U, W, bad);
    const double p = W[2 + DIM];
    const double r1c1 = ((kp.A[0] + 1.0) * p + kp.B[0]) / kp.A[0];
    const double r2c2 = ((kp.A[1] + 1.0) * p + kp.B[1]) / kp.A[1];
    double phi1 = a1, phi2 = 1.0 - a1;
    if (kp.nasg) {
        phi1 = nasg_phi(a1, U[0], kp.C[0], bad);
        phi2 = nasg_phi(1.0 - a1, U[1], kp.C[1], bad);
    }
    const double rho_c2 = 1.0 / (phi1 / r1c1 + phi2 / r2c2);
    const double K1 = rho_c2 / r1c1 - 1.0;
    return phi1 * (1.0 + K1);
}

// Ratio-form WENO5 (include/Operators_FV/WENO5.H:12-75, Reconstruction2Face_WENO5_D.H:12-37):
// p = 2, no epsilon, NaN ratios replaced by 1. FACE = +1: value at the high face of cell c.
template <int FACE>
__device__ __forceinline__ double weno5(double f_mm, double f_m, double f_c, double f_p, double f_pp)
{
    double f0, f1, f2, d0, d2;
    if (FACE == 1) {
        f0 = f_c + (2.0 / 3.0) * (f_p - f_c) - (1.0 / 6.0) * (f_pp - f_p);
        f1 = f_c + (1.0 / 6.0) * (f_c - f_m) + (1.0 / 3.0) * (f_p - f_c);
        f2 = f_c - (1.0 / 3.0) * (f_m - f_mm) + (5.0 / 6.0) * (f_c - f_m);
        d0 = 3.0 / 10.0;
        d2 = 1.0 / 10.0;
    } else {
        f0 = f_c - (5.0 / 6.0) * (f_p - f_c) + (1.0 / 3.0) * (f_pp - f_p);
        f1 = f_c - (1.0 / 3.0) * (f_c - f_m) - (1.0 / 6.0) * (f_p - f_c);
        f2 = f_c + (1.0 / 6.0) * (f_m - f_mm) - (2.0 / 3.0) * (f_c - f_m);
        d0 = 1.0 / 10.0;
        d2 = 3.0 / 10.0;
    }
    const double d1 = 1.0 - d0 - d2;
    const double a = f_p - f_c, b = f_pp - f_p, c = f_c - f_m, e = f_m - f_mm;
    const double IS0 = (10.0 / 3.0) * a * a - (11.0 / 3.0) * a * b + (4.0 / 3.0) * b * b;
    const double IS1 = -(5.0 / 3.0) * c * a + (4.0 / 3.0) * c * c + (4.0 / 3.0) * a * a;
    const double IS2 = (10.0 / 3.0) * c * c - (11.0 / 3.0) * e * c + (4.0 / 3.0) * e * e;
    double IS01 = IS0 / IS1;
    if (__builtin_isnan(IS01)) IS01 = 1.0;
    double IS02 = IS0 / IS2;
    if (__builtin_isnan(IS02)) IS02 = 1.0;
    double IS12 = IS1 / IS2;
    if (__builtin_isnan(IS12)) IS12 = 1.0;
    const double IS10 = 1.0 / IS01, IS20 = 1.0 / IS02, IS21 = 1.0 / IS12;
    const double w0 = d0 / (d0 + (IS01 * IS01) * d1 + (IS02 * IS02) * d2);
    const double w1 = d1 / ((IS10 * IS10) * d0 + d1 + (IS12 * IS12) * d2);
    const double w2 = d2 / ((IS20 * IS20) * d0 + (IS21 * IS21) * d1 + d2);
    return w0 * f0 + w1 * f1 + w2 * f2;
}

// Slope_TVD + Reconstruction2Face_MUSCL_D + Reconstruction2Face_Linear
// (include/Operators_FV/Slope_TVD.H:12-43, Reconstruction2Face_MUSCL_D.H:12-36, Limiter_*.H).
__device__ __forceinline__ double muscl(double f_m, double f_c, double f_p, double dx, int face,
                                        int limiter)
{
    const double sm = (f_c - f_m) / dx;
    const double sc = (f_p - f_m) / (2.0 * dx);
    const double sp = (f_p - f_c) / dx;
    double s = 0.0;
    if (limiter == 1) {
        s = smin(smax(sm, 0.0), smax(sp, 0.0)) + smax(smin(sm, 0.0), smin(sp, 0.0));
    } else if (limiter == 2) {
        s = (sm * sp > 0.0) ? 2.0 / (1.0 / sm + 1.0 / sp) : 0.0;
    } else if (limiter == 3) {
        const double am = fabs(sm), ap = fabs(sp);
        s = (dsign(sm) + dsign(sp)) * smin(smin(am, ap), 0.5 * smax(am, ap));
    } else if (limiter == 4) {
        const double hi = smin(smin(smax(sc, 0.0), smax(2.0 * sm, 0.0)), smax(2.0 * sp, 0.0));
        const double lo = smax(smax(smin(sc, 0.0), smin(2.0 * sm, 0.0)), smin(2.0 * sp, 0.0));
        s = hi + lo;
    }
    return f_c + double(face) * 0.5 * s * dx;
}

// One face's Riemann problem from conservative L/R states. HLLC follows
// Riemann_HLLC (include/Physics_Compressible5Eq/Riemann_HLLC_Compressible5Eq.H:16-194) but forms
// only the star state of the side sign(S*) selects (both halves when S* == 0); HLL follows
// Riemann_HLL (Riemann_HLL_Compressible5Eq.H:16-133). The states are taken through cons->prim
// first, as the reference does. Returns u*_n; F receives N fluxes; cmax the face wave speed.
template <int DIM, int RIEMANN>
__device__ __forceinline__ double riemann(const KParams& kp, const double* UL, const double* UR,
                                          int dir, double* F, double& cmax, int& bad)
{
    constexpr int N = 4 + DIM;
    double WL[N], WR[N];
    cons2prim<DIM>(kp, UL, WL, bad);
    cons2prim<DIM>(kp, UR, WR, bad);
    const double pL = WL[2 + DIM], pR = WR[2 + DIM];
    const double cL = sound_speed(kp, pL, WL[3 + DIM], WL[0], WL[1], bad);
    const double cR = sound_speed(kp, pR, WR[3 + DIM], WR[0], WR[1], bad);
    const double VL = WL[2 + dir], VR = WR[2 + dir];
    double qL = 0.0, qR = 0.0;
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
        qL = (m == 0) ? WL[2] * WL[2] : qL + WL[2 + m] * WL[2 + m];
        qR = (m == 0) ? WR[2] * WR[2] : qR + WR[2 + m] * WR[2 + m];
    }
    cmax = smax(cL + sqrt(qL), cR + sqrt(qR));

    if (RIEMANN == 2) {
        const double V_avg = 0.5 * (VL + VR);
        const double c_avg = 0.5 * (cL + cR);
        const double SL = smin(V_avg - c_avg, VL - cL);
        const double SR = smax(V_avg + c_avg, VR + cR);
        const double rhoL = WL[0] + WL[1];
        const double rhoR = WR[0] + WR[1];
        const double S_Star = (pR - pL + rhoL * VL * (SL - VL) - rhoR * VR * (SR - VR))
                            / (rhoL * (SL - VL) - rhoR * (SR - VR));
        // Physical flux of the upwind side (Conservative2Flux, Conservative2Flux_Compressible5Eq.H:15-80)
        // plus S_K (U*_K - U_K); the star state scales U_K by q = (S_K - V_K)/(S_K - S*).
        auto side = [&](const double* U, const double* W, double S, double V, double rho, double p,
                        double Sb, double* out) -> double {
            const double q = (S - V) / (S - S_Star);
            out[0] = W[0] * V + Sb * (q * U[0] - U[0]);
            out[1] = W[1] * V + Sb * (q * U[1] - U[1]);
#pragma unroll
            for (int m = 0; m < DIM; ++m) {
                double Fm = U[2 + m] * V;
                double Cs = q * U[2 + m];
                if (m == dir) {
                    Fm += p;
                    Cs += q * rho * (S_Star - V);
                }
                out[2 + m] = Fm + Sb * (Cs - U[2 + m]);
            }
            const double CsE = q * U[2 + DIM] + q * (S_Star - V) * (rho * S_Star + p / (S - V));
            out[2 + DIM] = (U[2 + DIM] + p) * V + Sb * (CsE - U[2 + DIM]);
            out[3 + DIM] = W[3 + DIM] * V + Sb * (q * U[3 + DIM] - U[3 + DIM]);
            return Sb * (q - 1.0);
        };
        double un;
        if (S_Star > 0.0) {
            const double Sm = smin(0.0, SL);
            un = VL + side(UL, WL, SL, VL, rhoL, pL, Sm, F);
        } else if (S_Star < 0.0) {
            const double Sp = smax(0.0, SR);
            un = VR + side(UR, WR, SR, VR, rhoR, pR, Sp, F);
        } else {
            double FR_[N];
            const double Sm = smin(0.0, SL), Sp = smax(0.0, SR);
            const double dL = side(UL, WL, SL, VL, rhoL, pL, Sm, F);
            const double dR = side(UR, WR, SR, VR, rhoR, pR, Sp, FR_);
#pragma unroll
            for (int n = 0; n < N; ++n) F[n] = 0.5 * F[n] + 0.5 * FR_[n];
            un = (0.5 * VL + 0.5 * VR) + (0.5 * dL + 0.5 * dR);
        }
        return un;
    } else {
        const double Sp = smax(smax(0.0, VL + cL), VR + cR);
        const double Sm = smin(smin(0.0, VL - cL), VR - cR);
        const double inv = 1.0 / (Sp - Sm);
        double FL[N], FR[N];
        FL[0] = WL[0] * VL; FR[0] = WR[0] * VR;
        FL[1] = WL[1] * VL; FR[1] = WR[1] * VR;
#pragma unroll
        for (int m = 0; m < DIM; ++m) {
            FL[2 + m] = UL[2 + m] * VL;
            FR[2 + m] = UR[2 + m] * VR;
        }
        FL[2 + dir] += pL;
        FR[2 + dir] += pR;
        FL[2 + DIM] = (UL[2 + DIM] + pL) * VL;
        FR[2 + DIM] = (UR[2 + DIM] + pR) * VR;
        FL[3 + DIM] = WL[3 + DIM] * VL;
        FR[3 + DIM] = WR[3 + DIM] * VR;
#pragma unroll
        for (int n = 0; n < N; ++n)
            F[n] = (Sp * FL[n] - Sm * FR[n] + Sp * Sm * (UR[n] - UL[n])) * inv;
        return (Sp * VL - Sm * VR) * inv;
    }
}

}  // namespace cmp
