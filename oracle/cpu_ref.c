/*
 * cpu_ref.c - CPU restatement (the oracle) of COMPAS's five-equation hyperbolic RHS, FP64.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg, never by the product library. It follows the reference's structure on purpose (per face:
 * cons->prim of every stencil cell, reconstruction, prim->cons, Riemann solver with its own
 * cons->prim; stored face fluxes; a separate divergence pass; separate RK LinComb/Saxpy passes),
 * so that it is both the checker and a faithful CPU baseline of the reference's algorithm.
 *
 * Operation order follows the cited reference lines; build with -ffp-contract=off (no FMA),
 * like x86-64 g++ builds the reference. std::min/std::max semantics (not fmin/fmax) are kept.
 *
 * The checks that pin each function independently (golden vectors, closed forms, invariants,
 * brute force) are the tests under tests/ named for it; every function here has at least one.
 * One part is parity unpinned against the reference: the composite implicit Grad_Q solve of a
 * hierarchy uses its own two-point coarse-fine flux where the reference's MLMG (AMReX, not
 * vendored) uses a higher-order coarse-fine stencil; it is pinned against its own equation,
 * symmetry and the single-level solve, not against the reference's numbers at the interfaces.
 */
#include "cpu_ref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* std::min(a,b) = (b < a) ? b : a ; std::max(a,b) = (a < b) ? b : a  */
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }

/* Sign (include/Tools/Sign.H:13-25) */
static inline double ref_sign(double x) { return (x > 0.0) ? 1.0 : ((x < 0.0) ? -1.0 : 0.0); }

#define ERR_SET(err, code) do { if ((err) && *(err) == 0) *(err) = (code); } while (0)

/* ------------------------------------------------------------------------------------------------
 * EOS and thermodynamics
 * ---------------------------------------------------------------------------------------------- */

/* NASG_Phi (include/EquationOfState_NASG.H:34-41): phi = alpha - b alpha rho for b > 0, else alpha;
 * the covolume check aborts in the reference, here it sets CMP_ERR_COVOLUME. */
double ref_nasg_phi(double alpha, double alpha_rho, double b, int* err)
{
    if (b > 0.0 && alpha > 0.0) {
        double xi = 1.0 - b * (alpha_rho / alpha);
        if (!(xi > 0.0)) ERR_SET(err, CMP_ERR_COVOLUME);
    }
    return (b > 0.0) ? alpha - b * alpha_rho : alpha;
}

/* EquationOfState_Mixture, pressure-based branch
 * (include/Physics_Compressible5Eq/EquationOfState_Mixture_Compressible5Eq.H:15-54). */
double ref_eos_mixture(const cmp_params* P, double input, double a1, double a1r1, double a2r2,
                       int output_internal_energy, int* err)
{
    double a2 = 1.0 - a1;
    double phi1 = ref_nasg_phi(a1, a1r1, P->eos_C[0], err);
    double phi2 = ref_nasg_phi(a2, a2r2, P->eos_C[1], err);
    double rho_D = a1r1 * P->eos_D[0] + a2r2 * P->eos_D[1];
    double AP = phi1 * P->eos_A[0] + phi2 * P->eos_A[1];
    double BP = phi1 * P->eos_B[0] + phi2 * P->eos_B[1];
    if (output_internal_energy)
        return AP * input + BP + rho_D;
    return (input - rho_D - BP) / AP;
}

/* Conservative2Primitive (include/Physics_Compressible5Eq/Conservative2Primitive_Compressible5Eq.H:15-61) */
void ref_cons2prim(const cmp_params* P, const double* U, double* W, int* err)
{
    const int d = P->dim;
    double a1r1 = U[0], a2r2 = U[1];
    double rE = U[2 + d], a1 = U[3 + d];
    double rho = a1r1 + a2r2;
    double u = U[2] / rho;
    double v = U[3] / rho;
    double ke;
    double w = 0.0;
    if (d == 3) {
        w = U[4] / rho;
        ke = u * u + v * v + w * w;
    } else {
        ke = u * u + v * v;
    }
    double rho_e = rE - (0.5) * rho * ke;
    double p = ref_eos_mixture(P, rho_e, a1, a1r1, a2r2, 0, err);
    W[0] = a1r1; W[1] = a2r2; W[2] = u; W[3] = v;
    if (d == 3) W[4] = w;
    W[2 + d] = p; W[3 + d] = a1;
}

/* Primitive2Conservative (include/Physics_Compressible5Eq/Primitive2Conservative_Compressible5Eq.H:15-62) */
void ref_prim2cons(const cmp_params* P, const double* W, double* U, int* err)
{
    const int d = P->dim;
    double a1r1 = W[0], a2r2 = W[1];
    double u = W[2], v = W[3], w = (d == 3) ? W[4] : 0.0;
    double p = W[2 + d], a1 = W[3 + d];
    double rho = a1r1 + a2r2;
    double ru = rho * u, rv = rho * v, rw = rho * w;
    double rho_e = ref_eos_mixture(P, p, a1, a1r1, a2r2, 1, err);
    double ke = (d == 3) ? (u * u + v * v + w * w) : (u * u + v * v);
    double rE = rho_e + (0.5) * rho * ke;
    U[0] = a1r1; U[1] = a2r2; U[2] = ru; U[3] = rv;
    if (d == 3) U[4] = rw;
    U[2 + d] = rE; U[3 + d] = a1;
}

/* rho_c2_Phase (include/Physics_Compressible5Eq/rho_c2_Phase_Compressible5Eq.H:17-23) */
double ref_rho_c2_phase(double p, double A, double B) { return ((A + 1.0) * p + B) / A; }

/* rho_c2_Mixture (include/Physics_Compressible5Eq/rho_c2_Mixture_Compressible5Eq.H:15-49) */
static double ref_rho_c2_mixture(const cmp_params* P, double r1c1, double r2c2, double a1,
                                 double a1r1, double a2r2, int id_k, int* err)
{
    double a2 = 1.0 - a1;
    double phi1 = ref_nasg_phi(a1, a1r1, P->eos_C[0], err);
    double phi2 = ref_nasg_phi(a2, a2r2, P->eos_C[1], err);
    if (id_k == 0) {
        double AP_rho_c2 = a1 * P->eos_A[0] * r1c1 + a2 * P->eos_A[1] * r2c2;
        double AP = phi1 * P->eos_A[0] + phi2 * P->eos_A[1];
        return AP_rho_c2 / AP;
    }
    return 1.0 / (phi1 / r1c1 + phi2 / r2c2);
}

/* SoundSpeed (include/Physics_Compressible5Eq/SoundSpeed_Compressible5Eq.H:15-54) */
double ref_sound_speed(const cmp_params* P, double p, double a1, double a1r1, double a2r2, int id_k,
                       int* err)
{
    double r1c1 = ref_rho_c2_phase(p, P->eos_A[0], P->eos_B[0]);
    double r2c2 = ref_rho_c2_phase(p, P->eos_A[1], P->eos_B[1]);
    double rho_c2 = ref_rho_c2_mixture(P, r1c1, r2c2, a1, a1r1, a2r2, id_k, err);
    double rho = a1r1 + a2r2;
    return sqrt(rho_c2 / rho);
}

/* NonConservative (include/Physics_Compressible5Eq/NonConservative_Compressible5Eq.H:15-58) */
void ref_nonconservative_K(const cmp_params* P, double p, double a1, double a1r1, double a2r2,
                           int id_k, double* K, int* err)
{
    if (id_k == 0) { K[0] = 0.0; K[1] = 0.0; return; }
    double r1c1 = ref_rho_c2_phase(p, P->eos_A[0], P->eos_B[0]);
    double r2c2 = ref_rho_c2_phase(p, P->eos_A[1], P->eos_B[1]);
    double rho_c2 = ref_rho_c2_mixture(P, r1c1, r2c2, a1, a1r1, a2r2, id_k, err);
    K[0] = (rho_c2 / r1c1 - 1.0);
    K[1] = (rho_c2 / r2c2 - 1.0);
}

/* Conservative2Flux (include/Physics_Compressible5Eq/Conservative2Flux_Compressible5Eq.H:15-80) */
void ref_cons2flux(const cmp_params* P, const double* U, int dir, double* F, int* err)
{
    const int d = P->dim;
    double W[REF_NMAX];
    ref_cons2prim(P, U, W, err);
    double p = W[2 + d];
    double V = W[2 + dir];
    F[0] = W[0] * V;
    F[1] = W[1] * V;
    for (int m = 0; m < d; ++m) F[2 + m] = U[2 + m] * V;
    F[2 + d] = (U[2 + d] + p) * V;
    F[3 + d] = W[3 + d] * V;
    F[2 + dir] += p;
}

/* ------------------------------------------------------------------------------------------------
 * Scalar reconstruction
 * ---------------------------------------------------------------------------------------------- */

/* WENO5 candidate values and ratio-form nonlinear weights (include/Operators_FV/WENO5.H:12-75):
 * p = 2, no epsilon, a NaN ratio (0/0) replaced by 1. id_face = +1 gives the value at the high
 * face of cell c, -1 at its low face. */
void ref_weno5_weights(double f_mm, double f_m, double f_c, double f_p, double f_pp, int id_face,
                       double* f, double* W)
{
    double d0 = 0.0, d1, d2 = 0.0;
    if (id_face == 1) {
        f[0] = f_c + (2.0 / 3.0) * (f_p - f_c) - (1.0 / 6.0) * (f_pp - f_p);
        f[1] = f_c + (1.0 / 6.0) * (f_c - f_m) + (1.0 / 3.0) * (f_p - f_c);
        f[2] = f_c - (1.0 / 3.0) * (f_m - f_mm) + (5.0 / 6.0) * (f_c - f_m);
        d0 = 3.0 / 10.0;
        d2 = 1.0 / 10.0;
    }
    if (id_face == -1) {
        f[0] = f_c - (5.0 / 6.0) * (f_p - f_c) + (1.0 / 3.0) * (f_pp - f_p);
        f[1] = f_c - (1.0 / 3.0) * (f_c - f_m) - (1.0 / 6.0) * (f_p - f_c);
        f[2] = f_c + (1.0 / 6.0) * (f_m - f_mm) - (2.0 / 3.0) * (f_c - f_m);
        d0 = 1.0 / 10.0;
        d2 = 3.0 / 10.0;
    }
    d1 = 1.0 - d0 - d2;

    double IS0 = (10.0 / 3.0) * (f_p - f_c) * (f_p - f_c) - (11.0 / 3.0) * (f_p - f_c) * (f_pp - f_p)
               + (4.0 / 3.0) * (f_pp - f_p) * (f_pp - f_p);
    double IS1 = -(5.0 / 3.0) * (f_c - f_m) * (f_p - f_c) + (4.0 / 3.0) * (f_c - f_m) * (f_c - f_m)
               + (4.0 / 3.0) * (f_p - f_c) * (f_p - f_c);
    double IS2 = (10.0 / 3.0) * (f_c - f_m) * (f_c - f_m) - (11.0 / 3.0) * (f_m - f_mm) * (f_c - f_m)
               + (4.0 / 3.0) * (f_m - f_mm) * (f_m - f_mm);

    double IS01 = IS0 / IS1;
    if (isnan(IS01)) IS01 = 1.0;
    double IS10 = 1.0 / IS01;
    double IS02 = IS0 / IS2;
    if (isnan(IS02)) IS02 = 1.0;
    double IS20 = 1.0 / IS02;
    double IS12 = IS1 / IS2;
    if (isnan(IS12)) IS12 = 1.0;
    double IS21 = 1.0 / IS12;

    W[0] = d0 / (d0 + (IS01 * IS01) * d1 + (IS02 * IS02) * d2);
    W[1] = d1 / ((IS10 * IS10) * d0 + d1 + (IS12 * IS12) * d2);
    W[2] = d2 / ((IS20 * IS20) * d0 + (IS21 * IS21) * d1 + d2);
}

/* Reconstruction2Face_WENO5_D + Reconstruction2Face_WENO
 * (include/Operators_FV/Reconstruction2Face_WENO5_D.H:12-37, Reconstruction2Face_WENO.H:13-22) */
double ref_weno5(double f_mm, double f_m, double f_c, double f_p, double f_pp, int id_face)
{
    double f[3], W[3];
    ref_weno5_weights(f_mm, f_m, f_c, f_p, f_pp, id_face, f, W);
    double face = 0.0;
    for (int s = 0; s < 3; ++s) face += W[s] * f[s];
    return face;
}

/* Limiters (include/Operators_FV/Limiter_{Minmod,vanLeer,Superbee,MC}.H) and Slope_TVD
 * (include/Operators_FV/Slope_TVD.H:12-43) */
static double ref_slope_tvd(double sm, double sc, double sp, int id_limiter)
{
    if (id_limiter == 1)
        return smin(smax(sm, 0.0), smax(sp, 0.0)) + smax(smin(sm, 0.0), smin(sp, 0.0));
    if (id_limiter == 2) {
        if (sm * sp > 0.0) return 2.0 / (1.0 / sm + 1.0 / sp);
        return 0.0;
    }
    if (id_limiter == 3) {
        double am = fabs(sm), ap = fabs(sp);
        double m = smin(smin(am, ap), (0.5) * smax(am, ap));
        return (ref_sign(sm) + ref_sign(sp)) * m;
    }
    if (id_limiter == 4) {
        /* std::min({c+, 2a+, 2b+}) + std::max({c-, 2a-, 2b-}); the initializer-list forms keep
         * the first of equal elements, which is value-identical */
        double hi = smin(smin(smax(sc, 0.0), smax(2.0 * sm, 0.0)), smax(2.0 * sp, 0.0));
        double lo = smax(smax(smin(sc, 0.0), smin(2.0 * sm, 0.0)), smin(2.0 * sp, 0.0));
        return hi + lo;
    }
    return 0.0;
}

/* Reconstruction2Face_MUSCL_D + Reconstruction2Face_Linear
 * (include/Operators_FV/Reconstruction2Face_MUSCL_D.H:12-36, Reconstruction2Face_Linear.H) */
double ref_muscl(double f_m, double f_c, double f_p, double dx, int id_face, int id_limiter)
{
    double sm = (f_c - f_m) / dx;
    double sc = (f_p - f_m) / (2.0 * dx);
    double sp = (f_p - f_c) / dx;
    double s = ref_slope_tvd(sm, sc, sp, id_limiter);
    return f_c + (double)id_face * 0.5 * s * dx;
}

/* fC_QuadratureRule + Limiter_BoundMin (include/Operators_FV/fC_QuadratureRule.H:12-19,
 * Limiter_BoundMin.H:12-22), called through Limiter_Bound with ID_BoundType = -1 (Limiter_Bound.H). */
static double ref_theta_min(double f_min, double f_c, double fL_fp, double fR_fm, double Weight)
{
    double fC = (f_c - Weight * fL_fp - Weight * fR_fm) / (1.0 - 2.0 * Weight);
    double min_f = smin(smin(fC, fL_fp), fR_fm);
    double th = smax(0.0, f_c - f_min) / smax(0.0, f_c - min_f);
    return smin(1.0, th);
}

/* FVM_Reconstruction2Face_Conservative_WENO5B_D, primitive part
 * (include/Operators_Compressible5Eq/FVM_Reconstruction2Face_Conservative_WENO5B_D_Compressible5Eq.H:15-320):
 * W[5][N] are the primitives of cells mm..pp; both faces of cell c are reconstructed for the
 * phase masses and volume fractions (alpha with the weights of alpha*rho), limited for positivity,
 * normalised by the sum of the volume fractions, limited again; velocity and pressure by WENO5. */
void ref_weno5b_side(const cmp_params* P, double W[][REF_NMAX], int id_face, double* Wf)
{
    const int d = P->dim;
    double ar[2][5], al[2][5];
    for (int s = 0; s < 5; ++s) {
        ar[0][s] = W[s][0];
        ar[1][s] = W[s][1];
        al[0][s] = W[s][3 + d];
        al[1][s] = 1.0 - W[s][3 + d];
    }
    double Weight = (P->id_poly == 2) ? 1.0 / 6.0 : 1.0 / 12.0;
    double arL[2], alL[2], arR[2], alR[2];
    double sumL = 0.0, sumR = 0.0;
    double cand[3], wr[3], cand2[3], wdummy[3];
    for (int k = 0; k < 2; ++k) {
        ref_weno5_weights(ar[k][0], ar[k][1], ar[k][2], ar[k][3], ar[k][4], 1, cand, wr);
        arL[k] = 0.0; for (int s = 0; s < 3; ++s) arL[k] += wr[s] * cand[s];
        ref_weno5_weights(al[k][0], al[k][1], al[k][2], al[k][3], al[k][4], 1, cand2, wdummy);
        alL[k] = 0.0; for (int s = 0; s < 3; ++s) alL[k] += wr[s] * cand2[s];
        ref_weno5_weights(ar[k][0], ar[k][1], ar[k][2], ar[k][3], ar[k][4], -1, cand, wr);
        arR[k] = 0.0; for (int s = 0; s < 3; ++s) arR[k] += wr[s] * cand[s];
        ref_weno5_weights(al[k][0], al[k][1], al[k][2], al[k][3], al[k][4], -1, cand2, wdummy);
        alR[k] = 0.0; for (int s = 0; s < 3; ++s) alR[k] += wr[s] * cand2[s];
        double th_ar = ref_theta_min(0.0, ar[k][2], arL[k], arR[k], Weight);
        double th_al = ref_theta_min(0.0, al[k][2], alL[k], alR[k], Weight);
        double th = smin(th_ar, th_al);
        arL[k] = ar[k][2] + th * (arL[k] - ar[k][2]);
        alL[k] = al[k][2] + th * (alL[k] - al[k][2]);
        arR[k] = ar[k][2] + th * (arR[k] - ar[k][2]);
        alR[k] = al[k][2] + th * (alR[k] - al[k][2]);
        sumL += alL[k];
        sumR += alR[k];
    }
    double theta = 1.0;
    for (int k = 0; k < 2; ++k) {
        arL[k] /= sumL; alL[k] /= sumL;
        arR[k] /= sumR; alR[k] /= sumR;
        double th_ar = ref_theta_min(0.0, ar[k][2], arL[k], arR[k], Weight);
        double th_al = ref_theta_min(0.0, al[k][2], alL[k], alR[k], Weight);
        theta = smin(smin(theta, th_ar), th_al);
    }
    double a1r1F = (id_face == 1) ? arL[0] : arR[0];
    double a2r2F = (id_face == 1) ? arL[1] : arR[1];
    double a1F = (id_face == 1) ? alL[0] : alR[0];
    Wf[0] = ar[0][2] + theta * (a1r1F - ar[0][2]);
    Wf[1] = ar[1][2] + theta * (a2r2F - ar[1][2]);
    Wf[3 + d] = al[0][2] + theta * (a1F - al[0][2]);
    for (int n = 2; n <= 2 + d; ++n)
        Wf[n] = ref_weno5(W[0][n], W[1][n], W[2][n], W[3][n], W[4][n], id_face);
}

/* Reconstruction2Face_THINC_D (include/Operators_FV/Reconstruction2Face_THINC_D.H:13-35):
 * id_face 1 gives the high face of the cell, anything else the low face. */
double ref_thinc(double beta, double sigma, double alpha_bar, int id_face)
{
    double a = (id_face == 1) ? 0.0 : 1.0;
    double b = (id_face == 1) ? 1.0 : -1.0;
    double arg_ln = (exp(2.0 * beta * alpha_bar) - 1.0) / (exp(2.0 * beta) - exp(2.0 * beta * alpha_bar));
    double c = 0.5 / beta * log(arg_ln);
    double arg_tanh = beta * (a + 0.5 * b * (sigma + 1.0) + c);
    double af = 0.5 * (1.0 + tanh(arg_tanh));
    /* amrex::min(1.0, amrex::max(0.0, af)) */
    af = (0.0 < af) ? af : 0.0;
    return (af < 1.0) ? af : 1.0;
}

/* ------------------------------------------------------------------------------------------------
 * Riemann solvers
 * ---------------------------------------------------------------------------------------------- */

/* Riemann_HLLC (include/Physics_Compressible5Eq/Riemann_HLLC_Compressible5Eq.H:16-194) and
 * Riemann_HLL (include/Physics_Compressible5Eq/Riemann_HLL_Compressible5Eq.H:16-133), dispatched as
 * RiemannSolver (RiemannSolver_Compressible5Eq.H:16-47). */
void ref_riemann(const cmp_params* P, const double* UL, const double* UR, int dir,
                 double* flux, double* vel, double* cmax, int* err)
{
    const int d = P->dim, N = P->nstate, id_k = P->source_term;
    double WL[REF_NMAX], WR[REF_NMAX], FL[REF_NMAX], FR[REF_NMAX];
    ref_cons2prim(P, UL, WL, err);
    ref_cons2prim(P, UR, WR, err);
    ref_cons2flux(P, UL, dir, FL, err);
    ref_cons2flux(P, UR, dir, FR, err);
    double pL = WL[2 + d], pR = WR[2 + d];
    double cL = ref_sound_speed(P, pL, WL[3 + d], WL[0], WL[1], id_k, err);
    double cR = ref_sound_speed(P, pR, WR[3 + d], WR[0], WR[1], id_k, err);
    double VL = WL[2 + dir], VR = WR[2 + dir];

    if (P->id_riemann == 2) {
        double V_avg = 0.5 * (VL + VR);
        double c_avg = 0.5 * (cL + cR);
        double SL = smin(V_avg - c_avg, VL - cL);
        double SR = smax(V_avg + c_avg, VR + cR);
        double Sm = smin(0.0, SL);
        double Sp = smax(0.0, SR);
        double rhoL = WL[0] + WL[1];
        double rhoR = WR[0] + WR[1];
        double S_Star = (pR - pL + rhoL * VL * (SL - VL) - rhoR * VR * (SR - VR))
                      / (rhoL * (SL - VL) - rhoR * (SR - VR));
        double s = ref_sign(S_Star);
        for (int m = 0; m < d; ++m)
            vel[m] = 0.5 * (1.0 + s) * WL[2 + m] + 0.5 * (1.0 - s) * WR[2 + m];
        vel[dir] += 0.5 * (1.0 + s) * Sm * ((SL - VL) / (SL - S_Star) - 1.0)
                  + 0.5 * (1.0 - s) * Sp * ((SR - VR) / (SR - S_Star) - 1.0);
        for (int n = 0; n < N; ++n) {
            double CLs = (SL - VL) / (SL - S_Star) * UL[n];
            double CRs = (SR - VR) / (SR - S_Star) * UR[n];
            if (n == 2 + dir) {
                CLs += (SL - VL) / (SL - S_Star) * rhoL * (S_Star - VL);
                CRs += (SR - VR) / (SR - S_Star) * rhoR * (S_Star - VR);
            } else if (n == 2 + d) {
                CLs += (SL - VL) / (SL - S_Star) * (S_Star - VL) * (rhoL * S_Star + pL / (SL - VL));
                CRs += (SR - VR) / (SR - S_Star) * (S_Star - VR) * (rhoR * S_Star + pR / (SR - VR));
            }
            flux[n] = 0.5 * (1.0 + s) * (FL[n] + Sm * (CLs - UL[n]))
                    + 0.5 * (1.0 - s) * (FR[n] + Sp * (CRs - UR[n]));
        }
    } else {
        /* std::max({0, VL+cL, VR+cR}) / std::min({0, VL-cL, VR-cR}) */
        double Sp = smax(smax(0.0, VL + cL), VR + cR);
        double Sm = smin(smin(0.0, VL - cL), VR - cR);
        for (int m = 0; m < d; ++m)
            vel[m] = (Sp * WL[2 + m] - Sm * WR[2 + m]) / (Sp - Sm);
        for (int n = 0; n < N; ++n)
            flux[n] = (Sp * FL[n] - Sm * FR[n] + Sp * Sm * (UR[n] - UL[n])) / (Sp - Sm);
    }
    double qL = (d == 3) ? (WL[2] * WL[2] + WL[3] * WL[3] + WL[4] * WL[4]) : (WL[2] * WL[2] + WL[3] * WL[3]);
    double qR = (d == 3) ? (WR[2] * WR[2] + WR[3] * WR[3] + WR[4] * WR[4]) : (WR[2] * WR[2] + WR[3] * WR[3]);
    *cmax = smax(cL + sqrt(qL), cR + sqrt(qR));
}

/* ------------------------------------------------------------------------------------------------
 * The N-phase five-equation model (PHYSICS == FIVEEQS_NPHASE, NPHASE = P->nphase): state
 * alpha_k rho_k (k < NP), rho u_d, rho E, alpha_k (k < NP), N = 2 NP + dim + 1
 * (include/Physics_Compressible5Eq_NPhase/Variables_Compressible5Eq_NPhase.H:21-29,
 *  Component2Vector_Compressible5Eq_NPhase.H); primitives alpha_k rho_k, u_d, p, alpha_k.
 * ---------------------------------------------------------------------------------------------- */

/* EquationOfState_Mixture, pressure-based branch
 * (include/Physics_Compressible5Eq_NPhase/EquationOfState_Mixture_Compressible5Eq_NPhase.H:15-67):
 * every phase's own alpha_k enters AP and BP. */
double ref_np_eos_mixture(const cmp_params* P, double input, const double* alpha, const double* arho,
                          int output_internal_energy, int* err)
{
    double rho_D = 0.0, AP = 0.0, BP = 0.0;
    for (int k = 0; k < P->nphase; ++k) rho_D += arho[k] * P->eos_D[k];
    for (int k = 0; k < P->nphase; ++k) {
        double phi = ref_nasg_phi(alpha[k], arho[k], P->eos_C[k], err);
        AP += phi * P->eos_A[k];
        BP += phi * P->eos_B[k];
    }
    if (output_internal_energy) return AP * input + BP + rho_D;
    return (input - rho_D - BP) / AP;
}

/* Conservative2Primitive (…Conservative2Primitive_Compressible5Eq_NPhase.H:16-60) */
void ref_np_cons2prim(const cmp_params* P, const double* U, double* W, int* err)
{
    const int d = P->dim, np = P->nphase;
    double rho = 0.0;
    for (int k = 0; k < np; ++k) rho += U[k];
    double vel[3] = {0.0, 0.0, 0.0};
    for (int m = 0; m < d; ++m) vel[m] = U[np + m] / rho;
    double ke = (d == 3) ? (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2])
                         : (vel[0] * vel[0] + vel[1] * vel[1]);
    double rho_e = U[np + d] - (0.5) * rho * ke;
    double p = ref_np_eos_mixture(P, rho_e, U + np + d + 1, U, 0, err);
    for (int k = 0; k < np; ++k) { W[k] = U[k]; W[np + d + 1 + k] = U[np + d + 1 + k]; }
    for (int m = 0; m < d; ++m) W[np + m] = vel[m];
    W[np + d] = p;
}

/* Primitive2Conservative (…Primitive2Conservative_Compressible5Eq_NPhase.H:15-62) */
void ref_np_prim2cons(const cmp_params* P, const double* W, double* U, int* err)
{
    const int d = P->dim, np = P->nphase;
    double rho = 0.0;
    for (int k = 0; k < np; ++k) rho += W[k];
    double mom[3];
    for (int m = 0; m < d; ++m) mom[m] = rho * W[np + m];
    double rho_e = ref_np_eos_mixture(P, W[np + d], W + np + d + 1, W, 1, err);
    double ke = (d == 3) ? (W[np] * W[np] + W[np + 1] * W[np + 1] + W[np + 2] * W[np + 2])
                         : (W[np] * W[np] + W[np + 1] * W[np + 1]);
    double rE = rho_e + (0.5) * rho * ke;
    for (int k = 0; k < np; ++k) { U[k] = W[k]; U[np + d + 1 + k] = W[np + d + 1 + k]; }
    for (int m = 0; m < d; ++m) U[np + m] = mom[m];
    U[np + d] = rE;
}

/* rho_c2_Mixture (…rho_c2_Mixture_Compressible5Eq_NPhase.H:15-52) of the phase values rc[k] =
 * rho_c2_Phase(p, A_k, B_k) */
static double ref_np_rho_c2_mixture(const cmp_params* P, const double* rc, const double* alpha,
                                    const double* arho, int id_k, int* err)
{
    if (id_k == 0) {
        double APrc = 0.0, AP = 0.0;
        for (int k = 0; k < P->nphase; ++k) {
            double phi = ref_nasg_phi(alpha[k], arho[k], P->eos_C[k], err);
            APrc += alpha[k] * P->eos_A[k] * rc[k];
            AP += phi * P->eos_A[k];
        }
        return APrc / AP;
    }
    double inv = 0.0;
    for (int k = 0; k < P->nphase; ++k) {
        double phi = ref_nasg_phi(alpha[k], arho[k], P->eos_C[k], err);
        inv += phi / rc[k];
    }
    return 1.0 / inv;
}

/* SoundSpeed (…SoundSpeed_Compressible5Eq_NPhase.H:15-52) */
double ref_np_sound_speed(const cmp_params* P, double p, const double* alpha, const double* arho,
                          int id_k, int* err)
{
    double rc[CMP_MAX_PHASES];
    for (int k = 0; k < P->nphase; ++k) rc[k] = ref_rho_c2_phase(p, P->eos_A[k], P->eos_B[k]);
    double rho_c2 = ref_np_rho_c2_mixture(P, rc, alpha, arho, id_k, err);
    double rho = 0.0;
    for (int k = 0; k < P->nphase; ++k) rho += arho[k];
    return sqrt(rho_c2 / rho);
}

/* NonConservative (…NonConservative_Compressible5Eq_NPhase.H:15-58): K_k */
void ref_np_K(const cmp_params* P, double p, const double* alpha, const double* arho, int id_k,
              double* K, int* err)
{
    if (id_k == 0) {
        for (int k = 0; k < P->nphase; ++k) K[k] = 0.0;
        return;
    }
    double rc[CMP_MAX_PHASES];
    for (int k = 0; k < P->nphase; ++k) rc[k] = ref_rho_c2_phase(p, P->eos_A[k], P->eos_B[k]);
    double rho_c2 = ref_np_rho_c2_mixture(P, rc, alpha, arho, id_k, err);
    for (int k = 0; k < P->nphase; ++k) K[k] = (rho_c2 / rc[k] - 1.0);
}

/* Conservative2Flux (…Conservative2Flux_Compressible5Eq_NPhase.H:15-80) */
void ref_np_cons2flux(const cmp_params* P, const double* U, int dir, double* F, int* err)
{
    const int d = P->dim, np = P->nphase;
    double W[REF_NMAX];
    ref_np_cons2prim(P, U, W, err);
    double p = W[np + d], V = W[np + dir];
    for (int k = 0; k < np; ++k) {
        F[k] = W[k] * V;
        F[np + d + 1 + k] = W[np + d + 1 + k] * V;
    }
    for (int m = 0; m < d; ++m) F[np + m] = U[np + m] * V;
    F[np + d] = (U[np + d] + p) * V;
    F[np + dir] += p;
}

/* Riemann_HLLC (…Riemann_HLLC_Compressible5Eq_NPhase.H:16-162) and Riemann_HLL
 * (…Riemann_HLL_Compressible5Eq_NPhase.H) */
void ref_np_riemann(const cmp_params* P, const double* UL, const double* UR, int dir, double* flux,
                    double* vel, double* cmax, int* err)
{
    const int d = P->dim, N = P->nstate, np = P->nphase, id_k = P->source_term;
    double WL[REF_NMAX], WR[REF_NMAX], FL[REF_NMAX], FR[REF_NMAX];
    ref_np_cons2prim(P, UL, WL, err);
    ref_np_cons2prim(P, UR, WR, err);
    ref_np_cons2flux(P, UL, dir, FL, err);
    ref_np_cons2flux(P, UR, dir, FR, err);
    double pL = WL[np + d], pR = WR[np + d];
    double cL = ref_np_sound_speed(P, pL, WL + np + d + 1, WL, id_k, err);
    double cR = ref_np_sound_speed(P, pR, WR + np + d + 1, WR, id_k, err);
    double VL = WL[np + dir], VR = WR[np + dir];
    if (P->id_riemann == 2) {
        double V_avg = 0.5 * (VL + VR);
        double c_avg = 0.5 * (cL + cR);
        double SL = smin(V_avg - c_avg, VL - cL);
        double SR = smax(V_avg + c_avg, VR + cR);
        double Sm = smin(0.0, SL);
        double Sp = smax(0.0, SR);
        double rhoL = 0.0, rhoR = 0.0;
        for (int k = 0; k < np; ++k) { rhoL += WL[k]; rhoR += WR[k]; }
        double S_Star = (pR - pL + rhoL * VL * (SL - VL) - rhoR * VR * (SR - VR))
                      / (rhoL * (SL - VL) - rhoR * (SR - VR));
        double s = ref_sign(S_Star);
        for (int m = 0; m < d; ++m)
            vel[m] = 0.5 * (1.0 + s) * WL[np + m] + 0.5 * (1.0 - s) * WR[np + m];
        vel[dir] += 0.5 * (1.0 + s) * Sm * ((SL - VL) / (SL - S_Star) - 1.0)
                  + 0.5 * (1.0 - s) * Sp * ((SR - VR) / (SR - S_Star) - 1.0);
        for (int n = 0; n < N; ++n) {
            double CLs = (SL - VL) / (SL - S_Star) * UL[n];
            double CRs = (SR - VR) / (SR - S_Star) * UR[n];
            if (n == np + dir) {
                CLs += (SL - VL) / (SL - S_Star) * rhoL * (S_Star - VL);
                CRs += (SR - VR) / (SR - S_Star) * rhoR * (S_Star - VR);
            } else if (n == np + d) {
                CLs += (SL - VL) / (SL - S_Star) * (S_Star - VL) * (rhoL * S_Star + pL / (SL - VL));
                CRs += (SR - VR) / (SR - S_Star) * (S_Star - VR) * (rhoR * S_Star + pR / (SR - VR));
            }
            flux[n] = 0.5 * (1.0 + s) * (FL[n] + Sm * (CLs - UL[n]))
                    + 0.5 * (1.0 - s) * (FR[n] + Sp * (CRs - UR[n]));
        }
    } else {
        double Sp = smax(smax(0.0, VL + cL), VR + cR);
        double Sm = smin(smin(0.0, VL - cL), VR - cR);
        for (int m = 0; m < d; ++m)
            vel[m] = (Sp * WL[np + m] - Sm * WR[np + m]) / (Sp - Sm);
        for (int n = 0; n < N; ++n)
            flux[n] = (Sp * FL[n] - Sm * FR[n] + Sp * Sm * (UR[n] - UL[n])) / (Sp - Sm);
    }
    double qL = 0.0, qR = 0.0;
    for (int m = 0; m < d; ++m) {
        qL = (m == 0) ? WL[np] * WL[np] : qL + WL[np + m] * WL[np + m];
        qR = (m == 0) ? WR[np] * WR[np] : qR + WR[np + m] * WR[np + m];
    }
    *cmax = smax(cL + sqrt(qL), cR + sqrt(qR));
}

/* Limiter_Consistent (include/Operators_FV/Limiter_Consistent.H:12-39): the first phase of largest
 * |slope| (std::max_element) takes theta = clip((s_max - S)/s_max), the others clip(s_max/(s_max - S)),
 * S the sum of the slopes; theta = 1 for all when |s_max| or |S| is below 1e-10. */
void ref_np_limiter_consistent(int np, const double* slope, double* theta)
{
    const double ep = 1.0e-10;
    double S = 0.0;
    int imax = 0;
    for (int k = 0; k < np; ++k) S += slope[k];
    for (int k = 1; k < np; ++k)
        if (fabs(slope[imax]) < fabs(slope[k])) imax = k;
    double smx = slope[imax];
    if (fabs(smx) < ep || fabs(S) < ep) {
        for (int k = 0; k < np; ++k) theta[k] = 1.0;
        return;
    }
    double smS = smx - S;
    for (int k = 0; k < np; ++k) theta[k] = smax(smin(smx / smS, 1.0), 0.0);
    theta[imax] = smax(smin(smS / smx, 1.0), 0.0);
}

/* FVM_Reconstruction2Face_Conservative_LimiterConsistent (…LimiterConsistent_Compressible5Eq_NPhase.H:15-73)
 * with FVM_Reconstruction2Face_LimiterConsistent (…:15-68) and Reconstruction2Face_Linear: the face
 * state Uf (unlimited, conservative) and the cell state Uc -> the limited conservative face state. */
void ref_np_limit_consistent(const cmp_params* P, const double* Uf, const double* Uc, double dx,
                             int id_face, double* Uo, int* err)
{
    const int d = P->dim, np = P->nphase, ia = np + d + 1;
    double Wf[REF_NMAX], Wc[REF_NMAX];
    ref_np_cons2prim(P, Uf, Wf, err);
    ref_np_cons2prim(P, Uc, Wc, err);
    double sr[CMP_MAX_PHASES], sa[CMP_MAX_PHASES], th[CMP_MAX_PHASES];
    for (int k = 0; k < np; ++k) {
        sr[k] = (Wf[k] - Wc[k]) / ((double)id_face * 0.5 * dx);
        sa[k] = (Wf[ia + k] - Wc[ia + k]) / ((double)id_face * 0.5 * dx);
    }
    ref_np_limiter_consistent(np, sa, th);
    for (int k = 0; k < np; ++k) {
        double slope = th[k] * sr[k];
        Wf[k] = Wc[k] + (double)id_face * 0.5 * slope * dx;
        slope = th[k] * sa[k];
        Wf[ia + k] = Wc[ia + k] + (double)id_face * 0.5 * slope * dx;
    }
    ref_np_prim2cons(P, Wf, Uo, err);
}

/* FVM_Reconstruction2FaceLR_Conservative_K of the N-phase model
 * (…FVM_Reconstruction2FaceLR_Conservative_K_Compressible5Eq_NPhase.H) with the per-side
 * FVM_Reconstruction2Face_Conservative_{WENO5,MUSCL}_D (…WENO5_D_Compressible5Eq_NPhase.H:15-73,
 * …MUSCL_D_…:15-74: scalar reconstruction of every primitive, prim->cons, the consistent limiter)
 * and …Godunov_D (cons->prim->cons, no limiter). Ust as in ref_recon_lr. */
static int ref_np_recon_lr(const cmp_params* P, double Ust[6][REF_NMAX], double dx_dir, double* UL,
                           double* UR, int* err)
{
    const int N = P->nstate;
    double Wst[6][REF_NMAX], WLf[REF_NMAX], WRf[REF_NMAX], Uu[REF_NMAX];
    if (P->id_weno == 1 && P->id_limiter == 5) {
        for (int s = -3; s <= 1; ++s) ref_np_cons2prim(P, Ust[s + 3], Wst[s + 3], err);
        for (int n = 0; n < N; ++n)
            WLf[n] = ref_weno5(Wst[0][n], Wst[1][n], Wst[2][n], Wst[3][n], Wst[4][n], 1);
        ref_np_prim2cons(P, WLf, Uu, err);
        ref_np_limit_consistent(P, Uu, Ust[2], dx_dir, 1, UL, err);
        for (int s = -2; s <= 2; ++s) ref_np_cons2prim(P, Ust[s + 3], Wst[s + 3], err);
        for (int n = 0; n < N; ++n)
            WRf[n] = ref_weno5(Wst[1][n], Wst[2][n], Wst[3][n], Wst[4][n], Wst[5][n], -1);
        ref_np_prim2cons(P, WRf, Uu, err);
        ref_np_limit_consistent(P, Uu, Ust[3], dx_dir, -1, UR, err);
    } else if (P->id_weno == 0 && P->id_limiter >= 1 && P->id_limiter <= 4) {
        for (int s = -2; s <= 0; ++s) ref_np_cons2prim(P, Ust[s + 3], Wst[s + 3], err);
        for (int n = 0; n < N; ++n)
            WLf[n] = ref_muscl(Wst[1][n], Wst[2][n], Wst[3][n], dx_dir, 1, P->id_limiter);
        ref_np_prim2cons(P, WLf, Uu, err);
        ref_np_limit_consistent(P, Uu, Ust[2], dx_dir, 1, UL, err);
        for (int s = -1; s <= 1; ++s) ref_np_cons2prim(P, Ust[s + 3], Wst[s + 3], err);
        for (int n = 0; n < N; ++n)
            WRf[n] = ref_muscl(Wst[2][n], Wst[3][n], Wst[4][n], dx_dir, -1, P->id_limiter);
        ref_np_prim2cons(P, WRf, Uu, err);
        ref_np_limit_consistent(P, Uu, Ust[3], dx_dir, -1, UR, err);
    } else if (P->id_weno == 0 && P->id_limiter == 0) {
        ref_np_cons2prim(P, Ust[2], WLf, err);
        ref_np_prim2cons(P, WLf, UL, err);
        ref_np_cons2prim(P, Ust[3], WRf, err);
        ref_np_prim2cons(P, WRf, UR, err);
    } else {
        return CMP_ERR_UNSUPPORTED;
    }
    return CMP_OK;
}

/* ------------------------------------------------------------------------------------------------
 * Fab access
 * ---------------------------------------------------------------------------------------------- */

typedef struct {
    int lo[3];      /* lowest stored index per dim (lo - ng, or 0 for inactive dims) */
    int ext[3];     /* stored extent per dim */
    long long sj, sk, sn;
} fab_t;

static fab_t make_fab(int dim, const cmp_box* b, int ng)
{
    fab_t f;
    for (int d = 0; d < 3; ++d) {
        if (d < dim) { f.lo[d] = b->lo[d] - ng; f.ext[d] = b->hi[d] - b->lo[d] + 1 + 2 * ng; }
        else         { f.lo[d] = b->lo[d];      f.ext[d] = 1; }
    }
    f.sj = f.ext[0];
    f.sk = f.sj * f.ext[1];
    f.sn = f.sk * f.ext[2];
    return f;
}

static inline long long fab_idx(const fab_t* f, int i, int j, int k)
{
    return (long long)(i - f->lo[0]) + f->sj * (j - f->lo[1]) + f->sk * (k - f->lo[2]);
}

long long cmp_fab_size(int dim, const cmp_box* box, int ng, int ncomp);

/* Normal of the volume fraction at the face (i,j,k) of direction dir (between cells c - e_dir and
 * c): FDM_Normal_{X,Y,Z}FaceFromCenter_{K,D} (include/Operators_FD/FDM_Normal_XFaceFromCenter_D.H:12-56
 * and its Y/Z twins) with FDM_Gradient*_*FaceFromCenter_D and Interpolation2Edge_Linear_D: the normal
 * derivative from the two cells, each transverse derivative from the two edge averages of four
 * cells (listed in axis order), normalised unless all derivatives are zero. */
static void ref_face_normal(const cmp_params* P, const cmp_geometry* G, const fab_t* f,
                            const double* U, int i, int j, int k, int dir, double* nrm)
{
    const int dim = P->dim;
    const long long sa = (long long)(3 + dim) * f->sn;
    int c[3] = {i, j, k};
    int e[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
#define AL(x, y, z) U[sa + fab_idx(f, (x), (y), (z))]
    double g[3] = {0.0, 0.0, 0.0};
    double fc = AL(c[0], c[1], c[2]);
    double fm = AL(c[0] - e[dir][0], c[1] - e[dir][1], c[2] - e[dir][2]);
    g[dir] = (fc - fm) / G->dx[dir];
    for (int t = 0; t < dim; ++t) {
        if (t == dir) continue;
        double edge[2];
        for (int s = 0; s < 2; ++s) {
            int sg = s ? 1 : -1;
            double A = fc;
            double B = fm;
            double C = AL(c[0] + sg * e[t][0], c[1] + sg * e[t][1], c[2] + sg * e[t][2]);
            double D = AL(c[0] - e[dir][0] + sg * e[t][0], c[1] - e[dir][1] + sg * e[t][1],
                          c[2] - e[dir][2] + sg * e[t][2]);
            if (dir < t) edge[s] = 0.25 * A + 0.25 * B + 0.25 * C + 0.25 * D;
            else         edge[s] = 0.25 * A + 0.25 * C + 0.25 * B + 0.25 * D;
        }
        g[t] = (edge[1] - edge[0]) / G->dx[t];
    }
#undef AL
    double sum = 0.0;
    for (int m = 0; m < dim; ++m) sum += g[m] * g[m];
    for (int m = 0; m < dim; ++m) nrm[m] = (sum == 0.0) ? 0.0 : g[m] / sqrt(sum);
}

/* Reconstruct_State_From_VolumeFraction
 * (include/Operators_Compressible5Eq/FVM_Reconstruction2Face_Conservative_THINC_K_Compressible5Eq.H:16-64):
 * the face state keeps the phase densities, velocity and pressure of Uin and takes alpha1 = a1n. */
static void ref_state_from_vf(const cmp_params* P, const double* Uin, double a1n, double* Uo,
                              int* err)
{
    const int d = P->dim;
    double W[REF_NMAX];
    ref_cons2prim(P, Uin, W, err);
    double a1 = W[3 + d], p = W[2 + d];
    double rho1 = W[0] / a1;
    double rho2 = W[1] / (1.0 - a1);
    double a1r1n = a1n * rho1;
    double a2r2n = (1.0 - a1n) * rho2;
    double rho_n = a1r1n + a2r2n;
    double ke = (d == 3) ? (W[2] * W[2] + W[3] * W[3] + W[4] * W[4]) : (W[2] * W[2] + W[3] * W[3]);
    double K = 0.5 * ke;
    double r1e1 = P->eos_A[0] * p + P->eos_B[0] + rho1 * P->eos_D[0];
    double r2e2 = P->eos_A[1] * p + P->eos_B[1] + rho2 * P->eos_D[1];
    if (P->eos_C[0] > 0.0) {
        double xi = 1.0 - P->eos_C[0] * rho1;
        if (!(xi > 0.0)) ERR_SET(err, CMP_ERR_COVOLUME);
        r1e1 = xi * (P->eos_A[0] * p + P->eos_B[0]) + rho1 * P->eos_D[0];
    }
    if (P->eos_C[1] > 0.0) {
        double xi = 1.0 - P->eos_C[1] * rho2;
        if (!(xi > 0.0)) ERR_SET(err, CMP_ERR_COVOLUME);
        r2e2 = xi * (P->eos_A[1] * p + P->eos_B[1]) + rho2 * P->eos_D[1];
    }
    Uo[0] = a1r1n;
    Uo[1] = a2r2n;
    for (int m = 0; m < d; ++m) Uo[2 + m] = W[2 + m] * rho_n;
    Uo[2 + d] = K * rho_n + r1e1 * a1n + r2e2 * (1.0 - a1n);
    Uo[3 + d] = a1n;
}

/* SignNZ (include/Tools/Sign.H:27-37) */
static inline double ref_signnz(double x) { return (x > 0.0) ? 1.0 : -1.0; }

/* FVM_Conservative2FluxHyperbolic_K, ID_Bound == 0 branch
 * (include/Operators_Compressible5Eq/FVM_Conservative2FluxHyperbolic_K_Compressible5Eq.H:15-170):
 * reconstruction dispatch FVM_Reconstruction2FaceLR_Conservative_K (…FaceLR_Conservative_K…H:15-82),
 * the per-side primitive reconstruction FVM_Reconstruction2Face_Conservative_{WENO5,MUSCL,Godunov}_D
 * (cons->prim of every stencil cell, scalar reconstruction per primitive, prim->cons), the Riemann
 * solver, and FluxNC[0] = u*_n (:165-168). L uses cells i-3..i+1 (id_face +1), R uses i-2..i+2 (-1). */
/* FVM_Reconstruction2FaceLR_Conservative_K (…FaceLR_Conservative_K_Compressible5Eq.H:15-82) with
 * the per-side primitive reconstructions FVM_Reconstruction2Face_Conservative_{WENO5,WENO5B,MUSCL,
 * Godunov}_D: cons->prim of every stencil cell of the side, scalar reconstruction per primitive,
 * prim->cons. Ust holds the conservative states of cells -3..+2 relative to the face's high cell;
 * L uses cells -3..+1 (id_face +1), R cells -2..+2 (id_face -1). */
static int ref_recon_lr(const cmp_params* P, double Ust[6][REF_NMAX], int id_weno, int id_limiter,
                        double dx_dir, double* UL, double* UR, int* err)
{
    const int N = P->nstate;
    double Wst[6][REF_NMAX], WLf[REF_NMAX], WRf[REF_NMAX];
    if (id_weno == 1 && (id_limiter == 5 || id_limiter == 6)) {
        for (int s = -3; s <= 1; ++s) ref_cons2prim(P, Ust[s + 3], Wst[s + 3], err);
        if (id_limiter == 5) {
            for (int n = 0; n < N; ++n)
                WLf[n] = ref_weno5(Wst[0][n], Wst[1][n], Wst[2][n], Wst[3][n], Wst[4][n], 1);
        } else {
            ref_weno5b_side(P, Wst, 1, WLf);
        }
        ref_prim2cons(P, WLf, UL, err);
        /* R: the cons->prim of the stencil cells is recomputed, as the reference does */
        for (int s = -2; s <= 2; ++s) ref_cons2prim(P, Ust[s + 3], Wst[s + 3], err);
        if (id_limiter == 5) {
            for (int n = 0; n < N; ++n)
                WRf[n] = ref_weno5(Wst[1][n], Wst[2][n], Wst[3][n], Wst[4][n], Wst[5][n], -1);
        } else {
            ref_weno5b_side(P, &Wst[1], -1, WRf);
        }
        ref_prim2cons(P, WRf, UR, err);
    } else if (id_weno == 0 && id_limiter >= 1 && id_limiter <= 4) {
        for (int s = -2; s <= 0; ++s) ref_cons2prim(P, Ust[s + 3], Wst[s + 3], err);
        for (int n = 0; n < N; ++n)
            WLf[n] = ref_muscl(Wst[1][n], Wst[2][n], Wst[3][n], dx_dir, 1, id_limiter);
        ref_prim2cons(P, WLf, UL, err);
        for (int s = -1; s <= 1; ++s) ref_cons2prim(P, Ust[s + 3], Wst[s + 3], err);
        for (int n = 0; n < N; ++n)
            WRf[n] = ref_muscl(Wst[2][n], Wst[3][n], Wst[4][n], dx_dir, -1, id_limiter);
        ref_prim2cons(P, WRf, UR, err);
    } else if (id_weno == 0 && id_limiter == 0) {
        ref_cons2prim(P, Ust[2], WLf, err);
        ref_prim2cons(P, WLf, UL, err);
        ref_cons2prim(P, Ust[3], WRf, err);
        ref_prim2cons(P, WRf, UR, err);
    } else {
        return CMP_ERR_UNSUPPORTED;
    }
    return CMP_OK;
}

/* FVM_Conservative2FluxHyperbolic_K, ID_Bound == 0 branch
 * (include/Operators_Compressible5Eq/FVM_Conservative2FluxHyperbolic_K_Compressible5Eq.H:15-170):
 * reconstruction (base scheme, or THINC, …THINC_K_Compressible5Eq.H:70-192), the Riemann solver and
 * FluxNC[0] = u*_n (:165-168). (i,j,k) is the face between cells c - e_dir and c = (i,j,k). */
int ref_face_flux(const cmp_params* P, const cmp_geometry* G, const cmp_box* box, int ng,
                  const double* U, int i, int j, int k, int dir,
                  double* flux, double* fluxnc, double* cmax)
{
    const int N = P->nstate, d = P->dim;
    const double dx_dir = G->dx[dir];
    int err = 0;
    fab_t f = make_fab(P->dim, box, ng);
    int off[3] = {0, 0, 0};
    off[dir] = 1;
    double Ust[6][REF_NMAX];
    int smin_ = -3, smax_ = 2;
    if (P->id_weno == 0 && P->id_limiter == 0 && !P->id_thinc) { smin_ = -1; smax_ = 0; }
    else if (P->id_weno == 0) { smin_ = -2; smax_ = 1; }
    for (int s = smin_; s <= smax_; ++s) {
        long long c = fab_idx(&f, i + s * off[0], j + s * off[1], k + s * off[2]);
        for (int n = 0; n < N; ++n) Ust[s + 3][n] = U[c + n * f.sn];
    }
    double UL[REF_NMAX], UR[REF_NMAX];
    int r;
    if (P->nphase) {
        /* FVM_Conservative2FluxHyperbolic_K of the N-phase model
         * (…FVM_Conservative2FluxHyperbolic_K_Compressible5Eq_NPhase.H:16-150), ID_Bound == 0 */
        if (P->id_thinc) return CMP_ERR_UNSUPPORTED;
        r = ref_np_recon_lr(P, Ust, dx_dir, UL, UR, &err);
        if (r) return r;
        double vel[3];
        ref_np_riemann(P, UL, UR, dir, flux, vel, cmax, &err);
        *fluxnc = vel[dir];
        return err;
    }
    if (P->id_thinc) {
        double nrm[3];
        ref_face_normal(P, G, &f, U, i, j, k, dir, nrm);
        double Wd[REF_NMAX];
        for (int s = -2; s <= 1; ++s) ref_cons2prim(P, Ust[s + 3], Wd, &err);   /* unused, as in the reference */
        const int iA = 3 + d;
        double Al[4] = {Ust[1][iA], Ust[2][iA], Ust[3][iA], Ust[4][iA]};   /* mm, m, c, p */
        int gL = (Al[2] - Al[1]) * (Al[1] - Al[0]) > 0.0;
        int gR = (Al[3] - Al[2]) * (Al[2] - Al[1]) > 0.0;
        int sym = (Al[0] == Al[3]) && (Al[1] == Al[2]) && (Al[1] != Al[0]);
        int gate = (!P->thinc_use_normal) || (fabs(nrm[dir]) > P->thinc_normal_threshold);
        double eps = P->thinc_eps;
        int intL = (Al[1] > eps && Al[1] < 1.0 - eps) && (gL || sym) && gate;
        int intR = (Al[2] > eps && Al[2] < 1.0 - eps) && (gR || sym) && gate;
        if (intL || intR) {
            r = ref_recon_lr(P, Ust, 0, 1, dx_dir, UL, UR, &err);   /* MUSCL-minmod base */
            if (r) return r;
            double Lb[REF_NMAX], Rb[REF_NMAX];
            for (int n = 0; n < N; ++n) { Lb[n] = UL[n]; Rb[n] = UR[n]; }
            double beta = P->thinc_use_normal ? (P->thinc_beta * fabs(nrm[dir]) + 0.01) : P->thinc_beta;
            if (intL) {
                double sigma = ref_signnz(Al[2] - Al[0]);
                ref_state_from_vf(P, Lb, ref_thinc(beta, sigma, Al[1], 1), UL, &err);
            }
            if (intR) {
                double sigma = ref_signnz(Al[3] - Al[1]);
                ref_state_from_vf(P, Rb, ref_thinc(beta, sigma, Al[2], 0), UR, &err);
            }
        } else {
            r = ref_recon_lr(P, Ust, P->id_weno, P->id_limiter, dx_dir, UL, UR, &err);
            if (r) return r;
        }
    } else {
        r = ref_recon_lr(P, Ust, P->id_weno, P->id_limiter, dx_dir, UL, UR, &err);
        if (r) return r;
    }
    double vel[3];
    ref_riemann(P, UL, UR, dir, flux, vel, cmax, &err);
    *fluxnc = vel[dir];
    return err;
}

/* ------------------------------------------------------------------------------------------------
 * One box: ComputeFaceFluxes (source/Class_Compressible_PhaseField.cpp:3152-3254) and
 * FluxDivergence (:3263-3391) with FVM_SurfaceIntegral (include/Operators_FV/FVM_SurfaceIntegral.H:17-33),
 * FVM_Divergence_{K,D} (include/Operators_FV/FVM_Divergence_K.H:12-42, _D.H:27-29) and
 * FVM_SurfaceIntegral_NC (include/Operators_Compressible5Eq/FVM_SurfaceIntegralNC_Compressible5Eq.H:15-52).
 * ---------------------------------------------------------------------------------------------- */

/* The faces of one box (ComputeFaceFluxes): F[d] (nface x N, x fastest) and FN[d] (u*_n), allocated
 * here; *cm the c_max of its valid cells. */
static int box_faces(const cmp_params* P, const cmp_geometry* G, const cmp_box* box, int ng,
                     const double* U, double** F, double** FN, double* cm_out)
{
    const int dim = P->dim, N = P->nstate;
    int nv[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) nv[d] = box->hi[d] - box->lo[d] + 1;
    int err = 0;
    double cm = 0.0;
    for (int d = 0; d < dim; ++d) {
        int fe[3] = {nv[0], nv[1], nv[2]};
        fe[d] += 1;
        long long nface = (long long)fe[0] * fe[1] * fe[2];
        F[d] = (double*)malloc(sizeof(double) * nface * N);
        FN[d] = (double*)malloc(sizeof(double) * nface);
        int e = 0;
        double cmd = 0.0;
        int lo2 = (dim == 3) ? box->lo[2] : 0;
#pragma omp parallel for collapse(2) reduction(max:cmd) reduction(max:e) schedule(static)
        for (int kk = 0; kk < fe[2]; ++kk)
            for (int jj = 0; jj < fe[1]; ++jj)
                for (int ii = 0; ii < fe[0]; ++ii) {
                    int i = box->lo[0] + ii, j = box->lo[1] + jj, k = lo2 + kk;
                    double fl[REF_NMAX], fnc, c;
                    int r = ref_face_flux(P, G, box, ng, U, i, j, k, d, fl, &fnc, &c);
                    if (r > e) e = r;
                    long long q = ii + (long long)fe[0] * (jj + (long long)fe[1] * kk);
                    for (int n = 0; n < N; ++n) F[d][q + n * nface] = fl[n];
                    FN[d][q] = fnc;
                    /* c_max(i,j,k) of a valid cell takes the faces (i,j,k) (:3212); the valid-cell
                     * reduction (:4490) therefore skips the box's high face */
                    int idx[3] = {ii, jj, kk};
                    if (idx[d] < nv[d]) {
                        if (isnan(c)) { if (CMP_ERR_NAN > e) e = CMP_ERR_NAN; }
                        else if (cmd < c) cmd = c;
                    }
                }
        if (e && !err) err = e;
        if (cm < cmd) cm = cmd;
    }
    *cm_out = cm;
    return err;
}

/* FluxDivergence of one box from its faces F, FN: dUdt (valid cells, x fastest, component-major). */
static int box_divergence(const cmp_params* P, const cmp_geometry* G, const cmp_box* box, int ng,
                          const double* U, double* const* F, double* const* FN, double* dUdt)
{
    const int dim = P->dim, N = P->nstate;
    int nv[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) nv[d] = box->hi[d] - box->lo[d] + 1;
    long long ncell = (long long)nv[0] * nv[1] * nv[2];
    /* divergence, then the non-conservative term, per valid cell */
    fab_t f = make_fab(dim, box, ng);
    long long sf[3][3];   /* face strides per direction */
    long long nf[3];
    for (int d = 0; d < dim; ++d) {
        int fe[3] = {nv[0], nv[1], nv[2]};
        fe[d] += 1;
        sf[d][0] = 1; sf[d][1] = fe[0]; sf[d][2] = (long long)fe[0] * fe[1];
        nf[d] = (long long)fe[0] * fe[1] * fe[2];
    }
    const int iA = 3 + dim;
    int e2 = 0;
#pragma omp parallel for collapse(2) reduction(max:e2) schedule(static)
    for (int kk = 0; kk < nv[2]; ++kk)
        for (int jj = 0; jj < nv[1]; ++jj)
            for (int ii = 0; ii < nv[0]; ++ii) {
                long long cell = ii + (long long)nv[0] * (jj + (long long)nv[1] * kk);
                int idx[3] = {ii, jj, kk};
                for (int n = 0; n < N; ++n) {
                    double div = 0.0;
                    for (int d = 0; d < dim; ++d) {
                        long long qm = idx[0] * sf[d][0] + idx[1] * sf[d][1] + idx[2] * sf[d][2];
                        long long qp = qm + sf[d][d];
                        double term = (F[d][qp + n * nf[d]] - F[d][qm + n * nf[d]]) / G->dx[d];
                        div = (d == 0) ? term : div + term;
                    }
                    dUdt[cell + n * ncell] = -div;
                }
                /* FVM_SurfaceIntegral_NC */
                int i = box->lo[0] + ii, j = box->lo[1] + jj, k = ((dim == 3) ? box->lo[2] : 0) + kk;
                long long c = fab_idx(&f, i, j, k);
                double Uc[REF_NMAX];
                for (int n = 0; n < N; ++n) Uc[n] = U[c + n * f.sn];
                int le = 0;
                int id_k = P->source_term;
                if (P->nphase) {
                    /* …FVM_SurfaceIntegralNC_Compressible5Eq_NPhase.H:16-56: every alpha_k */
                    const int np = P->nphase, ia = np + dim + 1;
                    double Kn[CMP_MAX_PHASES];
                    double Wc[REF_NMAX];
                    ref_np_cons2prim(P, Uc, Wc, &le);
                    ref_np_K(P, Wc[np + dim], Uc + ia, Uc, id_k, Kn, &le);
                    double divu = 0.0;
                    for (int d = 0; d < dim; ++d) {
                        long long qm = idx[0] * sf[d][0] + idx[1] * sf[d][1] + idx[2] * sf[d][2];
                        long long qp = qm + sf[d][d];
                        double term = (FN[d][qp] - FN[d][qm]) / G->dx[d];
                        divu = (d == 0) ? term : divu + term;
                    }
                    for (int q = 0; q < np; ++q) {
                        double ai = Uc[ia + q];
                        double phii = (id_k == 1) ? ref_nasg_phi(ai, Uc[q], P->eos_C[q], &le) : ai;
                        dUdt[cell + (ia + q) * ncell] += phii * (1.0 + Kn[q]) * divu;
                    }
                    if (le > e2) e2 = le;
                    continue;
                }
                double a1 = Uc[iA];
                double phi1 = (id_k == 1) ? ref_nasg_phi(a1, Uc[0], P->eos_C[0], &le) : a1;
                double K[2] = {0.0, 0.0};
                if (id_k == 1) {
                    double Wc[REF_NMAX];
                    ref_cons2prim(P, Uc, Wc, &le);
                    ref_nonconservative_K(P, Wc[2 + dim], Wc[3 + dim], Wc[0], Wc[1], id_k, K, &le);
                }
                double divu = 0.0;
                for (int d = 0; d < dim; ++d) {
                    long long qm = idx[0] * sf[d][0] + idx[1] * sf[d][1] + idx[2] * sf[d][2];
                    long long qp = qm + sf[d][d];
                    double term = (FN[d][qp] - FN[d][qm]) / G->dx[d];
                    divu = (d == 0) ? term : divu + term;
                }
                dUdt[cell + iA * ncell] += phi1 * (1.0 + K[0]) * divu;
                if (le > e2) e2 = le;
            }
    return e2;
}

int ref_rhs_box(const cmp_params* P, const cmp_geometry* G, const cmp_box* box, int ng,
                const double* U, double* dUdt, double* cmax, double** flux_out, double** fluxnc_out)
{
    const int dim = P->dim, N = P->nstate;
    double* F[3] = {NULL, NULL, NULL};
    double* FN[3] = {NULL, NULL, NULL};
    double cm = 0.0;
    int err = box_faces(P, G, box, ng, U, F, FN, &cm);
    for (int d = 0; d < dim; ++d) {
        int fe[3] = {1, 1, 1};
        for (int e = 0; e < dim; ++e) fe[e] = box->hi[e] - box->lo[e] + 1 + (e == d ? 1 : 0);
        const long long nface = (long long)fe[0] * fe[1] * fe[2];
        if (flux_out && flux_out[d]) memcpy(flux_out[d], F[d], sizeof(double) * nface * N);
        if (fluxnc_out && fluxnc_out[d]) memcpy(fluxnc_out[d], FN[d], sizeof(double) * nface);
    }
    int e2 = box_divergence(P, G, box, ng, U, F, FN, dUdt);
    if (e2 && !err) err = e2;
    for (int d = 0; d < dim; ++d) { free(F[d]); free(FN[d]); }
    if (cmax && (*cmax < cm || isnan(cm))) *cmax = cm;
    return err;
}

/* ------------------------------------------------------------------------------------------------
 * Level ghost fill: FillPatchSingleLevel semantics (source/COMPAS.cpp:1647-1678) with the BC map
 * of SetPhysicsBC (source/Class_Compressible5Eq_PhaseField.cpp:162-244) and AMReX's BCType rules:
 *   int_dir      periodic image or the valid cell of the box that owns it
 *   foextrap     nearest in-domain cell along that direction (Outflow)
 *   reflect_even mirror image across the boundary (Symmetry/SlipWall/NoSlipWall, non-normal comps)
 *   reflect_odd  mirror image with the sign flipped (normal momentum; tangential for NoSlipWall)
 * A ghost cell outside the domain in several directions composes the per-direction maps.
 * ---------------------------------------------------------------------------------------------- */

static int box_contains(int dim, const cmp_box* b, const int* c)
{
    for (int d = 0; d < dim; ++d)
        if (c[d] < b->lo[d] || c[d] > b->hi[d]) return 0;
    return 1;
}

/* pf_bcs != 0: the Phase-Field fields' rule (PF_bcs, LS_bcs and Normal_bcs of
 * source/Class_Compressible_PhaseField.cpp:471-488, 1516-1530, 1970-1986): int_dir on periodic
 * sides, foextrap on every other side, no sign changes. */
static int fill_level(int dim, int N, int mom0, const cmp_geometry* G, int nbox, const cmp_box* boxes,
                      int ng, double* U, int pf_bcs)
{
    int err = 0;
    for (int b = 0; b < nbox; ++b) {
        const cmp_box* bx = &boxes[b];
        fab_t f = make_fab(dim, bx, ng);
        double* base = U + bx->offset;
        int lo2 = f.lo[2], hi2 = f.lo[2] + f.ext[2] - 1;
        int e = 0;
#pragma omp parallel for collapse(2) reduction(max:e) schedule(static)
        for (int k = lo2; k <= hi2; ++k)
            for (int j = f.lo[1]; j <= f.lo[1] + f.ext[1] - 1; ++j)
                for (int i = f.lo[0]; i <= f.lo[0] + f.ext[0] - 1; ++i) {
                    int c[3] = {i, j, k};
                    if (box_contains(dim, bx, c)) continue;
                    /* map to a source cell inside the domain */
                    int src[3] = {i, j, k};
                    double sgn[REF_NMAX];
                    for (int n = 0; n < N; ++n) sgn[n] = 1.0;
                    int ok = 1;
                    for (int d = 0; d < dim; ++d) {
                        int dlo = G->domain_lo[d], dhi = G->domain_hi[d];
                        if (src[d] >= dlo && src[d] <= dhi) continue;
                        int per = dhi - dlo + 1;
                        if (G->is_periodic[d]) {
                            while (src[d] < dlo) src[d] += per;
                            while (src[d] > dhi) src[d] -= per;
                            continue;
                        }
                        int code = (src[d] < dlo) ? G->lo_bc[d] : G->hi_bc[d];
                        if (pf_bcs) code = 2;
                        if (code == 2 || code == 0) {          /* foextrap (Interior on a non-periodic side: treat as outflow) */
                            src[d] = (src[d] < dlo) ? dlo : dhi;
                        } else if (code >= 3 && code <= 5) { /* reflect */
                            src[d] = (src[d] < dlo) ? (2 * dlo - 1 - src[d]) : (2 * dhi + 1 - src[d]);
                            /* normal momentum odd; tangential odd only for NoSlipWall */
                            for (int m = 0; m < dim; ++m) {
                                int odd = (m == d) || (code == 5);
                                if (odd) sgn[mom0 + m] = -sgn[mom0 + m];
                            }
                        } else {
                            ok = 0;
                        }
                    }
                    if (!ok) { if (CMP_ERR_BC > e) e = CMP_ERR_BC; continue; }
                    int owner = -1;
                    for (int q = 0; q < nbox; ++q)
                        if (box_contains(dim, &boxes[q], src)) { owner = q; break; }
                    if (owner < 0) { if (CMP_ERR_LAYOUT > e) e = CMP_ERR_LAYOUT; continue; }
                    fab_t fs = make_fab(dim, &boxes[owner], ng);
                    const double* sb = U + boxes[owner].offset;
                    long long si = fab_idx(&fs, src[0], src[1], src[2]);
                    long long di = fab_idx(&f, i, j, k);
                    for (int n = 0; n < N; ++n) base[di + n * f.sn] = sgn[n] * sb[si + n * fs.sn];
                }
        if (e > err) err = e;
    }
    return err;
}

int ref_level_fill_ghosts(const cmp_params* P, const cmp_geometry* G, int nbox,
                          const cmp_box* boxes, int ng, double* U)
{
    return fill_level(P->dim, P->nstate, P->nphase ? P->nphase : 2, G, nbox, boxes, ng, U, 0);
}

int ref_level_fill_pf(int dim, const cmp_geometry* G, int nbox, const cmp_box* boxes, int ng,
                      int ncomp, double* A)
{
    return fill_level(dim, ncomp, 0, G, nbox, boxes, ng, A, 1);
}

/* ------------------------------------------------------------------------------------------------
 * Level RHS, RK3 and time step
 * ---------------------------------------------------------------------------------------------- */

static long long valid_cells(int dim, const cmp_box* b)
{
    long long n = 1;
    for (int d = 0; d < dim; ++d) n *= (b->hi[d] - b->lo[d] + 1);
    return n;
}

/* compute_dUdt_FV over all boxes; dUdt is a level buffer with U's layout (valid part written) */
int ref_level_rhs(const cmp_params* P, const cmp_geometry* G, int nbox, const cmp_box* boxes,
                  int ng, const double* U, double* dUdt, double* cmax)
{
    const int dim = P->dim, N = P->nstate;
    int err = 0;
    for (int b = 0; b < nbox; ++b) {
        long long nc = valid_cells(dim, &boxes[b]);
        double* tmp = (double*)malloc(sizeof(double) * nc * N);
        int e = ref_rhs_box(P, G, &boxes[b], ng, U + boxes[b].offset, tmp, cmax, NULL, NULL);
        if (e && !err) err = e;
        fab_t f = make_fab(dim, &boxes[b], ng);
        int nv[3] = {1, 1, 1};
        for (int d = 0; d < dim; ++d) nv[d] = boxes[b].hi[d] - boxes[b].lo[d] + 1;
        int lo2 = (dim == 3) ? boxes[b].lo[2] : 0;
        for (int n = 0; n < N; ++n)
            for (int kk = 0; kk < nv[2]; ++kk)
                for (int jj = 0; jj < nv[1]; ++jj)
                    for (int ii = 0; ii < nv[0]; ++ii) {
                        long long q = fab_idx(&f, boxes[b].lo[0] + ii, boxes[b].lo[1] + jj, lo2 + kk);
                        dUdt[boxes[b].offset + q + n * f.sn] = tmp[ii + (long long)nv[0] * (jj + (long long)nv[1] * kk) + n * nc];
                    }
        free(tmp);
    }
    return err;
}

/* ThirdOrderSSPRK (source/Class_Compressible_PhaseField.cpp:3938-4107) on one level:
 *   stage 1: Uborder = FillPatch(Un);  U1 = LinComb(1, Uborder, dt, dUdt)               (:4012)
 *   stage 2: Uborder = FillPatch(U1);  U2 = LinComb(1/4, Uborder, 3/4, Un); Saxpy(dt/4)  (:4046-4047)
 *   stage 3: Uborder = FillPatch(U2);  U3 = LinComb(2/3, Uborder, 1/3, Un); Saxpy(2dt/3) (:4083-4084)
 * AMReX LinComb: dst = a*x + b*y; Saxpy: dst += a*src. c_max accumulates over the three stages. */
/* The right-hand side a RK3 step evaluates: the flow's (compute_dUdt_FV) or the Phase-Field's
 * (PhaseField_compute_dUdt_FV with a fixed Grad_Q). */
typedef struct {
    const cmp_pf_params* pf;   /* NULL: flow right-hand side */
    const double* gradq;
} rk3_rhs_t;

int ref_level_pf_rhs(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                     int nbox, const cmp_box* boxes, int ng, const double* U, const double* gradq,
                     double* dUdt);

static int rk3_step(const cmp_params* P, const cmp_geometry* G, int nbox, const cmp_box* boxes,
                    int ng, double* Un, double dt, double* cmax, const rk3_rhs_t* R)
{
    const int dim = P->dim, N = P->nstate;
    long long total = 0;
    for (int b = 0; b < nbox; ++b) total += cmp_fab_size(dim, &boxes[b], ng, N);
    double* Ub = (double*)malloc(sizeof(double) * total);     /* Uborder */
    double* Unew = (double*)malloc(sizeof(double) * total);
    double* L = (double*)malloc(sizeof(double) * total);
    int err = 0, e;
    const double w2 = 0.25 * dt, w3 = (2.0 / 3.0) * dt;
    for (int stage = 0; stage < 3; ++stage) {
        memcpy(Ub, stage == 0 ? Un : Unew, sizeof(double) * total);
        e = ref_level_fill_ghosts(P, G, nbox, boxes, ng, Ub); if (e && !err) err = e;
        if (R && R->pf) e = ref_level_pf_rhs(P, R->pf, G, nbox, boxes, ng, Ub, R->gradq, L);
        else            e = ref_level_rhs(P, G, nbox, boxes, ng, Ub, L, cmax);
        if (e && !err) err = e;
        for (int b = 0; b < nbox; ++b) {
            fab_t f = make_fab(dim, &boxes[b], ng);
            int nv[3] = {1, 1, 1};
            for (int d = 0; d < dim; ++d) nv[d] = boxes[b].hi[d] - boxes[b].lo[d] + 1;
            int lo2 = (dim == 3) ? boxes[b].lo[2] : 0;
            long long o = boxes[b].offset;
            for (int n = 0; n < N; ++n)
#pragma omp parallel for collapse(2) schedule(static)
                for (int kk = 0; kk < nv[2]; ++kk)
                    for (int jj = 0; jj < nv[1]; ++jj)
                        for (int ii = 0; ii < nv[0]; ++ii) {
                            long long q = o + fab_idx(&f, boxes[b].lo[0] + ii, boxes[b].lo[1] + jj, lo2 + kk) + n * f.sn;
                            if (stage == 0) {
                                Unew[q] = 1.0 * Ub[q] + dt * L[q];
                            } else if (stage == 1) {
                                Unew[q] = 0.25 * Ub[q] + 0.75 * Un[q];
                                Unew[q] += w2 * L[q];
                            } else {
                                Unew[q] = (2.0 / 3.0) * Ub[q] + (1.0 / 3.0) * Un[q];
                                Unew[q] += w3 * L[q];
                            }
                        }
        }
    }
    /* the new state's valid cells replace Un's; Un's ghost cells are left as they were */
    for (int b = 0; b < nbox; ++b) {
        fab_t f = make_fab(dim, &boxes[b], ng);
        int nv[3] = {1, 1, 1};
        for (int d = 0; d < dim; ++d) nv[d] = boxes[b].hi[d] - boxes[b].lo[d] + 1;
        int lo2 = (dim == 3) ? boxes[b].lo[2] : 0;
        for (int n = 0; n < N; ++n)
            for (int kk = 0; kk < nv[2]; ++kk)
                for (int jj = 0; jj < nv[1]; ++jj)
                    for (int ii = 0; ii < nv[0]; ++ii) {
                        long long q = boxes[b].offset + fab_idx(&f, boxes[b].lo[0] + ii, boxes[b].lo[1] + jj, lo2 + kk) + n * f.sn;
                        Un[q] = Unew[q];
                    }
    }
    free(Ub); free(Unew); free(L);
    return err;
}

/* The other per-level integrators of run.TimeIntegrator (source/COMPAS.cpp:1306-1311):
 * ForwardEuler (source/Class_Compressible_PhaseField.cpp:3713-3800): U^{n+1} = 1 Ub(U^n) + dt L;
 * TVD-RK2 / SecondOrderSSPRK (:3805-3935): U1 = 1 Ub(U^n) + dt L1, U^{n+1} = 0.5 Ub(U1) + 0.5 U^n,
 *   += 0.5 dt L2;
 * RK4 / FourthOrderRK (:4111-4330): Ut = 1 Ub(U^n) + 0.5 dt L1, Un1 = 1 Ub(U^n) + (1/6) dt L1;
 *   Ut = 1 U^n + 0.5 dt L2(Ub(Ut)), Un1 += (1/3) dt L2; Ut = 1 U^n + 1 dt L3, Un1 += (1/3) dt L3;
 *   Un1 += (1/6) dt L4.
 * (LinComb: dst = a x + b y; Saxpy: dst += a y; every stage's Ub is the filled stage state.)
 * integrator: 1 ForwardEuler, 2 TVD-RK2, 3 TVD-RK3, 4 RK4. */
static int rk_step_any(const cmp_params* P, const cmp_geometry* G, int nbox, const cmp_box* boxes, int ng,
                       double* Un, double dt, double* cmax, int integrator, const rk3_rhs_t* R)
{
    if (integrator == 3) return rk3_step(P, G, nbox, boxes, ng, Un, dt, cmax, R);
    if (integrator != 1 && integrator != 2 && integrator != 4) return CMP_ERR_ARG;
    const int dim = P->dim, N = P->nstate;
    long long total = 0;
    for (int b = 0; b < nbox; ++b) total += cmp_fab_size(dim, &boxes[b], ng, N);
    double* Ub = (double*)malloc(sizeof(double) * total);
    double* L = (double*)malloc(sizeof(double) * total);
    double* Ut = (double*)malloc(sizeof(double) * total);
    double* Un1 = (double*)malloc(sizeof(double) * total);
    int err = 0, e;
    const int nst = (integrator == 4) ? 4 : integrator;
    for (int st = 0; st < nst; ++st) {
        const double* src = (st == 0) ? Un : ((integrator == 2) ? Un1 : Ut);
        memcpy(Ub, src, sizeof(double) * total);
        e = ref_level_fill_ghosts(P, G, nbox, boxes, ng, Ub); if (e && !err) err = e;
        if (R && R->pf) e = ref_level_pf_rhs(P, R->pf, G, nbox, boxes, ng, Ub, R->gradq, L);
        else            e = ref_level_rhs(P, G, nbox, boxes, ng, Ub, L, cmax);
        if (e && !err) err = e;
        for (int b = 0; b < nbox; ++b) {
            fab_t f = make_fab(dim, &boxes[b], ng);
            int nv[3] = {1, 1, 1};
            for (int d = 0; d < dim; ++d) nv[d] = boxes[b].hi[d] - boxes[b].lo[d] + 1;
            const int lo2 = (dim == 3) ? boxes[b].lo[2] : 0;
            for (int n = 0; n < N; ++n)
                for (int kk = 0; kk < nv[2]; ++kk)
                    for (int jj = 0; jj < nv[1]; ++jj)
                        for (int ii = 0; ii < nv[0]; ++ii) {
                            const long long q = boxes[b].offset + fab_idx(&f, boxes[b].lo[0] + ii, boxes[b].lo[1] + jj, lo2 + kk) + n * f.sn;
                            if (integrator == 1 || (integrator == 2 && st == 0)) {
                                Un1[q] = 1.0 * Ub[q] + dt * L[q];
                            } else if (integrator == 2) {
                                Un1[q] = 0.5 * Ub[q] + 0.5 * Un[q];
                                Un1[q] += (0.5 * dt) * L[q];
                            } else if (st == 0) {
                                Ut[q] = 1.0 * Ub[q] + (0.5 * dt) * L[q];
                                Un1[q] = 1.0 * Ub[q] + ((1.0 / 6.0) * dt) * L[q];
                            } else if (st == 1) {
                                Ut[q] = 1.0 * Un[q] + (0.5 * dt) * L[q];
                                Un1[q] += ((1.0 / 3.0) * dt) * L[q];
                            } else if (st == 2) {
                                Ut[q] = 1.0 * Un[q] + (1.0 * dt) * L[q];
                                Un1[q] += ((1.0 / 3.0) * dt) * L[q];
                            } else {
                                Un1[q] += ((1.0 / 6.0) * dt) * L[q];
                            }
                        }
        }
    }
    for (int b = 0; b < nbox; ++b) {
        fab_t f = make_fab(dim, &boxes[b], ng);
        int nv[3] = {1, 1, 1};
        for (int d = 0; d < dim; ++d) nv[d] = boxes[b].hi[d] - boxes[b].lo[d] + 1;
        const int lo2 = (dim == 3) ? boxes[b].lo[2] : 0;
        for (int n = 0; n < N; ++n)
            for (int kk = 0; kk < nv[2]; ++kk)
                for (int jj = 0; jj < nv[1]; ++jj)
                    for (int ii = 0; ii < nv[0]; ++ii) {
                        const long long q = boxes[b].offset + fab_idx(&f, boxes[b].lo[0] + ii, boxes[b].lo[1] + jj, lo2 + kk) + n * f.sn;
                        Un[q] = Un1[q];
                    }
    }
    free(Ub); free(L); free(Ut); free(Un1);
    return err;
}

int ref_level_rk_step(const cmp_params* P, const cmp_geometry* G, int nbox, const cmp_box* boxes, int ng,
                      double* Un, double dt, double* cmax, int integrator)
{
    return rk_step_any(P, G, nbox, boxes, ng, Un, dt, cmax, integrator, NULL);
}

/* The Phase-Field sub-step with run.TimeIntegrator (PhaseField_ForwardEuler :2426-2505,
 * PhaseField_SecondOrderSSPRK :2508-2617, PhaseField_ThirdOrderSSPRK, PhaseField_FourthOrderRK
 * :2756-): the stage combinations above with the Phase-Field right-hand side at a fixed Grad_Q. */
int ref_level_pf_rk_step(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G, int nbox,
                         const cmp_box* boxes, int ng, double* Un, const double* gradq, double pf_dt, int integrator)
{
    rk3_rhs_t R = {PF, gradq};
    return rk_step_any(P, G, nbox, boxes, ng, Un, pf_dt, NULL, integrator, &R);
}

int ref_level_rk3_step(const cmp_params* P, const cmp_geometry* G, int nbox,
                       const cmp_box* boxes, int ng, double* Un, double dt, double* cmax)
{
    return rk3_step(P, G, nbox, boxes, ng, Un, dt, cmax, NULL);
}

/* CalculateInitialcmax (source/Class_Compressible_PhaseField.cpp:4502-4605): FillPatch into a
 * bordered copy, then the face kernel over every box without storing fluxes. */
int ref_level_initial_cmax(const cmp_params* P, const cmp_geometry* G, int nbox,
                           const cmp_box* boxes, int ng, const double* U, double* cmax)
{
    const int dim = P->dim, N = P->nstate;
    long long total = 0;
    for (int b = 0; b < nbox; ++b) total += cmp_fab_size(dim, &boxes[b], ng, N);
    double* Ub = (double*)malloc(sizeof(double) * total);
    double* L = (double*)malloc(sizeof(double) * total);
    memcpy(Ub, U, sizeof(double) * total);
    *cmax = 0.0;
    int err = ref_level_fill_ghosts(P, G, nbox, boxes, ng, Ub);
    int e = ref_level_rhs(P, G, nbox, boxes, ng, Ub, L, cmax);
    if (e && !err) err = e;
    free(Ub); free(L);
    return err;
}

/* EstTimeStep (source/Class_Compressible_PhaseField.cpp:4473-4499): min over dims of cfl*dx/cmax */
double ref_est_dt(const cmp_geometry* G, double cmax, double cfl)
{
    double dt_est = DBL_MAX;
    for (int d = 0; d < G->dim; ++d) {
        double e = cfl * G->dx[d] / cmax;
        dt_est = (e < dt_est) ? e : dt_est;   /* amrex::min */
    }
    return dt_est;
}

/* ComputeDt for one level without stop_time clipping (source/COMPAS.cpp:2250-2302):
 * dt = min(est, 1.1 dt_prev) when the change limiter is on and dt_prev > 0. */
double ref_compute_dt(const cmp_geometry* G, double cmax, double cfl, int limiter, double dt_prev)
{
    double dt_tmp = ref_est_dt(G, cmax, cfl);
    double dt_0 = dt_tmp;
    const double change_max = 1.1;
    if (limiter && dt_prev > 0.0) dt_tmp = smin(dt_tmp, change_max * dt_prev);
    dt_0 = smin(dt_0, 1 * dt_tmp);
    return dt_0;
}

/* Evolve's step loop for a single level (source/Class_Compressible_PhaseField.cpp:202-263):
 * ComputeDt from the previous step's c_max (the initial state's at step 0), then one RK3 step. */
int ref_level_advance_int(const cmp_params* P, const cmp_geometry* G, int nbox, const cmp_box* boxes,
                          int ng, double* U, int nsteps, double cfl, int limiter, double* dt_hist, int integrator)
{
    double cmax = 0.0, dt = 0.0;
    int err = ref_level_initial_cmax(P, G, nbox, boxes, ng, U, &cmax);
    for (int s = 0; s < nsteps && !err; ++s) {
        /* at t = 0 the previous dt is set to the estimate itself (:2258-2260) */
        dt = ref_compute_dt(G, cmax, cfl, limiter, s == 0 ? ref_est_dt(G, cmax, cfl) : dt);
        if (dt_hist) dt_hist[s] = dt;
        cmax = 0.0;   /* c_max reset at stage 0 (:3161-3164) */
        err = ref_level_rk_step(P, G, nbox, boxes, ng, U, dt, &cmax, integrator);
        if (isnan(cmax)) err = CMP_ERR_NAN;
    }
    return err;
}

int ref_level_advance(const cmp_params* P, const cmp_geometry* G, int nbox, const cmp_box* boxes,
                      int ng, double* U, int nsteps, double cfl, int limiter, double* dt_hist)
{
    return ref_level_advance_int(P, G, nbox, boxes, ng, U, nsteps, cfl, limiter, dt_hist, 3);
}

/* ------------------------------------------------------------------------------------------------
 * Phase-field regularization: ACDI mechanism with the explicit RC, UD Phase-Field flux
 * ---------------------------------------------------------------------------------------------- */

/* Mapping_Backward (include/Physics_PhaseField/Mapping_Backward.H:12-29):
 * psi = sqrt(2) eta atanh(2 phi - 1), phi clipped to [1e-12, 1 - 1e-12]. */
double ref_pf_map_backward(double phi, double eta)
{
    const double ep = 1.0e-12;
    double t = phi;
    if (t < ep) t = ep;
    else if (t > 1.0 - ep) t = 1.0 - ep;
    return sqrt(2.0) * eta * atanh(2.0 * t - 1.0);
}

/* Mapping_Forward (include/Physics_PhaseField/Mapping_Forward.H:12-26):
 * deriv 0: phi = 1/2 + 1/2 tanh(psi/(sqrt(2) eta)); deriv 1: dphi/dpsi. */
double ref_pf_map_forward(double psi, double eta, int deriv)
{
    if (deriv == 0) return 0.5 + 0.5 * tanh(psi / (sqrt(2.0) * eta));
    double x = psi / (sqrt(2.0) * eta);
    return 0.5 / (sqrt(2.0) * eta) / cosh(x) / cosh(x);
}

/* FDM_PhaseFieldFlux_ACDI_D (include/Operators_PhaseField/FDM_PhaseFieldFlux_ACDI_D.H:12-45) with
 * PhaseFieldFlux_ACDI (include/Physics_PhaseField/PhaseFieldFlux_ACDI.H:12-26): the face flux
 * J = M (grad phi - F'(psi_f) n_f), psi and n averaged to the face, grad phi = (phi_c - phi_m)/dx. */
double ref_pf_acdi_flux(double pf_m, double pf_c, double n_m, double n_c, double mob, double eta,
                        double dx)
{
    double psi_c = ref_pf_map_backward(pf_c, eta);
    double psi_m = ref_pf_map_backward(pf_m, eta);
    double grad = (pf_c - pf_m) / dx;
    double psi = 0.5 * psi_m + 0.5 * psi_c;
    double nrm = 0.5 * n_m + 0.5 * n_c;
    double F1 = ref_pf_map_forward(psi, eta, 1);
    return mob * (grad - F1 * nrm);
}

/* The explicit Grad_Q of one face (source/Class_Compressible_PhaseField.cpp:2358-2391) with
 * PhaseFieldFlux2GradQPF (include/Physics_PhaseField/PhaseFieldFlux2GradQPF.H:12-28): alpha re-sharpened
 * through the mappings at eta and eta_factor*eta, averaged to the face, then
 * Grad_Q = Jhat/(phi(1-phi) + 1e-3), with Jhat taken as 0 where phi is outside [1e-3, 1 - 1e-3]. */
double ref_pf_gradq_face(double pf_m, double pf_c, double jhat, double eta, double eta_factor)
{
    const double epsilon = 1.0e-3;
    double psi_c = ref_pf_map_backward(pf_c, eta);
    double psi_m = ref_pf_map_backward(pf_m, eta);
    double c = ref_pf_map_forward(psi_c, eta_factor * eta, 0);
    double m = ref_pf_map_forward(psi_m, eta_factor * eta, 0);
    double pf = 0.5 * m + 0.5 * c;
    double flux = (pf < epsilon || pf > 1.0 - epsilon) ? 0.0 : jhat;
    double w = pf * (1.0 - pf);
    return flux / (w + epsilon);
}

/* f_Upwind, f_Downwind (include/Operators_FV/f_Upwind.H:12-27, f_Downwind.H:12-27) */
static inline double f_up(double fL, double fR, double V)
{
    return (V > 0.0) ? fL : ((V < 0.0) ? fR : 0.5 * (fL + fR));
}
static inline double f_down(double fL, double fR, double V)
{
    return (V > 0.0) ? fR : ((V < 0.0) ? fL : 0.5 * (fL + fR));
}

/* EquationOfState_Phase, pressure to alpha_k rho_k e_k
 * (include/Physics_Compressible5Eq/EquationOfState_Phase_Compressible5Eq.H:15-71) */
static double ref_eos_phase_e(double p, double alpha, double alpha_rho, double A, double B,
                              double C, double D, int* err)
{
    double phi = ref_nasg_phi(alpha, alpha_rho, C, err);
    return phi * A * p + phi * B + alpha_rho * D;
}

/* Conservative2ConservativePhase (include/Physics_Compressible5Eq/Conservative2ConservativePhase_Compressible5Eq.H:15-111):
 * the phasic vector [a1r1, a2r2, a1r1 u, a2r2 u, (v), (w), a1r1 E1, a2r2 E2, a1, a2] with the phasic
 * internal energies floored at p_min. */
static void ref_cons2phase(const cmp_params* P, const cmp_pf_params* PF, const double* U,
                           double* Cp, int* err)
{
    const int d = P->dim;
    double W[REF_NMAX];
    ref_cons2prim(P, U, W, err);
    double a1r1 = W[0], a2r2 = W[1], p = W[2 + d], a1 = W[3 + d];
    double a2 = 1.0 - a1;
    double e1 = ref_eos_phase_e(p, a1, a1r1, P->eos_A[0], P->eos_B[0], P->eos_C[0], P->eos_D[0], err);
    double e1m = ref_eos_phase_e(PF->p_min[0], a1, a1r1, P->eos_A[0], P->eos_B[0], P->eos_C[0],
                                 P->eos_D[0], err);
    double e2 = ref_eos_phase_e(p, a2, a2r2, P->eos_A[1], P->eos_B[1], P->eos_C[1], P->eos_D[1], err);
    double e2m = ref_eos_phase_e(PF->p_min[1], a2, a2r2, P->eos_A[1], P->eos_B[1], P->eos_C[1],
                                 P->eos_D[1], err);
    double rho_e = e1 + e2;
    if (e1 < e1m) { e1 = e1m; e2 = rho_e - e1; }
    else if (e2 < e2m) { e2 = e2m; e1 = rho_e - e2; }
    double ke = W[2] * W[2];
    for (int m = 1; m < d; ++m) ke = ke + W[2 + m] * W[2 + m];
    Cp[0] = a1r1;
    Cp[1] = a2r2;
    for (int m = 0; m < d; ++m) { Cp[2 + 2 * m] = a1r1 * W[2 + m]; Cp[3 + 2 * m] = a2r2 * W[2 + m]; }
    Cp[(1 + d) * 2] = e1 + 0.5 * a1r1 * ke;
    Cp[(1 + d) * 2 + 1] = e2 + 0.5 * a2r2 * ke;
    Cp[(2 + d) * 2] = a1;
    Cp[(2 + d) * 2 + 1] = a2;
}

/* FluxPhaseField (include/Physics_Compressible5Eq_PhaseField/FluxPhaseField_Compressible5Eq_PhaseField.H:12-65)
 * with the UD phasic flux FluxPhasePhaseField_UD
 * (include/Physics_Compressible_TwoPhase_PhaseField/FluxPhasePhaseField_UD_Compressible_TwoPhase_PhaseField.H:12-194):
 * Grad_Q1 = g, Grad_Q2 = -g; phase k's quantities are taken downwind of Grad_Qk and the other
 * phase's volume fraction upwind of it. Then either the phasic sum (VectorPhase2Vector,
 * include/Physics_Compressible5Eq/VectorPhase2Vector_Compressible5Eq.H:15-54) or the mixture form
 * FluxPhaseField_Mixture (include/Physics_Compressible5Eq_PhaseField/FluxPhaseField_Mixture_Compressible5Eq_PhaseField.H:12-116). */
void ref_pf_flux(const cmp_params* P, const cmp_pf_params* PF, const double* UL, const double* UR,
                 double g, double* flux, double* cmax, int* err)
{
    const int d = P->dim;
    const int nph = 2 * (3 + d);
    double CL[2 * (3 + 3)], CR[2 * (3 + 3)], Jp[2 * (3 + 3)];
    ref_cons2phase(P, PF, UL, CL, err);
    ref_cons2phase(P, PF, UR, CR, err);
    const int iA = (2 + d) * 2;
    double g1 = g, g2 = -g1;
    double c = 0.0;
    double a2u = f_up(CL[iA + 1], CR[iA + 1], g1);
    c = smax(c, fabs(a2u * g1));
    double a1u = f_up(CL[iA], CR[iA], g2);
    c = smax(c, fabs(a1u * g2));
    /* phase 1 components are the even entries, phase 2 the odd ones */
    for (int q = 0; q < nph; q += 2) {
        Jp[q] = f_down(CL[q], CR[q], g1) * a2u * g1;
        Jp[q + 1] = f_down(CL[q + 1], CR[q + 1], g2) * a1u * g2;
    }
    *cmax = c;
    const int N = P->nstate;
    if (PF->flux_phase2flux == 0) {
        flux[0] = Jp[0];
        flux[1] = Jp[1];
        for (int m = 0; m < d; ++m) flux[2 + m] = Jp[2 + 2 * m] + Jp[3 + 2 * m];
        flux[2 + d] = Jp[(1 + d) * 2] + Jp[(1 + d) * 2 + 1];
        flux[3 + d] = Jp[iA];
        return;
    }
    double J1r1 = Jp[0], J2r2 = Jp[1], J1 = Jp[iA], J2 = Jp[iA + 1];
    double WL[REF_NMAX], WR[REF_NMAX];
    ref_cons2prim(P, UL, WL, err);
    ref_cons2prim(P, UR, WR, err);
    double M = J1r1 + J2r2;
    double uD[3] = {0.0, 0.0, 0.0};
    for (int m = 0; m < d; ++m) uD[m] = f_down(WL[2 + m], WR[2 + m], M);
    double ke = uD[0] * uD[0];
    for (int m = 1; m < d; ++m) ke = ke + uD[m] * uD[m];
    double eK = 0.5 * ke;
    double MeK = M * eK;
    double p1 = f_down(WL[2 + d], WR[2 + d], J1);
    double J1phi = (P->eos_C[0] > 0.0) ? J1 - P->eos_C[0] * J1r1 : J1;
    double J1e1 = P->eos_A[0] * J1phi * p1 + P->eos_B[0] * J1phi + P->eos_D[0] * J1r1;
    double p2 = f_down(WL[2 + d], WR[2 + d], J2);
    double J2phi = (P->eos_C[1] > 0.0) ? J2 - P->eos_C[1] * J2r2 : J2;
    double J2e2 = P->eos_A[1] * J2phi * p2 + P->eos_B[1] * J2phi + P->eos_D[1] * J2r2;
    flux[0] = J1r1;
    flux[1] = J2r2;
    for (int m = 0; m < d; ++m) flux[2 + m] = M * uD[m];
    flux[2 + d] = MeK + J1e1 + J2e2;
    flux[3 + d] = J1;
    (void)N;
}

/* Conservative2ConservativePhase of the N-phase model
 * (include/Physics_Compressible5Eq_NPhase/Conservative2ConservativePhase_Compressible5Eq_NPhase.H):
 * Cp[q*NP + k] for q = alpha_k rho_k, alpha_k rho_k u_d (dim), alpha_k rho_k E_k, alpha_k; a phase
 * whose internal energy falls below its p_min value is clipped to it and the remainder shared among
 * the unclipped phases by volume fraction, repeated until none falls below. */
void ref_np_cons2phase(const cmp_params* P, const cmp_pf_params* PF, const double* U, double* Cp, int* err)
{
    const int d = P->dim, np = P->nphase, ia = np + d + 1;
    double W[REF_NMAX];
    ref_np_cons2prim(P, U, W, err);
    double p = W[np + d];
    double e[CMP_MAX_PHASES], em[CMP_MAX_PHASES];
    int clipped[CMP_MAX_PHASES] = {0, 0, 0};
    int neg = 0;
    double rho_e = 0.0;
    for (int k = 0; k < np; ++k) {
        e[k] = ref_eos_phase_e(p, W[ia + k], W[k], P->eos_A[k], P->eos_B[k], P->eos_C[k], P->eos_D[k], err);
        rho_e += e[k];
        em[k] = ref_eos_phase_e(PF->p_min[k], W[ia + k], W[k], P->eos_A[k], P->eos_B[k], P->eos_C[k],
                                P->eos_D[k], err);
        if (e[k] < em[k]) neg = 1;
    }
    int guard = 0;
    while (neg && guard++ < 2 * CMP_MAX_PHASES) {
        for (int k = 0; k < np; ++k)
            if (e[k] < em[k]) { e[k] = em[k]; clipped[k] = 1; }
        double se = 0.0, sa = 0.0;
        for (int k = 0; k < np; ++k) {
            se += e[k];
            if (!clipped[k]) sa += W[ia + k];
        }
        for (int k = 0; k < np; ++k)
            if (!clipped[k]) e[k] += W[ia + k] / sa * (rho_e - se);
        neg = 0;
        for (int k = 0; k < np; ++k)
            if (e[k] < em[k]) neg = 1;
    }
    double ke = W[np] * W[np];
    for (int m = 1; m < d; ++m) ke = ke + W[np + m] * W[np + m];
    for (int k = 0; k < np; ++k) {
        Cp[k] = W[k];
        for (int m = 0; m < d; ++m) Cp[(1 + m) * np + k] = W[k] * W[np + m];
        Cp[(1 + d) * np + k] = (0.5) * W[k] * ke + e[k];
        Cp[(2 + d) * np + k] = W[ia + k];
    }
}

/* FluxPhaseField of the N-phase model (…FluxPhaseField_Compressible5Eq_NPhase_PhaseField.H:12-79) with
 * the efficient UD phase flux (FluxPhasePhaseField_UD_Efficient,
 * include/Physics_Compressible_NPhase_PhaseField/FluxPhasePhaseField_UD_Efficient_…H:12-232):
 * per phase i, J_q,i = q_i (downwind of Grad_Q_i) (1 - alpha_i)(upwind of Grad_Q_i) Grad_Q_i; then
 * with Sum_J = sum_i J_alpha,i every J_q,i -= q_i (upwind of Sum_J) Sum_J. The phasic sum
 * (VectorPhase2Vector) or the mixture form (FluxPhaseField_Mixture …NPhase_PhaseField.H:12-104). */
void ref_np_pf_flux(const cmp_params* P, const cmp_pf_params* PF, const double* UL, const double* UR,
                    const double* g, double* flux, double* cmax, int* err)
{
    const int d = P->dim, np = P->nphase, nq = 3 + d;
    double CL[(3 + 3) * CMP_MAX_PHASES], CR[(3 + 3) * CMP_MAX_PHASES], Jp[(3 + 3) * CMP_MAX_PHASES];
    ref_np_cons2phase(P, PF, UL, CL, err);
    ref_np_cons2phase(P, PF, UR, CR, err);
    const int qa = (2 + d) * np;   /* alpha_k entries */
    double c = 0.0, SumJ = 0.0;
    for (int i = 0; i < np; ++i) {
        double gi = g[i];
        double aju = f_up(1.0 - CL[qa + i], 1.0 - CR[qa + i], gi);
        double ci = fabs(aju * gi);
        c = smax(c, ci);
        for (int q = 0; q < nq; ++q)
            Jp[q * np + i] = f_down(CL[q * np + i], CR[q * np + i], gi) * aju * gi;
        SumJ += Jp[qa + i];
    }
    c += fabs(SumJ);
    for (int i = 0; i < np; ++i)
        for (int q = 0; q < nq; ++q)
            Jp[q * np + i] += -f_up(CL[q * np + i], CR[q * np + i], SumJ) * SumJ;
    *cmax = c;
    if (PF->flux_phase2flux == 0) {
        for (int k = 0; k < np; ++k) {
            flux[k] = Jp[k];
            flux[np + d + 1 + k] = Jp[qa + k];
        }
        for (int m = 0; m < d; ++m) {
            double t = 0.0;
            for (int k = 0; k < np; ++k) t += Jp[(1 + m) * np + k];
            flux[np + m] = t;
        }
        double te = 0.0;
        for (int k = 0; k < np; ++k) te += Jp[(1 + d) * np + k];
        flux[np + d] = te;
        return;
    }
    double WL[REF_NMAX], WR[REF_NMAX];
    ref_np_cons2prim(P, UL, WL, err);
    ref_np_cons2prim(P, UR, WR, err);
    double M = 0.0;
    for (int k = 0; k < np; ++k) M += Jp[k];
    double uD[3] = {0.0, 0.0, 0.0};
    for (int m = 0; m < d; ++m) uD[m] = f_down(WL[np + m], WR[np + m], M);
    double ke = uD[0] * uD[0];
    for (int m = 1; m < d; ++m) ke = ke + uD[m] * uD[m];
    double MeK = M * (0.5 * ke);
    double Je = 0.0;
    for (int k = 0; k < np; ++k) {
        double Jk = Jp[qa + k];
        double pk = f_down(WL[np + d], WR[np + d], Jk);
        double Jphi = (P->eos_C[k] > 0.0) ? Jk - P->eos_C[k] * Jp[k] : Jk;
        Je += P->eos_A[k] * Jphi * pk + P->eos_B[k] * Jphi + P->eos_D[k] * Jp[k];
    }
    for (int k = 0; k < np; ++k) {
        flux[k] = Jp[k];
        flux[np + d + 1 + k] = Jp[qa + k];
    }
    for (int m = 0; m < d; ++m) flux[np + m] = M * uD[m];
    flux[np + d] = MeK + Je;
}

/* FVM_Conservative2FluxPhaseField_K, ID_Bound == 0
 * (include/Operators_Compressible5Eq_PhaseField/FVM_Conservative2FluxPhaseField_K_Compressible5Eq_PhaseField.H:15-74):
 * the base reconstruction of FVM_Reconstruction2FaceLR_Conservative_K (no THINC on this path),
 * then FluxPhaseField; FluxNC is zero. fhat is the flux before PhaseField_ComputeFaceFluxes
 * stores it negated (source/Class_Compressible_PhaseField.cpp:840). */
int ref_pf_face_flux(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                     const cmp_box* box, int ng, const double* U, int i, int j, int k, int dir,
                     const double* gradq, double* fhat, double* cmax)
{
    const int N = P->nstate;
    int err = 0;
    fab_t f = make_fab(P->dim, box, ng);
    int off[3] = {0, 0, 0};
    off[dir] = 1;
    double Ust[6][REF_NMAX];
    int smin_ = -3, smax_ = 2;
    if (P->id_weno == 0 && P->id_limiter == 0) { smin_ = -1; smax_ = 0; }
    else if (P->id_weno == 0) { smin_ = -2; smax_ = 1; }
    for (int s = smin_; s <= smax_; ++s) {
        long long c = fab_idx(&f, i + s * off[0], j + s * off[1], k + s * off[2]);
        for (int n = 0; n < N; ++n) Ust[s + 3][n] = U[c + n * f.sn];
    }
    double UL[REF_NMAX], UR[REF_NMAX];
    if (P->nphase) {
        /* …FVM_Conservative2FluxPhaseField_K_Compressible5Eq_NPhase_PhaseField.H:15-79 */
        int r = ref_np_recon_lr(P, Ust, G->dx[dir], UL, UR, &err);
        if (r) return r;
        ref_np_pf_flux(P, PF, UL, UR, gradq, fhat, cmax, &err);
        return err;
    }
    int r = ref_recon_lr(P, Ust, P->id_weno, P->id_limiter, G->dx[dir], UL, UR, &err);
    if (r) return r;
    ref_pf_flux(P, PF, UL, UR, gradq[0], fhat, cmax, &err);
    return err;
}

/* One box of PhaseField_ComputeFaceFluxes (:771-852) and PhaseField_FluxDivergence (:859-948):
 * flux = -fhat on every face, dUdt = -sum_d (flux_+ - flux_-)/dx_d (FVM_SurfaceIntegral); the
 * five-equation Phase-Field NC integral is empty
 * (include/Operators_Compressible5Eq_PhaseField/FVM_PhaseField_SurfaceIntegralNC_Compressible5Eq_PhaseField.H).
 * gq[d]: this box's nodal Grad_Q faces of direction d. */
int ref_pf_rhs_box(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                   const cmp_box* box, int ng, const double* U, const double* const* gq,
                   double* dUdt)
{
    const int dim = P->dim, N = P->nstate;
    const int npf = P->nphase ? P->nphase : 1;   /* Grad_Q components per face */
    int nv[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) nv[d] = box->hi[d] - box->lo[d] + 1;
    long long ncell = (long long)nv[0] * nv[1] * nv[2];
    double* F[3] = {NULL, NULL, NULL};
    long long sf[3][3], nf[3];
    int err = 0;
    for (int d = 0; d < dim; ++d) {
        int fe[3] = {nv[0], nv[1], nv[2]};
        fe[d] += 1;
        long long nface = (long long)fe[0] * fe[1] * fe[2];
        sf[d][0] = 1; sf[d][1] = fe[0]; sf[d][2] = (long long)fe[0] * fe[1];
        nf[d] = nface;
        F[d] = (double*)malloc(sizeof(double) * nface * N);
        int e = 0;
        int lo2 = (dim == 3) ? box->lo[2] : 0;
#pragma omp parallel for collapse(2) reduction(max:e) schedule(static)
        for (int kk = 0; kk < fe[2]; ++kk)
            for (int jj = 0; jj < fe[1]; ++jj)
                for (int ii = 0; ii < fe[0]; ++ii) {
                    long long q = ii + (long long)fe[0] * (jj + (long long)fe[1] * kk);
                    double fh[REF_NMAX], c, gv[CMP_MAX_PHASES];
                    for (int m = 0; m < npf; ++m) gv[m] = gq[d][q + m * nface];
                    int r = ref_pf_face_flux(P, PF, G, box, ng, U, box->lo[0] + ii, box->lo[1] + jj,
                                             lo2 + kk, d, gv, fh, &c);
                    if (r > e) e = r;
                    for (int n = 0; n < N; ++n) F[d][q + n * nface] = -fh[n];
                }
        if (e && !err) err = e;
    }
#pragma omp parallel for collapse(2) schedule(static)
    for (int kk = 0; kk < nv[2]; ++kk)
        for (int jj = 0; jj < nv[1]; ++jj)
            for (int ii = 0; ii < nv[0]; ++ii) {
                long long cell = ii + (long long)nv[0] * (jj + (long long)nv[1] * kk);
                int idx[3] = {ii, jj, kk};
                for (int n = 0; n < N; ++n) {
                    double div = 0.0;
                    for (int d = 0; d < dim; ++d) {
                        long long qm = idx[0] * sf[d][0] + idx[1] * sf[d][1] + idx[2] * sf[d][2];
                        long long qp = qm + sf[d][d];
                        double term = (F[d][qp + n * nf[d]] - F[d][qm + n * nf[d]]) / G->dx[d];
                        div = (d == 0) ? term : div + term;
                    }
                    dUdt[cell + n * ncell] = -div;
                }
            }
    for (int d = 0; d < dim; ++d) free(F[d]);
    return err;
}

int cmp_face_layout(int dim, int nbox, const cmp_box* boxes, int ncomp, long long* offsets,
                    long long* total);

/* PhaseField_compute_dUdt_FV over every box; gradq is a face buffer (cmp_face_layout with one
 * component per volume fraction: 1 for the two-phase model, nphase for the N-phase model). */
int ref_level_pf_rhs(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                     int nbox, const cmp_box* boxes, int ng, const double* U, const double* gradq,
                     double* dUdt)
{
    const int dim = P->dim, N = P->nstate;
    long long* foff = (long long*)malloc(sizeof(long long) * (size_t)(nbox * dim + 1));
    cmp_face_layout(dim, nbox, boxes, P->nphase ? P->nphase : 1, foff, NULL);
    int err = 0;
    for (int b = 0; b < nbox; ++b) {
        long long nc = valid_cells(dim, &boxes[b]);
        double* tmp = (double*)malloc(sizeof(double) * nc * N);
        const double* gq[3] = {NULL, NULL, NULL};
        for (int d = 0; d < dim; ++d) gq[d] = gradq + foff[b * dim + d];
        int e = ref_pf_rhs_box(P, PF, G, &boxes[b], ng, U + boxes[b].offset, gq, tmp);
        if (e && !err) err = e;
        fab_t f = make_fab(dim, &boxes[b], ng);
        int nv[3] = {1, 1, 1};
        for (int d = 0; d < dim; ++d) nv[d] = boxes[b].hi[d] - boxes[b].lo[d] + 1;
        int lo2 = (dim == 3) ? boxes[b].lo[2] : 0;
        for (int n = 0; n < N; ++n)
            for (int kk = 0; kk < nv[2]; ++kk)
                for (int jj = 0; jj < nv[1]; ++jj)
                    for (int ii = 0; ii < nv[0]; ++ii) {
                        long long q = fab_idx(&f, boxes[b].lo[0] + ii, boxes[b].lo[1] + jj, lo2 + kk);
                        dUdt[boxes[b].offset + q + n * f.sn] =
                            tmp[ii + (long long)nv[0] * (jj + (long long)nv[1] * kk) + n * nc];
                    }
        free(tmp);
    }
    free(foff);
    return err;
}

/* One Phase-Field RK3 sub-step (PhaseField_ThirdOrderSSPRK, :2620-2752): the stage combinations of
 * ThirdOrderSSPRK with the Phase-Field right-hand side; ghosts by the flow's FillPatch. */
int ref_level_pf_rk3_step(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                          int nbox, const cmp_box* boxes, int ng, double* Un, const double* gradq,
                          double pf_dt)
{
    rk3_rhs_t R = {PF, gradq};
    return rk3_step(P, G, nbox, boxes, ng, Un, pf_dt, NULL, &R);
}

/* MaxVelocity (:3541-3594): max over valid cells of |u|, u = rho u / (a1r1 + a2r2)
 * (Conservative2Velocity, include/Physics_Compressible5Eq/Conservative2Velocity_Compressible5Eq.H:15-40). */
double ref_level_max_velocity(const cmp_params* P, int nbox, const cmp_box* boxes, int ng,
                              const double* U)
{
    const int dim = P->dim;
    double mv = 0.0;
    for (int b = 0; b < nbox; ++b) {
        fab_t f = make_fab(dim, &boxes[b], ng);
        const double* B = U + boxes[b].offset;
        int lo2 = (dim == 3) ? boxes[b].lo[2] : 0, hi2 = (dim == 3) ? boxes[b].hi[2] : 0;
        for (int k = lo2; k <= hi2; ++k)
            for (int j = boxes[b].lo[1]; j <= boxes[b].hi[1]; ++j)
                for (int i = boxes[b].lo[0]; i <= boxes[b].hi[0]; ++i) {
                    long long c = fab_idx(&f, i, j, k);
                    const int nm = P->nphase ? P->nphase : 2;   /* masses, then momenta */
                    double rho = B[c];
                    for (int q = 1; q < nm; ++q) rho += B[c + q * f.sn];
                    double s = 0.0;
                    for (int m = 0; m < dim; ++m) {
                        double u = B[c + (nm + m) * f.sn] / rho;
                        s = (m == 0) ? u * u : s + u * u;
                    }
                    double V = sqrt(s);
                    if (mv < V || isnan(V)) mv = V;
                }
    }
    return mv;
}

/* Boxes of a Phase-Field field: the level's boxes with offsets for ncomp comps and ng ghosts. */
static cmp_box* pf_boxes(int dim, int nbox, const cmp_box* boxes, int ng, int ncomp, long long* total)
{
    cmp_box* bb = (cmp_box*)malloc(sizeof(cmp_box) * (size_t)nbox);
    memcpy(bb, boxes, sizeof(cmp_box) * (size_t)nbox);
    cmp_level_layout(dim, nbox, bb, ng, ncomp, total);
    return bb;
}

/* The CAC-Adv mechanism of one cell (FDM_PhaseFieldMechanism_ACAdv_{K,D},
 * include/Operators_PhaseField/FDM_PhaseFieldMechanism_ACAdv_D.H; FDM_ChemicalPotential_Adv_D.H:12-126;
 * ChemicalPotential_Adv.H; PhaseFieldMechanism_AC.H:12-17; PotentialFunction_DoubleWell.H:22-23):
 *   mu = Lambda (g'(phi)/eta^2 - lap phi - |grad phi| kappa), kappa = -div n (n averaged to faces),
 *   mech = -M mu.
 * f, fm[d], fp[d]: phi at the cell and its neighbours; nc[d], nm[d], np[d]: component d of the
 * normal at the cell and at its neighbours along d. */
double ref_pf_cac_mech(int dim, const double* dx, double f, const double* fm, const double* fp,
                       const double* nc, const double* nm, const double* np, double mob,
                       double lambda, double eta)
{
    double lap = 0.0, divn = 0.0, absg = 0.0;
    for (int d = 0; d < dim; ++d) {
        double gp = (fp[d] - f) / dx[d];
        double gm = (f - fm[d]) / dx[d];
        double t = (gp - gm) / dx[d];
        lap = (d == 0) ? t : lap + t;
        double nfp = 0.5 * nc[d] + 0.5 * np[d];
        double nfm = 0.5 * nm[d] + 0.5 * nc[d];
        double tn = (nfp - nfm) / dx[d];
        divn = (d == 0) ? tn : divn + tn;
        double g = 0.5 * gm + 0.5 * gp;
        absg += g * g;
    }
    absg = sqrt(absg);
    double curv = -divn;
    double g1 = 2.0 * f * (1.0 - f) * (1.0 - 2.0 * f);
    double mu = lambda * ((1.0 / eta / eta) * g1 - lap - absg * curv);
    return -mob * mu;
}

/* The mechanism over one flow step of size dt (PhaseFieldMechanism_Evolve, :597-709, with
 * PhaseField_Mechanism :1007-1152 (CAC) / :1155-1233 (flux based), PhaseField_ACDI :1495-1644 and
 * PhaseField_Normal :1958-2032):
 *   dt_m = 0.9/(2 M Lambda)/(sum_d 1/dx_d^2 + 1/eta^2), the last sub-step cut to end at dt;
 *   per sub-step: ghosts of alpha (PF_bcs), cell normals (FDM_Normal_CenterFromCenter_D,
 *   include/Operators_FD/FDM_Normal_CenterFromCenter_D.H:12-63) of psi = Mapping_Backward(alpha)
 *   for ACDI and of alpha itself for CAC, with their own ghost fill; then
 *     ACDI: J on every face, Jhat += dt_m J, alpha += dt_m div J (FVM_Divergence_D);
 *     CAC: mech and W = alpha (1 - alpha) per cell, their sums over the level, L = -sum mech / sum W
 *          (0 when sum W = 0, PhaseFieldMechanism_CAC.H:12-26), alpha += dt_m (mech + W L);
 *   finally Jhat *= 1/dt (explicit RC) or Div_Jhat = (alpha_new - alpha)/dt (implicit RC).
 * alpha: the level's alpha1 with one ghost cell (offsets of cmp_level_layout for 1 comp, 1 ghost),
 * ghosts filled; jhat: face buffer (1 comp) or NULL; divjhat: alpha layout (valid cells written)
 * or NULL. Returns the number of sub-steps in *nsub. */
int ref_level_pf_mechanism(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                           int nbox, const cmp_box* boxes, const double* alpha, double dt,
                           double mob, double eta, double lambda, double* jhat, double* divjhat,
                           int* nsub)
{
    const int dim = P->dim;
    const int cac = (PF->mechanism == CMP_PF_CAC);
    long long ta, tn, tf;
    cmp_box* ba = pf_boxes(dim, nbox, boxes, 1, 1, &ta);
    cmp_box* bn = pf_boxes(dim, nbox, boxes, 1, dim, &tn);
    long long* foff = (long long*)malloc(sizeof(long long) * (size_t)(nbox * dim + 1));
    cmp_face_layout(dim, nbox, boxes, 1, foff, &tf);
    double* A = (double*)malloc(sizeof(double) * ta);
    double* Nrm = (double*)calloc((size_t)tn, sizeof(double));
    double* J = (double*)malloc(sizeof(double) * tf);
    double* Mk = (double*)calloc((size_t)ta, sizeof(double));
    double* Wk = (double*)calloc((size_t)ta, sizeof(double));
    memcpy(A, alpha, sizeof(double) * ta);
    if (jhat)
        for (long long q = 0; q < tf; ++q) jhat[q] = 0.0;
    double s = 1.0 / G->dx[0] / G->dx[0];
    for (int d = 1; d < dim; ++d) s = s + 1.0 / G->dx[d] / G->dx[d];
    double dt_m = 0.9 / (2.0 * mob * lambda) / (s + 1.0 / eta / eta);
    double t = 0.0, T = dt;
    int n = 0, err = 0;
    while (t < T) {
        if (t + dt_m > T) dt_m = T - t;
        int e = ref_level_fill_pf(dim, G, nbox, ba, 1, 1, A);
        if (e && !err) err = e;
        /* normals on valid cells from psi of the cell and its neighbours */
        for (int b = 0; b < nbox; ++b) {
            fab_t f = make_fab(dim, &ba[b], 1);
            const double* Ab = A + ba[b].offset;
            double* Nb = Nrm + bn[b].offset;
            int lo2 = (dim == 3) ? ba[b].lo[2] : 0, hi2 = (dim == 3) ? ba[b].hi[2] : 0;
#pragma omp parallel for collapse(2) schedule(static)
            for (int k = lo2; k <= hi2; ++k)
                for (int j = ba[b].lo[1]; j <= ba[b].hi[1]; ++j)
                    for (int i = ba[b].lo[0]; i <= ba[b].hi[0]; ++i) {
                        int c3[3] = {i, j, k};
                        long long c = fab_idx(&f, i, j, k);
                        double fc = cac ? Ab[c] : ref_pf_map_backward(Ab[c], eta);
                        double g[3], sum = 0.0;
                        for (int d = 0; d < dim; ++d) {
                            int m3[3] = {c3[0], c3[1], c3[2]}, p3[3] = {c3[0], c3[1], c3[2]};
                            m3[d] -= 1; p3[d] += 1;
                            double am = Ab[fab_idx(&f, m3[0], m3[1], m3[2])];
                            double ap = Ab[fab_idx(&f, p3[0], p3[1], p3[2])];
                            double fm = cac ? am : ref_pf_map_backward(am, eta);
                            double fp = cac ? ap : ref_pf_map_backward(ap, eta);
                            double gm = (fc - fm) / G->dx[d];
                            double gp = (fp - fc) / G->dx[d];
                            g[d] = 0.5 * gm + 0.5 * gp;
                            sum += g[d] * g[d];
                        }
                        for (int d = 0; d < dim; ++d)
                            Nb[c + d * f.sn] = (sum == 0.0) ? 0.0 : g[d] / sqrt(sum);
                    }
        }
        e = ref_level_fill_pf(dim, G, nbox, bn, 1, dim, Nrm);
        if (e && !err) err = e;
        if (cac) {
            /* mech and W per valid cell, their level sums, then alpha += dt_m (mech + W L) */
            double sm = 0.0, sw = 0.0;
            for (int b = 0; b < nbox; ++b) {
                fab_t f = make_fab(dim, &ba[b], 1);
                const double* Ab = A + ba[b].offset;
                const double* Nb = Nrm + bn[b].offset;
                int lo2 = (dim == 3) ? ba[b].lo[2] : 0, hi2 = (dim == 3) ? ba[b].hi[2] : 0;
                for (int k = lo2; k <= hi2; ++k)
                    for (int j = ba[b].lo[1]; j <= ba[b].hi[1]; ++j)
                        for (int i = ba[b].lo[0]; i <= ba[b].hi[0]; ++i) {
                            int c3[3] = {i, j, k};
                            long long c = fab_idx(&f, i, j, k);
                            double fm[3], fp[3], nc[3], nm[3], np[3];
                            for (int d = 0; d < dim; ++d) {
                                int m3[3] = {c3[0], c3[1], c3[2]}, p3[3] = {c3[0], c3[1], c3[2]};
                                m3[d] -= 1; p3[d] += 1;
                                long long m = fab_idx(&f, m3[0], m3[1], m3[2]);
                                long long p = fab_idx(&f, p3[0], p3[1], p3[2]);
                                fm[d] = Ab[m];
                                fp[d] = Ab[p];
                                nc[d] = Nb[c + d * f.sn];
                                nm[d] = Nb[m + d * f.sn];
                                np[d] = Nb[p + d * f.sn];
                            }
                            double mk = ref_pf_cac_mech(dim, G->dx, Ab[c], fm, fp, nc, nm, np, mob, lambda, eta);
                            double wk = Ab[c] * (1.0 - Ab[c]);
                            Mk[ba[b].offset + c] = mk;
                            Wk[ba[b].offset + c] = wk;
                            sm += mk;
                            sw += wk;
                        }
            }
            double L = (sw == 0.0) ? 0.0 : -sm / sw;
            for (int b = 0; b < nbox; ++b) {
                fab_t f = make_fab(dim, &ba[b], 1);
                double* Ab = A + ba[b].offset;
                int lo2 = (dim == 3) ? ba[b].lo[2] : 0, hi2 = (dim == 3) ? ba[b].hi[2] : 0;
                for (int k = lo2; k <= hi2; ++k)
                    for (int j = ba[b].lo[1]; j <= ba[b].hi[1]; ++j)
                        for (int i = ba[b].lo[0]; i <= ba[b].hi[0]; ++i) {
                            long long c = fab_idx(&f, i, j, k);
                            double divj = Mk[ba[b].offset + c] + Wk[ba[b].offset + c] * L;
                            Ab[c] = 1.0 * Ab[c] + dt_m * divj;
                        }
            }
        } else {
        /* J on every face of every box, Jhat += dt_m J */
        for (int b = 0; b < nbox; ++b) {
            fab_t f = make_fab(dim, &ba[b], 1);
            const double* Ab = A + ba[b].offset;
            const double* Nb = Nrm + bn[b].offset;
            int nv[3] = {1, 1, 1};
            for (int d = 0; d < dim; ++d) nv[d] = ba[b].hi[d] - ba[b].lo[d] + 1;
            int lo2 = (dim == 3) ? ba[b].lo[2] : 0;
            for (int d = 0; d < dim; ++d) {
                int fe[3] = {nv[0], nv[1], nv[2]};
                fe[d] += 1;
                double* Jd = J + foff[b * dim + d];
                double* Hd = jhat ? jhat + foff[b * dim + d] : NULL;
#pragma omp parallel for collapse(2) schedule(static)
                for (int kk = 0; kk < fe[2]; ++kk)
                    for (int jj = 0; jj < fe[1]; ++jj)
                        for (int ii = 0; ii < fe[0]; ++ii) {
                            int c3[3] = {ba[b].lo[0] + ii, ba[b].lo[1] + jj, lo2 + kk};
                            int m3[3] = {c3[0], c3[1], c3[2]};
                            m3[d] -= 1;
                            long long c = fab_idx(&f, c3[0], c3[1], c3[2]);
                            long long m = fab_idx(&f, m3[0], m3[1], m3[2]);
                            double mob_face = 0.5 * mob + 0.5 * mob;
                            double Jf = ref_pf_acdi_flux(Ab[m], Ab[c], Nb[m + d * f.sn], Nb[c + d * f.sn],
                                                         mob_face, eta, G->dx[d]);
                            long long q = ii + (long long)fe[0] * (jj + (long long)fe[1] * kk);
                            Jd[q] = Jf;
                            if (Hd) Hd[q] += dt_m * Jf;
                        }
            }
        }
        /* alpha += dt_m div J on valid cells */
        for (int b = 0; b < nbox; ++b) {
            fab_t f = make_fab(dim, &ba[b], 1);
            double* Ab = A + ba[b].offset;
            int nv[3] = {1, 1, 1};
            for (int d = 0; d < dim; ++d) nv[d] = ba[b].hi[d] - ba[b].lo[d] + 1;
            int lo2 = (dim == 3) ? ba[b].lo[2] : 0;
            long long sf[3][3];
            for (int d = 0; d < dim; ++d) {
                int fe[3] = {nv[0], nv[1], nv[2]};
                fe[d] += 1;
                sf[d][0] = 1; sf[d][1] = fe[0]; sf[d][2] = (long long)fe[0] * fe[1];
            }
#pragma omp parallel for collapse(2) schedule(static)
            for (int kk = 0; kk < nv[2]; ++kk)
                for (int jj = 0; jj < nv[1]; ++jj)
                    for (int ii = 0; ii < nv[0]; ++ii) {
                        int idx[3] = {ii, jj, kk};
                        double div = 0.0;
                        for (int d = 0; d < dim; ++d) {
                            const double* Jd = J + foff[b * dim + d];
                            long long qm = idx[0] * sf[d][0] + idx[1] * sf[d][1] + idx[2] * sf[d][2];
                            long long qp = qm + sf[d][d];
                            double term = (Jd[qp] - Jd[qm]) / G->dx[d];
                            div = (d == 0) ? term : div + term;
                        }
                        long long c = fab_idx(&f, ba[b].lo[0] + ii, ba[b].lo[1] + jj, lo2 + kk);
                        Ab[c] = 1.0 * Ab[c] + dt_m * div;
                    }
        }
        }   /* flux-based mechanism */
        t += dt_m;
        ++n;
    }
    double sc = 1.0 / dt;
    if (jhat)
        for (long long q = 0; q < tf; ++q) jhat[q] *= sc;
    if (divjhat) {
        /* LinComb(Div_Jhat, 1/dt, PF_new, -1/dt, PF) (:703) */
        const double a = 1.0 / dt, bq = -1.0 / dt;
        for (int b = 0; b < nbox; ++b) {
            fab_t f = make_fab(dim, &ba[b], 1);
            int lo2 = (dim == 3) ? ba[b].lo[2] : 0, hi2 = (dim == 3) ? ba[b].hi[2] : 0;
            for (int k = lo2; k <= hi2; ++k)
                for (int j = ba[b].lo[1]; j <= ba[b].hi[1]; ++j)
                    for (int i = ba[b].lo[0]; i <= ba[b].hi[0]; ++i) {
                        long long c = ba[b].offset + fab_idx(&f, i, j, k);
                        divjhat[c] = a * A[c] + bq * alpha[c];
                    }
        }
    }
    if (nsub) *nsub = n;
    free(A); free(Nrm); free(J); free(Mk); free(Wk); free(foff); free(ba); free(bn);
    return err;
}

/* The implicit Grad_Q (PhaseField_GradQPF with Div_J, :2133-2310): b = W(phi_f) + 1e-3 on every
 * face, phi_f the face mean of the alpha re-sharpened through the mappings at eta and
 * eta_factor*eta; Q solves div(b grad Q) = Div_Jhat (MLMG's ABecLaplacian with a = 0, b = -1;
 * periodic sides periodic, the others homogeneous Neumann, Q = 0 initially); Grad_Q = grad Q on the
 * faces (getGradSolution), zero on Neumann domain faces.
 * The reference's solver is AMReX 26.01's geometric multigrid (MLMG, not vendored in the
 * reference) stopped at a relative residual of 1e-10 (LinearSystem_Parameter.H:31). Any converged
 * solve of the same discrete system gives the same Grad_Q to within that tolerance; this restates
 * the system and solves it by Jacobi-preconditioned conjugate gradients on the zero-mean subspace
 * down to the rounding floor (r.r <= 1e-28 of the initial r.r, or 400 iterations without a new
 * minimum of r.r), which the GPU solve also reaches.
 * alpha: 1 ghost, filled; divjhat: alpha layout; gradq: face buffer. *iters: CG iterations. */
int ref_level_pf_gradq_implicit(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                                int nbox, const cmp_box* boxes, const double* alpha,
                                const double* divjhat, double eta, double* gradq, int* iters)
{
    const int dim = P->dim;
    long long ta, tf;
    cmp_box* ba = pf_boxes(dim, nbox, boxes, 1, 1, &ta);
    long long* foff = (long long*)malloc(sizeof(long long) * (size_t)(nbox * dim + 1));
    cmp_face_layout(dim, nbox, boxes, 1, foff, &tf);
    double* Bf = (double*)malloc(sizeof(double) * tf);
    /* b on the faces */
    for (int b = 0; b < nbox; ++b) {
        fab_t f = make_fab(dim, &ba[b], 1);
        const double* Ab = alpha + ba[b].offset;
        int nv[3] = {1, 1, 1};
        for (int d = 0; d < dim; ++d) nv[d] = ba[b].hi[d] - ba[b].lo[d] + 1;
        int lo2 = (dim == 3) ? ba[b].lo[2] : 0;
        for (int d = 0; d < dim; ++d) {
            int fe[3] = {nv[0], nv[1], nv[2]};
            fe[d] += 1;
            for (int kk = 0; kk < fe[2]; ++kk)
                for (int jj = 0; jj < fe[1]; ++jj)
                    for (int ii = 0; ii < fe[0]; ++ii) {
                        int c3[3] = {ba[b].lo[0] + ii, ba[b].lo[1] + jj, lo2 + kk};
                        int m3[3] = {c3[0], c3[1], c3[2]};
                        m3[d] -= 1;
                        double c = ref_pf_map_forward(ref_pf_map_backward(Ab[fab_idx(&f, c3[0], c3[1], c3[2])], eta),
                                                      PF->eta_factor * eta, 0);
                        double m = ref_pf_map_forward(ref_pf_map_backward(Ab[fab_idx(&f, m3[0], m3[1], m3[2])], eta),
                                                      PF->eta_factor * eta, 0);
                        double pf = 0.5 * m + 0.5 * c;
                        long long q = ii + (long long)fe[0] * (jj + (long long)fe[1] * kk);
                        Bf[foff[b * dim + d] + q] = pf * (1.0 - pf) + 1.0e-3;
                    }
        }
    }
    double* X = (double*)calloc((size_t)ta, sizeof(double));
    double* R = (double*)calloc((size_t)ta, sizeof(double));
    double* Pv = (double*)calloc((size_t)ta, sizeof(double));
    double* AP = (double*)calloc((size_t)ta, sizeof(double));
    double* ID = (double*)calloc((size_t)ta, sizeof(double));
    /* Jacobi preconditioner: 1 / sum of b / dx^2 over the faces that carry a flux */
    for (int b = 0; b < nbox; ++b) {
        fab_t f = make_fab(dim, &ba[b], 1);
        int nv[3] = {1, 1, 1};
        for (int d = 0; d < dim; ++d) nv[d] = ba[b].hi[d] - ba[b].lo[d] + 1;
        int lo2 = (dim == 3) ? ba[b].lo[2] : 0;
        for (int kk = 0; kk < nv[2]; ++kk)
            for (int jj = 0; jj < nv[1]; ++jj)
                for (int ii = 0; ii < nv[0]; ++ii) {
                    int c3[3] = {ba[b].lo[0] + ii, ba[b].lo[1] + jj, lo2 + kk};
                    int idx[3] = {ii, jj, kk};
                    double sd = 0.0;
                    for (int d = 0; d < dim; ++d) {
                        int fe[3] = {nv[0], nv[1], nv[2]};
                        fe[d] += 1;
                        long long qm = idx[0] + (long long)fe[0] * (idx[1] + (long long)fe[1] * idx[2]);
                        long long qp = qm + (d == 0 ? 1 : (d == 1 ? fe[0] : (long long)fe[0] * fe[1]));
                        int lo_dom = (c3[d] == G->domain_lo[d]) && !G->is_periodic[d];
                        int hi_dom = (c3[d] == G->domain_hi[d]) && !G->is_periodic[d];
                        double h2 = G->dx[d] * G->dx[d];
                        if (!hi_dom) sd += Bf[foff[b * dim + d] + qp] / h2;
                        if (!lo_dom) sd += Bf[foff[b * dim + d] + qm] / h2;
                    }
                    ID[ba[b].offset + fab_idx(&f, c3[0], c3[1], c3[2])] = (sd > 0.0) ? 1.0 / sd : 0.0;
                }
    }
    long long ncell = 0;
    double mean = 0.0;
    for (int b = 0; b < nbox; ++b) {
        fab_t f = make_fab(dim, &ba[b], 1);
        int lo2 = (dim == 3) ? ba[b].lo[2] : 0, hi2 = (dim == 3) ? ba[b].hi[2] : 0;
        for (int k = lo2; k <= hi2; ++k)
            for (int j = ba[b].lo[1]; j <= ba[b].hi[1]; ++j)
                for (int i = ba[b].lo[0]; i <= ba[b].hi[0]; ++i) {
                    mean += divjhat[ba[b].offset + fab_idx(&f, i, j, k)];
                    ++ncell;
                }
    }
    mean /= (double)ncell;
    /* the operator is -div(b grad .): symmetric positive semi-definite; r = -(rhs - mean); Jacobi-
     * preconditioned conjugate gradients */
    double rr = 0.0, b2 = 0.0, rz = 0.0;
    for (int b = 0; b < nbox; ++b) {
        fab_t f = make_fab(dim, &ba[b], 1);
        int lo2 = (dim == 3) ? ba[b].lo[2] : 0, hi2 = (dim == 3) ? ba[b].hi[2] : 0;
        for (int k = lo2; k <= hi2; ++k)
            for (int j = ba[b].lo[1]; j <= ba[b].hi[1]; ++j)
                for (int i = ba[b].lo[0]; i <= ba[b].hi[0]; ++i) {
                    long long c = ba[b].offset + fab_idx(&f, i, j, k);
                    R[c] = -(divjhat[c] - mean);
                    Pv[c] = ID[c] * R[c];
                    rr += R[c] * R[c];
                    rz += R[c] * Pv[c];
                }
    }
    b2 = rr;
    int it = 0, err = 0;
    const int maxit = 100000;
    double best = rr;
    int stall = 0;
    while (rr > 1.0e-28 * b2 && it < maxit && b2 > 0.0) {
        int e = ref_level_fill_pf(dim, G, nbox, ba, 1, 1, Pv);
        if (e && !err) err = e;
        double pAp = 0.0;
        for (int b = 0; b < nbox; ++b) {
            fab_t f = make_fab(dim, &ba[b], 1);
            int nv[3] = {1, 1, 1};
            for (int d = 0; d < dim; ++d) nv[d] = ba[b].hi[d] - ba[b].lo[d] + 1;
            int lo2 = (dim == 3) ? ba[b].lo[2] : 0;
            for (int kk = 0; kk < nv[2]; ++kk)
                for (int jj = 0; jj < nv[1]; ++jj)
                    for (int ii = 0; ii < nv[0]; ++ii) {
                        int c3[3] = {ba[b].lo[0] + ii, ba[b].lo[1] + jj, lo2 + kk};
                        long long c = ba[b].offset + fab_idx(&f, c3[0], c3[1], c3[2]);
                        int idx[3] = {ii, jj, kk};
                        double acc = 0.0;
                        for (int d = 0; d < dim; ++d) {
                            int fe[3] = {nv[0], nv[1], nv[2]};
                            fe[d] += 1;
                            long long qm = idx[0] + (long long)fe[0] * (idx[1] + (long long)fe[1] * idx[2]);
                            long long qp = qm + (d == 0 ? 1 : (d == 1 ? fe[0] : (long long)fe[0] * fe[1]));
                            int m3[3] = {c3[0], c3[1], c3[2]}, p3[3] = {c3[0], c3[1], c3[2]};
                            m3[d] -= 1; p3[d] += 1;
                            double pm = Pv[ba[b].offset + fab_idx(&f, m3[0], m3[1], m3[2])];
                            double pp = Pv[ba[b].offset + fab_idx(&f, p3[0], p3[1], p3[2])];
                            int lo_dom = (c3[d] == G->domain_lo[d]) && !G->is_periodic[d];
                            int hi_dom = (c3[d] == G->domain_hi[d]) && !G->is_periodic[d];
                            double fp_ = hi_dom ? 0.0 : Bf[foff[b * dim + d] + qp] * (pp - Pv[c]) / G->dx[d];
                            double fm_ = lo_dom ? 0.0 : Bf[foff[b * dim + d] + qm] * (Pv[c] - pm) / G->dx[d];
                            double t = (fp_ - fm_) / G->dx[d];
                            acc = (d == 0) ? t : acc + t;
                        }
                        AP[c] = -acc;
                        pAp += Pv[c] * AP[c];
                    }
        }
        if (!(pAp > 0.0)) break;
        double al = rz / pAp, rr2 = 0.0, rz2 = 0.0;
        for (int b = 0; b < nbox; ++b) {
            fab_t f = make_fab(dim, &ba[b], 1);
            int lo2 = (dim == 3) ? ba[b].lo[2] : 0, hi2 = (dim == 3) ? ba[b].hi[2] : 0;
            for (int k = lo2; k <= hi2; ++k)
                for (int j = ba[b].lo[1]; j <= ba[b].hi[1]; ++j)
                    for (int i = ba[b].lo[0]; i <= ba[b].hi[0]; ++i) {
                        long long c = ba[b].offset + fab_idx(&f, i, j, k);
                        X[c] += al * Pv[c];
                        R[c] -= al * AP[c];
                        rr2 += R[c] * R[c];
                        rz2 += R[c] * ID[c] * R[c];
                    }
        }
        double be = rz2 / rz;
        rr = rr2;
        rz = rz2;
        for (int b = 0; b < nbox; ++b) {
            fab_t f = make_fab(dim, &ba[b], 1);
            int lo2 = (dim == 3) ? ba[b].lo[2] : 0, hi2 = (dim == 3) ? ba[b].hi[2] : 0;
            for (int k = lo2; k <= hi2; ++k)
                for (int j = ba[b].lo[1]; j <= ba[b].hi[1]; ++j)
                    for (int i = ba[b].lo[0]; i <= ba[b].hi[0]; ++i) {
                        long long c = ba[b].offset + fab_idx(&f, i, j, k);
                        Pv[c] = ID[c] * R[c] + be * Pv[c];
                    }
        }
        ++it;
        /* the rounding floor: no new minimum of r.r over a window long enough for CG's plateaus */
        if (rr < best) { best = rr; stall = 0; }
        else if (++stall > 400) break;
    }
    /* Grad_Q on every face */
    int e = ref_level_fill_pf(dim, G, nbox, ba, 1, 1, X);
    if (e && !err) err = e;
    for (int b = 0; b < nbox; ++b) {
        fab_t f = make_fab(dim, &ba[b], 1);
        int nv[3] = {1, 1, 1};
        for (int d = 0; d < dim; ++d) nv[d] = ba[b].hi[d] - ba[b].lo[d] + 1;
        int lo2 = (dim == 3) ? ba[b].lo[2] : 0;
        for (int d = 0; d < dim; ++d) {
            int fe[3] = {nv[0], nv[1], nv[2]};
            fe[d] += 1;
            for (int kk = 0; kk < fe[2]; ++kk)
                for (int jj = 0; jj < fe[1]; ++jj)
                    for (int ii = 0; ii < fe[0]; ++ii) {
                        int c3[3] = {ba[b].lo[0] + ii, ba[b].lo[1] + jj, lo2 + kk};
                        int m3[3] = {c3[0], c3[1], c3[2]};
                        m3[d] -= 1;
                        long long q = ii + (long long)fe[0] * (jj + (long long)fe[1] * kk);
                        int on_dom = !G->is_periodic[d] && (c3[d] == G->domain_lo[d] || c3[d] == G->domain_hi[d] + 1);
                        double xc = X[ba[b].offset + fab_idx(&f, c3[0], c3[1], c3[2])];
                        double xm = X[ba[b].offset + fab_idx(&f, m3[0], m3[1], m3[2])];
                        gradq[foff[b * dim + d] + q] = on_dom ? 0.0 : (xc - xm) / G->dx[d];
                    }
        }
    }
    if (iters) *iters = it;
    free(Bf); free(X); free(R); free(Pv); free(AP); free(ID); free(foff); free(ba);
    return err;
}

/* Grad_Q of every face from alpha (1 ghost, filled) and Jhat (explicit RC, :2315-2403). */
int ref_level_pf_gradq(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                       int nbox, const cmp_box* boxes, const double* alpha, const double* jhat,
                       double eta, double* gradq)
{
    const int dim = P->dim;
    long long ta;
    cmp_box* ba = pf_boxes(dim, nbox, boxes, 1, 1, &ta);
    long long* foff = (long long*)malloc(sizeof(long long) * (size_t)(nbox * dim + 1));
    cmp_face_layout(dim, nbox, boxes, 1, foff, NULL);
    (void)G;
    for (int b = 0; b < nbox; ++b) {
        fab_t f = make_fab(dim, &ba[b], 1);
        const double* Ab = alpha + ba[b].offset;
        int nv[3] = {1, 1, 1};
        for (int d = 0; d < dim; ++d) nv[d] = ba[b].hi[d] - ba[b].lo[d] + 1;
        int lo2 = (dim == 3) ? ba[b].lo[2] : 0;
        for (int d = 0; d < dim; ++d) {
            int fe[3] = {nv[0], nv[1], nv[2]};
            fe[d] += 1;
            const double* Hd = jhat + foff[b * dim + d];
            double* Qd = gradq + foff[b * dim + d];
            for (int kk = 0; kk < fe[2]; ++kk)
                for (int jj = 0; jj < fe[1]; ++jj)
                    for (int ii = 0; ii < fe[0]; ++ii) {
                        int c3[3] = {ba[b].lo[0] + ii, ba[b].lo[1] + jj, lo2 + kk};
                        int m3[3] = {c3[0], c3[1], c3[2]};
                        m3[d] -= 1;
                        long long q = ii + (long long)fe[0] * (jj + (long long)fe[1] * kk);
                        Qd[q] = ref_pf_gradq_face(Ab[fab_idx(&f, m3[0], m3[1], m3[2])],
                                                  Ab[fab_idx(&f, c3[0], c3[1], c3[2])], Hd[q], eta,
                                                  PF->eta_factor);
                    }
        }
    }
    free(foff); free(ba);
    return CMP_OK;
}

/* Resolved Lambda, eta and mobility of a Phase-Field step (PhaseField_Evolve, :427-465):
 * Lambda <= 0 -> 1; eta <= 0 -> eta_multiplier max_d dx_d; mobility 1.0 (reset by Evolve each step,
 * :236) times 0.5 eta max|u|. */
void ref_pf_coefficients(const cmp_pf_params* PF, const cmp_geometry* G, double max_v,
                         double* lambda, double* eta, double* mob)
{
    double L = PF->lambda;
    if (L <= 0.0) L = 1.0;
    double e = PF->eta;
    if (e <= 0.0) {
        double m = G->dx[0];
        for (int d = 1; d < G->dim; ++d) m = (m < G->dx[d]) ? G->dx[d] : m;
        e = PF->eta_multiplier * m;
    }
    double M = 1.0;
    M *= 0.5 * e * max_v;
    *lambda = L; *eta = e; *mob = M;
}

/* Grad_Q for a flow step of size dt: MaxVelocity, alpha1 with PF_bcs ghosts, the mechanism,
 * the explicit Grad_Q. gradq: face buffer; info[4] = max|u|, mobility, eta, mechanism sub-steps. */
int ref_level_pf_gradq_step(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                            int nbox, const cmp_box* boxes, int ng, const double* U, double dt,
                            double* gradq, double* gradq_max, double* info)
{
    const int dim = P->dim;
    /* a mechanism that is not flux based has no Jhat: PhaseField_Parameter.H:138-140 aborts */
    if (PF->mechanism == CMP_PF_CAC && PF->explicit_rc) return CMP_ERR_UNSUPPORTED;
    double mv = ref_level_max_velocity(P, nbox, boxes, ng, U);
    double lambda, eta, mob;
    ref_pf_coefficients(PF, G, mv, &lambda, &eta, &mob);
    /* one volume fraction (two-phase model) or one per phase (N-phase model); PhaseField_Mechanism
     * and PhaseField_ACDI loop over the phases (:1036-1150, :1495-1644), each phase's Grad_Q is the
     * phase's component of the face buffer */
    const int npf = P->nphase ? P->nphase : 1;
    const int ia0 = P->nphase ? P->nphase + dim + 1 : 3 + dim;
    long long ta, tf, tfn;
    cmp_box* ba = pf_boxes(dim, nbox, boxes, 1, 1, &ta);
    double* A = (double*)calloc((size_t)ta, sizeof(double));
    long long* foff = (long long*)malloc(sizeof(long long) * (size_t)(nbox * dim + 1));
    long long* foffn = (long long*)malloc(sizeof(long long) * (size_t)(nbox * dim + 1));
    cmp_face_layout(dim, nbox, boxes, 1, foff, &tf);
    cmp_face_layout(dim, nbox, boxes, npf, foffn, &tfn);
    double* H = (double*)malloc(sizeof(double) * tf);
    double* gq1 = (double*)malloc(sizeof(double) * tf);
    int err = 0, nsub = 0, e;
    for (int ph = 0; ph < npf; ++ph) {
        /* alpha of the valid cells, then its ghosts */
        for (int b = 0; b < nbox; ++b) {
            fab_t fu = make_fab(dim, &boxes[b], ng), fa = make_fab(dim, &ba[b], 1);
            int lo2 = (dim == 3) ? boxes[b].lo[2] : 0, hi2 = (dim == 3) ? boxes[b].hi[2] : 0;
            for (int k = lo2; k <= hi2; ++k)
                for (int j = boxes[b].lo[1]; j <= boxes[b].hi[1]; ++j)
                    for (int i = boxes[b].lo[0]; i <= boxes[b].hi[0]; ++i)
                        A[ba[b].offset + fab_idx(&fa, i, j, k)] =
                            U[boxes[b].offset + fab_idx(&fu, i, j, k) + (ia0 + ph) * fu.sn];
        }
        e = ref_level_fill_pf(dim, G, nbox, ba, 1, 1, A);
        if (e && !err) err = e;
        if (PF->explicit_rc) {
            e = ref_level_pf_mechanism(P, PF, G, nbox, boxes, A, dt, mob, eta, lambda, H, NULL, &nsub);
            if (e && !err) err = e;
            ref_level_pf_gradq(P, PF, G, nbox, boxes, A, H, eta, gq1);
        } else {
            double* DJ = (double*)calloc((size_t)ta, sizeof(double));
            e = ref_level_pf_mechanism(P, PF, G, nbox, boxes, A, dt, mob, eta, lambda, NULL, DJ, &nsub);
            if (e && !err) err = e;
            e = ref_level_pf_gradq_implicit(P, PF, G, nbox, boxes, A, DJ, eta, gq1, NULL);
            if (e && !err) err = e;
            free(DJ);
        }
        /* into component ph of the face buffer */
        for (int b = 0; b < nbox; ++b)
            for (int d = 0; d < dim; ++d) {
                long long nf = 1;
                for (int m = 0; m < dim; ++m) nf *= boxes[b].hi[m] - boxes[b].lo[m] + 1 + (m == d ? 1 : 0);
                for (long long q = 0; q < nf; ++q)
                    gradq[foffn[b * dim + d] + ph * nf + q] = gq1[foff[b * dim + d] + q];
            }
    }
    double gm = 0.0;
    for (long long q = 0; q < tfn; ++q) {
        double a = fabs(gradq[q]);
        if (gm < a || isnan(a)) gm = a;
    }
    if (gradq_max) *gradq_max = gm;
    if (info) { info[0] = mv; info[1] = mob; info[2] = eta; info[3] = nsub; }
    free(A); free(H); free(gq1); free(foff); free(foffn); free(ba);
    return err;
}

/* PhaseField_ComputeDt (:2037-2098) for one level at Phase-Field time pf_time:
 * tmp = cfl min_d dx_d/(max|Grad_Q| + 1e-14); with the change limiter tmp = min(tmp, 1.1 prev)
 * (prev = tmp itself at pf_time 0); the step is cut so that it ends at dt when it would come
 * within 1e-3 of its own size of dt. *prev receives the step. */
double ref_pf_compute_dt(const cmp_pf_params* PF, const cmp_geometry* G, double gradq_max,
                         double pf_time, double dt, double* prev)
{
    double dx = G->dx[0];
    for (int d = 1; d < G->dim; ++d) dx = (G->dx[d] < dx) ? G->dx[d] : dx;
    double tmp = PF->cfl * dx / (gradq_max + 1.0e-14);
    if (pf_time == 0.0) *prev = tmp;
    double dt0 = tmp;
    if (PF->change_limiter && *prev > 0.0) tmp = smin(tmp, 1.1 * *prev);
    dt0 = smin(dt0, 1 * tmp);
    const double eps = (1.0e-3) * dt0;
    if (pf_time + dt0 > dt - eps) dt0 = dt - pf_time;
    *prev = dt0;
    return dt0;
}

/* PhaseField_Evolve (:425-594) after a flow step of size dt, in place on U.
 * info[6] (or NULL): max|u|, mobility, eta, mechanism sub-steps, max|Grad_Q|, Phase-Field sub-steps. */
static int pf_evolve_int(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                         int nbox, const cmp_box* boxes, int ng, double* U, double dt, double* info, int integrator)
{
    const int dim = P->dim;
    long long tf;
    long long* foff = (long long*)malloc(sizeof(long long) * (size_t)(nbox * dim + 1));
    cmp_face_layout(dim, nbox, boxes, P->nphase ? P->nphase : 1, foff, &tf);
    double* Q = (double*)malloc(sizeof(double) * tf);
    double gm = 0.0, inf4[4];
    int err = ref_level_pf_gradq_step(P, PF, G, nbox, boxes, ng, U, dt, Q, &gm, inf4);
    double pf_time = 0.0, prev = 0.0;
    int n = 0;
    for (int step = 0; step < PF->max_step && pf_time < dt && !err; ++step) {
        double pdt = ref_pf_compute_dt(PF, G, gm, pf_time, dt, &prev);
        err = ref_level_pf_rk_step(P, PF, G, nbox, boxes, ng, U, Q, pdt, integrator);
        pf_time += pdt;
        ++n;
        if (pf_time >= dt - 1.0e-6 * dt) break;
    }
    if (info) {
        for (int q = 0; q < 4; ++q) info[q] = inf4[q];
        info[4] = gm;
        info[5] = n;
    }
    free(Q); free(foff);
    return err;
}

int ref_level_pf_evolve(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                        int nbox, const cmp_box* boxes, int ng, double* U, double dt, double* info)
{
    return pf_evolve_int(P, PF, G, nbox, boxes, ng, U, dt, info, 3);
}

/* Evolve's step loop with the Phase-Field step (:202-312): ComputeDt, the flow's RK3 step, then
 * PhaseField_Evolve with the same dt. */
int ref_level_advance_pf_int(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                             int nbox, const cmp_box* boxes, int ng, double* U, int nsteps, double cfl,
                             int limiter, double* dt_hist, int integrator)
{
    double cmax = 0.0, dt = 0.0;
    int err = ref_level_initial_cmax(P, G, nbox, boxes, ng, U, &cmax);
    for (int s = 0; s < nsteps && !err; ++s) {
        dt = ref_compute_dt(G, cmax, cfl, limiter, s == 0 ? ref_est_dt(G, cmax, cfl) : dt);
        if (dt_hist) dt_hist[s] = dt;
        cmax = 0.0;
        err = ref_level_rk_step(P, G, nbox, boxes, ng, U, dt, &cmax, integrator);
        if (isnan(cmax)) err = CMP_ERR_NAN;
        if (!err) err = pf_evolve_int(P, PF, G, nbox, boxes, ng, U, dt, NULL, integrator);
    }
    return err;
}

int ref_level_advance_pf(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                         int nbox, const cmp_box* boxes, int ng, double* U, int nsteps, double cfl,
                         int limiter, double* dt_hist)
{
    return ref_level_advance_pf_int(P, PF, G, nbox, boxes, ng, U, nsteps, cfl, limiter, dt_hist, 3);
}

/* ------------------------------------------------------------------------------------------------
 * Layout helpers (the same definitions as the product library's; the oracle keeps its own copy so
 * that it links nothing from it)
 * ---------------------------------------------------------------------------------------------- */

int cmp_face_layout(int dim, int nbox, const cmp_box* boxes, int ncomp, long long* offsets,
                    long long* total)
{
    long long off = 0;
    for (int b = 0; b < nbox; ++b)
        for (int d = 0; d < dim; ++d) {
            long long n = ncomp;
            for (int e = 0; e < dim; ++e) n *= (boxes[b].hi[e] - boxes[b].lo[e] + 1 + (e == d));
            if (offsets) offsets[b * dim + d] = off;
            off += n;
        }
    if (total) *total = off;
    return CMP_OK;
}

long long cmp_fab_size(int dim, const cmp_box* box, int ng, int ncomp)
{
    long long n = ncomp;
    for (int d = 0; d < dim; ++d) n *= (box->hi[d] - box->lo[d] + 1 + 2 * ng);
    return n;
}

int cmp_level_layout(int dim, int nbox, cmp_box* boxes, int ng, int ncomp, long long* total)
{
    long long off = 0;
    for (int b = 0; b < nbox; ++b) {
        boxes[b].offset = off;
        off += cmp_fab_size(dim, &boxes[b], ng, ncomp);
    }
    if (total) *total = off;
    return CMP_OK;
}

int ref_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void ref_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------------------------------
 * Refinement tagging: ErrorEst (source/COMPAS.cpp:769-932) with the criteria of include/Tagging.H
 * (state_grad :172-201, state_value_less / _greater / _range :118-160) on state components or on
 * the derived rho, u, v, w, P of compute_derived_var
 * (include/Physics_Compressible5Eq/ComputeDerivedVar_Compressible5Eq.H:110-160; the N-phase model's
 * …NPhase.H likewise: rho = sum of the partial densities).
 * ---------------------------------------------------------------------------------------------- */
static double tag_var(const cmp_params* P, const fab_t* f, const double* U, int i, int j, int k, int var)
{
    const long long c = fab_idx(f, i, j, k);
    if (var >= 0) return U[c + var * f->sn];
    double Uc[REF_NMAX], W[REF_NMAX];
    int err = 0;
    for (int n = 0; n < P->nstate; ++n) Uc[n] = U[c + n * f->sn];
    const int nm = P->nphase ? P->nphase : 2;
    if (var == CMP_TAG_RHO) {
        double rho = Uc[0];
        for (int q = 1; q < nm; ++q) rho += Uc[q];
        return rho;
    }
    if (P->nphase) ref_np_cons2prim(P, Uc, W, &err);
    else ref_cons2prim(P, Uc, W, &err);
    if (var == CMP_TAG_P) return W[nm + P->dim];
    const int m = CMP_TAG_U - var;          /* u, v, w */
    return (m < P->dim) ? W[nm + m] : 0.0;
}

int ref_level_tag(const cmp_params* P, int nbox, const cmp_box* boxes, int ng, const double* U,
                  const cmp_tag_criterion* crit, int ncrit, unsigned char* tags)
{
    const int dim = P->dim;
    long long toff = 0;
    for (int b = 0; b < nbox; ++b) {
        const cmp_box* B = &boxes[b];
        fab_t f = make_fab(dim, B, ng);
        const double* UB = U + B->offset;
        int nv[3] = {1, 1, 1};
        for (int d = 0; d < dim; ++d) nv[d] = B->hi[d] - B->lo[d] + 1;
        const int lo2 = (dim == 3) ? B->lo[2] : 0;
        for (int kk = 0; kk < nv[2]; ++kk)
            for (int jj = 0; jj < nv[1]; ++jj)
                for (int ii = 0; ii < nv[0]; ++ii) {
                    const int i = B->lo[0] + ii, j = B->lo[1] + jj, k = lo2 + kk;
                    unsigned char t = 0;
                    for (int q = 0; q < ncrit; ++q) {
                        const cmp_tag_criterion* C = &crit[q];
                        const double s = tag_var(P, &f, UB, i, j, k, C->var);
                        if (C->kind == CMP_TAG_GRAD) {
                            double ax = fabs(tag_var(P, &f, UB, i - 1, j, k, C->var) - s);
                            double ay = fabs(tag_var(P, &f, UB, i, j - 1, k, C->var) - s);
                            ax = smax(ax, fabs(s - tag_var(P, &f, UB, i + 1, j, k, C->var)));
                            ay = smax(ay, fabs(s - tag_var(P, &f, UB, i, j + 1, k, C->var)));
                            double az = 0.0;
                            if (dim == 3) {
                                az = fabs(tag_var(P, &f, UB, i, j, k - 1, C->var) - s);
                                az = smax(az, fabs(s - tag_var(P, &f, UB, i, j, k + 1, C->var)));
                            }
                            if (smax(smax(ax, ay), az) >= C->v0) t = 1;
                        } else if (C->kind == CMP_TAG_LESS) {
                            if (s < C->v0) t = 1;
                        } else if (C->kind == CMP_TAG_GREATER) {
                            if (s > C->v0) t = 1;
                        } else if (C->kind == CMP_TAG_RANGE) {
                            if (s > C->v0 && s < C->v1) t = 1;
                        }
                    }
                    tags[toff + ii + (long long)nv[0] * (jj + (long long)nv[1] * kk)] = t;
                }
        toff += (long long)nv[0] * nv[1] * nv[2];
    }
    return CMP_OK;
}

/* ------------------------------------------------------------------------------------------------
 * AverageDownTo (source/COMPAS.cpp:1550-1562): amrex::average_down of the fine level onto the coarse
 * cells it covers, for cell-centred data on a Cartesian grid: every coarse cell under a fine box
 * takes the mean of its ratio^dim children, summed k, then j, then i (slowest to fastest) and
 * multiplied by 1/ratio^dim (AMReX 26.01's amrex_avgdown, not vendored; its published form).
 * ---------------------------------------------------------------------------------------------- */
int ref_level_average_down(const cmp_params* P, int nbf, const cmp_box* fboxes, int ngf, const double* Uf,
                           int nbc, const cmp_box* cboxes, int ngc, double* Uc, int ratio)
{
    const int dim = P->dim, N = P->nstate;
    const int rz = (dim == 3) ? ratio : 1;
    double vol = 1.0;
    for (int d = 0; d < dim; ++d) vol *= ratio;
    const double volfrac = 1.0 / vol;
    for (int b = 0; b < nbf; ++b) {
        const cmp_box* F = &fboxes[b];
        for (int d = 0; d < dim; ++d)
            if (((F->lo[d] % ratio) + ratio) % ratio != 0 || ((F->hi[d] + 1) % ratio + ratio) % ratio != 0)
                return CMP_ERR_LAYOUT;
        fab_t ff = make_fab(dim, F, ngf);
        int clo[3] = {0, 0, 0}, chi[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) {
            clo[d] = F->lo[d] / ratio;
            chi[d] = (F->hi[d] + 1) / ratio - 1;
        }
        for (int kc = clo[2]; kc <= chi[2]; ++kc)
            for (int jc = clo[1]; jc <= chi[1]; ++jc)
                for (int ic = clo[0]; ic <= chi[0]; ++ic) {
                    int cb = -1;
                    for (int q = 0; q < nbc && cb < 0; ++q) {
                        int in = 1;
                        const int cc[3] = {ic, jc, kc};
                        for (int d = 0; d < dim; ++d)
                            if (cc[d] < cboxes[q].lo[d] || cc[d] > cboxes[q].hi[d]) in = 0;
                        if (in) cb = q;
                    }
                    if (cb < 0) return CMP_ERR_LAYOUT;
                    fab_t fc = make_fab(dim, &cboxes[cb], ngc);
                    const long long co = cboxes[cb].offset + fab_idx(&fc, ic, jc, kc);
                    for (int n = 0; n < N; ++n) {
                        double c = 0.0;
                        for (int kr = 0; kr < rz; ++kr)
                            for (int jr = 0; jr < ratio; ++jr)
                                for (int ir = 0; ir < ratio; ++ir) {
                                    const int i = ic * ratio + ir, j = jc * ratio + jr;
                                    const int k = (dim == 3) ? kc * ratio + kr : 0;
                                    c += Uf[F->offset + fab_idx(&ff, i, j, k) + n * ff.sn];
                                }
                        Uc[co + n * fc.sn] = volfrac * c;
                    }
                }
    }
    return CMP_OK;
}

/* ------------------------------------------------------------------------------------------------
 * FillPatchTwoLevels (source/COMPAS.cpp:1679-1724) with lincc_interp: every ghost cell of the fine
 * level is mapped to its source cell as the same-level fill maps it (periodic image, clamp or
 * mirror with odd momenta); a source cell in a fine box is copied, any other takes the coarse
 * level's cell-conservative linear interpolation at it. lincc_interp is AMReX 26.01's
 * CellConservativeLinear with linear limiting (not vendored; its published algorithm): per
 * direction the central coarse difference of every component, dc = (c+ - c-)/2, its MC-limited
 * value sx = sign(dc) min(|dc|, 2|c+ - c|, 2|c - c-|) (0 where the one-sided differences disagree in
 * sign), one factor per direction f = min over the components with dc != 0 of sx/dc (the slopes
 * of all components scaled alike), and the fine value c + sum_d off_d f_d dc_d with off_d the fine
 * cell's centre offset in coarse cells, (i - ic r + 1/2)/r - 1/2. The coarse neighbours come from
 * the coarse level by the same mapping.
 * ---------------------------------------------------------------------------------------------- */
static int map_source(int dim, const cmp_geometry* G, int N, int mom0, int* src, double* sgn)
{
    for (int n = 0; n < N; ++n) sgn[n] = 1.0;
    for (int d = 0; d < dim; ++d) {
        int dlo = G->domain_lo[d], dhi = G->domain_hi[d];
        if (src[d] >= dlo && src[d] <= dhi) continue;
        int per = dhi - dlo + 1;
        if (G->is_periodic[d]) {
            while (src[d] < dlo) src[d] += per;
            while (src[d] > dhi) src[d] -= per;
            continue;
        }
        int code = (src[d] < dlo) ? G->lo_bc[d] : G->hi_bc[d];
        if (code == 2 || code == 0) {
            src[d] = (src[d] < dlo) ? dlo : dhi;
        } else if (code >= 3 && code <= 5) {
            src[d] = (src[d] < dlo) ? (2 * dlo - 1 - src[d]) : (2 * dhi + 1 - src[d]);
            /* a field without momenta (mom0 outside [0, N - dim]) has no sign to flip */
            if (mom0 >= 0 && mom0 + dim <= N)
                for (int m = 0; m < dim; ++m)
                    if (m == d || code == 5) sgn[mom0 + m] = -sgn[mom0 + m];
        } else {
            return CMP_ERR_BC;
        }
    }
    return CMP_OK;
}

static int find_box(int dim, int nbox, const cmp_box* boxes, const int* c)
{
    for (int q = 0; q < nbox; ++q)
        if (box_contains(dim, &boxes[q], c)) return q;
    return -1;
}

static int coarse_cell(int dim, int N, int mom0, const cmp_geometry* Gc, int nbc, const cmp_box* cb, int ngc,
                       const double* Uc, const int* cell, double* v)
{
    int src[3] = {cell[0], cell[1], cell[2]};
    double sgn[REF_NMAX];
    int e = map_source(dim, Gc, N, mom0, src, sgn);
    if (e) return e;
    int q = find_box(dim, nbc, cb, src);
    if (q < 0) return CMP_ERR_LAYOUT;
    fab_t f = make_fab(dim, &cb[q], ngc);
    const long long i = cb[q].offset + fab_idx(&f, src[0], src[1], src[2]);
    for (int n = 0; n < N; ++n) v[n] = sgn[n] * Uc[i + n * f.sn];
    return CMP_OK;
}

static int floor_div(int a, int r) { return (a >= 0) ? a / r : -((-a + r - 1) / r); }

/* The two-level fill of the cells of the boxes dboxes (buffer Ud; all_cells: valid and ghost cells,
 * else the ghost cells only), the same-level sources taken from the boxes sboxes (buffer Us) of
 * the same level. */
static int fill_two_levels_nc(int dim, int N, int mom0, const cmp_geometry* Gf, int nbf, const cmp_box* fboxes,
                              int ngf, double* Uf, int nbs, const cmp_box* sboxes, int ngs, const double* Us,
                              int all_cells, const cmp_geometry* Gc, int nbc, const cmp_box* cboxes, int ngc,
                              const double* Uc, int ratio)
{
    int err = 0;
    for (int b = 0; b < nbf; ++b) {
        const cmp_box* bx = &fboxes[b];
        fab_t f = make_fab(dim, bx, ngf);
        double* base = Uf + bx->offset;
        for (int k = f.lo[2]; k <= f.lo[2] + f.ext[2] - 1; ++k)
            for (int j = f.lo[1]; j <= f.lo[1] + f.ext[1] - 1; ++j)
                for (int i = f.lo[0]; i <= f.lo[0] + f.ext[0] - 1; ++i) {
                    int c[3] = {i, j, k};
                    if (!all_cells && box_contains(dim, bx, c)) continue;
                    int src[3] = {i, j, k};
                    double sgn[REF_NMAX], v[REF_NMAX];
                    int e = map_source(dim, Gf, N, mom0, src, sgn);
                    if (e) { if (e > err) err = e; continue; }
                    const long long di = fab_idx(&f, i, j, k);
                    int q = find_box(dim, nbs, sboxes, src);
                    if (q >= 0) {
                        fab_t fs = make_fab(dim, &sboxes[q], ngs);
                        const long long si = sboxes[q].offset + fab_idx(&fs, src[0], src[1], src[2]);
                        for (int n = 0; n < N; ++n) base[di + n * f.sn] = sgn[n] * Us[si + n * fs.sn];
                        continue;
                    }
                    /* coarse-fine ghost: lincc at the source cell */
                    int ic[3] = {0, 0, 0};
                    for (int d = 0; d < 3; ++d) ic[d] = (d < dim) ? floor_div(src[d], ratio) : src[d];
                    double c0[REF_NMAX], cm[3][REF_NMAX], cp[3][REF_NMAX], slope[3][REF_NMAX];
                    e = coarse_cell(dim, N, mom0, Gc, nbc, cboxes, ngc, Uc, ic, c0);
                    for (int d = 0; d < dim && !e; ++d) {
                        int a[3] = {ic[0], ic[1], ic[2]}, z[3] = {ic[0], ic[1], ic[2]};
                        a[d] -= 1;
                        z[d] += 1;
                        e = coarse_cell(dim, N, mom0, Gc, nbc, cboxes, ngc, Uc, a, cm[d]);
                        if (!e) e = coarse_cell(dim, N, mom0, Gc, nbc, cboxes, ngc, Uc, z, cp[d]);
                    }
                    if (e) { if (e > err) err = e; continue; }
                    for (int d = 0; d < dim; ++d) {
                        double sf = 1.0;
                        for (int n = 0; n < N; ++n) {
                            const double dc = 0.5 * (cp[d][n] - cm[d][n]);
                            const double df = 2.0 * (cp[d][n] - c0[n]);
                            const double db = 2.0 * (c0[n] - cm[d][n]);
                            double sx = (df * db >= 0.0) ? smin(fabs(df), fabs(db)) : 0.0;
                            sx = copysign(1.0, dc) * smin(sx, fabs(dc));
                            if (dc != 0.0) sf = smin(sf, sx / dc);
                            slope[d][n] = dc;
                        }
                        for (int n = 0; n < N; ++n) slope[d][n] *= sf;
                    }
                    for (int n = 0; n < N; ++n) {
                        double val = c0[n];
                        for (int d = 0; d < dim; ++d) {
                            const double off = ((double)(src[d] - ic[d] * ratio) + 0.5) / (double)ratio - 0.5;
                            val = val + off * slope[d][n];
                        }
                        v[n] = val;
                    }
                    for (int n = 0; n < N; ++n) base[di + n * f.sn] = sgn[n] * v[n];
                }
    }
    return err;
}

static int fill_two_levels_impl(const cmp_params* P, const cmp_geometry* Gf, int nbf, const cmp_box* fboxes, int ngf,
                                double* Uf, int nbs, const cmp_box* sboxes, int ngs, const double* Us, int all_cells,
                                const cmp_geometry* Gc, int nbc, const cmp_box* cboxes, int ngc, const double* Uc,
                                int ratio)
{
    return fill_two_levels_nc(P->dim, P->nstate, P->nphase ? P->nphase : 2, Gf, nbf, fboxes, ngf, Uf, nbs, sboxes, ngs,
                              Us, all_cells, Gc, nbc, cboxes, ngc, Uc, ratio);
}

int ref_level_fill_two_levels(const cmp_params* P, const cmp_geometry* Gf, int nbf, const cmp_box* fboxes, int ngf,
                              double* Uf, const cmp_geometry* Gc, int nbc, const cmp_box* cboxes, int ngc,
                              const double* Uc, int ratio)
{
    return fill_two_levels_impl(P, Gf, nbf, fboxes, ngf, Uf, nbf, fboxes, ngf, Uf, 0, Gc, nbc, cboxes, ngc, Uc, ratio);
}

/* A level's new boxes after a regrid (RemakeLevel / MakeNewLevelFromCoarse, source/COMPAS.cpp,
 * FillPatch of the new MultiFab): every cell of the new boxes, valid and ghost, from the old
 * boxes of the level where one holds its source cell, else from the coarse level by the same
 * interpolation as the two-level fill. nbo = 0: a level made from the coarse level alone. */
int ref_level_fill_from(const cmp_params* P, const cmp_geometry* Gf, int nbn, const cmp_box* nboxes, int ngn,
                        double* Un, int nbo, const cmp_box* oboxes, int ngo, const double* Uo,
                        const cmp_geometry* Gc, int nbc, const cmp_box* cboxes, int ngc, const double* Uc, int ratio)
{
    return fill_two_levels_impl(P, Gf, nbn, nboxes, ngn, Un, nbo, oboxes, ngo, Uo, 1, Gc, nbc, cboxes, ngc, Uc, ratio);
}

/* ------------------------------------------------------------------------------------------------
 * All levels of a hierarchy, stage by stage, without subcycling: AdvanceAllLevelsStageCoupled
 * (source/Class_Compressible_PhaseField_StageCoupling.cpp:156-) with run.do_reflux = 1 and
 * StageRHSAllLevels (:84-153), then AverageDown (timeStepNoSubcycling, source/COMPAS.cpp:2139-2180).
 * Per stage: every level's stage state takes its ghost cells (level 0 the same-level fill, finer
 * levels the two-level fill from the next coarser stage state); every box computes its faces; from
 * the finest level down, the coarse faces under a finer level take the mean of the fine faces over
 * them (amrex::average_down_faces with periodicity: the face's value goes to every coarse box whose
 * face box holds the face or a periodic image of it; the fine faces summed with the lower transverse
 * index fastest, times 1/ratio^(dim-1)), fluxes and u*_n alike; every level then takes its
 * divergence and the stage combination of ThirdOrderSSPRK; the levels are averaged down (finest
 * first). ComputeDt over levels (:2250-2302, nsubsteps 1): dt = min over levels of the level's
 * estimate, each limited to 1.1 dt_prev.
 * ---------------------------------------------------------------------------------------------- */
static long long level_total(int dim, int nbox, const cmp_box* boxes, int ng, int N)
{
    long long t = 0;
    for (int b = 0; b < nbox; ++b) t += cmp_fab_size(dim, &boxes[b], ng, N);
    return t;
}

static void face_extents(int dim, const cmp_box* b, int d, int* fe)
{
    for (int e = 0; e < 3; ++e) fe[e] = 1;
    for (int e = 0; e < dim; ++e) fe[e] = b->hi[e] - b->lo[e] + 1 + (e == d ? 1 : 0);
}

/* faces F[lev][b*3+d], FN[lev][b*3+d] */
static void average_faces(const cmp_params* P, const cmp_geometry* Gc, int nbf, const cmp_box* fb, double** Ff,
                          double** FNf, int nbc, const cmp_box* cb, double** Fc, double** FNc, int ratio)
{
    const int dim = P->dim, N = P->nstate;
    int rt = 1;
    for (int e = 1; e < dim; ++e) rt *= ratio;
    const double facInv = 1.0 / (double)rt;
    for (int b = 0; b < nbf; ++b)
        for (int d = 0; d < dim; ++d) {
            int ffe[3];
            face_extents(dim, &fb[b], d, ffe);
            const long long nff = (long long)ffe[0] * ffe[1] * ffe[2];
            int clo[3] = {0, 0, 0}, chi[3] = {0, 0, 0};
            for (int e = 0; e < dim; ++e) {
                clo[e] = fb[b].lo[e] / ratio;
                chi[e] = (fb[b].hi[e] + 1) / ratio - 1 + (e == d ? 1 : 0);
            }
            /* transverse directions: t1 the lower (fastest), t2 the higher */
            int t1 = -1, t2 = -1;
            for (int e = 0; e < dim; ++e)
                if (e != d) { if (t1 < 0) t1 = e; else t2 = e; }
            const int r1 = (t1 >= 0) ? ratio : 1, r2 = (t2 >= 0) ? ratio : 1;
            for (int K = clo[2]; K <= chi[2]; ++K)
                for (int J = clo[1]; J <= chi[1]; ++J)
                    for (int I = clo[0]; I <= chi[0]; ++I) {
                        double avg[REF_NMAX], avgn = 0.0;
                        const int C[3] = {I, J, K};
                        for (int n = 0; n <= N; ++n) {
                            double c = 0.0;
                            for (int o2 = 0; o2 < r2; ++o2)
                                for (int o1 = 0; o1 < r1; ++o1) {
                                    int fcell[3] = {0, 0, 0};
                                    for (int e = 0; e < dim; ++e) fcell[e] = C[e] * ratio;
                                    if (t1 >= 0) fcell[t1] += o1;
                                    if (t2 >= 0) fcell[t2] += o2;
                                    long long q = 0;
                                    for (int e = dim - 1; e >= 0; --e) q = q * ffe[e] + (fcell[e] - fb[b].lo[e]);
                                    c += (n < N) ? Ff[b * 3 + d][q + n * nff] : FNf[b * 3 + d][q];
                                }
                            if (n < N) avg[n] = c * facInv; else avgn = c * facInv;
                        }
                        /* the face and its periodic images along d */
                        int nimg = 1;
                        int pos[3] = {C[d], C[d], C[d]};
                        const int ncd = Gc->domain_hi[d] - Gc->domain_lo[d] + 1;
                        if (Gc->is_periodic[d]) {
                            if (C[d] == Gc->domain_lo[d]) pos[nimg++] = C[d] + ncd;
                            else if (C[d] == Gc->domain_hi[d] + 1) pos[nimg++] = C[d] - ncd;
                        }
                        for (int m = 0; m < nimg; ++m) {
                            int F3[3] = {C[0], C[1], C[2]};
                            F3[d] = pos[m];
                            for (int q = 0; q < nbc; ++q) {
                                int cfe[3];
                                face_extents(dim, &cb[q], d, cfe);
                                int in = 1;
                                for (int e = 0; e < dim; ++e)
                                    if (F3[e] < cb[q].lo[e] || F3[e] > cb[q].lo[e] + cfe[e] - 1) in = 0;
                                if (!in) continue;
                                const long long ncf = (long long)cfe[0] * cfe[1] * cfe[2];
                                long long qq = 0;
                                for (int e = dim - 1; e >= 0; --e) qq = qq * cfe[e] + (F3[e] - cb[q].lo[e]);
                                for (int n = 0; n < N; ++n) Fc[q * 3 + d][qq + n * ncf] = avg[n];
                                FNc[q * 3 + d][qq] = avgn;
                            }
                        }
                    }
        }
}

static int amr_fill(const cmp_params* P, int nlev, const cmp_geometry* G, const int* nbox,
                    const cmp_box* const* boxes, int ng, double** Ub, int ratio)
{
    int err = ref_level_fill_ghosts(P, &G[0], nbox[0], boxes[0], ng, Ub[0]);
    for (int l = 1; l < nlev; ++l) {
        int e = ref_level_fill_two_levels(P, &G[l], nbox[l], boxes[l], ng, Ub[l], &G[l - 1], nbox[l - 1],
                                          boxes[l - 1], ng, Ub[l - 1], ratio);
        if (e && !err) err = e;
    }
    return err;
}

int ref_amr_rk3_step(const cmp_params* P, int nlev, const cmp_geometry* G, const int* nbox,
                     const cmp_box* const* boxes, int ng, double* const* Un, double dt, int ratio, double* cmax)
{
    const int dim = P->dim, N = P->nstate;
    if (nlev < 1 || nlev > 8) return CMP_ERR_ARG;
    double* Ub[8];
    double* Unew[8];
    double** F[8];
    double** FN[8];
    long long tot[8];
    for (int l = 0; l < nlev; ++l) {
        tot[l] = level_total(dim, nbox[l], boxes[l], ng, N);
        Ub[l] = (double*)malloc(sizeof(double) * tot[l]);
        Unew[l] = (double*)malloc(sizeof(double) * tot[l]);
        F[l] = (double**)calloc((size_t)nbox[l] * 3, sizeof(double*));
        FN[l] = (double**)calloc((size_t)nbox[l] * 3, sizeof(double*));
    }
    int err = 0, e;
    const double w2 = 0.25 * dt, w3 = (2.0 / 3.0) * dt;
    for (int stage = 0; stage < 3; ++stage) {
        for (int l = 0; l < nlev; ++l) memcpy(Ub[l], stage == 0 ? Un[l] : Unew[l], sizeof(double) * tot[l]);
        e = amr_fill(P, nlev, G, nbox, boxes, ng, Ub, ratio);
        if (e && !err) err = e;
        for (int l = 0; l < nlev; ++l)
            for (int b = 0; b < nbox[l]; ++b) {
                double cm = 0.0;
                e = box_faces(P, &G[l], &boxes[l][b], ng, Ub[l] + boxes[l][b].offset, &F[l][b * 3], &FN[l][b * 3], &cm);
                if (e && !err) err = e;
                if (cmax[l] < cm || isnan(cm)) cmax[l] = cm;
            }
        for (int l = nlev - 2; l >= 0; --l)
            average_faces(P, &G[l], nbox[l + 1], boxes[l + 1], F[l + 1], FN[l + 1], nbox[l], boxes[l], F[l], FN[l],
                          ratio);
        for (int l = 0; l < nlev; ++l)
            for (int b = 0; b < nbox[l]; ++b) {
                const cmp_box* bx = &boxes[l][b];
                const long long nc = valid_cells(dim, bx);
                double* L = (double*)malloc(sizeof(double) * nc * N);
                e = box_divergence(P, &G[l], bx, ng, Ub[l] + bx->offset, F[l] + b * 3, FN[l] + b * 3, L);
                if (e && !err) err = e;
                fab_t f = make_fab(dim, bx, ng);
                int nv[3] = {1, 1, 1};
                for (int d = 0; d < dim; ++d) nv[d] = bx->hi[d] - bx->lo[d] + 1;
                const int lo2 = (dim == 3) ? bx->lo[2] : 0;
                for (int n = 0; n < N; ++n)
                    for (int kk = 0; kk < nv[2]; ++kk)
                        for (int jj = 0; jj < nv[1]; ++jj)
                            for (int ii = 0; ii < nv[0]; ++ii) {
                                const long long q = bx->offset + fab_idx(&f, bx->lo[0] + ii, bx->lo[1] + jj, lo2 + kk) + n * f.sn;
                                const double Lq = L[ii + (long long)nv[0] * (jj + (long long)nv[1] * kk) + n * nc];
                                if (stage == 0) {
                                    Unew[l][q] = 1.0 * Ub[l][q] + dt * Lq;
                                } else if (stage == 1) {
                                    Unew[l][q] = 0.25 * Ub[l][q] + 0.75 * Un[l][q];
                                    Unew[l][q] += w2 * Lq;
                                } else {
                                    Unew[l][q] = (2.0 / 3.0) * Ub[l][q] + (1.0 / 3.0) * Un[l][q];
                                    Unew[l][q] += w3 * Lq;
                                }
                            }
                free(L);
            }
        for (int l = 0; l < nlev; ++l)
            for (int q = 0; q < nbox[l] * 3; ++q) { free(F[l][q]); free(FN[l][q]); F[l][q] = FN[l][q] = NULL; }
        for (int l = nlev - 2; l >= 0; --l) {
            e = ref_level_average_down(P, nbox[l + 1], boxes[l + 1], ng, Unew[l + 1], nbox[l], boxes[l], ng, Unew[l],
                                       ratio);
            if (e && !err) err = e;
        }
    }
    for (int l = 0; l < nlev; ++l) {
        for (int b = 0; b < nbox[l]; ++b) {
            const cmp_box* bx = &boxes[l][b];
            fab_t f = make_fab(dim, bx, ng);
            int nv[3] = {1, 1, 1};
            for (int d = 0; d < dim; ++d) nv[d] = bx->hi[d] - bx->lo[d] + 1;
            const int lo2 = (dim == 3) ? bx->lo[2] : 0;
            for (int n = 0; n < N; ++n)
                for (int kk = 0; kk < nv[2]; ++kk)
                    for (int jj = 0; jj < nv[1]; ++jj)
                        for (int ii = 0; ii < nv[0]; ++ii) {
                            const long long q = bx->offset + fab_idx(&f, bx->lo[0] + ii, bx->lo[1] + jj, lo2 + kk) + n * f.sn;
                            Un[l][q] = Unew[l][q];
                        }
        }
        free(Ub[l]); free(Unew[l]); free(F[l]); free(FN[l]);
    }
    return err;
}

double ref_amr_compute_dt(int nlev, const cmp_geometry* G, const double* cmax, double cfl, int limiter,
                          double dt_prev)
{
    double dt0 = 0.0;
    for (int l = 0; l < nlev; ++l) {
        double t = ref_est_dt(&G[l], cmax[l], cfl);
        if (limiter && dt_prev > 0.0) t = smin(t, 1.1 * dt_prev);
        dt0 = (l == 0) ? t : smin(dt0, t);
    }
    return dt0;
}

/* The initial c_max of every level (CalculateInitialcmax on the filled hierarchy), then nsteps
 * steps: ComputeDt over levels, one stage-coupled RK3 step. */
int ref_amr_advance(const cmp_params* P, int nlev, const cmp_geometry* G, const int* nbox,
                    const cmp_box* const* boxes, int ng, double* const* U, int nsteps, double cfl, int limiter,
                    int ratio, double* dt_hist)
{
    const int dim = P->dim, N = P->nstate;
    double cmax[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double* Ub[8];
    int err = 0;
    if (nlev < 1 || nlev > 8) return CMP_ERR_ARG;
    for (int l = 0; l < nlev; ++l) {
        const long long t = level_total(dim, nbox[l], boxes[l], ng, N);
        Ub[l] = (double*)malloc(sizeof(double) * t);
        memcpy(Ub[l], U[l], sizeof(double) * t);
    }
    err = amr_fill(P, nlev, G, nbox, boxes, ng, Ub, ratio);
    for (int l = 0; l < nlev && !err; ++l)
        for (int b = 0; b < nbox[l]; ++b) {
            double* Fx[3] = {NULL, NULL, NULL};
            double* FNx[3] = {NULL, NULL, NULL};
            double cm = 0.0;
            int e = box_faces(P, &G[l], &boxes[l][b], ng, Ub[l] + boxes[l][b].offset, Fx, FNx, &cm);
            if (e && !err) err = e;
            if (cmax[l] < cm || isnan(cm)) cmax[l] = cm;
            for (int d = 0; d < 3; ++d) { free(Fx[d]); free(FNx[d]); }
        }
    for (int l = 0; l < nlev; ++l) free(Ub[l]);
    double dt = 0.0;
    for (int s = 0; s < nsteps && !err; ++s) {
        dt = ref_amr_compute_dt(nlev, G, cmax, cfl, limiter, s == 0 ? 0.0 : dt);
        if (dt_hist) dt_hist[s] = dt;
        for (int l = 0; l < nlev; ++l) cmax[l] = 0.0;
        err = ref_amr_rk3_step(P, nlev, G, nbox, boxes, ng, U, dt, ratio, cmax);
        for (int l = 0; l < nlev; ++l)
            if (isnan(cmax[l])) err = CMP_ERR_NAN;
    }
    return err;
}

/* ------------------------------------------------------------------------------------------------
 * The Phase-Field step of a hierarchy without subcycling (run.do_subcycle = 0): PhaseField_Evolve
 * (source/Class_Compressible_PhaseField.cpp:425-594) over finest_level + 1 levels.
 *   - MaxVelocity over every level (:3577-3590); eta from the finest cell size (:433-443); mobility
 *     0.5 eta max|u|.
 *   - PFBorder: the step's volume fractions with their ghost cells, every phase in one FillPatch
 *     (:485-493): level 0 same-level, finer levels two-level (the coarse-fine ghosts by the
 *     interpolation of the flow's two-level fill, its limiter shared by the phases); PF_bcs int_dir
 *     on periodic sides, foextrap elsewhere.
 *   - PhaseFieldMechanism_Evolve (:597-709): sub-steps of dt_m = 0.9/(2 M Lambda)/(sum_d 1/dx_d^2 +
 *     1/eta^2) with the finest level's dx, every level advanced by each sub-step after the filled
 *     sub-step state (FillPatch of PF_new, all phases):
 *       CAC (PhaseField_Mechanism :1007-1152): per phase the cell normals of alpha (PhaseField_Normal
 *       :1958-2032, valid cells, then their own FillPatch), mech and W per level, AverageDown of both
 *       (the coarse cells under a finer level take the mean of their children), their sums over
 *       level 0, Div_J = mech + W L at every level;
 *       ACDI (:1155-1233, PhaseField_ACDI :1495-1644): psi of the valid cells and its FillPatch, per
 *       phase the normals of psi and J on every face of every level, then AverageDownFaces(J), Div_J
 *       = div J; Jhat += dt_m J after the averaging (explicit RC);
 *       alpha += dt_m Div_J on every level's valid cells.
 *   - Grad_Q (:2133-2403): explicit RC per level from Jhat/dt and PFBorder; implicit RC per phase:
 *     b = W(phi_f) + 1e-3 on every face of every level from PFBorder, AverageDownFaces(b), the
 *     composite solve below, AverageDownFaces(Grad_Q).
 *   - Phase-Field sub-steps: PhaseField_ComputeDt (:2037-2098) over the levels with the no-subcycling
 *     minimum, each sub-step the stage-coupled TVD-RK3 of the Phase-Field right-hand side over the
 *     levels (PhaseField_AdvanceAllLevels, Class_Compressible_PhaseField_StageCoupling.cpp:43-51, the
 *     coarse faces under a finer level taking the mean of its faces before the divergence) and
 *     AverageDown (PhaseField_TimeStepNoSubcycling :3010-3030).
 *
 * The composite solve (the reference calls AMReX's MLMG with composite_solve = 1,
 * include/Operators_LinearSystem/SolveABecLaplacianAllLevels_HomBC.H:11-62, MLMG not vendored):
 * the unknowns are Q on the cells no finer level covers; every face between two such cells carries
 * the flux b (Q_n - Q_c) / h; a coarse-fine face carries, per fine face on it, b_fine (Q_coarse -
 * Q_fine) / ((dx_fine + dx_coarse) / 2) (the fine ghost value linear between the two cell centres
 * along the normal), which is the coarse cell's flux through that face too (conservative and
 * symmetric). The volume-weighted system is solved by Jacobi-preconditioned conjugate gradients on
 * the volume-weighted zero-mean subspace to the rounding floor; covered cells then take the mean of
 * their children; Grad_Q is (Q_c - Q_m)/h on a level's faces between two of its cells, the two-point
 * difference on its coarse-fine faces, zero on Neumann domain faces. MLMG's coarse-fine stencil
 * (LinOp_maxorder 3, LinearSystem_Parameter.H:18) differs from this two-point flux at the
 * interfaces: parity of the hierarchy's implicit Grad_Q with the reference is unpinned there.
 * ---------------------------------------------------------------------------------------------- */
#define AMR_NLEV_MAX 8

typedef struct {
    int dim, nlev, ratio;
    cmp_geometry PG[AMR_NLEV_MAX];   /* PF_bcs: int_dir on periodic sides, foextrap elsewhere */
    int nbox[AMR_NLEV_MAX];
    const cmp_box* boxes[AMR_NLEV_MAX];
    cmp_box* a1[AMR_NLEV_MAX];       /* 1 ghost, 1 comp */
    cmp_box* ap[AMR_NLEV_MAX];       /* 1 ghost, npf comps */
    cmp_box* an[AMR_NLEV_MAX];       /* 1 ghost, dim comps */
    long long n1[AMR_NLEV_MAX], np[AMR_NLEV_MAX], nn[AMR_NLEV_MAX];
    long long* f1[AMR_NLEV_MAX];     /* face offsets, 1 comp */
    long long* fp[AMR_NLEV_MAX];     /* face offsets, npf comps */
    long long nf1[AMR_NLEV_MAX], nfp[AMR_NLEV_MAX];
    int npf;
} hier_t;

static void hier_init(hier_t* H, const cmp_params* P, int nlev, const cmp_geometry* G, const int* nbox,
                      const cmp_box* const* boxes, int ratio)
{
    memset(H, 0, sizeof(*H));
    H->dim = P->dim;
    H->nlev = nlev;
    H->ratio = ratio;
    H->npf = P->nphase ? P->nphase : 1;
    for (int l = 0; l < nlev; ++l) {
        H->PG[l] = G[l];
        for (int d = 0; d < 3; ++d)
            if (!H->PG[l].is_periodic[d]) { H->PG[l].lo_bc[d] = 2; H->PG[l].hi_bc[d] = 2; }
        H->nbox[l] = nbox[l];
        H->boxes[l] = boxes[l];
        H->a1[l] = pf_boxes(P->dim, nbox[l], boxes[l], 1, 1, &H->n1[l]);
        H->ap[l] = pf_boxes(P->dim, nbox[l], boxes[l], 1, H->npf, &H->np[l]);
        H->an[l] = pf_boxes(P->dim, nbox[l], boxes[l], 1, P->dim, &H->nn[l]);
        H->f1[l] = (long long*)malloc(sizeof(long long) * (size_t)(nbox[l] * P->dim + 1));
        H->fp[l] = (long long*)malloc(sizeof(long long) * (size_t)(nbox[l] * P->dim + 1));
        cmp_face_layout(P->dim, nbox[l], boxes[l], 1, H->f1[l], &H->nf1[l]);
        cmp_face_layout(P->dim, nbox[l], boxes[l], H->npf, H->fp[l], &H->nfp[l]);
    }
}

static void hier_free(hier_t* H)
{
    for (int l = 0; l < H->nlev; ++l) {
        free(H->a1[l]); free(H->ap[l]); free(H->an[l]); free(H->f1[l]); free(H->fp[l]);
    }
}

/* FillPatch of an nc-component Phase-Field field (1 ghost) on every level; lay: its layouts */
static int hier_fill(const hier_t* H, cmp_box* const* lay, int nc, double* const* F)
{
    int err = ref_level_fill_pf(H->dim, &H->PG[0], H->nbox[0], lay[0], 1, nc, F[0]);
    for (int l = 1; l < H->nlev; ++l) {
        int e = fill_two_levels_nc(H->dim, nc, 1000, &H->PG[l], H->nbox[l], lay[l], 1, F[l], H->nbox[l], lay[l], 1,
                                   F[l], 0, &H->PG[l - 1], H->nbox[l - 1], lay[l - 1], 1, F[l - 1], H->ratio);
        if (e && !err) err = e;
    }
    return err;
}

/* AverageDown of an nc-component cell field (lay: 1-ghost layouts), finest level first */
static int hier_avgdown(const hier_t* H, const cmp_params* P, cmp_box* const* lay, int nc, double* const* F)
{
    cmp_params Pn = *P;
    Pn.nstate = nc;
    int err = 0;
    for (int l = H->nlev - 2; l >= 0; --l) {
        int e = ref_level_average_down(&Pn, H->nbox[l + 1], lay[l + 1], 1, F[l + 1], H->nbox[l], lay[l], 1, F[l],
                                       H->ratio);
        if (e && !err) err = e;
    }
    return err;
}

/* AverageDownFaces of an nc-component face buffer (face offsets foff per level), finest level
 * first: every coarse face of the coarsened fine boxes (and its periodic images) takes the mean of
 * the fine faces on it, the lower transverse index fastest, times 1/ratio^(dim-1). */
static void hier_avgdown_faces(const hier_t* H, long long* const* foff, int nc, double* const* Fb)
{
    const int dim = H->dim, ratio = H->ratio;
    int rt = 1;
    for (int e = 1; e < dim; ++e) rt *= ratio;
    const double facInv = 1.0 / (double)rt;
    for (int l = H->nlev - 2; l >= 0; --l) {
        const cmp_geometry* Gc = &H->PG[l];
        const int nbf = H->nbox[l + 1], nbc = H->nbox[l];
        const cmp_box* fb = H->boxes[l + 1];
        const cmp_box* cb = H->boxes[l];
        for (int b = 0; b < nbf; ++b)
            for (int d = 0; d < dim; ++d) {
                int ffe[3];
                face_extents(dim, &fb[b], d, ffe);
                const long long nff = (long long)ffe[0] * ffe[1] * ffe[2];
                const double* Ff = Fb[l + 1] + foff[l + 1][b * dim + d];
                int clo[3] = {0, 0, 0}, chi[3] = {0, 0, 0};
                for (int e = 0; e < dim; ++e) {
                    clo[e] = fb[b].lo[e] / ratio;
                    chi[e] = (fb[b].hi[e] + 1) / ratio - 1 + (e == d ? 1 : 0);
                }
                int t1 = -1, t2 = -1;
                for (int e = 0; e < dim; ++e)
                    if (e != d) { if (t1 < 0) t1 = e; else t2 = e; }
                const int r1 = (t1 >= 0) ? ratio : 1, r2 = (t2 >= 0) ? ratio : 1;
                for (int K = clo[2]; K <= chi[2]; ++K)
                    for (int J = clo[1]; J <= chi[1]; ++J)
                        for (int I = clo[0]; I <= chi[0]; ++I) {
                            const int C[3] = {I, J, K};
                            double avg[REF_NMAX];
                            for (int n = 0; n < nc; ++n) {
                                double c = 0.0;
                                for (int o2 = 0; o2 < r2; ++o2)
                                    for (int o1 = 0; o1 < r1; ++o1) {
                                        int fcell[3] = {0, 0, 0};
                                        for (int e = 0; e < dim; ++e) fcell[e] = C[e] * ratio;
                                        if (t1 >= 0) fcell[t1] += o1;
                                        if (t2 >= 0) fcell[t2] += o2;
                                        long long q = 0;
                                        for (int e = dim - 1; e >= 0; --e) q = q * ffe[e] + (fcell[e] - fb[b].lo[e]);
                                        c += Ff[q + n * nff];
                                    }
                                avg[n] = c * facInv;
                            }
                            int nimg = 1;
                            int pos[3] = {C[d], C[d], C[d]};
                            const int ncd = Gc->domain_hi[d] - Gc->domain_lo[d] + 1;
                            if (Gc->is_periodic[d]) {
                                if (C[d] == Gc->domain_lo[d]) pos[nimg++] = C[d] + ncd;
                                else if (C[d] == Gc->domain_hi[d] + 1) pos[nimg++] = C[d] - ncd;
                            }
                            for (int m = 0; m < nimg; ++m) {
                                int F3[3] = {C[0], C[1], C[2]};
                                F3[d] = pos[m];
                                for (int q = 0; q < nbc; ++q) {
                                    int cfe[3];
                                    face_extents(dim, &cb[q], d, cfe);
                                    int in = 1;
                                    for (int e = 0; e < dim; ++e)
                                        if (F3[e] < cb[q].lo[e] || F3[e] > cb[q].lo[e] + cfe[e] - 1) in = 0;
                                    if (!in) continue;
                                    const long long ncf = (long long)cfe[0] * cfe[1] * cfe[2];
                                    long long qq = 0;
                                    for (int e = dim - 1; e >= 0; --e) qq = qq * cfe[e] + (F3[e] - cb[q].lo[e]);
                                    double* Fc = Fb[l] + foff[l][q * dim + d];
                                    for (int n = 0; n < nc; ++n) Fc[qq + n * ncf] = avg[n];
                                }
                            }
                        }
            }
    }
}

/* The box of level l holding cell c, after the periodic wrap (-1: none; *outside = 1 when c lies
 * beyond a non-periodic domain face). c is wrapped in place. */
static int hier_owner(const hier_t* H, int l, int* c, int* outside)
{
    *outside = 0;
    const cmp_geometry* G = &H->PG[l];
    for (int d = 0; d < H->dim; ++d) {
        const int lo = G->domain_lo[d], hi = G->domain_hi[d], n = hi - lo + 1;
        if (c[d] < lo || c[d] > hi) {
            if (!G->is_periodic[d]) { *outside = 1; return -1; }
            c[d] = (c[d] < lo) ? c[d] + n : c[d] - n;
        }
    }
    return find_box(H->dim, H->nbox[l], H->boxes[l], c);
}

static int hier_covered(const hier_t* H, int l, const int* c)
{
    if (l + 1 >= H->nlev) return 0;
    int f[3] = {0, 0, 0};
    for (int d = 0; d < H->dim; ++d) f[d] = c[d] * H->ratio;
    return find_box(H->dim, H->nbox[l + 1], H->boxes[l + 1], f) >= 0;
}

/* index (1-comp, 1-ghost layout) of cell c of box b of level l */
static long long hier_idx(const hier_t* H, int l, int b, const int* c)
{
    fab_t f = make_fab(H->dim, &H->a1[l][b], 1);
    return H->a1[l][b].offset + fab_idx(&f, c[0], c[1], c[2]);
}

/* the face index (1-comp face buffer) of the face of box b of level l between cell c - e_d and c */
static long long hier_face(const hier_t* H, int l, int b, int d, const int* c)
{
    int fe[3];
    face_extents(H->dim, &H->boxes[l][b], d, fe);
    long long q = 0;
    for (int e = H->dim - 1; e >= 0; --e) q = q * fe[e] + (c[e] - H->boxes[l][b].lo[e]);
    return H->f1[l][b * H->dim + d] + q;
}

/* One link of composite cell (l, b, c) across its side s of direction d: up to ratio^(dim-1)
 * neighbours (level, box, cell) with their transmissibilities b area / distance. Returns the count. */
typedef struct { int l, b, c[3]; double T; } link_t;

static int hier_links(const hier_t* H, double* const* Bf, int l, int b, const int* c, int d, int s, link_t* out)
{
    const int dim = H->dim, ratio = H->ratio;
    const cmp_geometry* G = &H->PG[l];
    int x[3] = {c[0], c[1], c[2]};
    x[d] += s ? 1 : -1;
    int outside;
    const int nb = hier_owner(H, l, x, &outside);
    if (outside) return 0;
    double area = 1.0;
    for (int e = 0; e < dim; ++e)
        if (e != d) area *= G->dx[e];
    if (nb >= 0 && !hier_covered(H, l, x)) {
        int fc[3] = {c[0], c[1], c[2]};
        fc[d] += s;
        out[0].l = l; out[0].b = nb;
        for (int e = 0; e < 3; ++e) out[0].c[e] = x[e];
        out[0].T = Bf[l][hier_face(H, l, b, d, fc)] * area / G->dx[d];
        return 1;
    }
    if (nb >= 0) {
        /* the finer level covers x: the fine cells next to this face, each with its fine face */
        const cmp_geometry* Gf = &H->PG[l + 1];
        double farea = 1.0;
        for (int e = 0; e < dim; ++e)
            if (e != d) farea *= Gf->dx[e];
        const double dist = 0.5 * (Gf->dx[d] + G->dx[d]);
        int t1 = -1, t2 = -1;
        for (int e = 0; e < dim; ++e)
            if (e != d) { if (t1 < 0) t1 = e; else t2 = e; }
        const int r1 = (t1 >= 0) ? ratio : 1, r2 = (t2 >= 0) ? ratio : 1;
        int n = 0;
        for (int o2 = 0; o2 < r2; ++o2)
            for (int o1 = 0; o1 < r1; ++o1) {
                int f[3] = {0, 0, 0}, ff[3];
                for (int e = 0; e < dim; ++e) f[e] = x[e] * ratio;
                f[d] = s ? x[d] * ratio : (x[d] + 1) * ratio - 1;
                if (t1 >= 0) f[t1] = c[t1] * ratio + o1;
                if (t2 >= 0) f[t2] = c[t2] * ratio + o2;
                for (int e = 0; e < 3; ++e) ff[e] = f[e];
                ff[d] += s ? 0 : 1;       /* the fine face between f and the coarse side */
                const int fb = find_box(dim, H->nbox[l + 1], H->boxes[l + 1], f);
                out[n].l = l + 1; out[n].b = fb;
                for (int e = 0; e < 3; ++e) out[n].c[e] = f[e];
                out[n].T = Bf[l + 1][hier_face(H, l + 1, fb, d, ff)] * farea / dist;
                ++n;
            }
        return n;
    }
    /* x is not on level l: the coarse cell holding it, through this box's face */
    const cmp_geometry* Gc = &H->PG[l - 1];
    int p[3] = {0, 0, 0};
    for (int e = 0; e < dim; ++e) p[e] = floor_div(x[e], ratio);
    int po;
    const int pb = hier_owner(H, l - 1, p, &po);
    int fc[3] = {c[0], c[1], c[2]};
    fc[d] += s;
    out[0].l = l - 1; out[0].b = pb;
    for (int e = 0; e < 3; ++e) out[0].c[e] = p[e];
    out[0].T = Bf[l][hier_face(H, l, b, d, fc)] * area / (0.5 * (G->dx[d] + Gc->dx[d]));
    return 1;
}

/* y = A x (volume weighted, A = -div(b grad .)) on the composite cells; diag (or NULL): its diagonal */
static void hier_apply(const hier_t* H, double* const* Bf, double* const* X, double* const* Y, double* const* Dg)
{
    const int dim = H->dim;
    for (int l = 0; l < H->nlev; ++l)
        for (int b = 0; b < H->nbox[l]; ++b) {
            const cmp_box* B = &H->boxes[l][b];
            const int lo2 = (dim == 3) ? B->lo[2] : 0, hi2 = (dim == 3) ? B->hi[2] : 0;
            for (int k = lo2; k <= hi2; ++k)
                for (int j = B->lo[1]; j <= B->hi[1]; ++j)
                    for (int i = B->lo[0]; i <= B->hi[0]; ++i) {
                        const int c[3] = {i, j, k};
                        const long long ci = hier_idx(H, l, b, c);
                        if (hier_covered(H, l, c)) { if (Y) Y[l][ci] = 0.0; if (Dg) Dg[l][ci] = 0.0; continue; }
                        double acc = 0.0, dsum = 0.0;
                        for (int d = 0; d < dim; ++d)
                            for (int s = 0; s < 2; ++s) {
                                link_t lk[4];
                                const int n = hier_links(H, Bf, l, b, c, d, s, lk);
                                for (int q = 0; q < n; ++q) {
                                    dsum += lk[q].T;
                                    if (X) acc += lk[q].T * (X[l][ci] - X[lk[q].l][hier_idx(H, lk[q].l, lk[q].b, lk[q].c)]);
                                }
                            }
                        if (Y) Y[l][ci] = acc;
                        if (Dg) Dg[l][ci] = dsum;
                    }
        }
}

static double hier_vol(const hier_t* H, int l)
{
    double v = 1.0;
    for (int d = 0; d < H->dim; ++d) v *= H->PG[l].dx[d];
    return v;
}

/* sum over composite cells of w(l) a b (a NULL: of w(l) b) */
static double hier_dot(const hier_t* H, double* const* A, double* const* B, int weighted)
{
    const int dim = H->dim;
    double s = 0.0;
    for (int l = 0; l < H->nlev; ++l) {
        const double w = weighted ? hier_vol(H, l) : 1.0;
        for (int b = 0; b < H->nbox[l]; ++b) {
            const cmp_box* Bx = &H->boxes[l][b];
            const int lo2 = (dim == 3) ? Bx->lo[2] : 0, hi2 = (dim == 3) ? Bx->hi[2] : 0;
            for (int k = lo2; k <= hi2; ++k)
                for (int j = Bx->lo[1]; j <= Bx->hi[1]; ++j)
                    for (int i = Bx->lo[0]; i <= Bx->hi[0]; ++i) {
                        const int c[3] = {i, j, k};
                        if (hier_covered(H, l, c)) continue;
                        const long long ci = hier_idx(H, l, b, c);
                        s += w * (A ? A[l][ci] : 1.0) * B[l][ci];
                    }
        }
    }
    return s;
}

/* The composite solve of div(b grad Q) = rhs (rhs and Q: 1-comp 1-ghost layouts of every level;
 * Bf: b on every face of every level, the coarse faces under a finer level averaged down), then
 * Grad_Q on every face of every level (averaged down). *iters: CG iterations. */
static int hier_solve(const hier_t* H, const cmp_params* P, double* const* Bf, double* const* rhs, double* const* Gq,
                      int* iters)
{
    const int dim = H->dim, nlev = H->nlev;
    double *X[AMR_NLEV_MAX] = {NULL}, *R[AMR_NLEV_MAX] = {NULL}, *Pv[AMR_NLEV_MAX] = {NULL};
    double *AP[AMR_NLEV_MAX] = {NULL}, *ID[AMR_NLEV_MAX] = {NULL};
    for (int l = 0; l < nlev; ++l) {
        X[l] = (double*)calloc((size_t)H->n1[l], sizeof(double));
        R[l] = (double*)calloc((size_t)H->n1[l], sizeof(double));
        Pv[l] = (double*)calloc((size_t)H->n1[l], sizeof(double));
        AP[l] = (double*)calloc((size_t)H->n1[l], sizeof(double));
        ID[l] = (double*)calloc((size_t)H->n1[l], sizeof(double));
    }
    hier_apply(H, Bf, NULL, NULL, ID);
    /* volume-weighted mean of rhs; r = V (-(rhs - mean)), the composite cells only */
    double sv = 0.0, srhs = 0.0;
    for (int l = 0; l < nlev; ++l) {
        const double v = hier_vol(H, l);
        for (int b = 0; b < H->nbox[l]; ++b) {
            const cmp_box* Bx = &H->boxes[l][b];
            const int lo2 = (dim == 3) ? Bx->lo[2] : 0, hi2 = (dim == 3) ? Bx->hi[2] : 0;
            for (int k = lo2; k <= hi2; ++k)
                for (int j = Bx->lo[1]; j <= Bx->hi[1]; ++j)
                    for (int i = Bx->lo[0]; i <= Bx->hi[0]; ++i) {
                        const int c[3] = {i, j, k};
                        if (hier_covered(H, l, c)) continue;
                        sv += v;
                        srhs += v * rhs[l][hier_idx(H, l, b, c)];
                    }
        }
    }
    const double mean = srhs / sv;
    double rr = 0.0, rz = 0.0;
    for (int l = 0; l < nlev; ++l) {
        const double v = hier_vol(H, l);
        for (int b = 0; b < H->nbox[l]; ++b) {
            const cmp_box* Bx = &H->boxes[l][b];
            const int lo2 = (dim == 3) ? Bx->lo[2] : 0, hi2 = (dim == 3) ? Bx->hi[2] : 0;
            for (int k = lo2; k <= hi2; ++k)
                for (int j = Bx->lo[1]; j <= Bx->hi[1]; ++j)
                    for (int i = Bx->lo[0]; i <= Bx->hi[0]; ++i) {
                        const int c[3] = {i, j, k};
                        if (hier_covered(H, l, c)) continue;
                        const long long ci = hier_idx(H, l, b, c);
                        const double dg = ID[l][ci];
                        ID[l][ci] = (dg > 0.0) ? 1.0 / dg : 0.0;
                        R[l][ci] = -(rhs[l][ci] - mean) * v;
                        Pv[l][ci] = ID[l][ci] * R[l][ci];
                        rr += R[l][ci] * R[l][ci];
                        rz += R[l][ci] * Pv[l][ci];
                    }
        }
    }
    const double b2 = rr;
    double best = rr;
    int it = 0, stall = 0;
    while (rr > 1.0e-28 * b2 && it < 100000 && b2 > 0.0) {
        hier_apply(H, Bf, Pv, AP, NULL);
        const double pAp = hier_dot(H, Pv, AP, 0);
        if (!(pAp > 0.0)) break;
        const double al = rz / pAp;
        double rr2 = 0.0, rz2 = 0.0;
        for (int l = 0; l < nlev; ++l)
            for (long long q = 0; q < H->n1[l]; ++q) {
                X[l][q] += al * Pv[l][q];
                R[l][q] -= al * AP[l][q];
            }
        rr2 = hier_dot(H, R, R, 0);
        for (int l = 0; l < nlev; ++l)
            for (long long q = 0; q < H->n1[l]; ++q) AP[l][q] = ID[l][q] * R[l][q];
        rz2 = hier_dot(H, R, AP, 0);
        const double be = rz2 / rz;
        for (int l = 0; l < nlev; ++l)
            for (long long q = 0; q < H->n1[l]; ++q) Pv[l][q] = AP[l][q] + be * Pv[l][q];
        rr = rr2;
        rz = rz2;
        ++it;
        if (rr < best) { best = rr; stall = 0; }
        else if (++stall > 400) break;
    }
    /* covered cells: the mean of their children (finest first) */
    int err = hier_avgdown(H, P, H->a1, 1, X);
    /* Grad_Q on every face of every level */
    for (int l = 0; l < nlev; ++l) {
        const cmp_geometry* G = &H->PG[l];
        for (int b = 0; b < H->nbox[l]; ++b) {
            const cmp_box* Bx = &H->boxes[l][b];
            for (int d = 0; d < dim; ++d) {
                int fe[3];
                face_extents(dim, Bx, d, fe);
                const int lo2 = (dim == 3) ? Bx->lo[2] : 0;
                for (int kk = 0; kk < fe[2]; ++kk)
                    for (int jj = 0; jj < fe[1]; ++jj)
                        for (int ii = 0; ii < fe[0]; ++ii) {
                            int c[3] = {Bx->lo[0] + ii, Bx->lo[1] + jj, lo2 + kk};
                            const long long fi = hier_face(H, l, b, d, c);
                            const int cd = c[d];
                            if (!G->is_periodic[d] && (cd == G->domain_lo[d] || cd == G->domain_hi[d] + 1)) {
                                Gq[l][fi] = 0.0;
                                continue;
                            }
                            int m[3] = {c[0], c[1], c[2]};
                            m[d] -= 1;
                            int oc, om;
                            int cc[3] = {c[0], c[1], c[2]};
                            const int bc = hier_owner(H, l, cc, &oc);
                            const int bm = hier_owner(H, l, m, &om);
                            double v;
                            if (bc >= 0 && bm >= 0) {
                                v = (X[l][hier_idx(H, l, bc, cc)] - X[l][hier_idx(H, l, bm, m)]) / G->dx[d];
                            } else {
                                /* a coarse-fine face of this box: the coarse cell beyond it */
                                const int in_c = (bc >= 0);
                                int o[3] = {in_c ? m[0] : cc[0], in_c ? m[1] : cc[1], in_c ? m[2] : cc[2]};
                                int p[3] = {0, 0, 0}, po;
                                for (int e = 0; e < dim; ++e) p[e] = floor_div(o[e], H->ratio);
                                const int pb = hier_owner(H, l - 1, p, &po);
                                const double xp = X[l - 1][hier_idx(H, l - 1, pb, p)];
                                const double dist = 0.5 * (G->dx[d] + H->PG[l - 1].dx[d]);
                                v = in_c ? (X[l][hier_idx(H, l, bc, cc)] - xp) / dist
                                         : (xp - X[l][hier_idx(H, l, bm, m)]) / dist;
                            }
                            Gq[l][fi] = v;
                        }
            }
        }
    }
    hier_avgdown_faces(H, H->f1, 1, Gq);
    if (iters) *iters = it;
    for (int l = 0; l < nlev; ++l) { free(X[l]); free(R[l]); free(Pv[l]); free(AP[l]); free(ID[l]); }
    return err;
}

/* the cell normals (component d of every valid cell, then the dim-component field's own FillPatch)
 * of component ph of the npf-component field F (central differences, 0 where the gradient is 0;
 * FDM_Normal_CenterFromCenter_D.H:12-63) on every level */
static int hier_normals(const hier_t* H, double* const* F, int ph, double* const* Nrm)
{
    const int dim = H->dim;
    for (int l = 0; l < H->nlev; ++l) {
        const cmp_geometry* G = &H->PG[l];
        for (int b = 0; b < H->nbox[l]; ++b) {
            fab_t fa = make_fab(dim, &H->ap[l][b], 1);
            fab_t fn = make_fab(dim, &H->an[l][b], 1);
            const double* Ab = F[l] + H->ap[l][b].offset + ph * fa.sn;
            double* Nb = Nrm[l] + H->an[l][b].offset;
            const cmp_box* B = &H->boxes[l][b];
            const int lo2 = (dim == 3) ? B->lo[2] : 0, hi2 = (dim == 3) ? B->hi[2] : 0;
            for (int k = lo2; k <= hi2; ++k)
                for (int j = B->lo[1]; j <= B->hi[1]; ++j)
                    for (int i = B->lo[0]; i <= B->hi[0]; ++i) {
                        const int c3[3] = {i, j, k};
                        const long long c = fab_idx(&fa, i, j, k);
                        double g[3], sum = 0.0;
                        for (int d = 0; d < dim; ++d) {
                            int m3[3] = {i, j, k}, p3[3] = {i, j, k};
                            m3[d] -= 1; p3[d] += 1;
                            const double gm = (Ab[c] - Ab[fab_idx(&fa, m3[0], m3[1], m3[2])]) / G->dx[d];
                            const double gp = (Ab[fab_idx(&fa, p3[0], p3[1], p3[2])] - Ab[c]) / G->dx[d];
                            g[d] = 0.5 * gm + 0.5 * gp;
                            sum += g[d] * g[d];
                        }
                        (void)c3;
                        const long long cn = fab_idx(&fn, i, j, k);
                        for (int d = 0; d < dim; ++d) Nb[cn + d * fn.sn] = (sum == 0.0) ? 0.0 : g[d] / sqrt(sum);
                    }
        }
    }
    return hier_fill(H, H->an, dim, Nrm);
}

/* the step's volume fractions (every phase) of the flow states U (ng ghosts) into A0 (npf comps,
 * 1 ghost), filled */
static int hier_extract(const hier_t* H, const cmp_params* P, int ng, double* const* U, double* const* A0)
{
    const int dim = H->dim;
    for (int l = 0; l < H->nlev; ++l)
        for (int b = 0; b < H->nbox[l]; ++b) {
            const cmp_box* B = &H->boxes[l][b];
            fab_t fu = make_fab(dim, B, ng);
            fab_t fa = make_fab(dim, &H->ap[l][b], 1);
            const int lo2 = (dim == 3) ? B->lo[2] : 0, hi2 = (dim == 3) ? B->hi[2] : 0;
            for (int ph = 0; ph < H->npf; ++ph) {
                const int comp = P->nphase ? P->nphase + dim + 1 + ph : 3 + dim;
                for (int k = lo2; k <= hi2; ++k)
                    for (int j = B->lo[1]; j <= B->hi[1]; ++j)
                        for (int i = B->lo[0]; i <= B->hi[0]; ++i)
                            A0[l][H->ap[l][b].offset + fab_idx(&fa, i, j, k) + ph * fa.sn] =
                                U[l][B->offset + fab_idx(&fu, i, j, k) + comp * fu.sn];
            }
        }
    return hier_fill(H, H->ap, H->npf, A0);
}

/* b = W(phi_f) + 1e-3 on every face of every level (phi_f the face mean of phase ph of the npf-comp
 * field A, its ghosts filled, re-sharpened through the mappings at eta and eta_factor eta;
 * :2205-2274), then AverageDownFaces(b) */
static void hier_bcoef(const hier_t* H, const cmp_pf_params* PF, double* const* A, int ph, double eta, double* const* Bf)
{
    const int dim = H->dim;
    for (int l = 0; l < H->nlev; ++l)
        for (int b = 0; b < H->nbox[l]; ++b) {
            fab_t fa = make_fab(dim, &H->ap[l][b], 1);
            const cmp_box* B = &H->boxes[l][b];
            const int lo2 = (dim == 3) ? B->lo[2] : 0;
            const double* Ab = A[l] + H->ap[l][b].offset + ph * fa.sn;
            for (int d = 0; d < dim; ++d) {
                int fe[3];
                face_extents(dim, B, d, fe);
                for (int kk = 0; kk < fe[2]; ++kk)
                    for (int jj = 0; jj < fe[1]; ++jj)
                        for (int ii = 0; ii < fe[0]; ++ii) {
                            int c3[3] = {B->lo[0] + ii, B->lo[1] + jj, lo2 + kk};
                            int m3[3] = {c3[0], c3[1], c3[2]};
                            m3[d] -= 1;
                            const double c = ref_pf_map_forward(
                                ref_pf_map_backward(Ab[fab_idx(&fa, c3[0], c3[1], c3[2])], eta), PF->eta_factor * eta, 0);
                            const double m = ref_pf_map_forward(
                                ref_pf_map_backward(Ab[fab_idx(&fa, m3[0], m3[1], m3[2])], eta), PF->eta_factor * eta, 0);
                            const double pf = 0.5 * m + 0.5 * c;
                            const long long q = ii + (long long)fe[0] * (jj + (long long)fe[1] * kk);
                            Bf[l][H->f1[l][b * dim + d] + q] = pf * (1.0 - pf) + 1.0e-3;
                        }
            }
        }
    hier_avgdown_faces(H, H->f1, 1, Bf);
}

/* The composite implicit solve alone, for a two-phase alpha (1 comp, 1 ghost, its ghosts filled
 * here) and a right-hand side (1 comp, 1 ghost layout) on every level: Grad_Q (1 comp face buffers)
 * and the face coefficients b it used (averaged down). */
int ref_amr_pf_gradq_implicit(const cmp_params* P, const cmp_pf_params* PF, int nlev, const cmp_geometry* G,
                              const int* nbox, const cmp_box* const* boxes, double* const* alpha,
                              double* const* rhs, double eta, int ratio, double* const* gq, double* const* bf,
                              int* iters)
{
    if (nlev < 1 || nlev > AMR_NLEV_MAX || P->nphase) return CMP_ERR_ARG;
    hier_t H;
    hier_init(&H, P, nlev, G, nbox, boxes, ratio);
    int err = hier_fill(&H, H.ap, 1, alpha);
    hier_bcoef(&H, PF, alpha, 0, eta, bf);
    int e = hier_solve(&H, P, bf, rhs, gq, iters);
    if (e && !err) err = e;
    hier_free(&H);
    return err;
}

/* Grad_Q of every level (face buffers gq[l] of cmp_face_layout with npf comps) for a flow step of
 * size dt from the hierarchy's states U; info: max|u|, mobility, eta, mechanism sub-steps, CG
 * iterations of the last implicit solve. */
int ref_amr_pf_gradq_step(const cmp_params* P, const cmp_pf_params* PF, int nlev, const cmp_geometry* G,
                          const int* nbox, const cmp_box* const* boxes, int ng, double* const* U, double dt,
                          int ratio, double* const* gq, double* info)
{
    if (nlev < 1 || nlev > AMR_NLEV_MAX) return CMP_ERR_ARG;
    const int dim = P->dim;
    hier_t H;
    hier_init(&H, P, nlev, G, nbox, boxes, ratio);
    const int npf = H.npf;
    const int cac = (PF->mechanism == CMP_PF_CAC);
    const int explicit_rc = (PF->explicit_rc == 1);
    int err = 0;
    /* MaxVelocity over the levels; Lambda, eta (finest cell size), mobility */
    double maxv = 0.0;
    for (int l = 0; l < nlev; ++l) {
        const double v = ref_level_max_velocity(P, nbox[l], boxes[l], ng, U[l]);
        if (maxv < v || isnan(v)) maxv = v;
    }
    double lambda, eta, mob;
    ref_pf_coefficients(PF, &G[nlev - 1], maxv, &lambda, &eta, &mob);
    double *A0[AMR_NLEV_MAX] = {NULL}, *A[AMR_NLEV_MAX] = {NULL}, *Ps[AMR_NLEV_MAX] = {NULL};
    double *Nrm[AMR_NLEV_MAX] = {NULL}, *J[AMR_NLEV_MAX] = {NULL};
    double *Jh[AMR_NLEV_MAX] = {NULL}, *Mk[AMR_NLEV_MAX] = {NULL}, *Wk[AMR_NLEV_MAX] = {NULL};
    for (int l = 0; l < nlev; ++l) {
        A0[l] = (double*)calloc((size_t)H.np[l], sizeof(double));
        A[l] = (double*)calloc((size_t)H.np[l], sizeof(double));
        Ps[l] = (double*)calloc((size_t)H.np[l], sizeof(double));
        Nrm[l] = (double*)calloc((size_t)H.nn[l], sizeof(double));
        J[l] = (double*)calloc((size_t)H.nfp[l], sizeof(double));
        Jh[l] = (double*)calloc((size_t)H.nfp[l], sizeof(double));
        Mk[l] = (double*)calloc((size_t)H.n1[l], sizeof(double));
        Wk[l] = (double*)calloc((size_t)H.n1[l], sizeof(double));
    }
    int e = hier_extract(&H, P, ng, U, A0);
    if (e && !err) err = e;
    for (int l = 0; l < nlev; ++l) memcpy(A[l], A0[l], sizeof(double) * (size_t)H.np[l]);
    /* PhaseFieldMechanism_Evolve */
    const double* dxf = G[nlev - 1].dx;
    double sdx = 1.0 / dxf[0] / dxf[0];
    for (int d = 1; d < dim; ++d) sdx = sdx + 1.0 / dxf[d] / dxf[d];
    double dt_m = 0.9 / (2.0 * mob * lambda) / (sdx + 1.0 / eta / eta);
    double t = 0.0;
    const double T = dt;
    int nsub = 0;
    while (t < T) {
        if (t + dt_m > T) dt_m = T - t;
        e = hier_fill(&H, H.ap, npf, A);
        if (e && !err) err = e;
        if (cac) {
            for (int ph = 0; ph < npf; ++ph) {
                e = hier_normals(&H, A, ph, Nrm);
                if (e && !err) err = e;
                for (int l = 0; l < nlev; ++l)
                    for (int b = 0; b < nbox[l]; ++b) {
                        fab_t fa = make_fab(dim, &H.ap[l][b], 1);
                        fab_t fn = make_fab(dim, &H.an[l][b], 1);
                        fab_t f1 = make_fab(dim, &H.a1[l][b], 1);
                        const double* Ab = A[l] + H.ap[l][b].offset + ph * fa.sn;
                        const double* Nb = Nrm[l] + H.an[l][b].offset;
                        const cmp_box* B = &boxes[l][b];
                        const int lo2 = (dim == 3) ? B->lo[2] : 0, hi2 = (dim == 3) ? B->hi[2] : 0;
                        for (int k = lo2; k <= hi2; ++k)
                            for (int j = B->lo[1]; j <= B->hi[1]; ++j)
                                for (int i = B->lo[0]; i <= B->hi[0]; ++i) {
                                    const long long c = fab_idx(&fa, i, j, k);
                                    double fm[3], fp[3], nc[3], nm[3], np_[3];
                                    for (int d = 0; d < dim; ++d) {
                                        int m3[3] = {i, j, k}, p3[3] = {i, j, k};
                                        m3[d] -= 1; p3[d] += 1;
                                        fm[d] = Ab[fab_idx(&fa, m3[0], m3[1], m3[2])];
                                        fp[d] = Ab[fab_idx(&fa, p3[0], p3[1], p3[2])];
                                        nc[d] = Nb[fab_idx(&fn, i, j, k) + d * fn.sn];
                                        nm[d] = Nb[fab_idx(&fn, m3[0], m3[1], m3[2]) + d * fn.sn];
                                        np_[d] = Nb[fab_idx(&fn, p3[0], p3[1], p3[2]) + d * fn.sn];
                                    }
                                    const long long c1 = H.a1[l][b].offset + fab_idx(&f1, i, j, k);
                                    Mk[l][c1] = ref_pf_cac_mech(dim, G[l].dx, Ab[c], fm, fp, nc, nm, np_, mob, lambda,
                                                                eta);
                                    Wk[l][c1] = Ab[c] * (1.0 - Ab[c]);
                                }
                    }
                e = hier_avgdown(&H, P, H.a1, 1, Mk);
                if (e && !err) err = e;
                e = hier_avgdown(&H, P, H.a1, 1, Wk);
                if (e && !err) err = e;
                double sm = 0.0, sw = 0.0;
                for (int b = 0; b < nbox[0]; ++b) {
                    fab_t f1 = make_fab(dim, &H.a1[0][b], 1);
                    const cmp_box* B = &boxes[0][b];
                    const int lo2 = (dim == 3) ? B->lo[2] : 0, hi2 = (dim == 3) ? B->hi[2] : 0;
                    for (int k = lo2; k <= hi2; ++k)
                        for (int j = B->lo[1]; j <= B->hi[1]; ++j)
                            for (int i = B->lo[0]; i <= B->hi[0]; ++i) {
                                const long long c1 = H.a1[0][b].offset + fab_idx(&f1, i, j, k);
                                sm += Mk[0][c1];
                                sw += Wk[0][c1];
                            }
                }
                const double L = (sw == 0.0) ? 0.0 : -sm / sw;
                /* Div_J of this phase; alpha advances after every phase's Div_J (LinComb) */
                for (int l = 0; l < nlev; ++l)
                    for (long long q = 0; q < H.n1[l]; ++q) Mk[l][q] = Mk[l][q] + Wk[l][q] * L;
                for (int l = 0; l < nlev; ++l)
                    for (int b = 0; b < nbox[l]; ++b) {
                        fab_t fa = make_fab(dim, &H.ap[l][b], 1);
                        fab_t f1 = make_fab(dim, &H.a1[l][b], 1);
                        const cmp_box* B = &boxes[l][b];
                        const int lo2 = (dim == 3) ? B->lo[2] : 0, hi2 = (dim == 3) ? B->hi[2] : 0;
                        for (int k = lo2; k <= hi2; ++k)
                            for (int j = B->lo[1]; j <= B->hi[1]; ++j)
                                for (int i = B->lo[0]; i <= B->hi[0]; ++i) {
                                    /* the phase's Div_J into Ps (a holding field of npf comps) */
                                    Ps[l][H.ap[l][b].offset + fab_idx(&fa, i, j, k) + ph * fa.sn] =
                                        Mk[l][H.a1[l][b].offset + fab_idx(&f1, i, j, k)];
                                }
                    }
            }
            for (int l = 0; l < nlev; ++l)
                for (int b = 0; b < nbox[l]; ++b) {
                    fab_t fa = make_fab(dim, &H.ap[l][b], 1);
                    const cmp_box* B = &boxes[l][b];
                    const int lo2 = (dim == 3) ? B->lo[2] : 0, hi2 = (dim == 3) ? B->hi[2] : 0;
                    for (int ph = 0; ph < npf; ++ph)
                        for (int k = lo2; k <= hi2; ++k)
                            for (int j = B->lo[1]; j <= B->hi[1]; ++j)
                                for (int i = B->lo[0]; i <= B->hi[0]; ++i) {
                                    const long long c = H.ap[l][b].offset + fab_idx(&fa, i, j, k) + ph * fa.sn;
                                    A[l][c] = 1.0 * A[l][c] + dt_m * Ps[l][c];
                                }
                }
        } else {
            /* psi of the valid cells, then its own FillPatch (LSBorder) */
            for (int l = 0; l < nlev; ++l)
                for (int b = 0; b < nbox[l]; ++b) {
                    fab_t fa = make_fab(dim, &H.ap[l][b], 1);
                    const cmp_box* B = &boxes[l][b];
                    const int lo2 = (dim == 3) ? B->lo[2] : 0, hi2 = (dim == 3) ? B->hi[2] : 0;
                    for (int ph = 0; ph < npf; ++ph)
                        for (int k = lo2; k <= hi2; ++k)
                            for (int j = B->lo[1]; j <= B->hi[1]; ++j)
                                for (int i = B->lo[0]; i <= B->hi[0]; ++i) {
                                    const long long c = H.ap[l][b].offset + fab_idx(&fa, i, j, k) + ph * fa.sn;
                                    Ps[l][c] = ref_pf_map_backward(A[l][c], eta);
                                }
                }
            e = hier_fill(&H, H.ap, npf, Ps);
            if (e && !err) err = e;
            for (int ph = 0; ph < npf; ++ph) {
                e = hier_normals(&H, Ps, ph, Nrm);
                if (e && !err) err = e;
                for (int l = 0; l < nlev; ++l)
                    for (int b = 0; b < nbox[l]; ++b) {
                        fab_t fa = make_fab(dim, &H.ap[l][b], 1);
                        fab_t fn = make_fab(dim, &H.an[l][b], 1);
                        const double* Ab = A[l] + H.ap[l][b].offset + ph * fa.sn;
                        const double* Nb = Nrm[l] + H.an[l][b].offset;
                        const cmp_box* B = &boxes[l][b];
                        const int lo2 = (dim == 3) ? B->lo[2] : 0;
                        for (int d = 0; d < dim; ++d) {
                            int fe[3];
                            face_extents(dim, B, d, fe);
                            const long long nface = (long long)fe[0] * fe[1] * fe[2];
                            double* Jd = J[l] + H.fp[l][b * dim + d] + ph * nface;
                            for (int kk = 0; kk < fe[2]; ++kk)
                                for (int jj = 0; jj < fe[1]; ++jj)
                                    for (int ii = 0; ii < fe[0]; ++ii) {
                                        int c3[3] = {B->lo[0] + ii, B->lo[1] + jj, lo2 + kk};
                                        int m3[3] = {c3[0], c3[1], c3[2]};
                                        m3[d] -= 1;
                                        const long long c = fab_idx(&fa, c3[0], c3[1], c3[2]);
                                        const long long m = fab_idx(&fa, m3[0], m3[1], m3[2]);
                                        const long long cn = fab_idx(&fn, c3[0], c3[1], c3[2]);
                                        const long long mn = fab_idx(&fn, m3[0], m3[1], m3[2]);
                                        const double mob_face = 0.5 * mob + 0.5 * mob;
                                        const long long q = ii + (long long)fe[0] * (jj + (long long)fe[1] * kk);
                                        Jd[q] = ref_pf_acdi_flux(Ab[m], Ab[c], Nb[mn + d * fn.sn], Nb[cn + d * fn.sn],
                                                                 mob_face, eta, G[l].dx[d]);
                                    }
                        }
                    }
            }
            hier_avgdown_faces(&H, H.fp, npf, J);
            for (int l = 0; l < nlev; ++l) {
                if (explicit_rc)
                    for (long long q = 0; q < H.nfp[l]; ++q) Jh[l][q] += dt_m * J[l][q];
                for (int b = 0; b < nbox[l]; ++b) {
                    fab_t fa = make_fab(dim, &H.ap[l][b], 1);
                    const cmp_box* B = &boxes[l][b];
                    int nv[3] = {1, 1, 1};
                    for (int d = 0; d < dim; ++d) nv[d] = B->hi[d] - B->lo[d] + 1;
                    const int lo2 = (dim == 3) ? B->lo[2] : 0;
                    for (int ph = 0; ph < npf; ++ph)
                        for (int kk = 0; kk < nv[2]; ++kk)
                            for (int jj = 0; jj < nv[1]; ++jj)
                                for (int ii = 0; ii < nv[0]; ++ii) {
                                    const int idx[3] = {ii, jj, kk};
                                    double div = 0.0;
                                    for (int d = 0; d < dim; ++d) {
                                        int fe[3] = {nv[0], nv[1], nv[2]};
                                        fe[d] += 1;
                                        const long long nface = (long long)fe[0] * fe[1] * fe[2];
                                        const double* Jd = J[l] + H.fp[l][b * dim + d] + ph * nface;
                                        const long long qm = idx[0] + (long long)fe[0] * (idx[1] + (long long)fe[1] * idx[2]);
                                        const long long qp = qm + (d == 0 ? 1 : (d == 1 ? fe[0] : (long long)fe[0] * fe[1]));
                                        const double term = (Jd[qp] - Jd[qm]) / G[l].dx[d];
                                        div = (d == 0) ? term : div + term;
                                    }
                                    const long long c = H.ap[l][b].offset +
                                                        fab_idx(&fa, B->lo[0] + ii, B->lo[1] + jj, lo2 + kk) + ph * fa.sn;
                                    A[l][c] = 1.0 * A[l][c] + dt_m * div;
                                }
                }
            }
        }
        t += dt_m;
        ++nsub;
    }
    /* Grad_Q */
    int iters = 0;
    if (explicit_rc) {
        const double sc = 1.0 / dt;
        for (int l = 0; l < nlev; ++l) {
            for (long long q = 0; q < H.nfp[l]; ++q) Jh[l][q] *= sc;
            for (int b = 0; b < nbox[l]; ++b) {
                fab_t fa = make_fab(dim, &H.ap[l][b], 1);
                const cmp_box* B = &boxes[l][b];
                const int lo2 = (dim == 3) ? B->lo[2] : 0;
                for (int d = 0; d < dim; ++d) {
                    int fe[3];
                    face_extents(dim, B, d, fe);
                    const long long nface = (long long)fe[0] * fe[1] * fe[2];
                    for (int ph = 0; ph < npf; ++ph) {
                        const double* Ab = A0[l] + H.ap[l][b].offset + ph * fa.sn;
                        const double* Hd = Jh[l] + H.fp[l][b * dim + d] + ph * nface;
                        double* Gd = gq[l] + H.fp[l][b * dim + d] + ph * nface;
                        for (int kk = 0; kk < fe[2]; ++kk)
                            for (int jj = 0; jj < fe[1]; ++jj)
                                for (int ii = 0; ii < fe[0]; ++ii) {
                                    int c3[3] = {B->lo[0] + ii, B->lo[1] + jj, lo2 + kk};
                                    int m3[3] = {c3[0], c3[1], c3[2]};
                                    m3[d] -= 1;
                                    const long long q = ii + (long long)fe[0] * (jj + (long long)fe[1] * kk);
                                    Gd[q] = ref_pf_gradq_face(Ab[fab_idx(&fa, m3[0], m3[1], m3[2])],
                                                              Ab[fab_idx(&fa, c3[0], c3[1], c3[2])], Hd[q], eta,
                                                              PF->eta_factor);
                                }
                    }
                }
            }
        }
    } else {
        double *Bf[AMR_NLEV_MAX], *Dj[AMR_NLEV_MAX], *Gx[AMR_NLEV_MAX];
        for (int l = 0; l < nlev; ++l) {
            Bf[l] = (double*)calloc((size_t)H.nf1[l], sizeof(double));
            Gx[l] = (double*)calloc((size_t)H.nf1[l], sizeof(double));
            Dj[l] = (double*)calloc((size_t)H.n1[l], sizeof(double));
        }
        for (int ph = 0; ph < npf; ++ph) {
            for (int l = 0; l < nlev; ++l)
                for (int b = 0; b < nbox[l]; ++b) {
                    fab_t fa = make_fab(dim, &H.ap[l][b], 1);
                    fab_t f1 = make_fab(dim, &H.a1[l][b], 1);
                    const cmp_box* B = &boxes[l][b];
                    const int lo2 = (dim == 3) ? B->lo[2] : 0, hi2 = (dim == 3) ? B->hi[2] : 0;
                    const double* Ab = A0[l] + H.ap[l][b].offset + ph * fa.sn;
                    /* Div_Jhat = LinComb(1/dt, PF_new, -1/dt, PF) (:703) */
                    for (int k = lo2; k <= hi2; ++k)
                        for (int j = B->lo[1]; j <= B->hi[1]; ++j)
                            for (int i = B->lo[0]; i <= B->hi[0]; ++i) {
                                const long long c = H.ap[l][b].offset + fab_idx(&fa, i, j, k) + ph * fa.sn;
                                Dj[l][H.a1[l][b].offset + fab_idx(&f1, i, j, k)] =
                                    (1.0 / dt) * A[l][c] + (-1.0 / dt) * A0[l][c];
                            }
                    (void)Ab;
                }
            hier_bcoef(&H, PF, A0, ph, eta, Bf);
            e = hier_solve(&H, P, Bf, Dj, Gx, &iters);
            if (e && !err) err = e;
            /* phase ph's component of Grad_Q */
            for (int l = 0; l < nlev; ++l)
                for (int b = 0; b < nbox[l]; ++b)
                    for (int d = 0; d < dim; ++d) {
                        int fe[3];
                        face_extents(dim, &boxes[l][b], d, fe);
                        const long long nface = (long long)fe[0] * fe[1] * fe[2];
                        for (long long q = 0; q < nface; ++q)
                            gq[l][H.fp[l][b * dim + d] + ph * nface + q] = Gx[l][H.f1[l][b * dim + d] + q];
                    }
        }
        for (int l = 0; l < nlev; ++l) { free(Bf[l]); free(Gx[l]); free(Dj[l]); }
    }
    if (info) {
        info[0] = maxv;
        info[1] = mob;
        info[2] = eta;
        info[3] = (double)nsub;
        info[4] = (double)iters;
    }
    for (int l = 0; l < nlev; ++l) {
        free(A0[l]); free(A[l]); free(Ps[l]); free(Nrm[l]); free(J[l]); free(Jh[l]); free(Mk[l]); free(Wk[l]);
    }
    hier_free(&H);
    return err;
}

/* One stage-coupled TVD-RK3 sub-step of the Phase-Field right-hand side over the levels
 * (PhaseField_AdvanceAllLevels -> AdvanceAllLevelsStageCoupled with Grad_Q,
 * Class_Compressible_PhaseField_StageCoupling.cpp:43-51, 84-153; PhaseField_ComputeFaceFluxes stores
 * -fhat, :840): per stage the flow fill of every level, the faces of every box, the coarse faces
 * under a finer level averaged down, -div F and the stage combination, AverageDown. gq: per level a
 * face buffer with npf comps. */
static int amr_pf_rk3_step(const cmp_params* P, const cmp_pf_params* PF, int nlev, const cmp_geometry* G,
                           const int* nbox, const cmp_box* const* boxes, int ng, double* const* Un,
                           double* const* gq, double dt, int ratio)
{
    const int dim = P->dim, N = P->nstate, npf = P->nphase ? P->nphase : 1;
    double *Ub[AMR_NLEV_MAX], *Unew[AMR_NLEV_MAX];
    double **F[AMR_NLEV_MAX], **FN[AMR_NLEV_MAX];
    long long tot[AMR_NLEV_MAX];
    long long* foff[AMR_NLEV_MAX];
    for (int l = 0; l < nlev; ++l) {
        tot[l] = level_total(dim, nbox[l], boxes[l], ng, N);
        Ub[l] = (double*)malloc(sizeof(double) * tot[l]);
        Unew[l] = (double*)malloc(sizeof(double) * tot[l]);
        F[l] = (double**)calloc((size_t)nbox[l] * 3, sizeof(double*));
        FN[l] = (double**)calloc((size_t)nbox[l] * 3, sizeof(double*));
        foff[l] = (long long*)malloc(sizeof(long long) * (size_t)(nbox[l] * dim + 1));
        cmp_face_layout(dim, nbox[l], boxes[l], npf, foff[l], NULL);
    }
    int err = 0, e;
    const double w2 = 0.25 * dt, w3 = (2.0 / 3.0) * dt;
    for (int stage = 0; stage < 3; ++stage) {
        for (int l = 0; l < nlev; ++l) memcpy(Ub[l], stage == 0 ? Un[l] : Unew[l], sizeof(double) * tot[l]);
        e = amr_fill(P, nlev, G, nbox, boxes, ng, Ub, ratio);
        if (e && !err) err = e;
        for (int l = 0; l < nlev; ++l)
            for (int b = 0; b < nbox[l]; ++b) {
                const cmp_box* bx = &boxes[l][b];
                for (int d = 0; d < dim; ++d) {
                    int fe[3];
                    face_extents(dim, bx, d, fe);
                    const long long nface = (long long)fe[0] * fe[1] * fe[2];
                    F[l][b * 3 + d] = (double*)malloc(sizeof(double) * nface * N);
                    FN[l][b * 3 + d] = (double*)calloc((size_t)nface, sizeof(double));
                    const double* gd = gq[l] + foff[l][b * dim + d];
                    const int lo2 = (dim == 3) ? bx->lo[2] : 0;
                    for (int kk = 0; kk < fe[2]; ++kk)
                        for (int jj = 0; jj < fe[1]; ++jj)
                            for (int ii = 0; ii < fe[0]; ++ii) {
                                const long long q = ii + (long long)fe[0] * (jj + (long long)fe[1] * kk);
                                double fh[REF_NMAX], c, gv[CMP_MAX_PHASES];
                                for (int m = 0; m < npf; ++m) gv[m] = gd[q + m * nface];
                                e = ref_pf_face_flux(P, PF, &G[l], bx, ng, Ub[l] + bx->offset, bx->lo[0] + ii,
                                                     bx->lo[1] + jj, lo2 + kk, d, gv, fh, &c);
                                if (e && !err) err = e;
                                for (int n = 0; n < N; ++n) F[l][b * 3 + d][q + n * nface] = -fh[n];
                            }
                }
            }
        for (int l = nlev - 2; l >= 0; --l)
            average_faces(P, &G[l], nbox[l + 1], boxes[l + 1], F[l + 1], FN[l + 1], nbox[l], boxes[l], F[l], FN[l],
                          ratio);
        for (int l = 0; l < nlev; ++l)
            for (int b = 0; b < nbox[l]; ++b) {
                const cmp_box* bx = &boxes[l][b];
                fab_t f = make_fab(dim, bx, ng);
                int nv[3] = {1, 1, 1};
                for (int d = 0; d < dim; ++d) nv[d] = bx->hi[d] - bx->lo[d] + 1;
                const int lo2 = (dim == 3) ? bx->lo[2] : 0;
                for (int n = 0; n < N; ++n)
                    for (int kk = 0; kk < nv[2]; ++kk)
                        for (int jj = 0; jj < nv[1]; ++jj)
                            for (int ii = 0; ii < nv[0]; ++ii) {
                                const int idx[3] = {ii, jj, kk};
                                double div = 0.0;
                                for (int d = 0; d < dim; ++d) {
                                    int fe[3] = {nv[0], nv[1], nv[2]};
                                    fe[d] += 1;
                                    const long long nface = (long long)fe[0] * fe[1] * fe[2];
                                    const long long qm = idx[0] + (long long)fe[0] * (idx[1] + (long long)fe[1] * idx[2]);
                                    const long long qp = qm + (d == 0 ? 1 : (d == 1 ? fe[0] : (long long)fe[0] * fe[1]));
                                    const double* Fd = F[l][b * 3 + d];
                                    const double term = (Fd[qp + n * nface] - Fd[qm + n * nface]) / G[l].dx[d];
                                    div = (d == 0) ? term : div + term;
                                }
                                const double Lq = -div;
                                const long long q = bx->offset + fab_idx(&f, bx->lo[0] + ii, bx->lo[1] + jj, lo2 + kk) + n * f.sn;
                                if (stage == 0) {
                                    Unew[l][q] = 1.0 * Ub[l][q] + dt * Lq;
                                } else if (stage == 1) {
                                    Unew[l][q] = 0.25 * Ub[l][q] + 0.75 * Un[l][q];
                                    Unew[l][q] += w2 * Lq;
                                } else {
                                    Unew[l][q] = (2.0 / 3.0) * Ub[l][q] + (1.0 / 3.0) * Un[l][q];
                                    Unew[l][q] += w3 * Lq;
                                }
                            }
            }
        for (int l = 0; l < nlev; ++l)
            for (int q = 0; q < nbox[l] * 3; ++q) { free(F[l][q]); free(FN[l][q]); F[l][q] = FN[l][q] = NULL; }
        for (int l = nlev - 2; l >= 0; --l) {
            e = ref_level_average_down(P, nbox[l + 1], boxes[l + 1], ng, Unew[l + 1], nbox[l], boxes[l], ng, Unew[l],
                                       ratio);
            if (e && !err) err = e;
        }
    }
    for (int l = 0; l < nlev; ++l) {
        for (int b = 0; b < nbox[l]; ++b) {
            const cmp_box* bx = &boxes[l][b];
            fab_t f = make_fab(dim, bx, ng);
            int nv[3] = {1, 1, 1};
            for (int d = 0; d < dim; ++d) nv[d] = bx->hi[d] - bx->lo[d] + 1;
            const int lo2 = (dim == 3) ? bx->lo[2] : 0;
            for (int n = 0; n < N; ++n)
                for (int kk = 0; kk < nv[2]; ++kk)
                    for (int jj = 0; jj < nv[1]; ++jj)
                        for (int ii = 0; ii < nv[0]; ++ii) {
                            const long long q = bx->offset + fab_idx(&f, bx->lo[0] + ii, bx->lo[1] + jj, lo2 + kk) + n * f.sn;
                            Un[l][q] = Unew[l][q];
                        }
        }
        free(Ub[l]); free(Unew[l]); free(F[l]); free(FN[l]); free(foff[l]);
    }
    return err;
}

/* The Phase-Field step of a hierarchy after a flow step of size dt (see above), in place on U.
 * info (7 doubles): max|u|, mobility, eta, mechanism sub-steps, max |Grad_Q| over the levels,
 * Phase-Field sub-steps, CG iterations of the last implicit solve. */
int ref_amr_pf_evolve(const cmp_params* P, const cmp_pf_params* PF, int nlev, const cmp_geometry* G, const int* nbox,
                      const cmp_box* const* boxes, int ng, double* const* U, double dt, int ratio, double* info)
{
    if (nlev < 1 || nlev > AMR_NLEV_MAX) return CMP_ERR_ARG;
    const int dim = P->dim, npf = P->nphase ? P->nphase : 1;
    double* gq[AMR_NLEV_MAX] = {NULL};
    long long nf[AMR_NLEV_MAX];
    for (int l = 0; l < nlev; ++l) {
        cmp_face_layout(dim, nbox[l], boxes[l], npf, NULL, &nf[l]);
        gq[l] = (double*)calloc((size_t)nf[l], sizeof(double));
    }
    double inf5[5] = {0, 0, 0, 0, 0};
    int err = ref_amr_pf_gradq_step(P, PF, nlev, G, nbox, boxes, ng, U, dt, ratio, gq, inf5);
    double gm[AMR_NLEV_MAX], gmax = 0.0;
    for (int l = 0; l < nlev; ++l) {
        gm[l] = 0.0;
        for (long long q = 0; q < nf[l]; ++q) gm[l] = smax(gm[l], fabs(gq[l][q]));
        gmax = smax(gmax, gm[l]);
    }
    /* sub-steps: PhaseField_ComputeDt over the levels (nsubsteps 1, run.do_subcycle = 0) */
    double prev[AMR_NLEV_MAX];
    double pf_time = 0.0;
    int n = 0;
    for (int step = 0; step < PF->max_step && pf_time < dt && !err; ++step) {
        double tmp[AMR_NLEV_MAX] = {0.0};
        for (int l = 0; l < nlev; ++l) {
            double dxm = G[l].dx[0];
            for (int d = 1; d < dim; ++d) dxm = smin(dxm, G[l].dx[d]);
            tmp[l] = PF->cfl * dxm / (gm[l] + 1.0e-14);
            if (pf_time == 0.0) prev[l] = tmp[l];
        }
        double dt0 = tmp[0];
        for (int l = 0; l < nlev; ++l) {
            if (PF->change_limiter && prev[l] > 0.0) tmp[l] = smin(tmp[l], 1.1 * prev[l]);
            dt0 = smin(dt0, tmp[l]);
        }
        const double eps = 1.0e-3 * dt0;
        if (pf_time + dt0 > dt - eps) dt0 = dt - pf_time;
        double dmin = dt0;
        for (int l = 0; l < nlev; ++l) dmin = smin(dmin, tmp[l]);
        for (int l = 0; l < nlev; ++l) prev[l] = dmin;
        err = amr_pf_rk3_step(P, PF, nlev, G, nbox, boxes, ng, U, gq, dmin, ratio);
        pf_time += dmin;
        ++n;
        if (pf_time >= dt - 1.0e-6 * dt) break;
    }
    if (info) {
        for (int q = 0; q < 4; ++q) info[q] = inf5[q];
        info[4] = gmax;
        info[5] = (double)n;
        info[6] = inf5[4];
    }
    for (int l = 0; l < nlev; ++l) free(gq[l]);
    return err;
}

/* nsteps steps of a hierarchy with the Phase-Field step after each flow step (Evolve with
 * timeStepNoSubcycling then PhaseField_Evolve): the initial c_max, then per step ComputeDt over
 * levels, the stage-coupled RK3 step, the Phase-Field step of that dt. */
int ref_amr_advance_pf(const cmp_params* P, const cmp_pf_params* PF, int nlev, const cmp_geometry* G, const int* nbox,
                       const cmp_box* const* boxes, int ng, double* const* U, int nsteps, double cfl, int limiter,
                       int ratio, double* dt_hist)
{
    const int dim = P->dim, N = P->nstate;
    double cmax[AMR_NLEV_MAX] = {0, 0, 0, 0, 0, 0, 0, 0};
    double* Ub[AMR_NLEV_MAX];
    int err = 0;
    if (nlev < 1 || nlev > AMR_NLEV_MAX) return CMP_ERR_ARG;
    for (int l = 0; l < nlev; ++l) {
        const long long t = level_total(dim, nbox[l], boxes[l], ng, N);
        Ub[l] = (double*)malloc(sizeof(double) * t);
        memcpy(Ub[l], U[l], sizeof(double) * t);
    }
    err = amr_fill(P, nlev, G, nbox, boxes, ng, Ub, ratio);
    for (int l = 0; l < nlev && !err; ++l)
        for (int b = 0; b < nbox[l]; ++b) {
            double* Fx[3] = {NULL, NULL, NULL};
            double* FNx[3] = {NULL, NULL, NULL};
            double cm = 0.0;
            int e = box_faces(P, &G[l], &boxes[l][b], ng, Ub[l] + boxes[l][b].offset, Fx, FNx, &cm);
            if (e && !err) err = e;
            if (cmax[l] < cm || isnan(cm)) cmax[l] = cm;
            for (int d = 0; d < 3; ++d) { free(Fx[d]); free(FNx[d]); }
        }
    for (int l = 0; l < nlev; ++l) free(Ub[l]);
    double dt = 0.0;
    for (int s = 0; s < nsteps && !err; ++s) {
        dt = ref_amr_compute_dt(nlev, G, cmax, cfl, limiter, s == 0 ? 0.0 : dt);
        if (dt_hist) dt_hist[s] = dt;
        for (int l = 0; l < nlev; ++l) cmax[l] = 0.0;
        err = ref_amr_rk3_step(P, nlev, G, nbox, boxes, ng, U, dt, ratio, cmax);
        for (int l = 0; l < nlev; ++l)
            if (isnan(cmax[l])) err = CMP_ERR_NAN;
        if (!err) err = ref_amr_pf_evolve(P, PF, nlev, G, nbox, boxes, ng, U, dt, ratio, NULL);
    }
    return err;
}
