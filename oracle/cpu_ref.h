/*
 * cpu_ref.h - CPU restatement (the oracle) of COMPAS's five-equation hyperbolic right-hand side.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * may load this library, and only as the checker or the timed CPU baseline. The product path
 * (libcompas_amd.so) never links or calls it.
 *
 * Every function restates one reference function in the same operation order (no FMA
 * contraction: built with -ffp-contract=off), and says which file:line it follows. Data layout
 * and the POD structs are those of include/compas_amd.h.
 */
#ifndef CPU_REF_H
#define CPU_REF_H

#include "../include/compas_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define REF_NMAX 10

/* pointwise physics */
double ref_nasg_phi(double alpha, double alpha_rho, double b, int* err);
double ref_eos_mixture(const cmp_params* P, double input, double a1, double a1r1, double a2r2,
                       int output_internal_energy, int* err);
void   ref_cons2prim(const cmp_params* P, const double* U, double* W, int* err);
void   ref_prim2cons(const cmp_params* P, const double* W, double* U, int* err);
double ref_rho_c2_phase(double p, double A, double B);
double ref_sound_speed(const cmp_params* P, double p, double a1, double a1r1, double a2r2, int id_k,
                       int* err);
void   ref_nonconservative_K(const cmp_params* P, double p, double a1, double a1r1, double a2r2,
                             int id_k, double* K, int* err);
void   ref_cons2flux(const cmp_params* P, const double* U, int dir, double* F, int* err);

/* the N-phase model (P->nphase = 2 or 3): state alpha_k rho_k, rho u_d, rho E, alpha_k */
double ref_np_eos_mixture(const cmp_params* P, double input, const double* alpha, const double* arho,
                          int output_internal_energy, int* err);
void   ref_np_cons2prim(const cmp_params* P, const double* U, double* W, int* err);
void   ref_np_prim2cons(const cmp_params* P, const double* W, double* U, int* err);
double ref_np_sound_speed(const cmp_params* P, double p, const double* alpha, const double* arho,
                          int id_k, int* err);
void   ref_np_K(const cmp_params* P, double p, const double* alpha, const double* arho, int id_k,
                double* K, int* err);
void   ref_np_cons2flux(const cmp_params* P, const double* U, int dir, double* F, int* err);
void   ref_np_riemann(const cmp_params* P, const double* UL, const double* UR, int dir, double* flux,
                      double* vel, double* cmax, int* err);
void   ref_np_limiter_consistent(int np, const double* slope, double* theta);
void   ref_np_limit_consistent(const cmp_params* P, const double* Uf, const double* Uc, double dx,
                               int id_face, double* Uo, int* err);

/* scalar reconstruction */
double ref_weno5(double f_mm, double f_m, double f_c, double f_p, double f_pp, int id_face);
void   ref_weno5_weights(double f_mm, double f_m, double f_c, double f_p, double f_pp, int id_face,
                         double* cand, double* w);
double ref_muscl(double f_m, double f_c, double f_p, double dx, int id_face, int id_limiter);
/* WENO5B face state of cell W[2] from the primitives of cells W[0..4] */
void   ref_weno5b_side(const cmp_params* P, double W[][REF_NMAX], int id_face, double* Wf);
double ref_thinc(double beta, double sigma, double alpha_bar, int id_face);

/* Riemann solvers: flux[nstate], vel[dim], *cmax */
void ref_riemann(const cmp_params* P, const double* UL, const double* UR, int dir,
                 double* flux, double* vel, double* cmax, int* err);

/* per-face kernel on a box fab: U is the fab base of `box` (ng ghosts); (i,j,k) the face index
 * (face i sits between cells i-1 and i in direction dir) */
int ref_face_flux(const cmp_params* P, const cmp_geometry* G, const cmp_box* box, int ng,
                  const double* U, int i, int j, int k, int dir,
                  double* flux, double* fluxnc, double* cmax);

/* one box: ComputeFaceFluxes + FluxDivergence. dUdt: valid box, nstate comps. cmax (scalar or NULL)
 * max-accumulated over low faces of valid cells. flux[d]/fluxnc[d] (NULL allowed) nodal arrays. */
int ref_rhs_box(const cmp_params* P, const cmp_geometry* G, const cmp_box* box, int ng,
                const double* U, double* dUdt, double* cmax, double** flux, double** fluxnc);

/* level operations on level buffers laid out by cmp_level_layout */
int ref_level_fill_ghosts(const cmp_params* P, const cmp_geometry* G, int nbox,
                          const cmp_box* boxes, int ng, double* U);
int ref_level_rhs(const cmp_params* P, const cmp_geometry* G, int nbox, const cmp_box* boxes,
                  int ng, const double* U, double* dUdt, double* cmax);
int ref_level_rk3_step(const cmp_params* P, const cmp_geometry* G, int nbox,
                       const cmp_box* boxes, int ng, double* Un, double dt, double* cmax);
/* One step of run.TimeIntegrator: 1 ForwardEuler, 2 TVD-RK2, 3 TVD-RK3, 4 RK4. */
int ref_level_rk_step(const cmp_params* P, const cmp_geometry* G, int nbox, const cmp_box* boxes, int ng,
                      double* Un, double dt, double* cmax, int integrator);
/* A Phase-Field sub-step of pf_dt at a fixed Grad_Q with run.TimeIntegrator (1..4 as above). */
int ref_level_pf_rk_step(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G, int nbox,
                         const cmp_box* boxes, int ng, double* Un, const double* gradq, double pf_dt, int integrator);
/* Evolve's step loop (flow only / with the Phase-Field step) with run.TimeIntegrator for the flow
 * steps and the Phase-Field sub-steps (1..4 as above). */
int ref_level_advance_int(const cmp_params* P, const cmp_geometry* G, int nbox, const cmp_box* boxes,
                          int ng, double* U, int nsteps, double cfl, int limiter, double* dt_hist, int integrator);
int ref_level_advance_pf_int(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                             int nbox, const cmp_box* boxes, int ng, double* U, int nsteps, double cfl,
                             int limiter, double* dt_hist, int integrator);
int ref_level_initial_cmax(const cmp_params* P, const cmp_geometry* G, int nbox,
                           const cmp_box* boxes, int ng, const double* U, double* cmax);
double ref_est_dt(const cmp_geometry* G, double cmax, double cfl);
double ref_compute_dt(const cmp_geometry* G, double cmax, double cfl, int limiter, double dt_prev);
int ref_level_advance(const cmp_params* P, const cmp_geometry* G, int nbox, const cmp_box* boxes,
                      int ng, double* U, int nsteps, double cfl, int limiter, double* dt_hist);

/* ghost fill of a Phase-Field field (ncomp comps, ng ghosts): periodic images or foextrap */
int ref_level_fill_pf(int dim, const cmp_geometry* G, int nbox, const cmp_box* boxes, int ng,
                      int ncomp, double* A);

/* phase field: pointwise pieces */
double ref_pf_map_backward(double phi, double eta);
double ref_pf_map_forward(double psi, double eta, int deriv);
double ref_pf_acdi_flux(double pf_m, double pf_c, double n_m, double n_c, double mob, double eta,
                        double dx);
double ref_pf_gradq_face(double pf_m, double pf_c, double jhat, double eta, double eta_factor);
void   ref_pf_flux(const cmp_params* P, const cmp_pf_params* PF, const double* UL, const double* UR,
                   double g, double* flux, double* cmax, int* err);
void   ref_np_cons2phase(const cmp_params* P, const cmp_pf_params* PF, const double* U, double* Cp,
                         int* err);
void   ref_np_pf_flux(const cmp_params* P, const cmp_pf_params* PF, const double* UL, const double* UR,
                      const double* g, double* flux, double* cmax, int* err);
int    ref_pf_face_flux(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                        const cmp_box* box, int ng, const double* U, int i, int j, int k, int dir,
                        const double* gradq, double* fhat, double* cmax);
int    ref_pf_rhs_box(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                      const cmp_box* box, int ng, const double* U, const double* const* gq,
                      double* dUdt);
void   ref_pf_coefficients(const cmp_pf_params* PF, const cmp_geometry* G, double max_v,
                           double* lambda, double* eta, double* mob);
double ref_pf_compute_dt(const cmp_pf_params* PF, const cmp_geometry* G, double gradq_max,
                         double pf_time, double dt, double* prev);

/* phase field: level operations. alpha buffers: 1 comp, 1 ghost (cmp_level_layout offsets);
 * face buffers: cmp_face_layout with 1 comp. */
int ref_level_pf_rhs(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                     int nbox, const cmp_box* boxes, int ng, const double* U, const double* gradq,
                     double* dUdt);
int ref_level_pf_rk3_step(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                          int nbox, const cmp_box* boxes, int ng, double* Un, const double* gradq,
                          double pf_dt);
double ref_level_max_velocity(const cmp_params* P, int nbox, const cmp_box* boxes, int ng,
                              const double* U);
double ref_pf_cac_mech(int dim, const double* dx, double f, const double* fm, const double* fp,
                       const double* nc, const double* nm, const double* np, double mob,
                       double lambda, double eta);
int ref_level_pf_mechanism(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                           int nbox, const cmp_box* boxes, const double* alpha, double dt,
                           double mob, double eta, double lambda, double* jhat, double* divjhat,
                           int* nsub);
int ref_level_pf_gradq_implicit(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                                int nbox, const cmp_box* boxes, const double* alpha,
                                const double* divjhat, double eta, double* gradq, int* iters);
int ref_level_pf_gradq(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                       int nbox, const cmp_box* boxes, const double* alpha, const double* jhat,
                       double eta, double* gradq);
int ref_level_pf_gradq_step(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                            int nbox, const cmp_box* boxes, int ng, const double* U, double dt,
                            double* gradq, double* gradq_max, double* info);
int ref_level_pf_evolve(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                        int nbox, const cmp_box* boxes, int ng, double* U, double dt, double* info);
int ref_level_advance_pf(const cmp_params* P, const cmp_pf_params* PF, const cmp_geometry* G,
                         int nbox, const cmp_box* boxes, int ng, double* U, int nsteps, double cfl,
                         int limiter, double* dt_hist);

/* refinement tags of every valid cell (cmp_level_tag's layout) */
int ref_level_tag(const cmp_params* P, int nbox, const cmp_box* boxes, int ng, const double* U,
                  const cmp_tag_criterion* crit, int ncrit, unsigned char* tags);

/* average_down of a fine level's boxes onto the coarse level's cells they cover (ratio >= 2) */
int ref_level_average_down(const cmp_params* P, int nbf, const cmp_box* fboxes, int ngf, const double* Uf,
                           int nbc, const cmp_box* cboxes, int ngc, double* Uc, int ratio);

/* FillPatchTwoLevels with lincc_interp: every ghost cell of the fine level (same-level sources
 * copied, coarse-fine ghosts interpolated from the coarse level) */
int ref_level_fill_two_levels(const cmp_params* P, const cmp_geometry* Gf, int nbf, const cmp_box* fboxes, int ngf,
                              double* Uf, const cmp_geometry* Gc, int nbc, const cmp_box* cboxes, int ngc,
                              const double* Uc, int ratio);
/* Every cell (valid and ghost) of a level's new boxes from its old boxes, else from the coarse
 * level (the regrid's FillPatch of a new level). */
int ref_level_fill_from(const cmp_params* P, const cmp_geometry* Gf, int nbn, const cmp_box* nboxes, int ngn,
                        double* Un, int nbo, const cmp_box* oboxes, int ngo, const double* Uo,
                        const cmp_geometry* Gc, int nbc, const cmp_box* cboxes, int ngc, const double* Uc, int ratio);
/* All levels of a hierarchy, stage by stage without subcycling (stage-coupled, do_reflux = 1),
 * then averaged down; levels coarsest first, ratio between consecutive levels, ng for all. */
int ref_amr_rk3_step(const cmp_params* P, int nlev, const cmp_geometry* G, const int* nbox,
                     const cmp_box* const* boxes, int ng, double* const* Un, double dt, int ratio, double* cmax);
double ref_amr_compute_dt(int nlev, const cmp_geometry* G, const double* cmax, double cfl, int limiter,
                          double dt_prev);
int ref_amr_advance(const cmp_params* P, int nlev, const cmp_geometry* G, const int* nbox,
                    const cmp_box* const* boxes, int ng, double* const* U, int nsteps, double cfl, int limiter,
                    int ratio, double* dt_hist);

/* number of OpenMP threads the level functions use (1 when built without OpenMP) */
/* The Phase-Field step of a hierarchy without subcycling (PhaseField_Evolve over finest_level + 1
 * levels, with the composite implicit solve), its Grad_Q alone, and the flow + Phase-Field loop. */
int ref_amr_pf_gradq_step(const cmp_params* P, const cmp_pf_params* PF, int nlev, const cmp_geometry* G,
                          const int* nbox, const cmp_box* const* boxes, int ng, double* const* U, double dt,
                          int ratio, double* const* gq, double* info);
int ref_amr_pf_gradq_implicit(const cmp_params* P, const cmp_pf_params* PF, int nlev, const cmp_geometry* G,
                              const int* nbox, const cmp_box* const* boxes, double* const* alpha,
                              double* const* rhs, double eta, int ratio, double* const* gq, double* const* bf,
                              int* iters);
int ref_amr_pf_evolve(const cmp_params* P, const cmp_pf_params* PF, int nlev, const cmp_geometry* G, const int* nbox,
                      const cmp_box* const* boxes, int ng, double* const* U, double dt, int ratio, double* info);
int ref_amr_advance_pf(const cmp_params* P, const cmp_pf_params* PF, int nlev, const cmp_geometry* G, const int* nbox,
                       const cmp_box* const* boxes, int ng, double* const* U, int nsteps, double cfl, int limiter,
                       int ratio, double* dt_hist);
int ref_num_threads(void);
void ref_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
