/*
 * compas_amd.h - C ABI of the MI355X (gfx950) five-equation finite-volume right-hand side.
 *
 * Plain C: POD structs, raw device pointers, sizes and an opaque stream (hipStream_t passed as
 * void*). Every entry point returns an int status (CMP_OK = 0) and never aborts the process.
 *
 * Memory layout (same as AMReX Array4, which the reference's kernels index): a "fab" of a box
 * with ng ghost cells stores (i,j,k,n) with i fastest, then j, k, then component n; the index
 * range is [lo-ng, hi+ng] in each active dimension and a single cell in inactive ones. A level
 * buffer concatenates the fabs of its boxes; cmp_box.offset is the element offset of a box's fab.
 * Face arrays (optional outputs) are nodal in their direction and carry no ghosts.
 *
 * State vector (PHYSICS == FIVEEQS, include/Physics_Compressible5Eq/Variables_Compressible5Eq.H:21-30):
 *   n = 0 a1rho1, 1 a2rho2, 2..1+dim rho*u_d, 2+dim rho*E, 3+dim a1      (nstate = 4+dim)
 * or, with cmp_params.nphase = NP in {2, 3} (PHYSICS == FIVEEQS_NPHASE,
 * include/Physics_Compressible5Eq_NPhase/Variables_Compressible5Eq_NPhase.H:21-29):
 *   n = 0..NP-1 alpha_k rho_k, NP..NP+dim-1 rho*u_d, NP+dim rho*E, NP+dim+1.. alpha_k
 *                                                                        (nstate = 2 NP + dim + 1)
 */
#ifndef COMPAS_AMD_H
#define COMPAS_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------------------- */
#define CMP_OK               0
#define CMP_ERR_ARG          1  /* bad argument (null pointer, size, dim, ghost width)                  */
#define CMP_ERR_UNSUPPORTED  2  /* scheme combination the reference aborts on, or not built here        */
#define CMP_ERR_HIP          3  /* a HIP runtime call failed                                            */
#define CMP_ERR_NAN          4  /* a NaN appeared in the state or the wave speed                        */
#define CMP_ERR_COVOLUME     5  /* NASG covolume bound 1 - b rho <= 0 (EquationOfState_NASG.H:28-32)   */
#define CMP_ERR_BC           6  /* boundary type without a fill rule here (Inflow/ext_dir)              */
#define CMP_ERR_LAYOUT       7  /* boxes overlap, leave the domain, or do not cover it                  */
#define CMP_ERR_NOMEM        8  /* a batch has no free case slot or no free range for a case's state    */
#define CMP_ERR_SCRATCH      9  /* larger than the fixed device scratch of an allocation-free entry point */
#define CMP_ERR_DENSITY     10  /* a cell's density sum_k alpha_k rho_k <= 0                              */
#define CMP_ERR_PRESSURE    11  /* c^2 <= 0: p + p_inf <= 0 in the mixture or a phase (the reference's
                                 * sound speed would be NaN or zero)                                     */
#define CMP_ERR_ALPHA       12  /* alpha1 of a stage state outside [0, 1] by more than 1e-12 with the
                                 * bounded limiter (WENO5B) on                                           */

/* Scheme IDs follow FiniteVolume_Parameter::Initialize (include/Operators_FV/FiniteVolume_Parameter.H:71-139):
 *   GODUNOV  id_weno=0 id_limiter=0 id_poly=0  ng>=1
 *   MUSCL-*  id_weno=0 id_limiter=1..4 (minmod, vanLeer, superbee, MC) id_poly=1  ng>=2
 *   WENO5    id_weno=1 id_limiter=5 id_poly=4  ng>=3
 *   WENO5B   id_weno=1 id_limiter=6 id_poly=4  ng>=3
 * Riemann IDs follow Physics_Parameter (include/Physics_Compressible5Eq/Physics_Parameter_Compressible5Eq.H:16-27):
 *   1 HLL, 2 HLLC.
 * Boundary codes follow run.lo_bc/run.hi_bc (source/Class_Compressible5Eq_PhaseField.cpp:169-183):
 *   0 Interior, 1 Inflow, 2 Outflow, 3 Symmetry, 4 SlipWall, 5 NoSlipWall. */

/* Parameters read by every kernel: the subset of the reference's Parm that this path uses. */
#define CMP_MAX_PHASES 3

typedef struct cmp_params {
    int    dim;            /* AMREX_SPACEDIM, 2 or 3                                                 */
    int    nstate;         /* 4 + dim (nphase 0), 2 nphase + dim + 1 (nphase 2 or 3)                  */
    int    nphase;         /* 0: the two-phase five-equation model (PHYSICS == FIVEEQS), state
                            * a1rho1 a2rho2 rho_u.. rhoE a1 (Variables_Compressible5Eq.H:21-30);
                            * 2 or 3: the N-phase model (PHYSICS == FIVEEQS_NPHASE, NPHASE = nphase),
                            * state alpha_k rho_k.. rho_u.. rhoE alpha_k..
                            * (Variables_Compressible5Eq_NPhase.H:21-29)                              */
    /* EOS coefficients per phase as set by Parm::Initialize (source/Parm.cpp:124-156, :187-272):
     *   rho_k e_k = (1 - b_k rho_k)(A_k p + B_k) + rho_k D_k, b_k = C_k (NASG covolume, 0 otherwise);
     * the first 2 (nphase 0) or nphase entries are read */
    double eos_A[CMP_MAX_PHASES];
    double eos_B[CMP_MAX_PHASES];
    double eos_C[CMP_MAX_PHASES];
    double eos_D[CMP_MAX_PHASES];
    int    id_weno;
    int    id_limiter;
    int    id_poly;
    int    id_bound;       /* 1 (bound-preserving WENO5BP/MUSCLBP) is not built here                 */
    int    id_thinc;       /* THINC interface sharpening (FiniteVolume.THINC)                         */
    int    thinc_use_normal;
    double thinc_eps;
    double thinc_beta;
    double thinc_normal_threshold;
    int    source_term;    /* Physics.source_term: ID_K, 0 or 1                                      */
    int    id_riemann;     /* 1 HLL, 2 HLLC                                                           */
} cmp_params;

typedef struct cmp_geometry {
    int    dim;
    int    domain_lo[3];
    int    domain_hi[3];
    int    is_periodic[3];
    int    lo_bc[3];
    int    hi_bc[3];
    double dx[3];
    double prob_lo[3];
} cmp_geometry;

typedef struct cmp_box {
    int       lo[3];
    int       hi[3];
    long long offset;      /* element offset of this box's fab inside the level buffer              */
} cmp_box;

/* Phase-field regularization parameters: the subset of PhaseField_Parameter
 * (include/Physics_PhaseField/PhaseField_Parameter.H:29-144) that the ACDI and CAC-Adv paths read. */
#define CMP_PF_CAC  1      /* PhaseField.Mechanism = "CAC-Adv" (ID_Mechanism 1, not flux based)      */
#define CMP_PF_ACDI 3      /* PhaseField.Mechanism = "ACDI" (ID_Mechanism 3, flux based)             */
typedef struct cmp_pf_params {
    int    mechanism;        /* ID_Mechanism: CMP_PF_CAC or CMP_PF_ACDI                              */
    int    explicit_rc;      /* ID_ExplicitRC: 1 = Grad_Q from the time-averaged flux Jhat (ACDI
                              * only; CAC with 1 is CMP_ERR_UNSUPPORTED, as the reference aborts,
                              * PhaseField_Parameter.H:138-140); 0 = Grad_Q from the elliptic solve
                              * div(b grad Q) = div Jhat                                             */
    int    id_flux;          /* ID_Flux: 0 = UD flux                                                  */
    int    flux_phase2flux;  /* ID_FluxPhase2Flux: 1 = mixture formulation, 0 = phasic              */
    double cfl;              /* PhaseField.cfl of the Phase-Field sub-steps                           */
    double lambda;           /* PhaseField.Lambda; <= 0 means 1                                       */
    double eta;              /* PhaseField.Eta; <= 0 means eta_multiplier * max_d dx_d                */
    double eta_factor;       /* PhaseField.Eta_Factor (re-sharpening width of Grad_Q's weight)        */
    double eta_multiplier;   /* PhaseField.Eta_Multiplier                                             */
    double p_min[CMP_MAX_PHASES]; /* phasic energy floors, one per phase (phasic formulation only;
                              * the reference sets -B_k/A_k, or -B_k/(A_k+1) with the source term) */
    int    max_step;         /* PhaseField.max_step: cap on Phase-Field sub-steps per flow step       */
    int    change_limiter;   /* timestep_change_limiter applied to the sub-step size                  */
} cmp_pf_params;

/* ---- layout helpers (host only, no GPU) ------------------------------------------------------- */

/* Elements of one fab: prod_d (hi_d - lo_d + 1 + 2 ng) over active dims, times ncomp. */
long long cmp_fab_size(int dim, const cmp_box* box, int ng, int ncomp);

/* Assigns consecutive offsets to the boxes of a level; *total receives the level buffer length. */
int cmp_level_layout(int dim, int nbox, cmp_box* boxes, int ng, int ncomp, long long* total);

/* Face-array layout of a level: for box b and direction d (d < dim) a nodal array of
 * prod_e (hi_e - lo_e + 1 + [e == d]) faces times ncomp, no ghosts, placed box-major:
 * offsets[b*dim + d]. *total receives the buffer length. */
int cmp_face_layout(int dim, int nbox, const cmp_box* boxes, int ncomp, long long* offsets,
                    long long* total);

/* sizeof(cmp_params), sizeof(cmp_geometry), sizeof(cmp_box): lets bindings check their mirrors. */
int cmp_abi_sizes(int* params_size, int* geometry_size, int* box_size);

/* sizeof(cmp_pf_params). */
int cmp_abi_pf_size(int* pf_params_size);

/* Library version string. */
const char* cmp_version(void);

/* ---- one box ---------------------------------------------------------------------------------- */

/* Right-hand side of one box from its ghost-filled state, i.e. COMPAS's per-box
 * ComputeFaceFluxes + FluxDivergence (source/Class_Compressible_PhaseField.cpp:3152-3391, called by
 * compute_dUdt_FV :3473-3537, per face FVM_Conservative2FluxHyperbolic_K,
 * include/Operators_Compressible5Eq/FVM_Conservative2FluxHyperbolic_K_Compressible5Eq.H:15-170).
 *   U        device, fab of `box` with ng ghosts, nstate comps (ghosts already filled)
 *   dUdt     device, valid box (no ghosts), nstate comps; overwritten
 *   cmax     device scalar or NULL: max-accumulated with the wave speed of every face that is the
 *            low face of a valid cell (the reference's c_max(i,j,k) = max over faces (i,j,k),
 *            reduced over valid cells, :3212 and :4490)
 *   flux_d   device or NULL: nodal face fluxes of direction d, nstate comps (the `fluxes` MultiFab)
 *   fluxnc_d device or NULL: nodal face normal velocity u*_n, 1 comp (the `fluxes_nc` MultiFab)
 *   stream   hipStream_t or NULL for the default stream */
int cmp_rhs_box(const cmp_params* params, const cmp_geometry* geom, const cmp_box* box, int ng,
                const double* U, double* dUdt, double* cmax,
                double* flux_x, double* flux_y, double* flux_z,
                double* fluxnc_x, double* fluxnc_y, double* fluxnc_z,
                void* stream);

/* The per-cell wave speed of one box from its ghost-filled state, the reference's c_max MultiFab
 * (ComputeFaceFluxes, source/Class_Compressible_PhaseField.cpp:3152-3254: c_max(i,j,k) = the largest
 * wave speed over the faces that are the low faces of cell (i,j,k), reduced over the valid cells by
 * ComputeDt :4490). cmax_cells: device, valid box (no ghosts), one double per cell, overwritten.
 * Its maximum is what cmp_rhs_box max-accumulates into its cmax. Two-phase model
 * (CMP_ERR_UNSUPPORTED for nphase > 0, whose face arrays are not built). */
int cmp_cmax_box(const cmp_params* params, const cmp_geometry* geom, const cmp_box* box, int ng, const double* U,
                 double* cmax_cells, void* stream);

/* Phase-Field right-hand side of one box from its ghost-filled state, i.e. the per-box body of
 * PhaseField_compute_dUdt_FV (source/Class_Compressible_PhaseField.cpp:951-1003): the faces of
 * PhaseField_ComputeFaceFluxes (:771-852) with FVM_Conservative2FluxPhaseField_K, their -fhat and
 * the conservative surface integral of PhaseField_FluxDivergence (:859-948); the same arithmetic as
 * cmp_level_pf_rhs on that box.
 *   pf      the flux form (flux_phase2flux 1 mixture, 0 phasic) and the phasic p_min; id_flux 0
 *   U       device, fab of `box` with ng ghosts, nstate comps (ghosts already filled)
 *   gradq   device, the box's Grad_Q faces in the layout of cmp_face_layout for this box alone
 *           (one component per volume fraction)
 *   dUdt    device, valid box (no ghosts), nstate comps; overwritten */
int cmp_pf_rhs_box(const cmp_params* params, const cmp_pf_params* pf, const cmp_geometry* geom, const cmp_box* box,
                   int ng, const double* U, const double* gradq, double* dUdt, void* stream);

/* ---- one level (all boxes of a level advance together) ---------------------------------------- */

typedef struct cmp_level cmp_level;

/* Builds the device work-lists of a level: tile list, ghost-fill map, box lookup. The boxes must
 * carry offsets from cmp_level_layout (nstate comps, ng ghosts). Allocates device memory for these
 * tables only; state buffers stay owned by the caller. */
int cmp_level_create(const cmp_params* params, const cmp_geometry* geom, int nbox,
                     const cmp_box* boxes, int ng, cmp_level** out);
int cmp_level_destroy(cmp_level* level);

/* Same-level ghost fill: neighbour boxes, periodic images, then the physical BCs
 * (FillPatchSingleLevel, source/COMPAS.cpp:1647-1678; BC map source/Class_Compressible5Eq_PhaseField.cpp:162-244). */
int cmp_level_fill_ghosts(cmp_level* level, double* U, void* stream);

/* Right-hand side of every box of the level (ghosts of U already filled); dUdt is a level buffer
 * with the same layout as U (its ghost cells are not written). */
int cmp_level_rhs(cmp_level* level, const double* U, double* dUdt, double* cmax, void* stream);

/* One TVD-RK3 stage (ThirdOrderSSPRK, source/Class_Compressible_PhaseField.cpp:3938-4107), fused:
 *   stage 0: Uout = Us + dt L(Us)
 *   stage 1: Uout = (1/4 Us + 3/4 Un) + (dt/4) L(Us)
 *   stage 2: Uout = (2/3 Us + 1/3 Un) + (2 dt/3) L(Us)
 * Us must have its ghosts filled; dt is read from device memory (*dt_dev); cmax (device) is
 * max-accumulated. Uout may alias Un (each cell reads Un before it writes Uout) but not Us. */
int cmp_level_rk3_stage(cmp_level* level, int stage, const double* Us, const double* Un,
                        double* Uout, const double* dt_dev, double* cmax, void* stream);

/* One full RK3 step in place on Un with two scratch level buffers. The ghost fill is folded into
 * the stages: cells beyond a box are read where cmp_level_fill_ghosts would take them from, so the
 * ghost cells of Un, U1 and U2 are neither read nor written (call cmp_level_fill_ghosts to refresh
 * Un's when the caller reads them). */
int cmp_level_rk3_step(cmp_level* level, double* Un, double* U1, double* U2,
                       const double* dt_dev, double* cmax, void* stream);

/* Wave speed of the initial state (CalculateInitialcmax, :4502-4605): fills ghosts of U, sets
 * *cmax to the max face wave speed over low faces of valid cells. */
int cmp_level_initial_cmax(cmp_level* level, double* U, double* cmax, void* stream);

/* Device-side ComputeDt (source/COMPAS.cpp:2250-2302, EstTimeStep :4473-4499) for one level:
 *   est = min_d cfl*dx_d/(*cmax);  dt = (limiter && *dt_dev > 0) ? min(est, 1.1 * *dt_dev) : est
 * writes *dt_dev and resets *cmax to 0 for the next step's stages. */
int cmp_level_compute_dt(cmp_level* level, double* cmax, double* dt_dev, double cfl,
                         int limiter, void* stream);

/* nsteps time steps: per step ComputeDt then one RK3 step. With resume == 0 the run starts at
 * t = 0: step 0 takes its wave speed from the initial state (CalculateInitialcmax) and no
 * previous dt limits it; with resume != 0 it continues from *cmax and *dt_dev of an earlier call
 * on the same buffers. dt_hist (device, nsteps doubles, or NULL) receives the dt of each step of
 * this call. With use_graph != 0 one step is captured into a hipGraph once and replayed.
 * After cmp_level_pf_setup every step ends with cmp_level_pf_evolve (Evolve,
 * source/Class_Compressible_PhaseField.cpp:307-312); the graph then holds the flow part only. */
int cmp_level_advance(cmp_level* level, double* Un, double* U1, double* U2, int nsteps,
                      double cfl, int limiter, double* dt_dev, double* cmax, double* dt_hist,
                      int use_graph, int resume, void* stream);

/* One step of run.TimeIntegrator (source/COMPAS.cpp:1306-1311): integrator 1 ForwardEuler
 * (source/Class_Compressible_PhaseField.cpp:3713-3800), 2 TVD-RK2 (SecondOrderSSPRK, :3805-3935),
 * 3 TVD-RK3 (cmp_level_rk3_step), 4 RK4 (FourthOrderRK, :4111-4330): per stage the stage state's
 * ghost cells filled, its right-hand side, and the reference's LinComb / Saxpy combinations in its
 * order. In place on Un; U1, U2 scratch level buffers; U3 a fourth one for RK4 (else unused). The
 * stages' c_max accumulate into *cmax. */
int cmp_level_rk_step(cmp_level* level, int integrator, double* Un, double* U1, double* U2, double* U3,
                      const double* dt_dev, double* cmax, void* stream);

/* The level's clock: every ComputeDt (cmp_level_compute_dt, cmp_level_advance) advances its time by
 * the dt it sets, and with stop_time >= 0 cuts that dt so that the step ends on stop_time
 * (source/COMPAS.cpp:2278-2283: t + dt > stop_time - 1e-3 dt gives dt = stop_time - t; the cut
 * follows the default run.do_subcycle = 1, where it can lengthen the step by up to 1e-3 dt). Once
 * the time has reached stop_time, where the reference's step loop ends, every further step has
 * dt = 0 and leaves the state and c_max as they are (no stage, no Phase-Field step). A new level has
 * time 0 and no stop_time (negative). cmp_level_set_time waits for the stream's queued work, then
 * writes the clock in the stream's order; a level with a stop_time cannot join a hierarchy
 * (cmp_amr_create). */
int cmp_level_set_time(cmp_level* level, double time, double stop_time, void* stream);
int cmp_level_get_time(const cmp_level* level, double* time, void* stream);

/* run.TimeIntegrator of the level's cmp_level_advance steps and cmp_level_pf_evolve sub-steps (1
 * ForwardEuler, 2 TVD-RK2, 3 TVD-RK3, the default, 4 RK4; RK4 allocates a fourth level buffer
 * here). */
int cmp_level_set_integrator(cmp_level* level, int integrator);

/* Reads and clears the level's device error word (CMP_OK, or the largest of CMP_ERR_NAN, CMP_ERR_COVOLUME,
 * CMP_ERR_DENSITY, CMP_ERR_PRESSURE, CMP_ERR_ALPHA raised since the last read).
 * Synchronises the stream. */
int cmp_level_check(cmp_level* level, void* stream);

/* Number of tiles in the level work-list and the tile shape used (for reporting). */
int cmp_level_info(const cmp_level* level, int* ntiles, int* tile_x, int* tile_y, int* tile_z);

/* ---- persistent work-list stage kernel and tile autotuner ---------------------------------------
 * These replace the per-box tiling of ComputeFaceFluxes' MFIter loop
 * (source/Class_Compressible_PhaseField.cpp:3152-3254; tiles of amrex::TilingIfNotGPU) with one
 * persistent launch per RK stage over a list of tiles, each box with a tile shape of its own. The
 * reference has no counterpart of the measured shape sweep.
 *
 * cmp_level_set_tiles: the flow stage of cmp_level_rk3_step / cmp_level_advance (TVD-RK3) through
 * the work-list kernel, box b cut into tiles of ext[3b..3b+2] cells (clipped to the box; at most 256
 * cells, CMP_ERR_ARG for a shape the kernel's LDS layout does not hold), the kernel built for
 * `waves` (2 or 3) resident waves per SIMD; ext NULL: back to the fixed-shape kernel. A box whose
 * extents are 0 0 0 stays with the fixed-shape kernel: each stage is then a fixed-shape launch over
 * those boxes' tiles and a work-list launch over the others (every box 0 0 0: the fixed kernel
 * alone). Two-phase and N-phase flow (not the Phase-Field flux). The results are those of the
 * fixed-shape kernel bit for bit.
 * cmp_level_tile_shapes: *use_list 1 when the work-list kernel is on, *waves its occupancy target
 * (0 when off); ext (3 ints per box, or NULL): each box's tile extents (0 0 0 for a box the
 * fixed-shape kernel keeps in a level that mixes the two). */
int cmp_level_set_tiles(cmp_level* level, const int* ext, int waves);
int cmp_level_tile_shapes(const cmp_level* level, int* use_list, int* waves, int* ext);

/* The tile autotuner: for every box shape of the level (and the nextra shapes of `extra`, 3 ints
 * each) not yet in this process's cache for the level's scheme (scheme, limiter, Riemann solver,
 * THINC, nphase, ghost cells), a synthetic periodic level of copies of that box, its cells sampled
 * from the level's largest box in U, takes one RK stage per candidate tile shape with the work-list
 * kernel at both occupancy targets and one with the fixed-shape kernel (HIP events, least of 3); the
 * times per cell are cached. The summed sweeps give up to two candidates besides the fixed kernel,
 * each kept only when its sum beats the fixed kernel's: the work-list kernel for every box (each box
 * its candidate of least time) and each box by the faster of the two kernels, each at its better
 * occupancy target; one RK stage of the level itself on U (HIP events, least of 3; the level's error
 * word is left as it was) picks the fastest of them and the fixed kernel (force 1: the work-list
 * kernel for every box, -1: the fixed kernel). A box thinner than the ghost width is not swept, and a level holding
 * one keeps the fixed kernel. Allocates device memory while it runs; synchronises. */
int cmp_level_tune_tiles(cmp_level* level, const double* U, int nextra, const int* extra, int force, void* stream);
/* The cached sweep of one box shape (n[3]) for the level's scheme: the fixed-shape kernel's ns per
 * cell, the candidates' extents (3 ints each) and their ns per cell at 2 and 3 waves per SIMD
 * (2 doubles each; arrays of at least 12 candidates, or NULL). CMP_ERR_ARG when not cached. */
int cmp_tile_tuning(const cmp_level* level, const int* n, double* ns_fixed, int* ncand, int* cand_ext,
                    double* cand_ns);
/* The cache as a text file (one line per box shape and scheme) and back; clear empties it. */
int cmp_tile_cache_save(const char* path);
int cmp_tile_cache_load(const char* path);
int cmp_tile_cache_clear(void);

/* ---- host clustering of refinement tags ------------------------------------------------------
 * No counterpart in the reference's sources: the reference hands its TagBoxArray to AMReX's
 * Berger-Rigoutsos clustering (AmrCore::MakeNewGrids, amr.grid_eff). This is the split step of this
 * library's own clustering (compas_amd/regrid.py): blocks (dim-dimensional, C order, shape[] with
 * the slowest axis first; nonzero = a tagged block) are covered by boxes, each at least grid_eff
 * tagged after shrinking to its tags' bounding box, else cut at the tag-free plane nearest its
 * middle, else where the second difference of the tag signature changes sign most strongly, else in
 * the middle of its longest side. Writes *nboxes boxes of 6 ints (lo[3], hi[3], block indices in the
 * axis order of shape, unused axes 0); CMP_ERR_ARG when they would exceed max_boxes (the number of
 * tagged blocks always suffices). Host memory only; no GPU work. */
int cmp_cluster_blocks(const unsigned char* blocks, int dim, const int* shape, double grid_eff, int max_boxes,
                       int* boxes, int* nboxes);

/* ---- phase-field interface regularization of a level ------------------------------------------
 * Grad_Q face buffers below follow cmp_face_layout with one component per volume fraction: one for
 * the two-phase model (NPHASE = 1), nphase for the N-phase model (component k is phase k's Grad_Q,
 * computed from alpha_k as the two-phase model computes it from alpha1). */

/* Enables the Phase-Field step on a level and allocates its device work buffers (two copies of
 * alpha1 and the cell normals with one ghost cell; J, Jhat and Grad_Q on the faces; for CAC the
 * per-cell mech and W; for the implicit RC Div_Jhat, the CG vectors and b on the faces). */
int cmp_level_pf_setup(cmp_level* level, const cmp_pf_params* pf);

/* Phase-Field right-hand side of every box, PhaseField_compute_dUdt_FV
 * (source/Class_Compressible_PhaseField.cpp:951-1003): faces by PhaseField_ComputeFaceFluxes
 * (:771-852) with FVM_Conservative2FluxPhaseField_K
 * (include/Operators_Compressible5Eq_PhaseField/FVM_Conservative2FluxPhaseField_K_Compressible5Eq_PhaseField.H:15-74),
 * stored as -fhat, and the conservative surface integral of PhaseField_FluxDivergence (:859-948).
 * U ghosts filled; dUdt is a level buffer laid out like U (ghost cells not written). */
int cmp_level_pf_rhs(cmp_level* level, const double* U, const double* gradq, double* dUdt,
                     void* stream);

/* Grad_Q for a flow step of size dt from the valid cells of U: MaxVelocity (:3541-3594), the
 * mobility 0.5 eta max|u| (PhaseField_Evolve :425-465), the mechanism sub-steps of
 * PhaseFieldMechanism_Evolve (:597-709; PhaseField_ACDI :1495-1644 or the CAC-Adv
 * PhaseField_Mechanism :1007-1152, with PhaseField_Normal :1958-2032) accumulating Jhat or
 * Div_Jhat, then PhaseField_GradQPF with the explicit RC (:2315-2403) or the implicit one
 * (:2133-2310: div(b grad Q) = Div_Jhat solved by conjugate gradients to the rounding floor where
 * the reference calls AMReX's MLMG); for the N-phase model once per phase.
 * gradq: device face buffer (written); *gradq_max (host) = max |Grad_Q| over all faces;
 * info (host, 4 doubles, or NULL) = max|u|, mobility, eta, number of mechanism sub-steps.
 * Synchronises the stream. */
int cmp_level_pf_gradq(cmp_level* level, const double* U, double dt, double* gradq,
                       double* gradq_max, double* info, void* stream);

/* Iterations of the last implicit Grad_Q solve of cmp_level_pf_gradq (0 with the explicit RC).
 * No counterpart in the reference (MLMG reports through its verbosity); for diagnostics. */
int cmp_level_pf_solver_iterations(const cmp_level* level, int* iterations);

/* One Phase-Field TVD-RK3 sub-step of size pf_dt in place on Un (PhaseField_ThirdOrderSSPRK,
 * :2620-2752): the stage combinations of cmp_level_rk3_stage with the Phase-Field right-hand side;
 * U1, U2 are scratch level buffers. */
int cmp_level_pf_rk3_step(cmp_level* level, double* Un, double* U1, double* U2,
                          const double* gradq, double pf_dt, void* stream);

/* One Phase-Field sub-step of pf_dt at a fixed Grad_Q with run.TimeIntegrator (PhaseField_ForwardEuler
 * :2426-2505, PhaseField_SecondOrderSSPRK :2508-2617, PhaseField_ThirdOrderSSPRK, PhaseField_FourthOrderRK
 * :2756-): integrator 1..4 as in cmp_level_rk_step, U3 for RK4. cmp_level_pf_evolve's sub-steps use
 * the level's integrator (cmp_level_set_integrator). */
int cmp_level_pf_rk_step(cmp_level* level, int integrator, double* Un, double* U1, double* U2, double* U3,
                         const double* gradq, double pf_dt, void* stream);

/* The Phase-Field step that follows a flow step of size dt (PhaseField_Evolve, :425-594):
 * cmp_level_pf_gradq, then sub-steps of PhaseField_ComputeDt (:2037-2098) size,
 * pf.cfl min_d dx_d / (max|Grad_Q| + 1e-14), the last one cut to end at dt.
 * dt is read from *dt_dev (device). info (host, 6 doubles, or NULL): the 4 of cmp_level_pf_gradq,
 * max|Grad_Q| and the number of sub-steps. Synchronises the stream. */
int cmp_level_pf_evolve(cmp_level* level, double* Un, double* U1, double* U2, const double* dt_dev,
                        double* info, void* stream);

/* ---- refinement tagging (ErrorEst, source/COMPAS.cpp:769-932; include/Tagging.H) -------------
 * One criterion of run.refine.<var>: var >= 0 is a state component; the derived variables of
 * compute_derived_var (include/Physics_Compressible5Eq/ComputeDerivedVar_Compressible5Eq.H:110-160)
 * built here are rho (sum of the partial densities), u, v, w and P. kind: state_grad (tag where the
 * largest one-sided difference to a face neighbour is >= v0, Tagging.H:172-201), state_value_less
 * (< v0), state_value_greater (> v0), state_value_range (v0 < value < v1). */
#define CMP_TAG_RHO   (-1)
#define CMP_TAG_P     (-2)
#define CMP_TAG_U     (-3)
#define CMP_TAG_V     (-4)
#define CMP_TAG_W     (-5)
#define CMP_TAG_GRAD    0
#define CMP_TAG_LESS    1
#define CMP_TAG_GREATER 2
#define CMP_TAG_RANGE   3
#define CMP_TAG_MAX_CRITERIA 16

typedef struct cmp_tag_criterion {
    int var;
    int kind;
    double v0, v1;
} cmp_tag_criterion;

/* Tags of every valid cell of a level (1 = refine, 0 = not): the OR of the criteria, in the
 * reference's per-variable loop. U: level buffer with its ghosts filled (the gradient reads the
 * face neighbours). tags: device bytes, one per valid cell, box by box in cmp_level_layout order
 * with no ghosts and one component (its offsets count bytes here). */
int cmp_level_tag(cmp_level* level, const double* U, const cmp_tag_criterion* criteria, int ncriteria,
                  unsigned char* tags, void* stream);

/* AverageDownTo (source/COMPAS.cpp:1550-1562; amrex::average_down): every coarse cell under a fine
 * box takes the mean of its ratio^dim children (sum over the children, k slowest, times
 * 1/ratio^dim). The fine level's domain is the coarse one refined by ratio and its boxes are
 * aligned to ratio; a covered coarse cell outside every coarse box is CMP_ERR_LAYOUT in the coarse
 * level's error word. Uf (fine level buffer) and Uc (coarse) are device buffers. The first call for a
 * pair of levels and ratio finds every covered cell's sources once (a device table the fine level
 * keeps: it allocates, so that call is not made under stream capture; under capture without the
 * table they are found per call); cmp_amr_advance builds its levels' tables before it captures. */
int cmp_level_average_down(const cmp_level* fine, const double* Uf, cmp_level* coarse, double* Uc, int ratio,
                           void* stream);

/* FillPatchTwoLevels (source/COMPAS.cpp:1679-1724, with the reference's interpolater
 * amrex::lincc_interp, CellConservativeLinear with linear limiting): every ghost cell of the fine
 * level's boxes is mapped to its source cell through the physical BCs and periodic wrap as the
 * same-level fill maps it; a source in a fine box is copied, any other is interpolated from the
 * coarse level (cell-conservative linear, MC-limited, one limiter factor per direction shared by
 * all components), the coarse neighbours read through the coarse level's own BCs. The fine domain
 * is the coarse one refined by ratio, with the same BCs; Uc's ghost cells are not read. A coarse
 * cell the coarse level does not hold is CMP_ERR_LAYOUT in the fine level's error word, and the
 * ghost keeps its value. Flow states only (two-phase or N-phase). The first call for a pair of
 * levels and ratio finds every ghost cell's sources once (a device table the fine level keeps: it
 * allocates, so that call is not made under stream capture; under capture without the table the
 * sources are found per call) and every call after it gathers through the table, with the same
 * values bit for bit; cmp_amr_advance builds the tables of its levels before it captures. */
int cmp_level_fill_two_levels(cmp_level* fine, double* Uf, const cmp_level* coarse, const double* Uc, int ratio,
                              void* stream);

/* A level's data after a regrid (RemakeLevel / MakeNewLevelFromCoarse, source/COMPAS.cpp: the
 * FillPatch of the new MultiFab): every cell of the new level's boxes, valid and ghost, is mapped
 * to its source cell as the ghost fill maps it; a source inside a box of the old level (old, Uold;
 * NULL for a level made from the coarse level alone) is copied, any other is interpolated from the
 * coarse level as cmp_level_fill_two_levels interpolates. The old level covers the same domain;
 * U and Uold are distinct device buffers. */
int cmp_level_fill_from(cmp_level* level, double* U, const cmp_level* old, const double* Uold,
                        const cmp_level* coarse, const double* Uc, int ratio, void* stream);

/* ---- a hierarchy of levels stepped together -------------------------------------------------
 * The reference's time step without subcycling (run.do_subcycle = 0, run.do_reflux = 1):
 * timeStepNoSubcycling (source/COMPAS.cpp:2139-2180) -> AdvanceAllLevelsStageCoupled
 * (source/Class_Compressible_PhaseField_StageCoupling.cpp:156-) and StageRHSAllLevels (:84-153):
 * every RK stage fills every level's ghost cells from the stage state (level 0 same-level, finer
 * levels two-level as cmp_level_fill_two_levels), computes every level's faces, gives the coarse
 * faces under a finer level the mean of the fine faces over them (fluxes and u*_n, finest level
 * first, AverageDownFacesTo :1575-1583), takes the divergence and the stage combination, and
 * averages the levels down (AverageDown). ComputeDt over the levels (:2250-2302, nsubsteps 1): the
 * smallest of the levels' estimates, each limited to 1.1 dt_prev. Two-phase or N-phase flow states; levels
 * coarsest first, with the same parameters and ghost width, each the refinement of the previous one
 * by ratio, boxes aligned to ratio; a third level must not touch the boundary of the second. The
 * caller's level buffers stay the caller's; every call takes arrays of nlev device pointers. */
typedef struct cmp_amr cmp_amr;

int cmp_amr_create(cmp_level* const* levels, int nlev, int ratio, cmp_amr** out);
/* The hierarchy refers to its levels until it is destroyed: destroy it before any of them (a regrid
 * destroys the hierarchy, remakes the levels, then creates a new one). */
int cmp_amr_destroy(cmp_amr* amr);
/* Coarse cells of `level` whose faces border the next finer level (0 for the finest). */
int cmp_amr_info(const cmp_amr* amr, int level, int* ncf_cells);
/* The flow stage of every level in one launch of the persistent work-list kernel, each level's boxes
 * cut into the tile shapes of the autotuner's cache (cmp_level_tune_tiles on every level first;
 * its cache holds the sweeps): force 1 on (CMP_ERR_UNSUPPORTED when a box shape has no sweep), -1
 * off, 0 on when the swept times summed over every level's boxes favour it, which cmp_amr_create
 * also applies. *use_list, *waves (or NULL): the outcome and the kernel's occupancy target. The
 * stage results are those of the per-level fixed-shape kernels bit for bit. */
int cmp_amr_tiles(cmp_amr* amr, int force, int* use_list, int* waves);
/* The hierarchy's current choice (as cmp_amr_tiles reports it), changing nothing. */
int cmp_amr_tile_info(const cmp_amr* amr, int* use_list, int* waves);
/* CalculateInitialcmax of every level on the filled hierarchy: cmax (device, nlev doubles). Writes
 * the ghost cells of U. */
int cmp_amr_initial_cmax(cmp_amr* amr, double* const* U, double* cmax, void* stream);
/* *dt_dev (device) = ComputeDt over the levels from cmax (device, nlev doubles; reset to 0 after),
 * with dt_prev = *dt_dev (0: a first step, no limiter). */
int cmp_amr_compute_dt(cmp_amr* amr, double* cmax, double* dt_dev, double cfl, int limiter, void* stream);
/* One RK3 step of size *dt_dev of every level, in place on Un[l] (U1[l], U2[l] scratch level
 * buffers); the levels' c_max accumulate into cmax[l]. */
int cmp_amr_rk3_step(cmp_amr* amr, double* const* Un, double* const* U1, double* const* U2, const double* dt_dev,
                     double* cmax, void* stream);
/* nsteps steps: initial c_max, then per step ComputeDt and the RK3 step; dt_hist (device, nsteps
 * doubles, or NULL) receives every step's dt. use_graph: one step captured in a hipGraph (kept
 * while the buffers and settings stay the same) and replayed. */
int cmp_amr_advance(cmp_amr* amr, double* const* Un, double* const* U1, double* const* U2, int nsteps, double cfl,
                    int limiter, double* dt_dev, double* cmax, double* dt_hist, int use_graph, void* stream);

/* The Phase-Field step of the hierarchy (PhaseField_Evolve over finest_level + 1 levels without
 * subcycling, source/Class_Compressible_PhaseField.cpp:425-594; the Phase-Field sub-steps by
 * PhaseField_TimeStepNoSubcycling -> PhaseField_AdvanceAllLevels,
 * source/Class_Compressible_PhaseField_StageCoupling.cpp:43-51, and AverageDown). cmp_amr_pf_setup
 * calls cmp_level_pf_setup on every level and builds the hierarchy's own work buffers; after it,
 * cmp_amr_advance ends every step with cmp_amr_pf_evolve. The mechanism sub-steps (dt_m from the
 * finest cell size) advance every level with the sub-step state filled on every level (all phases
 * at once, the coarse-fine ghosts interpolated as the two-level fill interpolates); CAC takes its two
 * sums over level 0 after the levels' mech and W are averaged down; ACDI's J is averaged down
 * (AverageDownFaces) before its divergence and before Jhat takes it. With the explicit RC every level
 * has its Grad_Q from its Jhat; with the implicit RC Grad_Q comes from one composite solve over the
 * cells no finer level covers: a face between two such cells of one level carries
 * b (Q_n - Q_c) / dx, a coarse-fine face per fine face on it b_fine (Q_coarse - Q_fine) /
 * ((dx_fine + dx_coarse) / 2) (a two-point flux where the reference's MLMG, composite_solve = 1,
 * LinearSystem_Parameter.H:16, uses its own higher-order coarse-fine stencil), solved by conjugate
 * gradients preconditioned by Jacobi on the finer levels plus a multigrid V-cycle of level 0 at every
 * cell's level-0 ancestor, to the rounding floor; covered cells then take the mean of their children
 * and the coarse faces under a finer level the mean of its faces. */
int cmp_amr_pf_setup(cmp_amr* amr, const cmp_pf_params* pf);
/* Grad_Q of every level for a flow step of size dt into every level's own Grad_Q buffer, copied into
 * gradq[l] (device face buffers of cmp_face_layout with one component per volume fraction) where
 * gradq and gradq[l] are non-null; gradq_max (host, nlev doubles, or NULL): max |Grad_Q| per level;
 * info (host, 5 doubles, or NULL): max|u| over the levels, mobility, eta, mechanism sub-steps, CG
 * iterations of the last implicit solve. Synchronises the stream. */
int cmp_amr_pf_gradq(cmp_amr* amr, double* const* U, double dt, double* const* gradq, double* gradq_max,
                     double* info, void* stream);
/* The Phase-Field step after a flow step of size *dt_dev (nothing when it is 0): Grad_Q, then
 * sub-steps of PhaseField_ComputeDt over the levels (the smallest of the levels' estimates, limited
 * to 1.1 the previous sub-step, the last one cut to end at dt), each a stage-coupled TVD-RK3 step of
 * the Phase-Field right-hand side of every level (U1, U2 scratch level buffers) with AverageDown.
 * info (host, 7 doubles, or NULL): the 5 of cmp_amr_pf_gradq in order max|u|, mobility, eta,
 * mechanism sub-steps, then max|Grad_Q| over the levels, the number of sub-steps, CG iterations.
 * Synchronises the stream. */
int cmp_amr_pf_evolve(cmp_amr* amr, double* const* Un, double* const* U1, double* const* U2, const double* dt_dev,
                      double* info, void* stream);
/* Iterations of the last composite implicit solve (0 with the explicit RC). */
int cmp_amr_pf_solver_iterations(const cmp_amr* amr, int* iterations);

/* ---- a batch of independent levels (the served-stream form) ---------------------------------
 * No counterpart in the reference, which runs one case per process (source/main.cpp:21-86): cases of
 * different grids and parameters are admitted while others run, keep their state in device memory,
 * and advance together, one launch per RK stage over the tiles of every admitted case. Every case
 * has its own parameters, geometry, boxes, dt, c_max and error word; each case's result is the one
 * the single-level entry points give for it alone. Two-phase model without THINC; all cases of a
 * batch share its dimension, scheme, limiter and Riemann solver. */
typedef struct cmp_batch cmp_batch;

/* A batch for cases like *like (dim, id_weno/id_limiter, id_riemann), with three device arenas of
 * `capacity` doubles each (a case's U^n and its two stage buffers sit at one offset in them) and up
 * to max_cases cases at once. */
int cmp_batch_create(const cmp_params* like, long long capacity, int max_cases, cmp_batch** out);
int cmp_batch_destroy(cmp_batch* batch);

/* Admits a case: a free slot and a free arena range for its level buffer (cmp_level_layout of its
 * boxes with nstate comps and ng ghosts; CMP_ERR_NOMEM when either is missing), its initial state
 * U0 (host, that layout) and its initial c_max; its steps then use ComputeDt with cfl and limiter
 * (the first one without the limiter, as cmp_level_advance). Synchronises the stream. */
int cmp_batch_admit(cmp_batch* batch, const cmp_params* params, const cmp_geometry* geom, int nbox,
                    const cmp_box* boxes, int ng, const double* U0, double cfl, int limiter, int* case_id,
                    void* stream);

/* Releases a case's slot and range (after the stream's queued work). */
int cmp_batch_retire(cmp_batch* batch, int case_id, void* stream);

/* One RK3 step of every admitted case, each with its own dt: one ComputeDt launch for all cases,
 * then three stage launches over all their tiles (cells beyond a box read from their source boxes,
 * as in cmp_level_rk3_step). */
int cmp_batch_step(cmp_batch* batch, void* stream);

/* A case's U^n (device pointer into the arena, its level layout, *len doubles) and, when time or
 * steps is non-null, its simulated time and step count (synchronises the stream). */
int cmp_batch_case_state(cmp_batch* batch, int case_id, double** Un, long long* len, double* time,
                         int* steps, void* stream);

/* Copies a case's U^n (its level layout, the *len of cmp_batch_case_state) into host memory;
 * synchronises the stream. */
int cmp_batch_case_download(cmp_batch* batch, int case_id, double* host, void* stream);

/* Reads and clears a case's error word (as cmp_level_check), its flow step's and its Phase-Field
 * step's together. */
int cmp_batch_check(cmp_batch* batch, int case_id, void* stream);

/* The Phase-Field step of the served cases (PhaseField_Evolve after every flow step, as each case's
 * own run has it, source/Class_Compressible_PhaseField.cpp:425-594): every admitted case, and every
 * case admitted later, gets the Phase-Field buffers of cmp_level_pf_setup with these parameters
 * (device memory outside the arenas). CAC with the explicit RC is CMP_ERR_UNSUPPORTED. */
int cmp_batch_pf_setup(cmp_batch* batch, const cmp_pf_params* pf);
/* After cmp_batch_step on `stream`: the Phase-Field step of every admitted case with the dt of its
 * flow step, the cases concurrently, each on a stream of its own driven by one of `nthreads` host
 * workers (0: 8), since a case's sub-step counts and solver iterations follow from its own numbers
 * (cmp_level_pf_evolve per case, so a case's result does not depend on what else is in the batch).
 * info (host, 6 doubles per case slot, or NULL): cmp_level_pf_evolve's info of each admitted case.
 * Returns when every case's step is done. */
int cmp_batch_pf_step(cmp_batch* batch, int nthreads, double* info, void* stream);

/* Admitted cases, tiles of the last rebuilt table (-1 before the next step rebuilds it) and free
 * arena doubles. */
int cmp_batch_info(const cmp_batch* batch, int* ncases, int* ntiles, long long* free_doubles);

/* ---- allocation-free entry point -----------------------------------------------------------
 * No counterpart in the reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). The same right-hand side as
 * cmp_level_create + cmp_level_rhs (every box of a level, U's ghost cells already filled, dUdt a
 * level buffer laid out like U, *cmax max-accumulated when cmax is non-null, the device error word
 * *err max-accumulated when err is non-null), with the box table and tile work-list that
 * cmp_level_create allocates kept instead in a fixed device array of the library, one per GPU,
 * reused by every call: no hipMalloc, hipFree or other allocating routine is called. Limits:
 * nbox <= CMP_NOALLOC_MAX_BOXES and at most CMP_NOALLOC_MAX_TILES tiles (16x15 cells in 2D, 8x8x3
 * in 3D, per box), else CMP_ERR_SCRATCH before any GPU work. The tables are written in stream order,
 * so calls on one stream follow each other; calls that overlap on one GPU (other streams, other
 * threads) race on the array. */
#define CMP_NOALLOC_MAX_BOXES 4096
#define CMP_NOALLOC_MAX_TILES 65536
int cmp_level_rhs_noalloc(const cmp_params* params, const cmp_geometry* geom, int nbox, const cmp_box* boxes,
                          int ng, const double* U, double* dUdt, double* cmax, int* err, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* COMPAS_AMD_H */
