// cmp_level.hpp - the level object behind the C ABI's opaque cmp_level handle (internal to the
// library: shared by the level entry points and the batch of levels).
#pragma once

#include <vector>

#include "../../include/compas_amd.h"
#include "cmp_kernels.hpp"

namespace cmp {

// HIP status -> CMP_OK / CMP_ERR_HIP (printing the HIP error)
int check_hip(hipError_t e);

// The halo sources of tile [tlo, thi] of box b on its box's faces (cmp_capi.hip); owner, nb, bs:
// the level's owner map of bs^dim blocks.
void tile_halo(int dim, const cmp_geometry* geom, const cmp_box* boxes, int b, const int* tlo, const int* thi,
               int ng, const std::vector<int>& owner, const int* nb, long long bs, int (*halo)[2]);

}  // namespace cmp
struct cmp_level;
namespace cmp {
// The work-list form of a level's flow stage (cmp_tune.hip): one stage through the level's list
// (fused ghost reads, or the filled ghost cells when fused is false), and the tiles of every box of a
// level by its box_ext with level index lev.
int level_list_stage(cmp_level* L, int mode, int stage, const double* Us, const double* Un, double* Uout,
                     const double* dt, double* cmax, hipStream_t s, bool fused = true);
void level_list_tiles(const cmp_level* L, int lev, std::vector<ListTile>& out);
// The tuned choices of L from the autotuner's cache: per occupancy target (2, 3 waves per SIMD) each
// box's tile extents and the level's summed time, and the fixed kernel's summed time; with ext_mixed
// and t_mixed, per occupancy target the per-box choice of the faster kernel (0 0 0 for a box the
// fixed kernel keeps) and its summed time. CMP_ERR_UNSUPPORTED when a box shape of L has no cached
// sweep.
int level_tuned(const cmp_level* L, std::vector<int>* ext, double* t_list, double* t_fixed,
                std::vector<int>* ext_mixed = nullptr, double* t_mixed = nullptr);
// The gather plan of fine level F's two-level fill from coarse level C at ratio, built when F holds
// none for that pair (cmp_capi.hip); allocates device memory, so not under stream capture.
int level_fill2_plan(cmp_level* F, const cmp_level* C, int ratio, hipStream_t s);
// The gather plan of fine level F's average-down onto coarse level C at ratio, built when F holds
// none for that pair (cmp_capi.hip); allocates device memory, so not under stream capture.
int level_avg_plan(const cmp_level* F, const cmp_level* C, int ratio, hipStream_t s);

}  // namespace cmp

using cmp::BoxDev;
using cmp::FillDev;
using cmp::KParams;
using cmp::LaunchCfg;
using cmp::PfGeom;
using cmp::TileDev;

#define CMP_TRY(x) do { int _r = cmp::check_hip(x); if (_r) return _r; } while (0)

struct cmp_level;
namespace cmp {
// the multigrid preconditioner of a level that covers its domain (cmp_capi.hip): face coefficients
// of every grid from the level's d_Bf, and one V-cycle from grid m's residual (its result in
// mg[m].x[mg[m].cur])
int mg_prepare(cmp_level* L, hipStream_t s);
int mg_vcycle(cmp_level* L, const double* sc, size_t m, hipStream_t s);
}  // namespace cmp

struct cmp_level {
    cmp_params P;
    cmp_geometry G;
    int ng;
    int nbox;
    std::vector<cmp_box> boxes;
    KParams kp;
    LaunchCfg lc;
    BoxDev* d_boxes = nullptr;
    FillDev* d_fd = nullptr;        // device copy of fd (the stage kernel's fused ghost reads)
    TileDev* d_tiles = nullptr;
    int ntiles = 0;
    long long* d_gprefix = nullptr;
    long long nghost = 0;
    int* d_owner = nullptr;
    FillDev fd;
    int* d_err = nullptr;
    int* d_step = nullptr;
    hipStream_t own_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipGraphExec_t cg_exec = nullptr;   // a chunk of the implicit Grad_Q's CG iterations (its buffers)
    const void* gkey[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    double gcfl = 0.0;
    int glimiter = -1;

    // Phase-Field step (cmp_level_pf_setup)
    bool pf_on = false;
    cmp_pf_params pf{};
    LaunchCfg pf_lc{};
    KParams pf_kp{};
    BoxDev* d_abox = nullptr;       // alpha1 layout: 1 ghost, 1 comp
    BoxDev* d_nbox = nullptr;       // normals layout: 1 ghost, dim comps
    long long* d_gprefix1 = nullptr;
    long long nghost1 = 0;
    FillDev fdA{}, fdN{};
    PfGeom pg{};
    long long maxcells = 0, maxfaces[3] = {0, 0, 0};
    long long na = 0, nfaces = 0;
    double* d_A0 = nullptr;         // alpha1 of the step, PF_bcs ghosts
    double* d_A = nullptr;          // alpha1 advanced by the mechanism sub-steps
    double* d_P = nullptr;          // psi of d_A; then the re-sharpened alpha1 of d_A0
    double* d_N = nullptr;          // cell normals
    double* d_J = nullptr;          // face J of the current sub-step
    double* d_H = nullptr;          // face Jhat
    double* d_gq = nullptr;         // face Grad_Q (cmp_level_pf_evolve), one component per volume fraction
    double* d_gq1 = nullptr;        // one phase's Grad_Q (N-phase model)
    unsigned long long* d_bits = nullptr;   // [0] max|u|, [1] max|Grad_Q| as bit patterns
    double* d_pfdt = nullptr;       // Phase-Field sub-step size read by the stage kernels
    double* h_pin = nullptr;        // pinned host scalars
    // CAC mechanism and the implicit RC's solve (alpha layout unless noted)
    double* d_Mk = nullptr;         // mech of every cell
    double* d_Wk = nullptr;         // alpha (1 - alpha)
    double* d_DJ = nullptr;         // Div_Jhat
    double* d_X = nullptr;          // Q
    double* d_R = nullptr;          // CG residual
    double* d_Pv = nullptr;         // CG search direction (its ghosts filled like alpha's)
    double* d_AP = nullptr;         // operator applied to d_Pv
    double* d_Z = nullptr;          // preconditioned residual
    double* d_ID = nullptr;         // inverse diagonal of the operator (Jacobi preconditioner)
    double* d_Bf = nullptr;         // face coefficients b (face layout)
    double* d_part = nullptr;       // block partial sums, 4 per block
    double* d_sc = nullptr;         // CG_NSC device scalars
    int nblocks = 0;                // blocks of a cell pass over the level
    long long ncell = 0;            // valid cells of the level
    int cg_iters = 0;               // iterations of the last implicit solve
    // Phase-Field sub-steps run on the boxes with a nonzero Grad_Q face and the boxes whose valid
    // cells fill their ghosts; every other box has a zero Phase-Field flux on all its faces
    std::vector<std::vector<int>> ghost_src;   // per box: boxes its ghost cells are filled from
    int* d_boxflag = nullptr;       // per box: a nonzero Grad_Q face (set by the Grad_Q kernels)
    int* h_boxflag = nullptr;       // pinned copy
    TileDev* d_pf_tiles = nullptr;  // tiles of the selected boxes
    cmp::BoxBlock* d_cmap = nullptr;  // compact block list of the level's cells (PfGeom::cmap)
    TileDev* h_pf_tiles = nullptr;  // pinned staging of the list
    int pf_ntiles = 0, pf_nfill = 0;

    // host copies of the level tables (tiles box by box, the owner map)
    std::vector<TileDev> h_tiles;
    std::vector<BoxDev> h_boxes;    // the box table as uploaded
    long long* d_toff = nullptr;    // each box's first tag (cmp_level_tag)
    long long tag_maxcells = 0;
    std::vector<int> tile_start;    // first tile of each box, nbox + 1 entries
    std::vector<int> h_owner;
    int integrator = 3;             // run.TimeIntegrator of cmp_level_advance and the Phase-Field sub-steps
    double* d_U3 = nullptr;         // RK4's fourth level buffer
    int ginteg = 0;                 // the integrator the captured step was built with
    // multigrid preconditioner of the implicit Grad_Q solve (a level that covers its domain): dense
    // grids from the level's resolution down, each with its iterate pair, residual, inverse diagonal
    // and face coefficients
    struct MgLev {
        cmp::MgGrid G;
        double* x[2] = {nullptr, nullptr};
        double* r = nullptr;
        double* id = nullptr;
        double* b[3] = {nullptr, nullptr, nullptr};
        int cur = 0;
    };
    std::vector<MgLev> mg;
    int mg_bottom_sweeps = 0;
    double mg_omega = 0.8;
    double* d_clock = nullptr;      // [0] the level's time (advanced by every ComputeDt), [1] stop_time
    double h_clock[2] = {0.0, -1.0};  // host staging of cmp_level_set_time
    double stop_time = -1.0;        // the last stop_time given to cmp_level_set_time
    long long bs = 1;
    int nb[3] = {1, 1, 1};

    // The flow stage as the persistent work-list kernel (cmp_level_set_tiles, cmp_level_tune_tiles):
    // per box its tile extents, the tiles, the one-level table (fused ghost reads) and the list's
    // device counters; use_list routes cmp_level_rk3_step and cmp_level_advance through it.
    bool use_list = false;
    std::vector<int> box_ext;       // 3 ints per box
    std::vector<cmp::ListTile> h_ltiles;
    cmp::ListTile* d_ltiles = nullptr;
    cmp::ListLevel* d_llev = nullptr;
    int* d_lctr = nullptr;
    int list_grid = 0;
    int list_waves = 2;             // the work-list kernel's occupancy target (waves per SIMD)
    // a level that mixes the two kernels: the tiles of the boxes the fixed-shape kernel keeps (box_ext
    // 0 0 0), stepped by a k_stage launch beside the work-list launch over the other boxes
    TileDev* d_ftiles = nullptr;
    int nftiles = 0;
    unsigned tiles_gen = 0;         // raised by every change of the above (a hierarchy's graph key)

    unsigned long long uid = 0;     // unique per level created in this process
    // the two-level fill's sources of every ghost cell (F2Cell, one per ghost cell in fd.gprefix
    // order), for the coarse level f2_uid at ratio f2_ratio (0: none built)
    void* d_f2plan = nullptr;
    unsigned long long f2_uid = 0;
    int f2_ratio = 0;
    // the average-down onto the coarser level avg_uid at ratio avg_ratio (0: none built): per box the
    // prefix sums of the coarse cells it covers (nbox + 1), avg_total in all, and each covered cell's
    // sources (AvgCell)
    mutable long long* d_avgprefix = nullptr;
    mutable cmp::AvgCell* d_avgplan = nullptr;
    mutable unsigned long long avg_uid = 0;
    mutable int avg_ratio = 0;
    mutable long long avg_total = 0;
};
