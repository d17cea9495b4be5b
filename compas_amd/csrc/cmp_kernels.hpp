// cmp_kernels.hpp - host-visible launch interface of the gfx950 kernels (internal to the library).
#pragma once

#include <hip/hip_runtime.h>
#include "cmp_device.hpp"

namespace cmp {

// One box of a level as the kernels see it: valid range, fab placement and strides.
struct BoxDev {
    int lo[3], hi[3];
    int flo[3];          // lowest stored index of the fab per dim
    int pad;             // the box's case in a batch of levels (0 in a single level)
    long long off;       // element offset of the fab in the level buffer
    long long sj, sk, sn;
    // output layout (cmp_rhs_box writes dUdt over the valid box only)
    long long ooff;
    int olo[3];
    int pad2;
    long long osj, osk, osn;
    // offsets of this box's face arrays in a level face buffer (cmp_face_layout with one
    // component per volume fraction)
    long long goff[3];
};

// Where the cells beyond a tile's box along direction d on side s (0 low, 1 high) come from when
// the stage kernel reads them itself (fused ghost fill): a box index (the cells lie in that box,
// after the periodic wrap along d), or one of the codes below; only a tile that touches its box's
// face on that side reads there.
enum HaloSrc { HALO_LOOKUP = -1, HALO_CLAMP = -2, HALO_MIRROR = -3, HALO_MIRROR_ALL = -4 };

struct TileDev {
    int box;
    int lo[3];
    int halo[3][2];
};

constexpr int NT = 256;   // threads per workgroup of every kernel here

// (x, y, z) of position p < n0 n1 n2 of a block of n0 x n1 x n2 positions, x fastest; 32-bit
// arithmetic (every box's and dense grid's position count is below 2^31, which level creation
// checks), where 64-bit division would cost an emulated sequence per thread
__device__ __forceinline__ void pos3(long long p, int n0, int n1, int& x, int& y, int& z)
{
    const int q = (int)p;
    const int a = q / n0;
    x = q - a * n0;
    z = a / n1;
    y = a - z * n1;
}

__device__ __forceinline__ long long fab_index(const BoxDev& bx, int i, int j, int k)
{
    return (long long)(i - bx.flo[0]) + bx.sj * (j - bx.flo[1]) + bx.sk * (k - bx.flo[2]);
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = smax(v, __shfl_xor(v, off, 64));
    return v;
}

// Ghost-fill tables of a level.
struct FillDev {
    const BoxDev* boxes;
    const long long* gprefix;   // nbox+1 prefix sums of ghost-cell counts
    const int* owner;           // box owning each bs^d block of the domain (-1 if none)
    int nbox;
    int ng;
    int bs;
    long long maxghost;         // ghost cells of the box with the most
    int nb[3];
    int dlo[3], dhi[3];
    int periodic[3];
    int lo_bc[3], hi_bc[3];
};

// The sources of one ghost cell of a fine level's two-level fill, found once for a pair of levels
// (k_fill2_plan): slot 0 the same-level source or the coarse cell under it, slots 1 + 2d / 2 + 2d the
// coarse neighbours below / above along d, each as its level's box index and the cell's offset in
// that box's fab; flip bits 0-2 the ghost cell's momentum flips, bits 3(k + 1).. those of slot k.
template <int DIM>
struct F2Cell {
    static constexpr int NS = 1 + 2 * DIM;
    int box[NS];
    int cell[NS];
    int dbox, dcell;   // the ghost cell itself: its box and its offset in that box's fab
    unsigned flip;
    int kind;   // bit 0: interpolated from the coarse level; bits 8 + 8d: offset in the coarse cell along d
    int err;    // nonzero: the fill's error code for this cell (nothing written)
};

// One covered coarse cell of an average-down, found once for a pair of levels (k_avgdown_plan): the
// fine box and the fab offset of the cell's first fine cell, the owning coarse box and the cell's
// offset there (cbox < 0: the error code -cbox, nothing written).
struct AvgCell {
    int fbox, fcell, cbox, ccell;
};

enum StageMode { MODE_RK = 0, MODE_RHS = 1, MODE_CMAX = 2 };

// One case of a batch of levels as the per-case ComputeDt sees it.
struct BatchSlot {
    double dx[3];
    double cfl;
    int dim;
    int limiter;
    int active;
    int pad;
};

// Scheme selector: 0 Godunov, 1 MUSCL (limiter from id_limiter), 5 WENO5, 6 WENO5B; THINC on top
// of WENO5 or WENO5B. riemann: 1 HLL, 2 HLLC for the flow's fluxes; RIEMANN_PF_MIX / RIEMANN_PF_PHASIC
// select the Phase-Field flux (UD, mixture or phasic formulation), read with a Grad_Q face buffer.
enum { RIEMANN_PF_MIX = 10, RIEMANN_PF_PHASIC = 11 };
struct LaunchCfg {
    int dim;
    int scheme;
    int limiter;
    int riemann;
    int thinc;
    int nphase;     // 0: two-phase five-equation model; 2, 3: N-phase model
};

hipError_t launch_stage(const LaunchCfg& lc, const KParams& kp, int mode, int stage,
                        const TileDev* tiles, int ntiles, const BoxDev* boxes,
                        const BoxDev* single_box, int single_ntiles,
                        const double* Us, const double* Un, double* Uout, const double* dt,
                        unsigned long long* cmax_bits, int* err, hipStream_t s,
                        const double* gradq = nullptr, const FillDev* fuse = nullptr);
// fuse: a FillDev in device memory; the stage then reads the cells beyond each tile's box where the
// ghost fill would take them from, instead of the box's ghost cells

// ---- persistent work-list stage kernel (k_stage_list) ----
// A tile of a work-list that may span several levels, each tile with its own shape: level (index
// into the launch's level table), box, lowest cell, extents, halo sources as in TileDev.
struct ListTile {
    int lev;
    int box;
    int lo[3];
    int ext[3];
    int halo[3][2];
};
// What the work-list kernel needs of one level: parameters, box table, fused ghost-fill table in
// device memory (null: the stage input's ghost cells are read) and error word.
struct ListLevel {
    KParams kp;
    const BoxDev* boxes;
    const FillDev* fd;
    int* err;
};
constexpr int LIST_MAXLEV = 8;
// The per-launch state pointers of the work-list kernel, by level.
struct ListCall {
    const double* Us[LIST_MAXLEV];
    const double* Un[LIST_MAXLEV];
    double* Uout[LIST_MAXLEV];
    unsigned long long* cmax[LIST_MAXLEV];
};
// ints of a work-list's device counters: one per XCD chunk of the list, then the workgroups done
constexpr int LIST_NCTR = 9;
// The flow stage of every tile of a work-list (two-phase or N-phase model, not the Phase-Field
// flux). waves: the kernel's occupancy target (2 or 3 waves per SIMD); grid: workgroups of the
// persistent launch (list_capacity); ctr: LIST_NCTR ints, zero before the first launch and left
// zero by every launch.
hipError_t launch_stage_list(const LaunchCfg& lc, int waves, const ListLevel* levs, const ListTile* tiles, int ntiles,
                             int* ctr, const ListCall& call, int mode, int stage, const double* dt, int grid,
                             hipStream_t s);
// Workgroups of the work-list kernel of lc (occupancy target waves) resident on the device at once
// (0: no such kernel).
int list_capacity(const LaunchCfg& lc, int waves);
// Whether a tile shape fits the work-list kernel's LDS and thread layout for lc.
bool tile_shape_ok(const LaunchCfg& lc, const int* ext);

// N = 4 + dim: the flow state (reflecting BCs flip momenta); N = 1 or dim: a Phase-Field field
// (alpha1 or the normals), filled without sign changes. With `list`, only the ghosts of the nlist
// boxes it names are filled.
// One RK stage of every tile of a batch of levels (BoxDev::pad = the box's case; per-case parameters,
// dt, c_max bits, error word and fill tables indexed by it). Two-phase model, no THINC.
hipError_t launch_stage_batch(const LaunchCfg& lc, const KParams* kps, int stage, const TileDev* tiles, int ntiles,
                              const BoxDev* boxes, const double* Us, const double* Un, double* Uout,
                              const double* dts, unsigned long long* cmax_bits, int* errs, const FillDev* fds,
                              hipStream_t s);
// ComputeDt of every active case of a batch; advances each case's time and step count
hipError_t launch_batch_dt(const BatchSlot* slots, int nslot, double* cmax, double* dt, double* time, int* steps,
                           int* errs, hipStream_t s);
// refinement tags of every valid cell (one grid row per box; the criteria travel as a kernel argument)
struct TagCriteria {
    cmp_tag_criterion c[CMP_TAG_MAX_CRITERIA];
    int n;
};
// A coarse cell next to a finer level: box, cell, bit 2 d + side set where its face of direction d
// (side 0 low, 1 high) borders a coarse cell the finer level covers, and the first of its face
// jobs (per bordering face, in (d, side) order: the coarse face, then the ratio^(dim-1) fine faces
// over it, the lower transverse index fastest).
struct CfCell {
    int box;
    int c[3];
    int mask;
    int job0;
};

// One face of the coarse-fine correction: the face between cells c - e_dir and c of box `box` of
// the coarse (fine = 0) or the finer (fine = 1) level.
struct CfFace {
    int box;
    int fine;
    int dir;
    int c[3];
};

// Up to this many levels in a hierarchy.
constexpr int AMR_MAX_LEVELS = 8;



struct AmrDt {
    int nlev;
    int dim;
    double dx[AMR_MAX_LEVELS][3];
};


// The coarse-fine face correction of one stage (cmp_amr_rk3_step): every listed coarse cell's
// stage output gains w dt times the change of its right-hand side when its faces over the finer
// level take the finer level's face means.
// Two launches: (part & 1) every face job (one thread each) into faces (N + 1 doubles a job: flux,
// u*_n), (part & 2) one thread per listed cell.
hipError_t launch_cf_correct(const LaunchCfg& lc, const KParams& kp, int part, const CfCell* list, int n,
                             const CfFace* jobs,
                             int njobs, double* faces, const BoxDev* cboxes, const double* Uc, const BoxDev* fboxes,
                             const double* Uf, int ratio, double* Uout, const double* dt, int stage, int* err,
                             hipStream_t s);
// One pair of levels of a hierarchy in the fused flow correction (launch_cf_apply_all): the coarse
// level's parameters, its listed cells (start: the pair's first thread of the launch), box table and
// error word, and the pair's face values.
struct CfPairDev {
    KParams kp;
    const CfCell* list;
    int n, start;
    const BoxDev* cboxes;
    const double* faces;
    int* err;
};
// Per pair of the launch, the coarse level's stage state and stage output.
struct CfStatePtrs {
    const double* Uc[LIST_MAXLEV];
    double* Uout[LIST_MAXLEV];
};
// The cell part (part 2) of launch_cf_correct for every pair of levels in one launch: ntotal threads,
// pairs[] in device memory in order of start. Flow stages only (not the Phase-Field flux).
hipError_t launch_cf_apply_all(const LaunchCfg& lc, const CfPairDev* pairs, int npairs, int ntotal,
                               const CfStatePtrs& st, int ratio, const double* dt, int stage, hipStream_t s);
// out = (p X + q Y) + (r dt) L on the valid cells of a level (Y, L optional), without contraction.
hipError_t launch_lincomb(const BoxDev* boxes, int nbox, long long maxcells, int N, double* out, double p,
                          const double* X, double q, const double* Y, double r, const double* L, const double* dt,
                          hipStream_t s);
// ComputeDt over the levels of a hierarchy (no subcycling); resets every level's c_max. stepc:
// device step counter and dt-history capacity.
hipError_t launch_amr_dt(const AmrDt& a, double* cmax, double* dt, double cfl, int limiter, double* hist, int* stepc,
                         int* err, hipStream_t s);

hipError_t launch_tag(const LaunchCfg& lc, const KParams& kp, const BoxDev* boxes, int nbox, long long maxcells,
                      const long long* toff, const double* U, const TagCriteria& tc, unsigned char* tags, int* err,
                      hipStream_t s);
hipError_t launch_fill(int dim, int N, const FillDev& fd, long long nghost, double* U, int* err,
                       hipStream_t s, const int* list = nullptr, int nlist = 0);

// ---- Phase-Field mechanism kernels (grid.y = box) ----
// PfGeom: per-level constants of the mechanism kernels.
// One workgroup of a compact launch over the cells of a level's boxes: the box and the block of NT
// cells inside it. Only blocks that hold cells are listed, so a level of many small boxes launches no
// empty workgroups (a launch sized by its largest box would, for every smaller box).
struct BoxBlock {
    int box;
    int blk;
};

struct PfGeom {
    const BoxDev* hbox;    // flow-state layout (ng ghosts, nstate comps)
    const BoxDev* abox;    // alpha layout (1 ghost, 1 comp); normals use DIM * off with the same strides
    int nbox;
    double dx[3];
    int dlo[3], dhi[3];    // domain index range
    int periodic[3];
    const BoxBlock* cmap;  // the compact block list of the level's cells (null: one grid row per box)
    int ncmap;
};

// The box and cell index of this thread in a launch over a level's cells: with a compact block list,
// workgroup x is its entry x; without, workgroup (x, y) is block x of box y (cell_grid sets the grid
// to match). Block sums index their partials by blockIdx.y * gridDim.x + blockIdx.x in both forms.
__device__ __forceinline__ int cell_block(const BoxBlock* map, long long& q)
{
    if (map) {
        const BoxBlock bb = map[blockIdx.x];
        q = (long long)bb.blk * NT + threadIdx.x;
        return bb.box;
    }
    q = (long long)blockIdx.x * NT + threadIdx.x;
    return (int)blockIdx.y;
}
inline dim3 cell_grid(const BoxBlock* map, int nmap, long long maxcells, int nbox)
{
    if (map) return dim3((unsigned)(nmap > 0 ? nmap : 1));
    const long long b = (maxcells + NT - 1) / NT;
    return dim3((unsigned)(b > 0 ? b : 1), (unsigned)nbox);
}

// A dense domain-shaped grid of the multigrid preconditioner (cmp_mg.hip): extents n (1 in inactive
// dimensions), cells i fastest; face arrays of direction d have extents n + e_d.
struct MgGrid {
    int n[3];
    int periodic[3];
    double ih2[3];      // 1 / h_d^2
    long long nc;       // cells
};

enum MgOp {
    MG_GATHER_R, MG_SCATTER, MG_GATHER_B, MG_RESTRICT_B, MG_DIAG, MG_SMOOTH, MG_RESID_RESTRICT, MG_PROLONG,
    MG_BOTTOM, MG_CG_RZ
};

// Arguments of one multigrid operation (the fields an operation does not use are ignored):
//   GATHER_R: src (box layout) -> dst (dense G); SCATTER: src (dense G) -> dst (box layout);
//   GATHER_B: src (box face buffer) -> b[d] (dense faces); RESTRICT_B: b[d] (G) -> bc[d] (Gc);
//   DIAG: b -> dst = 1/diag; SMOOTH: dst = src + omega id (r - A src) (zero_x: src = 0);
//   RESID_RESTRICT: dst (Gc) = mean of the children of r - A src (G); PROLONG: dst (G) += src (Gc);
//   BOTTOM: dst = `sweeps` Jacobi sweeps from 0 on r (G at most 1024 cells);
//   CG_RZ: r.z partials of src (R) and r (Z) into part's second slot, dst (P) = Z when non-null.
struct MgArgs {
    PfGeom g;
    long long maxcells;
    long long maxfaces[3];
    const double* sc;
    const double* src;
    double* dst;
    const double* r;
    const double* id;
    double* b[3];
    double* bc[3];
    double* part;
    MgGrid G, Gc;
    double omega;
    int zero_x;
    int sweeps;
};

template <int DIM>
hipError_t launch_mg_dim(int op, const MgArgs& a, hipStream_t s);

// The coarse end of a V-cycle in one workgroup: grids 0..nlev-1 of the tail (the coarsest last),
// the residual of grid 0 ready in r[0]; the same operations in the same order as the per-grid
// launches (two damped Jacobi sweeps, residual restriction, the coarser grids, prolongation, two
// sweeps; the last grid `sweeps` sweeps from zero), the result in x1[0] (x0[0] when nlev is 1).
// Grids lds_from..nlev-1 (lds_from >= 1) live in the workgroup's LDS for the whole tail (their
// iterates, residual, inverse diagonal and face coefficients at the double offsets loff[m][0..6]:
// x0, x1, r, id, b0, b1, b2; id and b copied in first), the others in device memory.
constexpr int MG_TAIL_MAX_LEVELS = 16;
constexpr long long MG_TAIL_MAX_CELLS = 4096;   // grids at most this size go to the tail
constexpr int MG_TAIL_LDS = 9216;               // doubles of LDS for the tail's coarsest grids
struct MgTail {
    int nlev;
    int sweeps;
    double omega;
    const double* sc;
    MgGrid G[MG_TAIL_MAX_LEVELS];
    double* x0[MG_TAIL_MAX_LEVELS];
    double* x1[MG_TAIL_MAX_LEVELS];
    double* r[MG_TAIL_MAX_LEVELS];
    const double* id[MG_TAIL_MAX_LEVELS];
    const double* b[MG_TAIL_MAX_LEVELS][3];
    int lds_from;
    int loff[MG_TAIL_MAX_LEVELS][7];
};
template <int DIM>
hipError_t launch_mg_tail_dim(const MgTail& t, hipStream_t s);

// ---- the composite implicit Grad_Q of a hierarchy (cmp_amr_pf.hip) ----------------------------------
// mask of a level's cell (alpha layout): covered by the next finer level, a regular cell (five- or
// seven-point stencil over the same-level ghosts), or >= 0: the index of its CompCell
constexpr int COMP_COVERED = -2;
constexpr int COMP_REGULAR = -1;

// one link of a cell with a coarse-fine face: the neighbour (level, index in its alpha layout), the
// face whose b it takes (level, index in that level's one-component face buffer) and area / distance
struct CompLink {
    int lev;
    int flev;
    long long idx;
    long long face;
    double geo;
};
struct CompCell {
    int nlink;
    int link0;
};
struct CompLev {
    const BoxDev* abox;     // alpha layout (1 ghost, 1 comp), goff = one-component face offsets
    const BoxBlock* cmap;   // the compact block list of its cells (PfGeom::cmap)
    int ncmap;
    FillDev fd;             // owner map and BCs of that layout
    const int* mask;
    double* v;              // the vector the operator reads (or the solution, for the gradient)
    const double* bf;       // b on the level's faces
    double dx[3];
    double area[3];         // face areas per direction
    double vol;             // cell volume
    const double* s;        // preconditioner: the residual per volume, covered cells the mean of
                            // their children
    const double* idl;      // preconditioner: 1 / the diagonal of the level's own operator, every cell
};
struct CompArgs {
    int nlev;
    int ratio;
    CompLev L[AMR_MAX_LEVELS];
    const CompCell* cells;
    const CompLink* links;
};

enum AmrPfOp {
    AP_AVGDOWN_FACES, AP_AXPY, AP_MAXABS, AP_COMP_AP, AP_COMP_DIAG, AP_COMP_INIT, AP_COMP_S,
    AP_COMP_Z, AP_COMP_GRAD
};

// Arguments of one hierarchy Phase-Field operation (unused fields ignored):
//   AVGDOWN_FACES: fboxes (fine alpha layout, nbox boxes, maxfaces) / cfd (coarse) src -> dst;
//   AXPY: dst += a src (n); MAXABS: bits = max(bits, max |src|) (n); SUM_LEVELS: dst[v] = sum_l src[l nvals + v];
//   COMP_AP / COMP_DIAG: level lev of *ca: dst = A v (partials of v.Av into part) / dst = 1/diag;
//   COMP_INIT: mean_sum null: partials of V rhs (src); else dst = X = 0, dst2 = R, partials of r.r;
//   COMP_S: dst = src / V; COMP_Z: dst = src (R) src2 (ID) [levels > 0] + src3 (x0 dense, G0) at the
//   ancestor; COMP_GRAD: dst = Grad_Q of the level's faces from ca->L[lev].v.
struct AmrPfArgs {
    const BoxDev* fboxes;
    FillDev cfd;
    int nbox;
    long long maxcells;
    long long maxfaces[3];
    int ratio;
    const double* src;
    const double* src2;
    const double* src3;
    double* dst;
    double* dst2;
    long long n;
    double a;
    unsigned long long* bits;
    int nlev, nvals;
    const double* sc;
    const CompArgs* ca;
    int lev;
    double* part;
    const double* mean_sum;
    MgGrid G0;
    int dlo0[3];
};

template <int DIM>
hipError_t launch_amrpf_dim(int op, const AmrPfArgs& p, hipStream_t s);


// volume fraction `comp` of the flow state -> an alpha-layout buffer (and a second one, A2, if given)
hipError_t launch_pf_extract(int dim, const PfGeom& g, long long maxcells, const double* U, double* A, double* A2,
                             int comp, hipStream_t s);
// a one-component face buffer (abox offsets) -> component comp of a multi-component one (hbox offsets)
hipError_t launch_pf_scatter(int dim, const PfGeom& g, const long long* maxfaces, const double* src, double* dst,
                             int comp, hipStream_t s);
// psi (mode 0) or the re-sharpened alpha1 (mode 1) of every element of an alpha-layout buffer
hipError_t launch_pf_map(const double* A, double* out, long long n, double eta, double eta2, int mode,
                         hipStream_t s);
hipError_t launch_pf_normals(int dim, const PfGeom& g, long long maxcells, const double* Psi,
                             double* Nrm, hipStream_t s);
hipError_t launch_pf_jfaces(int dim, const PfGeom& g, const long long* maxfaces, const double* A,
                            const double* Psi, const double* Nrm, double* J, double* H, double mob_face,
                            double eta, double dt_m, hipStream_t s);
hipError_t launch_pf_update(int dim, const PfGeom& g, long long maxcells, double* A, const double* J,
                            double dt_m, hipStream_t s);
// Grad_Q from the re-sharpened alpha S and Jhat = hscale * H; boxflag (nbox ints, or null): set to 1
// for every box with a nonzero Grad_Q on one of its faces
hipError_t launch_pf_gradq(int dim, const PfGeom& g, const long long* maxfaces, const double* S,
                           const double* H, double hscale, double* gq, unsigned long long* gmax_bits,
                           int* boxflag, hipStream_t s);
hipError_t launch_pf_maxvel(int dim, const PfGeom& g, long long maxcells, const double* U,
                            unsigned long long* bits, int nmass, hipStream_t s);

// CAC mechanism and the implicit Grad_Q. Block partial sums go to `part` (one or two per block,
// gridDim.x * nbox blocks), reduced in a fixed order by launch_reduce_sums into sums[0..n-1].
hipError_t launch_pf_cac_mech(int dim, const PfGeom& g, long long maxcells, const double* A,
                              const double* Nrm, double* Mk, double* Wk, double mob, double lambda,
                              double eta, double* part, hipStream_t s);
hipError_t launch_pf_cac_update(int dim, const PfGeom& g, long long maxcells, double* A,
                                const double* Mk, const double* Wk, const double* sums, double dt_m,
                                hipStream_t s);
hipError_t launch_pf_divjhat(int dim, const PfGeom& g, long long maxcells, const double* A,
                             const double* A0, double* DJ, double dt, hipStream_t s);
hipError_t launch_pf_bcoef(int dim, const PfGeom& g, const long long* maxfaces, const double* S,
                           double* Bf, hipStream_t s);
// block partial sums of a field over the valid cells
hipError_t launch_cell_sum(int dim, const PfGeom& g, long long maxcells, const double* F, double* part,
                           hipStream_t s);
// inverse diagonal of -div(b grad .) per cell (Jacobi preconditioner; Neumann domain faces excluded)
hipError_t launch_cg_idiag(int dim, const PfGeom& g, long long maxcells, const double* Bf, double* ID,
                           hipStream_t s);
// CG pieces: r = -(rhs - mean_sum[0]/ncell), z = r/diag, p = z, x = 0, partials of r.r and r.z
hipError_t launch_cg_init(int dim, const PfGeom& g, long long maxcells, const double* rhs,
                          const double* mean_sum, double ncell, const double* ID, double* X, double* R,
                          double* Z, double* P, double* part, hipStream_t s);
// Device scalars of the CG solve (a double array of CG_NSC entries). The iteration kernels return at
// once when sc[CG_DONE] is set, so the host only polls the flag every few iterations.
enum {
    CG_RR = 0,      // r.r of the current residual
    CG_RZ = 1,      // r.z (z the preconditioned residual); the pair is reduced together
    CG_PAP = 2,     // p.Ap of this iteration
    CG_RRNEW = 3,   // r.r after this iteration's update
    CG_RZNEW = 4,   // r.z after this iteration's update
    CG_B2 = 5,      // r.r of the initial residual
    CG_DONE = 6,    // 1 once the solve has stopped
    CG_BEST = 7,    // smallest r.r so far
    CG_STALL = 8,   // iterations since the last new minimum
    CG_IT = 9,      // iterations done
    CG_SUM = 10,    // two slots for level sums (CAC's sum mech, sum W; the mean of the rhs)
    CG_ZS = 12,     // sum of z over the unknowns (the preconditioned residual's constant part)
    CG_RS = 13,     // sum of r over the unknowns; the pair is reduced together
    CG_NSC = 16
};
// r.z of the residual against the projected z - mean(z) when inv_n = 1/(number of unknowns) > 0:
// r.z - (sum z)(sum r)/n; the plain r.z when inv_n is 0
__host__ __device__ inline double cg_rz_projected(const double* sc, int slot, double inv_n)
{
    return (inv_n > 0.0) ? sc[slot] - sc[CG_ZS] * inv_n * sc[CG_RS] : sc[slot];
}
// ap = -div(b grad p) (p ghosts filled; Neumann domain faces carry no flux), partial of p.ap
hipError_t launch_cg_ap(int dim, const PfGeom& g, long long maxcells, const double* sc, const double* Bf,
                        const double* P, double* AP, double* part, hipStream_t s);
// al = rz/pAp; x += al p; r -= al ap; z = r/diag; partials of r.r and r.z (nothing when pAp <= 0)
hipError_t launch_cg_xr(int dim, const PfGeom& g, long long maxcells, const double* sc, const double* ID,
                        double* X, double* R, double* Z, const double* P, const double* AP, double* part,
                        hipStream_t s);
// be = rz_new/rz; p = z' + be p, z' = z - sum(z) inv_n on every cell not masked as covered (z' = z
// with inv_n 0; mask null: no covered cells); first: p = z' (the first direction, after cg_start)
hipError_t launch_cg_p(int dim, const PfGeom& g, long long maxcells, const double* sc, const double* Z,
                       double* P, hipStream_t s, double inv_n = 0.0, const int* mask = nullptr, int first = 0);
// the initial r.r and r.z (projected with inv_n > 0) start the stopping test
hipError_t launch_cg_start(double* sc, hipStream_t s, double inv_n = 0.0);
hipError_t launch_cg_end(double* sc, int maxit, hipStream_t s, double inv_n = 0.0);
// face gradient of x (ghosts filled), zero on Neumann domain faces, with max |.|
hipError_t launch_pf_gradx(int dim, const PfGeom& g, const long long* maxfaces, const double* X,
                           double* gq, unsigned long long* gmax_bits, int* boxflag, hipStream_t s);
// sums[v] = the sum of the nblocks partials of value v (fixed order); sc non-null: nothing once the
// CG's stop flag sc[CG_DONE] is set
hipError_t launch_reduce_sums(const double* part, int nblocks, int nvals, double* sums, hipStream_t s,
                              const double* sc = nullptr);
// The partial sums of every level of a composite solve in one launch: per level the same fixed-order
// reduction as launch_reduce_sums (into per[l * nvals + v]), then sums[v] = the levels' values added
// in level order.
struct LevelParts {
    const double* part[AMR_MAX_LEVELS];
    int nblocks[AMR_MAX_LEVELS];
    int nlev;
};
hipError_t launch_reduce_levels(const LevelParts& lp, int nvals, double* per, double* sums, hipStream_t s,
                                const double* sc = nullptr);
hipError_t launch_scale(double* p, long long n, double s_, hipStream_t s);
hipError_t launch_setd(double* p, double v, hipStream_t s);
hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t s);

hipError_t launch_compute_dt(int dim, const double* dx, double* cmax, double* dt, double cfl,
                             int limiter, double* hist, int* step, int* err, hipStream_t s,
                             double* clock = nullptr);

hipError_t launch_faces(const LaunchCfg& lc, const KParams& kp, const BoxDev& box, int dir,
                        const double* U, double* flux, double* fluxnc, double* cmc, hipStream_t s);

hipError_t launch_set2(int* p, int a, int b, hipStream_t s);

int tile_dims(int dim, int* tx, int* ty, int* tz);
int scheme_ghosts(int scheme);

}  // namespace cmp
