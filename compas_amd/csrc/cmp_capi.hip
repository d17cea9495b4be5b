// cmp_capi.hip - host runtime and C ABI of the five-equation RHS library.
//
// A level plan (cmp_level) owns only device tables: the box table, the tile work-list (tiles of
// 16x16 cells in 2D, 8x8x4 in 3D, clipped to uneven boxes), the ghost-count prefix sums and the
// block->box owner map the ghost fill uses. State buffers are the caller's, as in the reference,
// where every kernel receives non-owning Array4 views of caller-owned MultiFabs.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "cmp_kernels.hpp"
#include "cmp_level.hpp"

namespace cmp {

template <int DIM>
hipError_t launch_stage_dim(const LaunchCfg&, const KParams&, int, int, const TileDev*, int,
                            const BoxDev*, const BoxDev*, int, const double*, const double*,
                            double*, const double*, unsigned long long*, int*, hipStream_t,
                            const double*, const FillDev*);
template <int DIM>
hipError_t launch_faces_dim(const LaunchCfg&, const KParams&, const BoxDev&, int, const double*,
                            double*, double*, double*, hipStream_t);
template <int DIM>
hipError_t launch_fill_dim(int, const FillDev&, long long, const int*, int, double*, int*, hipStream_t);
template <int DIM>
hipError_t launch_dt_dim(const double*, double*, double*, double, int, double*, int*, int*,
                         hipStream_t, double*);
template <int DIM>
hipError_t launch_pf_extract_dim(const PfGeom&, long long, const double*, double*, double*, int, hipStream_t);
template <int DIM>
hipError_t launch_pf_scatter_dim(const PfGeom&, const long long*, const double*, double*, int, hipStream_t);
template <int DIM>
hipError_t launch_pf_normals_dim(const PfGeom&, long long, const double*, double*, hipStream_t);
template <int DIM>
hipError_t launch_pf_jfaces_dim(const PfGeom&, const long long*, const double*, const double*,
                                const double*, double*, double*, double, double, double, hipStream_t);
template <int DIM>
hipError_t launch_pf_update_dim(const PfGeom&, long long, double*, const double*, double, hipStream_t);
template <int DIM>
hipError_t launch_pf_gradq_dim(const PfGeom&, const long long*, const double*, const double*, double, double*,
                               unsigned long long*, int*, hipStream_t);
template <int DIM>
hipError_t launch_pf_maxvel_dim(const PfGeom&, long long, const double*, unsigned long long*, int,
                                hipStream_t);

hipError_t launch_stage(const LaunchCfg& lc, const KParams& kp, int mode, int stage,
                        const TileDev* tiles, int ntiles, const BoxDev* boxes,
                        const BoxDev* single_box, int single_ntiles, const double* Us,
                        const double* Un, double* Uout, const double* dt,
                        unsigned long long* cmax_bits, int* err, hipStream_t s, const double* gradq,
                        const FillDev* fuse)
{
    if (lc.dim == 2)
        return launch_stage_dim<2>(lc, kp, mode, stage, tiles, ntiles, boxes, single_box,
                                   single_ntiles, Us, Un, Uout, dt, cmax_bits, err, s, gradq, fuse);
    return launch_stage_dim<3>(lc, kp, mode, stage, tiles, ntiles, boxes, single_box, single_ntiles,
                               Us, Un, Uout, dt, cmax_bits, err, s, gradq, fuse);
}

template <int DIM>
hipError_t launch_fill2_dim(int, const FillDev&, const FillDev*, const FillDev&, int, double*, const double*,
                            const double*, long long, int, int*, hipStream_t);
template <int DIM>
hipError_t launch_avgdown_plan_dim(const BoxDev*, const long long*, int, long long, const FillDev&, int, AvgCell*,
                                   hipStream_t);
template <int DIM>
hipError_t launch_avgdown_apply_dim(const BoxDev*, const BoxDev*, const AvgCell*, long long, const double*, double*,
                                    int, int, double, int*, hipStream_t);
template <int DIM>
hipError_t launch_fill2_plan_dim(const FillDev&, const FillDev&, int, void*, hipStream_t);
template <int DIM>
hipError_t launch_fill2_apply_dim(int, const FillDev&, const FillDev&, long long, int, const void*, double*,
                                  const double*, int*, hipStream_t);
template <int DIM>
hipError_t launch_avgdown_dim(const BoxDev*, int, long long, const FillDev&, const double*, double*, int, int, double,
                              int*, hipStream_t);

template <int DIM>
hipError_t launch_tag_dim(const LaunchCfg&, const KParams&, const BoxDev*, int, long long, const long long*,
                          const double*, const TagCriteria&, unsigned char*, int*, hipStream_t);

hipError_t launch_tag(const LaunchCfg& lc, const KParams& kp, const BoxDev* boxes, int nbox, long long maxcells,
                      const long long* toff, const double* U, const TagCriteria& tc, unsigned char* tags, int* err,
                      hipStream_t s)
{
    if (lc.dim == 2) return launch_tag_dim<2>(lc, kp, boxes, nbox, maxcells, toff, U, tc, tags, err, s);
    return launch_tag_dim<3>(lc, kp, boxes, nbox, maxcells, toff, U, tc, tags, err, s);
}

template <int DIM>
hipError_t launch_stage_batch_dim(const LaunchCfg&, const KParams*, int, const TileDev*, int, const BoxDev*,
                                  const double*, const double*, double*, const double*, unsigned long long*, int*,
                                  const FillDev*, hipStream_t);

hipError_t launch_stage_batch(const LaunchCfg& lc, const KParams* kps, int stage, const TileDev* tiles, int ntiles,
                              const BoxDev* boxes, const double* Us, const double* Un, double* Uout,
                              const double* dts, unsigned long long* cmax_bits, int* errs, const FillDev* fds,
                              hipStream_t s)
{
    if (lc.dim == 2)
        return launch_stage_batch_dim<2>(lc, kps, stage, tiles, ntiles, boxes, Us, Un, Uout, dts, cmax_bits, errs,
                                         fds, s);
    return launch_stage_batch_dim<3>(lc, kps, stage, tiles, ntiles, boxes, Us, Un, Uout, dts, cmax_bits, errs, fds,
                                     s);
}

hipError_t launch_fill(int dim, int N, const FillDev& fd, long long nghost, double* U, int* err,
                       hipStream_t s, const int* list, int nlist)
{
    return dim == 2 ? launch_fill_dim<2>(N, fd, nghost, list, nlist, U, err, s)
                    : launch_fill_dim<3>(N, fd, nghost, list, nlist, U, err, s);
}

#define CMP_BY_DIM(dim, FN, ...) ((dim) == 2 ? FN<2>(__VA_ARGS__) : FN<3>(__VA_ARGS__))

hipError_t launch_pf_extract(int dim, const PfGeom& g, long long maxcells, const double* U, double* A, double* A2,
                             int comp, hipStream_t s)
{
    return CMP_BY_DIM(dim, launch_pf_extract_dim, g, maxcells, U, A, A2, comp, s);
}

hipError_t launch_pf_scatter(int dim, const PfGeom& g, const long long* maxfaces, const double* src, double* dst,
                             int comp, hipStream_t s)
{
    return CMP_BY_DIM(dim, launch_pf_scatter_dim, g, maxfaces, src, dst, comp, s);
}

hipError_t launch_pf_normals(int dim, const PfGeom& g, long long maxcells, const double* Psi,
                             double* Nrm, hipStream_t s)
{
    return CMP_BY_DIM(dim, launch_pf_normals_dim, g, maxcells, Psi, Nrm, s);
}

hipError_t launch_pf_jfaces(int dim, const PfGeom& g, const long long* maxfaces, const double* A,
                            const double* Psi, const double* Nrm, double* J, double* H, double mob_face,
                            double eta, double dt_m, hipStream_t s)
{
    return CMP_BY_DIM(dim, launch_pf_jfaces_dim, g, maxfaces, A, Psi, Nrm, J, H, mob_face, eta, dt_m, s);
}

hipError_t launch_pf_update(int dim, const PfGeom& g, long long maxcells, double* A, const double* J,
                            double dt_m, hipStream_t s)
{
    return CMP_BY_DIM(dim, launch_pf_update_dim, g, maxcells, A, J, dt_m, s);
}

hipError_t launch_pf_gradq(int dim, const PfGeom& g, const long long* maxfaces, const double* S,
                           const double* H, double hscale, double* gq, unsigned long long* gmax_bits,
                           int* boxflag, hipStream_t s)
{
    return CMP_BY_DIM(dim, launch_pf_gradq_dim, g, maxfaces, S, H, hscale, gq, gmax_bits, boxflag, s);
}

hipError_t launch_pf_maxvel(int dim, const PfGeom& g, long long maxcells, const double* U,
                            unsigned long long* bits, int nmass, hipStream_t s)
{
    return CMP_BY_DIM(dim, launch_pf_maxvel_dim, g, maxcells, U, bits, nmass, s);
}

template <int DIM>
hipError_t launch_pf_cac_mech_dim(const PfGeom&, long long, const double*, const double*, double*, double*,
                                  double, double, double, double*, hipStream_t);
template <int DIM>
hipError_t launch_pf_cac_update_dim(const PfGeom&, long long, double*, const double*, const double*,
                                    const double*, double, hipStream_t);
template <int DIM>
hipError_t launch_pf_divjhat_dim(const PfGeom&, long long, const double*, const double*, double*, double,
                                 hipStream_t);
template <int DIM>
hipError_t launch_pf_bcoef_dim(const PfGeom&, const long long*, const double*, double*, hipStream_t);
template <int DIM>
hipError_t launch_cell_sum_dim(const PfGeom&, long long, const double*, double*, hipStream_t);
template <int DIM>
hipError_t launch_cg_idiag_dim(const PfGeom&, long long, const double*, double*, hipStream_t);
template <int DIM>
hipError_t launch_cg_init_dim(const PfGeom&, long long, const double*, const double*, double, const double*,
                              double*, double*, double*, double*, double*, hipStream_t);
template <int DIM>
hipError_t launch_cg_ap_dim(const PfGeom&, long long, const double*, const double*, const double*, double*,
                            double*, hipStream_t);
template <int DIM>
hipError_t launch_cg_xr_dim(const PfGeom&, long long, const double*, const double*, double*, double*, double*,
                            const double*, const double*, double*, hipStream_t);
template <int DIM>
hipError_t launch_cg_p_dim(const PfGeom&, long long, const double*, const double*, double*, double, const int*, int,
                           hipStream_t);
template <int DIM>
hipError_t launch_pf_gradx_dim(const PfGeom&, const long long*, const double*, double*, unsigned long long*,
                               int*, hipStream_t);

hipError_t launch_pf_cac_mech(int dim, const PfGeom& g, long long maxcells, const double* A,
                              const double* Nrm, double* Mk, double* Wk, double mob, double lambda,
                              double eta, double* part, hipStream_t s)
{
    return CMP_BY_DIM(dim, launch_pf_cac_mech_dim, g, maxcells, A, Nrm, Mk, Wk, mob, lambda, eta, part, s);
}

hipError_t launch_pf_cac_update(int dim, const PfGeom& g, long long maxcells, double* A, const double* Mk,
                                const double* Wk, const double* sums, double dt_m, hipStream_t s)
{
    return CMP_BY_DIM(dim, launch_pf_cac_update_dim, g, maxcells, A, Mk, Wk, sums, dt_m, s);
}

hipError_t launch_pf_divjhat(int dim, const PfGeom& g, long long maxcells, const double* A, const double* A0,
                             double* DJ, double dt, hipStream_t s)
{
    return CMP_BY_DIM(dim, launch_pf_divjhat_dim, g, maxcells, A, A0, DJ, dt, s);
}

hipError_t launch_pf_bcoef(int dim, const PfGeom& g, const long long* maxfaces, const double* S, double* Bf,
                           hipStream_t s)
{
    return CMP_BY_DIM(dim, launch_pf_bcoef_dim, g, maxfaces, S, Bf, s);
}

hipError_t launch_cell_sum(int dim, const PfGeom& g, long long maxcells, const double* F, double* part,
                           hipStream_t s)
{
    return CMP_BY_DIM(dim, launch_cell_sum_dim, g, maxcells, F, part, s);
}

hipError_t launch_cg_idiag(int dim, const PfGeom& g, long long maxcells, const double* Bf, double* ID,
                           hipStream_t s)
{
    return CMP_BY_DIM(dim, launch_cg_idiag_dim, g, maxcells, Bf, ID, s);
}

hipError_t launch_cg_init(int dim, const PfGeom& g, long long maxcells, const double* rhs,
                          const double* mean_sum, double ncell, const double* ID, double* X, double* R,
                          double* Z, double* P, double* part, hipStream_t s)
{
    return CMP_BY_DIM(dim, launch_cg_init_dim, g, maxcells, rhs, mean_sum, ncell, ID, X, R, Z, P, part, s);
}

hipError_t launch_cg_ap(int dim, const PfGeom& g, long long maxcells, const double* sc, const double* Bf,
                        const double* P, double* AP, double* part, hipStream_t s)
{
    return CMP_BY_DIM(dim, launch_cg_ap_dim, g, maxcells, sc, Bf, P, AP, part, s);
}

hipError_t launch_cg_xr(int dim, const PfGeom& g, long long maxcells, const double* sc, const double* ID,
                        double* X, double* R, double* Z, const double* P, const double* AP, double* part,
                        hipStream_t s)
{
    return CMP_BY_DIM(dim, launch_cg_xr_dim, g, maxcells, sc, ID, X, R, Z, P, AP, part, s);
}

hipError_t launch_cg_p(int dim, const PfGeom& g, long long maxcells, const double* sc, const double* Z,
                       double* P, hipStream_t s, double inv_n, const int* mask, int first)
{
    return CMP_BY_DIM(dim, launch_cg_p_dim, g, maxcells, sc, Z, P, inv_n, mask, first, s);
}

hipError_t launch_pf_gradx(int dim, const PfGeom& g, const long long* maxfaces, const double* X, double* gq,
                           unsigned long long* gmax_bits, int* boxflag, hipStream_t s)
{
    return CMP_BY_DIM(dim, launch_pf_gradx_dim, g, maxfaces, X, gq, gmax_bits, boxflag, s);
}

hipError_t launch_compute_dt(int dim, const double* dx, double* cmax, double* dt, double cfl,
                             int limiter, double* hist, int* step, int* err, hipStream_t s, double* clock)
{
    return dim == 2 ? launch_dt_dim<2>(dx, cmax, dt, cfl, limiter, hist, step, err, s, clock)
                    : launch_dt_dim<3>(dx, cmax, dt, cfl, limiter, hist, step, err, s, clock);
}

hipError_t launch_faces(const LaunchCfg& lc, const KParams& kp, const BoxDev& box, int dir,
                        const double* U, double* flux, double* fluxnc, double* cmc, hipStream_t s)
{
    return lc.dim == 2 ? launch_faces_dim<2>(lc, kp, box, dir, U, flux, fluxnc, cmc, s)
                       : launch_faces_dim<3>(lc, kp, box, dir, U, flux, fluxnc, cmc, s);
}

template <int DIM>
void tile_dims_dim(int* tx, int* ty, int* tz);

int tile_dims(int dim, int* tx, int* ty, int* tz)
{
    if (dim == 2) tile_dims_dim<2>(tx, ty, tz);
    else          tile_dims_dim<3>(tx, ty, tz);
    return 0;
}

// Ghost cells a scheme reads (FiniteVolume_Parameter.H:71-139 Num_Grow; THINC needs 2, :194-196).
int scheme_ghosts(int scheme) { return scheme >= 5 ? 3 : (scheme == 1 ? 2 : 1); }

}  // namespace cmp

using namespace cmp;


int cmp::check_hip(hipError_t e)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "compas_amd: HIP error %d (%s)\n", (int)e, hipGetErrorString(e));
        return CMP_ERR_HIP;
    }
    return CMP_OK;
}

namespace {

// FiniteVolume_Parameter IDs -> kernel scheme. THINC is built on the WENO5 / WENO5B base schemes
// (the ones the benchmark cases use); THINC on MUSCL/Godunov and the bound-preserving schemes are
// not built here.
int scheme_of(const cmp_params* P, int* scheme)
{
    if (P->id_bound != 0) return CMP_ERR_UNSUPPORTED;
    if (P->id_thinc != 0 && P->id_thinc != 1) return CMP_ERR_ARG;
    if (P->id_weno == 1 && P->id_limiter == 5) { *scheme = 5; return CMP_OK; }
    if (P->id_weno == 1 && P->id_limiter == 6) {
        if (P->id_poly != 2 && P->id_poly != 4) return CMP_ERR_ARG;
        *scheme = 6;
        return CMP_OK;
    }
    if (P->id_thinc) return CMP_ERR_UNSUPPORTED;
    if (P->id_weno == 0 && P->id_limiter >= 1 && P->id_limiter <= 4) { *scheme = 1; return CMP_OK; }
    if (P->id_weno == 0 && P->id_limiter == 0) { *scheme = 0; return CMP_OK; }
    return CMP_ERR_UNSUPPORTED;
}

int validate(const cmp_params* P, const cmp_geometry* G, int ng, LaunchCfg* lc, KParams* kp)
{
    if (!P || !G) return CMP_ERR_ARG;
    if (P->dim != 2 && P->dim != 3) return CMP_ERR_ARG;
    if (G->dim != P->dim) return CMP_ERR_ARG;
    if (P->nphase == 0) {
        if (P->nstate != 4 + P->dim) return CMP_ERR_ARG;
    } else if (P->nphase == 2 || P->nphase == 3) {
        if (P->nstate != 2 * P->nphase + P->dim + 1) return CMP_ERR_ARG;
        // the N-phase path here: Godunov, MUSCL and WENO5 with the consistent limiter, HLL/HLLC
        if (P->id_thinc || !(P->id_weno == 0 || P->id_limiter == 5)) return CMP_ERR_UNSUPPORTED;
    } else {
        return CMP_ERR_ARG;
    }
    if (P->id_riemann != 1 && P->id_riemann != 2) return CMP_ERR_UNSUPPORTED;
    if (P->source_term != 0 && P->source_term != 1) return CMP_ERR_ARG;
    int scheme = 0;
    int r = scheme_of(P, &scheme);
    if (r) return r;
    if (ng < scheme_ghosts(scheme)) return CMP_ERR_ARG;
    for (int d = 0; d < P->dim; ++d)
        if (!(G->dx[d] > 0.0)) return CMP_ERR_ARG;
    const int nph = P->nphase ? P->nphase : 2;
    for (int k = 0; k < nph; ++k)
        if (!(P->eos_A[k] > 0.0)) return CMP_ERR_ARG;
    if (P->id_thinc && !(P->thinc_beta > 0.0)) return CMP_ERR_ARG;
    lc->dim = P->dim;
    lc->scheme = scheme;
    lc->limiter = P->id_limiter;
    lc->riemann = P->id_riemann;
    lc->thinc = P->id_thinc;
    lc->nphase = P->nphase;
    std::memset(kp, 0, sizeof(*kp));
    kp->weno5b_weight = (P->id_poly == 2) ? 1.0 / 6.0 : 1.0 / 12.0;
    for (int k = 0; k < nph; ++k) {
        kp->A[k] = P->eos_A[k];
        kp->B[k] = P->eos_B[k];
        kp->C[k] = P->eos_C[k];
        kp->D[k] = P->eos_D[k];
        if (P->eos_C[k] > 0.0) kp->nasg = 1;
    }
    kp->source_term = P->source_term;
    kp->thinc = P->id_thinc;
    kp->thinc_use_normal = P->thinc_use_normal;
    kp->thinc_eps = P->thinc_eps;
    kp->thinc_beta = P->thinc_beta;
    kp->thinc_thr = P->thinc_normal_threshold;
    for (int d = 0; d < 3; ++d) {
        kp->dx[d] = (d < P->dim) ? G->dx[d] : 1.0;
        kp->inv_dx[d] = 1.0 / kp->dx[d];
    }
    return CMP_OK;
}

BoxDev make_boxdev(int dim, const cmp_box& b, int ng)
{
    BoxDev d{};
    long long ext[3];
    for (int m = 0; m < 3; ++m) {
        d.lo[m] = b.lo[m];
        d.hi[m] = (m < dim) ? b.hi[m] : b.lo[m];
        d.flo[m] = (m < dim) ? b.lo[m] - ng : b.lo[m];
        ext[m] = (m < dim) ? (b.hi[m] - b.lo[m] + 1 + 2 * ng) : 1;
        d.olo[m] = d.lo[m];
    }
    d.off = b.offset;
    d.sj = ext[0];
    d.sk = ext[0] * ext[1];
    d.sn = ext[0] * ext[1] * ext[2];
    // default output layout: same as the state (level dUdt buffers)
    d.ooff = b.offset;
    for (int m = 0; m < 3; ++m) d.olo[m] = d.flo[m];
    d.osj = d.sj;
    d.osk = d.sk;
    d.osn = d.sn;
    return d;
}

// Face-buffer offsets of cmp_face_layout: box-major, nodal in each direction.
long long face_offsets(int dim, int nbox, const cmp_box* boxes, int ncomp, long long* off)
{
    long long o = 0;
    for (int b = 0; b < nbox; ++b)
        for (int d = 0; d < dim; ++d) {
            long long n = ncomp;
            for (int e = 0; e < dim; ++e) n *= (boxes[b].hi[e] - boxes[b].lo[e] + 1 + (e == d ? 1 : 0));
            if (off) off[b * dim + d] = o;
            o += n;
        }
    return o;
}

long long gcd_ll(long long a, long long b)
{
    a = a < 0 ? -a : a;
    b = b < 0 ? -b : b;
    while (b) { long long t = a % b; a = b; b = t; }
    return a;
}

}  // namespace

extern "C" {

const char* cmp_version(void) { return "compas_amd 0.1 (gfx950, FP64)"; }

int cmp_abi_sizes(int* params_size, int* geometry_size, int* box_size)
{
    if (params_size) *params_size = (int)sizeof(cmp_params);
    if (geometry_size) *geometry_size = (int)sizeof(cmp_geometry);
    if (box_size) *box_size = (int)sizeof(cmp_box);
    return CMP_OK;
}

int cmp_abi_pf_size(int* pf_params_size)
{
    if (!pf_params_size) return CMP_ERR_ARG;
    *pf_params_size = (int)sizeof(cmp_pf_params);
    return CMP_OK;
}

int cmp_face_layout(int dim, int nbox, const cmp_box* boxes, int ncomp, long long* offsets,
                    long long* total)
{
    if (!boxes || nbox < 0 || dim < 2 || dim > 3 || ncomp < 1) return CMP_ERR_ARG;
    for (int b = 0; b < nbox; ++b)
        for (int d = 0; d < dim; ++d)
            if (boxes[b].hi[d] < boxes[b].lo[d]) return CMP_ERR_ARG;
    const long long t = face_offsets(dim, nbox, boxes, ncomp, offsets);
    if (total) *total = t;
    return CMP_OK;
}

long long cmp_fab_size(int dim, const cmp_box* box, int ng, int ncomp)
{
    if (!box || dim < 1 || dim > 3) return -1;
    long long n = ncomp;
    for (int d = 0; d < dim; ++d) n *= (box->hi[d] - box->lo[d] + 1 + 2 * ng);
    return n;
}

int cmp_level_layout(int dim, int nbox, cmp_box* boxes, int ng, int ncomp, long long* total)
{
    if (!boxes || nbox < 0 || dim < 2 || dim > 3 || ng < 0 || ncomp < 1) return CMP_ERR_ARG;
    long long off = 0;
    for (int b = 0; b < nbox; ++b) {
        for (int d = 0; d < dim; ++d)
            if (boxes[b].hi[d] < boxes[b].lo[d]) return CMP_ERR_ARG;
        boxes[b].offset = off;
        off += cmp_fab_size(dim, &boxes[b], ng, ncomp);
    }
    if (total) *total = off;
    return CMP_OK;
}

int cmp_rhs_box(const cmp_params* params, const cmp_geometry* geom, const cmp_box* box, int ng,
                const double* U, double* dUdt, double* cmax,
                double* flux_x, double* flux_y, double* flux_z,
                double* fluxnc_x, double* fluxnc_y, double* fluxnc_z, void* stream)
{
    LaunchCfg lc;
    KParams kp;
    int r = validate(params, geom, ng, &lc, &kp);
    if (r) return r;
    if (!box || !U || !dUdt) return CMP_ERR_ARG;
    const int dim = params->dim;
    for (int d = 0; d < dim; ++d)
        if (box->hi[d] < box->lo[d]) return CMP_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    cmp_box b0 = *box;
    b0.offset = 0;
    BoxDev bd = make_boxdev(dim, b0, ng);
    // output: valid box only, comp-major
    bd.ooff = 0;
    long long nv[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) nv[d] = box->hi[d] - box->lo[d] + 1;
    for (int m = 0; m < 3; ++m) bd.olo[m] = bd.lo[m];
    bd.osj = nv[0];
    bd.osk = nv[0] * nv[1];
    bd.osn = nv[0] * nv[1] * nv[2];
    int tx, ty, tz;
    tile_dims(dim, &tx, &ty, &tz);
    const int ntx = (int)((nv[0] + tx - 1) / tx), nty = (int)((nv[1] + ty - 1) / ty);
    const int ntz = (dim == 3) ? (int)((nv[2] + tz - 1) / tz) : 1;
    CMP_TRY(launch_stage(lc, kp, MODE_RHS, 0, nullptr, 0, nullptr, &bd, ntx * nty * ntz, U, nullptr,
                         dUdt, nullptr, (unsigned long long*)cmax, nullptr, s));
    double* fx[3] = {flux_x, flux_y, flux_z};
    double* fn[3] = {fluxnc_x, fluxnc_y, fluxnc_z};
    for (int d = 0; d < dim; ++d) {
        if (!fx[d] && !fn[d]) continue;
        if (lc.nphase) return CMP_ERR_UNSUPPORTED;   // face arrays of the N-phase model are not built
        CMP_TRY(launch_faces(lc, kp, bd, d, U, fx[d], fn[d], nullptr, s));
    }
    return CMP_OK;
}

int cmp_cmax_box(const cmp_params* params, const cmp_geometry* geom, const cmp_box* box, int ng, const double* U,
                 double* cmax_cells, void* stream)
{
    LaunchCfg lc;
    KParams kp;
    int r = validate(params, geom, ng, &lc, &kp);
    if (r) return r;
    if (!box || !U || !cmax_cells) return CMP_ERR_ARG;
    if (lc.nphase) return CMP_ERR_UNSUPPORTED;
    const int dim = params->dim;
    for (int d = 0; d < dim; ++d)
        if (box->hi[d] < box->lo[d]) return CMP_ERR_ARG;
    cmp_box b0 = *box;
    b0.offset = 0;
    const BoxDev bd = make_boxdev(dim, b0, ng);
    // direction 0 writes every cell, the later directions max-combine, in stream order
    for (int d = 0; d < dim; ++d)
        CMP_TRY(launch_faces(lc, kp, bd, d, U, nullptr, nullptr, cmax_cells, (hipStream_t)stream));
    return CMP_OK;
}

int cmp_pf_rhs_box(const cmp_params* params, const cmp_pf_params* pf, const cmp_geometry* geom, const cmp_box* box,
                   int ng, const double* U, const double* gradq, double* dUdt, void* stream)
{
    LaunchCfg lc;
    KParams kp;
    int r = validate(params, geom, ng, &lc, &kp);
    if (r) return r;
    if (!pf || !box || !U || !gradq || !dUdt) return CMP_ERR_ARG;
    if (pf->id_flux != 0) return CMP_ERR_UNSUPPORTED;
    if (pf->flux_phase2flux != 0 && pf->flux_phase2flux != 1) return CMP_ERR_ARG;
    const int dim = params->dim;
    for (int d = 0; d < dim; ++d)
        if (box->hi[d] < box->lo[d]) return CMP_ERR_ARG;
    // the Phase-Field form of the stage kernel (as cmp_level_pf_setup configures it)
    lc.thinc = 0;
    lc.riemann = (pf->flux_phase2flux == 1) ? RIEMANN_PF_MIX : RIEMANN_PF_PHASIC;
    kp.thinc = 0;
    for (int k = 0; k < CMP_MAX_PHASES; ++k) kp.pf_pmin[k] = pf->p_min[k];
    cmp_box b0 = *box;
    b0.offset = 0;
    BoxDev bd = make_boxdev(dim, b0, ng);
    long long foff[3] = {0, 0, 0};
    face_offsets(dim, 1, &b0, params->nphase ? params->nphase : 1, foff);
    for (int d = 0; d < dim; ++d) bd.goff[d] = foff[d];
    // output: valid box only, comp-major
    bd.ooff = 0;
    long long nv[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) nv[d] = box->hi[d] - box->lo[d] + 1;
    for (int m = 0; m < 3; ++m) bd.olo[m] = bd.lo[m];
    bd.osj = nv[0];
    bd.osk = nv[0] * nv[1];
    bd.osn = nv[0] * nv[1] * nv[2];
    int tx, ty, tz;
    tile_dims(dim, &tx, &ty, &tz);
    const int ntx = (int)((nv[0] + tx - 1) / tx), nty = (int)((nv[1] + ty - 1) / ty);
    const int ntz = (dim == 3) ? (int)((nv[2] + tz - 1) / tz) : 1;
    CMP_TRY(launch_stage(lc, kp, MODE_RHS, 0, nullptr, 0, nullptr, &bd, ntx * nty * ntz, U, nullptr, dUdt, nullptr,
                         nullptr, nullptr, (hipStream_t)stream, gradq));
    return CMP_OK;
}

}  // extern "C"

// The halo sources of tile [tlo, thi] of box b on its box's faces (the ghost fill's rules, resolved
// per tile side): the ng cells beyond the face, across the tile's transverse extent, lie in one box
// (after the periodic wrap), or beyond a non-periodic domain face whose BC clamps (codes 0, 2) or
// mirrors (3, 4: normal momentum negated; 5: all momenta) within the tile's own box; anything else
// is looked up cell by cell. Sides that are not on the box's face keep their entry.
void cmp::tile_halo(int dim, const cmp_geometry* geom, const cmp_box* boxes, int b, const int* tlo, const int* thi,
                    int ng, const std::vector<int>& owner, const int* nb, long long bs, int (*halo)[2])
{
    auto owner_of = [&](const int* c) -> int {
        long long q = 0;
        for (int d = dim - 1; d >= 0; --d) q = q * nb[d] + (c[d] - geom->domain_lo[d]) / bs;
        return owner[q];
    };
    const cmp_box& B = boxes[b];
    for (int d = 0; d < dim; ++d)
        for (int side = 0; side < 2; ++side) {
            if ((side == 0 && tlo[d] != B.lo[d]) || (side == 1 && thi[d] != B.hi[d])) continue;
            const int dlo = geom->domain_lo[d], dhi = geom->domain_hi[d];
            int rlo[3], rhi[3];
            for (int e = 0; e < 3; ++e) {
                rlo[e] = (e < dim) ? tlo[e] : 0;
                rhi[e] = (e < dim) ? thi[e] : 0;
            }
            if (side == 0) { rlo[d] = B.lo[d] - ng; rhi[d] = B.lo[d] - 1; }
            else { rlo[d] = B.hi[d] + 1; rhi[d] = B.hi[d] + ng; }
            const bool below = rhi[d] < dlo, above = rlo[d] > dhi;
            const bool inside = rlo[d] >= dlo && rhi[d] <= dhi;
            int code = HALO_LOOKUP;
            if (inside || ((below || above) && geom->is_periodic[d])) {
                if (below) { rlo[d] += dhi - dlo + 1; rhi[d] += dhi - dlo + 1; }
                if (above) { rlo[d] -= dhi - dlo + 1; rhi[d] -= dhi - dlo + 1; }
                const int o = owner_of(rlo);
                bool ok = o >= 0;
                for (int e = 0; e < dim && ok; ++e)
                    ok = rlo[e] >= boxes[o].lo[e] && rhi[e] <= boxes[o].hi[e];
                if (ok) code = o;
            } else if (below || above) {
                const int bc = below ? geom->lo_bc[d] : geom->hi_bc[d];
                const bool fits = B.hi[d] - B.lo[d] + 1 >= ng;
                if (bc == 0 || bc == 2) code = HALO_CLAMP;
                else if ((bc == 3 || bc == 4) && fits) code = HALO_MIRROR;
                else if (bc == 5 && fits) code = HALO_MIRROR_ALL;
            }
            halo[d][side] = code;
        }
}

extern "C" {

int cmp_level_create(const cmp_params* params, const cmp_geometry* geom, int nbox,
                     const cmp_box* boxes, int ng, cmp_level** out)
{
    if (!out) return CMP_ERR_ARG;
    *out = nullptr;
    LaunchCfg lc;
    KParams kp;
    int r = validate(params, geom, ng, &lc, &kp);
    if (r) return r;
    if (nbox <= 0 || !boxes) return CMP_ERR_ARG;
    const int dim = params->dim;
    // boxes inside the domain; offsets consistent with cmp_level_layout
    long long expect = 0;
    for (int b = 0; b < nbox; ++b) {
        for (int d = 0; d < dim; ++d) {
            if (boxes[b].lo[d] < geom->domain_lo[d] || boxes[b].hi[d] > geom->domain_hi[d] ||
                boxes[b].hi[d] < boxes[b].lo[d])
                return CMP_ERR_LAYOUT;
            if (boxes[b].hi[d] - boxes[b].lo[d] + 1 < ng && !geom->is_periodic[d]) {
                // a reflecting side mirrors up to ng cells of the box itself
            }
        }
        if (boxes[b].offset != expect) return CMP_ERR_LAYOUT;
        // the kernels index the cells of one fab with 32-bit integers
        if (cmp_fab_size(dim, &boxes[b], ng, 1) >= (1LL << 31)) return CMP_ERR_UNSUPPORTED;
        expect += cmp_fab_size(dim, &boxes[b], ng, params->nstate);
    }
    for (int d = 0; d < dim; ++d)
        if (geom->domain_hi[d] - geom->domain_lo[d] + 1 < ng && geom->is_periodic[d]) return CMP_ERR_LAYOUT;

    cmp_level* L = new cmp_level();
    static std::atomic<unsigned long long> next_uid{1};
    L->uid = next_uid++;
    L->P = *params;
    L->G = *geom;
    L->ng = ng;
    L->nbox = nbox;
    L->boxes.assign(boxes, boxes + nbox);
    L->kp = kp;
    L->lc = lc;

    // box table, with each box's place in a level face buffer (Phase-Field Grad_Q)
    std::vector<BoxDev> hb(nbox);
    std::vector<long long> foff((size_t)nbox * dim);
    face_offsets(dim, nbox, boxes, params->nphase ? params->nphase : 1, foff.data());
    for (int b = 0; b < nbox; ++b) {
        hb[b] = make_boxdev(dim, boxes[b], ng);
        for (int d = 0; d < dim; ++d) hb[b].goff[d] = foff[(size_t)b * dim + d];
    }
    // tile work-list, box by box
    int tx, ty, tz;
    tile_dims(dim, &tx, &ty, &tz);
    std::vector<TileDev> ht;
    for (int b = 0; b < nbox; ++b) {
        const cmp_box& B = boxes[b];
        const int klo = (dim == 3) ? B.lo[2] : 0, khi = (dim == 3) ? B.hi[2] : 0;
        for (int k = klo; k <= khi; k += (dim == 3 ? tz : 1))
            for (int j = B.lo[1]; j <= B.hi[1]; j += ty)
                for (int i = B.lo[0]; i <= B.hi[0]; i += tx)
                    ht.push_back(TileDev{b, {i, j, k}, {{HALO_LOOKUP, HALO_LOOKUP}, {HALO_LOOKUP, HALO_LOOKUP},
                                                         {HALO_LOOKUP, HALO_LOOKUP}}});
    }
    L->ntiles = (int)ht.size();
    L->tile_start.assign(nbox + 1, 0);
    for (const TileDev& t : ht) ++L->tile_start[t.box + 1];
    for (int b = 0; b < nbox; ++b) L->tile_start[b + 1] += L->tile_start[b];
    // ghost-count prefix
    std::vector<long long> gp(nbox + 1, 0);
    for (int b = 0; b < nbox; ++b) {
        long long fab = 1, val = 1;
        for (int d = 0; d < dim; ++d) {
            const long long n = boxes[b].hi[d] - boxes[b].lo[d] + 1;
            fab *= n + 2 * ng;
            val *= n;
        }
        gp[b + 1] = gp[b] + (fab - val);
    }
    L->nghost = gp[nbox];
    // owner map at the coarsest block size every box edge is aligned to
    long long bs = 0;
    for (int d = 0; d < dim; ++d) {
        bs = gcd_ll(bs, geom->domain_hi[d] - geom->domain_lo[d] + 1);
        for (int b = 0; b < nbox; ++b) {
            bs = gcd_ll(bs, boxes[b].lo[d] - geom->domain_lo[d]);
            bs = gcd_ll(bs, boxes[b].hi[d] + 1 - geom->domain_lo[d]);
        }
    }
    if (bs <= 0) bs = 1;
    int nb[3] = {1, 1, 1};
    long long nblocks = 1;
    for (int d = 0; d < dim; ++d) {
        nb[d] = (int)((geom->domain_hi[d] - geom->domain_lo[d] + 1) / bs);
        nblocks *= nb[d];
    }
    std::vector<int> owner(nblocks, -1);
    for (int b = 0; b < nbox; ++b) {
        int blo[3] = {0, 0, 0}, bhi[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) {
            blo[d] = (int)((boxes[b].lo[d] - geom->domain_lo[d]) / bs);
            bhi[d] = (int)((boxes[b].hi[d] + 1 - geom->domain_lo[d]) / bs) - 1;
        }
        for (int k = blo[2]; k <= bhi[2]; ++k)
            for (int j = blo[1]; j <= bhi[1]; ++j)
                for (int i = blo[0]; i <= bhi[0]; ++i) {
                    const long long q = i + (long long)nb[0] * (j + (long long)nb[1] * k);
                    if (owner[q] >= 0) { delete L; return CMP_ERR_LAYOUT; }   // overlap
                    owner[q] = b;
                }
    }
    // halo sources of the tiles on box faces
    for (TileDev& t : ht) {
        const cmp_box& B = boxes[t.box];
        int thi[3] = {0, 0, 0};
        const int te[3] = {tx, ty, tz};
        for (int d = 0; d < dim; ++d) thi[d] = std::min(t.lo[d] + te[d] - 1, B.hi[d]);
        tile_halo(dim, geom, boxes, t.box, t.lo, thi, ng, owner, nb, bs, t.halo);
    }

    int rc = CMP_OK;
#define CMP_TRYL(x) do { rc = check_hip(x); if (rc) { cmp_level_destroy(L); return rc; } } while (0)
    CMP_TRYL(hipMalloc(&L->d_boxes, sizeof(BoxDev) * nbox));
    CMP_TRYL(hipMemcpy(L->d_boxes, hb.data(), sizeof(BoxDev) * nbox, hipMemcpyHostToDevice));
    CMP_TRYL(hipMalloc(&L->d_tiles, sizeof(TileDev) * std::max<size_t>(1, ht.size())));
    CMP_TRYL(hipMemcpy(L->d_tiles, ht.data(), sizeof(TileDev) * ht.size(), hipMemcpyHostToDevice));
    CMP_TRYL(hipMalloc(&L->d_gprefix, sizeof(long long) * (nbox + 1)));
    CMP_TRYL(hipMemcpy(L->d_gprefix, gp.data(), sizeof(long long) * (nbox + 1), hipMemcpyHostToDevice));
    CMP_TRYL(hipMalloc(&L->d_owner, sizeof(int) * nblocks));
    CMP_TRYL(hipMemcpy(L->d_owner, owner.data(), sizeof(int) * nblocks, hipMemcpyHostToDevice));
    // d_err[0] error word, d_err[1..2] step counter and dt-history capacity of the current call
    CMP_TRYL(hipMalloc(&L->d_err, sizeof(int) * 4));
    CMP_TRYL(hipMemset(L->d_err, 0, sizeof(int) * 4));
    L->d_step = L->d_err + 1;
#undef CMP_TRYL

    FillDev& fd = L->fd;
    fd.boxes = L->d_boxes;
    fd.gprefix = L->d_gprefix;
    fd.owner = L->d_owner;
    L->h_tiles = std::move(ht);
    L->h_boxes = hb;
    L->h_owner = std::move(owner);
    L->bs = bs;
    for (int d = 0; d < 3; ++d) L->nb[d] = nb[d];
    fd.nbox = nbox;
    fd.ng = ng;
    fd.bs = (int)bs;
    fd.maxghost = 0;
    for (int b = 0; b < nbox; ++b) fd.maxghost = std::max(fd.maxghost, gp[b + 1] - gp[b]);
    for (int d = 0; d < 3; ++d) {
        fd.nb[d] = nb[d];
        fd.dlo[d] = geom->domain_lo[d];
        fd.dhi[d] = (d < dim) ? geom->domain_hi[d] : geom->domain_lo[d];
        fd.periodic[d] = geom->is_periodic[d];
        fd.lo_bc[d] = geom->lo_bc[d];
        fd.hi_bc[d] = geom->hi_bc[d];
    }
    // the level's clock: time 0, no stop_time
    const double clk[2] = {0.0, -1.0};
    int rc2 = check_hip(hipMalloc(&L->d_clock, sizeof(double) * 2));
    if (!rc2) rc2 = check_hip(hipMemcpy(L->d_clock, clk, sizeof(double) * 2, hipMemcpyHostToDevice));
    if (rc2) { cmp_level_destroy(L); return rc2; }
    rc2 = check_hip(hipMalloc(&L->d_fd, sizeof(FillDev)));
    if (!rc2) rc2 = check_hip(hipMemcpy(L->d_fd, &fd, sizeof(FillDev), hipMemcpyHostToDevice));
    if (rc2) { cmp_level_destroy(L); return rc2; }
    *out = L;
    return CMP_OK;
}

// Releases what cmp_level_pf_setup allocated (errors of the frees are not actionable here).
static void pf_free(cmp_level* L)
{
    if (L->cg_exec) (void)hipGraphExecDestroy(L->cg_exec);
    L->cg_exec = nullptr;
    void* bufs[] = {L->d_abox, L->d_nbox, L->d_gprefix1, L->d_A0, L->d_A, L->d_P, L->d_N,
                    L->d_J, L->d_H, L->d_gq, L->d_gq1, L->d_bits, L->d_pfdt, L->d_Mk, L->d_Wk, L->d_DJ,
                    L->d_X, L->d_R, L->d_Pv, L->d_AP, L->d_Z, L->d_ID, L->d_Bf, L->d_part, L->d_sc,
                    L->d_boxflag,
                    L->d_pf_tiles, L->d_cmap};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    void* pins[] = {L->h_pin, L->h_boxflag, L->h_pf_tiles};
    for (void* p : pins)
        if (p) (void)hipHostFree(p);
    L->d_boxflag = L->h_boxflag = nullptr;
    L->d_cmap = nullptr;
    L->pg.cmap = nullptr;
    L->pg.ncmap = 0;
    L->d_pf_tiles = L->h_pf_tiles = nullptr;
    L->ghost_src.clear();
    L->d_abox = L->d_nbox = nullptr;
    L->d_gprefix1 = nullptr;
    L->d_A0 = L->d_A = L->d_P = L->d_N = L->d_J = L->d_H = L->d_gq = L->d_gq1 = L->d_pfdt = nullptr;
    L->d_Mk = L->d_Wk = L->d_DJ = L->d_X = L->d_R = L->d_Pv = L->d_AP = L->d_Bf = nullptr;
    L->d_Z = L->d_ID = nullptr;
    L->d_part = L->d_sc = nullptr;
    L->d_bits = nullptr;
    L->h_pin = nullptr;
    for (auto& V : L->mg) {
        void* mb[] = {V.x[0], V.x[1], V.r, V.id, V.b[0], V.b[1], V.b[2]};
        for (void* p : mb)
            if (p) (void)hipFree(p);
    }
    L->mg.clear();
    L->pf_on = false;
}

// The dense multigrid grids of a level that covers its whole domain: its own resolution, then halved
// in every active direction while every active extent is even; none (Jacobi preconditioning) when
// the level leaves part of its domain uncovered or its coarsest grid would exceed 1024 cells.
static int mg_setup(cmp_level* L)
{
    const int dim = L->P.dim;
    long long dom = 1;
    for (int d = 0; d < dim; ++d) dom *= L->G.domain_hi[d] - L->G.domain_lo[d] + 1;
    if (dom != L->ncell) return CMP_OK;
    cmp::MgGrid G{};
    for (int d = 0; d < 3; ++d) {
        G.n[d] = (d < dim) ? L->G.domain_hi[d] - L->G.domain_lo[d] + 1 : 1;
        G.periodic[d] = (d < dim) ? L->G.is_periodic[d] : 1;
        G.ih2[d] = (d < dim) ? 1.0 / (L->G.dx[d] * L->G.dx[d]) : 0.0;
    }
    std::vector<cmp::MgGrid> grids;
    for (;;) {
        G.nc = (long long)G.n[0] * G.n[1] * G.n[2];
        grids.push_back(G);
        bool even = true;
        for (int d = 0; d < dim; ++d) even = even && (G.n[d] % 2 == 0);
        if (!even) break;
        for (int d = 0; d < dim; ++d) {
            G.n[d] /= 2;
            G.ih2[d] *= 0.25;
        }
    }
    const cmp::MgGrid& B = grids.back();
    if (B.nc > 1024) return CMP_OK;
    int mx = 1;
    for (int d = 0; d < dim; ++d) mx = std::max(mx, B.n[d]);
    L->mg_bottom_sweeps = std::min(256, std::max(8, 2 * mx * mx));
    L->mg_omega = (dim == 3) ? 6.0 / 7.0 : 0.8;
    L->mg.resize(grids.size());
    for (size_t m = 0; m < grids.size(); ++m) {
        auto& V = L->mg[m];
        V.G = grids[m];
        const size_t nc = (size_t)V.G.nc;
        CMP_TRY(hipMalloc(&V.x[0], sizeof(double) * nc));
        CMP_TRY(hipMalloc(&V.x[1], sizeof(double) * nc));
        CMP_TRY(hipMalloc(&V.r, sizeof(double) * nc));
        CMP_TRY(hipMalloc(&V.id, sizeof(double) * nc));
        for (int d = 0; d < dim; ++d) {
            size_t nf = 1;
            for (int e = 0; e < dim; ++e) nf *= (size_t)(V.G.n[e] + (e == d));
            CMP_TRY(hipMalloc(&V.b[d], sizeof(double) * nf));
        }
    }
    return CMP_OK;
}

int cmp_level_destroy(cmp_level* L)
{
    if (!L) return CMP_OK;
    pf_free(L);
    if (L->gexec) (void)hipGraphExecDestroy(L->gexec);
    if (L->graph) (void)hipGraphDestroy(L->graph);
    if (L->own_stream) (void)hipStreamDestroy(L->own_stream);
    void* bufs[] = {L->d_boxes, L->d_tiles, L->d_gprefix, L->d_owner, L->d_err, L->d_fd, L->d_toff, L->d_U3,
                    L->d_clock, L->d_ltiles, L->d_llev, L->d_lctr, L->d_ftiles, L->d_f2plan, L->d_avgprefix, L->d_avgplan};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    delete L;
    return CMP_OK;
}

int cmp_level_info(const cmp_level* L, int* ntiles, int* tile_x, int* tile_y, int* tile_z)
{
    if (!L) return CMP_ERR_ARG;
    int tx, ty, tz;
    tile_dims(L->P.dim, &tx, &ty, &tz);
    if (ntiles) *ntiles = L->ntiles;
    if (tile_x) *tile_x = tx;
    if (tile_y) *tile_y = ty;
    if (tile_z) *tile_z = tz;
    return CMP_OK;
}

int cmp_level_fill_ghosts(cmp_level* L, double* U, void* stream)
{
    if (!L || !U) return CMP_ERR_ARG;
    CMP_TRY(launch_fill(L->P.dim, L->P.nstate, L->fd, L->nghost, U, L->d_err, (hipStream_t)stream));
    return CMP_OK;
}

int cmp_level_rhs(cmp_level* L, const double* U, double* dUdt, double* cmax, void* stream)
{
    if (!L || !U || !dUdt) return CMP_ERR_ARG;
    CMP_TRY(launch_stage(L->lc, L->kp, MODE_RHS, 0, L->d_tiles, L->ntiles, L->d_boxes, nullptr, 0, U,
                         nullptr, dUdt, nullptr, (unsigned long long*)cmax, L->d_err,
                         (hipStream_t)stream));
    return CMP_OK;
}

int cmp_level_rk3_stage(cmp_level* L, int stage, const double* Us, const double* Un, double* Uout,
                        const double* dt_dev, double* cmax, void* stream)
{
    if (!L || !Us || !Uout || !dt_dev || stage < 0 || stage > 2) return CMP_ERR_ARG;
    if (stage > 0 && !Un) return CMP_ERR_ARG;
    if (Us == Uout) return CMP_ERR_ARG;
    CMP_TRY(launch_stage(L->lc, L->kp, MODE_RK, stage, L->d_tiles, L->ntiles, L->d_boxes, nullptr, 0,
                         Us, Un, Uout, dt_dev, (unsigned long long*)cmax, L->d_err,
                         (hipStream_t)stream));
    return CMP_OK;
}

// The three stages read the cells beyond their boxes where the ghost fill would take them from
// (the fill fused into the stage kernel's staging), so no stage input has its ghosts written.
int cmp_level_rk3_step(cmp_level* L, double* Un, double* U1, double* U2, const double* dt_dev,
                       double* cmax, void* stream)
{
    if (!L || !Un || !U1 || !U2 || !dt_dev) return CMP_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    double* Us[3] = {Un, U1, U2};
    double* Uo[3] = {U1, U2, Un};
    for (int stage = 0; stage < 3; ++stage) {
        if (L->use_list) {
            const int r = level_list_stage(L, MODE_RK, stage, Us[stage], Un, Uo[stage], dt_dev, cmax, s);
            if (r) return r;
            continue;
        }
        CMP_TRY(launch_stage(L->lc, L->kp, MODE_RK, stage, L->d_tiles, L->ntiles, L->d_boxes, nullptr, 0,
                             Us[stage], Un, Uo[stage], dt_dev, (unsigned long long*)cmax, L->d_err, s, nullptr,
                             L->d_fd));
    }
    return CMP_OK;
}

// ForwardEuler, TVD-RK2 and RK4 (run.TimeIntegrator): per stage the filled stage state, its
// right-hand side into a level buffer, then the reference's LinComb / Saxpy sequence.
// gradq non-null: the Phase-Field right-hand side at that Grad_Q (no c_max).
static int rk_step_impl(cmp_level* L, int integrator, double* Un, double* U1, double* U2, double* U3,
                        const double* dt_dev, double* cmax, const double* gradq, void* stream)
{
    if (integrator != 1 && integrator != 2 && integrator != 4) return CMP_ERR_ARG;
    if (integrator == 4 && !U3) return CMP_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const int N = L->P.nstate;
    long long maxc = 0;
    for (const cmp_box& b : L->boxes) {
        long long v = 1;
        for (int d = 0; d < L->P.dim; ++d) v *= b.hi[d] - b.lo[d] + 1;
        maxc = std::max(maxc, v);
    }
    // U1: the stage state (U^(1) of TVD-RK2, U_temp of RK4); U2: the right-hand side; U3: RK4's U^(n+1)
    double* Lb = U2;
    auto rhs = [&](double* src) -> int {
        int r = cmp_level_fill_ghosts(L, src, stream);
        if (r) return r;
        if (gradq)
            CMP_TRY(launch_stage(L->pf_lc, L->pf_kp, MODE_RHS, 0, L->d_tiles, L->ntiles, L->d_boxes, nullptr, 0, src,
                                 nullptr, Lb, nullptr, nullptr, L->d_err, s, gradq));
        else
            CMP_TRY(launch_stage(L->lc, L->kp, MODE_RHS, 0, L->d_tiles, L->ntiles, L->d_boxes, nullptr, 0, src,
                                 nullptr, Lb, nullptr, (unsigned long long*)cmax, L->d_err, s));
        return CMP_OK;
    };
    auto lc = [&](double* out, double p, const double* X, double q, const double* Y, double r, const double* Lr) -> int {
        CMP_TRY(launch_lincomb(L->d_boxes, L->nbox, maxc, N, out, p, X, q, Y, r, Lr, dt_dev, s));
        return CMP_OK;
    };
    int r;
    if (integrator == 1) {
        if ((r = rhs(Un)) || (r = lc(Un, 1.0, Un, 0.0, nullptr, 1.0, Lb))) return r;
        return CMP_OK;
    }
    if (integrator == 2) {
        if ((r = rhs(Un)) || (r = lc(U1, 1.0, Un, 0.0, nullptr, 1.0, Lb))) return r;
        if ((r = rhs(U1)) || (r = lc(Un, 0.5, U1, 0.5, Un, 0.5, Lb))) return r;
        return CMP_OK;
    }
    double* Ut = U1;
    double* Un1 = U3;
    if ((r = rhs(Un)) || (r = lc(Ut, 1.0, Un, 0.0, nullptr, 0.5, Lb)) || (r = lc(Un1, 1.0, Un, 0.0, nullptr, 1.0 / 6.0, Lb)))
        return r;
    if ((r = rhs(Ut)) || (r = lc(Ut, 1.0, Un, 0.0, nullptr, 0.5, Lb)) || (r = lc(Un1, 1.0, Un1, 0.0, nullptr, 1.0 / 3.0, Lb)))
        return r;
    if ((r = rhs(Ut)) || (r = lc(Ut, 1.0, Un, 0.0, nullptr, 1.0, Lb)) || (r = lc(Un1, 1.0, Un1, 0.0, nullptr, 1.0 / 3.0, Lb)))
        return r;
    if ((r = rhs(Ut)) || (r = lc(Un, 1.0, Un1, 0.0, nullptr, 1.0 / 6.0, Lb))) return r;
    return CMP_OK;
}

int cmp_level_rk_step(cmp_level* L, int integrator, double* Un, double* U1, double* U2, double* U3,
                      const double* dt_dev, double* cmax, void* stream)
{
    if (integrator == 3) return cmp_level_rk3_step(L, Un, U1, U2, dt_dev, cmax, stream);
    if (!L || !Un || !U1 || !U2 || !dt_dev || !cmax) return CMP_ERR_ARG;
    return rk_step_impl(L, integrator, Un, U1, U2, U3, dt_dev, cmax, nullptr, stream);
}

int cmp_level_initial_cmax(cmp_level* L, double* U, double* cmax, void* stream)
{
    if (!L || !U || !cmax) return CMP_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    CMP_TRY(hipMemsetAsync(cmax, 0, sizeof(double), s));
    int r = cmp_level_fill_ghosts(L, U, stream);
    if (r) return r;
    CMP_TRY(launch_stage(L->lc, L->kp, MODE_CMAX, 0, L->d_tiles, L->ntiles, L->d_boxes, nullptr, 0, U,
                         nullptr, nullptr, nullptr, (unsigned long long*)cmax, L->d_err, s));
    return CMP_OK;
}

int cmp_level_compute_dt(cmp_level* L, double* cmax, double* dt_dev, double cfl, int limiter,
                         void* stream)
{
    if (!L || !cmax || !dt_dev) return CMP_ERR_ARG;
    // no history is written by this entry point (hist == nullptr)
    CMP_TRY(launch_compute_dt(L->P.dim, L->G.dx, cmax, dt_dev, cfl, limiter, nullptr, L->d_step,
                              L->d_err, (hipStream_t)stream, L->d_clock));
    return CMP_OK;
}

static int level_step_with_hist(cmp_level* L, double* Un, double* U1, double* U2, double cfl,
                                int limiter, double* dt_dev, double* cmax, double* hist,
                                hipStream_t s)
{
    CMP_TRY(launch_compute_dt(L->P.dim, L->G.dx, cmax, dt_dev, cfl, limiter, hist, L->d_step,
                              L->d_err, s, L->d_clock));
    if (L->integrator != 3) return rk_step_impl(L, L->integrator, Un, U1, U2, L->d_U3, dt_dev, cmax, nullptr, s);
    return cmp_level_rk3_step(L, Un, U1, U2, dt_dev, cmax, s);
}

int cmp_level_set_time(cmp_level* L, double time, double stop_time, void* stream)
{
    if (!L) return CMP_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    // in the order of the stream's queued work (a queued ComputeDt advances the clock)
    double* clk = L->h_clock;
    CMP_TRY(hipStreamSynchronize(s));
    clk[0] = time;
    clk[1] = stop_time;
    CMP_TRY(hipMemcpyAsync(L->d_clock, clk, sizeof(double) * 2, hipMemcpyHostToDevice, s));
    CMP_TRY(hipStreamSynchronize(s));
    L->stop_time = stop_time;
    return CMP_OK;
}

int cmp_level_get_time(const cmp_level* L, double* time, void* stream)
{
    if (!L || !time) return CMP_ERR_ARG;
    CMP_TRY(hipStreamSynchronize((hipStream_t)stream));
    CMP_TRY(hipMemcpy(time, L->d_clock, sizeof(double), hipMemcpyDeviceToHost));
    return CMP_OK;
}

int cmp_level_set_integrator(cmp_level* L, int integrator)
{
    if (!L || integrator < 1 || integrator > 4) return CMP_ERR_ARG;
    if (integrator == 4 && !L->d_U3) {
        long long total = 0;
        for (const cmp_box& b : L->boxes)
            total = std::max(total, b.offset + cmp_fab_size(L->P.dim, &b, L->ng, L->P.nstate));
        CMP_TRY(hipMalloc(&L->d_U3, sizeof(double) * (size_t)std::max(1LL, total)));
    }
    L->integrator = integrator;
    return CMP_OK;
}

int cmp_level_advance(cmp_level* L, double* Un, double* U1, double* U2, int nsteps, double cfl,
                      int limiter, double* dt_dev, double* cmax, double* dt_hist, int use_graph,
                      int resume, void* stream)
{
    if (!L || !Un || !U1 || !U2 || !dt_dev || !cmax || nsteps < 0) return CMP_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    int r;
    if (!resume) {
        // step 0 takes the wave speed of the initial state; dt_prev = 0 disables the limiter there
        if ((r = cmp_level_initial_cmax(L, Un, cmax, stream))) return r;
        CMP_TRY(hipMemsetAsync(dt_dev, 0, sizeof(double), s));
    }
    // dt_hist is indexed by the steps of this call and holds at most nsteps entries
    CMP_TRY(launch_set2(L->d_step, 0, dt_hist ? nsteps : 0, s));
    if (!use_graph || nsteps < 2) {
        for (int n = 0; n < nsteps; ++n) {
            if ((r = level_step_with_hist(L, Un, U1, U2, cfl, limiter, dt_dev, cmax, dt_hist, s))) return r;
            if (L->pf_on && (r = cmp_level_pf_evolve(L, Un, U1, U2, dt_dev, nullptr, stream))) return r;
        }
        return CMP_OK;
    }
    // graph of one step, replayed; captured on a private stream when the caller gave the legacy
    // default stream, which cannot be captured
    hipStream_t cs = s;
    if (!cs) {
        if (!L->own_stream) CMP_TRY(hipStreamCreateWithFlags(&L->own_stream, hipStreamNonBlocking));
        cs = L->own_stream;
        CMP_TRY(hipStreamSynchronize(nullptr));
    }
    const void* key[6] = {Un, U1, U2, dt_dev, cmax, dt_hist};
    bool same = L->gexec && L->gcfl == cfl && L->glimiter == limiter && L->ginteg == L->integrator;
    for (int q = 0; q < 6 && same; ++q) same = (L->gkey[q] == key[q]);
    if (!same) {
        if (L->gexec) { CMP_TRY(hipGraphExecDestroy(L->gexec)); L->gexec = nullptr; }
        if (L->graph) { CMP_TRY(hipGraphDestroy(L->graph)); L->graph = nullptr; }
        CMP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        r = level_step_with_hist(L, Un, U1, U2, cfl, limiter, dt_dev, cmax, dt_hist, cs);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(cs, &g);
        if (r) { if (g) (void)hipGraphDestroy(g); return r; }
        CMP_TRY(e);
        L->graph = g;
        CMP_TRY(hipGraphInstantiate(&L->gexec, g, nullptr, nullptr, 0));
        for (int q = 0; q < 6; ++q) L->gkey[q] = key[q];
        L->gcfl = cfl;
        L->glimiter = limiter;
        L->ginteg = L->integrator;
    }
    for (int n = 0; n < nsteps; ++n) {
        CMP_TRY(hipGraphLaunch(L->gexec, cs));
        if (L->pf_on && (r = cmp_level_pf_evolve(L, Un, U1, U2, dt_dev, nullptr, cs))) return r;
    }
    if (cs != s) CMP_TRY(hipStreamSynchronize(cs));
    return CMP_OK;
}

// ---- Phase-Field step ---------------------------------------------------------------------------

// For every box, the boxes whose valid cells its ng-deep ghost cells are filled from: the owners of
// the blocks covering the box grown by ng, the grown range wrapped on periodic sides and clamped to
// the domain elsewhere (foextrap and reflect sources lie within ng cells of the domain face).
static std::vector<std::vector<int>> ghost_sources(const cmp_level* L)
{
    const int dim = L->P.dim, ng = L->ng;
    std::vector<std::vector<int>> src(L->nbox);
    std::vector<char> seen(L->nbox, 0);
    for (int b = 0; b < L->nbox; ++b) {
        std::vector<int> blk[3];
        for (int d = 0; d < 3; ++d) {
            if (d >= dim) { blk[d].push_back(0); continue; }
            const int dlo = L->G.domain_lo[d], dhi = L->G.domain_hi[d], per = dhi - dlo + 1;
            std::vector<char> hit(L->nb[d], 0);
            for (int c = L->boxes[b].lo[d] - ng; c <= L->boxes[b].hi[d] + ng; ++c) {
                int x = c;
                if (L->G.is_periodic[d]) {
                    x = ((x - dlo) % per + per) % per + dlo;
                } else {
                    x = std::min(std::max(x, dlo), dhi);
                }
                hit[(x - dlo) / L->bs] = 1;
            }
            for (int q = 0; q < L->nb[d]; ++q)
                if (hit[q]) blk[d].push_back(q);
        }
        for (int k : blk[2])
            for (int j : blk[1])
                for (int i : blk[0]) {
                    const long long q = i + (long long)L->nb[0] * (j + (long long)L->nb[1] * k);
                    const int o = L->h_owner[q];
                    if (o >= 0 && !seen[o]) { seen[o] = 1; src[b].push_back(o); }
                }
        for (int o : src[b]) seen[o] = 0;
    }
    return src;
}

int cmp_level_pf_setup(cmp_level* L, const cmp_pf_params* pf)
{
    if (!L || !pf) return CMP_ERR_ARG;
    if ((pf->mechanism != CMP_PF_ACDI && pf->mechanism != CMP_PF_CAC) || pf->id_flux != 0)
        return CMP_ERR_UNSUPPORTED;
    if (pf->explicit_rc != 0 && pf->explicit_rc != 1) return CMP_ERR_ARG;
    // a mechanism that is not flux based has no Jhat (PhaseField_Parameter.H:138-140 aborts)
    if (pf->mechanism == CMP_PF_CAC && pf->explicit_rc == 1) return CMP_ERR_UNSUPPORTED;
    if (pf->flux_phase2flux != 0 && pf->flux_phase2flux != 1) return CMP_ERR_ARG;
    if (!(pf->cfl > 0.0) || !(pf->eta_factor > 0.0) || pf->max_step < 1) return CMP_ERR_ARG;
    if (!(pf->eta > 0.0) && !(pf->eta_multiplier > 0.0)) return CMP_ERR_ARG;
    if (L->nbox > 65535) return CMP_ERR_UNSUPPORTED;   // one grid row per box
    pf_free(L);
    const int dim = L->P.dim, nbox = L->nbox;
    std::vector<cmp_box> ba(L->boxes), bn(L->boxes);
    long long na = 0, nn = 0;
    cmp_level_layout(dim, nbox, ba.data(), 1, 1, &na);
    cmp_level_layout(dim, nbox, bn.data(), 1, dim, &nn);
    std::vector<long long> foff((size_t)nbox * dim);
    const long long nf = face_offsets(dim, nbox, L->boxes.data(), 1, foff.data());
    std::vector<BoxDev> ha(nbox), hn(nbox);
    std::vector<long long> gp(nbox + 1, 0);
    long long maxcells = 0, maxfaces[3] = {0, 0, 0};
    for (int b = 0; b < nbox; ++b) {
        ha[b] = make_boxdev(dim, ba[b], 1);
        hn[b] = make_boxdev(dim, bn[b], 1);
        for (int d = 0; d < dim; ++d) ha[b].goff[d] = hn[b].goff[d] = foff[(size_t)b * dim + d];
        long long fab = 1, val = 1;
        for (int d = 0; d < dim; ++d) {
            const long long n = ba[b].hi[d] - ba[b].lo[d] + 1;
            fab *= n + 2;
            val *= n;
        }
        gp[b + 1] = gp[b] + (fab - val);
        maxcells = std::max(maxcells, val);
        for (int d = 0; d < dim; ++d) {
            long long f = 1;
            for (int e = 0; e < dim; ++e) f *= ba[b].hi[e] - ba[b].lo[e] + 1 + (e == d ? 1 : 0);
            maxfaces[d] = std::max(maxfaces[d], f);
        }
    }
    int rc = CMP_OK;
#define CMP_TRYP(x) do { rc = check_hip(x); if (rc) { pf_free(L); return rc; } } while (0)
    CMP_TRYP(hipMalloc(&L->d_abox, sizeof(BoxDev) * nbox));
    CMP_TRYP(hipMemcpy(L->d_abox, ha.data(), sizeof(BoxDev) * nbox, hipMemcpyHostToDevice));
    CMP_TRYP(hipMalloc(&L->d_nbox, sizeof(BoxDev) * nbox));
    CMP_TRYP(hipMemcpy(L->d_nbox, hn.data(), sizeof(BoxDev) * nbox, hipMemcpyHostToDevice));
    CMP_TRYP(hipMalloc(&L->d_gprefix1, sizeof(long long) * (nbox + 1)));
    CMP_TRYP(hipMemcpy(L->d_gprefix1, gp.data(), sizeof(long long) * (nbox + 1), hipMemcpyHostToDevice));
    CMP_TRYP(hipMalloc(&L->d_A0, sizeof(double) * na));
    CMP_TRYP(hipMalloc(&L->d_A, sizeof(double) * na));
    CMP_TRYP(hipMalloc(&L->d_P, sizeof(double) * na));
    CMP_TRYP(hipMalloc(&L->d_N, sizeof(double) * nn));
    CMP_TRYP(hipMalloc(&L->d_J, sizeof(double) * nf));
    CMP_TRYP(hipMalloc(&L->d_H, sizeof(double) * nf));
    // Grad_Q of the level: one component per volume fraction; the N-phase model computes each
    // phase's into d_gq1 first
    const int npf = L->P.nphase ? L->P.nphase : 1;
    CMP_TRYP(hipMalloc(&L->d_gq, sizeof(double) * nf * npf));
    if (npf > 1) CMP_TRYP(hipMalloc(&L->d_gq1, sizeof(double) * nf));
    CMP_TRYP(hipMalloc(&L->d_bits, sizeof(unsigned long long) * 2));
    CMP_TRYP(hipMalloc(&L->d_pfdt, sizeof(double)));
    CMP_TRYP(hipHostMalloc((void**)&L->h_pin, sizeof(double) * 4, hipHostMallocDefault));
    // the unfilled ghost cells of the normals are never read, but keep them defined
    CMP_TRYP(hipMemset(L->d_N, 0, sizeof(double) * nn));
    // the compact block list of the level's cells: every cell kernel of the Phase-Field step (and the
    // composite solve) launches one workgroup per listed block, and its block partial sums have
    // that many slots
    std::vector<cmp::BoxBlock> cmap;
    for (int b = 0; b < nbox; ++b) {
        long long cells = 1;
        for (int d = 0; d < dim; ++d) cells *= L->boxes[b].hi[d] - L->boxes[b].lo[d] + 1;
        for (long long k = 0; k < (cells + NT - 1) / NT; ++k) cmap.push_back(cmp::BoxBlock{b, (int)k});
    }
    const long long nblk = (long long)cmap.size();
    if (nblk < 1 || nblk > (1LL << 30)) { pf_free(L); return CMP_ERR_UNSUPPORTED; }
    L->nblocks = (int)nblk;
    CMP_TRYP(hipMalloc(&L->d_cmap, sizeof(cmp::BoxBlock) * cmap.size()));
    CMP_TRYP(hipMemcpy(L->d_cmap, cmap.data(), sizeof(cmp::BoxBlock) * cmap.size(), hipMemcpyHostToDevice));
    CMP_TRYP(hipMalloc(&L->d_sc, sizeof(double) * CG_NSC));
    CMP_TRYP(hipMemset(L->d_sc, 0, sizeof(double) * CG_NSC));
    CMP_TRYP(hipMalloc(&L->d_part, sizeof(double) * 4 * nblk));
    CMP_TRYP(hipMemset(L->d_part, 0, sizeof(double) * 4 * nblk));
    if (pf->mechanism == CMP_PF_CAC) {
        CMP_TRYP(hipMalloc(&L->d_Mk, sizeof(double) * na));
        CMP_TRYP(hipMalloc(&L->d_Wk, sizeof(double) * na));
    }
    if (pf->explicit_rc == 0) {
        double** cell[] = {&L->d_DJ, &L->d_X, &L->d_R, &L->d_Pv, &L->d_AP, &L->d_Z, &L->d_ID};
        for (double** p : cell) {
            CMP_TRYP(hipMalloc(p, sizeof(double) * na));
            CMP_TRYP(hipMemset(*p, 0, sizeof(double) * na));
        }
        CMP_TRYP(hipMalloc(&L->d_Bf, sizeof(double) * nf));
    }
    CMP_TRYP(hipMalloc(&L->d_boxflag, sizeof(int) * nbox));
    CMP_TRYP(hipMemset(L->d_boxflag, 0, sizeof(int) * nbox));
    CMP_TRYP(hipHostMalloc((void**)&L->h_boxflag, sizeof(int) * nbox, hipHostMallocDefault));
    CMP_TRYP(hipMalloc(&L->d_pf_tiles, sizeof(TileDev) * std::max(1, L->ntiles)));
    CMP_TRYP(hipHostMalloc((void**)&L->h_pf_tiles, sizeof(TileDev) * std::max(1, L->ntiles), hipHostMallocDefault));
#undef CMP_TRYP
    L->ghost_src = ghost_sources(L);
    long long ncell = 0;
    for (int b = 0; b < nbox; ++b) {
        long long v = 1;
        for (int d = 0; d < dim; ++d) v *= ba[b].hi[d] - ba[b].lo[d] + 1;
        ncell += v;
    }
    L->ncell = ncell;
    L->na = na;
    L->nfaces = nf;
    L->maxcells = maxcells;
    for (int d = 0; d < 3; ++d) L->maxfaces[d] = maxfaces[d];
    L->nghost1 = gp[nbox];
    // PF_bcs / LS_bcs / Normal_bcs: int_dir on periodic sides, foextrap elsewhere
    // (source/Class_Compressible_PhaseField.cpp:471-488, 1516-1530, 1970-1986)
    L->fdA = L->fd;
    L->fdA.boxes = L->d_abox;
    L->fdA.gprefix = L->d_gprefix1;
    L->fdA.ng = 1;
    L->fdA.maxghost = 0;
    for (int b = 0; b < nbox; ++b) L->fdA.maxghost = std::max(L->fdA.maxghost, gp[b + 1] - gp[b]);
    for (int d = 0; d < 3; ++d) {
        L->fdA.lo_bc[d] = 2;
        L->fdA.hi_bc[d] = 2;
    }
    L->fdN = L->fdA;
    L->fdN.boxes = L->d_nbox;
    L->pg.hbox = L->d_boxes;
    L->pg.abox = L->d_abox;
    L->pg.nbox = nbox;
    L->pg.cmap = L->d_cmap;
    L->pg.ncmap = L->nblocks;
    for (int d = 0; d < 3; ++d) {
        L->pg.dx[d] = (d < dim) ? L->G.dx[d] : 1.0;
        L->pg.dlo[d] = (d < dim) ? L->G.domain_lo[d] : 0;
        L->pg.dhi[d] = (d < dim) ? L->G.domain_hi[d] : 0;
        L->pg.periodic[d] = (d < dim) ? L->G.is_periodic[d] : 1;
    }
    L->pf = *pf;
    L->pf_lc = L->lc;
    L->pf_lc.thinc = 0;
    L->pf_lc.riemann = (pf->flux_phase2flux == 1) ? RIEMANN_PF_MIX : RIEMANN_PF_PHASIC;
    L->pf_kp = L->kp;
    L->pf_kp.thinc = 0;
    for (int k = 0; k < CMP_MAX_PHASES; ++k) L->pf_kp.pf_pmin[k] = pf->p_min[k];
    if (pf->explicit_rc == 0) {
        rc = mg_setup(L);
        if (rc) { pf_free(L); return rc; }
    }
    L->pf_on = true;
    return CMP_OK;
}

int cmp_level_pf_rhs(cmp_level* L, const double* U, const double* gradq, double* dUdt, void* stream)
{
    if (!L || !U || !gradq || !dUdt) return CMP_ERR_ARG;
    if (!L->pf_on) return CMP_ERR_ARG;
    CMP_TRY(launch_stage(L->pf_lc, L->pf_kp, MODE_RHS, 0, L->d_tiles, L->ntiles, L->d_boxes, nullptr, 0,
                         U, nullptr, dUdt, nullptr, nullptr, L->d_err, (hipStream_t)stream, gradq));
    return CMP_OK;
}

static inline double bits_to_double(const void* p)
{
    double v;
    std::memcpy(&v, p, sizeof(double));
    return v;
}

}  // extern "C"

using cmp::mg_prepare;
using cmp::mg_vcycle;

static int mg_launch(int dim, int op, const cmp::MgArgs& a, hipStream_t s)
{
    CMP_TRY(dim == 2 ? cmp::launch_mg_dim<2>(op, a, s) : cmp::launch_mg_dim<3>(op, a, s));
    return CMP_OK;
}

static cmp::MgArgs mg_args(const cmp_level* L, const double* sc)
{
    cmp::MgArgs a{};
    a.g = L->pg;
    a.maxcells = L->maxcells;
    for (int d = 0; d < 3; ++d) a.maxfaces[d] = L->maxfaces[d];
    a.sc = sc;
    a.omega = L->mg_omega;
    return a;
}

// Face coefficients of every multigrid grid (the level's b gathered, then restricted) and their
// inverse diagonals; once per solve.
int cmp::mg_prepare(cmp_level* L, hipStream_t s)
{
    const int dim = L->P.dim;
    cmp::MgArgs a = mg_args(L, nullptr);
    a.src = L->d_Bf;
    for (int d = 0; d < 3; ++d) a.b[d] = L->mg[0].b[d];
    a.G = L->mg[0].G;
    int r = mg_launch(dim, cmp::MG_GATHER_B, a, s);
    for (size_t m = 0; !r && m + 1 < L->mg.size(); ++m) {
        for (int d = 0; d < 3; ++d) {
            a.b[d] = L->mg[m].b[d];
            a.bc[d] = L->mg[m + 1].b[d];
        }
        a.G = L->mg[m].G;
        a.Gc = L->mg[m + 1].G;
        r = mg_launch(dim, cmp::MG_RESTRICT_B, a, s);
    }
    for (size_t m = 0; !r && m < L->mg.size(); ++m) {
        for (int d = 0; d < 3; ++d) a.b[d] = L->mg[m].b[d];
        a.G = L->mg[m].G;
        a.dst = L->mg[m].id;
        r = mg_launch(dim, cmp::MG_DIAG, a, s);
    }
    return r;
}

// One V-cycle on grid m from its residual r, the result left in x[cur]: two damped Jacobi sweeps
// (the first from zero), the residual restricted, the coarser grid's V-cycle prolonged and added,
// two more sweeps; the coarsest grid takes its fixed number of sweeps in one workgroup.
int cmp::mg_vcycle(cmp_level* L, const double* sc, size_t m, hipStream_t s)
{
    const int dim = L->P.dim;
    auto& V = L->mg[m];
    if (V.G.nc <= cmp::MG_TAIL_MAX_CELLS && L->mg.size() - m <= (size_t)cmp::MG_TAIL_MAX_LEVELS) {
        // this grid and every coarser one in a single workgroup (one launch instead of six a grid)
        cmp::MgTail t{};
        t.nlev = (int)(L->mg.size() - m);
        t.sweeps = L->mg_bottom_sweeps;
        t.omega = L->mg_omega;
        t.sc = sc;
        for (int q = 0; q < t.nlev; ++q) {
            const auto& W = L->mg[m + q];
            t.G[q] = W.G;
            t.x0[q] = W.x[0];
            t.x1[q] = W.x[1];
            t.r[q] = W.r;
            t.id[q] = W.id;
            for (int d = 0; d < 3; ++d) t.b[q][d] = W.b[d];
        }
        // the coarsest grids that fit the tail's LDS together (never the first, whose result the
        // caller reads from device memory)
        auto need = [&](int q) -> long long {
            long long n = 4 * t.G[q].nc;
            for (int d = 0; d < dim; ++d) {
                long long nf = 1;
                for (int e = 0; e < dim; ++e) nf *= t.G[q].n[e] + (e == d);
                n += nf;
            }
            return n;
        };
        t.lds_from = t.nlev;
        long long used = 0;
        for (int q = t.nlev - 1; q >= 1 && used + need(q) <= cmp::MG_TAIL_LDS; --q) {
            used += need(q);
            t.lds_from = q;
        }
        long long off = 0;
        for (int q = t.lds_from; q < t.nlev; ++q) {
            const long long nc = t.G[q].nc;
            for (int a = 0; a < 4; ++a) t.loff[q][a] = (int)(off + a * nc);
            off += 4 * nc;
            for (int d = 0; d < 3; ++d) {
                t.loff[q][4 + d] = (int)off;
                if (d < dim) {
                    long long nf = 1;
                    for (int e = 0; e < dim; ++e) nf *= t.G[q].n[e] + (e == d);
                    off += nf;
                }
            }
        }
        CMP_TRY(dim == 2 ? cmp::launch_mg_tail_dim<2>(t, s) : cmp::launch_mg_tail_dim<3>(t, s));
        V.cur = (t.nlev == 1) ? 0 : 1;
        return CMP_OK;
    }
    cmp::MgArgs a = mg_args(L, sc);
    for (int d = 0; d < 3; ++d) a.b[d] = V.b[d];
    a.G = V.G;
    a.r = V.r;
    a.id = V.id;
    int r;
    if (m + 1 == L->mg.size()) {
        a.dst = V.x[0];
        a.sweeps = L->mg_bottom_sweeps;
        V.cur = 0;
        return mg_launch(dim, cmp::MG_BOTTOM, a, s);
    }
    a.zero_x = 1;
    a.dst = V.x[0];
    if ((r = mg_launch(dim, cmp::MG_SMOOTH, a, s))) return r;
    a.zero_x = 0;
    a.src = V.x[0];
    a.dst = V.x[1];
    if ((r = mg_launch(dim, cmp::MG_SMOOTH, a, s))) return r;
    auto& C = L->mg[m + 1];
    a.src = V.x[1];
    a.dst = C.r;
    a.Gc = C.G;
    if ((r = mg_launch(dim, cmp::MG_RESID_RESTRICT, a, s))) return r;
    if ((r = mg_vcycle(L, sc, m + 1, s))) return r;
    a.src = C.x[C.cur];
    a.dst = V.x[1];
    if ((r = mg_launch(dim, cmp::MG_PROLONG, a, s))) return r;
    a.src = V.x[1];
    a.dst = V.x[0];
    if ((r = mg_launch(dim, cmp::MG_SMOOTH, a, s))) return r;
    a.src = V.x[0];
    a.dst = V.x[1];
    if ((r = mg_launch(dim, cmp::MG_SMOOTH, a, s))) return r;
    V.cur = 1;
    return CMP_OK;
}

// z = V(r) for the CG's residual (level buffers R -> Z), then the block partials of r.z, sum z and
// sum r into the second to fourth slots of d_part.
static int mg_precondition(cmp_level* L, const double* sc, hipStream_t s)
{
    const int dim = L->P.dim;
    cmp::MgArgs a = mg_args(L, sc);
    a.G = L->mg[0].G;
    a.src = L->d_R;
    a.dst = L->mg[0].r;
    int r = mg_launch(dim, cmp::MG_GATHER_R, a, s);
    if (!r) r = mg_vcycle(L, sc, 0, s);
    if (r) return r;
    a.src = L->mg[0].x[L->mg[0].cur];
    a.dst = L->d_Z;
    if ((r = mg_launch(dim, cmp::MG_SCATTER, a, s))) return r;
    a.src = L->d_R;
    a.r = L->d_Z;
    a.part = L->d_part;
    return mg_launch(dim, cmp::MG_CG_RZ, a, s);
}

extern "C" {

// The implicit Grad_Q (PhaseField_GradQPF with Div_J, source/Class_Compressible_PhaseField.cpp:2133-2310):
// Div_Jhat = (alpha_new - alpha)/dt (:703); b = W(phi_f) + 1e-3 on the faces from the re-sharpened
// alpha; Q solves div(b grad Q) = Div_Jhat with periodic / homogeneous Neumann sides; Grad_Q = grad Q
// on the faces, zero on Neumann domain faces. The reference hands the system to AMReX's MLMG (relative
// residual 1e-10); here conjugate gradients on the zero-mean subspace, preconditioned by a multigrid
// V-cycle when the level covers its domain (else by the Jacobi diagonal), run to the rounding floor
// (r.r <= 1e-28 of the initial r.r, or 400 iterations without a new minimum), with fixed-order
// reductions and the stopping test on the device. With the V-cycle the preconditioned residual is
// projected off the constant vector (the operator's null space) before it enters the direction.
static int pf_gradq_implicit(cmp_level* L, double dt, double eta, double* gradq, hipStream_t s)
{
    const int dim = L->P.dim;
    const int maxit = 100000;
    double* sc = L->d_sc;
    CMP_TRY(launch_pf_divjhat(dim, L->pg, L->maxcells, L->d_A, L->d_A0, L->d_DJ, dt, s));
    CMP_TRY(launch_pf_map(L->d_A0, L->d_P, L->na, eta, L->pf.eta_factor * eta, 1, s));
    CMP_TRY(launch_pf_bcoef(dim, L->pg, L->maxfaces, L->d_P, L->d_Bf, s));
    CMP_TRY(launch_cell_sum(dim, L->pg, L->maxcells, L->d_DJ, L->d_part, s));
    CMP_TRY(launch_reduce_sums(L->d_part, L->nblocks, 1, sc + CG_SUM, s));
    CMP_TRY(launch_cg_idiag(dim, L->pg, L->maxcells, L->d_Bf, L->d_ID, s));
    CMP_TRY(launch_cg_init(dim, L->pg, L->maxcells, L->d_DJ, sc + CG_SUM, (double)L->ncell, L->d_ID, L->d_X,
                           L->d_R, L->d_Z, L->d_Pv, L->d_part, s));
    const bool mg = !L->mg.empty();
    const double inv_n = mg ? 1.0 / (double)L->ncell : 0.0;
    if (mg) {
        // the V-cycle's face coefficients and diagonals for this b, then z = V(r)
        int r = mg_prepare(L, s);
        if (!r) r = mg_precondition(L, nullptr, s);
        if (r) return r;
    }
    CMP_TRY(launch_reduce_sums(L->d_part, L->nblocks, 2, sc + CG_RR, s));
    if (mg) {
        // p = z - mean(z)
        CMP_TRY(launch_reduce_sums(L->d_part + 2 * (long long)L->nblocks, L->nblocks, 2, sc + CG_ZS, s));
        CMP_TRY(launch_cg_start(sc, s, inv_n));
        CMP_TRY(launch_cg_p(dim, L->pg, L->maxcells, sc, L->d_Z, L->d_Pv, s, inv_n, nullptr, 1));
    } else {
        CMP_TRY(launch_cg_start(sc, s));
    }
    // the host looks at the stop flag every `chunk` iterations; iterations queued past the stop
    // return at once. A chunk's launches depend only on the level's buffers, so it is captured once
    // into a graph and replayed (one launch instead of 60-80 per iteration).
    const int chunk = mg ? 8 : 16;
    auto enqueue_chunk = [&](hipStream_t cs) -> int {
        for (int q = 0; q < chunk; ++q) {
            CMP_TRY(launch_fill(dim, 1, L->fdA, L->nghost1, L->d_Pv, L->d_err, cs));
            CMP_TRY(launch_cg_ap(dim, L->pg, L->maxcells, sc, L->d_Bf, L->d_Pv, L->d_AP, L->d_part, cs));
            CMP_TRY(launch_reduce_sums(L->d_part, L->nblocks, 1, sc + CG_PAP, cs, sc));
            CMP_TRY(launch_cg_xr(dim, L->pg, L->maxcells, sc, mg ? nullptr : L->d_ID, L->d_X, L->d_R, L->d_Z,
                                 L->d_Pv, L->d_AP, L->d_part, cs));
            if (mg) {
                int r = mg_precondition(L, sc, cs);
                if (r) return r;
            }
            CMP_TRY(launch_reduce_sums(L->d_part, L->nblocks, 2, sc + CG_RRNEW, cs, sc));
            if (mg)
                CMP_TRY(launch_reduce_sums(L->d_part + 2 * (long long)L->nblocks, L->nblocks, 2, sc + CG_ZS, cs,
                                           sc));
            CMP_TRY(launch_cg_p(dim, L->pg, L->maxcells, sc, L->d_Z, L->d_Pv, cs, inv_n));
            CMP_TRY(launch_cg_end(sc, maxit, cs, inv_n));
        }
        return CMP_OK;
    };
    // captured on a private stream when the caller gave the legacy default stream
    hipStream_t cs = s;
    if (!cs) {
        if (!L->own_stream) CMP_TRY(hipStreamCreateWithFlags(&L->own_stream, hipStreamNonBlocking));
        cs = L->own_stream;
        CMP_TRY(hipStreamSynchronize(nullptr));
    }
    if (!L->cg_exec) {
        CMP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        const int r = enqueue_chunk(cs);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(cs, &g);
        if (r) { if (g) (void)hipGraphDestroy(g); return r; }
        CMP_TRY(e);
        const hipError_t ei = hipGraphInstantiate(&L->cg_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        CMP_TRY(ei);
    }
    for (long long queued = 0;; queued += chunk) {
        CMP_TRY(hipGraphLaunch(L->cg_exec, cs));
        CMP_TRY(hipMemcpyAsync(L->h_pin + 2, sc + CG_DONE, sizeof(double), hipMemcpyDeviceToHost, cs));
        CMP_TRY(hipMemcpyAsync(L->h_pin + 3, sc + CG_IT, sizeof(double), hipMemcpyDeviceToHost, cs));
        CMP_TRY(hipStreamSynchronize(cs));
        if (L->h_pin[2] != 0.0) break;
        if (queued > maxit) return CMP_ERR_NAN;   // the device stop test failed to fire
    }
    L->cg_iters = (int)L->h_pin[3];
    CMP_TRY(launch_fill(dim, 1, L->fdA, L->nghost1, L->d_X, L->d_err, s));
    CMP_TRY(launch_pf_gradx(dim, L->pg, L->maxfaces, L->d_X, gradq, L->d_bits + 1, L->d_boxflag, s));
    return CMP_OK;
}

int cmp_level_pf_solver_iterations(const cmp_level* L, int* iterations)
{
    if (!L || !iterations) return CMP_ERR_ARG;
    *iterations = L->cg_iters;
    return CMP_OK;
}

int cmp_level_pf_gradq(cmp_level* L, const double* U, double dt, double* gradq, double* gradq_max,
                       double* info, void* stream)
{
    if (!L || !U || !gradq) return CMP_ERR_ARG;
    if (!L->pf_on) return CMP_ERR_ARG;
    if (!(dt > 0.0)) return CMP_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const int dim = L->P.dim;
    const cmp_pf_params& pf = L->pf;
    CMP_TRY(launch_zero_u64(L->d_bits, 2, s));
    CMP_TRY(hipMemsetAsync(L->d_boxflag, 0, sizeof(int) * L->nbox, s));
    const int nph = L->P.nphase;
    CMP_TRY(launch_pf_maxvel(dim, L->pg, L->maxcells, U, L->d_bits, nph ? nph : 2, s));
    CMP_TRY(hipMemcpyAsync(L->h_pin, L->d_bits, sizeof(double), hipMemcpyDeviceToHost, s));
    CMP_TRY(hipStreamSynchronize(s));
    const double maxv = bits_to_double(L->h_pin);
    if (__builtin_isnan(maxv)) return CMP_ERR_NAN;
    // Lambda, eta, mobility (PhaseField_Evolve, source/Class_Compressible_PhaseField.cpp:427-465)
    double lambda = pf.lambda;
    if (lambda <= 0.0) lambda = 1.0;
    double eta = pf.eta;
    if (eta <= 0.0) {
        double m = L->G.dx[0];
        for (int d = 1; d < dim; ++d) m = (m < L->G.dx[d]) ? L->G.dx[d] : m;
        eta = pf.eta_multiplier * m;
    }
    double mob = 1.0;
    mob *= 0.5 * eta * maxv;
    // mechanism sub-steps (PhaseFieldMechanism_Evolve, :642-692)
    double sdx = 1.0 / L->G.dx[0] / L->G.dx[0];
    for (int d = 1; d < dim; ++d) sdx = sdx + 1.0 / L->G.dx[d] / L->G.dx[d];
    double dt_m = 0.9 / (2.0 * mob * lambda) / (sdx + 1.0 / eta / eta);
    const double mob_face = 0.5 * mob + 0.5 * mob;
    const bool cac = (pf.mechanism == CMP_PF_CAC);
    const bool explicit_rc = (pf.explicit_rc == 1);
    // one volume fraction (two-phase model) or each phase's in turn (N-phase model: PhaseField_Mechanism
    // and PhaseField_ACDI loop over the phases, :1036-1150, :1495-1644); Grad_Q's max and the box
    // flags gather over the phases
    const int npf = nph ? nph : 1;
    long long n = 0;
    for (int ph = 0; ph < npf; ++ph) {
        double* out = (npf == 1) ? gradq : L->d_gq1;
        // alpha of the step (A0) and the mechanism's working copy (A, its ghosts filled by each sub-step)
        CMP_TRY(launch_pf_extract(dim, L->pg, L->maxcells, U, L->d_A0, L->d_A, nph ? nph + dim + 1 + ph : 3 + dim,
                                  s));
        CMP_TRY(launch_fill(dim, 1, L->fdA, L->nghost1, L->d_A0, L->d_err, s));
        if (explicit_rc) CMP_TRY(hipMemsetAsync(L->d_H, 0, sizeof(double) * L->nfaces, s));
        double t = 0.0, dt_s = dt_m;
        const double T = dt;
        n = 0;
        while (t < T) {
            if (t + dt_s > T) dt_s = T - t;
            if (!(dt_s > 0.0) || t + dt_s == t || n >= 10000000) return CMP_ERR_NAN;   // no progress
            CMP_TRY(launch_fill(dim, 1, L->fdA, L->nghost1, L->d_A, L->d_err, s));
            if (cac) {
                // normals of alpha itself; mech and W per cell, their level sums, alpha += dt_m (mech + W L)
                CMP_TRY(launch_pf_normals(dim, L->pg, L->maxcells, L->d_A, L->d_N, s));
                CMP_TRY(launch_fill(dim, dim, L->fdN, L->nghost1, L->d_N, L->d_err, s));
                CMP_TRY(launch_pf_cac_mech(dim, L->pg, L->maxcells, L->d_A, L->d_N, L->d_Mk, L->d_Wk, mob, lambda,
                                           eta, L->d_part, s));
                CMP_TRY(launch_reduce_sums(L->d_part, L->nblocks, 2, L->d_sc + CG_SUM, s));
                CMP_TRY(launch_pf_cac_update(dim, L->pg, L->maxcells, L->d_A, L->d_Mk, L->d_Wk, L->d_sc + CG_SUM,
                                             dt_s, s));
            } else {
                CMP_TRY(launch_pf_map(L->d_A, L->d_P, L->na, eta, eta, 0, s));
                CMP_TRY(launch_pf_normals(dim, L->pg, L->maxcells, L->d_P, L->d_N, s));
                CMP_TRY(launch_fill(dim, dim, L->fdN, L->nghost1, L->d_N, L->d_err, s));
                // Jhat is accumulated for the explicit RC only
                CMP_TRY(launch_pf_jfaces(dim, L->pg, L->maxfaces, L->d_A, L->d_P, L->d_N, L->d_J,
                                         explicit_rc ? L->d_H : nullptr, mob_face, eta, dt_s, s));
                CMP_TRY(launch_pf_update(dim, L->pg, L->maxcells, L->d_A, L->d_J, dt_s, s));
            }
            t += dt_s;
            ++n;
        }
        if (explicit_rc) {
            // explicit Grad_Q from the step's alpha (PhaseField_GradQPF, :2315-2403), re-sharpened at
            // eta_factor * eta, and Jhat = H (1/dt)
            CMP_TRY(launch_pf_map(L->d_A0, L->d_P, L->na, eta, pf.eta_factor * eta, 1, s));
            CMP_TRY(launch_pf_gradq(dim, L->pg, L->maxfaces, L->d_P, L->d_H, 1.0 / dt, out, L->d_bits + 1,
                                    L->d_boxflag, s));
        } else {
            int r = pf_gradq_implicit(L, dt, eta, out, s);
            if (r) return r;
        }
        if (npf > 1) CMP_TRY(launch_pf_scatter(dim, L->pg, L->maxfaces, out, gradq, ph, s));
    }
    CMP_TRY(hipMemcpyAsync(L->h_pin + 1, L->d_bits + 1, sizeof(double), hipMemcpyDeviceToHost, s));
    CMP_TRY(hipMemcpyAsync(L->h_boxflag, L->d_boxflag, sizeof(int) * L->nbox, hipMemcpyDeviceToHost, s));
    CMP_TRY(hipStreamSynchronize(s));
    if (gradq_max) *gradq_max = bits_to_double(L->h_pin + 1);
    if (info) {
        info[0] = maxv;
        info[1] = mob;
        info[2] = eta;
        info[3] = (double)n;
    }
    return CMP_OK;
}

// One Phase-Field RK3 sub-step over the given tiles; cells beyond a tile's box are read where the
// ghost fill would take them from.
static int pf_rk3(cmp_level* L, double* Un, double* U1, double* U2, const double* gradq, double pf_dt,
                  const TileDev* tiles, int ntiles, hipStream_t s)
{
    CMP_TRY(launch_setd(L->d_pfdt, pf_dt, s));
    double* Us[3] = {Un, U1, U2};
    double* Uo[3] = {U1, U2, Un};
    for (int stage = 0; stage < 3; ++stage) {
        if (ntiles > 0)
            CMP_TRY(launch_stage(L->pf_lc, L->pf_kp, MODE_RK, stage, tiles, ntiles, L->d_boxes, nullptr, 0,
                                 Us[stage], Un, Uo[stage], L->d_pfdt, nullptr, L->d_err, s, gradq, L->d_fd));
    }
    return CMP_OK;
}

int cmp_level_pf_rk3_step(cmp_level* L, double* Un, double* U1, double* U2, const double* gradq,
                          double pf_dt, void* stream)
{
    if (!L || !Un || !U1 || !U2 || !gradq) return CMP_ERR_ARG;
    if (!L->pf_on) return CMP_ERR_ARG;
    return pf_rk3(L, Un, U1, U2, gradq, pf_dt, L->d_tiles, L->ntiles, (hipStream_t)stream);
}

int cmp_level_pf_rk_step(cmp_level* L, int integrator, double* Un, double* U1, double* U2, double* U3,
                         const double* gradq, double pf_dt, void* stream)
{
    if (!L || !Un || !U1 || !U2 || !gradq) return CMP_ERR_ARG;
    if (!L->pf_on) return CMP_ERR_ARG;
    if (integrator == 3) return pf_rk3(L, Un, U1, U2, gradq, pf_dt, L->d_tiles, L->ntiles, (hipStream_t)stream);
    CMP_TRY(launch_setd(L->d_pfdt, pf_dt, (hipStream_t)stream));
    return rk_step_impl(L, integrator, Un, U1, U2, U3, L->d_pfdt, nullptr, gradq, stream);
}

// The boxes the Phase-Field sub-steps of one step touch, from the box flags of the last Grad_Q:
// the flagged boxes (ghosts filled, fluxes computed) and the boxes their ghosts are filled from
// (RK combination with a zero flux, so that their stage states hold what the filled ghosts read).
// A box outside both has a zero flux on every face; its state is left as it is.
static int pf_select_boxes(cmp_level* L, hipStream_t s)
{
    std::vector<char> need(L->nbox, 0);
    int nfill = 0;
    for (int b = 0; b < L->nbox; ++b) {
        if (!L->h_boxflag[b]) continue;
        ++nfill;
        need[b] = 1;
        for (int o : L->ghost_src[b]) need[o] = 1;
    }
    int nt = 0;
    for (int b = 0; b < L->nbox; ++b) {
        if (!need[b]) continue;
        for (int t = L->tile_start[b]; t < L->tile_start[b + 1]; ++t) L->h_pf_tiles[nt++] = L->h_tiles[t];
    }
    L->pf_nfill = nfill;
    L->pf_ntiles = nt;
    if (nt > 0) CMP_TRY(hipMemcpyAsync(L->d_pf_tiles, L->h_pf_tiles, sizeof(TileDev) * nt, hipMemcpyHostToDevice, s));
    return CMP_OK;
}

int cmp_level_pf_evolve(cmp_level* L, double* Un, double* U1, double* U2, const double* dt_dev,
                        double* info, void* stream)
{
    if (!L || !Un || !U1 || !U2 || !dt_dev) return CMP_ERR_ARG;
    if (!L->pf_on) return CMP_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    CMP_TRY(hipMemcpyAsync(L->h_pin + 2, dt_dev, sizeof(double), hipMemcpyDeviceToHost, s));
    CMP_TRY(hipStreamSynchronize(s));
    const double dt = L->h_pin[2];
    double gm = 0.0, inf4[4] = {0.0, 0.0, 0.0, 0.0};
    if (dt == 0.0) {
        // a zero step (the level's clock past its stop_time) has no Phase-Field step either
        if (info)
            for (int q = 0; q < 6; ++q) info[q] = 0.0;
        return CMP_OK;
    }
    int r = cmp_level_pf_gradq(L, Un, dt, L->d_gq, &gm, inf4, stream);
    if (r) return r;
    if ((r = pf_select_boxes(L, s))) return r;
    // Phase-Field sub-steps (PhaseField_Evolve :541-590, PhaseField_ComputeDt :2037-2098)
    const int dim = L->P.dim;
    double dxm = L->G.dx[0];
    for (int d = 1; d < dim; ++d) dxm = (L->G.dx[d] < dxm) ? L->G.dx[d] : dxm;
    double pf_time = 0.0, prev = 0.0;
    int n = 0;
    for (int step = 0; step < L->pf.max_step && pf_time < dt; ++step) {
        double tmp = L->pf.cfl * dxm / (gm + 1.0e-14);
        if (pf_time == 0.0) prev = tmp;
        double dt0 = tmp;
        if (L->pf.change_limiter && prev > 0.0) tmp = (1.1 * prev < tmp) ? 1.1 * prev : tmp;
        dt0 = (tmp < dt0) ? tmp : dt0;
        const double eps = (1.0e-3) * dt0;
        if (pf_time + dt0 > dt - eps) dt0 = dt - pf_time;
        prev = dt0;
        if (L->pf_nfill > 0 &&
            (r = (L->integrator == 3)
                     ? pf_rk3(L, Un, U1, U2, L->d_gq, dt0, L->d_pf_tiles, L->pf_ntiles, s)
                     : cmp_level_pf_rk_step(L, L->integrator, Un, U1, U2, L->d_U3, L->d_gq, dt0, s)))
            return r;
        pf_time += dt0;
        ++n;
        if (pf_time >= dt - 1.0e-6 * dt) break;
    }
    if (info) {
        for (int q = 0; q < 4; ++q) info[q] = inf4[q];
        info[4] = gm;
        info[5] = (double)n;
    }
    return CMP_OK;
}

int cmp_level_tag(cmp_level* L, const double* U, const cmp_tag_criterion* criteria, int ncriteria,
                  unsigned char* tags, void* stream)
{
    if (!L || !U || !tags || (ncriteria > 0 && !criteria)) return CMP_ERR_ARG;
    if (ncriteria < 0 || ncriteria > CMP_TAG_MAX_CRITERIA) return CMP_ERR_ARG;
    TagCriteria tc{};
    tc.n = ncriteria;
    for (int q = 0; q < ncriteria; ++q) {
        const cmp_tag_criterion& c = criteria[q];
        if (c.var >= L->P.nstate || c.var < CMP_TAG_W || c.kind < CMP_TAG_GRAD || c.kind > CMP_TAG_RANGE)
            return CMP_ERR_ARG;
        tc.c[q] = c;
    }
    if (L->ng < 1) return CMP_ERR_ARG;   // the gradient reads the face neighbours
    if (!L->d_toff) {
        // each box's first tag: the level layout with no ghosts and one component
        std::vector<cmp_box> b(L->boxes);
        long long total = 0;
        int r = cmp_level_layout(L->P.dim, L->nbox, b.data(), 0, 1, &total);
        if (r) return r;
        std::vector<long long> off(L->nbox);
        long long m = 0;
        for (int q = 0; q < L->nbox; ++q) {
            off[q] = b[q].offset;
            long long v = 1;
            for (int d = 0; d < L->P.dim; ++d) v *= b[q].hi[d] - b[q].lo[d] + 1;
            m = std::max(m, v);
        }
        CMP_TRY(hipMalloc(&L->d_toff, sizeof(long long) * L->nbox));
        CMP_TRY(hipMemcpy(L->d_toff, off.data(), sizeof(long long) * L->nbox, hipMemcpyHostToDevice));
        L->tag_maxcells = m;
    }
    CMP_TRY(launch_tag(L->lc, L->kp, L->d_boxes, L->nbox, L->tag_maxcells, L->d_toff, U, tc, tags, L->d_err,
                       (hipStream_t)stream));
    return CMP_OK;
}

int cmp_level_average_down(const cmp_level* F, const double* Uf, cmp_level* C, double* Uc, int ratio,
                           void* stream)
{
    if (!F || !C || !Uf || !Uc || ratio < 2) return CMP_ERR_ARG;
    const int dim = F->P.dim;
    if (C->P.dim != dim || C->P.nstate != F->P.nstate) return CMP_ERR_ARG;
    for (int d = 0; d < dim; ++d) {
        if (F->G.domain_lo[d] != C->G.domain_lo[d] * ratio ||
            F->G.domain_hi[d] + 1 != (C->G.domain_hi[d] + 1) * ratio || C->G.domain_lo[d] < 0)
            return CMP_ERR_LAYOUT;
    }
    long long maxc = 0;
    for (const cmp_box& b : F->boxes) {
        long long v = 1;
        for (int d = 0; d < dim; ++d) {
            if (b.lo[d] % ratio != 0 || (b.hi[d] + 1) % ratio != 0) return CMP_ERR_LAYOUT;
            v *= (b.hi[d] - b.lo[d] + 1) / ratio;
        }
        maxc = std::max(maxc, v);
    }
    double vol = 1.0;
    for (int d = 0; d < dim; ++d) vol *= ratio;
    const double volfrac = 1.0 / vol;
    hipStream_t s = (hipStream_t)stream;
    const bool have = F->d_avgplan && F->avg_uid == C->uid && F->avg_ratio == ratio;
    if (!have) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        CMP_TRY(hipStreamIsCapturing(s, &cs));
        int r;
        if (cs == hipStreamCaptureStatusNone && (r = level_avg_plan(F, C, ratio, s))) return r;
    }
    if (F->d_avgplan && F->avg_uid == C->uid && F->avg_ratio == ratio) {
        // one thread per covered coarse cell and component, through the plan
        CMP_TRY((dim == 2 ? launch_avgdown_apply_dim<2> : launch_avgdown_apply_dim<3>)(
            F->d_boxes, C->d_boxes, F->d_avgplan, F->avg_total, Uf, Uc, F->P.nstate, ratio, volfrac, C->d_err, s));
        return CMP_OK;
    }
    CMP_TRY((dim == 2 ? launch_avgdown_dim<2> : launch_avgdown_dim<3>)(F->d_boxes, F->nbox, maxc, C->fd, Uf, Uc,
                                                                        F->P.nstate, ratio, volfrac, C->d_err, s));
    return CMP_OK;
}

// The checks shared by the two-level fills: C's domain refined by ratio is F's, with the same BCs.
static int two_level_args(const cmp_level* F, const cmp_level* C, int ratio)
{
    const int dim = F->P.dim;
    if (C->P.dim != dim || C->P.nstate != F->P.nstate) return CMP_ERR_ARG;
    const int N = F->P.nstate;
    if (N != 4 + dim && N != 5 + dim && N != 7 + dim) return CMP_ERR_ARG;
    for (int d = 0; d < dim; ++d) {
        if (F->G.domain_lo[d] != C->G.domain_lo[d] * ratio ||
            F->G.domain_hi[d] + 1 != (C->G.domain_hi[d] + 1) * ratio || C->G.domain_lo[d] < 0 ||
            F->G.is_periodic[d] != C->G.is_periodic[d] || F->G.lo_bc[d] != C->G.lo_bc[d] ||
            F->G.hi_bc[d] != C->G.hi_bc[d])
            return CMP_ERR_LAYOUT;
        if (F->ng > F->G.domain_hi[d] - F->G.domain_lo[d] + 1) return CMP_ERR_LAYOUT;
    }
    return CMP_OK;
}

}  // extern "C"

// The gather plan of F's average-down onto C at ratio (k_avgdown_plan; the boxes are aligned to ratio:
// checked by the caller), built when F has none for that pair.
int cmp::level_avg_plan(const cmp_level* F, const cmp_level* C, int ratio, hipStream_t s)
{
    if (F->d_avgplan && F->avg_uid == C->uid && F->avg_ratio == ratio) return CMP_OK;
    for (void** p : {(void**)&F->d_avgprefix, (void**)&F->d_avgplan})
        if (*p) {
            CMP_TRY(hipFree(*p));
            *p = nullptr;
        }
    F->avg_uid = 0;
    F->avg_ratio = 0;
    std::vector<long long> cp(F->nbox + 1, 0);
    for (int b = 0; b < F->nbox; ++b) {
        long long v = 1;
        for (int d = 0; d < F->P.dim; ++d) v *= (F->boxes[b].hi[d] - F->boxes[b].lo[d] + 1) / ratio;
        cp[b + 1] = cp[b] + v;
    }
    CMP_TRY(hipMalloc(&F->d_avgprefix, sizeof(long long) * cp.size()));
    CMP_TRY(hipMemcpy(F->d_avgprefix, cp.data(), sizeof(long long) * cp.size(), hipMemcpyHostToDevice));
    F->avg_total = cp[F->nbox];
    CMP_TRY(hipMalloc(&F->d_avgplan, sizeof(AvgCell) * (size_t)std::max(1LL, F->avg_total)));
    CMP_TRY((F->P.dim == 2 ? launch_avgdown_plan_dim<2> : launch_avgdown_plan_dim<3>)(
        F->d_boxes, F->d_avgprefix, F->nbox, F->avg_total, C->fd, ratio, F->d_avgplan, s));
    F->avg_uid = C->uid;
    F->avg_ratio = ratio;
    return CMP_OK;
}

// The gather plan of F's two-level fill from C at ratio (k_fill2_plan), built when F has none for that
// pair; allocates device memory (not while the stream is captured).
int cmp::level_fill2_plan(cmp_level* F, const cmp_level* C, int ratio, hipStream_t s)
{
    if (F->d_f2plan && F->f2_uid == C->uid && F->f2_ratio == ratio) return CMP_OK;
    if (ratio > 255) return CMP_ERR_UNSUPPORTED;
    if (F->d_f2plan) {
        CMP_TRY(hipFree(F->d_f2plan));
        F->d_f2plan = nullptr;
    }
    F->f2_uid = 0;
    F->f2_ratio = 0;
    const int dim = F->P.dim;
    const size_t bytes = (dim == 2 ? sizeof(F2Cell<2>) : sizeof(F2Cell<3>)) * (size_t)std::max(1LL, F->nghost);
    CMP_TRY(hipMalloc(&F->d_f2plan, bytes));
    CMP_TRY((dim == 2 ? launch_fill2_plan_dim<2> : launch_fill2_plan_dim<3>)(F->fd, C->fd, ratio, F->d_f2plan, s));
    F->f2_uid = C->uid;
    F->f2_ratio = ratio;
    return CMP_OK;
}

extern "C" {

int cmp_level_fill_two_levels(cmp_level* F, double* Uf, const cmp_level* C, const double* Uc, int ratio,
                              void* stream)
{
    if (!F || !C || !Uf || !Uc || ratio < 2) return CMP_ERR_ARG;
    int r = two_level_args(F, C, ratio);
    if (r) return r;
    hipStream_t s = (hipStream_t)stream;
    const int dim = F->P.dim;
    if (!(F->d_f2plan && F->f2_uid == C->uid && F->f2_ratio == ratio)) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        CMP_TRY(hipStreamIsCapturing(s, &cs));
        if (cs == hipStreamCaptureStatusNone && (r = level_fill2_plan(F, C, ratio, s))) return r;
    }
    if (F->d_f2plan && F->f2_uid == C->uid && F->f2_ratio == ratio) {
        CMP_TRY((dim == 2 ? launch_fill2_apply_dim<2> : launch_fill2_apply_dim<3>)(F->P.nstate, F->fd, C->fd, F->nghost,
                                                                                 ratio, F->d_f2plan, Uf, Uc, F->d_err,
                                                                                 s));
        return CMP_OK;
    }
    // under capture without a plan: the search form (the same-level sources are the level's own boxes)
    CMP_TRY((dim == 2 ? launch_fill2_dim<2> : launch_fill2_dim<3>)(F->P.nstate, F->fd, &F->fd, C->fd, ratio, Uf, Uf, Uc,
                                                                   0, 0, F->d_err, s));
    return CMP_OK;
}

int cmp_level_fill_from(cmp_level* F, double* Uf, const cmp_level* O, const double* Uo, const cmp_level* C,
                        const double* Uc, int ratio, void* stream)
{
    if (!F || !C || !Uf || !Uc || ratio < 2 || (O && !Uo)) return CMP_ERR_ARG;
    int r = two_level_args(F, C, ratio);
    if (r) return r;
    const int dim = F->P.dim;
    if (O) {
        if (O->P.dim != dim || O->P.nstate != F->P.nstate) return CMP_ERR_ARG;
        for (int d = 0; d < dim; ++d)
            if (O->G.domain_lo[d] != F->G.domain_lo[d] || O->G.domain_hi[d] != F->G.domain_hi[d]) return CMP_ERR_LAYOUT;
        if (Uo == Uf) return CMP_ERR_ARG;
    }
    long long maxc = 0;
    for (const cmp_box& b : F->boxes) {
        long long v = 1;
        for (int d = 0; d < dim; ++d) v *= b.hi[d] - b.lo[d] + 1 + 2 * F->ng;
        maxc = std::max(maxc, v);
    }
    CMP_TRY((dim == 2 ? launch_fill2_dim<2> : launch_fill2_dim<3>)(F->P.nstate, F->fd, O ? &O->fd : nullptr, C->fd,
                                                                   ratio, Uf, O ? Uo : Uf, Uc, maxc, 1, F->d_err,
                                                                   (hipStream_t)stream));
    return CMP_OK;
}

int cmp_level_check(cmp_level* L, void* stream)
{
    if (!L) return CMP_ERR_ARG;
    CMP_TRY(hipStreamSynchronize((hipStream_t)stream));
    int h = 0;
    CMP_TRY(hipMemcpy(&h, L->d_err, sizeof(int), hipMemcpyDeviceToHost));
    CMP_TRY(hipMemset(L->d_err, 0, sizeof(int)));
    return h;
}

}  // extern "C"

// ---- the allocation-free right-hand side of a box list ------------------------------------------
// cmp_level_rhs_noalloc keeps the device tables cmp_level_create would allocate (the box table,
// then the tile work-list) in this file-scope array, which each GPU holds once for the life of the
// process. In host code the array's name is not a device address: every host use goes through the
// per-device address that scratch_base() looks up with hipGetSymbolAddress.
namespace {
constexpr size_t NOALLOC_BOX_BYTES = ((sizeof(BoxDev) * CMP_NOALLOC_MAX_BOXES + 255) / 256) * 256;
constexpr size_t NOALLOC_BYTES = NOALLOC_BOX_BYTES + sizeof(TileDev) * CMP_NOALLOC_MAX_TILES;
}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch[NOALLOC_MAX_DEVICES] = {};

// The current device's copy of `scratch`, looked up on the first call made with that device current.
int scratch_base(void** out)
{
    int dev = -1;
    CMP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return CMP_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        CMP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return CMP_ERR_HIP;
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return CMP_OK;
}
}  // namespace

extern "C" {

int cmp_level_rhs_noalloc(const cmp_params* params, const cmp_geometry* geom, int nbox, const cmp_box* boxes,
                          int ng, const double* U, double* dUdt, double* cmax, int* err, void* stream)
{
    LaunchCfg lc;
    KParams kp;
    int r = validate(params, geom, ng, &lc, &kp);
    if (r) return r;
    if (nbox <= 0 || !boxes || !U || !dUdt) return CMP_ERR_ARG;
    if (nbox > CMP_NOALLOC_MAX_BOXES) return CMP_ERR_SCRATCH;
    const int dim = params->dim;
    int tx, ty, tz;
    tile_dims(dim, &tx, &ty, &tz);
    const int te[3] = {tx, ty, dim == 3 ? tz : 1};
    // the checks of cmp_level_create: boxes inside the domain, offsets of cmp_level_layout
    long long expect = 0, ntiles = 0;
    for (int b = 0; b < nbox; ++b) {
        long long nt = 1;
        for (int d = 0; d < dim; ++d) {
            if (boxes[b].lo[d] < geom->domain_lo[d] || boxes[b].hi[d] > geom->domain_hi[d] ||
                boxes[b].hi[d] < boxes[b].lo[d])
                return CMP_ERR_LAYOUT;
            nt *= (boxes[b].hi[d] - boxes[b].lo[d] + te[d]) / te[d];
        }
        if (boxes[b].offset != expect) return CMP_ERR_LAYOUT;
        if (cmp_fab_size(dim, &boxes[b], ng, 1) >= (1LL << 31)) return CMP_ERR_UNSUPPORTED;
        expect += cmp_fab_size(dim, &boxes[b], ng, params->nstate);
        ntiles += nt;
    }
    if (ntiles > CMP_NOALLOC_MAX_TILES) return CMP_ERR_SCRATCH;
    void* base = nullptr;
    if ((r = scratch_base(&base))) return r;
    BoxDev* d_boxes = (BoxDev*)base;
    TileDev* d_tiles = (TileDev*)((unsigned char*)base + NOALLOC_BOX_BYTES);
    hipStream_t s = (hipStream_t)stream;
    // the tables go up in fixed-size pieces from the stack, in stream order behind the previous
    // call's kernels (a copy from pageable host memory is staged before hipMemcpyAsync returns)
    constexpr int CH = 64;
    BoxDev hb[CH];
    for (int b0 = 0; b0 < nbox; b0 += CH) {
        const int n = std::min(CH, nbox - b0);
        for (int q = 0; q < n; ++q) hb[q] = make_boxdev(dim, boxes[b0 + q], ng);
        CMP_TRY(hipMemcpyAsync(d_boxes + b0, hb, sizeof(BoxDev) * n, hipMemcpyHostToDevice, s));
    }
    // tiles box by box, in cmp_level_create's order
    TileDev ht[4 * CH];
    int nh = 0;
    long long t0 = 0;
    for (int b = 0; b < nbox; ++b) {
        const cmp_box& B = boxes[b];
        const int klo = (dim == 3) ? B.lo[2] : 0, khi = (dim == 3) ? B.hi[2] : 0;
        for (int k = klo; k <= khi; k += te[2])
            for (int j = B.lo[1]; j <= B.hi[1]; j += ty)
                for (int i = B.lo[0]; i <= B.hi[0]; i += tx) {
                    ht[nh++] = TileDev{b, {i, j, k}, {{HALO_LOOKUP, HALO_LOOKUP}, {HALO_LOOKUP, HALO_LOOKUP},
                                                      {HALO_LOOKUP, HALO_LOOKUP}}};
                    if (nh == 4 * CH) {
                        CMP_TRY(hipMemcpyAsync(d_tiles + t0, ht, sizeof(TileDev) * nh, hipMemcpyHostToDevice, s));
                        t0 += nh;
                        nh = 0;
                    }
                }
    }
    if (nh > 0) CMP_TRY(hipMemcpyAsync(d_tiles + t0, ht, sizeof(TileDev) * nh, hipMemcpyHostToDevice, s));
    CMP_TRY(launch_stage(lc, kp, MODE_RHS, 0, d_tiles, (int)ntiles, d_boxes, nullptr, 0, U, nullptr, dUdt, nullptr,
                         (unsigned long long*)cmax, err, s));
    return CMP_OK;
}

}  // extern "C"
