// cmp_capi.hip - host runtime and C ABI of the five-equation RHS library.
//
// A level plan (cmp_level) owns only device tables: the box table, the tile work-list (tiles of
// 16x16 cells in 2D, 8x8x4 in 3D, clipped to uneven boxes), the ghost-count prefix sums and the
// block->box owner map the ghost fill uses. State buffers are the caller's, as in the reference,
// where every kernel receives non-owning Array4 views of caller-owned MultiFabs.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "cmp_kernels.hpp"

namespace cmp {

template <int DIM>
hipError_t launch_stage_dim(const LaunchCfg&, const KParams&, int, int, const TileDev*, int,
                            const BoxDev*, const BoxDev*, int, const double*, const double*,
                            double*, const double*, unsigned long long*, int*, hipStream_t);
template <int DIM>
hipError_t launch_faces_dim(const LaunchCfg&, const KParams&, const BoxDev&, int, const double*,
                            double*, double*, hipStream_t);
template <int DIM>
hipError_t launch_fill_dim(const FillDev&, long long, double*, int*, hipStream_t);
template <int DIM>
hipError_t launch_dt_dim(const double*, double*, double*, double, int, double*, int*, int*,
                         hipStream_t);

hipError_t launch_stage(const LaunchCfg& lc, const KParams& kp, int mode, int stage,
                        const TileDev* tiles, int ntiles, const BoxDev* boxes,
                        const BoxDev* single_box, int single_ntiles, const double* Us,
                        const double* Un, double* Uout, const double* dt,
                        unsigned long long* cmax_bits, int* err, hipStream_t s)
{
    if (lc.dim == 2)
        return launch_stage_dim<2>(lc, kp, mode, stage, tiles, ntiles, boxes, single_box,
                                   single_ntiles, Us, Un, Uout, dt, cmax_bits, err, s);
    return launch_stage_dim<3>(lc, kp, mode, stage, tiles, ntiles, boxes, single_box, single_ntiles,
                               Us, Un, Uout, dt, cmax_bits, err, s);
}

hipError_t launch_fill(int dim, int /*N*/, const FillDev& fd, long long nghost, double* U, int* err,
                       hipStream_t s)
{
    return dim == 2 ? launch_fill_dim<2>(fd, nghost, U, err, s) : launch_fill_dim<3>(fd, nghost, U, err, s);
}

hipError_t launch_compute_dt(int dim, const double* dx, double* cmax, double* dt, double cfl,
                             int limiter, double* hist, int* step, int* err, hipStream_t s)
{
    return dim == 2 ? launch_dt_dim<2>(dx, cmax, dt, cfl, limiter, hist, step, err, s)
                    : launch_dt_dim<3>(dx, cmax, dt, cfl, limiter, hist, step, err, s);
}

hipError_t launch_faces(const LaunchCfg& lc, const KParams& kp, const BoxDev& box, int dir,
                        const double* U, double* flux, double* fluxnc, hipStream_t s)
{
    return lc.dim == 2 ? launch_faces_dim<2>(lc, kp, box, dir, U, flux, fluxnc, s)
                       : launch_faces_dim<3>(lc, kp, box, dir, U, flux, fluxnc, s);
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

struct cmp_level {
    cmp_params P;
    cmp_geometry G;
    int ng;
    int nbox;
    std::vector<cmp_box> boxes;
    KParams kp;
    LaunchCfg lc;
    BoxDev* d_boxes = nullptr;
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
    const void* gkey[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    double gcfl = 0.0;
    int glimiter = -1;
};

namespace {

int check_hip(hipError_t e)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "compas_amd: HIP error %d (%s)\n", (int)e, hipGetErrorString(e));
        return CMP_ERR_HIP;
    }
    return CMP_OK;
}

#define CMP_TRY(x) do { int _r = check_hip(x); if (_r) return _r; } while (0)

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
    if (G->dim != P->dim || P->nstate != 4 + P->dim) return CMP_ERR_ARG;
    if (P->id_riemann != 1 && P->id_riemann != 2) return CMP_ERR_UNSUPPORTED;
    if (P->source_term != 0 && P->source_term != 1) return CMP_ERR_ARG;
    int scheme = 0;
    int r = scheme_of(P, &scheme);
    if (r) return r;
    if (ng < scheme_ghosts(scheme)) return CMP_ERR_ARG;
    for (int d = 0; d < P->dim; ++d)
        if (!(G->dx[d] > 0.0)) return CMP_ERR_ARG;
    for (int k = 0; k < 2; ++k)
        if (!(P->eos_A[k] > 0.0)) return CMP_ERR_ARG;
    if (P->id_thinc && !(P->thinc_beta > 0.0)) return CMP_ERR_ARG;
    lc->dim = P->dim;
    lc->scheme = scheme;
    lc->limiter = P->id_limiter;
    lc->riemann = P->id_riemann;
    lc->thinc = P->id_thinc;
    std::memset(kp, 0, sizeof(*kp));
    kp->weno5b_weight = (P->id_poly == 2) ? 1.0 / 6.0 : 1.0 / 12.0;
    for (int k = 0; k < 2; ++k) {
        kp->A[k] = P->eos_A[k];
        kp->B[k] = P->eos_B[k];
        kp->C[k] = P->eos_C[k];
        kp->D[k] = P->eos_D[k];
    }
    kp->nasg = (P->eos_C[0] > 0.0 || P->eos_C[1] > 0.0) ? 1 : 0;
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
    for (int d = 0; d < dim; ++d)
        if (fx[d] || fn[d]) CMP_TRY(launch_faces(lc, kp, bd, d, U, fx[d], fn[d], s));
    return CMP_OK;
}

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
        expect += cmp_fab_size(dim, &boxes[b], ng, params->nstate);
    }
    for (int d = 0; d < dim; ++d)
        if (geom->domain_hi[d] - geom->domain_lo[d] + 1 < ng && geom->is_periodic[d]) return CMP_ERR_LAYOUT;

    cmp_level* L = new cmp_level();
    L->P = *params;
    L->G = *geom;
    L->ng = ng;
    L->nbox = nbox;
    L->boxes.assign(boxes, boxes + nbox);
    L->kp = kp;
    L->lc = lc;

    // box table
    std::vector<BoxDev> hb(nbox);
    for (int b = 0; b < nbox; ++b) hb[b] = make_boxdev(dim, boxes[b], ng);
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
                    ht.push_back(TileDev{b, {i, j, k}});
    }
    L->ntiles = (int)ht.size();
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
    fd.nbox = nbox;
    fd.ng = ng;
    fd.bs = (int)bs;
    for (int d = 0; d < 3; ++d) {
        fd.nb[d] = nb[d];
        fd.dlo[d] = geom->domain_lo[d];
        fd.dhi[d] = (d < dim) ? geom->domain_hi[d] : geom->domain_lo[d];
        fd.periodic[d] = geom->is_periodic[d];
        fd.lo_bc[d] = geom->lo_bc[d];
        fd.hi_bc[d] = geom->hi_bc[d];
    }
    *out = L;
    return CMP_OK;
}

int cmp_level_destroy(cmp_level* L)
{
    if (!L) return CMP_OK;
    if (L->gexec) hipGraphExecDestroy(L->gexec);
    if (L->graph) hipGraphDestroy(L->graph);
    if (L->own_stream) hipStreamDestroy(L->own_stream);
    hipFree(L->d_boxes);
    hipFree(L->d_tiles);
    hipFree(L->d_gprefix);
    hipFree(L->d_owner);
    hipFree(L->d_err);
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

int cmp_level_rk3_step(cmp_level* L, double* Un, double* U1, double* U2, const double* dt_dev,
                       double* cmax, void* stream)
{
    if (!L || !Un || !U1 || !U2 || !dt_dev) return CMP_ERR_ARG;
    int r;
    if ((r = cmp_level_fill_ghosts(L, Un, stream))) return r;
    if ((r = cmp_level_rk3_stage(L, 0, Un, Un, U1, dt_dev, cmax, stream))) return r;
    if ((r = cmp_level_fill_ghosts(L, U1, stream))) return r;
    if ((r = cmp_level_rk3_stage(L, 1, U1, Un, U2, dt_dev, cmax, stream))) return r;
    if ((r = cmp_level_fill_ghosts(L, U2, stream))) return r;
    if ((r = cmp_level_rk3_stage(L, 2, U2, Un, Un, dt_dev, cmax, stream))) return r;
    return CMP_OK;
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
                              L->d_err, (hipStream_t)stream));
    return CMP_OK;
}

static int level_step_with_hist(cmp_level* L, double* Un, double* U1, double* U2, double cfl,
                                int limiter, double* dt_dev, double* cmax, double* hist,
                                hipStream_t s)
{
    CMP_TRY(launch_compute_dt(L->P.dim, L->G.dx, cmax, dt_dev, cfl, limiter, hist, L->d_step,
                              L->d_err, s));
    return cmp_level_rk3_step(L, Un, U1, U2, dt_dev, cmax, s);
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
        for (int n = 0; n < nsteps; ++n)
            if ((r = level_step_with_hist(L, Un, U1, U2, cfl, limiter, dt_dev, cmax, dt_hist, s))) return r;
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
    bool same = L->gexec && L->gcfl == cfl && L->glimiter == limiter;
    for (int q = 0; q < 6 && same; ++q) same = (L->gkey[q] == key[q]);
    if (!same) {
        if (L->gexec) { hipGraphExecDestroy(L->gexec); L->gexec = nullptr; }
        if (L->graph) { hipGraphDestroy(L->graph); L->graph = nullptr; }
        CMP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        r = level_step_with_hist(L, Un, U1, U2, cfl, limiter, dt_dev, cmax, dt_hist, cs);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(cs, &g);
        if (r) { if (g) hipGraphDestroy(g); return r; }
        CMP_TRY(e);
        L->graph = g;
        CMP_TRY(hipGraphInstantiate(&L->gexec, g, nullptr, nullptr, 0));
        for (int q = 0; q < 6; ++q) L->gkey[q] = key[q];
        L->gcfl = cfl;
        L->glimiter = limiter;
    }
    for (int n = 0; n < nsteps; ++n) CMP_TRY(hipGraphLaunch(L->gexec, cs));
    if (cs != s) CMP_TRY(hipStreamSynchronize(cs));
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
