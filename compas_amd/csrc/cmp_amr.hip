// cmp_amr.hip - a hierarchy of levels advanced together, stage by stage, without subcycling: the
// reference's AdvanceAllLevelsStageCoupled (source/Class_Compressible_PhaseField_StageCoupling.cpp)
// with run.do_reflux = 1, followed by AverageDown (timeStepNoSubcycling, source/COMPAS.cpp:2139-2180).
//
// One RK stage of the hierarchy, on one stream:
//   1. ghost cells of every level's stage state: level 0 by the same-level fill, finer levels by the
//      two-level fill from the next coarser level's stage state;
//   2. the stage kernel of every level (ghosts read from the buffers);
//   3. per coarse level, the coarse cells next to the finer level (listed here at creation, with
//      which of their faces border it) take the change of their right-hand side when those faces
//      take the mean of the fine faces over them;
//   4. the levels averaged down, finest first.
// The step is the reference's to rounding: the reference replaces the coarse faces before its
// divergence, this library adds the difference after the stage kernel's update.
#include <algorithm>
#include <vector>

#include "cmp_kernels.hpp"
#include "cmp_level.hpp"

using namespace cmp;

namespace cmp {
template <int DIM>
hipError_t launch_cf_correct_dim(const LaunchCfg&, const KParams&, const CfCell*, int, const BoxDev*, const double*,
                                 const FillDev&, const FillDev&, const double*, int, double*, const double*, int,
                                 int*, hipStream_t);

hipError_t launch_cf_correct(const LaunchCfg& lc, const KParams& kp, const CfCell* list, int n,
                             const BoxDev* cboxes, const double* Uc, const FillDev& cfd, const FillDev& ffd,
                             const double* Uf, int ratio, double* Uout, const double* dt, int stage, int* err,
                             hipStream_t s)
{
    return lc.dim == 2 ? launch_cf_correct_dim<2>(lc, kp, list, n, cboxes, Uc, cfd, ffd, Uf, ratio, Uout, dt, stage,
                                                  err, s)
                       : launch_cf_correct_dim<3>(lc, kp, list, n, cboxes, Uc, cfd, ffd, Uf, ratio, Uout, dt, stage,
                                                  err, s);
}
}  // namespace cmp

struct cmp_amr {
    std::vector<cmp_level*> L;
    int ratio = 2;
    std::vector<CfCell*> d_cf;   // per coarse level (all but the finest)
    std::vector<int> ncf;
    AmrDt adt{};
};

namespace {

// The box of level L owning cell c (-1: none, or outside the domain).
int owner_of(const cmp_level* L, const int* c)
{
    const int dim = L->P.dim;
    long long q = 0;
    for (int d = dim - 1; d >= 0; --d) {
        if (c[d] < L->G.domain_lo[d] || c[d] > L->G.domain_hi[d]) return -1;
        q = q * L->nb[d] + (c[d] - L->G.domain_lo[d]) / L->bs;
    }
    return L->h_owner[q];
}

// The coarse cells of level C next to level F (its refinement by ratio), with their bordering faces.
int build_cf(const cmp_level* C, const cmp_level* F, const cmp_level* FF, int ratio, std::vector<CfCell>& out)
{
    const int dim = C->P.dim;
    auto covered = [&](const int* c) -> bool {
        int f[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) f[d] = c[d] * ratio;
        return owner_of(F, f) >= 0;
    };
    for (int b = 0; b < C->nbox; ++b) {
        const cmp_box& B = C->boxes[b];
        const int klo = (dim == 3) ? B.lo[2] : 0, khi = (dim == 3) ? B.hi[2] : 0;
        for (int k = klo; k <= khi; ++k)
            for (int j = B.lo[1]; j <= B.hi[1]; ++j)
                for (int i = B.lo[0]; i <= B.hi[0]; ++i) {
                    const int c[3] = {i, j, k};
                    if (covered(c)) continue;
                    int mask = 0;
                    for (int d = 0; d < dim; ++d)
                        for (int side = 0; side < 2; ++side) {
                            int x[3] = {i, j, k};
                            x[d] += side ? 1 : -1;
                            const int dlo = C->G.domain_lo[d], dhi = C->G.domain_hi[d];
                            if (x[d] < dlo || x[d] > dhi) {
                                if (!C->G.is_periodic[d]) continue;
                                x[d] += (x[d] < dlo) ? dhi - dlo + 1 : -(dhi - dlo + 1);
                            }
                            if (!covered(x)) continue;
                            mask |= 1 << (2 * d + side);
                            // the fine faces used must be the finer level's own (not under a third
                            // level, whose face means would replace them first)
                            if (FF) {
                                int g[3] = {0, 0, 0};
                                for (int a = 0; a < dim; ++a) g[a] = x[a] * ratio;
                                g[d] = side ? x[d] * ratio : (x[d] + 1) * ratio - 1;
                                const int r1 = ratio, r2 = (dim == 3) ? ratio : 1;
                                for (int o2 = 0; o2 < r2; ++o2)
                                    for (int o1 = 0; o1 < r1; ++o1) {
                                        int h[3] = {g[0], g[1], g[2]};
                                        int t = 0;
                                        for (int a = 0; a < dim; ++a) {
                                            if (a == d) continue;
                                            h[a] += (t++ == 0) ? o1 : o2;
                                        }
                                        int hh[3] = {0, 0, 0};
                                        for (int a = 0; a < dim; ++a) hh[a] = h[a] * ratio;
                                        if (owner_of(FF, hh) >= 0) return CMP_ERR_LAYOUT;
                                    }
                            }
                        }
                    if (mask) out.push_back(CfCell{b, {i, j, k}, mask});
                }
    }
    return CMP_OK;
}

void amr_free(cmp_amr* A)
{
    for (CfCell* p : A->d_cf)
        if (p) (void)hipFree(p);
    delete A;
}

}  // namespace

extern "C" {

int cmp_amr_create(cmp_level* const* levels, int nlev, int ratio, cmp_amr** out)
{
    if (!levels || !out || nlev < 1 || nlev > AMR_MAX_LEVELS || ratio < 2) return CMP_ERR_ARG;
    *out = nullptr;
    const cmp_level* L0 = levels[0];
    if (!L0) return CMP_ERR_ARG;
    const int dim = L0->P.dim;
    const int N = L0->P.nstate;
    if (L0->P.nphase) return CMP_ERR_UNSUPPORTED;
    for (int l = 0; l < nlev; ++l) {
        const cmp_level* L = levels[l];
        if (!L || L->P.dim != dim || L->P.nstate != N || L->ng != L0->ng) return CMP_ERR_ARG;
        if (L->P.id_weno != L0->P.id_weno || L->P.id_limiter != L0->P.id_limiter ||
            L->P.id_riemann != L0->P.id_riemann || L->P.id_thinc != L0->P.id_thinc)
            return CMP_ERR_ARG;
        if (l == 0) continue;
        const cmp_level* C = levels[l - 1];
        for (int d = 0; d < dim; ++d) {
            if (L->G.domain_lo[d] != C->G.domain_lo[d] * ratio ||
                L->G.domain_hi[d] + 1 != (C->G.domain_hi[d] + 1) * ratio || C->G.domain_lo[d] < 0 ||
                L->G.is_periodic[d] != C->G.is_periodic[d] || L->G.lo_bc[d] != C->G.lo_bc[d] ||
                L->G.hi_bc[d] != C->G.hi_bc[d])
                return CMP_ERR_LAYOUT;
            for (const cmp_box& b : L->boxes)
                if (b.lo[d] % ratio != 0 || (b.hi[d] + 1) % ratio != 0) return CMP_ERR_LAYOUT;
        }
    }
    cmp_amr* A = new cmp_amr;
    A->L.assign(levels, levels + nlev);
    A->ratio = ratio;
    A->adt.nlev = nlev;
    A->adt.dim = dim;
    for (int l = 0; l < nlev; ++l)
        for (int d = 0; d < 3; ++d) A->adt.dx[l][d] = levels[l]->G.dx[d];
    for (int l = 0; l + 1 < nlev; ++l) {
        std::vector<CfCell> cf;
        const int rc = build_cf(levels[l], levels[l + 1], (l + 2 < nlev) ? levels[l + 2] : nullptr, ratio, cf);
        if (rc) { amr_free(A); return rc; }
        CfCell* d = nullptr;
        if (!cf.empty()) {
            int r = check_hip(hipMalloc(&d, sizeof(CfCell) * cf.size()));
            if (!r) r = check_hip(hipMemcpy(d, cf.data(), sizeof(CfCell) * cf.size(), hipMemcpyHostToDevice));
            if (r) { if (d) (void)hipFree(d); amr_free(A); return r; }
        }
        A->d_cf.push_back(d);
        A->ncf.push_back((int)cf.size());
    }
    *out = A;
    return CMP_OK;
}

int cmp_amr_destroy(cmp_amr* A)
{
    if (!A) return CMP_ERR_ARG;
    amr_free(A);
    return CMP_OK;
}

int cmp_amr_info(const cmp_amr* A, int level, int* ncf_cells)
{
    if (!A || !ncf_cells || level < 0 || level >= (int)A->L.size()) return CMP_ERR_ARG;
    *ncf_cells = (level + 1 < (int)A->L.size()) ? A->ncf[level] : 0;
    return CMP_OK;
}

// Every ghost cell of every level's buffer (level 0 same-level, finer levels two-level).
static int amr_fill(cmp_amr* A, double* const* U, hipStream_t s)
{
    int r = cmp_level_fill_ghosts(A->L[0], U[0], s);
    if (r) return r;
    for (size_t l = 1; l < A->L.size(); ++l) {
        r = cmp_level_fill_two_levels(A->L[l], U[l], A->L[l - 1], U[l - 1], A->ratio, s);
        if (r) return r;
    }
    return CMP_OK;
}

int cmp_amr_initial_cmax(cmp_amr* A, double* const* U, double* cmax, void* stream)
{
    if (!A || !U || !cmax) return CMP_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const int nlev = (int)A->L.size();
    CMP_TRY(hipMemsetAsync(cmax, 0, sizeof(double) * nlev, s));
    int r = amr_fill(A, U, s);
    if (r) return r;
    for (int l = 0; l < nlev; ++l) {
        cmp_level* L = A->L[l];
        CMP_TRY(launch_stage(L->lc, L->kp, MODE_CMAX, 0, L->d_tiles, L->ntiles, L->d_boxes, nullptr, 0, U[l],
                             nullptr, nullptr, nullptr, (unsigned long long*)(cmax + l), L->d_err, s));
    }
    return CMP_OK;
}

int cmp_amr_compute_dt(cmp_amr* A, double* cmax, double* dt_dev, double cfl, int limiter, void* stream)
{
    if (!A || !cmax || !dt_dev) return CMP_ERR_ARG;
    CMP_TRY(launch_amr_dt(A->adt, cmax, dt_dev, cfl, limiter, nullptr, 0, A->L[0]->d_err, (hipStream_t)stream));
    return CMP_OK;
}

int cmp_amr_rk3_step(cmp_amr* A, double* const* Un, double* const* U1, double* const* U2, const double* dt_dev,
                     double* cmax, void* stream)
{
    if (!A || !Un || !U1 || !U2 || !dt_dev || !cmax) return CMP_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const int nlev = (int)A->L.size();
    for (int stage = 0; stage < 3; ++stage) {
        double* const* src = (stage == 0) ? Un : ((stage == 1) ? U1 : U2);
        double* const* dst = (stage == 0) ? U1 : ((stage == 1) ? U2 : Un);
        int r = amr_fill(A, src, s);
        if (r) return r;
        for (int l = 0; l < nlev; ++l) {
            cmp_level* L = A->L[l];
            CMP_TRY(launch_stage(L->lc, L->kp, MODE_RK, stage, L->d_tiles, L->ntiles, L->d_boxes, nullptr, 0,
                                 src[l], Un[l], dst[l], dt_dev, (unsigned long long*)(cmax + l), L->d_err, s));
        }
        for (int l = 0; l + 1 < nlev; ++l) {
            cmp_level* C = A->L[l];
            cmp_level* F = A->L[l + 1];
            CMP_TRY(launch_cf_correct(C->lc, C->kp, A->d_cf[l], A->ncf[l], C->d_boxes, src[l], C->fd, F->fd,
                                      src[l + 1], A->ratio, dst[l], dt_dev, stage, C->d_err, s));
        }
        for (int l = nlev - 2; l >= 0; --l) {
            r = cmp_level_average_down(A->L[l + 1], dst[l + 1], A->L[l], dst[l], A->ratio, s);
            if (r) return r;
        }
    }
    return CMP_OK;
}

int cmp_amr_advance(cmp_amr* A, double* const* Un, double* const* U1, double* const* U2, int nsteps, double cfl,
                    int limiter, double* dt_dev, double* cmax, double* dt_hist, void* stream)
{
    if (!A || nsteps < 0 || !dt_dev || !cmax) return CMP_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    int r = cmp_amr_initial_cmax(A, Un, cmax, stream);
    if (r) return r;
    CMP_TRY(hipMemsetAsync(dt_dev, 0, sizeof(double), s));
    for (int step = 0; step < nsteps; ++step) {
        CMP_TRY(launch_amr_dt(A->adt, cmax, dt_dev, cfl, limiter, dt_hist, step, A->L[0]->d_err, s));
        r = cmp_amr_rk3_step(A, Un, U1, U2, dt_dev, cmax, stream);
        if (r) return r;
    }
    return CMP_OK;
}

}  // extern "C"
