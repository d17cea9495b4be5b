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
#include <cstring>
#include <vector>

#include "cmp_kernels.hpp"
#include "cmp_level.hpp"

using namespace cmp;

namespace cmp {
template <int DIM>
hipError_t launch_cf_correct_dim(const LaunchCfg&, const KParams&, int, const CfCell*, int, const CfFace*, int,
                                 double*, const BoxDev*, const double*, const BoxDev*, const double*, int, double*,
                                 const double*, int, int*, hipStream_t);

template <int DIM>
hipError_t launch_cf_apply_all_dim(const LaunchCfg&, const CfPairDev*, int, int, const CfStatePtrs&, int, const double*,
                                   int, hipStream_t);

hipError_t launch_cf_apply_all(const LaunchCfg& lc, const CfPairDev* pairs, int npairs, int ntotal,
                               const CfStatePtrs& st, int ratio, const double* dt, int stage, hipStream_t s)
{
    return lc.dim == 2 ? launch_cf_apply_all_dim<2>(lc, pairs, npairs, ntotal, st, ratio, dt, stage, s)
                       : launch_cf_apply_all_dim<3>(lc, pairs, npairs, ntotal, st, ratio, dt, stage, s);
}

hipError_t launch_cf_correct(const LaunchCfg& lc, const KParams& kp, int part, const CfCell* list, int n,
                             const CfFace* jobs, int njobs, double* faces, const BoxDev* cboxes, const double* Uc,
                             const BoxDev* fboxes, const double* Uf, int ratio, double* Uout, const double* dt,
                             int stage, int* err, hipStream_t s)
{
    return lc.dim == 2 ? launch_cf_correct_dim<2>(lc, kp, part, list, n, jobs, njobs, faces, cboxes, Uc, fboxes, Uf,
                                                  ratio, Uout, dt, stage, err, s)
                       : launch_cf_correct_dim<3>(lc, kp, part, list, n, jobs, njobs, faces, cboxes, Uc, fboxes, Uf,
                                                  ratio, Uout, dt, stage, err, s);
}
}  // namespace cmp

namespace cmp {
template <int DIM>
hipError_t launch_cf_correct_pf_dim(const LaunchCfg&, const KParams&, int, const CfCell*, int, const CfFace*, int,
                                    double*, const BoxDev*, const double*, const double*, const BoxDev*,
                                    const double*, const double*, int, double*, const double*, int, int*,
                                    hipStream_t);
template <int DIM>
hipError_t launch_fill2_dim(int, const FillDev&, const FillDev*, const FillDev&, int, double*, const double*,
                            const double*, long long, int, int*, hipStream_t);
template <int DIM>
hipError_t launch_avgdown_dim(const BoxDev*, int, long long, const FillDev&, const double*, double*, int, int, double,
                              int*, hipStream_t);
}  // namespace cmp

// The Phase-Field step of a hierarchy: per level the phase-major fields of the mechanism (component
// k of a phase-major field is k times the level's one-component alpha-layout length further on) and
// the composite implicit solve's tables.
struct AmrPf {
    bool on = false;
    cmp_pf_params pf{};
    int npf = 1;
    std::vector<BoxDev*> d_pbox;          // phase-major alpha layout (npf > 1; else the level's d_abox)
    std::vector<FillDev> fdP;
    std::vector<double*> A0, A, Ps, H, S, IDL;   // IDL: the composite preconditioner's level diagonals
    std::vector<int*> d_mask;             // composite masks (implicit RC)
    CompCell* d_cells = nullptr;
    CompLink* d_links = nullptr;
    int ncells = 0, nlinks = 0;
    double vol_total = 0.0;
    long long ncomp = 0;                  // unknowns of the composite solve (cells no finer level covers)
    double* d_lsum = nullptr;             // per-level sums, 2 per level
    unsigned long long* d_bits = nullptr; // [0] max|u|, [1 + l] max|Grad_Q| of level l
    double* h_pin = nullptr;              // pinned host copies
    int cg_iters = 0;
    hipGraphExec_t cg_exec = nullptr;     // a chunk of the composite CG's iterations, captured once
};

struct cmp_amr {
    std::vector<cmp_level*> L;
    int ratio = 2;
    std::vector<CfCell*> d_cf;   // per coarse level (all but the finest)
    std::vector<int> ncf;
    std::vector<CfFace*> d_jobs; // the face jobs of each coarse level's list
    std::vector<int> njobs;
    std::vector<double*> d_faces;   // their results, N + 1 doubles a job
    // the flow corrections of every pair in one launch: the pairs with listed cells (device table),
    // each pair's coarse level, and the cells in all
    CfPairDev* d_cfpairs = nullptr;
    std::vector<int> cfpair_level;
    int ncftotal = 0;
    // each finer level's work of a stage runs on a stream of its own beside level 0's (fork / join):
    // q[l - 1] for level l; ev_fill[l] after level l's ghost fill, ev_done[l] after its last work
    std::vector<hipStream_t> q;
    hipEvent_t ev_fork = nullptr;
    std::vector<hipEvent_t> ev_fill, ev_done;
    AmrDt adt{};
    int* d_step = nullptr;               // step counter and dt-history capacity of an advance call
    // the replayed step of cmp_amr_advance, keyed by its buffers and settings
    hipStream_t own_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<const void*> gkey;
    double gcfl = 0.0;
    int glimiter = -1;
    AmrPf pf;
    // the flow stage of every level in one launch of the persistent work-list kernel (the levels'
    // tuned tile shapes; ghost cells read as filled by the level and two-level fills)
    bool use_list = false;
    int list_waves = 2, list_grid = 0, nltiles = 0;
    ListTile* d_ltiles = nullptr;
    ListLevel* d_llev = nullptr;
    int* d_lctr = nullptr;
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
int build_cf(const cmp_level* C, const cmp_level* F, const cmp_level* FF, int ratio, std::vector<CfCell>& out,
             std::vector<CfFace>& jobs)
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
                    const int job0 = (int)jobs.size();
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
                            // the coarse face, then the fine faces over it (lower transverse index fastest)
                            CfFace cj{b, 0, d, {i, j, k}};
                            cj.c[d] += side;
                            jobs.push_back(cj);
                            int g0[3] = {0, 0, 0};
                            for (int a = 0; a < dim; ++a) g0[a] = x[a] * ratio;
                            const int fpos = side ? x[d] * ratio : (x[d] + 1) * ratio;
                            int cov[3] = {g0[0], g0[1], g0[2]};
                            cov[d] = side ? fpos : fpos - 1;
                            const int fb = owner_of(F, cov);
                            if (fb < 0) return CMP_ERR_LAYOUT;
                            int t1 = -1, t2 = -1;
                            for (int a = 0; a < dim; ++a)
                                if (a != d) { if (t1 < 0) t1 = a; else t2 = a; }
                            const int n1 = (t1 >= 0) ? ratio : 1, n2 = (t2 >= 0) ? ratio : 1;
                            for (int o2 = 0; o2 < n2; ++o2)
                                for (int o1 = 0; o1 < n1; ++o1) {
                                    CfFace fj{fb, 1, d, {g0[0], g0[1], g0[2]}};
                                    fj.c[d] = fpos;
                                    if (t1 >= 0) fj.c[t1] += o1;
                                    if (t2 >= 0) fj.c[t2] += o2;
                                    jobs.push_back(fj);
                                }
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
                    if (mask) out.push_back(CfCell{b, {i, j, k}, mask, job0});
                }
    }
    return CMP_OK;
}

void amr_pf_free(cmp_amr* A)
{
    AmrPf& P = A->pf;
    for (size_t l = 0; l < P.d_pbox.size(); ++l)
        if (P.d_pbox[l] && P.d_pbox[l] != A->L[l]->d_abox) (void)hipFree(P.d_pbox[l]);
    for (auto* v : {&P.A0, &P.A, &P.Ps, &P.H, &P.S, &P.IDL})
        for (double* p : *v)
            if (p) (void)hipFree(p);
    for (int* p : P.d_mask)
        if (p) (void)hipFree(p);
    if (P.d_cells) (void)hipFree(P.d_cells);
    if (P.d_links) (void)hipFree(P.d_links);
    if (P.d_lsum) (void)hipFree(P.d_lsum);
    if (P.d_bits) (void)hipFree(P.d_bits);
    if (P.h_pin) (void)hipHostFree(P.h_pin);
    if (P.cg_exec) (void)hipGraphExecDestroy(P.cg_exec);
    P = AmrPf();
}

void amr_free(cmp_amr* A)
{
    amr_pf_free(A);
    for (CfCell* p : A->d_cf)
        if (p) (void)hipFree(p);
    for (CfFace* p : A->d_jobs)
        if (p) (void)hipFree(p);
    for (double* p : A->d_faces)
        if (p) (void)hipFree(p);
    if (A->d_step) (void)hipFree(A->d_step);
    if (A->d_cfpairs) (void)hipFree(A->d_cfpairs);
    void* lb[] = {A->d_ltiles, A->d_llev, A->d_lctr};
    for (void* p : lb)
        if (p) (void)hipFree(p);
    if (A->gexec) (void)hipGraphExecDestroy(A->gexec);
    if (A->graph) (void)hipGraphDestroy(A->graph);
    if (A->own_stream) (void)hipStreamDestroy(A->own_stream);
    for (hipStream_t x : A->q)
        if (x) (void)hipStreamDestroy(x);
    if (A->ev_fork) (void)hipEventDestroy(A->ev_fork);
    for (hipEvent_t e : A->ev_fill)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : A->ev_done)
        if (e) (void)hipEventDestroy(e);
    delete A;
}

// The hierarchy's work-list from the levels' cached tile sweeps (force: 1 always, -1 never, 0 when
// the summed times over every level's boxes favour it): the occupancy target of least summed time,
// every level's tiles with their tuned shapes in one list, one table entry per level.
int amr_tiles(cmp_amr* A, int force)
{
    const int nlev = (int)A->L.size();
    A->use_list = false;
    A->gkey.clear();   // a captured step is captured again
    if (force < 0 || nlev > LIST_MAXLEV) return force > 0 ? CMP_ERR_UNSUPPORTED : CMP_OK;
    const cmp_level* L0 = A->L[0];
    if (L0->lc.riemann != 1 && L0->lc.riemann != 2) return force > 0 ? CMP_ERR_UNSUPPORTED : CMP_OK;
    std::vector<std::vector<int>> ext[2];
    double t_list[2] = {0.0, 0.0}, t_fixed = 0.0;
    for (int l = 0; l < nlev; ++l) {
        std::vector<int> e[2];
        double tl[2], tf;
        const int r = level_tuned(A->L[l], e, tl, &tf);
        if (r) return (force > 0) ? r : CMP_OK;
        for (int w = 0; w < 2; ++w) {
            t_list[w] += tl[w];
            ext[w].push_back(std::move(e[w]));
        }
        t_fixed += tf;
    }
    const int w = (t_list[1] < t_list[0]) ? 1 : 0;
    if (force == 0 && !(t_list[w] < t_fixed)) return CMP_OK;
    std::vector<ListTile> ht;
    std::vector<ListLevel> lv(nlev);
    for (int l = 0; l < nlev; ++l) {
        cmp_level* L = A->L[l];
        // the level's own list choice is left as it was: its box_ext is borrowed for the build
        std::vector<int> keep = std::move(L->box_ext);
        L->box_ext = std::move(ext[w][l]);
        level_list_tiles(L, l, ht);
        L->box_ext = std::move(keep);
        lv[l] = ListLevel{L->kp, L->d_boxes, nullptr, L->d_err};
    }
    for (void* p : {(void*)A->d_ltiles, (void*)A->d_llev})
        if (p) CMP_TRY(hipFree(p));
    A->d_ltiles = nullptr;
    A->d_llev = nullptr;
    CMP_TRY(hipMalloc(&A->d_ltiles, sizeof(ListTile) * std::max<size_t>(1, ht.size())));
    CMP_TRY(hipMemcpy(A->d_ltiles, ht.data(), sizeof(ListTile) * ht.size(), hipMemcpyHostToDevice));
    CMP_TRY(hipMalloc(&A->d_llev, sizeof(ListLevel) * nlev));
    CMP_TRY(hipMemcpy(A->d_llev, lv.data(), sizeof(ListLevel) * nlev, hipMemcpyHostToDevice));
    if (!A->d_lctr) {
        CMP_TRY(hipMalloc(&A->d_lctr, sizeof(int) * LIST_NCTR));
        CMP_TRY(hipMemset(A->d_lctr, 0, sizeof(int) * LIST_NCTR));
    }
    A->list_waves = (w == 1) ? 3 : 2;
    A->list_grid = list_capacity(L0->lc, A->list_waves);
    if (A->list_grid <= 0) return CMP_ERR_UNSUPPORTED;
    A->nltiles = (int)ht.size();
    A->use_list = true;
    return CMP_OK;
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
    for (int l = 0; l < nlev; ++l) {
        const cmp_level* L = levels[l];
        if (!L || L->P.dim != dim || L->P.nstate != N || L->ng != L0->ng) return CMP_ERR_ARG;
        if (L->P.id_weno != L0->P.id_weno || L->P.id_limiter != L0->P.id_limiter ||
            L->P.id_riemann != L0->P.id_riemann || L->P.id_thinc != L0->P.id_thinc)
            return CMP_ERR_ARG;
        // the hierarchy step is the stage-coupled TVD-RK3 without a stop_time cut; a level set to
        // another integrator or given a stop_time would silently be stepped otherwise
        if (L->integrator != 3 || L->stop_time >= 0.0) return CMP_ERR_UNSUPPORTED;
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
        std::vector<CfFace> jobs;
        const int rc = build_cf(levels[l], levels[l + 1], (l + 2 < nlev) ? levels[l + 2] : nullptr, ratio, cf, jobs);
        if (rc) { amr_free(A); return rc; }
        A->d_cf.push_back(nullptr);
        A->d_jobs.push_back(nullptr);
        A->d_faces.push_back(nullptr);
        A->ncf.push_back((int)cf.size());
        A->njobs.push_back((int)jobs.size());
        if (!cf.empty()) {
            int r = check_hip(hipMalloc(&A->d_cf.back(), sizeof(CfCell) * cf.size()));
            if (!r) r = check_hip(hipMemcpy(A->d_cf.back(), cf.data(), sizeof(CfCell) * cf.size(), hipMemcpyHostToDevice));
            if (!r) r = check_hip(hipMalloc(&A->d_jobs.back(), sizeof(CfFace) * jobs.size()));
            if (!r) r = check_hip(hipMemcpy(A->d_jobs.back(), jobs.data(), sizeof(CfFace) * jobs.size(),
                                            hipMemcpyHostToDevice));
            if (!r) r = check_hip(hipMalloc(&A->d_faces.back(), sizeof(double) * (size_t)(N + 1) * jobs.size()));
            if (r) { amr_free(A); return r; }
        }
    }
    int r = CMP_OK;
    {
        std::vector<CfPairDev> pairs;
        for (int l = 0; l + 1 < nlev; ++l) {
            if (A->ncf[l] <= 0) continue;
            const cmp_level* C = levels[l];
            pairs.push_back(CfPairDev{C->kp, A->d_cf[l], A->ncf[l], A->ncftotal, C->d_boxes, A->d_faces[l], C->d_err});
            A->cfpair_level.push_back(l);
            A->ncftotal += A->ncf[l];
        }
        if (!pairs.empty() && pairs.size() <= (size_t)LIST_MAXLEV) {
            r = check_hip(hipMalloc(&A->d_cfpairs, sizeof(CfPairDev) * pairs.size()));
            if (!r) r = check_hip(hipMemcpy(A->d_cfpairs, pairs.data(), sizeof(CfPairDev) * pairs.size(),
                                            hipMemcpyHostToDevice));
        }
    }
    if (!r) r = check_hip(hipMalloc(&A->d_step, sizeof(int) * 2));
    if (!r) r = check_hip(hipMemset(A->d_step, 0, sizeof(int) * 2));
    if (!r && nlev > 1) {
        r = check_hip(hipEventCreateWithFlags(&A->ev_fork, hipEventDisableTiming));
        A->q.assign(nlev - 1, nullptr);
        A->ev_fill.assign(nlev, nullptr);
        A->ev_done.assign(nlev, nullptr);
        for (int l = 0; l < nlev && !r; ++l) {
            if (l > 0) r = check_hip(hipStreamCreateWithFlags(&A->q[l - 1], hipStreamNonBlocking));
            if (!r) r = check_hip(hipEventCreateWithFlags(&A->ev_fill[l], hipEventDisableTiming));
            if (!r) r = check_hip(hipEventCreateWithFlags(&A->ev_done[l], hipEventDisableTiming));
        }
    }
    if (!r) r = amr_tiles(A, 0);
    if (r) { amr_free(A); return r; }
    *out = A;
    return CMP_OK;
}

int cmp_amr_tiles(cmp_amr* A, int force, int* use_list, int* waves)
{
    if (!A || force < -1 || force > 1) return CMP_ERR_ARG;
    const int r = amr_tiles(A, force);
    if (r) return r;
    if (use_list) *use_list = A->use_list ? 1 : 0;
    if (waves) *waves = A->use_list ? A->list_waves : 0;
    return CMP_OK;
}

int cmp_amr_tile_info(const cmp_amr* A, int* use_list, int* waves)
{
    if (!A) return CMP_ERR_ARG;
    if (use_list) *use_list = A->use_list ? 1 : 0;
    if (waves) *waves = A->use_list ? A->list_waves : 0;
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
    CMP_TRY(launch_amr_dt(A->adt, cmax, dt_dev, cfl, limiter, nullptr, A->d_step, A->L[0]->d_err,
                          (hipStream_t)stream));
    return CMP_OK;
}

// One stage's work graph: level 0's fill and stage kernel on the caller's stream; on a stream per
// finer level, its two-level fill (it reads the coarser stage state's valid cells, which that
// level's own fill leaves alone) and stage kernel, then the coarse-fine face jobs between it and the
// next coarser level (after that level's fill); back on the caller's stream, the corrections (after
// level 0's stage kernel and every face job) and the average-down (after every stage kernel),
// finest first. With the work-list kernel (A->use_list) every level's stage is one launch on the
// caller's stream after all the fills, and the face jobs run beside it on the finer levels' streams.
// The stages of one TVD-RK3 step of the hierarchy: the flow's (pf false: c_max accumulated into
// cmax[l]) or the Phase-Field sub-step's (pf true: every level's right-hand side the Phase-Field
// flux at its Grad_Q, the levels' d_gq; no c_max).
static int amr_stages(cmp_amr* A, double* const* Un, double* const* U1, double* const* U2, const double* dt_dev,
                      double* cmax, bool pf, hipStream_t s)
{
    const int nlev = (int)A->L.size();
    auto stage_of = [&](int l, int stage, double* const* src, double* const* dst, hipStream_t q) -> int {
        cmp_level* L = A->L[l];
        if (pf)
            CMP_TRY(launch_stage(L->pf_lc, L->pf_kp, MODE_RK, stage, L->d_tiles, L->ntiles, L->d_boxes, nullptr, 0,
                                 src[l], Un[l], dst[l], dt_dev, nullptr, L->d_err, q, L->d_gq));
        else if (L->use_list)   // the level's tuned tiles (work-list kernel, or both kernels)
            return level_list_stage(L, MODE_RK, stage, src[l], Un[l], dst[l], dt_dev, cmax + l, q, false);
        else
            CMP_TRY(launch_stage(L->lc, L->kp, MODE_RK, stage, L->d_tiles, L->ntiles, L->d_boxes, nullptr, 0, src[l],
                                 Un[l], dst[l], dt_dev, (unsigned long long*)(cmax + l), L->d_err, q));
        return CMP_OK;
    };
    auto cf = [&](int l, int part, int stage, double* const* src, double* const* dst, hipStream_t q) -> int {
        cmp_level* C = A->L[l];
        cmp_level* F = A->L[l + 1];
        if (pf)
            CMP_TRY((C->P.dim == 2 ? launch_cf_correct_pf_dim<2> : launch_cf_correct_pf_dim<3>)(
                C->pf_lc, C->pf_kp, part, A->d_cf[l], A->ncf[l], A->d_jobs[l], A->njobs[l], A->d_faces[l], C->d_boxes,
                src[l], C->d_gq, F->d_boxes, src[l + 1], F->d_gq, A->ratio, dst[l], dt_dev, stage, C->d_err, q));
        else
            CMP_TRY(launch_cf_correct(C->lc, C->kp, part, A->d_cf[l], A->ncf[l], A->d_jobs[l], A->njobs[l],
                                      A->d_faces[l], C->d_boxes, src[l], F->d_boxes, src[l + 1], A->ratio, dst[l],
                                      dt_dev, stage, C->d_err, q));
        return CMP_OK;
    };
    // every pair's cell corrections on the caller's stream: the flow's in one launch
    auto cf_cells = [&](int stage, double* const* src, double* const* dst) -> int {
        if (pf || !A->d_cfpairs) {
            for (int l = 0; l + 1 < nlev; ++l) {
                const int r = cf(l, 2, stage, src, dst, s);
                if (r) return r;
            }
            return CMP_OK;
        }
        CfStatePtrs st{};
        for (size_t k = 0; k < A->cfpair_level.size(); ++k) {
            st.Uc[k] = src[A->cfpair_level[k]];
            st.Uout[k] = dst[A->cfpair_level[k]];
        }
        CMP_TRY(launch_cf_apply_all(A->L[0]->lc, A->d_cfpairs, (int)A->cfpair_level.size(), A->ncftotal, st, A->ratio,
                                    dt_dev, stage, s));
        return CMP_OK;
    };
    for (int stage = 0; stage < 3; ++stage) {
        double* const* src = (stage == 0) ? Un : ((stage == 1) ? U1 : U2);
        double* const* dst = (stage == 0) ? U1 : ((stage == 1) ? U2 : Un);
        int r;
        if (nlev == 1) {
            if ((r = cmp_level_fill_ghosts(A->L[0], src[0], s))) return r;
            if ((r = stage_of(0, stage, src, dst, s))) return r;
            continue;
        }
        CMP_TRY(hipEventRecord(A->ev_fork, s));
        for (int l = 1; l < nlev; ++l) CMP_TRY(hipStreamWaitEvent(A->q[l - 1], A->ev_fork, 0));
        if (!pf && A->use_list) {
            // every level's fill (the finer ones on their streams), then the face jobs on those
            // streams beside one work-list launch for every level's tiles on the caller's stream
            if ((r = cmp_level_fill_ghosts(A->L[0], src[0], s))) return r;
            CMP_TRY(hipEventRecord(A->ev_fill[0], s));
            for (int l = 1; l < nlev; ++l) {
                hipStream_t q = A->q[l - 1];
                if ((r = cmp_level_fill_two_levels(A->L[l], src[l], A->L[l - 1], src[l - 1], A->ratio, q))) return r;
                CMP_TRY(hipEventRecord(A->ev_fill[l], q));
            }
            for (int l = 1; l < nlev; ++l) {
                hipStream_t q = A->q[l - 1];
                CMP_TRY(hipStreamWaitEvent(q, A->ev_fill[l - 1], 0));
                if ((r = cf(l - 1, 1, stage, src, dst, q))) return r;
                CMP_TRY(hipEventRecord(A->ev_done[l], q));
                CMP_TRY(hipStreamWaitEvent(s, A->ev_fill[l], 0));
            }
            ListCall c{};
            for (int l = 0; l < nlev; ++l) {
                c.Us[l] = src[l];
                c.Un[l] = Un[l];
                c.Uout[l] = dst[l];
                c.cmax[l] = (unsigned long long*)(cmax + l);
            }
            CMP_TRY(launch_stage_list(A->L[0]->lc, A->list_waves, A->d_llev, A->d_ltiles, A->nltiles, A->d_lctr, c,
                                      MODE_RK, stage, dt_dev, A->list_grid, s));
            for (int l = 1; l < nlev; ++l) CMP_TRY(hipStreamWaitEvent(s, A->ev_done[l], 0));
            if ((r = cf_cells(stage, src, dst))) return r;
            for (int l = nlev - 2; l >= 0; --l)
                if ((r = cmp_level_average_down(A->L[l + 1], dst[l + 1], A->L[l], dst[l], A->ratio, s))) return r;
            continue;
        }
        // level 0
        if ((r = cmp_level_fill_ghosts(A->L[0], src[0], s))) return r;
        CMP_TRY(hipEventRecord(A->ev_fill[0], s));
        if ((r = stage_of(0, stage, src, dst, s))) return r;
        // each finer level: fill, stage kernel, then the face jobs with the next coarser level
        for (int l = 1; l < nlev; ++l) {
            hipStream_t q = A->q[l - 1];
            if ((r = cmp_level_fill_two_levels(A->L[l], src[l], A->L[l - 1], src[l - 1], A->ratio, q))) return r;
            CMP_TRY(hipEventRecord(A->ev_fill[l], q));
        }
        for (int l = 1; l < nlev; ++l) {
            hipStream_t q = A->q[l - 1];
            if ((r = stage_of(l, stage, src, dst, q))) return r;
            CMP_TRY(hipStreamWaitEvent(q, A->ev_fill[l - 1], 0));
            if ((r = cf(l - 1, 1, stage, src, dst, q))) return r;
            CMP_TRY(hipEventRecord(A->ev_done[l], q));
            CMP_TRY(hipStreamWaitEvent(s, A->ev_done[l], 0));
        }
        if ((r = cf_cells(stage, src, dst))) return r;
        for (int l = nlev - 2; l >= 0; --l)
            if ((r = cmp_level_average_down(A->L[l + 1], dst[l + 1], A->L[l], dst[l], A->ratio, s))) return r;
    }
    return CMP_OK;
}

int cmp_amr_rk3_step(cmp_amr* A, double* const* Un, double* const* U1, double* const* U2, const double* dt_dev,
                     double* cmax, void* stream)
{
    if (!A || !Un || !U1 || !U2 || !dt_dev || !cmax) return CMP_ERR_ARG;
    return amr_stages(A, Un, U1, U2, dt_dev, cmax, false, (hipStream_t)stream);
}

static int amr_step(cmp_amr* A, double* const* Un, double* const* U1, double* const* U2, double cfl, int limiter,
                    double* dt_dev, double* cmax, double* dt_hist, hipStream_t s)
{
    CMP_TRY(launch_amr_dt(A->adt, cmax, dt_dev, cfl, limiter, dt_hist, A->d_step, A->L[0]->d_err, s));
    return cmp_amr_rk3_step(A, Un, U1, U2, dt_dev, cmax, s);
}

int cmp_amr_advance(cmp_amr* A, double* const* Un, double* const* U1, double* const* U2, int nsteps, double cfl,
                    int limiter, double* dt_dev, double* cmax, double* dt_hist, int use_graph, void* stream)
{
    if (!A || !Un || !U1 || !U2 || nsteps < 0 || !dt_dev || !cmax) return CMP_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    int r = cmp_amr_initial_cmax(A, Un, cmax, stream);
    if (r) return r;
    CMP_TRY(hipMemsetAsync(dt_dev, 0, sizeof(double), s));
    CMP_TRY(launch_set2(A->d_step, 0, dt_hist ? nsteps : 0, s));
    if (!use_graph || nsteps < 2) {
        for (int step = 0; step < nsteps; ++step) {
            if ((r = amr_step(A, Un, U1, U2, cfl, limiter, dt_dev, cmax, dt_hist, s))) return r;
            if (A->pf.on && (r = cmp_amr_pf_evolve(A, Un, U1, U2, dt_dev, nullptr, s))) return r;
        }
        return CMP_OK;
    }
    // one step captured and replayed (on a private stream when the caller gave the legacy default
    // stream, which cannot be captured)
    hipStream_t cs = s;
    if (!cs) {
        if (!A->own_stream) CMP_TRY(hipStreamCreateWithFlags(&A->own_stream, hipStreamNonBlocking));
        cs = A->own_stream;
        CMP_TRY(hipStreamSynchronize(nullptr));
    }
    const int nlev = (int)A->L.size();
    // the finer levels' two-level fill and average-down plans (they allocate, so before the capture)
    for (int l = 1; l < nlev; ++l)
        if ((r = level_fill2_plan(A->L[l], A->L[l - 1], A->ratio, cs)) || (r = level_avg_plan(A->L[l], A->L[l - 1], A->ratio, cs)))
            return r;
    std::vector<const void*> key;
    for (int l = 0; l < nlev; ++l) {
        key.push_back(Un[l]);
        key.push_back(U1[l]);
        key.push_back(U2[l]);
    }
    key.push_back(dt_dev);
    key.push_back(cmax);
    key.push_back(dt_hist);
    // the levels' tile choices (cmp_level_set_tiles, cmp_level_tune_tiles) and fill plans are baked
    // into the graph
    for (int l = 0; l < nlev; ++l) {
        key.push_back((const void*)(uintptr_t)A->L[l]->tiles_gen);
        key.push_back(A->L[l]->d_f2plan);
        key.push_back((const void*)(uintptr_t)A->L[l]->f2_uid);
        key.push_back(A->L[l]->d_avgplan);
        key.push_back((const void*)(uintptr_t)A->L[l]->avg_uid);
    }
    if (!(A->gexec && A->gkey == key && A->gcfl == cfl && A->glimiter == limiter)) {
        if (A->gexec) { CMP_TRY(hipGraphExecDestroy(A->gexec)); A->gexec = nullptr; }
        if (A->graph) { CMP_TRY(hipGraphDestroy(A->graph)); A->graph = nullptr; }
        CMP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        r = amr_step(A, Un, U1, U2, cfl, limiter, dt_dev, cmax, dt_hist, cs);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(cs, &g);
        if (r) { if (g) (void)hipGraphDestroy(g); return r; }
        CMP_TRY(e);
        A->graph = g;
        CMP_TRY(hipGraphInstantiate(&A->gexec, g, nullptr, nullptr, 0));
        A->gkey = key;
        A->gcfl = cfl;
        A->glimiter = limiter;
    }
    for (int step = 0; step < nsteps; ++step) {
        CMP_TRY(hipGraphLaunch(A->gexec, cs));
        if (A->pf.on && (r = cmp_amr_pf_evolve(A, Un, U1, U2, dt_dev, nullptr, cs))) return r;
    }
    if (cs != s) CMP_TRY(hipStreamSynchronize(cs));
    return CMP_OK;
}


}  // extern "C"

// ---- the Phase-Field step of a hierarchy -----------------------------------------------------------
// PhaseField_Evolve over finest_level + 1 levels without subcycling (source/Class_Compressible_PhaseField.cpp:
// 425-594): MaxVelocity over the levels, eta from the finest cell size, the mechanism sub-steps of
// PhaseFieldMechanism_Evolve (:597-709) on every level with the finest level's dt_m (the sub-step
// state filled on every level, all phases at once; CAC's sums over the averaged-down level 0; ACDI's
// J averaged down before its divergence and before Jhat takes it), Grad_Q per level (explicit RC) or
// from the composite implicit solve, then PhaseField_ComputeDt over the levels and stage-coupled
// Phase-Field RK3 sub-steps with AverageDown.

namespace {

// the alpha layout's box offsets of a level (1 ghost, 1 comp): nbox + 1 prefix sums of the fab sizes
std::vector<long long> alpha_offsets(const cmp_level* L)
{
    std::vector<long long> off(L->nbox + 1, 0);
    for (int q = 0; q < L->nbox; ++q) {
        long long f = 1;
        for (int d = 0; d < L->P.dim; ++d) f *= L->boxes[q].hi[d] - L->boxes[q].lo[d] + 3;
        off[q + 1] = off[q] + f;
    }
    return off;
}

// valid cells (host): the index of cell c of box b of a level in its alpha layout (1 ghost, 1 comp);
// aoff: that level's alpha_offsets
long long aidx(const cmp_level* L, const std::vector<long long>& aoff, int b, const int* c)
{
    const cmp_box& B = L->boxes[b];
    const int dim = L->P.dim;
    const long long off = aoff[b];
    long long idx = 0, st = 1;
    for (int d = 0; d < dim; ++d) {
        idx += (long long)(c[d] - B.lo[d] + 1) * st;
        st *= B.hi[d] - B.lo[d] + 3;
    }
    return off + idx;
}

// offset of box b's direction-d block in a level's one-component face buffer, and the face index of
// the face at the low side of cell c along d inside it
long long aface(const cmp_level* L, const std::vector<long long>& foff, int b, int d, const int* c)
{
    const cmp_box& B = L->boxes[b];
    const int dim = L->P.dim;
    long long idx = 0, st = 1;
    for (int e = 0; e < dim; ++e) {
        idx += (long long)(c[e] - B.lo[e]) * st;
        st *= B.hi[e] - B.lo[e] + 1 + (e == d ? 1 : 0);
    }
    return foff[(size_t)b * dim + d] + idx;
}

// periodic wrap of c into level L's domain; false when c lies beyond a non-periodic face
bool wrap_cell(const cmp_level* L, int* c)
{
    for (int d = 0; d < L->P.dim; ++d) {
        const int lo = L->G.domain_lo[d], hi = L->G.domain_hi[d], n = hi - lo + 1;
        if (c[d] < lo || c[d] > hi) {
            if (!L->G.is_periodic[d]) return false;
            c[d] = (c[d] < lo) ? c[d] + n : c[d] - n;
        }
    }
    return true;
}

int floor_div_h(int a, int r) { return (a >= 0) ? a / r : -((-a + r - 1) / r); }

// The composite masks of every level and the link lists of the cells with a coarse-fine face.
int amr_comp_build(cmp_amr* A)
{
    AmrPf& P = A->pf;
    const int nlev = (int)A->L.size();
    const int ratio = A->ratio;
    const int dim = A->L[0]->P.dim;
    std::vector<CompCell> cells;
    std::vector<CompLink> links;
    std::vector<std::vector<long long>> foff(nlev), aoff(nlev);
    for (int l = 0; l < nlev; ++l) {
        const cmp_level* L = A->L[l];
        aoff[l] = alpha_offsets(L);
        foff[l].resize((size_t)L->nbox * dim);
        cmp_face_layout(dim, L->nbox, L->boxes.data(), 1, foff[l].data(), nullptr);
    }
    auto geo_of = [&](int l, int d) {
        double a = 1.0;
        for (int e = 0; e < dim; ++e)
            if (e != d) a *= A->L[l]->G.dx[e];
        return a;
    };
    double vol_total = 0.0;
    long long ncomp = 0;
    P.d_mask.assign(nlev, nullptr);
    std::vector<CompLink> lk;   // one cell's links (its storage kept from cell to cell)
    lk.reserve(64);
    for (int l = 0; l < nlev; ++l) {
        const cmp_level* L = A->L[l];
        std::vector<int> mask((size_t)L->na, COMP_REGULAR);
        double vol = 1.0;
        for (int d = 0; d < dim; ++d) vol *= L->G.dx[d];
        for (int b = 0; b < L->nbox; ++b) {
            const cmp_box& B = L->boxes[b];
            const int klo = (dim == 3) ? B.lo[2] : 0, khi = (dim == 3) ? B.hi[2] : 0;
            for (int k = klo; k <= khi; ++k)
                for (int j = B.lo[1]; j <= B.hi[1]; ++j)
                    for (int i = B.lo[0]; i <= B.hi[0]; ++i) {
                        const int c[3] = {i, j, k};
                        const long long ci = aidx(L, aoff[l], b, c);
                        int f[3] = {0, 0, 0};
                        for (int d = 0; d < dim; ++d) f[d] = c[d] * ratio;
                        if (l + 1 < nlev && owner_of(A->L[l + 1], f) >= 0) {
                            mask[ci] = COMP_COVERED;
                            continue;
                        }
                        vol_total += vol;
                        ++ncomp;
                        lk.clear();
                        bool special = false;
                        for (int d = 0; d < dim; ++d)
                            for (int side = 0; side < 2; ++side) {
                                int x[3] = {i, j, k};
                                x[d] += side ? 1 : -1;
                                if (!wrap_cell(L, x)) continue;   // Neumann domain face
                                int fc[3] = {i, j, k};
                                fc[d] += side;
                                const int nb = owner_of(L, x);
                                int fx[3] = {0, 0, 0};
                                for (int e = 0; e < dim; ++e) fx[e] = x[e] * ratio;
                                const bool xcov = nb >= 0 && l + 1 < nlev && owner_of(A->L[l + 1], fx) >= 0;
                                if (nb >= 0 && !xcov) {
                                    lk.push_back(CompLink{l, l, aidx(L, aoff[l], nb, x), aface(L, foff[l], b, d, fc),
                                                          geo_of(l, d) / L->G.dx[d]});
                                    continue;
                                }
                                special = true;
                                if (nb >= 0) {
                                    // the fine cells next to this face, each with its fine face
                                    const cmp_level* F = A->L[l + 1];
                                    const double dist = 0.5 * (F->G.dx[d] + L->G.dx[d]);
                                    int t1 = -1, t2 = -1;
                                    for (int e = 0; e < dim; ++e)
                                        if (e != d) { if (t1 < 0) t1 = e; else t2 = e; }
                                    const int r1 = (t1 >= 0) ? ratio : 1, r2 = (t2 >= 0) ? ratio : 1;
                                    for (int o2 = 0; o2 < r2; ++o2)
                                        for (int o1 = 0; o1 < r1; ++o1) {
                                            int g[3] = {0, 0, 0};
                                            for (int e = 0; e < dim; ++e) g[e] = x[e] * ratio;
                                            g[d] = side ? x[d] * ratio : (x[d] + 1) * ratio - 1;
                                            if (t1 >= 0) g[t1] = c[t1] * ratio + o1;
                                            if (t2 >= 0) g[t2] = c[t2] * ratio + o2;
                                            int gf[3] = {g[0], g[1], g[2]};
                                            gf[d] += side ? 0 : 1;
                                            const int fb = owner_of(F, g);
                                            if (fb < 0) return CMP_ERR_LAYOUT;
                                            lk.push_back(CompLink{l + 1, l + 1, aidx(F, aoff[l + 1], fb, g), aface(F, foff[l + 1], fb, d, gf),
                                                                  geo_of(l + 1, d) / dist});
                                        }
                                } else {
                                    // x is not on this level: the coarse cell holding it
                                    if (l == 0) return CMP_ERR_LAYOUT;
                                    const cmp_level* C = A->L[l - 1];
                                    int p[3] = {0, 0, 0};
                                    for (int e = 0; e < dim; ++e) p[e] = floor_div_h(x[e], ratio);
                                    if (!wrap_cell(C, p)) return CMP_ERR_LAYOUT;
                                    const int pb = owner_of(C, p);
                                    if (pb < 0) return CMP_ERR_LAYOUT;
                                    lk.push_back(CompLink{l - 1, l, aidx(C, aoff[l - 1], pb, p), aface(L, foff[l], b, d, fc),
                                                          geo_of(l, d) / (0.5 * (L->G.dx[d] + C->G.dx[d]))});
                                }
                            }
                        if (special) {
                            mask[ci] = (int)cells.size();
                            cells.push_back(CompCell{(int)lk.size(), (int)links.size()});
                            links.insert(links.end(), lk.begin(), lk.end());
                        }
                    }
        }
        CMP_TRY(hipMalloc(&P.d_mask[l], sizeof(int) * mask.size()));
        CMP_TRY(hipMemcpy(P.d_mask[l], mask.data(), sizeof(int) * mask.size(), hipMemcpyHostToDevice));
    }
    P.ncells = (int)cells.size();
    P.nlinks = (int)links.size();
    CMP_TRY(hipMalloc(&P.d_cells, sizeof(CompCell) * std::max<size_t>(1, cells.size())));
    CMP_TRY(hipMalloc(&P.d_links, sizeof(CompLink) * std::max<size_t>(1, links.size())));
    if (!cells.empty())
        CMP_TRY(hipMemcpy(P.d_cells, cells.data(), sizeof(CompCell) * cells.size(), hipMemcpyHostToDevice));
    if (!links.empty())
        CMP_TRY(hipMemcpy(P.d_links, links.data(), sizeof(CompLink) * links.size(), hipMemcpyHostToDevice));
    P.vol_total = vol_total;
    P.ncomp = ncomp;
    return CMP_OK;
}

CompArgs comp_args(const cmp_amr* A, bool solution)
{
    const AmrPf& P = A->pf;
    CompArgs ca{};
    ca.nlev = (int)A->L.size();
    ca.ratio = A->ratio;
    ca.cells = P.d_cells;
    ca.links = P.d_links;
    for (int l = 0; l < ca.nlev; ++l) {
        const cmp_level* L = A->L[l];
        CompLev& c = ca.L[l];
        c.abox = L->d_abox;
        c.cmap = L->pg.cmap;
        c.ncmap = L->pg.ncmap;
        c.fd = L->fdA;
        c.mask = P.d_mask[l];
        c.v = solution ? L->d_X : L->d_Pv;
        c.bf = L->d_Bf;
        c.s = (l < (int)P.S.size()) ? P.S[l] : nullptr;
        c.idl = (l < (int)P.IDL.size()) ? P.IDL[l] : nullptr;
        c.vol = 1.0;
        for (int d = 0; d < 3; ++d) {
            c.dx[d] = L->pg.dx[d];
            if (d < L->P.dim) c.vol *= L->G.dx[d];
        }
        for (int d = 0; d < 3; ++d) {
            double a = 1.0;
            for (int e = 0; e < L->P.dim; ++e)
                if (e != d) a *= L->G.dx[e];
            c.area[d] = a;
        }
    }
    return ca;
}

int ap_launch(int dim, int op, const AmrPfArgs& a, hipStream_t s)
{
    CMP_TRY(dim == 2 ? launch_amrpf_dim<2>(op, a, s) : launch_amrpf_dim<3>(op, a, s));
    return CMP_OK;
}

AmrPfArgs level_args(const cmp_level* L, int l, const CompArgs* ca)
{
    AmrPfArgs a{};
    a.nbox = L->nbox;
    a.maxcells = L->maxcells;
    for (int d = 0; d < 3; ++d) a.maxfaces[d] = L->maxfaces[d];
    a.ca = ca;
    a.lev = l;
    a.part = L->d_part;
    return a;
}

// AverageDownFaces of one-component face buffers F[l] (finest level first)
int amr_avgdown_faces(cmp_amr* A, double* const* F, hipStream_t s)
{
    const int nlev = (int)A->L.size();
    for (int l = nlev - 2; l >= 0; --l) {
        const cmp_level* Fl = A->L[l + 1];
        const cmp_level* C = A->L[l];
        AmrPfArgs a{};
        a.fboxes = Fl->d_abox;
        a.cfd = C->fdA;
        a.nbox = Fl->nbox;
        // coarse faces of a coarsened fine box: at most (n/ratio + 1) along d
        for (int d = 0; d < Fl->P.dim; ++d) {
            long long m = 0;
            for (const cmp_box& b : Fl->boxes) {
                long long v = 1;
                for (int e = 0; e < Fl->P.dim; ++e) v *= (b.hi[e] - b.lo[e] + 1) / A->ratio + (e == d ? 1 : 0);
                m = std::max(m, v);
            }
            a.maxfaces[d] = m;
        }
        a.ratio = A->ratio;
        a.src = F[l + 1];
        a.dst = F[l];
        int r = ap_launch(C->P.dim, AP_AVGDOWN_FACES, a, s);
        if (r) return r;
    }
    return CMP_OK;
}

// AverageDown of one-component alpha-layout cell fields X[l] (finest level first)
int amr_avgdown_cells(cmp_amr* A, double* const* X, hipStream_t s)
{
    const int nlev = (int)A->L.size();
    for (int l = nlev - 2; l >= 0; --l) {
        cmp_level* F = A->L[l + 1];
        cmp_level* C = A->L[l];
        const int dim = F->P.dim;
        long long maxc = 0;
        for (const cmp_box& b : F->boxes) {
            long long v = 1;
            for (int d = 0; d < dim; ++d) v *= (b.hi[d] - b.lo[d] + 1) / A->ratio;
            maxc = std::max(maxc, v);
        }
        double vol = 1.0;
        for (int d = 0; d < dim; ++d) vol *= A->ratio;
        CMP_TRY((dim == 2 ? launch_avgdown_dim<2> : launch_avgdown_dim<3>)(F->d_abox, F->nbox, maxc, C->fdA, X[l + 1],
                                                                            X[l], 1, A->ratio, 1.0 / vol, C->d_err, s));
    }
    return CMP_OK;
}

// FillPatch of an nc-component Phase-Field field of every level (1 ghost): level 0 same-level, finer
// levels two-level. F[l]: the field; fd[l]: its tables (phase-major alpha layout or the normals').
int amr_fill_pf(cmp_amr* A, int nc, double* const* F, const std::vector<FillDev>& fd, hipStream_t s)
{
    const int nlev = (int)A->L.size();
    cmp_level* L0 = A->L[0];
    CMP_TRY(launch_fill(L0->P.dim, nc, fd[0], L0->nghost1, F[0], L0->d_err, s));
    for (int l = 1; l < nlev; ++l) {
        cmp_level* L = A->L[l];
        CMP_TRY((L->P.dim == 2 ? launch_fill2_dim<2> : launch_fill2_dim<3>)(nc, fd[l], &fd[l], fd[l - 1], A->ratio,
                                                                            F[l], F[l], F[l - 1], 0, 0, L->d_err, s));
    }
    return CMP_OK;
}

// The composite implicit solve of one phase: rhs in every level's d_DJ, b in d_Bf (averaged down);
// Grad_Q into out[l] (one-component face buffers), averaged down.
int amr_comp_solve(cmp_amr* A, double* const* out, hipStream_t s0)
{
    // s: the stream every launch below goes to (the caller's; the capture stream while a chunk of
    // CG iterations is captured)
    hipStream_t s = s0;
    AmrPf& P = A->pf;
    const int nlev = (int)A->L.size();
    cmp_level* L0 = A->L[0];
    const int dim = L0->P.dim;
    double* sc = L0->d_sc;
    const CompArgs caP = comp_args(A, false), caX = comp_args(A, true);
    const bool mg = !L0->mg.empty();
    // with the level-0 V-cycle the preconditioned residual is projected off the constant vector over
    // the composite cells (the operator's null space)
    const double inv_n = (mg && P.ncomp > 0) ? 1.0 / (double)P.ncomp : 0.0;
    int r;
    // the level sums of nv partial slots from slot slot0 on, and their sum over the levels (one launch)
    auto sum_levels = [&](int nv, double* dst, const double* scx, int slot0 = 0) -> int {
        LevelParts lp{};
        lp.nlev = nlev;
        for (int l = 0; l < nlev; ++l) {
            const cmp_level* L = A->L[l];
            lp.part[l] = L->d_part + (long long)slot0 * L->nblocks;
            lp.nblocks[l] = L->nblocks;
        }
        CMP_TRY(launch_reduce_levels(lp, nv, P.d_lsum, dst, s, scx));
        return CMP_OK;
    };
    // z = R ID (levels above 0, Jacobi) + the level-0 V-cycle of the restricted R/V at every cell's
    // level-0 ancestor; partials of r.z, sum z and sum r into each level's second to fourth slots
    auto precondition = [&](const double* scx) -> int {
        for (int l = 0; l < nlev; ++l) {
            AmrPfArgs a = level_args(A->L[l], l, &caP);
            a.sc = scx;
            a.src = A->L[l]->d_R;
            a.dst = P.S[l];
            if ((r = ap_launch(dim, AP_COMP_S, a, s))) return r;
        }
        if ((r = amr_avgdown_cells(A, P.S.data(), s))) return r;
        const double* x0 = nullptr;
        if (mg) {
            cmp::MgArgs m{};
            m.g = L0->pg;
            m.maxcells = L0->maxcells;
            m.sc = scx;
            m.src = P.S[0];
            m.dst = L0->mg[0].r;
            m.G = L0->mg[0].G;
            CMP_TRY(dim == 2 ? launch_mg_dim<2>(MG_GATHER_R, m, s) : launch_mg_dim<3>(MG_GATHER_R, m, s));
            if ((r = mg_vcycle(L0, scx, 0, s))) return r;
            x0 = L0->mg[0].x[L0->mg[0].cur];
        }
        for (int l = 0; l < nlev; ++l) {
            cmp_level* L = A->L[l];
            AmrPfArgs a = level_args(L, l, &caP);
            a.sc = scx;
            a.src = L->d_R;
            a.src2 = L->d_ID;
            a.src3 = x0;
            if (mg) a.G0 = L0->mg[0].G;
            for (int d = 0; d < 3; ++d) a.dlo0[d] = (d < dim) ? L0->G.domain_lo[d] : 0;
            a.dst = L->d_Z;
            if ((r = ap_launch(dim, AP_COMP_Z, a, s))) return r;
            cmp::MgArgs m{};
            m.g = L->pg;
            m.maxcells = L->maxcells;
            m.sc = scx;
            m.src = L->d_R;
            m.r = L->d_Z;
            m.part = L->d_part;
            CMP_TRY(dim == 2 ? launch_mg_dim<2>(MG_CG_RZ, m, s) : launch_mg_dim<3>(MG_CG_RZ, m, s));
        }
        return CMP_OK;
    };
    // every level's own operator diagonal for the preconditioner's intermediate levels
    for (int l = 1; l + 1 < nlev; ++l) {
        cmp_level* L = A->L[l];
        CMP_TRY(launch_cg_idiag(dim, L->pg, L->maxcells, L->d_Bf, P.IDL[l], s));
    }
    // Jacobi diagonal; the rhs mean over the composite cells; r, x
    for (int l = 0; l < nlev; ++l) {
        AmrPfArgs a = level_args(A->L[l], l, &caP);
        a.dst = A->L[l]->d_ID;
        if ((r = ap_launch(dim, AP_COMP_DIAG, a, s))) return r;
        a.src = A->L[l]->d_DJ;
        a.mean_sum = nullptr;
        if ((r = ap_launch(dim, AP_COMP_INIT, a, s))) return r;
    }
    if ((r = sum_levels(1, sc + CG_SUM, nullptr))) return r;
    for (int l = 0; l < nlev; ++l) {
        AmrPfArgs a = level_args(A->L[l], l, &caP);
        a.src = A->L[l]->d_DJ;
        a.mean_sum = sc + CG_SUM;
        a.a = P.vol_total;
        a.dst = A->L[l]->d_X;
        a.dst2 = A->L[l]->d_R;
        if ((r = ap_launch(dim, AP_COMP_INIT, a, s))) return r;
    }
    if (mg && (r = mg_prepare(L0, s))) return r;
    if ((r = precondition(nullptr))) return r;
    if ((r = sum_levels(2, sc + CG_RR, nullptr))) return r;
    if ((r = sum_levels(2, sc + CG_ZS, nullptr, 2))) return r;
    CMP_TRY(launch_cg_start(sc, s, inv_n));
    for (int l = 0; l < nlev; ++l) {
        cmp_level* L = A->L[l];
        CMP_TRY(launch_cg_p(dim, L->pg, L->maxcells, sc, L->d_Z, L->d_Pv, s, inv_n, P.d_mask[l], 1));
    }
    const int maxit = 100000;
    const int chunk = 8;
    // a chunk of iterations depends only on the hierarchy's buffers: captured once into a graph and
    // replayed (one launch instead of about a hundred per iteration)
    auto enqueue_chunk = [&]() -> int {
        for (int q = 0; q < chunk; ++q) {
            // the search direction's ghost cells for the regular stencil (a finer level's coarse-fine
            // ghosts are interpolated but never read: those cells take their link lists)
            {
                std::vector<double*> Pv(nlev);
                std::vector<FillDev> fda(nlev);
                for (int l = 0; l < nlev; ++l) {
                    Pv[l] = A->L[l]->d_Pv;
                    fda[l] = A->L[l]->fdA;
                }
                if ((r = amr_fill_pf(A, 1, Pv.data(), fda, s))) return r;
            }
            for (int l = 0; l < nlev; ++l) {
                AmrPfArgs a = level_args(A->L[l], l, &caP);
                a.sc = sc;
                a.dst = A->L[l]->d_AP;
                if ((r = ap_launch(dim, AP_COMP_AP, a, s))) return r;
            }
            if ((r = sum_levels(1, sc + CG_PAP, sc))) return r;
            for (int l = 0; l < nlev; ++l) {
                cmp_level* L = A->L[l];
                CMP_TRY(launch_cg_xr(dim, L->pg, L->maxcells, sc, nullptr, L->d_X, L->d_R, L->d_Z, L->d_Pv, L->d_AP,
                                     L->d_part, s));
            }
            if ((r = precondition(sc))) return r;
            if ((r = sum_levels(2, sc + CG_RRNEW, sc))) return r;
            if (mg && (r = sum_levels(2, sc + CG_ZS, sc, 2))) return r;
            for (int l = 0; l < nlev; ++l) {
                cmp_level* L = A->L[l];
                CMP_TRY(launch_cg_p(dim, L->pg, L->maxcells, sc, L->d_Z, L->d_Pv, s, inv_n, P.d_mask[l]));
            }
            CMP_TRY(launch_cg_end(sc, maxit, s, inv_n));
        }
        return CMP_OK;
    };
    hipStream_t cs = s0;
    if (!cs) {
        if (!A->own_stream) CMP_TRY(hipStreamCreateWithFlags(&A->own_stream, hipStreamNonBlocking));
        cs = A->own_stream;
        CMP_TRY(hipStreamSynchronize(nullptr));
    }
    if (!P.cg_exec) {
        CMP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        s = cs;
        const int rc = enqueue_chunk();
        s = s0;
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(cs, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        CMP_TRY(e);
        const hipError_t ei = hipGraphInstantiate(&P.cg_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        CMP_TRY(ei);
    }
    for (long long queued = 0;; queued += chunk) {
        CMP_TRY(hipGraphLaunch(P.cg_exec, cs));
        CMP_TRY(hipMemcpyAsync(P.h_pin + 0, sc + CG_DONE, sizeof(double), hipMemcpyDeviceToHost, cs));
        CMP_TRY(hipMemcpyAsync(P.h_pin + 1, sc + CG_IT, sizeof(double), hipMemcpyDeviceToHost, cs));
        CMP_TRY(hipStreamSynchronize(cs));
        if (P.h_pin[0] != 0.0) break;
        if (queued > maxit) return CMP_ERR_NAN;
    }
    P.cg_iters = (int)P.h_pin[1];
    // covered cells take the mean of their children; Grad_Q of every level, averaged down
    std::vector<double*> X(nlev);
    for (int l = 0; l < nlev; ++l) X[l] = A->L[l]->d_X;
    if ((r = amr_avgdown_cells(A, X.data(), s))) return r;
    for (int l = 0; l < nlev; ++l) {
        AmrPfArgs a = level_args(A->L[l], l, &caX);
        a.dst = out[l];
        if ((r = ap_launch(dim, AP_COMP_GRAD, a, s))) return r;
    }
    return amr_avgdown_faces(A, out, s);
}

}  // namespace

extern "C" {

int cmp_amr_pf_setup(cmp_amr* A, const cmp_pf_params* pf)
{
    if (!A || !pf) return CMP_ERR_ARG;
    amr_pf_free(A);
    const int nlev = (int)A->L.size();
    for (int l = 0; l < nlev; ++l) {
        int r = cmp_level_pf_setup(A->L[l], pf);
        if (r) return r;
    }
    AmrPf& P = A->pf;
    P.pf = *pf;
    const int nph = A->L[0]->P.nphase;
    P.npf = nph ? nph : 1;
    const bool implicit = (pf->explicit_rc == 0);
    int r = CMP_OK;
#define CMP_TRYA(x) do { r = cmp::check_hip(x); if (r) { amr_pf_free(A); return r; } } while (0)
    for (int l = 0; l < nlev; ++l) {
        cmp_level* L = A->L[l];
        BoxDev* pb = L->d_abox;
        if (P.npf > 1) {
            // the phase-major table: the alpha layout with the component stride of the whole level buffer
            std::vector<BoxDev> hb(L->nbox);
            CMP_TRYA(hipMemcpy(hb.data(), L->d_abox, sizeof(BoxDev) * L->nbox, hipMemcpyDeviceToHost));
            for (BoxDev& b : hb) b.sn = L->na;
            CMP_TRYA(hipMalloc(&pb, sizeof(BoxDev) * L->nbox));
            CMP_TRYA(hipMemcpy(pb, hb.data(), sizeof(BoxDev) * L->nbox, hipMemcpyHostToDevice));
        }
        P.d_pbox.push_back(pb);
        FillDev fd = L->fdA;
        fd.boxes = pb;
        P.fdP.push_back(fd);
        double *a0 = nullptr, *a = nullptr, *ps = nullptr, *h = nullptr, *sv = nullptr, *idl = nullptr;
        CMP_TRYA(hipMalloc(&a0, sizeof(double) * L->na * P.npf));
        P.A0.push_back(a0);
        CMP_TRYA(hipMalloc(&a, sizeof(double) * L->na * P.npf));
        P.A.push_back(a);
        CMP_TRYA(hipMalloc(&ps, sizeof(double) * L->na * P.npf));
        P.Ps.push_back(ps);
        CMP_TRYA(hipMalloc(&h, sizeof(double) * L->nfaces * P.npf));
        P.H.push_back(h);
        if (implicit) {
            CMP_TRYA(hipMalloc(&sv, sizeof(double) * L->na));
            CMP_TRYA(hipMemset(sv, 0, sizeof(double) * L->na));
            CMP_TRYA(hipMalloc(&idl, sizeof(double) * L->na));
            CMP_TRYA(hipMemset(idl, 0, sizeof(double) * L->na));
        }
        P.S.push_back(sv);
        P.IDL.push_back(idl);
        // the unfilled ghost cells stay defined
        CMP_TRYA(hipMemset(a0, 0, sizeof(double) * L->na * P.npf));
        CMP_TRYA(hipMemset(a, 0, sizeof(double) * L->na * P.npf));
        CMP_TRYA(hipMemset(ps, 0, sizeof(double) * L->na * P.npf));
    }
    CMP_TRYA(hipMalloc(&P.d_lsum, sizeof(double) * 2 * AMR_MAX_LEVELS));
    CMP_TRYA(hipMalloc(&P.d_bits, sizeof(unsigned long long) * (1 + AMR_MAX_LEVELS)));
    CMP_TRYA(hipHostMalloc((void**)&P.h_pin, sizeof(double) * (4 + AMR_MAX_LEVELS), hipHostMallocDefault));
#undef CMP_TRYA
    if (implicit && (r = amr_comp_build(A))) { amr_pf_free(A); return r; }
    P.on = true;
    return CMP_OK;
}

int cmp_amr_pf_solver_iterations(const cmp_amr* A, int* iterations)
{
    if (!A || !iterations) return CMP_ERR_ARG;
    *iterations = A->pf.cg_iters;
    return CMP_OK;
}

// Grad_Q of every level into the levels' own face buffers (and the caller's, when given).
int cmp_amr_pf_gradq(cmp_amr* A, double* const* U, double dt, double* const* gradq, double* gradq_max, double* info,
                     void* stream)
{
    if (!A || !U || !A->pf.on) return CMP_ERR_ARG;
    if (!(dt > 0.0)) return CMP_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    AmrPf& P = A->pf;
    const cmp_pf_params& pf = P.pf;
    const int nlev = (int)A->L.size();
    cmp_level* L0 = A->L[0];
    cmp_level* Lf = A->L[nlev - 1];
    const int dim = L0->P.dim;
    const int nph = L0->P.nphase;
    const int npf = P.npf;
    const bool cac = (pf.mechanism == CMP_PF_CAC);
    const bool explicit_rc = (pf.explicit_rc == 1);
    int r;
    CMP_TRY(launch_zero_u64(P.d_bits, 1 + nlev, s));
    // MaxVelocity over the levels
    for (int l = 0; l < nlev; ++l) {
        cmp_level* L = A->L[l];
        CMP_TRY(launch_pf_maxvel(dim, L->pg, L->maxcells, U[l], P.d_bits, nph ? nph : 2, s));
    }
    CMP_TRY(hipMemcpyAsync(P.h_pin, P.d_bits, sizeof(double), hipMemcpyDeviceToHost, s));
    CMP_TRY(hipStreamSynchronize(s));
    double maxv;
    std::memcpy(&maxv, P.h_pin, sizeof(double));
    if (__builtin_isnan(maxv)) return CMP_ERR_NAN;
    double lambda = pf.lambda;
    if (lambda <= 0.0) lambda = 1.0;
    double eta = pf.eta;
    if (eta <= 0.0) {
        double m = Lf->G.dx[0];
        for (int d = 1; d < dim; ++d) m = (m < Lf->G.dx[d]) ? Lf->G.dx[d] : m;
        eta = pf.eta_multiplier * m;
    }
    double mob = 1.0;
    mob *= 0.5 * eta * maxv;
    double sdx = 1.0 / Lf->G.dx[0] / Lf->G.dx[0];
    for (int d = 1; d < dim; ++d) sdx = sdx + 1.0 / Lf->G.dx[d] / Lf->G.dx[d];
    double dt_m = 0.9 / (2.0 * mob * lambda) / (sdx + 1.0 / eta / eta);
    const double mob_face = 0.5 * mob + 0.5 * mob;
    // the step's volume fractions (PFBorder) and the mechanism's copy
    for (int l = 0; l < nlev; ++l) {
        cmp_level* L = A->L[l];
        for (int ph = 0; ph < npf; ++ph)
            CMP_TRY(launch_pf_extract(dim, L->pg, L->maxcells, U[l], P.A0[l] + ph * L->na, nullptr,
                                      nph ? nph + dim + 1 + ph : 3 + dim, s));
    }
    if ((r = amr_fill_pf(A, npf, P.A0.data(), P.fdP, s))) return r;
    for (int l = 0; l < nlev; ++l) {
        cmp_level* L = A->L[l];
        CMP_TRY(hipMemcpyAsync(P.A[l], P.A0[l], sizeof(double) * L->na * npf, hipMemcpyDeviceToDevice, s));
        if (explicit_rc)
            CMP_TRY(hipMemsetAsync(P.H[l], 0, sizeof(double) * L->nfaces * npf, s));
    }
    std::vector<double*> Jl(nlev), Nl(nlev), Mk(nlev), Wk(nlev);
    std::vector<FillDev> fdN(nlev);
    for (int l = 0; l < nlev; ++l) {
        Jl[l] = A->L[l]->d_J;
        Nl[l] = A->L[l]->d_N;
        Mk[l] = A->L[l]->d_Mk;
        Wk[l] = A->L[l]->d_Wk;
        fdN[l] = A->L[l]->fdN;
    }
    double t = 0.0;
    const double T = dt;
    long long n = 0;
    while (t < T) {
        if (t + dt_m > T) dt_m = T - t;
        if (!(dt_m > 0.0) || t + dt_m == t || n >= 10000000) return CMP_ERR_NAN;
        if ((r = amr_fill_pf(A, npf, P.A.data(), P.fdP, s))) return r;
        if (cac) {
            for (int ph = 0; ph < npf; ++ph) {
                for (int l = 0; l < nlev; ++l) {
                    cmp_level* L = A->L[l];
                    CMP_TRY(launch_pf_normals(dim, L->pg, L->maxcells, P.A[l] + ph * L->na, L->d_N, s));
                }
                if ((r = amr_fill_pf(A, dim, Nl.data(), fdN, s))) return r;
                for (int l = 0; l < nlev; ++l) {
                    cmp_level* L = A->L[l];
                    CMP_TRY(launch_pf_cac_mech(dim, L->pg, L->maxcells, P.A[l] + ph * L->na, L->d_N, L->d_Mk,
                                               L->d_Wk, mob, lambda, eta, L->d_part, s));
                }
                if ((r = amr_avgdown_cells(A, Mk.data(), s))) return r;
                if ((r = amr_avgdown_cells(A, Wk.data(), s))) return r;
                CMP_TRY(launch_cell_sum(dim, L0->pg, L0->maxcells, L0->d_Mk, L0->d_part, s));
                CMP_TRY(launch_cell_sum(dim, L0->pg, L0->maxcells, L0->d_Wk, L0->d_part + L0->nblocks, s));
                CMP_TRY(launch_reduce_sums(L0->d_part, L0->nblocks, 2, L0->d_sc + CG_SUM, s));
                for (int l = 0; l < nlev; ++l) {
                    cmp_level* L = A->L[l];
                    CMP_TRY(launch_pf_cac_update(dim, L->pg, L->maxcells, P.A[l] + ph * L->na, L->d_Mk, L->d_Wk,
                                                 L0->d_sc + CG_SUM, dt_m, s));
                }
            }
        } else {
            for (int l = 0; l < nlev; ++l) {
                cmp_level* L = A->L[l];
                CMP_TRY(launch_pf_map(P.A[l], P.Ps[l], L->na * npf, eta, eta, 0, s));
            }
            if ((r = amr_fill_pf(A, npf, P.Ps.data(), P.fdP, s))) return r;
            for (int ph = 0; ph < npf; ++ph) {
                for (int l = 0; l < nlev; ++l) {
                    cmp_level* L = A->L[l];
                    CMP_TRY(launch_pf_normals(dim, L->pg, L->maxcells, P.Ps[l] + ph * L->na, L->d_N, s));
                }
                if ((r = amr_fill_pf(A, dim, Nl.data(), fdN, s))) return r;
                for (int l = 0; l < nlev; ++l) {
                    cmp_level* L = A->L[l];
                    // the face psi from alpha of the two cells (FDM_PhaseFieldFlux_ACDI_D maps alpha)
                    CMP_TRY(launch_pf_map(P.A[l] + ph * L->na, L->d_P, L->na, eta, eta, 0, s));
                    CMP_TRY(launch_pf_jfaces(dim, L->pg, L->maxfaces, P.A[l] + ph * L->na, L->d_P, L->d_N, L->d_J,
                                             nullptr, mob_face, eta, dt_m, s));
                }
                if ((r = amr_avgdown_faces(A, Jl.data(), s))) return r;
                for (int l = 0; l < nlev; ++l) {
                    cmp_level* L = A->L[l];
                    if (explicit_rc) {
                        AmrPfArgs a{};
                        a.n = L->nfaces;
                        a.a = dt_m;
                        a.src = L->d_J;
                        a.dst = P.H[l] + ph * L->nfaces;
                        if ((r = ap_launch(dim, AP_AXPY, a, s))) return r;
                    }
                    CMP_TRY(launch_pf_update(dim, L->pg, L->maxcells, P.A[l] + ph * L->na, L->d_J, dt_m, s));
                }
            }
        }
        t += dt_m;
        ++n;
    }
    // Grad_Q of every level
    for (int ph = 0; ph < npf; ++ph) {
        std::vector<double*> out(nlev);
        for (int l = 0; l < nlev; ++l) out[l] = (npf == 1) ? A->L[l]->d_gq : A->L[l]->d_gq1;
        if (explicit_rc) {
            for (int l = 0; l < nlev; ++l) {
                cmp_level* L = A->L[l];
                CMP_TRY(launch_pf_map(P.A0[l] + ph * L->na, L->d_P, L->na, eta, pf.eta_factor * eta, 1, s));
                CMP_TRY(launch_pf_gradq(dim, L->pg, L->maxfaces, L->d_P, P.H[l] + ph * L->nfaces, 1.0 / dt, out[l],
                                        L->d_bits + 1, nullptr, s));
            }
        } else {
            std::vector<double*> Bf(nlev);
            for (int l = 0; l < nlev; ++l) {
                cmp_level* L = A->L[l];
                Bf[l] = L->d_Bf;
                CMP_TRY(launch_pf_divjhat(dim, L->pg, L->maxcells, P.A[l] + ph * L->na, P.A0[l] + ph * L->na, L->d_DJ,
                                          dt, s));
                CMP_TRY(launch_pf_map(P.A0[l] + ph * L->na, L->d_P, L->na, eta, pf.eta_factor * eta, 1, s));
                CMP_TRY(launch_pf_bcoef(dim, L->pg, L->maxfaces, L->d_P, L->d_Bf, s));
            }
            if ((r = amr_avgdown_faces(A, Bf.data(), s))) return r;
            if ((r = amr_comp_solve(A, out.data(), s))) return r;
        }
        if (npf > 1)
            for (int l = 0; l < nlev; ++l) {
                cmp_level* L = A->L[l];
                CMP_TRY(launch_pf_scatter(dim, L->pg, L->maxfaces, out[l], L->d_gq, ph, s));
            }
    }
    // max |Grad_Q| of every level, every box flagged for the sub-steps
    for (int l = 0; l < nlev; ++l) {
        cmp_level* L = A->L[l];
        AmrPfArgs a{};
        a.n = L->nfaces * npf;
        a.src = L->d_gq;
        a.bits = P.d_bits + 1 + l;
        if ((r = ap_launch(dim, AP_MAXABS, a, s))) return r;
        if (gradq && gradq[l])
            CMP_TRY(hipMemcpyAsync(gradq[l], L->d_gq, sizeof(double) * L->nfaces * npf, hipMemcpyDeviceToDevice, s));
    }
    CMP_TRY(hipMemcpyAsync(P.h_pin + 4, P.d_bits + 1, sizeof(double) * nlev, hipMemcpyDeviceToHost, s));
    CMP_TRY(hipStreamSynchronize(s));
    if (gradq_max)
        for (int l = 0; l < nlev; ++l) std::memcpy(&gradq_max[l], P.h_pin + 4 + l, sizeof(double));
    if (info) {
        info[0] = maxv;
        info[1] = mob;
        info[2] = eta;
        info[3] = (double)n;
        info[4] = explicit_rc ? 0.0 : (double)P.cg_iters;
    }
    return CMP_OK;
}

int cmp_amr_pf_evolve(cmp_amr* A, double* const* Un, double* const* U1, double* const* U2, const double* dt_dev,
                      double* info, void* stream)
{
    if (!A || !Un || !U1 || !U2 || !dt_dev || !A->pf.on) return CMP_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    AmrPf& P = A->pf;
    const int nlev = (int)A->L.size();
    cmp_level* L0 = A->L[0];
    CMP_TRY(hipMemcpyAsync(P.h_pin + 2, dt_dev, sizeof(double), hipMemcpyDeviceToHost, s));
    CMP_TRY(hipStreamSynchronize(s));
    const double dt = P.h_pin[2];
    if (dt == 0.0) {
        if (info)
            for (int q = 0; q < 7; ++q) info[q] = 0.0;
        return CMP_OK;
    }
    double gm[AMR_MAX_LEVELS], inf5[5];
    int r = cmp_amr_pf_gradq(A, Un, dt, nullptr, gm, inf5, stream);
    if (r) return r;
    // sub-steps: PhaseField_ComputeDt over the levels (nsubsteps 1, run.do_subcycle = 0, :2037-2098)
    const int dim = L0->P.dim;
    double prev[AMR_MAX_LEVELS] = {0.0};
    double pf_time = 0.0, gmax = 0.0;
    for (int l = 0; l < nlev; ++l) gmax = (gm[l] > gmax) ? gm[l] : gmax;
    int n = 0;
    for (int step = 0; step < P.pf.max_step && pf_time < dt; ++step) {
        double tmp[AMR_MAX_LEVELS];
        for (int l = 0; l < nlev; ++l) {
            double dxm = A->L[l]->G.dx[0];
            for (int d = 1; d < dim; ++d) dxm = (A->L[l]->G.dx[d] < dxm) ? A->L[l]->G.dx[d] : dxm;
            tmp[l] = P.pf.cfl * dxm / (gm[l] + 1.0e-14);
            if (pf_time == 0.0) prev[l] = tmp[l];
        }
        double dt0 = tmp[0];
        for (int l = 0; l < nlev; ++l) {
            if (P.pf.change_limiter && prev[l] > 0.0) tmp[l] = (1.1 * prev[l] < tmp[l]) ? 1.1 * prev[l] : tmp[l];
            dt0 = (tmp[l] < dt0) ? tmp[l] : dt0;
        }
        const double eps = 1.0e-3 * dt0;
        if (pf_time + dt0 > dt - eps) dt0 = dt - pf_time;
        double dmin = dt0;
        for (int l = 0; l < nlev; ++l) dmin = (tmp[l] < dmin) ? tmp[l] : dmin;
        for (int l = 0; l < nlev; ++l) prev[l] = dmin;
        CMP_TRY(launch_setd(L0->d_pfdt, dmin, s));
        if ((r = amr_stages(A, Un, U1, U2, L0->d_pfdt, nullptr, true, s))) return r;
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
    return CMP_OK;
}

}  // extern "C"
