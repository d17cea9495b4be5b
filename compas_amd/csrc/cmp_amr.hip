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
hipError_t launch_cf_correct_dim(const LaunchCfg&, const KParams&, int, const CfCell*, int, const CfFace*, int,
                                 double*, const BoxDev*, const double*, const BoxDev*, const double*, int, double*,
                                 const double*, int, int*, hipStream_t);

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

struct cmp_amr {
    std::vector<cmp_level*> L;
    int ratio = 2;
    std::vector<CfCell*> d_cf;   // per coarse level (all but the finest)
    std::vector<int> ncf;
    std::vector<CfFace*> d_jobs; // the face jobs of each coarse level's list
    std::vector<int> njobs;
    std::vector<double*> d_faces;   // their results, N + 1 doubles a job
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

void amr_free(cmp_amr* A)
{
    for (CfCell* p : A->d_cf)
        if (p) (void)hipFree(p);
    for (CfFace* p : A->d_jobs)
        if (p) (void)hipFree(p);
    for (double* p : A->d_faces)
        if (p) (void)hipFree(p);
    if (A->d_step) (void)hipFree(A->d_step);
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
    int r = check_hip(hipMalloc(&A->d_step, sizeof(int) * 2));
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
    if (r) { amr_free(A); return r; }
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
    CMP_TRY(launch_amr_dt(A->adt, cmax, dt_dev, cfl, limiter, nullptr, A->d_step, A->L[0]->d_err,
                          (hipStream_t)stream));
    return CMP_OK;
}

// One stage's work graph: level 0's fill and stage kernel on the caller's stream; on a stream per
// finer level, its two-level fill (it reads the coarser stage state's valid cells, which that
// level's own fill leaves alone) and stage kernel, then the coarse-fine face jobs between it and the
// next coarser level (after that level's fill); back on the caller's stream, the corrections (after
// level 0's stage kernel and every face job) and the average-down (after every stage kernel),
// finest first.
int cmp_amr_rk3_step(cmp_amr* A, double* const* Un, double* const* U1, double* const* U2, const double* dt_dev,
                     double* cmax, void* stream)
{
    if (!A || !Un || !U1 || !U2 || !dt_dev || !cmax) return CMP_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const int nlev = (int)A->L.size();
    auto stage_of = [&](int l, int stage, double* const* src, double* const* dst, hipStream_t q) -> int {
        cmp_level* L = A->L[l];
        CMP_TRY(launch_stage(L->lc, L->kp, MODE_RK, stage, L->d_tiles, L->ntiles, L->d_boxes, nullptr, 0, src[l],
                             Un[l], dst[l], dt_dev, (unsigned long long*)(cmax + l), L->d_err, q));
        return CMP_OK;
    };
    auto cf = [&](int l, int part, int stage, double* const* src, double* const* dst, hipStream_t q) -> int {
        cmp_level* C = A->L[l];
        cmp_level* F = A->L[l + 1];
        CMP_TRY(launch_cf_correct(C->lc, C->kp, part, A->d_cf[l], A->ncf[l], A->d_jobs[l], A->njobs[l],
                                  A->d_faces[l], C->d_boxes, src[l], F->d_boxes, src[l + 1], A->ratio, dst[l], dt_dev,
                                  stage, C->d_err, q));
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
        for (int l = 0; l + 1 < nlev; ++l)
            if ((r = cf(l, 2, stage, src, dst, s))) return r;
        for (int l = nlev - 2; l >= 0; --l)
            if ((r = cmp_level_average_down(A->L[l + 1], dst[l + 1], A->L[l], dst[l], A->ratio, s))) return r;
    }
    return CMP_OK;
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
        for (int step = 0; step < nsteps; ++step)
            if ((r = amr_step(A, Un, U1, U2, cfl, limiter, dt_dev, cmax, dt_hist, s))) return r;
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
    std::vector<const void*> key;
    for (int l = 0; l < nlev; ++l) {
        key.push_back(Un[l]);
        key.push_back(U1[l]);
        key.push_back(U2[l]);
    }
    key.push_back(dt_dev);
    key.push_back(cmax);
    key.push_back(dt_hist);
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
    for (int step = 0; step < nsteps; ++step) CMP_TRY(hipGraphLaunch(A->gexec, cs));
    if (cs != s) CMP_TRY(hipStreamSynchronize(cs));
    return CMP_OK;
}

}  // extern "C"
