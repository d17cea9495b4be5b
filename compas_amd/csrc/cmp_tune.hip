// cmp_tune.hip - the persistent work-list form of the flow stage and its tile autotuner.
//
// A level (or every level of a hierarchy, cmp_amr.hip) can run its flow stage as k_stage_list: a
// fixed grid of workgroups walks one list of tiles, each box cut into tiles of its own shape. The
// shape of a box is chosen by a measured sweep, cached per (box shape, scheme): a synthetic level of
// many copies of the box (its cells sampled from a real level's state) is stepped once per candidate
// shape with the work-list kernel, and once with the fixed-shape kernel k_stage, each timed with
// HIP events; the candidate of least time per cell wins. A level then uses the work-list kernel when
// the sum over its boxes of the best candidates' times is below the fixed kernel's.
#include <algorithm>
#include <array>
#include <cstdio>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "cmp_level.hpp"

namespace cmp {

template <int DIM>
hipError_t launch_stage_list_dim(const LaunchCfg&, int, int, int*, const ListLevel*, const ListTile*, int, int*,
                                 const ListCall&, int, int, const double*, int, hipStream_t);
template <int DIM>
bool tile_shape_ok_dim(const LaunchCfg&, const int*);

hipError_t launch_stage_list(const LaunchCfg& lc, int waves, const ListLevel* levs, const ListTile* tiles, int ntiles,
                             int* ctr, const ListCall& call, int mode, int stage, const double* dt, int grid,
                             hipStream_t s)
{
    return lc.dim == 2 ? launch_stage_list_dim<2>(lc, waves, 0, nullptr, levs, tiles, ntiles, ctr, call, mode, stage,
                                                  dt, grid, s)
                       : launch_stage_list_dim<3>(lc, waves, 0, nullptr, levs, tiles, ntiles, ctr, call, mode, stage,
                                                  dt, grid, s);
}

int list_capacity(const LaunchCfg& lc, int waves)
{
    int cap = 0;
    const ListCall c{};
    const hipError_t e = lc.dim == 2
                             ? launch_stage_list_dim<2>(lc, waves, 1, &cap, nullptr, nullptr, 0, nullptr, c, 0, 0,
                                                        nullptr, 0, nullptr)
                             : launch_stage_list_dim<3>(lc, waves, 1, &cap, nullptr, nullptr, 0, nullptr, c, 0, 0,
                                                        nullptr, 0, nullptr);
    return e == hipSuccess ? cap : 0;
}

bool tile_shape_ok(const LaunchCfg& lc, const int* ext)
{
    return lc.dim == 2 ? tile_shape_ok_dim<2>(lc, ext) : tile_shape_ok_dim<3>(lc, ext);
}

// The tiles of box b with extents ext (clipped to the box), lowest corner first along x.
static void box_tiles(const cmp_level* L, int lev, int b, const int* ext, std::vector<ListTile>& out)
{
    const int dim = L->P.dim;
    const cmp_box& B = L->boxes[b];
    int e[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) e[d] = std::min(ext[d], B.hi[d] - B.lo[d] + 1);
    const int klo = (dim == 3) ? B.lo[2] : 0, khi = (dim == 3) ? B.hi[2] : 0;
    for (int k = klo; k <= khi; k += e[2])
        for (int j = B.lo[1]; j <= B.hi[1]; j += e[1])
            for (int i = B.lo[0]; i <= B.hi[0]; i += e[0]) {
                ListTile t{};
                t.lev = lev;
                t.box = b;
                t.lo[0] = i;
                t.lo[1] = j;
                t.lo[2] = k;
                for (int d = 0; d < 3; ++d) t.ext[d] = e[d];
                int thi[3] = {0, 0, 0};
                for (int d = 0; d < dim; ++d) thi[d] = std::min(t.lo[d] + e[d] - 1, B.hi[d]);
                for (int d = 0; d < 3; ++d) t.halo[d][0] = t.halo[d][1] = HALO_LOOKUP;
                tile_halo(dim, &L->G, L->boxes.data(), b, t.lo, thi, L->ng, L->h_owner, L->nb, L->bs, t.halo);
                out.push_back(t);
            }
}

// The tiles of every box of L with the extents of box_ext (3 ints per box), level index lev; a box
// whose extents are 0 0 0 is left to the fixed-shape kernel and has no tiles here.
void level_list_tiles(const cmp_level* L, int lev, std::vector<ListTile>& out)
{
    for (int b = 0; b < L->nbox; ++b)
        if (L->box_ext[3 * (size_t)b] > 0) box_tiles(L, lev, b, &L->box_ext[3 * (size_t)b], out);
}

// Builds L's work-list from L->box_ext: tiles, one-level table (fused ghost reads), counters, and the
// fixed-shape kernel's tiles of the boxes it keeps.
static int level_build_list(cmp_level* L)
{
    ++L->tiles_gen;
    std::vector<ListTile> ht;
    level_list_tiles(L, 0, ht);
    std::vector<TileDev> ft;
    for (int b = 0; b < L->nbox; ++b)
        if (L->box_ext[3 * (size_t)b] == 0)
            for (int t = L->tile_start[b]; t < L->tile_start[b + 1]; ++t) ft.push_back(L->h_tiles[t]);
    if (ht.empty()) return CMP_ERR_ARG;
    if (L->d_ltiles) { CMP_TRY(hipFree(L->d_ltiles)); L->d_ltiles = nullptr; }
    CMP_TRY(hipMalloc(&L->d_ltiles, sizeof(ListTile) * ht.size()));
    CMP_TRY(hipMemcpy(L->d_ltiles, ht.data(), sizeof(ListTile) * ht.size(), hipMemcpyHostToDevice));
    if (L->d_ftiles) { CMP_TRY(hipFree(L->d_ftiles)); L->d_ftiles = nullptr; }
    L->nftiles = 0;
    if (!ft.empty()) {
        CMP_TRY(hipMalloc(&L->d_ftiles, sizeof(TileDev) * ft.size()));
        CMP_TRY(hipMemcpy(L->d_ftiles, ft.data(), sizeof(TileDev) * ft.size(), hipMemcpyHostToDevice));
        L->nftiles = (int)ft.size();
    }
    if (!L->d_llev) {
        // [0] the stage reads the cells beyond its boxes where the ghost fill takes them from (a
        // level stepped alone); [1] it reads the filled ghost cells (a level of a hierarchy)
        CMP_TRY(hipMalloc(&L->d_llev, 2 * sizeof(ListLevel)));
        const ListLevel lv[2] = {{L->kp, L->d_boxes, L->d_fd, L->d_err}, {L->kp, L->d_boxes, nullptr, L->d_err}};
        CMP_TRY(hipMemcpy(L->d_llev, lv, 2 * sizeof(ListLevel), hipMemcpyHostToDevice));
    }
    if (!L->d_lctr) {
        CMP_TRY(hipMalloc(&L->d_lctr, sizeof(int) * LIST_NCTR));
        CMP_TRY(hipMemset(L->d_lctr, 0, sizeof(int) * LIST_NCTR));
    }
    L->list_grid = list_capacity(L->lc, L->list_waves);
    if (L->list_grid <= 0) return CMP_ERR_UNSUPPORTED;
    L->h_ltiles = std::move(ht);
    return CMP_OK;
}

// One flow stage of L through its work-list (fused: the ghost reads of cmp_level_rk3_step, else the
// filled ghost cells, as in a hierarchy's stage); the boxes the
// fixed-shape kernel keeps by a k_stage launch over their tiles first, on the same stream (the two
// launches write disjoint cells and max-accumulate one c_max).
int level_list_stage(cmp_level* L, int mode, int stage, const double* Us, const double* Un, double* Uout,
                     const double* dt, double* cmax, hipStream_t s, bool fused)
{
    if (L->nftiles > 0)
        CMP_TRY(launch_stage(L->lc, L->kp, mode, stage, L->d_ftiles, L->nftiles, L->d_boxes, nullptr, 0, Us, Un, Uout,
                             dt, (unsigned long long*)cmax, L->d_err, s, nullptr, fused ? L->d_fd : nullptr));
    ListCall c{};
    c.Us[0] = Us;
    c.Un[0] = Un;
    c.Uout[0] = Uout;
    c.cmax[0] = (unsigned long long*)cmax;
    CMP_TRY(launch_stage_list(L->lc, L->list_waves, L->d_llev + (fused ? 0 : 1), L->d_ltiles, (int)L->h_ltiles.size(),
                              L->d_lctr, c, mode, stage, dt, L->list_grid, s));
    return CMP_OK;
}

// ---- the tile autotuner -------------------------------------------------------------------------

namespace {

constexpr int MAX_CAND = 12;
constexpr int NWAVES = 2;            // occupancy targets of the work-list kernel: 2, 3 waves per SIMD
constexpr int WAVES[NWAVES] = {2, 3};
// what a level that mixes the two kernels pays for its second launch per stage (ns), added to its
// summed swept time when the tuner weighs it against the fixed or the work-list kernel alone
constexpr double MIX_NS = 4000.0;

struct TuneKey {
    int dim, scheme, limiter, riemann, thinc, nphase, ng, n[3];
    bool operator<(const TuneKey& o) const
    {
        return std::tie(dim, scheme, limiter, riemann, thinc, nphase, ng, n[0], n[1], n[2]) <
               std::tie(o.dim, o.scheme, o.limiter, o.riemann, o.thinc, o.nphase, o.ng, o.n[0], o.n[1], o.n[2]);
    }
};

// The measured sweep of one box shape: the fixed-shape kernel's time and, per candidate shape (clipped
// to the box) and occupancy target, the work-list kernel's, in ns per cell.
struct TuneEntry {
    double ns_fixed = 0.0;
    int ncand = 0;
    int cand[MAX_CAND][3] = {};
    double ns[MAX_CAND][NWAVES] = {};
    // the candidate of least time for occupancy target w (-1: none)
    int best(int w) const
    {
        int b = -1;
        for (int k = 0; k < ncand; ++k)
            if (b < 0 || ns[k][w] < ns[b][w]) b = k;
        return b;
    }
};

std::mutex g_mu;
std::map<TuneKey, TuneEntry> g_cache;

TuneKey key_of(const cmp_level* L, const int* n)
{
    TuneKey k{L->lc.dim, L->lc.scheme, L->lc.limiter, L->lc.riemann, L->lc.thinc, L->lc.nphase, L->ng, {1, 1, 1}};
    for (int d = 0; d < L->P.dim; ++d) k.n[d] = n[d];
    return k;
}

// Candidate shapes per dimension (each kept where it fits the work-list kernel's layout).
const int CAND2[][3] = {{16, 15, 1}, {16, 16, 1}, {32, 8, 1}, {8, 32, 1}, {24, 8, 1}, {8, 24, 1}, {24, 10, 1},
                        {16, 8, 1}, {8, 16, 1}, {12, 16, 1}, {16, 12, 1}, {8, 8, 1}};
const int CAND3[][3] = {{8, 8, 3}, {8, 8, 4}, {16, 4, 4}, {4, 8, 8}, {8, 4, 8}, {4, 4, 16}, {8, 8, 2}, {16, 4, 3}};

// Every valid cell of every box of a synthetic level takes the state of the sample box's cell at
// the same position modulo the sample box's extents.
__global__ void k_sample_fill(const BoxDev* __restrict__ dboxes, double* Ud, BoxDev sb, const double* __restrict__ Us,
                              int N, long long cells)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= cells) return;
    const BoxDev db = dboxes[blockIdx.y];
    const int nx = db.hi[0] - db.lo[0] + 1, ny = db.hi[1] - db.lo[1] + 1;
    const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((long long)nx * ny));
    const int sx = sb.hi[0] - sb.lo[0] + 1, sy = sb.hi[1] - sb.lo[1] + 1, sz = sb.hi[2] - sb.lo[2] + 1;
    const long long di = db.off + fab_index(db, db.lo[0] + i, db.lo[1] + j, db.lo[2] + k);
    const long long si = sb.off + fab_index(sb, sb.lo[0] + i % sx, sb.lo[1] + j % sy, sb.lo[2] + k % sz);
    for (int n = 0; n < N; ++n) Ud[di + n * db.sn] = Us[si + n * sb.sn];
}

struct Synthetic {
    cmp_level* S = nullptr;
    double* U = nullptr;
    double* V = nullptr;
    double* dt = nullptr;
    long long cells = 0;
    ~Synthetic()
    {
        if (S) cmp_level_destroy(S);
        void* b[] = {U, V, dt};
        for (void* p : b)
            if (p) (void)hipFree(p);
    }
};

// A synthetic periodic level of M copies of an n[0] x n[1] x n[2] box in a row along x (about 2^19
// cells), its valid cells sampled from box sbox of level L (state U), its ghosts filled.
int make_synthetic(const cmp_level* L, const double* U, int sbox, const int* n, Synthetic& Y, hipStream_t s)
{
    const int dim = L->P.dim;
    long long cells = 1;
    for (int d = 0; d < dim; ++d) cells *= n[d];
    const int M = (int)std::max(1LL, std::min(4096LL, (1LL << 19) / cells));
    cmp_geometry g = L->G;
    for (int d = 0; d < 3; ++d) {
        g.domain_lo[d] = 0;
        g.domain_hi[d] = (d < dim) ? (d == 0 ? M * n[0] : n[d]) - 1 : 0;
        g.is_periodic[d] = 1;
    }
    std::vector<cmp_box> bx(M);
    long long off = 0;
    for (int b = 0; b < M; ++b) {
        for (int d = 0; d < 3; ++d) {
            bx[b].lo[d] = (d == 0) ? b * n[0] : 0;
            bx[b].hi[d] = (d < dim) ? bx[b].lo[d] + n[d] - 1 : 0;
        }
        bx[b].offset = off;
        off += cmp_fab_size(dim, &bx[b], L->ng, L->P.nstate);
    }
    int r = cmp_level_create(&L->P, &g, M, bx.data(), L->ng, &Y.S);
    if (r) return r;
    CMP_TRY(hipMalloc(&Y.U, sizeof(double) * off));
    CMP_TRY(hipMalloc(&Y.V, sizeof(double) * off));
    CMP_TRY(hipMalloc(&Y.dt, sizeof(double)));
    const double dt = 1.0e-9;
    CMP_TRY(hipMemcpyAsync(Y.dt, &dt, sizeof(double), hipMemcpyHostToDevice, s));
    CMP_TRY(hipMemsetAsync(Y.U, 0, sizeof(double) * off, s));
    const BoxDev sb = L->h_boxes[sbox];
    hipLaunchKernelGGL(k_sample_fill, dim3((unsigned)((cells + NT - 1) / NT), (unsigned)M), dim3(NT), 0, s,
                       Y.S->d_boxes, Y.U, sb, U, L->P.nstate, cells);
    CMP_TRY(hipGetLastError());
    r = cmp_level_fill_ghosts(Y.S, Y.U, s);
    if (r) return r;
    Y.cells = cells * M;
    return CMP_OK;
}

// Least of `reps` timed launches of f (after one untimed), in ms.
template <class F>
int time_min(F&& f, int reps, hipStream_t s, double* ms)
{
    hipEvent_t a = nullptr, b = nullptr;
    CMP_TRY(hipEventCreate(&a));
    CMP_TRY(hipEventCreate(&b));
    int r = f();
    double best = 1e300;
    for (int i = 0; i < reps && !r; ++i) {
        (void)hipEventRecord(a, s);
        r = f();
        (void)hipEventRecord(b, s);
        if (!r) r = check_hip(hipEventSynchronize(b));
        float t = 0.f;
        if (!r) r = check_hip(hipEventElapsedTime(&t, a, b));
        best = std::min(best, (double)t);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = best;
    return r;
}

// The measured sweep of one box shape n of level L (state U, sample box sbox).
int tune_one(const cmp_level* L, const double* U, int sbox, const int* n, int reps, TuneEntry& E, hipStream_t s)
{
    const int dim = L->P.dim;
    Synthetic Y;
    int r = make_synthetic(L, U, sbox, n, Y, s);
    if (r) return r;
    cmp_level* S = Y.S;
    double ms = 0.0;
    // the fixed-shape kernel, fused ghost reads (cmp_level_rk3_step's stage 0)
    r = time_min([&]() -> int {
        CMP_TRY(launch_stage(S->lc, S->kp, MODE_RK, 0, S->d_tiles, S->ntiles, S->d_boxes, nullptr, 0, Y.U, nullptr,
                             Y.V, Y.dt, nullptr, S->d_err, s, nullptr, S->d_fd));
        return CMP_OK;
    }, reps, s, &ms);
    if (r) return r;
    E.ns_fixed = ms * 1e6 / (double)Y.cells;
    // the work-list kernel with every candidate shape that fits (clipped to the box; duplicates once)
    // at both occupancy targets
    const int (*cand)[3] = (dim == 2) ? CAND2 : CAND3;
    const int ncand = (dim == 2) ? (int)(sizeof(CAND2) / sizeof(CAND2[0])) : (int)(sizeof(CAND3) / sizeof(CAND3[0]));
    E.ncand = 0;
    S->box_ext.assign(3 * (size_t)S->nbox, 1);
    for (int c = 0; c < ncand && E.ncand < MAX_CAND; ++c) {
        int e[3] = {1, 1, 1};
        for (int d = 0; d < dim; ++d) e[d] = std::min(cand[c][d], n[d]);
        if (!tile_shape_ok(S->lc, e)) continue;
        bool dup = false;
        for (int k = 0; k < E.ncand && !dup; ++k)
            dup = E.cand[k][0] == e[0] && E.cand[k][1] == e[1] && E.cand[k][2] == e[2];
        if (dup) continue;
        for (int b = 0; b < S->nbox; ++b)
            for (int d = 0; d < 3; ++d) S->box_ext[3 * (size_t)b + d] = e[d];
        for (int w = 0; w < NWAVES; ++w) {
            S->list_waves = WAVES[w];
            if ((r = level_build_list(S))) return r;
            r = time_min([&]() -> int { return level_list_stage(S, MODE_RK, 0, Y.U, nullptr, Y.V, Y.dt, nullptr, s); },
                         reps, s, &ms);
            if (r) return r;
            E.ns[E.ncand][w] = ms * 1e6 / (double)Y.cells;
        }
        for (int d = 0; d < 3; ++d) E.cand[E.ncand][d] = e[d];
        ++E.ncand;
    }
    return CMP_OK;
}

void box_shape(const cmp_level* L, int b, int* n)
{
    n[0] = n[1] = n[2] = 1;
    for (int d = 0; d < L->P.dim; ++d) n[d] = L->boxes[b].hi[d] - L->boxes[b].lo[d] + 1;
}

}  // namespace

// The cached choices for L (every box shape of L must be in the cache): per occupancy target, each
// box's candidate of least time and the level's summed time; the fixed kernel's summed time; with
// ext_mixed / t_mixed, per occupancy target each box by the faster of the two kernels (0 0 0: the
// fixed one) and that summed time, plus MIX_NS for the second launch when the level has boxes of both.
// CMP_ERR_UNSUPPORTED when a box shape has no usable entry.
int level_tuned(const cmp_level* L, std::vector<int>* ext, double* t_list, double* t_fixed,
                std::vector<int>* ext_mixed, double* t_mixed)
{
    std::lock_guard<std::mutex> g(g_mu);
    *t_fixed = 0.0;
    int nfix[NWAVES] = {};
    for (int w = 0; w < NWAVES; ++w) {
        ext[w].assign(3 * (size_t)L->nbox, 0);
        t_list[w] = 0.0;
        if (ext_mixed) {
            ext_mixed[w].assign(3 * (size_t)L->nbox, 0);
            t_mixed[w] = 0.0;
        }
    }
    for (int b = 0; b < L->nbox; ++b) {
        int n[3];
        box_shape(L, b, n);
        auto it = g_cache.find(key_of(L, n));
        if (it == g_cache.end()) return CMP_ERR_UNSUPPORTED;
        const TuneEntry& E = it->second;
        const double cells = (double)n[0] * n[1] * n[2];
        *t_fixed += cells * E.ns_fixed;
        for (int w = 0; w < NWAVES; ++w) {
            const int k = E.best(w);
            // an entry loaded from a file of another build is checked against this build's layout
            if (k < 0 || !tile_shape_ok(L->lc, E.cand[k])) return CMP_ERR_UNSUPPORTED;
            t_list[w] += cells * E.ns[k][w];
            for (int d = 0; d < 3; ++d) ext[w][3 * (size_t)b + d] = E.cand[k][d];
            if (!ext_mixed) continue;
            if (E.ns[k][w] < E.ns_fixed) {
                t_mixed[w] += cells * E.ns[k][w];
                for (int d = 0; d < 3; ++d) ext_mixed[w][3 * (size_t)b + d] = E.cand[k][d];
            } else {
                t_mixed[w] += cells * E.ns_fixed;
                ++nfix[w];
            }
        }
    }
    if (ext_mixed)
        for (int w = 0; w < NWAVES; ++w)
            if (nfix[w] > 0 && nfix[w] < L->nbox) t_mixed[w] += MIX_NS;
    return CMP_OK;
}

// Applies the cached choices to L (force: 1 the work-list kernel for every box at its better
// occupancy target, -1 the fixed kernel). force 0: the summed sweeps give up to two candidates
// besides the fixed kernel - the work-list kernel for every box, and each box by the faster of the
// two kernels, each at its better occupancy target, each only when its sum beats the fixed kernel's -
// and one RK stage of the level itself on U (HIP events, least of 3) decides among them and the
// fixed kernel: boxes of other shapes beside each other, and two launches in place of one, do not
// cost what the sums of isolated sweeps say.
static int level_apply_tuning(cmp_level* L, int force, const double* U, hipStream_t s)
{
    std::vector<int> ext[NWAVES], mix[NWAVES];
    double t_list[NWAVES], t_mix[NWAVES], t_fixed = 0.0;
    int r = level_tuned(L, ext, t_list, &t_fixed, mix, t_mix);
    if (r) return r;
    L->gkey[0] = nullptr;   // a captured step is captured again
    ++L->tiles_gen;
    const int w = (t_list[1] < t_list[0]) ? 1 : 0;
    const int wm = (t_mix[1] < t_mix[0]) ? 1 : 0;
    L->use_list = false;
    if (force < 0) return CMP_OK;
    if (force > 0) {
        L->box_ext = std::move(ext[w]);
        L->list_waves = WAVES[w];
        if ((r = level_build_list(L))) return r;
        L->use_list = true;
        return CMP_OK;
    }
    struct Opt {
        std::vector<int> e;
        int w;
    };
    // the candidates whose summed sweeps beat the fixed kernel's; none: the fixed kernel, unmeasured
    std::vector<Opt> opts;
    if (t_list[w] < t_fixed) opts.push_back({ext[w], w});
    bool any = false;
    for (size_t b = 0; b < mix[wm].size(); b += 3) any = any || mix[wm][b] > 0;
    if (any && t_mix[wm] < t_fixed && (opts.empty() || mix[wm] != ext[w] || wm != w)) opts.push_back({mix[wm], wm});
    if (opts.empty()) return CMP_OK;
    long long total = 0;
    for (const cmp_box& b : L->boxes) total += cmp_fab_size(L->P.dim, &b, L->ng, L->P.nstate);
    // the level's error word as the measurement finds it, put back afterwards (a finer level's cells
    // beyond its boxes are not all filled in U, so the measured stages may flag them)
    int err0 = 0;
    CMP_TRY(hipMemcpy(&err0, L->d_err, sizeof(int), hipMemcpyDeviceToHost));
    double *V = nullptr, *dt = nullptr;
    CMP_TRY(hipMalloc(&V, sizeof(double) * (size_t)total));
    r = check_hip(hipMalloc(&dt, sizeof(double)));
    const double dtv = 1.0e-9;
    if (!r) r = check_hip(hipMemcpy(dt, &dtv, sizeof(double), hipMemcpyHostToDevice));
    double best = 0.0, ms = 0.0;
    int pick = -1;
    if (!r)
        r = time_min([&]() -> int {
            CMP_TRY(launch_stage(L->lc, L->kp, MODE_RK, 0, L->d_tiles, L->ntiles, L->d_boxes, nullptr, 0, U, nullptr, V,
                                 dt, nullptr, L->d_err, s, nullptr, L->d_fd));
            return CMP_OK;
        }, 3, s, &best);
    for (size_t k = 0; k < opts.size() && !r; ++k) {
        L->box_ext = opts[k].e;
        L->list_waves = WAVES[opts[k].w];
        if ((r = level_build_list(L))) break;
        r = time_min([&]() -> int { return level_list_stage(L, MODE_RK, 0, U, nullptr, V, dt, nullptr, s); }, 3, s, &ms);
        if (!r && ms < best) {
            best = ms;
            pick = (int)k;
        }
    }
    if (!r) r = check_hip(hipStreamSynchronize(s));
    if (!r) r = check_hip(hipMemcpy(L->d_err, &err0, sizeof(int), hipMemcpyHostToDevice));
    (void)hipFree(V);
    if (dt) (void)hipFree(dt);
    if (r) return r;
    if (pick < 0) return CMP_OK;
    L->box_ext = std::move(opts[pick].e);
    L->list_waves = WAVES[opts[pick].w];
    if ((r = level_build_list(L))) return r;
    L->use_list = true;
    return CMP_OK;
}

}  // namespace cmp

using namespace cmp;

extern "C" {

int cmp_level_set_tiles(cmp_level* L, const int* ext, int waves)
{
    if (!L) return CMP_ERR_ARG;
    if (!ext) {
        L->use_list = false;
        ++L->tiles_gen;
        L->gkey[0] = nullptr;
        return CMP_OK;
    }
    if (waves != 2 && waves != 3) return CMP_ERR_ARG;
    if (L->lc.riemann != 1 && L->lc.riemann != 2) return CMP_ERR_UNSUPPORTED;
    std::vector<int> e(3 * (size_t)L->nbox, 1);
    int nlist = 0;
    for (int b = 0; b < L->nbox; ++b) {
        const int* x = &ext[3 * (size_t)b];
        if (x[0] == 0 && x[1] == 0 && x[2] == 0) {   // the fixed-shape kernel keeps this box
            e[3 * (size_t)b] = e[3 * (size_t)b + 1] = e[3 * (size_t)b + 2] = 0;
            continue;
        }
        for (int d = 0; d < L->P.dim; ++d) {
            const int n = L->boxes[b].hi[d] - L->boxes[b].lo[d] + 1;
            e[3 * (size_t)b + d] = std::min(x[d], n);
        }
        if (!tile_shape_ok(L->lc, &e[3 * (size_t)b])) return CMP_ERR_ARG;
        ++nlist;
    }
    if (nlist == 0) {
        L->use_list = false;
        ++L->tiles_gen;
        L->gkey[0] = nullptr;
        return CMP_OK;
    }
    L->box_ext = std::move(e);
    L->list_waves = waves;
    int r = level_build_list(L);
    if (r) return r;
    L->use_list = true;
    L->gkey[0] = nullptr;
    return CMP_OK;
}

int cmp_level_tile_shapes(const cmp_level* L, int* use_list, int* waves, int* ext)
{
    if (!L) return CMP_ERR_ARG;
    if (use_list) *use_list = L->use_list ? 1 : 0;
    if (waves) *waves = L->use_list ? L->list_waves : 0;
    if (ext) {
        int tx, ty, tz;
        tile_dims(L->P.dim, &tx, &ty, &tz);
        const int te[3] = {tx, ty, tz};
        for (int b = 0; b < L->nbox; ++b)
            for (int d = 0; d < 3; ++d)
                ext[3 * (size_t)b + d] = L->use_list ? L->box_ext[3 * (size_t)b + d] : ((d < L->P.dim) ? te[d] : 1);
    }
    return CMP_OK;
}

int cmp_level_tune_tiles(cmp_level* L, const double* U, int nextra, const int* extra, int force, void* stream)
{
    if (!L || !U || nextra < 0 || (nextra > 0 && !extra) || force < -1 || force > 1) return CMP_ERR_ARG;
    if (L->lc.riemann != 1 && L->lc.riemann != 2) return CMP_ERR_UNSUPPORTED;
    if (stream) CMP_TRY(hipStreamSynchronize((hipStream_t)stream));
    // the sample: the level's largest box
    int sbox = 0;
    long long most = -1;
    std::vector<std::array<int, 3>> shapes;
    for (int b = 0; b < L->nbox; ++b) {
        std::array<int, 3> n;
        box_shape(L, b, n.data());
        const long long c = (long long)n[0] * n[1] * n[2];
        if (c > most) { most = c; sbox = b; }
        shapes.push_back(n);
    }
    for (int k = 0; k < nextra; ++k) {
        std::array<int, 3> n{1, 1, 1};
        for (int d = 0; d < L->P.dim; ++d) n[d] = extra[3 * k + d];
        shapes.push_back(n);
    }
    std::sort(shapes.begin(), shapes.end());
    shapes.erase(std::unique(shapes.begin(), shapes.end()), shapes.end());
    hipStream_t s = nullptr;
    CMP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    int r = CMP_OK;
    for (const auto& n : shapes) {
        bool have;
        {
            std::lock_guard<std::mutex> g(g_mu);
            have = g_cache.count(key_of(L, n.data())) > 0;
        }
        if (have) continue;
        bool small = false;
        for (int d = 0; d < L->P.dim; ++d) small = small || n[d] < L->ng;
        // a box thinner than the ghost width has no synthetic periodic level: it keeps no entry,
        // and a level holding one keeps the fixed kernel
        if (small) continue;
        TuneEntry E;
        if ((r = tune_one(L, U, sbox, n.data(), 3, E, s))) break;
        std::lock_guard<std::mutex> g(g_mu);
        g_cache[key_of(L, n.data())] = E;
    }
    const bool swept = (r == CMP_OK);
    if (swept) r = level_apply_tuning(L, force, U, s);
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    if (swept && r == CMP_ERR_UNSUPPORTED) {
        L->use_list = false;
        ++L->tiles_gen;
        return CMP_OK;
    }
    return r;
}

int cmp_tile_tuning(const cmp_level* L, const int* n, double* ns_fixed, int* ncand, int* cand_ext, double* cand_ns)
{
    if (!L || !n) return CMP_ERR_ARG;
    std::lock_guard<std::mutex> g(g_mu);
    auto it = g_cache.find(key_of(L, n));
    if (it == g_cache.end()) return CMP_ERR_ARG;
    const TuneEntry& E = it->second;
    if (ns_fixed) *ns_fixed = E.ns_fixed;
    if (ncand) *ncand = E.ncand;
    for (int k = 0; k < E.ncand; ++k) {
        if (cand_ext)
            for (int d = 0; d < 3; ++d) cand_ext[3 * k + d] = E.cand[k][d];
        if (cand_ns)
            for (int w = 0; w < NWAVES; ++w) cand_ns[NWAVES * k + w] = E.ns[k][w];
    }
    return CMP_OK;
}

// File format: one line per entry, the key (dim scheme limiter riemann thinc nphase ng nx ny nz),
// the fixed kernel's ns per cell, the number of candidates, then per candidate its extents and its
// ns per cell at 2 and 3 waves per SIMD.
int cmp_tile_cache_save(const char* path)
{
    if (!path) return CMP_ERR_ARG;
    FILE* f = std::fopen(path, "w");
    if (!f) return CMP_ERR_ARG;
    std::lock_guard<std::mutex> g(g_mu);
    for (const auto& kv : g_cache) {
        const TuneKey& k = kv.first;
        const TuneEntry& E = kv.second;
        std::fprintf(f, "%d %d %d %d %d %d %d %d %d %d %.9g %d", k.dim, k.scheme, k.limiter, k.riemann, k.thinc,
                     k.nphase, k.ng, k.n[0], k.n[1], k.n[2], E.ns_fixed, E.ncand);
        for (int c = 0; c < E.ncand; ++c)
            std::fprintf(f, "  %d %d %d %.9g %.9g", E.cand[c][0], E.cand[c][1], E.cand[c][2], E.ns[c][0], E.ns[c][1]);
        std::fprintf(f, "\n");
    }
    std::fclose(f);
    return CMP_OK;
}

int cmp_tile_cache_load(const char* path)
{
    if (!path) return CMP_ERR_ARG;
    FILE* f = std::fopen(path, "r");
    if (!f) return CMP_ERR_ARG;
    std::lock_guard<std::mutex> g(g_mu);
    int rc = CMP_OK;
    for (;;) {
        TuneKey k{};
        TuneEntry E;
        const int got = std::fscanf(f, "%d %d %d %d %d %d %d %d %d %d %lf %d", &k.dim, &k.scheme, &k.limiter,
                                    &k.riemann, &k.thinc, &k.nphase, &k.ng, &k.n[0], &k.n[1], &k.n[2], &E.ns_fixed,
                                    &E.ncand);
        if (got == EOF) break;
        if (got != 12 || E.ncand < 0 || E.ncand > MAX_CAND) { rc = CMP_ERR_ARG; break; }
        bool ok = true;
        for (int c = 0; c < E.ncand && ok; ++c)
            ok = std::fscanf(f, "%d %d %d %lf %lf", &E.cand[c][0], &E.cand[c][1], &E.cand[c][2], &E.ns[c][0],
                             &E.ns[c][1]) == 5;
        if (!ok) { rc = CMP_ERR_ARG; break; }
        g_cache[k] = E;
    }
    std::fclose(f);
    return rc;
}

int cmp_tile_cache_clear(void)
{
    std::lock_guard<std::mutex> g(g_mu);
    g_cache.clear();
    return CMP_OK;
}

}  // extern "C"
