// cmp_batch.hip - a batch of independent levels ("cases") advanced together: the served-stream
// form of the solver. Every admitted case keeps its state in device memory for its whole life, in
// three arenas (U^n, U^(1), U^(2)) at the same offset; an RK3 step of the batch is one per-case
// ComputeDt launch and three stage launches over the tiles of every case. Each case has its own
// parameters, geometry, boxes, dt, c_max and error word, so a case's result does not depend on what
// else is in the batch (the stage kernel computes every tile from its own case's data only).
#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
#include <thread>
#include <vector>

#include "cmp_kernels.hpp"
#include "cmp_level.hpp"

using namespace cmp;

namespace {

struct BatchCase {
    cmp_level* L = nullptr;   // the case as a level: its tables, with box offsets inside its range
    long long off = 0;        // its range in the arenas (doubles)
    long long len = 0;
    double cfl = 0.0;
    int limiter = 1;
    hipStream_t pf_stream = nullptr;   // the stream of its Phase-Field step (with the Phase-Field step on)
};

}  // namespace

struct cmp_batch {
    LaunchCfg lc{};
    int dim = 0;
    long long capacity = 0;    // doubles per arena
    int max_cases = 0;
    double* d_arena[3] = {nullptr, nullptr, nullptr};   // U^n, U^(1), U^(2)
    std::map<long long, long long> free_ranges;          // offset -> length, coalesced
    std::vector<BatchCase> cases;                        // by slot
    // per-slot device arrays
    KParams* d_kps = nullptr;
    FillDev* d_fds = nullptr;
    BatchSlot* d_slots = nullptr;
    double* d_cmax = nullptr;
    double* d_dt = nullptr;
    double* d_time = nullptr;
    int* d_steps = nullptr;
    int* d_err = nullptr;
    // the batch's box and tile tables (all admitted cases), rebuilt when the set changes
    BoxDev* d_boxes = nullptr;
    TileDev* d_tiles = nullptr;
    long long box_cap = 0, tile_cap = 0;
    int ntiles = 0;
    bool dirty = true;
    std::vector<KParams> h_kps;
    std::vector<FillDev> h_fds;
    std::vector<BatchSlot> h_slots;
    // the Phase-Field step of every case (cmp_batch_pf_setup): each case's level carries its buffers
    bool pf_on = false;
    cmp_pf_params pf{};
    hipEvent_t ev_flow = nullptr;       // the flow step queued on the batch stream
};

namespace {

void batch_free(cmp_batch* B)
{
    for (BatchCase& c : B->cases) {
        if (c.L) cmp_level_destroy(c.L);
        if (c.pf_stream) (void)hipStreamDestroy(c.pf_stream);
    }
    if (B->ev_flow) (void)hipEventDestroy(B->ev_flow);
    void* bufs[] = {B->d_arena[0], B->d_arena[1], B->d_arena[2], B->d_kps, B->d_fds, B->d_slots, B->d_cmax,
                    B->d_dt, B->d_time, B->d_steps, B->d_err, B->d_boxes, B->d_tiles};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    delete B;
}

// first fit of len doubles in the arenas; -1 when no free range is long enough
long long arena_take(cmp_batch* B, long long len)
{
    for (auto it = B->free_ranges.begin(); it != B->free_ranges.end(); ++it) {
        if (it->second < len) continue;
        const long long off = it->first, rest = it->second - len;
        B->free_ranges.erase(it);
        if (rest > 0) B->free_ranges[off + len] = rest;
        return off;
    }
    return -1;
}

void arena_give(cmp_batch* B, long long off, long long len)
{
    auto it = B->free_ranges.emplace(off, len).first;
    auto nx = std::next(it);
    if (nx != B->free_ranges.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        B->free_ranges.erase(nx);
    }
    if (it != B->free_ranges.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            B->free_ranges.erase(it);
        }
    }
}

template <typename T>
int grow(T** p, long long* cap, long long need)
{
    if (need <= *cap) return CMP_OK;
    long long n = std::max(need, 2 * *cap);
    if (*p) CMP_TRY(hipFree(*p));
    *p = nullptr;
    *cap = 0;
    CMP_TRY(hipMalloc(p, sizeof(T) * n));
    *cap = n;
    return CMP_OK;
}

// The batch's tables from its cases: boxes with the case's arena offset and slot, tiles with box
// indices (and neighbour-box halo sources) into the batch table, each case's fill tables pointing
// at its boxes there.
int batch_rebuild(cmp_batch* B, hipStream_t s)
{
    std::vector<BoxDev> hb;
    std::vector<TileDev> ht;
    std::vector<int> first(B->cases.size(), 0);   // each case's first box in the batch table
    for (int slot = 0; slot < (int)B->cases.size(); ++slot) {
        BatchCase& c = B->cases[slot];
        if (!c.L) continue;
        const int base = (int)hb.size();
        first[slot] = base;
        for (BoxDev bx : c.L->h_boxes) {
            bx.off += c.off;
            bx.ooff += c.off;
            bx.pad = slot;
            hb.push_back(bx);
        }
        for (TileDev t : c.L->h_tiles) {
            t.box += base;
            for (int d = 0; d < 3; ++d)
                for (int sd = 0; sd < 2; ++sd)
                    if (t.halo[d][sd] >= 0) t.halo[d][sd] += base;
            ht.push_back(t);
        }
    }
    int r;
    if ((r = grow(&B->d_boxes, &B->box_cap, std::max<long long>(1, (long long)hb.size())))) return r;
    if ((r = grow(&B->d_tiles, &B->tile_cap, std::max<long long>(1, (long long)ht.size())))) return r;
    for (int slot = 0; slot < (int)B->cases.size(); ++slot) {
        BatchCase& c = B->cases[slot];
        if (!c.L) continue;
        B->h_fds[slot] = c.L->fd;
        B->h_fds[slot].boxes = B->d_boxes + first[slot];
    }
    if (!hb.empty()) CMP_TRY(hipMemcpyAsync(B->d_boxes, hb.data(), sizeof(BoxDev) * hb.size(), hipMemcpyHostToDevice, s));
    if (!ht.empty()) CMP_TRY(hipMemcpyAsync(B->d_tiles, ht.data(), sizeof(TileDev) * ht.size(), hipMemcpyHostToDevice, s));
    const size_t ns = B->cases.size();
    CMP_TRY(hipMemcpyAsync(B->d_kps, B->h_kps.data(), sizeof(KParams) * ns, hipMemcpyHostToDevice, s));
    CMP_TRY(hipMemcpyAsync(B->d_fds, B->h_fds.data(), sizeof(FillDev) * ns, hipMemcpyHostToDevice, s));
    CMP_TRY(hipMemcpyAsync(B->d_slots, B->h_slots.data(), sizeof(BatchSlot) * ns, hipMemcpyHostToDevice, s));
    // the host vectors are read by the copies above: finish them before the vectors go
    CMP_TRY(hipStreamSynchronize(s));
    B->ntiles = (int)ht.size();
    B->dirty = false;
    return CMP_OK;
}

}  // namespace

extern "C" {

int cmp_batch_create(const cmp_params* like, long long capacity, int max_cases, cmp_batch** out)
{
    if (!like || !out || capacity <= 0 || max_cases <= 0) return CMP_ERR_ARG;
    *out = nullptr;
    if (like->nphase != 0 || like->id_thinc != 0) return CMP_ERR_UNSUPPORTED;
    // the scheme / Riemann selection of a level with these parameters (validated by a level create
    // of the first case; here only what the batch kernel is built for)
    if (like->dim != 2 && like->dim != 3) return CMP_ERR_ARG;
    cmp_batch* B = new cmp_batch();
    B->dim = like->dim;
    B->capacity = capacity;
    B->max_cases = max_cases;
    B->cases.resize(max_cases);
    B->h_kps.resize(max_cases);
    B->h_fds.resize(max_cases);
    B->h_slots.assign(max_cases, BatchSlot{});
    B->free_ranges[0] = capacity;
    B->lc.dim = like->dim;
    B->lc.limiter = like->id_limiter;
    B->lc.riemann = like->id_riemann;
    B->lc.thinc = 0;
    B->lc.nphase = 0;
    if (like->id_weno == 1 && like->id_limiter == 5) B->lc.scheme = 5;
    else if (like->id_weno == 1 && like->id_limiter == 6) B->lc.scheme = 6;
    else if (like->id_weno == 0 && like->id_limiter >= 1 && like->id_limiter <= 4) B->lc.scheme = 1;
    else if (like->id_weno == 0 && like->id_limiter == 0) B->lc.scheme = 0;
    else { delete B; return CMP_ERR_UNSUPPORTED; }
    if (B->lc.riemann != 1 && B->lc.riemann != 2) { delete B; return CMP_ERR_UNSUPPORTED; }
    int rc = CMP_OK;
#define CMP_TRYB(x) do { rc = check_hip(x); if (rc) { batch_free(B); return rc; } } while (0)
    for (int a = 0; a < 3; ++a) CMP_TRYB(hipMalloc(&B->d_arena[a], sizeof(double) * capacity));
    CMP_TRYB(hipMalloc(&B->d_kps, sizeof(KParams) * max_cases));
    CMP_TRYB(hipMalloc(&B->d_fds, sizeof(FillDev) * max_cases));
    CMP_TRYB(hipMalloc(&B->d_slots, sizeof(BatchSlot) * max_cases));
    CMP_TRYB(hipMemset(B->d_slots, 0, sizeof(BatchSlot) * max_cases));
    double** dv[] = {&B->d_cmax, &B->d_dt, &B->d_time};
    for (double** p : dv) {
        CMP_TRYB(hipMalloc(p, sizeof(double) * max_cases));
        CMP_TRYB(hipMemset(*p, 0, sizeof(double) * max_cases));
    }
    int** iv[] = {&B->d_steps, &B->d_err};
    for (int** p : iv) {
        CMP_TRYB(hipMalloc(p, sizeof(int) * max_cases));
        CMP_TRYB(hipMemset(*p, 0, sizeof(int) * max_cases));
    }
#undef CMP_TRYB
    *out = B;
    return CMP_OK;
}

int cmp_batch_destroy(cmp_batch* B)
{
    if (!B) return CMP_OK;
    (void)hipDeviceSynchronize();
    batch_free(B);
    return CMP_OK;
}

int cmp_batch_admit(cmp_batch* B, const cmp_params* P, const cmp_geometry* G, int nbox, const cmp_box* boxes,
                    int ng, const double* U0, double cfl, int limiter, int* case_id, void* stream)
{
    if (!B || !P || !G || !boxes || nbox <= 0 || !U0 || !case_id || !(cfl > 0.0)) return CMP_ERR_ARG;
    *case_id = -1;
    hipStream_t s = (hipStream_t)stream;
    if (P->dim != B->dim || P->nphase != 0 || P->id_thinc != 0 || P->id_riemann != B->lc.riemann ||
        P->id_limiter != B->lc.limiter)
        return CMP_ERR_UNSUPPORTED;
    int slot = -1;
    for (int q = 0; q < B->max_cases; ++q)
        if (!B->cases[q].L) { slot = q; break; }
    if (slot < 0) return CMP_ERR_NOMEM;
    std::vector<cmp_box> bx(boxes, boxes + nbox);
    long long len = 0;
    int r = cmp_level_layout(P->dim, nbox, bx.data(), ng, P->nstate, &len);
    if (r) return r;
    const long long off = arena_take(B, len);
    if (off < 0) return CMP_ERR_NOMEM;
    cmp_level* L = nullptr;
    if ((r = cmp_level_create(P, G, nbox, bx.data(), ng, &L))) { arena_give(B, off, len); return r; }
    if (L->lc.scheme != B->lc.scheme || L->lc.thinc) {
        cmp_level_destroy(L);
        arena_give(B, off, len);
        return CMP_ERR_UNSUPPORTED;
    }
    double* Un = B->d_arena[0] + off;
    // the case's state, its initial wave speed and a fresh dt / time / step count / error word
    if ((r = check_hip(hipMemcpyAsync(Un, U0, sizeof(double) * len, hipMemcpyHostToDevice, s))) ||
        (r = cmp_level_initial_cmax(L, Un, B->d_cmax + slot, stream)) ||
        (r = check_hip(hipMemsetAsync(B->d_dt + slot, 0, sizeof(double), s))) ||
        (r = check_hip(hipMemsetAsync(B->d_time + slot, 0, sizeof(double), s))) ||
        (r = check_hip(hipMemsetAsync(B->d_steps + slot, 0, sizeof(int), s))) ||
        (r = check_hip(hipMemsetAsync(B->d_err + slot, 0, sizeof(int), s))) ||
        (r = check_hip(hipStreamSynchronize(s)))) {
        cmp_level_destroy(L);
        arena_give(B, off, len);
        return r;
    }
    // the level's own error word holds what its initial c_max found
    int e = cmp_level_check(L, stream);
    if (!e && B->pf_on) e = cmp_level_pf_setup(L, &B->pf);
    if (e) {
        cmp_level_destroy(L);
        arena_give(B, off, len);
        return e;
    }
    BatchCase& c = B->cases[slot];
    c.L = L;
    c.off = off;
    c.len = len;
    c.cfl = cfl;
    c.limiter = limiter;
    B->h_kps[slot] = L->kp;
    BatchSlot& sl = B->h_slots[slot];
    for (int d = 0; d < 3; ++d) sl.dx[d] = (d < P->dim) ? G->dx[d] : 1.0;
    sl.cfl = cfl;
    sl.dim = P->dim;
    sl.limiter = limiter;
    sl.active = 1;
    B->dirty = true;
    *case_id = slot;
    return CMP_OK;
}

int cmp_batch_retire(cmp_batch* B, int case_id, void* stream)
{
    if (!B || case_id < 0 || case_id >= B->max_cases || !B->cases[case_id].L) return CMP_ERR_ARG;
    // work already queued on the stream may still read the case's range
    CMP_TRY(hipStreamSynchronize((hipStream_t)stream));
    BatchCase& c = B->cases[case_id];
    cmp_level_destroy(c.L);
    arena_give(B, c.off, c.len);
    c = BatchCase{};
    B->h_slots[case_id].active = 0;
    B->dirty = true;
    return CMP_OK;
}

int cmp_batch_step(cmp_batch* B, void* stream)
{
    if (!B) return CMP_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    int r;
    if (B->dirty && (r = batch_rebuild(B, s))) return r;
    if (B->ntiles == 0) return CMP_OK;
    CMP_TRY(launch_batch_dt(B->d_slots, B->max_cases, B->d_cmax, B->d_dt, B->d_time, B->d_steps, B->d_err, s));
    double* const* A = B->d_arena;
    double* Us[3] = {A[0], A[1], A[2]};
    double* Uo[3] = {A[1], A[2], A[0]};
    for (int stage = 0; stage < 3; ++stage)
        CMP_TRY(launch_stage_batch(B->lc, B->d_kps, stage, B->d_tiles, B->ntiles, B->d_boxes, Us[stage], A[0],
                                   Uo[stage], B->d_dt, (unsigned long long*)B->d_cmax, B->d_err, B->d_fds, s));
    return CMP_OK;
}

int cmp_batch_case_state(cmp_batch* B, int case_id, double** Un, long long* len, double* time, int* steps,
                         void* stream)
{
    if (!B || case_id < 0 || case_id >= B->max_cases || !B->cases[case_id].L) return CMP_ERR_ARG;
    const BatchCase& c = B->cases[case_id];
    if (Un) *Un = B->d_arena[0] + c.off;
    if (len) *len = c.len;
    if (time || steps) {
        hipStream_t s = (hipStream_t)stream;
        double t = 0.0;
        int n = 0;
        CMP_TRY(hipMemcpyAsync(&t, B->d_time + case_id, sizeof(double), hipMemcpyDeviceToHost, s));
        CMP_TRY(hipMemcpyAsync(&n, B->d_steps + case_id, sizeof(int), hipMemcpyDeviceToHost, s));
        CMP_TRY(hipStreamSynchronize(s));
        if (time) *time = t;
        if (steps) *steps = n;
    }
    return CMP_OK;
}

int cmp_batch_case_download(cmp_batch* B, int case_id, double* host, void* stream)
{
    if (!B || !host || case_id < 0 || case_id >= B->max_cases || !B->cases[case_id].L) return CMP_ERR_ARG;
    const BatchCase& c = B->cases[case_id];
    hipStream_t s = (hipStream_t)stream;
    CMP_TRY(hipMemcpyAsync(host, B->d_arena[0] + c.off, sizeof(double) * c.len, hipMemcpyDeviceToHost, s));
    CMP_TRY(hipStreamSynchronize(s));
    return CMP_OK;
}

int cmp_batch_check(cmp_batch* B, int case_id, void* stream)
{
    if (!B || case_id < 0 || case_id >= B->max_cases || !B->cases[case_id].L) return CMP_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    int h = 0;
    CMP_TRY(hipMemcpyAsync(&h, B->d_err + case_id, sizeof(int), hipMemcpyDeviceToHost, s));
    CMP_TRY(hipStreamSynchronize(s));
    if (h) CMP_TRY(hipMemsetAsync(B->d_err + case_id, 0, sizeof(int), s));
    // the case's level word: its Phase-Field step's kernels
    const int l = cmp_level_check(B->cases[case_id].L, stream);
    return std::max(h, l);
}

int cmp_batch_pf_setup(cmp_batch* B, const cmp_pf_params* pf)
{
    if (!B || !pf) return CMP_ERR_ARG;
    if (pf->mechanism == CMP_PF_CAC && pf->explicit_rc == 1) return CMP_ERR_UNSUPPORTED;
    if (!B->ev_flow) CMP_TRY(hipEventCreateWithFlags(&B->ev_flow, hipEventDisableTiming));
    for (BatchCase& c : B->cases) {
        if (!c.L) continue;
        int r = cmp_level_pf_setup(c.L, pf);
        if (r) return r;
    }
    B->pf = *pf;
    B->pf_on = true;
    return CMP_OK;
}

int cmp_batch_pf_step(cmp_batch* B, int nthreads, double* info, void* stream)
{
    if (!B || !B->pf_on) return CMP_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    std::vector<int> slots;
    for (int q = 0; q < B->max_cases; ++q)
        if (B->cases[q].L) slots.push_back(q);
    if (slots.empty()) return CMP_OK;
    int dev = 0;
    CMP_TRY(hipGetDevice(&dev));
    for (int q : slots)
        if (!B->cases[q].pf_stream) CMP_TRY(hipStreamCreateWithFlags(&B->cases[q].pf_stream, hipStreamNonBlocking));
    // every case's Phase-Field step starts after the flow step the batch stream holds
    CMP_TRY(hipEventRecord(B->ev_flow, s));
    std::atomic<int> next{0}, err{0};
    auto work = [&]() {
        if (hipSetDevice(dev) != hipSuccess) { err.store(CMP_ERR_HIP); return; }
        for (int k = next.fetch_add(1); k < (int)slots.size(); k = next.fetch_add(1)) {
            const int q = slots[k];
            BatchCase& c = B->cases[q];
            double inf6[6] = {0, 0, 0, 0, 0, 0};
            int r = check_hip(hipStreamWaitEvent(c.pf_stream, B->ev_flow, 0));
            if (!r) r = cmp_level_pf_evolve(c.L, B->d_arena[0] + c.off, B->d_arena[1] + c.off,
                                            B->d_arena[2] + c.off, B->d_dt + q, inf6, c.pf_stream);
            if (!r) r = check_hip(hipStreamSynchronize(c.pf_stream));
            if (info)
                for (int m = 0; m < 6; ++m) info[6 * q + m] = inf6[m];
            if (r) {
                int z = 0;
                err.compare_exchange_strong(z, r);
            }
        }
    };
    const int nw = std::max(1, std::min(nthreads > 0 ? nthreads : 8, (int)slots.size()));
    std::vector<std::thread> pool;
    for (int w = 1; w < nw; ++w) pool.emplace_back(work);
    work();
    for (std::thread& t : pool) t.join();
    return err.load();
}

int cmp_batch_info(const cmp_batch* B, int* ncases, int* ntiles, long long* free_doubles)
{
    if (!B) return CMP_ERR_ARG;
    int n = 0;
    for (const BatchCase& c : B->cases) n += c.L ? 1 : 0;
    long long f = 0;
    for (const auto& kv : B->free_ranges) f += kv.second;
    if (ncases) *ncases = n;
    if (ntiles) *ntiles = B->dirty ? -1 : B->ntiles;
    if (free_doubles) *free_doubles = f;
    return CMP_OK;
}

}  // extern "C"
