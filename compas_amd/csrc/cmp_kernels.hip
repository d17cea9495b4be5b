// cmp_kernels.hip - gfx950 kernels of the five-equation RHS, RK3 stage, ghost fill and CFL.
//
// Built twice, with CMP_KERNEL_DIM=2 and CMP_KERNEL_DIM=3, so that the two dimensions compile in
// parallel; cmp_capi.hip routes a call to the right instantiation.
//
// Stage kernel (one workgroup = one tile of a box):
//   for each direction d:
//     1. the primitives of the tile plus G cells on both sides along d are staged in LDS
//        (cons->prim once per staged cell),
//     2. every face of the tile along d is computed once, one face per thread (reconstruction,
//        prim->cons, Riemann); flux and u*_n then go to LDS over the staged primitives,
//     3. every cell accumulates (F_+ - F_-)/dx_d and (u*_+ - u*_-)/dx_d in registers;
//   then dU = -div F, dU[alpha1] += phi1 (1+K1) div u*, and the RK3 combination writes the new
//   state: dU never goes to HBM. The face wave speeds of the tile are max-reduced with wave
//   shuffles and one 64-bit atomicMax per workgroup.
#include "cmp_kernels.hpp"

#ifndef CMP_KERNEL_DIM
#error "CMP_KERNEL_DIM must be 2 or 3"
#endif

namespace cmp {

// Tile shapes chosen so that every direction's face count fits one pass of the 256 threads:
// 2D 16x15 (255 x-faces, 256 y-faces), 3D 8x8x3 (216, 216, 256).
template <int DIM> struct Tile;
template <> struct Tile<2> { static constexpr int X = 16, Y = 15, Z = 1; };
template <> struct Tile<3> { static constexpr int X = 8, Y = 8, Z = 3; };

template <int SCHEME> struct Stencil { static constexpr int G = (SCHEME >= 5) ? 3 : (SCHEME == 1 ? 2 : 1); };

constexpr int cmax3(int a, int b, int c) { return (a > b ? (a > c ? a : c) : (b > c ? b : c)); }

template <int DIM, int G>
struct Lds {
    static constexpr int TX = Tile<DIM>::X, TY = Tile<DIM>::Y, TZ = Tile<DIM>::Z;
    static constexpr int RMAX = cmax3((TX + 2 * G) * TY * TZ, TX * (TY + 2 * G) * TZ,
                                      DIM == 3 ? TX * TY * (TZ + 2 * G) : 0);
    static constexpr int FMAX = cmax3((TX + 1) * TY * TZ, TX * (TY + 1) * TZ,
                                      DIM == 3 ? TX * TY * (TZ + 1) : 0);
};

// The tile shape the stage code is instantiated for. ShapeFixed: k_stage's one shape, whose
// extents fold into the index arithmetic as constants. ShapeRT: a shape read per tile at run time
// (k_stage_list), any TX x TY x TZ of at most NT cells whose staged region along every direction
// fits LdsRT::RMAX cells and whose faces along every direction fit NT + LdsRT::XS (the faces beyond
// NT take a second pass and LDS slots of their own); the host checks a shape against these bounds
// (tile_shape_ok) before it enters a work-list.
template <int DIM>
struct ShapeFixed {
    static constexpr bool RT = false;
    __device__ __forceinline__ int t(int d) const
    {
        return d == 0 ? Tile<DIM>::X : (d == 1 ? Tile<DIM>::Y : Tile<DIM>::Z);
    }
};
struct ShapeRT {
    static constexpr bool RT = true;
    int e[3];
    __device__ __forceinline__ int t(int d) const { return e[d]; }
};

template <int DIM, int G>
struct LdsRT {
    // 2D: 32 cells across times a pencil of 8 + halo (32 x 8, 8 x 32, 16 x 16, ...); 3D: 64 cells
    // across times a pencil of 4 + halo (8 x 8 x 4, 16 x 4 x 4, ...)
    static constexpr int RMAX = (DIM == 2) ? 32 * (8 + 2 * G) : 64 * (4 + 2 * G);
    static constexpr int XS = (DIM == 2) ? 32 : 64;
};

// LDS layout of a shape policy: RMAX cells of staged primitives per component, the first FS faces'
// fluxes over them, XS more faces' fluxes after them.
template <int DIM, int G, class SH> struct LdsOf;
template <int DIM, int G>
struct LdsOf<DIM, G, ShapeFixed<DIM>> {
    static constexpr int RMAX = Lds<DIM, G>::RMAX, FS = Lds<DIM, G>::FMAX, XS = 0;
};
template <int DIM, int G>
struct LdsOf<DIM, G, ShapeRT> {
    static constexpr int RMAX = LdsRT<DIM, G>::RMAX, FS = NT, XS = LdsRT<DIM, G>::XS;
};
// doubles of a stage workgroup's LDS for N components: the staged region (or the main flux slots,
// N fluxes and u*_n each), then the extra flux slots
template <int N, class LO>
constexpr int lds_doubles()
{
    return ((N * LO::RMAX > (N + 1) * LO::FS) ? N * LO::RMAX : (N + 1) * LO::FS) + (N + 1) * LO::XS;
}

// (x, y, z) of index r in a block of e0 x e1 x * positions. Run-time extents: the quotients by a
// float reciprocal (rcp0 = 1/e0, rcp01 = 1/(e0 e1)): the exact quotient (r + 1/2)/e lies at least
// 1/(2e) from an integer and the float product errs by less than 3e-7 of it, so truncation gives
// the integer quotient for every r below 10^6.
template <class SH>
__device__ __forceinline__ void split3(int r, int e0, int e1, float rcp0, float rcp01, int& x, int& y, int& z)
{
    if constexpr (SH::RT) {
        const float rh = (float)r + 0.5f;
        const int q = (int)(rh * rcp0);
        z = (int)(rh * rcp01);
        x = r - q * e0;
        y = q - z * e1;
    } else {
        x = r % e0;
        y = (r / e0) % e1;
        z = r / (e0 * e1);
    }
}

// Maps a workgroup to a tile so that the workgroups one XCD receives (blocks b, b+8, b+16, ...
// under round-robin dispatch) hold consecutive tiles, whose halos overlap in that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// What one workgroup's direction passes share.
template <int DIM>
struct StageCtx {
    const KParams& kp;
    int limiter;
    const BoxDev& bx;
    const double* __restrict__ Ub;
    int tlo[3];
    int n[3];
    int tid, cx, cy, cz;
    bool valid;
    const double* __restrict__ gq;   // Grad_Q face buffer (Phase-Field flux only)
    // fused ghost fill (fd non-null, in device memory): cells outside the box are read where the
    // ghost fill would take them from (Ul: the level buffer), so the ghost cells of the stage's
    // input are never read
    const FillDev* __restrict__ fd;
    const double* __restrict__ Ul;
    const BoxDev* __restrict__ boxes;
    int halo[3][2];                  // the tile's halo sources (HaloSrc), by direction and side
};

// map_cell / fill_source: where the ghost fill takes cell (ci, cj, ck) of a level from
// (FillPatchSingleLevel and the physical BCs): the cell itself wrapped on periodic sides; on the others foextrap / first-order extrapolation
// (codes 0, 2) clamps to the domain, the reflecting codes (3 symmetry, 4 slip wall, 5 no-slip wall)
// mirror it and flip the normal momentum (all momenta for code 5). si: the source cell's index in the
// level buffer, ssn: its box's component stride, flip: bit m set = negate momentum m. Returns 0, or
// CMP_ERR_BC (a code without a fill here) / CMP_ERR_LAYOUT (no box owns the source cell).
template <int DIM>
__device__ __forceinline__ int map_cell(const FillDev& fd, int* src, unsigned& flip)
{
    flip = 0u;
    int ok = 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const int dlo = fd.dlo[d], dhi = fd.dhi[d];
        if (src[d] >= dlo && src[d] <= dhi) continue;
        if (fd.periodic[d]) {
            const int per = dhi - dlo + 1;
            src[d] += (src[d] < dlo) ? per : -per;
            continue;
        }
        const int code = (src[d] < dlo) ? fd.lo_bc[d] : fd.hi_bc[d];
        if (code == 0 || code == 2) {
            src[d] = (src[d] < dlo) ? dlo : dhi;
        } else if (code >= 3 && code <= 5) {
            src[d] = (src[d] < dlo) ? (2 * dlo - 1 - src[d]) : (2 * dhi + 1 - src[d]);
            for (int m = 0; m < DIM; ++m)
                if (m == d || code == 5) flip ^= (1u << m);
        } else {
            ok = 0;
        }
    }
    if (!ok) return CMP_ERR_BC;
#pragma unroll
    for (int d = 0; d < DIM; ++d)
        if (src[d] < fd.dlo[d] || src[d] > fd.dhi[d]) return CMP_ERR_LAYOUT;
    return 0;
}

// The box of the level owning domain cell src (-1: none).
template <int DIM>
__device__ __forceinline__ int cell_owner(const FillDev& fd, const int* src)
{
    int bi = 0;
    for (int d = DIM - 1; d >= 0; --d)
        bi = bi * fd.nb[d] + (src[d] - fd.dlo[d]) / fd.bs;
    return fd.owner[bi];
}

template <int DIM>
__device__ __forceinline__ int fill_source(const FillDev& fd, int ci, int cj, int ck, long long& si, long long& ssn,
                                           unsigned& flip)
{
    int src[3] = {ci, cj, ck};
    const int e = map_cell<DIM>(fd, src, flip);
    if (e) return e;
    const int own = cell_owner<DIM>(fd, src);
    if (own < 0) return CMP_ERR_LAYOUT;
    const BoxDev so = fd.boxes[own];
    si = so.off + fab_index(so, src[0], src[1], src[2]);
    ssn = so.sn;
    return 0;
}

// fill_source for the stage kernel: out of line (only the tiles on box faces take it, and inlined
// at every read site it would crowd the face computation's registers), the table read from device
// memory.
struct GhostSrc {
    long long si, sn;
    unsigned flip;
    int err;
};
template <int DIM>
__device__ __noinline__ GhostSrc ghost_source(const FillDev* __restrict__ fd, int ci, int cj, int ck)
{
    GhostSrc g;
    g.err = fill_source<DIM>(*fd, ci, cj, ck, g.si, g.sn, g.flip);
    return g;
}

// A cell outside the tile's box, read as the ghost fill would write it (the box's own stale ghost
// value when the fill has no source for it; the error is then reported as the fill reports it).
template <int DIM, int N>
__device__ __forceinline__ void read_ghost(const StageCtx<DIM>& cx_, int i, int j, int k, long long c, double* Uc,
                                           int& bad)
{
    constexpr int MOM = (N == 4 + DIM) ? 2 : (N - DIM - 1) / 2;
    const GhostSrc g = ghost_source<DIM>(cx_.fd, i, j, k);
    if (g.err) {
        bad = g.err;
#pragma unroll
        for (int q = 0; q < N; ++q) Uc[q] = cx_.Ub[c + q * cx_.bx.sn];
        return;
    }
#pragma unroll
    for (int q = 0; q < N; ++q) {
        const double v = cx_.Ul[g.si + q * g.sn];
        Uc[q] = (q >= MOM && q < MOM + DIM && ((g.flip >> (q - MOM)) & 1u)) ? -v : v;
    }
}

// A cell beyond the tile's box along D (side 0 low, 1 high) by the tile's halo source: a
// neighbour box's cell (periodic image along D), the own box's boundary cell (clamp) or mirror
// image (normal or all momenta negated), or the cell-by-cell lookup.
template <int DIM, int N, int D>
__device__ __forceinline__ void read_halo(const StageCtx<DIM>& cx_, int side, int i, int j, int k, long long c,
                                          double* Uc, int& bad)
{
    constexpr int MOM = (N == 4 + DIM) ? 2 : (N - DIM - 1) / 2;
    const int code = cx_.halo[D][side];
    int s[3] = {i, j, k};
    const int dlo = cx_.fd->dlo[D], dhi = cx_.fd->dhi[D];
    if (code >= 0) {
        const BoxDev& nb = cx_.boxes[code];
        if (s[D] < dlo) s[D] += dhi - dlo + 1;
        else if (s[D] > dhi) s[D] -= dhi - dlo + 1;
        const double* __restrict__ src = cx_.Ul + nb.off + fab_index(nb, s[0], s[1], s[2]);
        const long long sn = nb.sn;
#pragma unroll
        for (int q = 0; q < N; ++q) Uc[q] = src[q * sn];
    } else if (code == HALO_LOOKUP) {
        read_ghost<DIM, N>(cx_, i, j, k, c, Uc, bad);
    } else {
        if (code == HALO_CLAMP) s[D] = (s[D] < dlo) ? dlo : dhi;
        else s[D] = (s[D] < dlo) ? 2 * dlo - 1 - s[D] : 2 * dhi + 1 - s[D];
        const BoxDev& bx = cx_.bx;
        const long long cc = fab_index(bx, s[0], s[1], s[2]);
#pragma unroll
        for (int q = 0; q < N; ++q) {
            const double v = cx_.Ub[cc + q * bx.sn];
            const bool neg = (code == HALO_MIRROR && q == MOM + D) ||
                             (code == HALO_MIRROR_ALL && q >= MOM && q < MOM + DIM);
            Uc[q] = neg ? -v : v;
        }
    }
}

// One direction of the stage kernel (D known at compile time): stage the primitives of the tile
// plus G cells on both sides along D, compute every face along D once, and accumulate the flux and
// u*_n differences of the thread's cell. sh: the tile shape; sW / sF: the staged primitives and,
// over them, the fluxes of the first NT faces; sX: the fluxes of the faces beyond NT.
template <int DIM, int SCHEME, int RIEMANN, bool THINC, int D, int NP, class SH>
__device__ __forceinline__ void dir_pass(const StageCtx<DIM>& cx_, const SH& sh, double* sW, double* sF, double* sX,
                                         double* acc, double& divu, double& cmax_loc, int& bad)
{
    constexpr int N = NP ? 2 * NP + DIM + 1 : 4 + DIM;
    constexpr int G = Stencil<SCHEME>::G;
    using LO = LdsOf<DIM, G, SH>;
    constexpr int RMAX = LO::RMAX, FS = LO::FS, XS = LO::XS;
    const KParams& kp = cx_.kp;
    const BoxDev& bx = cx_.bx;
    const int tid = cx_.tid;
    // Region (tile plus G cells on both sides along D) and face index spaces have the full tile's
    // extents (for the fixed shape known at compile time: divisions by constants, constant LDS
    // strides); a tile clipped by its box leaves the positions beyond its cells unused.
    const int E0 = sh.t(0) + (D == 0 ? 2 * G : 0);
    const int E1 = sh.t(1) + (D == 1 ? 2 * G : 0);
    const int E2 = sh.t(2) + (D == 2 ? 2 * G : 0);
    const int R = E0 * E1 * E2;
    const int F0 = sh.t(0) + (D == 0 ? 1 : 0);
    const int F1 = sh.t(1) + (D == 1 ? 1 : 0);
    const int F2 = sh.t(2) + (D == 2 ? 1 : 0);
    const int F = F0 * F1 * F2;
    float rE0 = 0.f, rE01 = 0.f, rF0 = 0.f, rF01 = 0.f;
    if constexpr (SH::RT) {
        rE0 = __builtin_amdgcn_rcpf((float)E0);
        rE01 = __builtin_amdgcn_rcpf((float)(E0 * E1));
        rF0 = __builtin_amdgcn_rcpf((float)F0);
        rF01 = __builtin_amdgcn_rcpf((float)(F0 * F1));
    } else {
        static_assert(SH::RT || XS == 0, "a fixed shape has no second face pass");
    }
    // extents actually used by this tile
    const int e0 = cx_.n[0] + (D == 0 ? 2 * G : 0);
    const int e1 = cx_.n[1] + (D == 1 ? 2 * G : 0);
    const int e2 = cx_.n[2] + (D == 2 ? 2 * G : 0);
    const int f0 = cx_.n[0] + (D == 0 ? 1 : 0);
    const int f1 = cx_.n[1] + (D == 1 ? 1 : 0);
    const int f2 = cx_.n[2] + (D == 2 ? 1 : 0);
    // Phase-Field flux: a face with Grad_Q == 0 has the flux 0 exactly (every term carries the
    // factor Grad_Q), so its reconstruction is skipped, and a tile direction without any nonzero
    // Grad_Q skips the staging as well (its flux differences are all zero)
    // Grad_Q has one component per volume fraction (1, or NP for the N-phase model), each the
    // box's whole face block of direction D
    constexpr int NG = NP ? NP : 1;
    const int gnv0 = bx.hi[0] - bx.lo[0] + 1 + (D == 0 ? 1 : 0);
    const int gnv1 = bx.hi[1] - bx.lo[1] + 1 + (D == 1 ? 1 : 0);
    const int gnv2 = (DIM == 3) ? bx.hi[2] - bx.lo[2] + 1 + (D == 2 ? 1 : 0) : 1;
    const long long gcs = (long long)gnv0 * gnv1 * gnv2;
    auto gface = [&](int x, int y, int z, int m) -> double {
        const long long gi = bx.goff[D] + (long long)(cx_.tlo[0] + x - bx.lo[0])
                           + (long long)gnv0 * ((cx_.tlo[1] + y - bx.lo[1])
                                                + (long long)gnv1 * (cx_.tlo[2] + z - bx.lo[2]));
        return cx_.gq[gi + m * gcs];
    };
    if constexpr (RIEMANN >= RIEMANN_PF_MIX) {
        int nz = 0;
        for (int f = tid; f < F; f += NT) {
            int x, y, z;
            split3<SH>(f, F0, F1, rF0, rF01, x, y, z);
            if (x < f0 && y < f1 && z < f2)
#pragma unroll
                for (int m = 0; m < NG; ++m) nz |= (gface(x, y, z, m) != 0.0);
        }
        if (!__syncthreads_or(nz)) return;
    }
    // 1. primitives of the region, staged in LDS
    for (int r = tid; r < R; r += NT) {
        int x, y, z;
        split3<SH>(r, E0, E1, rE0, rE01, x, y, z);
        if (x >= e0 || y >= e1 || z >= e2) continue;
        const int i = cx_.tlo[0] + x - (D == 0 ? G : 0);
        const int j = cx_.tlo[1] + y - (D == 1 ? G : 0);
        const int k = cx_.tlo[2] + z - (D == 2 ? G : 0);
        const long long c = fab_index(bx, i, j, k);
        double Uc[N], Wc[N];
        const int cd = (D == 0) ? i : ((D == 1) ? j : k);
        if (cx_.fd && (cd < bx.lo[D] || cd > bx.hi[D])) {
            read_halo<DIM, N, D>(cx_, cd > bx.hi[D] ? 1 : 0, i, j, k, c, Uc, bad);
        } else {
#pragma unroll
            for (int q = 0; q < N; ++q) Uc[q] = cx_.Ub[c + q * bx.sn];
        }
        if constexpr (NP > 0)
            np_cons2prim<DIM, NP>(kp, Uc, Wc, bad);
        else
            cons2prim<DIM>(kp, Uc, Wc, bad);
#pragma unroll
        for (int q = 0; q < N; ++q) sW[q * RMAX + r] = Wc[q];
    }
    __syncthreads();
    // 2. faces along D: one face per thread and pass (the fixed shape has at most NT faces per
    // direction, a run-time shape at most NT + XS). A face of the second pass (index NT + tid) is
    // computed first and its flux stored in its own slots right away; the face of the first pass is
    // computed into registers while the staged primitives are read, and after a barrier its flux is
    // written over the staged region, whose LDS the first NT flux slots share.
    const int rs = (D == 0) ? 1 : ((D == 1) ? E0 : E0 * E1);
    const int npass = SH::RT ? (F > NT ? 2 : 1) : 1;
    double Fl[N];
    double un = 0.0;
    bool has_face = false;
#pragma unroll 1
    for (int pass = npass - 1; pass >= 0; --pass) {
    const int fi = pass * NT + tid;
    int x, y, z;
    split3<SH>(fi, F0, F1, rF0, rF01, x, y, z);
    has_face = (fi < F) && (x < f0) && (y < f1) && (z < f2);
    un = 0.0;
    if (has_face) {
        const int a = (D == 0) ? x : ((D == 1) ? y : z);
        const int base = x + E0 * (y + E1 * z);
        auto st = [&](int t, int q) -> double { return sW[q * RMAX + base + t * rs]; };
        if constexpr (RIEMANN >= RIEMANN_PF_MIX) {
            double g[NG];
            bool any = false;
#pragma unroll
            for (int m = 0; m < NG; ++m) {
                g[m] = gface(x, y, z, m);
                any = any || (g[m] != 0.0);
            }
            if (any) {
                if constexpr (NP > 0)
                    np_pf_face_flux<DIM, SCHEME, RIEMANN - RIEMANN_PF_MIX, (NP > 0 ? NP : 2)>(kp, cx_.limiter, D,
                                                                                          st, g, Fl, bad);
                else
                    pf_face_flux<DIM, SCHEME, RIEMANN - RIEMANN_PF_MIX>(kp, cx_.limiter, D, st, g[0], Fl, bad);
            } else {
#pragma unroll
                for (int q = 0; q < N; ++q) Fl[q] = 0.0;
            }
        } else if constexpr (NP > 0) {
            double c;
            un = np_face_flux<DIM, SCHEME, RIEMANN, NP>(kp, cx_.limiter, D, st, Fl, c, bad);
            if (a < cx_.n[D]) {
                if (__builtin_isnan(c)) bad = CMP_ERR_NAN;
                cmax_loc = smax(cmax_loc, c);
            }
        } else {
            double aL = -1.0, aR = -1.0;
            if constexpr (THINC) {
                // THINC decision (gates, face normal, tanh profile) ahead of the base reconstruction.
                // The face normal is needed only where the volume-fraction gates admit THINC, so its
                // alpha1 stencil (the face's two cells and their transverse neighbours, transverse
                // axes in ascending order) is read from the state itself, which the staging pass has
                // just brought into L2, instead of being staged for every face.
                const int hi_ = cx_.tlo[0] + x, hj = cx_.tlo[1] + y, hk = cx_.tlo[2] + ((DIM == 3) ? z : 0);
                auto aa = [&](int od, int o1, int o2) -> double {
                    const int lo = (od == -1) ? 1 : 0;
                    int pi = hi_ - (D == 0 ? lo : 0), pj = hj - (D == 1 ? lo : 0), pk = hk - (D == 2 ? lo : 0);
                    if (D == 0) { pj += o1; pk += (DIM == 3) ? o2 : 0; }
                    else if (D == 1) { pi += o1; pk += (DIM == 3) ? o2 : 0; }
                    else { pi += o1; pj += o2; }
                    if (cx_.fd && (pi < bx.lo[0] || pi > bx.hi[0] || pj < bx.lo[1] || pj > bx.hi[1] ||
                                   (DIM == 3 && (pk < bx.lo[2] || pk > bx.hi[2])))) {
                        const GhostSrc gs = ghost_source<DIM>(cx_.fd, pi, pj, pk);
                        if (gs.err == 0) return cx_.Ul[gs.si + (3 + DIM) * gs.sn];
                        bad = gs.err;
                    }
                    return cx_.Ub[fab_index(bx, pi, pj, pk) + (3 + DIM) * bx.sn];
                };
                auto normal = [&]() -> double { return face_normal_component<DIM>(kp, D, aa); };
                thinc_decide<DIM, G>(kp, st, normal, aL, aR);
            }
            double c;
            un = face_flux_decided<DIM, SCHEME, RIEMANN, THINC>(kp, cx_.limiter, D, st, aL, aR, Fl, c, bad);
            if (a < cx_.n[D]) {
                if (__builtin_isnan(c)) bad = CMP_ERR_NAN;
                cmax_loc = smax(cmax_loc, c);
            }
        }
    }
    if (pass > 0 && has_face) {
        if constexpr (XS > 0) {
#pragma unroll
            for (int q = 0; q < N; ++q) sX[q * XS + tid] = Fl[q];
            if constexpr (RIEMANN < RIEMANN_PF_MIX) sX[N * XS + tid] = un;
        }
    }
    }
    __syncthreads();
    if (has_face) {
#pragma unroll
        for (int q = 0; q < N; ++q) sF[q * FS + tid] = Fl[q];
        if constexpr (RIEMANN < RIEMANN_PF_MIX) sF[N * FS + tid] = un;
    }
    __syncthreads();
    // 3. flux differences of this direction
    auto flux = [&](int q, int f) -> double {
        if constexpr (XS > 0)
            return (f < NT) ? sF[q * FS + f] : sX[q * XS + (f - NT)];
        else
            return sF[q * FS + f];
    };
    if (cx_.valid) {
        const int fm = cx_.cx + F0 * (cx_.cy + F1 * cx_.cz);
        const int fp = fm + ((D == 0) ? 1 : ((D == 1) ? F0 : F0 * F1));
        const double idx = kp.inv_dx[D];
#pragma unroll
        for (int q = 0; q < N; ++q) {
            const double t = (flux(q, fp) - flux(q, fm)) * idx;
            acc[q] = (D == 0) ? t : acc[q] + t;
        }
        if constexpr (RIEMANN < RIEMANN_PF_MIX) {
            const double t = (flux(N, fp) - flux(N, fm)) * idx;
            divu = (D == 0) ? t : divu + t;
        }
    }
    __syncthreads();
}

// Minimum resident waves per SIMD the register allocator must leave room for (0: compiler's choice).
#ifndef CMP_STAGE_WAVES
#define CMP_STAGE_WAVES 2
#endif
#if CMP_STAGE_WAVES > 0
#define CMP_STAGE_ATTR __attribute__((amdgpu_waves_per_eu(CMP_STAGE_WAVES)))
#else
#define CMP_STAGE_ATTR
#endif

// The stage of one tile (box bx, lowest cell tlo, shape sh, halo sources halo): staging, faces and
// flux differences per direction, then the cell update of `mode`; c_max max-accumulated into
// *cmax_bits, the error word into *err. sWF / sRed: the workgroup's LDS (sWF: lds_doubles of the
// shape policy's layout).
template <int DIM, int SCHEME, int RIEMANN, bool THINC, int NP, class SH>
__device__ __forceinline__ void stage_tile(const KParams& kp, int limiter, int mode, int stage, const BoxDev& bx,
                                           const int* tlo, const SH& sh, const int (&halo)[3][2],
                                           const BoxDev* __restrict__ boxes,
                                           const double* __restrict__ Us, const double* Un, double* Uout,
                                           const double* __restrict__ dtp, unsigned long long* cmax_bits, int* err,
                                           const double* __restrict__ gq, const FillDev* __restrict__ fd,
                                           double* sWF, double* sRed, int tid = -1)
{
    constexpr int N = NP ? 2 * NP + DIM + 1 : 4 + DIM;
    constexpr int IA = 3 + DIM;
    constexpr int G = Stencil<SCHEME>::G;
    using LO = LdsOf<DIM, G, SH>;
    double* const sW = sWF;
    double* const sF = sWF;
    double* const sX = sWF + (lds_doubles<N, LO>() - (N + 1) * LO::XS);

    if (tid < 0) tid = threadIdx.x;
    const int TX = sh.t(0), TY = sh.t(1), TZ = sh.t(2);
    int n[3];
    n[0] = min(TX, bx.hi[0] - tlo[0] + 1);
    n[1] = min(TY, bx.hi[1] - tlo[1] + 1);
    n[2] = (DIM == 3) ? min(TZ, bx.hi[2] - tlo[2] + 1) : 1;

    int cx, cy, cz;
    if constexpr (SH::RT)
        split3<SH>(tid, TX, TY, __builtin_amdgcn_rcpf((float)TX), __builtin_amdgcn_rcpf((float)(TX * TY)), cx, cy, cz);
    else
        split3<SH>(tid, TX, TY, 0.f, 0.f, cx, cy, cz);
    const bool valid = (cx < n[0]) && (cy < n[1]) && (cz < n[2]);
    const double* __restrict__ Ub = Us + bx.off;

    int bad = 0;
    double cmax_loc = 0.0;
    double acc[N];
    double divu = 0.0;
#pragma unroll
    for (int q = 0; q < N; ++q) acc[q] = 0.0;

    StageCtx<DIM> ctx{kp, limiter, bx, Ub, {tlo[0], tlo[1], tlo[2]}, {n[0], n[1], n[2]}, tid, cx, cy, cz,
                      valid, gq, fd, Us, boxes,
                      {{halo[0][0], halo[0][1]}, {halo[1][0], halo[1][1]}, {halo[2][0], halo[2][1]}}};
    dir_pass<DIM, SCHEME, RIEMANN, THINC, 0, NP>(ctx, sh, sW, sF, sX, acc, divu, cmax_loc, bad);
    if (DIM > 1) dir_pass<DIM, SCHEME, RIEMANN, THINC, 1, NP>(ctx, sh, sW, sF, sX, acc, divu, cmax_loc, bad);
    if (DIM > 2)
        dir_pass<DIM, SCHEME, RIEMANN, THINC, (DIM > 2 ? 2 : 1), NP>(ctx, sh, sW, sF, sX, acc, divu, cmax_loc, bad);

    if (valid) {
        const int i = tlo[0] + cx, j = tlo[1] + cy, k = tlo[2] + cz;
        const long long c = fab_index(bx, i, j, k);
        double Uc[N], dU[N];
        if constexpr (RIEMANN >= RIEMANN_PF_MIX) {
            // faces hold fhat and the reference stores -fhat: -div(-fhat) = +div(fhat); the
            // five-equation Phase-Field NC term is empty
#pragma unroll
            for (int q = 0; q < N; ++q) {
                Uc[q] = Ub[c + q * bx.sn];
                dU[q] = acc[q];
            }
        } else if constexpr (NP > 0) {
            // every alpha_k takes phi_k (1 + K_k) div u* (…FVM_SurfaceIntegralNC_Compressible5Eq_NPhase.H)
#pragma unroll
            for (int q = 0; q < N; ++q) {
                Uc[q] = Ub[c + q * bx.sn];
                dU[q] = -acc[q];
            }
            double f[NP > 0 ? NP : 1];
            np_nc_coefficients<DIM, NP>(kp, Uc, f, bad);
#pragma unroll
            for (int k = 0; k < NP; ++k) dU[NpIdx<DIM, NP>::A + k] += f[k] * divu;
        } else {
#pragma unroll
            for (int q = 0; q < N; ++q) {
                Uc[q] = Ub[c + q * bx.sn];
                dU[q] = -acc[q];
            }
            dU[IA] += nc_coefficient<DIM>(kp, Uc, bad) * divu;
        }
        if (mode == MODE_RK) {
            const double dt = *dtp;
            double* Ob = Uout + bx.off;
            const double* Nb = Un + bx.off;
            if (stage == 0) {
#pragma unroll
                for (int q = 0; q < N; ++q) {
                    const double v = Uc[q] + dt * dU[q];
                    if (__builtin_isnan(v)) bad = CMP_ERR_NAN;
                    Ob[c + q * bx.sn] = v;
                }
            } else {
                const double a = (stage == 1) ? 0.25 : (2.0 / 3.0);
                const double b = (stage == 1) ? 0.75 : (1.0 / 3.0);
                const double w = (stage == 1) ? 0.25 * dt : (2.0 / 3.0) * dt;
#pragma unroll
                for (int q = 0; q < N; ++q) {
                    const double v = (a * Uc[q] + b * Nb[c + q * bx.sn]) + w * dU[q];
                    if (__builtin_isnan(v)) bad = CMP_ERR_NAN;
                    Ob[c + q * bx.sn] = v;
                }
            }
            if constexpr (SCHEME == 6 && NP == 0 && RIEMANN < RIEMANN_PF_MIX) {
                // the bounded limiter keeps alpha1 in [0, 1]; a stage state beyond it (by more than
                // rounding) is reported
                const double an = Ob[c + IA * bx.sn];
                if (an < -1.0e-12 || an > 1.0 + 1.0e-12) bad = CMP_ERR_ALPHA;
            }
        } else if (mode == MODE_RHS) {
            const long long o = bx.ooff + (long long)(i - bx.olo[0]) + bx.osj * (j - bx.olo[1]) + bx.osk * (k - bx.olo[2]);
#pragma unroll
            for (int q = 0; q < N; ++q) Uout[o + q * bx.osn] = dU[q];
        }
    }

    if (cmax_bits) {
        const double m = wave_max(cmax_loc);
        if ((tid & 63) == 0) sRed[tid >> 6] = m;
        __syncthreads();
        if (tid == 0) {
            double v = sRed[0];
#pragma unroll
            for (int w = 1; w < NT / 64; ++w) v = smax(v, sRed[w]);
            // non-negative doubles order like their bit patterns; skip the atomic when the
            // stored maximum is already as large
            const unsigned long long vb = (unsigned long long)__double_as_longlong(v);
            if (vb > __hip_atomic_load(cmax_bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                atomicMax(cmax_bits, vb);
        }
    }
    if (bad && err) atomicMax(err, bad);
}

template <int DIM, int SCHEME, int RIEMANN, bool THINC, int NP>
__global__ __launch_bounds__(NT) CMP_STAGE_ATTR
void k_stage(KParams kp, int limiter, int mode, int stage,
             const TileDev* __restrict__ tiles, int ntiles, const BoxDev* __restrict__ boxes,
             BoxDev sbox, int sntx, int snty, int single,
             const double* __restrict__ Us, const double* Un, double* Uout,
             const double* __restrict__ dtp, unsigned long long* cmax_bits, int* err,
             const double* __restrict__ gq, const FillDev* __restrict__ fd)
{
    constexpr int N = NP ? 2 * NP + DIM + 1 : 4 + DIM;
    constexpr int G = Stencil<SCHEME>::G;
    using L = Lds<DIM, G>;
    constexpr int TX = L::TX, TY = L::TY, TZ = L::TZ;
    // the staged primitives of a direction and, once its faces are computed, their fluxes
    __shared__ double sWF[lds_doubles<N, LdsOf<DIM, G, ShapeFixed<DIM>>>()];
    __shared__ double sRed[NT / 64];
    // a step of size zero (a level clock past its stop_time) leaves every state as it is
    if (mode == MODE_RK && *dtp == 0.0) return;

    int tlo[3];
    int halo[3][2] = {{HALO_LOOKUP, HALO_LOOKUP}, {HALO_LOOKUP, HALO_LOOKUP}, {HALO_LOOKUP, HALO_LOOKUP}};
    BoxDev bx;
    if (single) {
        int b = blockIdx.x;
        const int tx = b % sntx;
        b /= sntx;
        const int ty = b % snty;
        const int tz = b / snty;
        bx = sbox;
        tlo[0] = sbox.lo[0] + tx * TX;
        tlo[1] = sbox.lo[1] + ty * TY;
        tlo[2] = sbox.lo[2] + (DIM == 3 ? tz * TZ : 0);
    } else {
        const TileDev td = tiles[xcd_tile(blockIdx.x, ntiles)];
        bx = boxes[td.box];
        tlo[0] = td.lo[0];
        tlo[1] = td.lo[1];
        tlo[2] = td.lo[2];
        if (fd) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                halo[d][0] = td.halo[d][0];
                halo[d][1] = td.halo[d][1];
            }
        }
    }
    stage_tile<DIM, SCHEME, RIEMANN, THINC, NP>(kp, limiter, mode, stage, bx, tlo, ShapeFixed<DIM>{}, halo, boxes, Us,
                                                Un, Uout, dtp, cmax_bits, err, gq, fd, sWF, sRed);
}

// One RK stage of a batch of independent levels: every tile of every admitted case in one launch.
// A box's case is its `pad` field; each case has its own parameters, dt, c_max, error word and
// ghost-fill tables (the halo lookup); the cells beyond a box are read from their source boxes.
template <int DIM, int SCHEME, int RIEMANN>
__global__ __launch_bounds__(NT) CMP_STAGE_ATTR
void k_stage_batch(const KParams* __restrict__ kps, int limiter, int stage, const TileDev* __restrict__ tiles,
                   int ntiles, const BoxDev* __restrict__ boxes, const double* __restrict__ Us, const double* Un,
                   double* Uout, const double* __restrict__ dts, unsigned long long* cmax_bits, int* errs,
                   const FillDev* __restrict__ fds)
{
    constexpr int N = 4 + DIM;
    constexpr int G = Stencil<SCHEME>::G;
    __shared__ double sWF[lds_doubles<N, LdsOf<DIM, G, ShapeFixed<DIM>>>()];
    __shared__ double sRed[NT / 64];

    const TileDev td = tiles[xcd_tile(blockIdx.x, ntiles)];
    const BoxDev bx = boxes[td.box];
    const int cs = bx.pad;
    const int tlo[3] = {td.lo[0], td.lo[1], td.lo[2]};
    const int halo[3][2] = {{td.halo[0][0], td.halo[0][1]}, {td.halo[1][0], td.halo[1][1]},
                            {td.halo[2][0], td.halo[2][1]}};
    stage_tile<DIM, SCHEME, RIEMANN, false, 0>(kps[cs], limiter, MODE_RK, stage, bx, tlo, ShapeFixed<DIM>{}, halo,
                                               boxes, Us, Un, Uout, dts + cs, cmax_bits + cs, errs + cs, nullptr,
                                               fds + cs, sWF, sRed);
}

#ifndef CMP_LIST_UNIT
// The face flux and u*_n of the face between cells (ci, cj, ck) - e_DIR and (ci, cj, ck) of box bx
// (Ub: the box's fab, its ghost cells filled), every stencil cell taken through cons->prim here.
// NP > 0: the N-phase model (no THINC).
template <int DIM, int SCHEME, int RIEMANN, bool THINC, int DIR, int NP = 0>
__device__ __forceinline__ double face_at(const KParams& kp, int limiter, const BoxDev& bx,
                                          const double* __restrict__ Ub, int ci, int cj, int ck, double* Fl,
                                          int& bad, double* cm_out = nullptr)
{
    constexpr int N = (NP > 0) ? 2 * NP + DIM + 1 : 4 + DIM;
    constexpr int IA = 3 + DIM;
    constexpr int G = Stencil<SCHEME>::G;
    const long long sd = (DIR == 0) ? 1 : ((DIR == 1) ? bx.sj : bx.sk);
    const long long c0 = fab_index(bx, ci, cj, ck);
    double W[2 * G][N];
#pragma unroll
    for (int t = 0; t < 2 * G; ++t) {
        double Uc[N];
        const long long c = c0 + (t - G) * sd;
#pragma unroll
        for (int q = 0; q < N; ++q) Uc[q] = Ub[c + q * bx.sn];
        if constexpr (NP > 0) np_cons2prim<DIM, NP>(kp, Uc, W[t], bad);
        else cons2prim<DIM>(kp, Uc, W[t], bad);
    }
    auto st = [&](int t, int q) -> double { return W[t][q]; };
    double cm;
    if constexpr (NP > 0) {
        const double un = np_face_flux<DIM, SCHEME, RIEMANN, NP>(kp, limiter, DIR, st, Fl, cm, bad);
        if (cm_out) *cm_out = cm;
        return un;
    } else {
        auto aa = [&](int od, int o1, int o2) -> double {
            const int lo = (od == -1) ? 1 : 0;
            int px = ci - (DIR == 0 ? lo : 0), py = cj - (DIR == 1 ? lo : 0), pz = ck - (DIR == 2 ? lo : 0);
            if (DIR == 0) { py += o1; pz += (DIM == 3) ? o2 : 0; }
            else if (DIR == 1) { px += o1; pz += (DIM == 3) ? o2 : 0; }
            else { px += o1; py += o2; }
            return Ub[fab_index(bx, px, py, pz) + IA * bx.sn];
        };
        auto normal = [&]() -> double { return face_normal_component<DIM>(kp, DIR, aa); };
        const double un = face_flux<DIM, SCHEME, RIEMANN, THINC>(kp, limiter, DIR, st, normal, Fl, cm, bad);
        if (cm_out) *cm_out = cm;
        return un;
    }
}

// Face fluxes of one box in the reference's shape (one thread per face, every stencil cell taken
// through cons->prim by the thread that uses it), for callers that need the face arrays (flux
// registers, face averaging) and as a second GPU path the fused kernel is checked against. cmc
// (valid cells of the box): the wave speed of the face that is the low face of a cell along DIR,
// written by the DIR 0 launch and max-combined by the later ones (the reference's per-cell c_max).
template <int DIM, int SCHEME, int RIEMANN, bool THINC, int DIR>
__global__ __launch_bounds__(NT)
void k_faces(KParams kp, int limiter, BoxDev bx, const double* __restrict__ U,
             double* __restrict__ flux, double* __restrict__ fluxnc, double* __restrict__ cmc)
{
    constexpr int N = 4 + DIM;
    const int fe0 = bx.hi[0] - bx.lo[0] + 1 + (DIR == 0 ? 1 : 0);
    const int fe1 = bx.hi[1] - bx.lo[1] + 1 + (DIR == 1 ? 1 : 0);
    const int fe2 = (DIM == 3) ? bx.hi[2] - bx.lo[2] + 1 + (DIR == 2 ? 1 : 0) : 1;
    const long long F = (long long)fe0 * fe1 * fe2;
    const long long f = (long long)blockIdx.x * NT + threadIdx.x;
    if (f >= F) return;
    int x, y, z;
    pos3(f, fe0, fe1, x, y, z);
    int bad = 0;
    double Fl[N];
    double cm = 0.0;
    const double un = face_at<DIM, SCHEME, RIEMANN, THINC, DIR>(kp, limiter, bx, U + bx.off, bx.lo[0] + x,
                                                                bx.lo[1] + y, bx.lo[2] + z, Fl, bad, &cm);
    if (flux) {
#pragma unroll
        for (int q = 0; q < N; ++q) flux[f + q * F] = Fl[q];
    }
    if (fluxnc) fluxnc[f] = un;
    if (cmc) {
        const int n0 = bx.hi[0] - bx.lo[0] + 1, n1 = bx.hi[1] - bx.lo[1] + 1;
        const int co = (DIR == 0) ? x : ((DIR == 1) ? y : z);
        const int nd = (DIR == 0) ? n0 : ((DIR == 1) ? n1 : bx.hi[2] - bx.lo[2] + 1);
        if (co < nd) {
            const long long c = (long long)x + (long long)n0 * (y + (long long)n1 * z);
            cmc[c] = (DIR == 0) ? cm : ((cmc[c] < cm) ? cm : cmc[c]);
        }
    }
}

// Ghost cell g (0 <= g < the box's ghost count) of a box with nv1 valid rows and fab extents
// ex, ey, ez (ng ghosts a side), as fab coordinates: in 3D first the two z slabs (whole xy planes),
// then per interior z plane the y-low rows, the y-high rows and the x ghosts of the interior rows.
template <int DIM>
__device__ __forceinline__ void ghost_xyz(int g, int ng, int ex, int ey, int ez, int nv1, int& x, int& y, int& z)
{
    x = 0;
    y = 0;
    z = 0;
    int r = g;
    bool slab = false;
    if (DIM == 3) {
        const int exy = ex * ey;
        const int A = ng * exy;
        if (r < 2 * A) {
            // z-low (r < A) or z-high slab: whole xy planes
            const bool high = r >= A;
            if (high) r -= A;
            z = (high ? ez - ng : 0) + r / exy;
            r -= (r / exy) * exy;
            y = r / ex;
            x = r - y * ex;
            slab = true;
        } else {
            r -= 2 * A;
            const int Rg = 2 * ng * ex + 2 * ng * nv1;
            const int q = r / Rg;
            z = ng + q;
            r -= q * Rg;
        }
    }
    if (!slab) {
        if (r < ng * ex) {
            y = r / ex;
            x = r - y * ex;
        } else if (r < 2 * ng * ex) {
            r -= ng * ex;
            const int q = r / ex;
            y = ey - ng + q;
            x = r - q * ex;
        } else {
            r -= 2 * ng * ex;
            const int q = r / (2 * ng);
            y = ng + q;
            const int cc = r - q * (2 * ng);
            x = (cc < ng) ? cc : (ex - 2 * ng + cc);
        }
    }
}

// Same-level ghost fill. Thread g handles ghost cell g of the level: the box comes from a binary
// search over ghost-count prefixes, the cell from the ghost shell's slab decomposition; the source
// is mapped through the physical BCs (foextrap clamp, reflect mirror with odd momentum signs) and
// periodic wrap, and read from the valid cells of the box owning it.
template <int DIM, int N>
__global__ __launch_bounds__(NT)
void k_fill(FillDev fd, const int* __restrict__ list, int box0, double* U, int* err)
{
    // first momentum component of a flow state (two-phase: after 2 masses; N-phase: after NP
    // masses, N = 2 NP + DIM + 1); a Phase-Field field (1 or DIM comps) has none
    constexpr int MOM = (N == 4 + DIM) ? 2 : ((N > 4 + DIM) ? (N - DIM - 1) / 2 : -1);
    // one grid row per box (of the level, or of `list`); 32-bit ghost indices inside the box
    const int b = list ? list[box0 + (int)blockIdx.y] : box0 + (int)blockIdx.y;
    const BoxDev bx = fd.boxes[b];
    const int ng = fd.ng;
    const int nv0 = bx.hi[0] - bx.lo[0] + 1, nv1 = bx.hi[1] - bx.lo[1] + 1;
    const int nv2 = (DIM == 3) ? bx.hi[2] - bx.lo[2] + 1 : 1;
    const int ex = nv0 + 2 * ng, ey = nv1 + 2 * ng;
    const int ez = (DIM == 3) ? nv2 + 2 * ng : 1;
    const int nghost = ex * ey * ez - nv0 * nv1 * nv2;
    for (int g = (int)(blockIdx.x * NT + threadIdx.x); g < nghost; g += (int)(gridDim.x * NT)) {
        int x, y, z;
        ghost_xyz<DIM>(g, ng, ex, ey, ez, nv1, x, y, z);
        const int cell[3] = {bx.flo[0] + x, bx.flo[1] + y, bx.flo[2] + z};
        long long si, ssn;
        unsigned flip;
        const int e = fill_source<DIM>(fd, cell[0], cell[1], cell[2], si, ssn, flip);
        if (e) { atomicMax(err, e); continue; }
        const long long di = bx.off + fab_index(bx, cell[0], cell[1], cell[2]);
#pragma unroll
        for (int c = 0; c < N; ++c) {
            double v = U[si + c * ssn];
            if (MOM >= 0 && c >= MOM && c < MOM + DIM && ((flip >> (c - MOM)) & 1u)) v = -v;
            U[di + c * bx.sn] = v;
        }
    }
}

// ComputeDt on the device (source/COMPAS.cpp:2250-2302, EstTimeStep
// source/Class_Compressible_PhaseField.cpp:4473-4499). One thread. step[0] counts the steps of the
// current advance call, step[1] is the capacity of hist.
template <int DIM>
// clock (or null): [0] the level's time, advanced by every dt; [1] stop_time (negative: none), at
// which the dt is cut so that the step ends on it (source/COMPAS.cpp:2278-2283: when
// t + dt > stop_time - 1e-3 dt, dt = stop_time - t). Once the time has reached stop_time the
// reference's step loop ends (Evolve's `cur_time < stop_time`); here every later step has dt = 0,
// its stage kernels return without writing, and c_max keeps the last real step's value.
__global__ void k_compute_dt(double dx0, double dx1, double dx2, double* cmax, double* dt, double cfl,
                             int limiter, double* hist, int* step, int* err, double* clock)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (clock && clock[1] >= 0.0 && clock[0] >= clock[1]) {
        *dt = 0.0;
        const int s0 = step[0];
        if (hist && s0 >= 0 && s0 < step[1]) hist[s0] = 0.0;
        step[0] = s0 + 1;
        return;
    }
    const double cm = *cmax;
    const double dxs[3] = {dx0, dx1, dx2};
    double est = 1.7976931348623157e308;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const double e = cfl * dxs[d] / cm;
        est = (e < est) ? e : est;
    }
    const double prev = *dt;
    double dtn = est;
    if (limiter && prev > 0.0) dtn = smin(est, 1.1 * prev);
    if (!(cm > 0.0) || __builtin_isnan(dtn)) atomicMax(err, CMP_ERR_NAN);
    if (clock) {
        const double t = clock[0], stop = clock[1];
        if (stop >= 0.0) {
            const double eps = 1.0e-3 * dtn;
            if (t + dtn > stop - eps) dtn = stop - t;
            if (!(dtn > 0.0)) dtn = 0.0;
        }
        clock[0] = t + dtn;
    }
    *dt = dtn;
    const int s = step[0];
    if (hist && s >= 0 && s < step[1]) hist[s] = dtn;
    step[0] = s + 1;
    *cmax = 0.0;
}

#if CMP_KERNEL_DIM == 2
// ComputeDt of every case of a batch (one thread per case; k_compute_dt's rule), then the case's
// time and step count advance by it and its c_max restarts at zero. Inactive slots are left alone.
__global__ void k_batch_dt(const BatchSlot* __restrict__ slots, int nslot, double* cmax, double* dt,
                           double* time, int* steps, int* errs)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nslot || !slots[c].active) return;
    const BatchSlot sl = slots[c];
    const double cm = cmax[c];
    double est = 1.7976931348623157e308;
    for (int d = 0; d < sl.dim; ++d) {
        const double e = sl.cfl * sl.dx[d] / cm;
        est = (e < est) ? e : est;
    }
    const double prev = dt[c];
    double dtn = est;
    if (sl.limiter && prev > 0.0) dtn = smin(est, 1.1 * prev);
    if (!(cm > 0.0) || __builtin_isnan(dtn)) atomicMax(errs + c, CMP_ERR_NAN);
    dt[c] = dtn;
    time[c] += dtn;
    steps[c] += 1;
    cmax[c] = 0.0;
}

hipError_t launch_batch_dt(const BatchSlot* slots, int nslot, double* cmax, double* dt, double* time, int* steps,
                           int* errs, hipStream_t s)
{
    if (nslot <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_dt, dim3((nslot + 63) / 64), dim3(64), 0, s, slots, nslot, cmax, dt, time, steps,
                       errs);
    return hipGetLastError();
}

// ComputeDt over the levels of a hierarchy without subcycling (source/COMPAS.cpp:2250-2302 with
// nsubsteps 1): each level's estimate min_d cfl dx_d / c_max, limited to 1.1 dt_prev when the
// limiter is on and dt_prev > 0 (a first step passes dt_prev = 0), the smallest over the levels;
// every level's c_max then restarts at zero. stepc[0] counts the steps of the current advance call
// (hist[stepc[0]] takes the dt while below the capacity stepc[1]). One thread.
__global__ void k_amr_dt(AmrDt a, double* cmax, double* dt, double cfl, int limiter, double* hist, int* stepc,
                         int* err)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double prev = *dt;
    double dt0 = 0.0;
    int bad = 0;
    for (int l = 0; l < a.nlev; ++l) {
        const double cm = cmax[l];
        if (!(cm > 0.0)) bad = 1;
        double est = 1.7976931348623157e308;
        for (int d = 0; d < a.dim; ++d) {
            const double e = cfl * a.dx[l][d] / cm;
            est = (e < est) ? e : est;
        }
        if (limiter && prev > 0.0) est = smin(est, 1.1 * prev);
        dt0 = (l == 0) ? est : smin(dt0, est);
    }
    if (bad || __builtin_isnan(dt0)) atomicMax(err, CMP_ERR_NAN);
    *dt = dt0;
    const int st = stepc[0];
    if (hist && st >= 0 && st < stepc[1]) hist[st] = dt0;
    stepc[0] = st + 1;
    for (int l = 0; l < a.nlev; ++l) cmax[l] = 0.0;
}

hipError_t launch_amr_dt(const AmrDt& a, double* cmax, double* dt, double cfl, int limiter, double* hist, int* stepc,
                         int* err, hipStream_t s)
{
    hipLaunchKernelGGL(k_amr_dt, dim3(1), dim3(64), 0, s, a, cmax, dt, cfl, limiter, hist, stepc, err);
    return hipGetLastError();
}

// The stage combinations of the integrators other than TVD-RK3 (AMReX LinComb / Saxpy on the valid
// cells of a level): out = (p X + q Y) + (r dt) L, a term present where its array is given (no Y:
// p X + (r dt) L; no L: p X + q Y), in that order and without FMA contraction, as the reference's
// LinComb (dst = a x + b y) then Saxpy (dst += a y) round. One thread per valid cell and component
// of one box per grid row.
__global__ __launch_bounds__(NT)
void k_lincomb(const BoxDev* __restrict__ boxes, int N, double* out, double p, const double* X, double q,
               const double* Y, double r, const double* L, const double* __restrict__ dtp)
{
#pragma clang fp contract(off)
    const BoxDev bx = boxes[blockIdx.y];
    const int n0 = bx.hi[0] - bx.lo[0] + 1, n1 = bx.hi[1] - bx.lo[1] + 1, n2 = bx.hi[2] - bx.lo[2] + 1;
    const long long ncell = (long long)n0 * n1 * n2;
    // a step of size zero (a level clock past its stop_time) leaves every state as it is
    if (L && *dtp == 0.0) return;
    const double rdt = L ? r * (*dtp) : 0.0;
    for (long long t = (long long)blockIdx.x * NT + threadIdx.x; t < ncell * N; t += (long long)gridDim.x * NT) {
        const int c = (int)(t / ncell);
        const long long v = t - (long long)c * ncell;
        int i, j, k;
        pos3(v, n0, n1, i, j, k);
        i += bx.lo[0];
        j += bx.lo[1];
        k += bx.lo[2];
        const long long o = bx.off + fab_index(bx, i, j, k) + c * bx.sn;
        double w;
        if (Y && L) {
            w = p * X[o] + q * Y[o];
            w = w + rdt * L[o];
        } else if (Y) {
            w = p * X[o] + q * Y[o];
        } else {
            w = p * X[o] + rdt * L[o];
        }
        out[o] = w;
    }
}

hipError_t launch_lincomb(const BoxDev* boxes, int nbox, long long maxcells, int N, double* out, double p,
                          const double* X, double q, const double* Y, double r, const double* L, const double* dt,
                          hipStream_t s)
{
    long long cols = (maxcells * N + NT - 1) / NT;
    if (cols <= 0 || nbox <= 0) return hipSuccess;
    if (cols > 4096) cols = 4096;
    for (int b0 = 0; b0 < nbox; b0 += 65535) {
        const int rows = (nbox - b0 < 65535) ? nbox - b0 : 65535;
        hipLaunchKernelGGL(k_lincomb, dim3((unsigned)cols, (unsigned)rows), dim3(NT), 0, s, boxes + b0, N, out, p, X,
                           q, Y, r, L, dt);
    }
    return hipGetLastError();
}

// Sets two ints of a device table from kernel arguments (keeps graph-capturable code free of
// memset nodes and of host-to-device copies).
__global__ void k_set2(int* p, int a, int b)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        p[0] = a;
        p[1] = b;
    }
}

hipError_t launch_set2(int* p, int a, int b, hipStream_t s)
{
    hipLaunchKernelGGL(k_set2, dim3(1), dim3(64), 0, s, p, a, b);
    return hipGetLastError();
}
#endif

// ---- launchers for this dimension ---------------------------------------------------------------

template <int DIM, int SCHEME, int RIEMANN, bool THINC, int NP = 0>
static hipError_t stage4(const KParams& kp, int limiter, int mode, int stage, const TileDev* tiles,
                         int ntiles, const BoxDev* boxes, const BoxDev* sbox, int sntiles,
                         const double* Us, const double* Un, double* Uout, const double* dt,
                         unsigned long long* cmax_bits, int* err, hipStream_t s, const double* gq,
                         const FillDev* fuse)
{
    BoxDev sb{};
    int sntx = 1, snty = 1, single = 0, grid = ntiles;
    if (sbox) {
        sb = *sbox;
        single = 1;
        sntx = (sb.hi[0] - sb.lo[0] + Tile<DIM>::X) / Tile<DIM>::X;
        snty = (sb.hi[1] - sb.lo[1] + Tile<DIM>::Y) / Tile<DIM>::Y;
        grid = sntiles;
    }
    if (grid <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_stage<DIM, SCHEME, RIEMANN, THINC, NP>), dim3(grid), dim3(NT), 0, s, kp, limiter,
                       mode, stage, tiles, ntiles, boxes, sb, sntx, snty, single, Us, Un, Uout, dt,
                       cmax_bits, err, gq, fuse);
    return hipGetLastError();
}

template <int DIM, int SCHEME, int RIEMANN, bool THINC>
static hipError_t faces4(const KParams& kp, int limiter, const BoxDev& box, int dir, const double* U,
                         double* flux, double* fluxnc, double* cmc, hipStream_t s)
{
    long long F = 1;
    for (int m = 0; m < DIM; ++m) F *= (box.hi[m] - box.lo[m] + 1 + (m == dir ? 1 : 0));
    const int grid = (int)((F + NT - 1) / NT);
    if (dir == 0)
        hipLaunchKernelGGL((k_faces<DIM, SCHEME, RIEMANN, THINC, 0>), dim3(grid), dim3(NT), 0, s, kp,
                           limiter, box, U, flux, fluxnc, cmc);
    else if (dir == 1)
        hipLaunchKernelGGL((k_faces<DIM, SCHEME, RIEMANN, THINC, 1>), dim3(grid), dim3(NT), 0, s, kp,
                           limiter, box, U, flux, fluxnc, cmc);
    else
        hipLaunchKernelGGL((k_faces<DIM, SCHEME, RIEMANN, THINC, (DIM == 3 ? 2 : 1)>), dim3(grid),
                           dim3(NT), 0, s, kp, limiter, box, U, flux, fluxnc, cmc);
    return hipGetLastError();
}

// Scheme x Riemann x THINC dispatch shared by the two launchers.
#define CMP_DISPATCH(FN, ...)                                                         \
    do {                                                                              \
        const bool th = lc.thinc != 0;                                                \
        if (lc.riemann == 2) {                                                        \
            if (lc.scheme == 5) return th ? FN<DIM, 5, 2, true>(__VA_ARGS__) : FN<DIM, 5, 2, false>(__VA_ARGS__); \
            if (lc.scheme == 6) return th ? FN<DIM, 6, 2, true>(__VA_ARGS__) : FN<DIM, 6, 2, false>(__VA_ARGS__); \
            if (lc.scheme == 1 && !th) return FN<DIM, 1, 2, false>(__VA_ARGS__);      \
            if (lc.scheme == 0 && !th) return FN<DIM, 0, 2, false>(__VA_ARGS__);      \
        } else {                                                                      \
            if (lc.scheme == 5) return th ? FN<DIM, 5, 1, true>(__VA_ARGS__) : FN<DIM, 5, 1, false>(__VA_ARGS__); \
            if (lc.scheme == 6) return th ? FN<DIM, 6, 1, true>(__VA_ARGS__) : FN<DIM, 6, 1, false>(__VA_ARGS__); \
            if (lc.scheme == 1 && !th) return FN<DIM, 1, 1, false>(__VA_ARGS__);      \
            if (lc.scheme == 0 && !th) return FN<DIM, 0, 1, false>(__VA_ARGS__);      \
        }                                                                             \
    } while (0)

template <int DIM>
hipError_t launch_stage_dim(const LaunchCfg& lc, const KParams& kp, int mode, int stage,
                            const TileDev* tiles, int ntiles, const BoxDev* boxes,
                            const BoxDev* sbox, int sntiles, const double* Us, const double* Un,
                            double* Uout, const double* dt, unsigned long long* cmax_bits, int* err,
                            hipStream_t s, const double* gq, const FillDev* fuse)
{
    if (fuse && sbox) return hipErrorInvalidValue;   // a single box has no level to read ghosts from
    if (lc.riemann == RIEMANN_PF_MIX || lc.riemann == RIEMANN_PF_PHASIC) {
        if (!gq || lc.thinc) return hipErrorInvalidValue;
        const bool mix = lc.riemann == RIEMANN_PF_MIX;
#define CMP_PF_ARGS kp, lc.limiter, mode, stage, tiles, ntiles, boxes, sbox, sntiles, Us, Un, Uout, dt, \
                    cmax_bits, err, s, gq, fuse
        if (lc.nphase) {
            // the N-phase model's Phase-Field flux: its schemes (Godunov, MUSCL, WENO5 with the
            // consistent limiter), mixture or phasic form
#define CMP_NPF_CASE(SC)                                                                              \
            if (lc.scheme == SC) {                                                                    \
                if (lc.nphase == 3) return mix ? stage4<DIM, SC, RIEMANN_PF_MIX, false, 3>(CMP_PF_ARGS)    \
                                               : stage4<DIM, SC, RIEMANN_PF_PHASIC, false, 3>(CMP_PF_ARGS); \
                if (lc.nphase == 2) return mix ? stage4<DIM, SC, RIEMANN_PF_MIX, false, 2>(CMP_PF_ARGS)    \
                                               : stage4<DIM, SC, RIEMANN_PF_PHASIC, false, 2>(CMP_PF_ARGS); \
            }
            CMP_NPF_CASE(5)
            CMP_NPF_CASE(1)
            CMP_NPF_CASE(0)
#undef CMP_NPF_CASE
            return hipErrorInvalidValue;
        }
        if (lc.scheme == 5) return mix ? stage4<DIM, 5, RIEMANN_PF_MIX, false>(CMP_PF_ARGS)
                                       : stage4<DIM, 5, RIEMANN_PF_PHASIC, false>(CMP_PF_ARGS);
        if (lc.scheme == 6) return mix ? stage4<DIM, 6, RIEMANN_PF_MIX, false>(CMP_PF_ARGS)
                                       : stage4<DIM, 6, RIEMANN_PF_PHASIC, false>(CMP_PF_ARGS);
        if (lc.scheme == 1) return mix ? stage4<DIM, 1, RIEMANN_PF_MIX, false>(CMP_PF_ARGS)
                                       : stage4<DIM, 1, RIEMANN_PF_PHASIC, false>(CMP_PF_ARGS);
        if (lc.scheme == 0) return mix ? stage4<DIM, 0, RIEMANN_PF_MIX, false>(CMP_PF_ARGS)
                                       : stage4<DIM, 0, RIEMANN_PF_PHASIC, false>(CMP_PF_ARGS);
#undef CMP_PF_ARGS
        return hipErrorInvalidValue;
    }
    if (lc.nphase) {
        // the N-phase model: Godunov, MUSCL, WENO5 (consistent limiter) with HLL or HLLC
        if (lc.thinc || (lc.riemann != 1 && lc.riemann != 2)) return hipErrorInvalidValue;
#define CMP_NP_ARGS kp, lc.limiter, mode, stage, tiles, ntiles, boxes, sbox, sntiles, Us, Un, Uout, dt, \
                    cmax_bits, err, s, nullptr, fuse
#define CMP_NP_CASE(SC)                                                                               \
        if (lc.scheme == SC) {                                                                        \
            if (lc.nphase == 3) return lc.riemann == 2 ? stage4<DIM, SC, 2, false, 3>(CMP_NP_ARGS)    \
                                                       : stage4<DIM, SC, 1, false, 3>(CMP_NP_ARGS);   \
            if (lc.nphase == 2) return lc.riemann == 2 ? stage4<DIM, SC, 2, false, 2>(CMP_NP_ARGS)    \
                                                       : stage4<DIM, SC, 1, false, 2>(CMP_NP_ARGS);   \
        }
        CMP_NP_CASE(5)
        CMP_NP_CASE(1)
        CMP_NP_CASE(0)
#undef CMP_NP_CASE
#undef CMP_NP_ARGS
        return hipErrorInvalidValue;
    }
    CMP_DISPATCH(stage4, kp, lc.limiter, mode, stage, tiles, ntiles, boxes, sbox, sntiles, Us, Un, Uout,
                 dt, cmax_bits, err, s, nullptr, fuse);
    return hipErrorInvalidValue;
}

template <int DIM>
hipError_t launch_stage_batch_dim(const LaunchCfg& lc, const KParams* kps, int stage, const TileDev* tiles,
                                  int ntiles, const BoxDev* boxes, const double* Us, const double* Un, double* Uout,
                                  const double* dts, unsigned long long* cmax_bits, int* errs, const FillDev* fds,
                                  hipStream_t s)
{
    if (ntiles <= 0) return hipSuccess;
    if (lc.thinc || lc.nphase) return hipErrorInvalidValue;
#define CMP_BATCH_LAUNCH(SC, RI)                                                                        \
    do {                                                                                                \
        hipLaunchKernelGGL((k_stage_batch<DIM, SC, RI>), dim3(ntiles), dim3(NT), 0, s, kps, lc.limiter,  \
                           stage, tiles, ntiles, boxes, Us, Un, Uout, dts, cmax_bits, errs, fds);      \
        return hipGetLastError();                                                                       \
    } while (0)
    if (lc.riemann == 2) {
        if (lc.scheme == 5) CMP_BATCH_LAUNCH(5, 2);
        if (lc.scheme == 6) CMP_BATCH_LAUNCH(6, 2);
        if (lc.scheme == 1) CMP_BATCH_LAUNCH(1, 2);
        if (lc.scheme == 0) CMP_BATCH_LAUNCH(0, 2);
    } else if (lc.riemann == 1) {
        if (lc.scheme == 5) CMP_BATCH_LAUNCH(5, 1);
        if (lc.scheme == 6) CMP_BATCH_LAUNCH(6, 1);
        if (lc.scheme == 1) CMP_BATCH_LAUNCH(1, 1);
        if (lc.scheme == 0) CMP_BATCH_LAUNCH(0, 1);
    }
#undef CMP_BATCH_LAUNCH
    return hipErrorInvalidValue;
}

// ---- refinement tagging ----------------------------------------------------------------------
// The tagging variable at a cell: a state component, or a derived rho / P / u / v / w
// (compute_derived_var, include/Physics_Compressible5Eq/ComputeDerivedVar_Compressible5Eq.H:110-160).
template <int DIM, int NP>
__device__ __forceinline__ double tag_var(const KParams& kp, const BoxDev& bx, const double* __restrict__ Ub, int i,
                                          int j, int k, int var, int& bad)
{
    constexpr int N = NP ? 2 * NP + DIM + 1 : 4 + DIM;
    constexpr int NM = NP ? NP : 2;
    const long long c = fab_index(bx, i, j, k);
    if (var >= 0) return Ub[c + var * bx.sn];
    double Uc[N], W[N];
#pragma unroll
    for (int q = 0; q < N; ++q) Uc[q] = Ub[c + q * bx.sn];
    if (var == CMP_TAG_RHO) {
        double rho = Uc[0];
#pragma unroll
        for (int q = 1; q < NM; ++q) rho += Uc[q];
        return rho;
    }
    if constexpr (NP > 0) np_cons2prim<DIM, NP>(kp, Uc, W, bad);
    else cons2prim<DIM>(kp, Uc, W, bad);
    if (var == CMP_TAG_P) return W[NM + DIM];
    const int m = CMP_TAG_U - var;
    return (m < DIM) ? W[NM + m] : 0.0;
}

// ErrorEst's criteria loop (source/COMPAS.cpp:839-918; include/Tagging.H:118-201) for every valid
// cell; one grid row per box, tags box by box (toff: each box's first tag).
template <int DIM, int NP>
__global__ __launch_bounds__(NT) void k_tag(KParams kp, const BoxDev* __restrict__ boxes,
                                            const long long* __restrict__ toff, const double* __restrict__ U,
                                            TagCriteria tc, unsigned char* tags, int* err)
{
    const BoxDev bx = boxes[blockIdx.y];
    const int n0 = bx.hi[0] - bx.lo[0] + 1, n1 = bx.hi[1] - bx.lo[1] + 1;
    const int n2 = (DIM == 3) ? bx.hi[2] - bx.lo[2] + 1 : 1;
    const long long q = (long long)blockIdx.x * NT + threadIdx.x;
    if (q >= (long long)n0 * n1 * n2) return;
    int ii, jj, kk;
    pos3(q, n0, n1, ii, jj, kk);
    const int i = bx.lo[0] + ii, j = bx.lo[1] + jj, k = bx.lo[2] + kk;
    const double* __restrict__ Ub = U + bx.off;
    int bad = 0;
    unsigned char t = 0;
    for (int m = 0; m < tc.n; ++m) {
        const cmp_tag_criterion C = tc.c[m];
        const double s = tag_var<DIM, NP>(kp, bx, Ub, i, j, k, C.var, bad);
        if (C.kind == CMP_TAG_GRAD) {
            double ax = fabs(tag_var<DIM, NP>(kp, bx, Ub, i - 1, j, k, C.var, bad) - s);
            double ay = fabs(tag_var<DIM, NP>(kp, bx, Ub, i, j - 1, k, C.var, bad) - s);
            ax = smax(ax, fabs(s - tag_var<DIM, NP>(kp, bx, Ub, i + 1, j, k, C.var, bad)));
            ay = smax(ay, fabs(s - tag_var<DIM, NP>(kp, bx, Ub, i, j + 1, k, C.var, bad)));
            double az = 0.0;
            if (DIM == 3) {
                az = fabs(tag_var<DIM, NP>(kp, bx, Ub, i, j, k - 1, C.var, bad) - s);
                az = smax(az, fabs(s - tag_var<DIM, NP>(kp, bx, Ub, i, j, k + 1, C.var, bad)));
            }
            if (smax(smax(ax, ay), az) >= C.v0) t = 1;
        } else if (C.kind == CMP_TAG_LESS) {
            if (s < C.v0) t = 1;
        } else if (C.kind == CMP_TAG_GREATER) {
            if (s > C.v0) t = 1;
        } else if (C.kind == CMP_TAG_RANGE) {
            if (s > C.v0 && s < C.v1) t = 1;
        }
    }
    tags[toff[blockIdx.y] + q] = t;
    if (bad && err) atomicMax(err, bad);
}

template <int DIM>
hipError_t launch_tag_dim(const LaunchCfg& lc, const KParams& kp, const BoxDev* boxes, int nbox, long long maxcells,
                          const long long* toff, const double* U, const TagCriteria& tc, unsigned char* tags,
                          int* err, hipStream_t s)
{
    const dim3 grid((unsigned)((maxcells + NT - 1) / NT), (unsigned)nbox);
    if (lc.nphase == 3)
        hipLaunchKernelGGL((k_tag<DIM, 3>), grid, dim3(NT), 0, s, kp, boxes, toff, U, tc, tags, err);
    else if (lc.nphase == 2)
        hipLaunchKernelGGL((k_tag<DIM, 2>), grid, dim3(NT), 0, s, kp, boxes, toff, U, tc, tags, err);
    else
        hipLaunchKernelGGL((k_tag<DIM, 0>), grid, dim3(NT), 0, s, kp, boxes, toff, U, tc, tags, err);
    return hipGetLastError();
}

// ---- average_down (AverageDownTo, source/COMPAS.cpp:1550-1562) -----------------------------------
// Every coarse cell under a fine box takes the mean of its ratio^DIM children (summed k, j, i, times
// volfrac = 1/ratio^DIM); the coarse cell's box comes from the coarse level's owner map. One grid row
// per fine box, one thread per coarse cell.
template <int DIM>
__global__ __launch_bounds__(NT) void k_avgdown(const BoxDev* __restrict__ fboxes, FillDev cfd,
                                                const double* __restrict__ Uf, double* Uc, int N, int ratio,
                                                double volfrac, int* err)
{
    const BoxDev fb = fboxes[blockIdx.y];
    int clo[3] = {0, 0, 0}, ce[3] = {1, 1, 1};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        clo[d] = fb.lo[d] / ratio;
        ce[d] = (fb.hi[d] + 1) / ratio - clo[d];
    }
    const long long q = (long long)blockIdx.x * NT + threadIdx.x;
    if (q >= (long long)ce[0] * ce[1] * ce[2]) return;
    int ic, jc, kc;
    pos3(q, ce[0], ce[1], ic, jc, kc);
    ic += clo[0];
    jc += clo[1];
    kc += clo[2];
    const int cc[3] = {ic, jc, kc};
    int bi = 0;
    for (int d = DIM - 1; d >= 0; --d) {
        if (cc[d] < cfd.dlo[d] || cc[d] > cfd.dhi[d]) { atomicMax(err, CMP_ERR_LAYOUT); return; }
        bi = bi * cfd.nb[d] + (cc[d] - cfd.dlo[d]) / cfd.bs;
    }
    const int own = cfd.owner[bi];
    if (own < 0) { atomicMax(err, CMP_ERR_LAYOUT); return; }
    const BoxDev cb = cfd.boxes[own];
    const long long co = cb.off + fab_index(cb, ic, jc, kc);
    const int rz = (DIM == 3) ? ratio : 1;
    for (int n = 0; n < N; ++n) {
        double c = 0.0;
        for (int kr = 0; kr < rz; ++kr)
            for (int jr = 0; jr < ratio; ++jr)
                for (int ir = 0; ir < ratio; ++ir) {
                    const int i = ic * ratio + ir, j = jc * ratio + jr, k = (DIM == 3) ? kc * ratio + kr : 0;
                    c += Uf[fb.off + fab_index(fb, i, j, k) + n * fb.sn];
                }
        Uc[co + n * cb.sn] = volfrac * c;
    }
}

// The average-down as a gather: k_avgdown_plan finds, once for a pair of levels, what k_avgdown finds
// per call for each covered coarse cell (numbered box by box, cprefix: nfbox + 1 prefix sums of each
// fine box's covered coarse cells, x fastest inside a box) - its fine box and the fab offset of its
// first fine cell, its owning coarse box and its offset there (cbox < 0: error code -cbox) - and
// k_avgdown_apply, one thread per covered cell and component, sums the same fine cells in the same
// order as k_avgdown.
template <int DIM>
__global__ __launch_bounds__(NT) void k_avgdown_plan(const BoxDev* __restrict__ fboxes,
                                                     const long long* __restrict__ cprefix, int nfbox,
                                                     long long total, FillDev cfd, int ratio,
                                                     AvgCell* __restrict__ plan)
{
    const long long qg = (long long)blockIdx.x * NT + threadIdx.x;
    if (qg >= total) return;
    int lo = 0, hi = nfbox - 1;
    while (lo < hi) {   // the fine box whose covered cells hold qg
        const int mid = (lo + hi + 1) >> 1;
        if (cprefix[mid] <= qg) lo = mid;
        else hi = mid - 1;
    }
    const BoxDev fb = fboxes[lo];
    int clo[3] = {0, 0, 0}, ce[3] = {1, 1, 1};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        clo[d] = fb.lo[d] / ratio;
        ce[d] = (fb.hi[d] + 1) / ratio - clo[d];
    }
    int ic, jc, kc;
    pos3(qg - cprefix[lo], ce[0], ce[1], ic, jc, kc);
    ic += clo[0];
    jc += clo[1];
    kc += clo[2];
    AvgCell a;
    a.fbox = lo;
    a.fcell = (int)fab_index(fb, ic * ratio, jc * ratio, (DIM == 3) ? kc * ratio : 0);
    a.cbox = -CMP_ERR_LAYOUT;
    a.ccell = 0;
    const int cc[3] = {ic, jc, kc};
    int bi = 0;
    bool in = true;
    for (int d = DIM - 1; d >= 0; --d) {
        if (cc[d] < cfd.dlo[d] || cc[d] > cfd.dhi[d]) in = false;
        else bi = bi * cfd.nb[d] + (cc[d] - cfd.dlo[d]) / cfd.bs;
    }
    if (in) {
        const int own = cfd.owner[bi];
        if (own >= 0) {
            a.cbox = own;
            a.ccell = (int)fab_index(cfd.boxes[own], ic, jc, kc);
        }
    }
    plan[qg] = a;
}

template <int DIM>
__global__ __launch_bounds__(NT) void k_avgdown_apply(const BoxDev* __restrict__ fboxes,
                                                      const BoxDev* __restrict__ cboxes,
                                                      const AvgCell* __restrict__ plan, long long total,
                                                      const double* __restrict__ Uf, double* Uc, int N, int ratio,
                                                      double volfrac, int* err)
{
    const long long t = (long long)blockIdx.x * NT + threadIdx.x;
    if (t >= total * N) return;
    const int n = (int)(t / total);
    const AvgCell a = plan[t - (long long)n * total];
    if (a.cbox < 0) { atomicMax(err, -a.cbox); return; }
    const BoxDev fb = fboxes[a.fbox];
    const BoxDev cb = cboxes[a.cbox];
    const double* __restrict__ f = Uf + fb.off + a.fcell + (long long)n * fb.sn;
    const int rz = (DIM == 3) ? ratio : 1;
    double c = 0.0;
    for (int kr = 0; kr < rz; ++kr)
        for (int jr = 0; jr < ratio; ++jr)
            for (int ir = 0; ir < ratio; ++ir) c += f[ir + (long long)jr * fb.sj + (long long)kr * fb.sk];
    Uc[cb.off + a.ccell + (long long)n * cb.sn] = volfrac * c;
}

// A coarse level's cell c (any index within one cell of its domain) as the coarse ghost fill would
// give it: mapped through the coarse level's BCs, read from the box owning it, momenta flipped.
template <int DIM, int N, int MOM>
__device__ __forceinline__ int coarse_value(const FillDev& cd, const double* __restrict__ Uc, int c0, int c1, int c2,
                                            double* v)
{
    int src[3] = {c0, c1, c2};
    unsigned flip;
    const int e = map_cell<DIM>(cd, src, flip);
    if (e) return e;
    const int own = cell_owner<DIM>(cd, src);
    if (own < 0) return CMP_ERR_LAYOUT;
    const BoxDev cb = cd.boxes[own];
    const long long ci = cb.off + fab_index(cb, src[0], src[1], src[2]);
#pragma unroll
    for (int n = 0; n < N; ++n) {
        const double x = Uc[ci + n * cb.sn];
        v[n] = (n >= MOM && n < MOM + DIM && ((flip >> (n - MOM)) & 1u)) ? -x : x;
    }
    return 0;
}

// Two-level ghost fill of a fine level (FillPatchTwoLevels, source/COMPAS.cpp:1679-1724). Thread g
// handles ghost cell g of one fine box: its source cell is mapped through the fine level's BCs as
// the same-level fill maps it; a source in a fine box is copied, any other is the cell-conservative
// linear interpolation of the coarse level at it (AMReX's CellConservativeLinear with linear
// limiting): per direction d the central coarse difference dc = (c+ - c-)/2 of every component, its
// MC-limited value sx = sign(dc) min(|dc|, 2|c+ - c|, 2|c - c-|) (0 where the one-sided differences
// have opposite signs), a factor f_d = min over the components with dc != 0 of sx/dc that scales
// the slopes of all components alike, and the value c + sum_d off_d f_d dc_d, off_d the fine
// cell's centre offset from the coarse cell's in coarse cells, (i - ic r + 1/2)/r - 1/2. Products
// and sums are not contracted to FMAs, so the result is bit for bit the one of plain IEEE
// arithmetic in this order.
// ALL: every cell of the boxes of fd (valid and ghost), the same-level sources taken from the boxes
// of sfd (has_src 0: none) in Us - the data of a level's new boxes after a regrid; otherwise the
// ghost cells only, sfd = fd and Us = Uf.
// One direction of that interpolation: the central differences dc of every component (returned in
// cm, which holds the minus neighbour on entry) and the factor sf that scales them all.
template <int N>
__device__ __forceinline__ double lin_slopes(const double* c0, double* cm, const double* cp)
{
#pragma clang fp contract(off)
    double sf = 1.0;
#pragma unroll
    for (int n = 0; n < N; ++n) {
        const double dc = 0.5 * (cp[n] - cm[n]);
        const double df = 2.0 * (cp[n] - c0[n]);
        const double db = 2.0 * (c0[n] - cm[n]);
        const double af = fabs(df), ab = fabs(db);
        double sx = (df * db >= 0.0) ? ((ab < af) ? ab : af) : 0.0;
        const double adc = fabs(dc);
        sx = copysign(1.0, dc) * ((adc < sx) ? adc : sx);
        if (dc != 0.0) {
            const double r = sx / dc;
            sf = (r < sf) ? r : sf;
        }
        cm[n] = dc;
    }
    return sf;
}

template <int DIM, int N, bool ALL>
__global__ __launch_bounds__(NT)
void k_fill2(FillDev fd, FillDev sfd, int has_src, FillDev cd, int box0, int ratio, double* Uf,
             const double* Us, const double* __restrict__ Uc, int* err)
{
#pragma clang fp contract(off)
    // first momentum component of a flow state; a Phase-Field field (1 to 3 comps) has none
    constexpr int MOM = (N == 4 + DIM) ? 2 : ((N > 3) ? (N - DIM - 1) / 2 : -100);
    const int b = box0 + (int)blockIdx.y;
    const BoxDev bx = fd.boxes[b];
    const int ng = fd.ng;
    const int nv0 = bx.hi[0] - bx.lo[0] + 1, nv1 = bx.hi[1] - bx.lo[1] + 1;
    const int nv2 = (DIM == 3) ? bx.hi[2] - bx.lo[2] + 1 : 1;
    const int ex = nv0 + 2 * ng, ey = nv1 + 2 * ng;
    const int ez = (DIM == 3) ? nv2 + 2 * ng : 1;
    const int ncell = ALL ? ex * ey * ez : ex * ey * ez - nv0 * nv1 * nv2;
    for (int g = (int)(blockIdx.x * NT + threadIdx.x); g < ncell; g += (int)(gridDim.x * NT)) {
        int x, y, z;
        if (ALL) {
            x = g % ex;
            y = (g / ex) % ey;
            z = g / (ex * ey);
        } else {
            ghost_xyz<DIM>(g, ng, ex, ey, ez, nv1, x, y, z);
        }
        const int cell[3] = {bx.flo[0] + x, bx.flo[1] + y, bx.flo[2] + z};
        const long long di = bx.off + fab_index(bx, cell[0], cell[1], cell[2]);
        int src[3] = {cell[0], cell[1], cell[2]};
        unsigned flip;
        int e = map_cell<DIM>(fd, src, flip);
        if (e) { atomicMax(err, e); continue; }
        const int own = has_src ? cell_owner<DIM>(sfd, src) : -1;
        double v[N];
        if (own >= 0) {
            const BoxDev so = sfd.boxes[own];
            const long long si = so.off + fab_index(so, src[0], src[1], src[2]);
#pragma unroll
            for (int n = 0; n < N; ++n) v[n] = Us[si + n * so.sn];
        } else {
            int ic[3] = {src[0], src[1], src[2]};
#pragma unroll
            for (int d = 0; d < DIM; ++d)
                ic[d] = (src[d] >= 0) ? src[d] / ratio : -((-src[d] + ratio - 1) / ratio);
            double c0[N];
            e = coarse_value<DIM, N, MOM>(cd, Uc, ic[0], ic[1], ic[2], c0);
#pragma unroll
            for (int n = 0; n < N; ++n) v[n] = c0[n];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                if (e) break;
                double cm[N], cp[N];
                int a[3] = {ic[0], ic[1], ic[2]}, z3[3] = {ic[0], ic[1], ic[2]};
                a[d] -= 1;
                z3[d] += 1;
                e = coarse_value<DIM, N, MOM>(cd, Uc, a[0], a[1], a[2], cm);
                if (!e) e = coarse_value<DIM, N, MOM>(cd, Uc, z3[0], z3[1], z3[2], cp);
                if (e) break;
                const double sf = lin_slopes<N>(c0, cm, cp);
                const double off = ((double)(src[d] - ic[d] * ratio) + 0.5) / (double)ratio - 0.5;
#pragma unroll
                for (int n = 0; n < N; ++n) v[n] = v[n] + off * (cm[n] * sf);
            }
            if (e) { atomicMax(err, e); continue; }
        }
#pragma unroll
        for (int n = 0; n < N; ++n) {
            const bool neg = n >= MOM && n < MOM + DIM && ((flip >> (n - MOM)) & 1u);
            Uf[di + n * bx.sn] = neg ? -v[n] : v[n];
        }
    }
}

// The two-level ghost fill of a fine level as a gather: k_fill2_plan finds, once for a pair of
// levels, what k_fill2 (ghost cells only, the same-level sources the level's own boxes) finds per
// call for each ghost cell - the BC map and momentum flips of the cell, its same-level source, or
// the coarse cell under it and that cell's neighbours along each direction (each mapped through the
// coarse level's BCs, with its flips and owning box) - and k_fill2_apply reads those sources and
// does k_fill2's arithmetic, so its values are k_fill2's bit for bit. Plan entry g of box b at
// fd.gprefix[b] + g. Slot 0: the same-level source or the coarse centre; slots 1 + 2d and 2 + 2d the
// coarse neighbours below and above along d.
template <int DIM>
__device__ __forceinline__ int f2_source(const FillDev& fd, int* c, int& box, int& cell, unsigned& flip)
{
    const int e = map_cell<DIM>(fd, c, flip);
    if (e) return e;
    box = cell_owner<DIM>(fd, c);
    if (box < 0) return CMP_ERR_LAYOUT;
    cell = (int)fab_index(fd.boxes[box], c[0], c[1], c[2]);
    return 0;
}

template <int DIM>
__global__ __launch_bounds__(NT)
void k_fill2_plan(FillDev fd, FillDev cd, int box0, int ratio, F2Cell<DIM>* __restrict__ plan)
{
    const int b = box0 + (int)blockIdx.y;
    const BoxDev bx = fd.boxes[b];
    const int ng = fd.ng;
    const int nv0 = bx.hi[0] - bx.lo[0] + 1, nv1 = bx.hi[1] - bx.lo[1] + 1;
    const int nv2 = (DIM == 3) ? bx.hi[2] - bx.lo[2] + 1 : 1;
    const int ex = nv0 + 2 * ng, ey = nv1 + 2 * ng;
    const int ez = (DIM == 3) ? nv2 + 2 * ng : 1;
    const int ncell = ex * ey * ez - nv0 * nv1 * nv2;
    for (int g = (int)(blockIdx.x * NT + threadIdx.x); g < ncell; g += (int)(gridDim.x * NT)) {
        int x, y, z;
        ghost_xyz<DIM>(g, ng, ex, ey, ez, nv1, x, y, z);
        F2Cell<DIM> p;
        for (int k = 0; k < F2Cell<DIM>::NS; ++k) p.box[k] = p.cell[k] = 0;
        p.kind = 0;
        p.err = 0;
        int src[3] = {bx.flo[0] + x, bx.flo[1] + y, bx.flo[2] + z};
        p.dbox = b;
        p.dcell = (int)fab_index(bx, src[0], src[1], src[2]);
        unsigned flip;
        int e = map_cell<DIM>(fd, src, flip);
        p.flip = flip & 7u;
        if (!e) {
            const int own = cell_owner<DIM>(fd, src);
            if (own >= 0) {
                p.box[0] = own;
                p.cell[0] = (int)fab_index(fd.boxes[own], src[0], src[1], src[2]);
            } else {
                int ic[3] = {src[0], src[1], src[2]};
                p.kind = 1;
                for (int d = 0; d < DIM; ++d) {
                    ic[d] = (src[d] >= 0) ? src[d] / ratio : -((-src[d] + ratio - 1) / ratio);
                    p.kind |= (src[d] - ic[d] * ratio) << (8 + 8 * d);
                }
                for (int k = 0; k < 1 + 2 * DIM && !e; ++k) {
                    int c[3] = {ic[0], ic[1], ic[2]};
                    if (k > 0) c[(k - 1) / 2] += ((k - 1) % 2) ? 1 : -1;
                    unsigned fl;
                    e = f2_source<DIM>(cd, c, p.box[k], p.cell[k], fl);
                    p.flip |= (fl & 7u) << (3 * (k + 1));
                }
            }
        }
        p.err = e;
        plan[fd.gprefix[b] + g] = p;
    }
}

template <int DIM, int N>
__global__ __launch_bounds__(NT)
void k_fill2_apply(FillDev fd, FillDev cd, long long nghost, int ratio, const F2Cell<DIM>* __restrict__ plan,
                   double* Uf, const double* __restrict__ Uc, int* err)
{
#pragma clang fp contract(off)
    constexpr int MOM = (N == 4 + DIM) ? 2 : ((N > 3) ? (N - DIM - 1) / 2 : -100);
    for (long long g = (long long)blockIdx.x * NT + threadIdx.x; g < nghost; g += (long long)gridDim.x * NT) {
        const F2Cell<DIM> p = plan[g];
        if (p.err) { atomicMax(err, p.err); continue; }
        const BoxDev bx = fd.boxes[p.dbox];
        const long long di = bx.off + p.dcell;
        // the N components of source k of the entry (momenta flipped by its bits)
        auto load = [&](const FillDev& f, const double* __restrict__ U, int k, double* v) {
            const BoxDev so = f.boxes[p.box[k]];
            const long long si = so.off + p.cell[k];
            const unsigned fl = (p.flip >> (3 * (k + 1))) & 7u;
#pragma unroll
            for (int n = 0; n < N; ++n) {
                const double w = U[si + n * so.sn];
                v[n] = (n >= MOM && n < MOM + DIM && ((fl >> (n - MOM)) & 1u)) ? -w : w;
            }
        };
        double v[N];
        if (!(p.kind & 1)) {
            const BoxDev so = fd.boxes[p.box[0]];
            const long long si = so.off + p.cell[0];
#pragma unroll
            for (int n = 0; n < N; ++n) v[n] = Uf[si + n * so.sn];
        } else {
            double c0[N];
            load(cd, Uc, 0, c0);
#pragma unroll
            for (int n = 0; n < N; ++n) v[n] = c0[n];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                double cm[N], cp[N];
                load(cd, Uc, 1 + 2 * d, cm);
                load(cd, Uc, 2 + 2 * d, cp);
                const double sf = lin_slopes<N>(c0, cm, cp);
                const int rem = (p.kind >> (8 + 8 * d)) & 255;
                const double off = ((double)rem + 0.5) / (double)ratio - 0.5;
#pragma unroll
                for (int n = 0; n < N; ++n) v[n] = v[n] + off * (cm[n] * sf);
            }
        }
#pragma unroll
        for (int n = 0; n < N; ++n) {
            const bool neg = n >= MOM && n < MOM + DIM && ((p.flip >> (n - MOM)) & 1u);
            Uf[di + n * bx.sn] = neg ? -v[n] : v[n];
        }
    }
}

// The coarse-fine face correction of one RK stage of a hierarchy (StageRHSAllLevels,
// source/Class_Compressible_PhaseField_StageCoupling.cpp:84-153, with run.do_reflux = 1): a coarse
// face over the finer level takes the mean of the fine faces over it (amrex::average_down_faces:
// the lower transverse index fastest, times 1/ratio^(dim-1)), fluxes and u*_n alike, before the
// divergence. The stage kernel has already written the coarse cell with its own faces; the
// correction adds w dt times the change: -(F_mean - F)/dx for a high face, +(F_mean - F)/dx for a
// low face, and the nc coefficient of the cell times the change of div u*. Uc and Uf are the stage
// states with every ghost cell filled.
//
// k_cf_faces: one face job per thread (a coarse face, or one of the fine faces over it).
template <int DIM, int SCHEME, int RIEMANN, bool THINC, int NP>
__global__ __launch_bounds__(64)
void k_cf_faces(KParams kp, int limiter, const CfFace* __restrict__ jobs, int njobs, const BoxDev* __restrict__ cboxes,
                const double* __restrict__ Uc, const BoxDev* __restrict__ fboxes, const double* __restrict__ Uf,
                double* __restrict__ faces, int* err)
{
    constexpr int N = (NP > 0) ? 2 * NP + DIM + 1 : 4 + DIM;
    const int j = (int)(blockIdx.x * 64 + threadIdx.x);
    if (j >= njobs) return;
    const CfFace fj = jobs[j];
    const BoxDev bx = fj.fine ? fboxes[fj.box] : cboxes[fj.box];
    const double* Ub = (fj.fine ? Uf : Uc) + bx.off;
    int bad = 0;
    double F[N], u;
    if (fj.dir == 0)
        u = face_at<DIM, SCHEME, RIEMANN, THINC, 0, NP>(kp, limiter, bx, Ub, fj.c[0], fj.c[1], fj.c[2], F, bad);
    else if (fj.dir == 1)
        u = face_at<DIM, SCHEME, RIEMANN, THINC, 1, NP>(kp, limiter, bx, Ub, fj.c[0], fj.c[1], fj.c[2], F, bad);
    else
        u = face_at<DIM, SCHEME, RIEMANN, THINC, (DIM == 3 ? 2 : 1), NP>(kp, limiter, bx, Ub, fj.c[0], fj.c[1],
                                                                          fj.c[2], F, bad);
    double* o = faces + (long long)j * (N + 1);
#pragma unroll
    for (int q = 0; q < N; ++q) o[q] = F[q];
    o[N] = u;
    if (bad) atomicMax(err, bad);
}

// The correction of one listed coarse cell, its face jobs combined in their fixed order. PF: a
// Phase-Field stage (faces hold -fhat; no non-conservative term).
template <int DIM, int NP, bool PF>
__device__ __forceinline__ void cf_apply_cell(const KParams& kp, const CfCell& cc, const BoxDev* __restrict__ cboxes,
                                              const double* __restrict__ Uc, const double* __restrict__ faces,
                                              int ratio, double* Uout, const double* __restrict__ dtp, int stage,
                                              int* err)
{
    constexpr int N = (NP > 0) ? 2 * NP + DIM + 1 : 4 + DIM;
    constexpr int IA = 3 + DIM;
    const BoxDev cb = cboxes[cc.box];
    int bad = 0;
    double dL[N];
#pragma unroll
    for (int q = 0; q < N; ++q) dL[q] = 0.0;
    double ddiv = 0.0;
    int jb = cc.job0;
    int nsub = 1;
    for (int a = 1; a < DIM; ++a) nsub *= ratio;
    const double facInv = 1.0 / (double)nsub;
    for (int d = 0; d < DIM; ++d)
        for (int side = 0; side < 2; ++side) {
            if (!((cc.mask >> (2 * d + side)) & 1)) continue;
            const double* fc = faces + (long long)jb * (N + 1);
            double sF[N], su = 0.0;
#pragma unroll
            for (int q = 0; q < N; ++q) sF[q] = 0.0;
            for (int m = 1; m <= nsub; ++m) {
                const double* ff = faces + (long long)(jb + m) * (N + 1);
#pragma unroll
                for (int q = 0; q < N; ++q) sF[q] += ff[q];
                su += ff[N];
            }
            jb += 1 + nsub;
            const double sgn = side ? -1.0 : 1.0;
            const double idx = kp.inv_dx[d];
#pragma unroll
            for (int q = 0; q < N; ++q) dL[q] += sgn * ((sF[q] * facInv - fc[q]) * idx);
            ddiv -= sgn * ((su * facInv - fc[N]) * idx);
        }
    const long long c = cb.off + fab_index(cb, cc.c[0], cc.c[1], cc.c[2]);
    double Ucell[N];
#pragma unroll
    for (int q = 0; q < N; ++q) Ucell[q] = Uc[c + q * cb.sn];
    if constexpr (PF) {
        (void)ddiv;
    } else if constexpr (NP > 0) {
        // every alpha_k takes phi_k (1 + K_k) times the change of div u*
        double f[NP > 0 ? NP : 1];
        np_nc_coefficients<DIM, (NP > 0 ? NP : 2)>(kp, Ucell, f, bad);
#pragma unroll
        for (int k = 0; k < NP; ++k) dL[NpIdx<DIM, (NP > 0 ? NP : 2)>::A + k] += f[k] * ddiv;
    } else {
        dL[IA] += nc_coefficient<DIM>(kp, Ucell, bad) * ddiv;
    }
    const double dt = *dtp;
    const double w = (stage == 0) ? dt : ((stage == 1) ? 0.25 * dt : (2.0 / 3.0) * dt);
#pragma unroll
    for (int q = 0; q < N; ++q) {
        const double v = Uout[c + q * cb.sn] + w * dL[q];
        if (__builtin_isnan(v)) bad = CMP_ERR_NAN;
        Uout[c + q * cb.sn] = v;
    }
    if (bad) atomicMax(err, bad);
}

// k_cf_apply: one thread per listed coarse cell of one pair of levels.
template <int DIM, int NP, bool PF = false>
__global__ __launch_bounds__(64)
void k_cf_apply(KParams kp, const CfCell* __restrict__ list, int n, const BoxDev* __restrict__ cboxes,
                const double* __restrict__ Uc, const double* __restrict__ faces, int ratio, double* Uout,
                const double* __restrict__ dtp, int stage, int* err)
{
    const int e = (int)(blockIdx.x * 64 + threadIdx.x);
    if (e >= n) return;
    cf_apply_cell<DIM, NP, PF>(kp, list[e], cboxes, Uc, faces, ratio, Uout, dtp, stage, err);
}

// k_cf_apply_all: the flow corrections of every pair of levels of a hierarchy in one launch (they
// write different levels), thread e taking cell e - start of the last pair that starts at or before e.
template <int DIM, int NP>
__global__ __launch_bounds__(64)
void k_cf_apply_all(const CfPairDev* __restrict__ pairs, int npairs, int ntotal, CfStatePtrs st, int ratio,
                    const double* __restrict__ dtp, int stage)
{
    const int e = (int)(blockIdx.x * 64 + threadIdx.x);
    if (e >= ntotal) return;
    int k = npairs - 1;
    while (k > 0 && pairs[k].start > e) --k;
    const CfPairDev& P = pairs[k];
    cf_apply_cell<DIM, NP, false>(P.kp, P.list[e - P.start], P.cboxes, st.Uc[k], P.faces, ratio, st.Uout[k], dtp,
                                  stage, P.err);
}

template <int DIM>
hipError_t launch_cf_apply_all_dim(const LaunchCfg& lc, const CfPairDev* pairs, int npairs, int ntotal,
                                   const CfStatePtrs& st, int ratio, const double* dt, int stage, hipStream_t s)
{
    if (ntotal <= 0 || npairs <= 0) return hipSuccess;
    if (lc.riemann == RIEMANN_PF_MIX || lc.riemann == RIEMANN_PF_PHASIC) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((ntotal + 63) / 64));
    if (lc.nphase == 3)
        hipLaunchKernelGGL((k_cf_apply_all<DIM, 3>), grid, dim3(64), 0, s, pairs, npairs, ntotal, st, ratio, dt, stage);
    else if (lc.nphase == 2)
        hipLaunchKernelGGL((k_cf_apply_all<DIM, 2>), grid, dim3(64), 0, s, pairs, npairs, ntotal, st, ratio, dt, stage);
    else
        hipLaunchKernelGGL((k_cf_apply_all<DIM, 0>), grid, dim3(64), 0, s, pairs, npairs, ntotal, st, ratio, dt, stage);
    return hipGetLastError();
}

// The Phase-Field flux of the face between cells c - e_DIR and c of box bx (Ub its fab, ghosts
// filled) at the Grad_Q values g of that face, every stencil cell taken through cons->prim here.
template <int DIM, int SCHEME, int FORM, int DIR, int NP>
__device__ __forceinline__ void face_at_pf(const KParams& kp, int limiter, const BoxDev& bx,
                                           const double* __restrict__ Ub, int ci, int cj, int ck, const double* g,
                                           double* Fl, int& bad)
{
    constexpr int N = (NP > 0) ? 2 * NP + DIM + 1 : 4 + DIM;
    constexpr int G = Stencil<SCHEME>::G;
    const long long sd = (DIR == 0) ? 1 : ((DIR == 1) ? bx.sj : bx.sk);
    const long long c0 = fab_index(bx, ci, cj, ck);
    double W[2 * G][N];
#pragma unroll
    for (int t = 0; t < 2 * G; ++t) {
        double Uc[N];
        const long long c = c0 + (t - G) * sd;
#pragma unroll
        for (int q = 0; q < N; ++q) Uc[q] = Ub[c + q * bx.sn];
        if constexpr (NP > 0) np_cons2prim<DIM, NP>(kp, Uc, W[t], bad);
        else cons2prim<DIM>(kp, Uc, W[t], bad);
    }
    auto st = [&](int t, int q) -> double { return W[t][q]; };
    if constexpr (NP > 0)
        np_pf_face_flux<DIM, SCHEME, FORM, (NP > 0 ? NP : 2)>(kp, limiter, DIR, st, g, Fl, bad);
    else
        pf_face_flux<DIM, SCHEME, FORM>(kp, limiter, DIR, st, g[0], Fl, bad);
}

// The coarse-fine face jobs of a Phase-Field stage (PhaseField_AdvanceAllLevels, the Grad_Q form of
// StageRHSAllLevels): the Phase-Field flux at the face's Grad_Q (each box's face block of cmp_face_layout
// with one component per volume fraction), stored as -fhat as PhaseField_ComputeFaceFluxes stores it
// (source/Class_Compressible_PhaseField.cpp:840), so that k_cf_apply's -div F is the stage's +div fhat;
// a face with Grad_Q = 0 has the flux 0.
template <int DIM, int SCHEME, int FORM, int NP>
__global__ __launch_bounds__(64)
void k_cf_faces_pf(KParams kp, int limiter, const CfFace* __restrict__ jobs, int njobs, const BoxDev* __restrict__ cboxes,
                   const double* __restrict__ Uc, const double* __restrict__ gqc, const BoxDev* __restrict__ fboxes,
                   const double* __restrict__ Uf, const double* __restrict__ gqf, double* __restrict__ faces, int* err)
{
    constexpr int N = (NP > 0) ? 2 * NP + DIM + 1 : 4 + DIM;
    constexpr int NG = (NP > 0) ? NP : 1;
    const int j = (int)(blockIdx.x * 64 + threadIdx.x);
    if (j >= njobs) return;
    const CfFace fj = jobs[j];
    const BoxDev bx = fj.fine ? fboxes[fj.box] : cboxes[fj.box];
    const double* Ub = (fj.fine ? Uf : Uc) + bx.off;
    const double* gq = fj.fine ? gqf : gqc;
    const int d = fj.dir;
    const int n0 = bx.hi[0] - bx.lo[0] + 1 + (d == 0), n1 = bx.hi[1] - bx.lo[1] + 1 + (d == 1);
    const int n2 = (DIM == 3) ? bx.hi[2] - bx.lo[2] + 1 + (d == 2) : 1;
    const long long nface = (long long)n0 * n1 * n2;
    const long long q = (long long)(fj.c[0] - bx.lo[0]) +
                        (long long)n0 * ((fj.c[1] - bx.lo[1]) + (long long)n1 * ((DIM == 3) ? fj.c[2] - bx.lo[2] : 0));
    double g[NG];
    bool any = false;
#pragma unroll
    for (int m = 0; m < NG; ++m) {
        g[m] = gq[bx.goff[d] + q + m * nface];
        any = any || (g[m] != 0.0);
    }
    int bad = 0;
    double F[N];
#pragma unroll
    for (int n = 0; n < N; ++n) F[n] = 0.0;
    if (any) {
        if (d == 0)
            face_at_pf<DIM, SCHEME, FORM, 0, NP>(kp, limiter, bx, Ub, fj.c[0], fj.c[1], fj.c[2], g, F, bad);
        else if (d == 1)
            face_at_pf<DIM, SCHEME, FORM, 1, NP>(kp, limiter, bx, Ub, fj.c[0], fj.c[1], fj.c[2], g, F, bad);
        else
            face_at_pf<DIM, SCHEME, FORM, (DIM == 3 ? 2 : 1), NP>(kp, limiter, bx, Ub, fj.c[0], fj.c[1], fj.c[2], g,
                                                                  F, bad);
    }
    double* o = faces + (long long)j * (N + 1);
#pragma unroll
    for (int n = 0; n < N; ++n) o[n] = -F[n];
    o[N] = 0.0;
    if (bad) atomicMax(err, bad);
}

template <int DIM, int SCHEME, int FORM, int NP>
static hipError_t cfcorr_pf(const KParams& kp, int limiter, int part, const CfCell* list, int n, const CfFace* jobs,
                            int njobs, double* faces, const BoxDev* cboxes, const double* Uc, const double* gqc,
                            const BoxDev* fboxes, const double* Uf, const double* gqf, int ratio, double* Uout,
                            const double* dt, int stage, int* err, hipStream_t s)
{
    if (part & 1)
        hipLaunchKernelGGL((k_cf_faces_pf<DIM, SCHEME, FORM, NP>), dim3((njobs + 63) / 64), dim3(64), 0, s, kp, limiter,
                           jobs, njobs, cboxes, Uc, gqc, fboxes, Uf, gqf, faces, err);
    if (part & 2)
        hipLaunchKernelGGL((k_cf_apply<DIM, NP, true>), dim3((n + 63) / 64), dim3(64), 0, s, kp, list, n, cboxes, Uc,
                           faces, ratio, Uout, dt, stage, err);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_cf_correct_pf_dim(const LaunchCfg& lc, const KParams& kp, int part, const CfCell* list, int n,
                                    const CfFace* jobs, int njobs, double* faces, const BoxDev* cboxes,
                                    const double* Uc, const double* gqc, const BoxDev* fboxes, const double* Uf,
                                    const double* gqf, int ratio, double* Uout, const double* dt, int stage, int* err,
                                    hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (lc.riemann != RIEMANN_PF_MIX && lc.riemann != RIEMANN_PF_PHASIC) return hipErrorInvalidValue;
    const int form = lc.riemann - RIEMANN_PF_MIX;
#define CMP_CFPF_ARGS kp, lc.limiter, part, list, n, jobs, njobs, faces, cboxes, Uc, gqc, fboxes, Uf, gqf, ratio, Uout, \
                      dt, stage, err, s
#define CMP_CFPF(SC, NPV)                                                                                       \
    if (lc.scheme == SC) return form ? cfcorr_pf<DIM, SC, 1, NPV>(CMP_CFPF_ARGS) : cfcorr_pf<DIM, SC, 0, NPV>(CMP_CFPF_ARGS);
    if (lc.nphase == 3) { CMP_CFPF(5, 3) CMP_CFPF(1, 3) CMP_CFPF(0, 3) }
    else if (lc.nphase == 2) { CMP_CFPF(5, 2) CMP_CFPF(1, 2) CMP_CFPF(0, 2) }
    else { CMP_CFPF(5, 0) CMP_CFPF(6, 0) CMP_CFPF(1, 0) CMP_CFPF(0, 0) }
#undef CMP_CFPF
#undef CMP_CFPF_ARGS
    return hipErrorInvalidValue;
}

template <int DIM, int SCHEME, int RIEMANN, bool THINC, int NP = 0>
static hipError_t cfcorr4(const KParams& kp, int limiter, int part, const CfCell* list, int n, const CfFace* jobs,
                          int njobs, double* faces, const BoxDev* cboxes, const double* Uc, const BoxDev* fboxes,
                          const double* Uf, int ratio, double* Uout, const double* dt, int stage, int* err,
                          hipStream_t s)
{
    if (part & 1)
        hipLaunchKernelGGL((k_cf_faces<DIM, SCHEME, RIEMANN, THINC, NP>), dim3((njobs + 63) / 64), dim3(64), 0, s,
                           kp, limiter, jobs, njobs, cboxes, Uc, fboxes, Uf, faces, err);
    if (part & 2)
        hipLaunchKernelGGL((k_cf_apply<DIM, NP>), dim3((n + 63) / 64), dim3(64), 0, s, kp, list, n, cboxes, Uc,
                           faces, ratio, Uout, dt, stage, err);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_cf_correct_dim(const LaunchCfg& lc, const KParams& kp, int part, const CfCell* list, int n,
                                 const CfFace* jobs, int njobs, double* faces, const BoxDev* cboxes,
                                 const double* Uc, const BoxDev* fboxes, const double* Uf, int ratio, double* Uout,
                                 const double* dt, int stage, int* err, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (lc.riemann == RIEMANN_PF_MIX || lc.riemann == RIEMANN_PF_PHASIC) return hipErrorInvalidValue;
#define CMP_CF_ARGS kp, lc.limiter, part, list, n, jobs, njobs, faces, cboxes, Uc, fboxes, Uf, ratio, Uout, dt, stage, \
                    err, s
    if (lc.nphase) {
        // the N-phase model's schemes (Godunov, MUSCL, WENO5 with the consistent limiter), HLLC or HLL
        if (lc.thinc) return hipErrorInvalidValue;
#define CMP_CF_NP(SC, RS)                                                                                  \
        if (lc.scheme == SC && lc.riemann == RS) {                                                        \
            if (lc.nphase == 3) return cfcorr4<DIM, SC, RS, false, 3>(CMP_CF_ARGS);                       \
            if (lc.nphase == 2) return cfcorr4<DIM, SC, RS, false, 2>(CMP_CF_ARGS);                       \
        }
        CMP_CF_NP(5, 2) CMP_CF_NP(5, 1) CMP_CF_NP(1, 2) CMP_CF_NP(1, 1) CMP_CF_NP(0, 2) CMP_CF_NP(0, 1)
#undef CMP_CF_NP
        return hipErrorInvalidValue;
    }
    CMP_DISPATCH(cfcorr4, CMP_CF_ARGS);
#undef CMP_CF_ARGS
    return hipErrorInvalidValue;
}

template <int DIM>
hipError_t launch_fill2_plan_dim(const FillDev& fd, const FillDev& cd, int ratio, void* plan, hipStream_t s)
{
    long long cols = (fd.maxghost + NT - 1) / NT;
    if (cols > 65536) cols = 65536;
    if (cols <= 0) return hipSuccess;
    for (int box0 = 0; box0 < fd.nbox; box0 += 65535) {
        const int rows = (fd.nbox - box0 < 65535) ? fd.nbox - box0 : 65535;
        hipLaunchKernelGGL((k_fill2_plan<DIM>), dim3((unsigned)cols, (unsigned)rows), dim3(NT), 0, s, fd, cd, box0, ratio,
                           (F2Cell<DIM>*)plan);
    }
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_fill2_apply_dim(int N, const FillDev& fd, const FillDev& cd, long long nghost, int ratio,
                                  const void* plan, double* Uf, const double* Uc, int* err, hipStream_t s)
{
    if (N != 4 + DIM && N != 5 + DIM && N != 7 + DIM) return hipErrorInvalidValue;
    if (nghost <= 0) return hipSuccess;
    long long blocks = (nghost + NT - 1) / NT;
    if (blocks > (1LL << 20)) blocks = 1LL << 20;
    const dim3 grid((unsigned)blocks);
    const F2Cell<DIM>* pl = (const F2Cell<DIM>*)plan;
    if (N == 4 + DIM)
        hipLaunchKernelGGL((k_fill2_apply<DIM, 4 + DIM>), grid, dim3(NT), 0, s, fd, cd, nghost, ratio, pl, Uf, Uc, err);
    else if (N == 5 + DIM)
        hipLaunchKernelGGL((k_fill2_apply<DIM, 5 + DIM>), grid, dim3(NT), 0, s, fd, cd, nghost, ratio, pl, Uf, Uc, err);
    else
        hipLaunchKernelGGL((k_fill2_apply<DIM, 7 + DIM>), grid, dim3(NT), 0, s, fd, cd, nghost, ratio, pl, Uf, Uc, err);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_fill2_dim(int N, const FillDev& fd, const FillDev* sfd, const FillDev& cd, int ratio, double* Uf,
                            const double* Us, const double* Uc, long long maxcells, int all, int* err, hipStream_t s)
{
    if (N != 4 + DIM && N != 5 + DIM && N != 7 + DIM && (N < 1 || N > 3)) return hipErrorInvalidValue;
    long long cols = ((all ? maxcells : fd.maxghost) + NT - 1) / NT;
    if (cols <= 0) return hipSuccess;
    if (cols > 65536) cols = 65536;
    // sfd null: no same-level source (every cell interpolated from the coarse level)
    const FillDev& sf = sfd ? *sfd : fd;
    const int has = sfd ? 1 : 0;
    for (int box0 = 0; box0 < fd.nbox; box0 += 65535) {
        const int rows = (fd.nbox - box0 < 65535) ? fd.nbox - box0 : 65535;
        const dim3 grid((unsigned)cols, (unsigned)rows);
#define CMP_F2(NN)                                                                                              \
        if (all) hipLaunchKernelGGL((k_fill2<DIM, NN, true>), grid, dim3(NT), 0, s, fd, sf, has, cd, box0, ratio, Uf, \
                                    Us, Uc, err);                                                               \
        else hipLaunchKernelGGL((k_fill2<DIM, NN, false>), grid, dim3(NT), 0, s, fd, sf, has, cd, box0, ratio, Uf,    \
                                Us, Uc, err);
        if (N == 4 + DIM) { CMP_F2(4 + DIM) }
        else if (N == 5 + DIM) { CMP_F2(5 + DIM) }
        else if (N == 7 + DIM) { CMP_F2(7 + DIM) }
        else if (N == 3) { CMP_F2(3) }
        else if (N == 2) { CMP_F2(2) }
        else { CMP_F2(1) }
#undef CMP_F2
    }
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_avgdown_dim(const BoxDev* fboxes, int nfbox, long long maxcells, const FillDev& cfd,
                              const double* Uf, double* Uc, int N, int ratio, double volfrac, int* err, hipStream_t s)
{
    const dim3 grid((unsigned)((maxcells + NT - 1) / NT), (unsigned)nfbox);
    hipLaunchKernelGGL((k_avgdown<DIM>), grid, dim3(NT), 0, s, fboxes, cfd, Uf, Uc, N, ratio, volfrac, err);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_avgdown_plan_dim(const BoxDev* fboxes, const long long* cprefix, int nfbox, long long total,
                                   const FillDev& cfd, int ratio, AvgCell* plan, hipStream_t s)
{
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_avgdown_plan<DIM>), dim3((unsigned)((total + NT - 1) / NT)), dim3(NT), 0, s, fboxes, cprefix,
                       nfbox, total, cfd, ratio, plan);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_avgdown_apply_dim(const BoxDev* fboxes, const BoxDev* cboxes, const AvgCell* plan, long long total,
                                    const double* Uf, double* Uc, int N, int ratio, double volfrac, int* err,
                                    hipStream_t s)
{
    const long long n = total * N;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_avgdown_apply<DIM>), dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, s, fboxes, cboxes, plan,
                       total, Uf, Uc, N, ratio, volfrac, err);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_faces_dim(const LaunchCfg& lc, const KParams& kp, const BoxDev& box, int dir,
                            const double* U, double* flux, double* fluxnc, double* cmc, hipStream_t s)
{
    if (lc.nphase) return hipErrorInvalidValue;   // face arrays of the N-phase model are not built
    CMP_DISPATCH(faces4, kp, lc.limiter, box, dir, U, flux, fluxnc, cmc, s);
    return hipErrorInvalidValue;
}

template <int DIM>
hipError_t launch_fill_dim(int N, const FillDev& fd, long long nghost, const int* list, int nlist, double* U,
                           int* err,
                           hipStream_t s)
{
    if (nghost <= 0) return hipSuccess;
    if (N != 4 + DIM && N != 5 + DIM && N != 7 + DIM && (N < 1 || N > 3)) return hipErrorInvalidValue;
    // grid rows are boxes (at most 65535 per launch); grid columns cover the largest box's ghosts
    long long cols = (fd.maxghost + NT - 1) / NT;
    if (cols > 65536) cols = 65536;
    const int nb = list ? nlist : fd.nbox;
    for (int box0 = 0; box0 < nb; box0 += 65535) {
        const int rows = (nb - box0 < 65535) ? nb - box0 : 65535;
        const dim3 grid((unsigned)cols, (unsigned)rows);
        if (N == 4 + DIM)
            hipLaunchKernelGGL((k_fill<DIM, 4 + DIM>), grid, dim3(NT), 0, s, fd, list, box0, U, err);
        else if (N == 5 + DIM)
            hipLaunchKernelGGL((k_fill<DIM, 5 + DIM>), grid, dim3(NT), 0, s, fd, list, box0, U, err);
        else if (N == 7 + DIM)
            hipLaunchKernelGGL((k_fill<DIM, 7 + DIM>), grid, dim3(NT), 0, s, fd, list, box0, U, err);
        // a Phase-Field field: 1 to 3 components (volume fractions, or the cell normals)
        else if (N == 3)
            hipLaunchKernelGGL((k_fill<DIM, 3>), grid, dim3(NT), 0, s, fd, list, box0, U, err);
        else if (N == 2)
            hipLaunchKernelGGL((k_fill<DIM, 2>), grid, dim3(NT), 0, s, fd, list, box0, U, err);
        else
            hipLaunchKernelGGL((k_fill<DIM, 1>), grid, dim3(NT), 0, s, fd, list, box0, U, err);
    }
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_dt_dim(const double* dx, double* cmax, double* dt, double cfl, int limiter,
                         double* hist, int* step, int* err, hipStream_t s, double* clock)
{
    hipLaunchKernelGGL((k_compute_dt<DIM>), dim3(1), dim3(64), 0, s, dx[0], dx[1], dx[2], cmax, dt,
                       cfl, limiter, hist, step, err, clock);
    return hipGetLastError();
}

template <int DIM>
void tile_dims_dim(int* tx, int* ty, int* tz)
{
    *tx = Tile<DIM>::X;
    *ty = Tile<DIM>::Y;
    *tz = Tile<DIM>::Z;
}

template void tile_dims_dim<CMP_KERNEL_DIM>(int*, int*, int*);
template hipError_t launch_cf_apply_all_dim<CMP_KERNEL_DIM>(const LaunchCfg&, const CfPairDev*, int, int,
                                                           const CfStatePtrs&, int, const double*, int, hipStream_t);
template hipError_t launch_cf_correct_dim<CMP_KERNEL_DIM>(const LaunchCfg&, const KParams&, int, const CfCell*, int,
                                                         const CfFace*, int, double*, const BoxDev*,
                                                         const double*, const BoxDev*, const double*, int,
                                                         double*, const double*, int, int*, hipStream_t);
template hipError_t launch_cf_correct_pf_dim<CMP_KERNEL_DIM>(const LaunchCfg&, const KParams&, int, const CfCell*, int,
                                                            const CfFace*, int, double*, const BoxDev*, const double*,
                                                            const double*, const BoxDev*, const double*, const double*,
                                                            int, double*, const double*, int, int*, hipStream_t);
template hipError_t launch_fill2_dim<CMP_KERNEL_DIM>(int, const FillDev&, const FillDev*, const FillDev&, int,
                                                     double*, const double*, const double*, long long, int, int*,
                                                     hipStream_t);
template hipError_t launch_fill2_plan_dim<CMP_KERNEL_DIM>(const FillDev&, const FillDev&, int, void*, hipStream_t);
template hipError_t launch_fill2_apply_dim<CMP_KERNEL_DIM>(int, const FillDev&, const FillDev&, long long, int,
                                                           const void*, double*, const double*, int*, hipStream_t);
template hipError_t launch_avgdown_dim<CMP_KERNEL_DIM>(const BoxDev*, int, long long, const FillDev&, const double*,
                                                       double*, int, int, double, int*, hipStream_t);
template hipError_t launch_avgdown_plan_dim<CMP_KERNEL_DIM>(const BoxDev*, const long long*, int, long long,
                                                            const FillDev&, int, AvgCell*, hipStream_t);
template hipError_t launch_avgdown_apply_dim<CMP_KERNEL_DIM>(const BoxDev*, const BoxDev*, const AvgCell*, long long,
                                                             const double*, double*, int, int, double, int*,
                                                             hipStream_t);
template hipError_t launch_tag_dim<CMP_KERNEL_DIM>(const LaunchCfg&, const KParams&, const BoxDev*, int, long long,
                                                   const long long*, const double*, const TagCriteria&,
                                                   unsigned char*, int*, hipStream_t);
template hipError_t launch_stage_batch_dim<CMP_KERNEL_DIM>(const LaunchCfg&, const KParams*, int, const TileDev*, int,
                                                           const BoxDev*, const double*, const double*, double*,
                                                           const double*, unsigned long long*, int*,
                                                           const FillDev*, hipStream_t);
template hipError_t launch_stage_dim<CMP_KERNEL_DIM>(const LaunchCfg&, const KParams&, int, int,
                                                     const TileDev*, int, const BoxDev*, const BoxDev*,
                                                     int, const double*, const double*, double*,
                                                     const double*, unsigned long long*, int*,
                                                     hipStream_t, const double*, const FillDev*);
template hipError_t launch_faces_dim<CMP_KERNEL_DIM>(const LaunchCfg&, const KParams&, const BoxDev&,
                                                     int, const double*, double*, double*, double*, hipStream_t);
template hipError_t launch_fill_dim<CMP_KERNEL_DIM>(int, const FillDev&, long long, const int*, int, double*, int*,
                                                    hipStream_t);
template hipError_t launch_dt_dim<CMP_KERNEL_DIM>(const double*, double*, double*, double, int,
                                                  double*, int*, int*, hipStream_t, double*);
#else   // CMP_LIST_UNIT

// ---- persistent work-list stage kernel ------------------------------------------------------------
// A fixed grid of workgroups (as many as are resident at once) walks one list of tiles, each with
// its own level and shape (the tile shape the autotuner chose for its box's shape); the tiles of
// every level of a hierarchy go into one launch. The list is cut into 8 consecutive chunks, one per
// XCD: workgroup b starts on chunk b mod 8 (workgroups are dispatched round-robin over the XCDs, so
// a chunk's tiles, whose halos overlap, meet in one XCD's L2), takes its tiles one at a time from
// the chunk's device counter and, once the chunk is exhausted, moves on to the following chunks.
// Every workgroup leaves after it has found all 8 chunks exhausted; the last one to leave sets the
// counters back to zero for the next launch. W: the resident waves per SIMD the register allocation
// must leave room for (2, or 3 at the price of some register spills: the autotuner measures both).
template <int DIM, int SCHEME, int RIEMANN, bool THINC, int NP, int W>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(W)))
void k_stage_list(const ListLevel* __restrict__ levs, const ListTile* __restrict__ tiles, int ntiles, int* ctr,
                  ListCall call, int limiter, int mode, int stage, const double* __restrict__ dtp)
{
    constexpr int N = NP ? 2 * NP + DIM + 1 : 4 + DIM;
    constexpr int G = Stencil<SCHEME>::G;
    __shared__ double sWF[lds_doubles<N, LdsOf<DIM, G, ShapeRT>>()];
    __shared__ double sRed[NT / 64];
    __shared__ int sT;
    // a step of size zero leaves every state as it is (and the counters untouched)
    if (mode == MODE_RK && *dtp == 0.0) return;
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x & 7;
    int tried = 0;   // chunks found exhausted (thread 0)
    for (;;) {
        if (tid == 0) {
            int t = -1;
            while (tried < 8) {
                const int c = (x0 + tried) & 7;
                const int lo = (int)(((long long)ntiles * c) >> 3);
                const int hi = (int)(((long long)ntiles * (c + 1)) >> 3);
                if (lo < hi) {
                    const int k = atomicAdd(ctr + c, 1);
                    if (k < hi - lo) {
                        t = lo + k;
                        break;
                    }
                }
                ++tried;
            }
            sT = t;
        }
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(sT);
        __syncthreads();
        if (t < 0) break;
        const ListTile td = tiles[t];
        const int l = td.lev;
        const ListLevel& lv = levs[l];
        const BoxDev bx = lv.boxes[td.box];
        ShapeRT sh;
        sh.e[0] = td.ext[0];
        sh.e[1] = td.ext[1];
        sh.e[2] = td.ext[2];
        const int tlo[3] = {td.lo[0], td.lo[1], td.lo[2]};
        int halo[3][2] = {{HALO_LOOKUP, HALO_LOOKUP}, {HALO_LOOKUP, HALO_LOOKUP}, {HALO_LOOKUP, HALO_LOOKUP}};
        if (lv.fd) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                halo[d][0] = td.halo[d][0];
                halo[d][1] = td.halo[d][1];
            }
        }
        stage_tile<DIM, SCHEME, RIEMANN, THINC, NP>(lv.kp, limiter, mode, stage, bx, tlo, sh, halo, lv.boxes,
                                                    call.Us[l], call.Un[l], call.Uout[l], dtp, call.cmax[l], lv.err,
                                                    nullptr, lv.fd, sWF, sRed);
    }
    if (tid == 0) {
        __threadfence();
        if (atomicAdd(ctr + 8, 1) == (int)gridDim.x - 1) {
            // every other workgroup has left: no counter is taken from any more
#pragma unroll
            for (int c = 0; c < LIST_NCTR; ++c) atomicExch(ctr + c, 0);
        }
    }
}

template <int DIM, int SCHEME, int RIEMANN, bool THINC, int NP = 0>
static hipError_t list4(int waves, int query, int* cap, const ListLevel* levs, const ListTile* tiles, int ntiles,
                        int* ctr, const ListCall& call, int limiter, int mode, int stage, const double* dt, int grid,
                        hipStream_t s)
{
    if (waves != 2 && waves != 3) return hipErrorInvalidValue;
    auto kern = (waves == 3) ? k_stage_list<DIM, SCHEME, RIEMANN, THINC, NP, 3>
                             : k_stage_list<DIM, SCHEME, RIEMANN, THINC, NP, 2>;
    if (query) {
        int per_cu = 0, dev = 0, cus = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, NT, 0);
        if (e == hipSuccess) e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        *cap = (e == hipSuccess) ? per_cu * cus : 0;
        return e;
    }
    if (ntiles <= 0) return hipSuccess;
    if (grid <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kern, dim3(grid < ntiles ? grid : ntiles), dim3(NT), 0, s, levs, tiles, ntiles, ctr, call,
                       limiter, mode, stage, dt);
    return hipGetLastError();
}

// query != 0: the resident-workgroup capacity of lc's kernel into *cap, nothing launched
template <int DIM>
hipError_t launch_stage_list_dim(const LaunchCfg& lc, int waves, int query, int* cap, const ListLevel* levs,
                                 const ListTile* tiles, int ntiles, int* ctr, const ListCall& call, int mode,
                                 int stage, const double* dt, int grid, hipStream_t s)
{
    if (lc.riemann != 1 && lc.riemann != 2) return hipErrorInvalidValue;
#define CMP_LIST_ARGS waves, query, cap, levs, tiles, ntiles, ctr, call, lc.limiter, mode, stage, dt, grid, s
    if (lc.nphase) {
        if (lc.thinc) return hipErrorInvalidValue;
#define CMP_NPL_CASE(SC)                                                                                \
        if (lc.scheme == SC) {                                                                          \
            if (lc.nphase == 3) return lc.riemann == 2 ? list4<DIM, SC, 2, false, 3>(CMP_LIST_ARGS)     \
                                                       : list4<DIM, SC, 1, false, 3>(CMP_LIST_ARGS);    \
            if (lc.nphase == 2) return lc.riemann == 2 ? list4<DIM, SC, 2, false, 2>(CMP_LIST_ARGS)     \
                                                       : list4<DIM, SC, 1, false, 2>(CMP_LIST_ARGS);    \
        }
        CMP_NPL_CASE(5)
        CMP_NPL_CASE(1)
        CMP_NPL_CASE(0)
#undef CMP_NPL_CASE
        return hipErrorInvalidValue;
    }
    const bool th = lc.thinc != 0;
    if (lc.riemann == 2) {
        if (lc.scheme == 5) return th ? list4<DIM, 5, 2, true>(CMP_LIST_ARGS) : list4<DIM, 5, 2, false>(CMP_LIST_ARGS);
        if (lc.scheme == 6) return th ? list4<DIM, 6, 2, true>(CMP_LIST_ARGS) : list4<DIM, 6, 2, false>(CMP_LIST_ARGS);
        if (lc.scheme == 1 && !th) return list4<DIM, 1, 2, false>(CMP_LIST_ARGS);
        if (lc.scheme == 0 && !th) return list4<DIM, 0, 2, false>(CMP_LIST_ARGS);
    } else {
        if (lc.scheme == 5) return th ? list4<DIM, 5, 1, true>(CMP_LIST_ARGS) : list4<DIM, 5, 1, false>(CMP_LIST_ARGS);
        if (lc.scheme == 6) return th ? list4<DIM, 6, 1, true>(CMP_LIST_ARGS) : list4<DIM, 6, 1, false>(CMP_LIST_ARGS);
        if (lc.scheme == 1 && !th) return list4<DIM, 1, 1, false>(CMP_LIST_ARGS);
        if (lc.scheme == 0 && !th) return list4<DIM, 0, 1, false>(CMP_LIST_ARGS);
    }
#undef CMP_LIST_ARGS
    return hipErrorInvalidValue;
}

// Whether ext fits k_stage_list's layout: at most NT cells, every direction's staged region within
// LdsRT::RMAX cells and its faces within NT + LdsRT::XS.
template <int DIM>
bool tile_shape_ok_dim(const LaunchCfg& lc, const int* ext)
{
    const int G = (lc.scheme >= 5) ? 3 : (lc.scheme == 1 ? 2 : 1);
    const int rmax = (DIM == 2) ? 32 * (8 + 2 * G) : 64 * (4 + 2 * G);
    const int xs = (DIM == 2) ? 32 : 64;
    const int e[3] = {ext[0], ext[1], DIM == 3 ? ext[2] : 1};
    for (int d = 0; d < 3; ++d)
        if (e[d] < 1 || e[d] > 255) return false;
    if (e[0] * e[1] * e[2] > NT) return false;
    for (int d = 0; d < DIM; ++d) {
        int r = 1, f = 1;
        for (int m = 0; m < 3; ++m) {
            r *= e[m] + (m == d ? 2 * G : 0);
            f *= e[m] + (m == d ? 1 : 0);
        }
        if (r > rmax || f > NT + xs) return false;
    }
    return true;
}

template hipError_t launch_stage_list_dim<CMP_KERNEL_DIM>(const LaunchCfg&, int, int, int*, const ListLevel*,
                                                          const ListTile*, int, int*, const ListCall&, int, int,
                                                          const double*, int, hipStream_t);
template bool tile_shape_ok_dim<CMP_KERNEL_DIM>(const LaunchCfg&, const int*);
#endif  // CMP_LIST_UNIT

}  // namespace cmp
