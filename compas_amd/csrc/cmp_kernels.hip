// cmp_kernels.hip - gfx950 kernels of the five-equation RHS, RK3 stage, ghost fill and CFL.
//
// Built twice, with CMP_KERNEL_DIM=2 and CMP_KERNEL_DIM=3, so that the two dimensions compile in
// parallel; cmp_dispatch.hip routes a call to the right instantiation.
//
// Stage kernel design (one workgroup = one tile of a box, 256 threads = 4 wave64):
//   for each direction d:
//     1. stage the primitives of the tile plus G cells on both sides along d in LDS
//        (cons->prim once per staged cell instead of once per stencil use),
//     2. each face of the tile along d is computed once: WENO5/MUSCL/Godunov on the staged
//        primitives, prim->cons, HLLC/HLL; flux and u*_n go to LDS,
//     3. each cell accumulates (F_{+} - F_{-})/dx_d and (u*_{+} - u*_{-})/dx_d in registers;
//   then dU = -div F, dU[alpha1] += phi1 (1+K1) div u*, and the RK3 combination writes the new
//   state directly: dU never goes to HBM. The face wave speeds of the tile are max-reduced with
//   wave shuffles and one 64-bit atomicMax per workgroup.
#include "cmp_kernels.hpp"

#ifndef CMP_KERNEL_DIM
#error "CMP_KERNEL_DIM must be 2 or 3"
#endif

namespace cmp {

template <int DIM> struct Tile;
template <> struct Tile<2> { static constexpr int X = 16, Y = 16, Z = 1; };
template <> struct Tile<3> { static constexpr int X = 8, Y = 8, Z = 4; };

template <int SCHEME> struct Stencil;
template <> struct Stencil<0> { static constexpr int G = 1; };   // Godunov
template <> struct Stencil<1> { static constexpr int G = 2; };   // MUSCL
template <> struct Stencil<5> { static constexpr int G = 3; };   // WENO5

constexpr int NT = 256;

constexpr int cmax3(int a, int b, int c) { return (a > b ? (a > c ? a : c) : (b > c ? b : c)); }

template <int DIM, int G>
struct Lds {
    static constexpr int TX = Tile<DIM>::X, TY = Tile<DIM>::Y, TZ = Tile<DIM>::Z;
    static constexpr int RMAX = cmax3((TX + 2 * G) * TY * TZ, TX * (TY + 2 * G) * TZ,
                                      DIM == 3 ? TX * TY * (TZ + 2 * G) : 0);
    static constexpr int FMAX = cmax3((TX + 1) * TY * TZ, TX * (TY + 1) * TZ,
                                      DIM == 3 ? TX * TY * (TZ + 1) : 0);
};

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = smax(v, __shfl_xor(v, off, 64));
    return v;
}

// Maps a workgroup to a tile so that the workgroups one XCD receives (blocks b, b+8, b+16, ...
// under round-robin dispatch) hold consecutive tiles, whose halos overlap in that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

template <int DIM, int SCHEME, int RIEMANN>
__global__ __launch_bounds__(NT)
void k_stage(KParams kp, int limiter, int mode, int stage,
             const TileDev* __restrict__ tiles, int ntiles, const BoxDev* __restrict__ boxes,
             BoxDev sbox, int sntx, int snty, int single,
             const double* __restrict__ Us, const double* Un, double* Uout,
             const double* __restrict__ dtp, unsigned long long* cmax_bits, int* err)
{
    constexpr int N = 4 + DIM;
    constexpr int G = Stencil<SCHEME>::G;
    using L = Lds<DIM, G>;
    constexpr int TX = L::TX, TY = L::TY, TZ = L::TZ;
    constexpr int RMAX = L::RMAX, FMAX = L::FMAX;

    __shared__ double sW[N * RMAX];
    __shared__ double sF[(N + 1) * FMAX];
    __shared__ double sRed[NT / 64];

    const int tid = threadIdx.x;
    int bidx;
    int tlo[3];
    const BoxDev* bxp;
    if (single) {
        int b = blockIdx.x;
        const int tx = b % sntx;
        b /= sntx;
        const int ty = b % snty;
        const int tz = b / snty;
        bxp = &sbox;
        tlo[0] = sbox.lo[0] + tx * TX;
        tlo[1] = sbox.lo[1] + ty * TY;
        tlo[2] = sbox.lo[2] + (DIM == 3 ? tz * TZ : 0);
        bidx = 0;
    } else {
        const int t = xcd_tile(blockIdx.x, ntiles);
        const TileDev td = tiles[t];
        bidx = td.box;
        bxp = &boxes[bidx];
        tlo[0] = td.lo[0];
        tlo[1] = td.lo[1];
        tlo[2] = td.lo[2];
    }
    const BoxDev bx = *bxp;
    int n[3];
    n[0] = min(TX, bx.hi[0] - tlo[0] + 1);
    n[1] = min(TY, bx.hi[1] - tlo[1] + 1);
    n[2] = (DIM == 3) ? min(TZ, bx.hi[2] - tlo[2] + 1) : 1;

    const int cx = tid % TX, cy = (tid / TX) % TY, cz = tid / (TX * TY);
    const bool valid = (cx < n[0]) && (cy < n[1]) && (cz < n[2]);
    const double* __restrict__ Ub = Us + bx.off;

    int bad = 0;
    double cmax_loc = 0.0;
    double acc[N];
    double divu = 0.0;
#pragma unroll
    for (int q = 0; q < N; ++q) acc[q] = 0.0;

#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        int e[3] = {n[0], n[1], n[2]};
        e[d] += 2 * G;
        const int R = e[0] * e[1] * e[2];
        // 1. primitives of the region, staged in LDS
        for (int r = tid; r < R; r += NT) {
            const int x = r % e[0];
            const int y = (r / e[0]) % e[1];
            const int z = r / (e[0] * e[1]);
            const int i = tlo[0] + x - (d == 0 ? G : 0);
            const int j = tlo[1] + y - (d == 1 ? G : 0);
            const int k = tlo[2] + z - (d == 2 ? G : 0);
            const long long c = (long long)(i - bx.flo[0]) + bx.sj * (j - bx.flo[1]) + bx.sk * (k - bx.flo[2]);
            double Uc[N], Wc[N];
#pragma unroll
            for (int q = 0; q < N; ++q) Uc[q] = Ub[c + q * bx.sn];
            cons2prim<DIM>(kp, Uc, Wc, bad);
#pragma unroll
            for (int q = 0; q < N; ++q) sW[q * RMAX + r] = Wc[q];
        }
        __syncthreads();
        // 2. faces along d
        int fe[3] = {n[0], n[1], n[2]};
        fe[d] += 1;
        const int F = fe[0] * fe[1] * fe[2];
        const int rs = (d == 0) ? 1 : ((d == 1) ? e[0] : e[0] * e[1]);
        for (int f = tid; f < F; f += NT) {
            const int x = f % fe[0];
            const int y = (f / fe[0]) % fe[1];
            const int z = f / (fe[0] * fe[1]);
            const int a = (d == 0) ? x : ((d == 1) ? y : z);
            const int base = x + e[0] * (y + e[1] * z);
            double WL[N], WR[N];
#pragma unroll
            for (int q = 0; q < N; ++q) {
                const double* s = sW + q * RMAX + base;
                if (SCHEME == 5) {
                    const double v0 = s[0], v1 = s[rs], v2 = s[2 * rs], v3 = s[3 * rs], v4 = s[4 * rs],
                                 v5 = s[5 * rs];
                    WL[q] = weno5<1>(v0, v1, v2, v3, v4);
                    WR[q] = weno5<-1>(v1, v2, v3, v4, v5);
                } else if (SCHEME == 1) {
                    const double v0 = s[0], v1 = s[rs], v2 = s[2 * rs], v3 = s[3 * rs];
                    WL[q] = muscl(v0, v1, v2, kp.dx[d], 1, limiter);
                    WR[q] = muscl(v1, v2, v3, kp.dx[d], -1, limiter);
                } else {
                    WL[q] = s[0];
                    WR[q] = s[rs];
                }
            }
            double UL[N], UR[N], Fl[N], c;
            prim2cons<DIM>(kp, WL, UL, bad);
            prim2cons<DIM>(kp, WR, UR, bad);
            const double un = riemann<DIM, RIEMANN>(kp, UL, UR, d, Fl, c, bad);
#pragma unroll
            for (int q = 0; q < N; ++q) sF[q * FMAX + f] = Fl[q];
            sF[N * FMAX + f] = un;
            if (a < n[d]) {
                if (__builtin_isnan(c)) bad = CMP_ERR_NAN;
                cmax_loc = smax(cmax_loc, c);
            }
        }
        __syncthreads();
        // 3. flux differences of this direction
        if (valid) {
            const int fm = cx + fe[0] * (cy + fe[1] * cz);
            const int fs = (d == 0) ? 1 : ((d == 1) ? fe[0] : fe[0] * fe[1]);
            const int fp = fm + fs;
#pragma unroll
            for (int q = 0; q < N; ++q) {
                const double t = (sF[q * FMAX + fp] - sF[q * FMAX + fm]) / kp.dx[d];
                acc[q] = (d == 0) ? t : acc[q] + t;
            }
            const double t = (sF[N * FMAX + fp] - sF[N * FMAX + fm]) / kp.dx[d];
            divu = (d == 0) ? t : divu + t;
        }
        __syncthreads();
    }

    if (valid) {
        const int i = tlo[0] + cx, j = tlo[1] + cy, k = tlo[2] + cz;
        const long long c = (long long)(i - bx.flo[0]) + bx.sj * (j - bx.flo[1]) + bx.sk * (k - bx.flo[2]);
        double Uc[N], dU[N];
#pragma unroll
        for (int q = 0; q < N; ++q) {
            Uc[q] = Ub[c + q * bx.sn];
            dU[q] = -acc[q];
        }
        dU[3 + DIM] += nc_coefficient<DIM>(kp, Uc, bad) * divu;
        if (mode == MODE_RK) {
            const double dt = *dtp;
            double* Ob = Uout + bx.off;
            const double* Nb = Un + bx.off;
            if (stage == 0) {
#pragma unroll
                for (int q = 0; q < N; ++q) {
                    const double v = 1.0 * Uc[q] + dt * dU[q];
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
            atomicMax(cmax_bits, (unsigned long long)__double_as_longlong(v));
        }
    }
    if (bad && err) atomicMax(err, bad);
}

// Face fluxes of one box in the reference's shape (one thread per face, every stencil cell taken
// through cons->prim by the thread that uses it), for callers that need the face arrays (flux
// registers, face averaging) and as a second GPU path to check the fused kernel against.
template <int DIM, int SCHEME, int RIEMANN>
__global__ __launch_bounds__(NT)
void k_faces(KParams kp, int limiter, BoxDev bx, int dir, const double* __restrict__ U,
             double* __restrict__ flux, double* __restrict__ fluxnc)
{
    constexpr int N = 4 + DIM;
    constexpr int G = Stencil<SCHEME>::G;
    int fe[3];
    for (int m = 0; m < 3; ++m) fe[m] = (m < DIM) ? bx.hi[m] - bx.lo[m] + 1 : 1;
    fe[dir] += 1;
    const long long F = (long long)fe[0] * fe[1] * fe[2];
    const long long f = (long long)blockIdx.x * NT + threadIdx.x;
    if (f >= F) return;
    const int x = (int)(f % fe[0]);
    const int y = (int)((f / fe[0]) % fe[1]);
    const int z = (int)(f / ((long long)fe[0] * fe[1]));
    const int i = bx.lo[0] + x, j = bx.lo[1] + y, k = bx.lo[2] + z;
    const long long sd = (dir == 0) ? 1 : ((dir == 1) ? bx.sj : bx.sk);
    const long long c0 = (long long)(i - bx.flo[0]) + bx.sj * (j - bx.flo[1]) + bx.sk * (k - bx.flo[2]);
    const double* Ub = U + bx.off;
    double W[2 * G][N];
    int bad = 0;
#pragma unroll
    for (int t = 0; t < 2 * G; ++t) {
        double Uc[N];
        const long long c = c0 + (t - G) * sd;
#pragma unroll
        for (int q = 0; q < N; ++q) Uc[q] = Ub[c + q * bx.sn];
        cons2prim<DIM>(kp, Uc, W[t], bad);
    }
    double WL[N], WR[N];
#pragma unroll
    for (int q = 0; q < N; ++q) {
        if (SCHEME == 5) {
            WL[q] = weno5<1>(W[0][q], W[1][q], W[2][q], W[3][q], W[4][q]);
            WR[q] = weno5<-1>(W[1][q], W[2][q], W[3][q], W[4][q], W[5][q]);
        } else if (SCHEME == 1) {
            WL[q] = muscl(W[0][q], W[1][q], W[2][q], kp.dx[dir], 1, limiter);
            WR[q] = muscl(W[1][q], W[2][q], W[3][q], kp.dx[dir], -1, limiter);
        } else {
            WL[q] = W[0][q];
            WR[q] = W[1][q];
        }
    }
    double UL[N], UR[N], Fl[N], cm;
    prim2cons<DIM>(kp, WL, UL, bad);
    prim2cons<DIM>(kp, WR, UR, bad);
    const double un = riemann<DIM, RIEMANN>(kp, UL, UR, dir, Fl, cm, bad);
    if (flux) {
#pragma unroll
        for (int q = 0; q < N; ++q) flux[f + q * F] = Fl[q];
    }
    if (fluxnc) fluxnc[f] = un;
}

// Same-level ghost fill. Thread g handles ghost cell g of the level: the box comes from a binary
// search over ghost-count prefixes, the cell from the ghost shell's slab decomposition; the source
// is mapped through the physical BCs (foextrap clamp, reflect mirror with odd momentum signs) and
// periodic wrap, and read from the valid cells of the box owning it.
template <int DIM>
__global__ __launch_bounds__(NT)
void k_fill(FillDev fd, long long nghost, double* U, int* err)
{
    constexpr int N = 4 + DIM;
    for (long long g = (long long)blockIdx.x * NT + threadIdx.x; g < nghost;
         g += (long long)gridDim.x * NT) {
        int lo = 0, hi = fd.nbox;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (fd.gprefix[mid] <= g) lo = mid; else hi = mid;
        }
        const int b = lo;
        long long q = g - fd.gprefix[b];
        const BoxDev bx = fd.boxes[b];
        const int ng = fd.ng;
        int nv[3] = {bx.hi[0] - bx.lo[0] + 1, bx.hi[1] - bx.lo[1] + 1, DIM == 3 ? bx.hi[2] - bx.lo[2] + 1 : 1};
        const int ex = nv[0] + 2 * ng, ey = nv[1] + 2 * ng, ez = (DIM == 3) ? nv[2] + 2 * ng : 1;
        int x, y, z;
        long long r = q;
        if (DIM == 3) {
            const long long A = (long long)ng * ex * ey;
            if (r < A) {
                z = (int)(r / ((long long)ex * ey)); r %= (long long)ex * ey;
                y = (int)(r / ex); x = (int)(r % ex);
                goto mapped;
            }
            if (r < 2 * A) {
                r -= A;
                z = ez - ng + (int)(r / ((long long)ex * ey)); r %= (long long)ex * ey;
                y = (int)(r / ex); x = (int)(r % ex);
                goto mapped;
            }
            r -= 2 * A;
            const long long Rg = 2LL * ng * ex + 2LL * ng * nv[1];
            z = ng + (int)(r / Rg);
            r %= Rg;
        } else {
            z = 0;
        }
        if (r < (long long)ng * ex) {
            y = (int)(r / ex); x = (int)(r % ex);
        } else if (r < 2LL * ng * ex) {
            r -= (long long)ng * ex;
            y = ey - ng + (int)(r / ex); x = (int)(r % ex);
        } else {
            r -= 2LL * ng * ex;
            y = ng + (int)(r / (2 * ng));
            const int cc = (int)(r % (2 * ng));
            x = (cc < ng) ? cc : (ex - 2 * ng + cc);
        }
    mapped:
        int cell[3] = {bx.flo[0] + x, bx.flo[1] + y, bx.flo[2] + z};
        int src[3] = {cell[0], cell[1], cell[2]};
        unsigned flip = 0u;
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
        if (!ok) { atomicMax(err, CMP_ERR_BC); continue; }
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            if (src[d] < fd.dlo[d] || src[d] > fd.dhi[d]) ok = 0;
        if (!ok) { atomicMax(err, CMP_ERR_LAYOUT); continue; }
        long long bi = 0;
        for (int d = DIM - 1; d >= 0; --d)
            bi = bi * fd.nb[d] + (src[d] - fd.dlo[d]) / fd.bs;
        const int own = fd.owner[bi];
        if (own < 0) { atomicMax(err, CMP_ERR_LAYOUT); continue; }
        const BoxDev so = fd.boxes[own];
        const long long si = so.off + (src[0] - so.flo[0]) + so.sj * (src[1] - so.flo[1]) + so.sk * (src[2] - so.flo[2]);
        const long long di = bx.off + (cell[0] - bx.flo[0]) + bx.sj * (cell[1] - bx.flo[1]) + bx.sk * (cell[2] - bx.flo[2]);
#pragma unroll
        for (int c = 0; c < N; ++c) {
            double v = U[si + c * so.sn];
            if (c >= 2 && c < 2 + DIM && ((flip >> (c - 2)) & 1u)) v = -v;
            U[di + c * bx.sn] = v;
        }
    }
}

// ComputeDt on the device (source/COMPAS.cpp:2250-2302, EstTimeStep
// source/Class_Compressible_PhaseField.cpp:4473-4499). One thread.
// step[0] counts the steps of the current advance call, step[1] is the capacity of hist.
template <int DIM>
__global__ void k_compute_dt(double dx0, double dx1, double dx2, double* cmax, double* dt, double cfl,
                             int limiter, double* hist, int* step, int* err)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
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
    *dt = dtn;
    const int s = step[0];
    if (hist && s >= 0 && s < step[1]) hist[s] = dtn;
    step[0] = s + 1;
    *cmax = 0.0;
}

#if CMP_KERNEL_DIM == 2
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

template <int DIM, int SCHEME, int RIEMANN>
static hipError_t stage3(const KParams& kp, int limiter, int mode, int stage, const TileDev* tiles,
                         int ntiles, const BoxDev* boxes, const BoxDev* sbox, int sntiles,
                         const double* Us, const double* Un, double* Uout, const double* dt,
                         unsigned long long* cmax_bits, int* err, hipStream_t s)
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
    hipLaunchKernelGGL((k_stage<DIM, SCHEME, RIEMANN>), dim3(grid), dim3(NT), 0, s, kp, limiter, mode,
                       stage, tiles, ntiles, boxes, sb, sntx, snty, single, Us, Un, Uout, dt,
                       cmax_bits, err);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_stage_dim(const LaunchCfg& lc, const KParams& kp, int mode, int stage,
                            const TileDev* tiles, int ntiles, const BoxDev* boxes,
                            const BoxDev* sbox, int sntiles, const double* Us, const double* Un,
                            double* Uout, const double* dt, unsigned long long* cmax_bits, int* err,
                            hipStream_t s)
{
#define CMP_STAGE(SCH, RIE) \
    return stage3<DIM, SCH, RIE>(kp, lc.limiter, mode, stage, tiles, ntiles, boxes, sbox, sntiles, Us, Un, Uout, dt, cmax_bits, err, s)
    if (lc.riemann == 2) {
        if (lc.scheme == 5) CMP_STAGE(5, 2);
        if (lc.scheme == 1) CMP_STAGE(1, 2);
        if (lc.scheme == 0) CMP_STAGE(0, 2);
    } else {
        if (lc.scheme == 5) CMP_STAGE(5, 1);
        if (lc.scheme == 1) CMP_STAGE(1, 1);
        if (lc.scheme == 0) CMP_STAGE(0, 1);
    }
#undef CMP_STAGE
    return hipErrorInvalidValue;
}

template <int DIM>
hipError_t launch_faces_dim(const LaunchCfg& lc, const KParams& kp, const BoxDev& box, int dir,
                            const double* U, double* flux, double* fluxnc, hipStream_t s)
{
    long long F = 1;
    for (int m = 0; m < DIM; ++m) F *= (box.hi[m] - box.lo[m] + 1 + (m == dir ? 1 : 0));
    const int grid = (int)((F + NT - 1) / NT);
#define CMP_FACES(SCH, RIE)                                                                          \
    {                                                                                                \
        hipLaunchKernelGGL((k_faces<DIM, SCH, RIE>), dim3(grid), dim3(NT), 0, s, kp, lc.limiter, box, \
                           dir, U, flux, fluxnc);                                                    \
        return hipGetLastError();                                                                    \
    }
    if (lc.riemann == 2) {
        if (lc.scheme == 5) CMP_FACES(5, 2);
        if (lc.scheme == 1) CMP_FACES(1, 2);
        if (lc.scheme == 0) CMP_FACES(0, 2);
    } else {
        if (lc.scheme == 5) CMP_FACES(5, 1);
        if (lc.scheme == 1) CMP_FACES(1, 1);
        if (lc.scheme == 0) CMP_FACES(0, 1);
    }
#undef CMP_FACES
    return hipErrorInvalidValue;
}

template <int DIM>
hipError_t launch_fill_dim(const FillDev& fd, long long nghost, double* U, int* err, hipStream_t s)
{
    if (nghost <= 0) return hipSuccess;
    long long blocks = (nghost + NT - 1) / NT;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL((k_fill<DIM>), dim3((unsigned)blocks), dim3(NT), 0, s, fd, nghost, U, err);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_dt_dim(const double* dx, double* cmax, double* dt, double cfl, int limiter,
                         double* hist, int* step, int* err, hipStream_t s)
{
    hipLaunchKernelGGL((k_compute_dt<DIM>), dim3(1), dim3(64), 0, s, dx[0], dx[1], dx[2], cmax, dt,
                       cfl, limiter, hist, step, err);
    return hipGetLastError();
}

template hipError_t launch_stage_dim<CMP_KERNEL_DIM>(const LaunchCfg&, const KParams&, int, int,
                                                     const TileDev*, int, const BoxDev*, const BoxDev*,
                                                     int, const double*, const double*, double*,
                                                     const double*, unsigned long long*, int*,
                                                     hipStream_t);
template hipError_t launch_faces_dim<CMP_KERNEL_DIM>(const LaunchCfg&, const KParams&, const BoxDev&,
                                                     int, const double*, double*, double*, hipStream_t);
template hipError_t launch_fill_dim<CMP_KERNEL_DIM>(const FillDev&, long long, double*, int*,
                                                    hipStream_t);
template hipError_t launch_dt_dim<CMP_KERNEL_DIM>(const double*, double*, double*, double, int,
                                                  double*, int*, int*, hipStream_t);

}  // namespace cmp
