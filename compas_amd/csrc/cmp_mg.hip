// cmp_mg.hip - geometric multigrid V-cycle on dense domain-shaped grids, the preconditioner of the
// implicit Grad_Q solve (div(b grad Q) = Div_Jhat, the system the reference hands to AMReX's MLMG,
// source/Class_Compressible_PhaseField.cpp:2133-2310, include/Operators_LinearSystem/Solver_MLMG_AllLevels.H).
//
// Built twice, with CMP_KERNEL_DIM=2 and CMP_KERNEL_DIM=3. A level that covers its whole domain is
// gathered into a dense array (i fastest, no ghosts; the operator's periodic and Neumann sides are
// applied by index arithmetic, so no fill is needed), then coarsened by 2 in every direction while
// every extent is even: face coefficients of a coarse face are the mean of the fine faces on it,
// residuals restrict as the mean of the 2^DIM children, corrections prolong piecewise constant, the
// coarse operators are the same five- (seven-) point discretisation at the coarse spacing. The
// smoother is damped Jacobi, the same number of sweeps before and after the coarse correction, and
// the coarsest grid (at most 1024 cells) takes a fixed number of sweeps in one workgroup: the
// V-cycle is then a fixed symmetric positive definite operator, as a conjugate-gradient
// preconditioner must be. Every kernel returns at once when the solve's stop flag is set.
#include "cmp_kernels.hpp"

#ifndef CMP_KERNEL_DIM
#error "CMP_KERNEL_DIM must be 2 or 3"
#endif

namespace cmp {

namespace {

__device__ __forceinline__ double mg_wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int DIM>
__device__ __forceinline__ void mg_cell(const MgGrid& G, long long q, int& i, int& j, int& k)
{
    pos3(q, G.n[0], G.n[1], i, j, k);
    if (DIM != 3) k = 0;
}

__device__ __forceinline__ long long mg_idx(const MgGrid& G, int i, int j, int k)
{
    return (long long)i + (long long)G.n[0] * (j + (long long)G.n[1] * k);
}

// the face of direction d at the low side of cell (i,j,k): nodal extents n + e_d
__device__ __forceinline__ long long mg_face(const MgGrid& G, int d, int i, int j, int k)
{
    const int f0 = G.n[0] + (d == 0), f1 = G.n[1] + (d == 1);
    return (long long)i + (long long)f0 * (j + (long long)f1 * k);
}

// (A x)_c = -div(b grad x) at cell (i,j,k): sum_d (b_lo (x_c - x_m) - b_hi (x_p - x_c)) / h_d^2, with
// no flux through a Neumann domain face and periodic neighbours wrapped
template <int DIM>
__device__ __forceinline__ double mg_apply(const MgGrid& G, const double* __restrict__ x, const double* const* b,
                                           int i, int j, int k)
{
    const int co[3] = {i, j, k};
    const long long c = mg_idx(G, i, j, k);
    const double xc = x[c];
    double acc = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        int m[3] = {i, j, k}, p[3] = {i, j, k};
        const bool lo_dom = co[d] == 0, hi_dom = co[d] == G.n[d] - 1;
        m[d] = lo_dom ? G.n[d] - 1 : co[d] - 1;
        p[d] = hi_dom ? 0 : co[d] + 1;
        double flo = 0.0, fhi = 0.0;
        if (!lo_dom || G.periodic[d]) flo = b[d][mg_face(G, d, i, j, k)] * (xc - x[mg_idx(G, m[0], m[1], m[2])]);
        if (!hi_dom || G.periodic[d]) {
            int f[3] = {i, j, k};
            f[d] += 1;
            fhi = b[d][mg_face(G, d, f[0], f[1], f[2])] * (x[mg_idx(G, p[0], p[1], p[2])] - xc);
        }
        acc += (flo - fhi) * G.ih2[d];
    }
    return acc;
}

template <int DIM>
__global__ __launch_bounds__(NT) void k_mg_gather_r(PfGeom g, const double* __restrict__ sc,
                                                    const double* __restrict__ R, double* r0, MgGrid G)
{
    if (sc && sc[CG_DONE] != 0.0) return;
    long long q;
    const BoxDev ab = g.abox[cell_block(g.cmap, q)];
    const int n0 = ab.hi[0] - ab.lo[0] + 1, n1 = ab.hi[1] - ab.lo[1] + 1;
    const int n2 = (DIM == 3) ? ab.hi[2] - ab.lo[2] + 1 : 1;
    if (q >= (long long)n0 * n1 * n2) return;
    int i, j, k;
    pos3(q, n0, n1, i, j, k);
    i += ab.lo[0];
    j += ab.lo[1];
    k += ab.lo[2];
    r0[mg_idx(G, i - g.dlo[0], j - g.dlo[1], (DIM == 3) ? k - g.dlo[2] : 0)] = R[ab.off + fab_index(ab, i, j, k)];
}

template <int DIM>
__global__ __launch_bounds__(NT) void k_mg_scatter(PfGeom g, const double* __restrict__ sc,
                                                   const double* __restrict__ x0, double* Z, MgGrid G)
{
    if (sc && sc[CG_DONE] != 0.0) return;
    long long q;
    const BoxDev ab = g.abox[cell_block(g.cmap, q)];
    const int n0 = ab.hi[0] - ab.lo[0] + 1, n1 = ab.hi[1] - ab.lo[1] + 1;
    const int n2 = (DIM == 3) ? ab.hi[2] - ab.lo[2] + 1 : 1;
    if (q >= (long long)n0 * n1 * n2) return;
    int i, j, k;
    pos3(q, n0, n1, i, j, k);
    i += ab.lo[0];
    j += ab.lo[1];
    k += ab.lo[2];
    Z[ab.off + fab_index(ab, i, j, k)] = x0[mg_idx(G, i - g.dlo[0], j - g.dlo[1], (DIM == 3) ? k - g.dlo[2] : 0)];
}

// the box face buffer of direction D (cmp_face_layout, one component) -> the dense face array
template <int DIM, int D>
__global__ __launch_bounds__(NT) void k_mg_gather_b(PfGeom g, const double* __restrict__ Bf, double* b0, MgGrid G)
{
    const BoxDev ab = g.abox[blockIdx.y];
    const int n0 = ab.hi[0] - ab.lo[0] + 1 + (D == 0), n1 = ab.hi[1] - ab.lo[1] + 1 + (D == 1);
    const int n2 = (DIM == 3) ? ab.hi[2] - ab.lo[2] + 1 + (D == 2) : 1;
    const long long q = (long long)blockIdx.x * NT + threadIdx.x;
    if (q >= (long long)n0 * n1 * n2) return;
    int i, j, k;
    pos3(q, n0, n1, i, j, k);
    i += ab.lo[0];
    j += ab.lo[1];
    k += ab.lo[2];
    b0[mg_face(G, D, i - g.dlo[0], j - g.dlo[1], (DIM == 3) ? k - g.dlo[2] : 0)] = Bf[ab.goff[D] + q];
}

// a coarse face of direction d: the mean of the 2^(DIM-1) fine faces on it (the fine faces at twice
// its position along d, the two (four) transverse children)
template <int DIM>
__global__ __launch_bounds__(NT) void k_mg_restrict_b(const double* __restrict__ bf, double* bc, MgGrid Gf,
                                                      MgGrid Gc, int d)
{
    const int f0 = Gc.n[0] + (d == 0), f1 = Gc.n[1] + (d == 1), f2 = (DIM == 3) ? Gc.n[2] + (d == 2) : 1;
    const long long nf = (long long)f0 * f1 * f2;
    const int t1 = (d == 0) ? 1 : 0;                   // the transverse axes, t1 < t2
    const int t2 = (d == 2) ? 1 : 2;
    constexpr int NS = (DIM == 3) ? 4 : 2;
    for (long long q = (long long)blockIdx.x * NT + threadIdx.x; q < nf; q += (long long)gridDim.x * NT) {
        int co[3];
        pos3(q, f0, f1, co[0], co[1], co[2]);
        if (DIM != 3) co[2] = 0;
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < NS; ++m) {
            int f[3] = {2 * co[0], 2 * co[1], 2 * co[2]};
            f[t1] += m & 1;
            if (DIM == 3) f[t2] += m >> 1;
            s += bf[mg_face(Gf, d, f[0], f[1], f[2])];
        }
        bc[q] = s * (1.0 / (double)NS);
    }
}

// 1 / diag(A): sum of b / h_d^2 over the cell's faces that carry a flux
template <int DIM>
__global__ __launch_bounds__(NT) void k_mg_diag(const double* __restrict__ b0, const double* __restrict__ b1,
                                                const double* __restrict__ b2, double* id, MgGrid G)
{
    const double* b[3] = {b0, b1, b2};
    for (long long q = (long long)blockIdx.x * NT + threadIdx.x; q < G.nc; q += (long long)gridDim.x * NT) {
        int i, j, k;
        mg_cell<DIM>(G, q, i, j, k);
        const int co[3] = {i, j, k};
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            int f[3] = {i, j, k};
            f[d] += 1;
            if (co[d] > 0 || G.periodic[d]) s += b[d][mg_face(G, d, i, j, k)] * G.ih2[d];
            if (co[d] < G.n[d] - 1 || G.periodic[d]) s += b[d][mg_face(G, d, f[0], f[1], f[2])] * G.ih2[d];
        }
        id[q] = (s > 0.0) ? 1.0 / s : 0.0;
    }
}

// one damped Jacobi sweep y = x + omega D^-1 (r - A x); zero_x: x = 0 (y = omega D^-1 r)
template <int DIM>
__global__ __launch_bounds__(NT) void k_mg_smooth(const double* __restrict__ sc, const double* __restrict__ x,
                                                  double* __restrict__ y, const double* __restrict__ r,
                                                  const double* __restrict__ b0, const double* __restrict__ b1,
                                                  const double* __restrict__ b2, const double* __restrict__ id,
                                                  MgGrid G, double omega, int zero_x)
{
    if (sc && sc[CG_DONE] != 0.0) return;
    const double* b[3] = {b0, b1, b2};
    for (long long q = (long long)blockIdx.x * NT + threadIdx.x; q < G.nc; q += (long long)gridDim.x * NT) {
        if (zero_x) {
            y[q] = omega * id[q] * r[q];
        } else {
            int i, j, k;
            mg_cell<DIM>(G, q, i, j, k);
            y[q] = x[q] + omega * id[q] * (r[q] - mg_apply<DIM>(G, x, b, i, j, k));
        }
    }
}

// rc = mean over the 2^DIM children of (r - A x)
template <int DIM>
__global__ __launch_bounds__(NT) void k_mg_resid_restrict(const double* __restrict__ sc, const double* __restrict__ x,
                                                          const double* __restrict__ r, const double* __restrict__ b0,
                                                          const double* __restrict__ b1, const double* __restrict__ b2,
                                                          MgGrid Gf, double* rc, MgGrid Gc)
{
    if (sc && sc[CG_DONE] != 0.0) return;
    const double* b[3] = {b0, b1, b2};
    constexpr int NCH = (DIM == 3) ? 8 : 4;
    for (long long q = (long long)blockIdx.x * NT + threadIdx.x; q < Gc.nc; q += (long long)gridDim.x * NT) {
        int I, J, K;
        mg_cell<DIM>(Gc, q, I, J, K);
        double s = 0.0;
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            const int i = 2 * I + (ch & 1), j = 2 * J + ((ch >> 1) & 1), k = (DIM == 3) ? 2 * K + (ch >> 2) : 0;
            const long long c = mg_idx(Gf, i, j, k);
            s += r[c] - mg_apply<DIM>(Gf, x, b, i, j, k);
        }
        rc[q] = s * (1.0 / (double)NCH);
    }
}

// x_fine += x_coarse of the parent cell
template <int DIM>
__global__ __launch_bounds__(NT) void k_mg_prolong(const double* __restrict__ sc, const double* __restrict__ xc,
                                                   double* xf, MgGrid Gf, MgGrid Gc)
{
    if (sc && sc[CG_DONE] != 0.0) return;
    for (long long q = (long long)blockIdx.x * NT + threadIdx.x; q < Gf.nc; q += (long long)gridDim.x * NT) {
        int i, j, k;
        mg_cell<DIM>(Gf, q, i, j, k);
        xf[q] += xc[mg_idx(Gc, i >> 1, j >> 1, (DIM == 3) ? (k >> 1) : 0)];
    }
}

// the coarsest grid: `sweeps` damped Jacobi sweeps from x = 0 in one workgroup (one cell a thread)
template <int DIM>
__global__ __launch_bounds__(1024) void k_mg_bottom(const double* __restrict__ sc, const double* __restrict__ r,
                                                    double* x, const double* __restrict__ b0,
                                                    const double* __restrict__ b1, const double* __restrict__ b2,
                                                    const double* __restrict__ id, MgGrid G, double omega,
                                                    int sweeps)
{
    if (sc && sc[CG_DONE] != 0.0) return;
    __shared__ double xs[2][1024];
    const double* b[3] = {b0, b1, b2};
    const long long q = threadIdx.x;
    const bool on = q < G.nc;
    int i = 0, j = 0, k = 0;
    double rq = 0.0, iq = 0.0;
    if (on) {
        mg_cell<DIM>(G, q, i, j, k);
        rq = r[q];
        iq = id[q];
        xs[0][q] = omega * iq * rq;
    }
    __syncthreads();
    int cur = 0;
    for (int s = 1; s < sweeps; ++s) {
        if (on) xs[1 - cur][q] = xs[cur][q] + omega * iq * (rq - mg_apply<DIM>(G, xs[cur], b, i, j, k));
        __syncthreads();
        cur = 1 - cur;
    }
    if (on) x[q] = xs[cur][q];
}

// The coarse end of the V-cycle (MgTail) in one workgroup of 1024 threads, every operation a
// strided loop over the grid's cells followed by a barrier. The coarsest grids (from t.lds_from on)
// are kept in LDS, so that their many short phases wait on LDS rather than on device memory; the
// arithmetic is the same wherever a grid lives.
constexpr int NTAIL = 1024;
template <int DIM>
__global__ __launch_bounds__(NTAIL) void k_mg_tail(MgTail t)
{
    if (t.sc && t.sc[CG_DONE] != 0.0) return;
    __shared__ double lds[MG_TAIL_LDS];
    const int tid = threadIdx.x;
    const double om = t.omega;
    const int last = t.nlev - 1;
    auto in_lds = [&](int m) { return m >= t.lds_from; };
    auto X0 = [&](int m) -> double* { return in_lds(m) ? lds + t.loff[m][0] : t.x0[m]; };
    auto X1 = [&](int m) -> double* { return in_lds(m) ? lds + t.loff[m][1] : t.x1[m]; };
    auto Rr = [&](int m) -> double* { return in_lds(m) ? lds + t.loff[m][2] : t.r[m]; };
    auto ID = [&](int m) -> const double* { return in_lds(m) ? lds + t.loff[m][3] : t.id[m]; };
    auto Bf = [&](int m, int d) -> const double* { return in_lds(m) ? lds + t.loff[m][4 + d] : t.b[m][d]; };
    // the LDS grids' inverse diagonals and face coefficients
    for (int m = t.lds_from; m < t.nlev; ++m) {
        const MgGrid& G = t.G[m];
        for (long long q = tid; q < G.nc; q += NTAIL) lds[t.loff[m][3] + q] = t.id[m][q];
        for (int d = 0; d < DIM; ++d) {
            long long nf = 1;
            for (int e = 0; e < DIM; ++e) nf *= G.n[e] + (e == d);
            for (long long q = tid; q < nf; q += NTAIL) lds[t.loff[m][4 + d] + q] = t.b[m][d][q];
        }
    }
    __syncthreads();
    // y = x + omega id (r - A x) on grid m (zero_x: y = omega id r)
    auto smooth = [&](int m, const double* x, double* y, bool zero_x) {
        const MgGrid& G = t.G[m];
        const double* r = Rr(m);
        const double* id = ID(m);
        const double* bb[3] = {Bf(m, 0), Bf(m, 1), Bf(m, DIM == 3 ? 2 : 1)};
        for (long long q = tid; q < G.nc; q += NTAIL) {
            if (zero_x) {
                y[q] = om * id[q] * r[q];
            } else {
                int i, j, k;
                mg_cell<DIM>(G, q, i, j, k);
                y[q] = x[q] + om * id[q] * (r[q] - mg_apply<DIM>(G, x, bb, i, j, k));
            }
        }
        __syncthreads();
    };
    for (int m = 0; m < last; ++m) {
        smooth(m, nullptr, X0(m), true);
        smooth(m, X0(m), X1(m), false);
        // r of grid m + 1: the mean of the children of r - A x1
        const MgGrid& Gf = t.G[m];
        const MgGrid& Gc = t.G[m + 1];
        const double* rf = Rr(m);
        const double* x1 = X1(m);
        const double* bb[3] = {Bf(m, 0), Bf(m, 1), Bf(m, DIM == 3 ? 2 : 1)};
        double* rc = Rr(m + 1);
        constexpr int NCH = (DIM == 3) ? 8 : 4;
        for (long long q = tid; q < Gc.nc; q += NTAIL) {
            int I, J, K;
            mg_cell<DIM>(Gc, q, I, J, K);
            double sum = 0.0;
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) {
                const int i = 2 * I + (ch & 1), j = 2 * J + ((ch >> 1) & 1), k = (DIM == 3) ? 2 * K + (ch >> 2) : 0;
                const long long c = mg_idx(Gf, i, j, k);
                sum += rf[c] - mg_apply<DIM>(Gf, x1, bb, i, j, k);
            }
            rc[q] = sum * (1.0 / (double)NCH);
        }
        __syncthreads();
    }
    // the coarsest grid: `sweeps` damped Jacobi sweeps from zero, alternating x0 / x1, the result
    // copied to x0
    {
        const MgGrid& G = t.G[last];
        double* xa = X0(last);
        double* xb = X1(last);
        smooth(last, nullptr, xa, true);
        for (int sw = 1; sw < t.sweeps; ++sw) {
            smooth(last, xa, xb, false);
            double* tmp = xa;
            xa = xb;
            xb = tmp;
        }
        if (xa != X0(last)) {
            double* x0l = X0(last);
            for (long long q = tid; q < G.nc; q += NTAIL) x0l[q] = xa[q];
            __syncthreads();
        }
    }
    for (int m = last - 1; m >= 0; --m) {
        // x1 += the coarser grid's result at the parent cell, then two sweeps
        const MgGrid& Gf = t.G[m];
        const MgGrid& Gc = t.G[m + 1];
        const double* xc = (m + 1 == last) ? X0(m + 1) : X1(m + 1);
        double* x1 = X1(m);
        for (long long q = tid; q < Gf.nc; q += NTAIL) {
            int i, j, k;
            mg_cell<DIM>(Gf, q, i, j, k);
            x1[q] += xc[mg_idx(Gc, i >> 1, j >> 1, (DIM == 3) ? (k >> 1) : 0)];
        }
        __syncthreads();
        smooth(m, X1(m), X0(m), false);
        smooth(m, X0(m), X1(m), false);
    }
}

// the partial sums of one CG iteration's r.z, sum z and sum r (the second, third and fourth partial
// slots; the first holds r.r)
template <int DIM>
__global__ __launch_bounds__(NT) void k_cg_rz(PfGeom g, const double* __restrict__ sc, const double* __restrict__ R,
                                              const double* __restrict__ Z, double* part)
{
    if (sc && sc[CG_DONE] != 0.0) return;
    long long q;
    const BoxDev ab = g.abox[cell_block(g.cmap, q)];
    const int n0 = ab.hi[0] - ab.lo[0] + 1, n1 = ab.hi[1] - ab.lo[1] + 1;
    const int n2 = (DIM == 3) ? ab.hi[2] - ab.lo[2] + 1 : 1;
    double w[3] = {0.0, 0.0, 0.0};
    if (q < (long long)n0 * n1 * n2) {
        int i, j, k;
        pos3(q, n0, n1, i, j, k);
        i += ab.lo[0];
        j += ab.lo[1];
        k += ab.lo[2];
        const long long c = ab.off + fab_index(ab, i, j, k);
        const double r = R[c], z = Z[c];
        w[0] = r * z;
        w[1] = z;
        w[2] = r;
    }
    __shared__ double red[3][NT / 64];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const double t = mg_wave_sum(w[m]);
        if ((threadIdx.x & 63) == 0) red[m][threadIdx.x >> 6] = t;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int m = threadIdx.x;
        double s = red[m][0];
#pragma unroll
        for (int v = 1; v < NT / 64; ++v) s += red[m][v];
        const long long nb = (long long)gridDim.x * gridDim.y;
        part[(m + 1) * nb + (long long)blockIdx.y * gridDim.x + blockIdx.x] = s;
    }
}

dim3 mg_grid(long long n)
{
    long long b = (n + NT - 1) / NT;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

dim3 box_grid(long long maxn, int nbox)
{
    long long b = (maxn + NT - 1) / NT;
    if (b < 1) b = 1;
    return dim3((unsigned)b, (unsigned)nbox);
}

}  // namespace

template <int DIM>
hipError_t launch_mg_dim(int op, const MgArgs& a, hipStream_t s)
{
    switch (op) {
    case MG_GATHER_R:
        hipLaunchKernelGGL((k_mg_gather_r<DIM>), cell_grid(a.g.cmap, a.g.ncmap, a.maxcells, a.g.nbox), dim3(NT), 0, s, a.g, a.sc, a.src,
                           a.dst, a.G);
        break;
    case MG_SCATTER:
        hipLaunchKernelGGL((k_mg_scatter<DIM>), cell_grid(a.g.cmap, a.g.ncmap, a.maxcells, a.g.nbox), dim3(NT), 0, s, a.g, a.sc, a.src,
                           a.dst, a.G);
        break;
    case MG_GATHER_B:
        hipLaunchKernelGGL((k_mg_gather_b<DIM, 0>), box_grid(a.maxfaces[0], a.g.nbox), dim3(NT), 0, s, a.g, a.src,
                           a.b[0], a.G);
        hipLaunchKernelGGL((k_mg_gather_b<DIM, 1>), box_grid(a.maxfaces[1], a.g.nbox), dim3(NT), 0, s, a.g, a.src,
                           a.b[1], a.G);
        if (DIM == 3)
            hipLaunchKernelGGL((k_mg_gather_b<DIM, (DIM == 3 ? 2 : 1)>), box_grid(a.maxfaces[2], a.g.nbox), dim3(NT),
                               0, s, a.g, a.src, a.b[2], a.G);
        break;
    case MG_RESTRICT_B:
        for (int d = 0; d < DIM; ++d) {
            long long nf = 1;
            for (int e = 0; e < DIM; ++e) nf *= a.Gc.n[e] + (e == d);
            hipLaunchKernelGGL((k_mg_restrict_b<DIM>), mg_grid(nf), dim3(NT), 0, s, a.b[d], a.bc[d], a.G, a.Gc, d);
        }
        break;
    case MG_DIAG:
        hipLaunchKernelGGL((k_mg_diag<DIM>), mg_grid(a.G.nc), dim3(NT), 0, s, a.b[0], a.b[1], a.b[2], a.dst, a.G);
        break;
    case MG_SMOOTH:
        hipLaunchKernelGGL((k_mg_smooth<DIM>), mg_grid(a.G.nc), dim3(NT), 0, s, a.sc, a.src, a.dst, a.r, a.b[0],
                           a.b[1], a.b[2], a.id, a.G, a.omega, a.zero_x);
        break;
    case MG_RESID_RESTRICT:
        hipLaunchKernelGGL((k_mg_resid_restrict<DIM>), mg_grid(a.Gc.nc), dim3(NT), 0, s, a.sc, a.src, a.r, a.b[0],
                           a.b[1], a.b[2], a.G, a.dst, a.Gc);
        break;
    case MG_PROLONG:
        hipLaunchKernelGGL((k_mg_prolong<DIM>), mg_grid(a.G.nc), dim3(NT), 0, s, a.sc, a.src, a.dst, a.G, a.Gc);
        break;
    case MG_BOTTOM:
        if (a.G.nc > 1024) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_mg_bottom<DIM>), dim3(1), dim3(1024), 0, s, a.sc, a.r, a.dst, a.b[0], a.b[1], a.b[2],
                           a.id, a.G, a.omega, a.sweeps);
        break;
    case MG_CG_RZ:
        hipLaunchKernelGGL((k_cg_rz<DIM>), cell_grid(a.g.cmap, a.g.ncmap, a.maxcells, a.g.nbox), dim3(NT), 0, s, a.g, a.sc, a.src, a.r,
                           a.part);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template hipError_t launch_mg_dim<CMP_KERNEL_DIM>(int, const MgArgs&, hipStream_t);

template <int DIM>
hipError_t launch_mg_tail_dim(const MgTail& t, hipStream_t s)
{
    if (t.nlev < 1 || t.nlev > MG_TAIL_MAX_LEVELS) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_mg_tail<DIM>), dim3(1), dim3(NTAIL), 0, s, t);
    return hipGetLastError();
}

template hipError_t launch_mg_tail_dim<CMP_KERNEL_DIM>(const MgTail&, hipStream_t);

}  // namespace cmp
