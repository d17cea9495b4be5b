// cmp_amr_pf.hip - gfx950 kernels of the Phase-Field step of a hierarchy (PhaseField_Evolve over
// finest_level + 1 levels, source/Class_Compressible_PhaseField.cpp:425-594, without subcycling):
// AverageDownFaces of face fields, the composite implicit Grad_Q (the cell-centred operator over the
// cells no finer level covers, its Jacobi diagonal, the pieces of its multigrid preconditioner, the
// face gradients) and small vector kernels.
//
// Built twice, with CMP_KERNEL_DIM=2 and CMP_KERNEL_DIM=3. Fields are one-component level buffers
// with one ghost cell (the alpha layout) and one-component nodal face buffers; blockIdx.y is the box.
//
// The composite operator (volume weighted, A = -div(b grad .)): a face between two cells of one level
// that no finer level covers carries b (Q_n - Q_c) area / dx; a coarse-fine face carries, per fine
// face on it, b_fine (Q_coarse - Q_fine) area_fine / ((dx_fine + dx_coarse) / 2). The host lists the
// cells with a coarse-fine face (CompCell) and every link of theirs (CompLink); all other cells of a
// level take the five- (seven-) point stencil, their neighbours read from the ghost cells of the
// same-level fill.
#include "cmp_kernels.hpp"

#ifndef CMP_KERNEL_DIM
#error "CMP_KERNEL_DIM must be 2 or 3"
#endif

namespace cmp {

namespace {

__device__ __forceinline__ double ap_wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// block sum of v into part[idx] (every thread of the block calls it)
__device__ __forceinline__ void ap_block_sum(double v, double* part, long long idx)
{
    __shared__ double red[NT / 64];
    v = ap_wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = red[0];
#pragma unroll
        for (int w = 1; w < NT / 64; ++w) s += red[w];
        part[idx] = s;
    }
}

template <int DIM>
__device__ __forceinline__ bool ap_cell(const BoxDev& bx, long long q, int& i, int& j, int& k)
{
    const int n0 = bx.hi[0] - bx.lo[0] + 1, n1 = bx.hi[1] - bx.lo[1] + 1;
    const int n2 = (DIM == 3) ? bx.hi[2] - bx.lo[2] + 1 : 1;
    if (q >= (long long)n0 * n1 * n2) return false;
    pos3(q, n0, n1, i, j, k);
    i += bx.lo[0];
    j += bx.lo[1];
    k += bx.lo[2];
    return true;
}

// face index (one-component face buffer of the box, direction d) of cell (i,j,k)'s low (0) / high (1) face
__device__ __forceinline__ long long ap_face(const BoxDev& bx, int d, int i, int j, int k, int off)
{
    const int n0 = bx.hi[0] - bx.lo[0] + 1 + (d == 0), n1 = bx.hi[1] - bx.lo[1] + 1 + (d == 1);
    const int x = i - bx.lo[0] + (d == 0 ? off : 0);
    const int y = j - bx.lo[1] + (d == 1 ? off : 0);
    const int z = k - bx.lo[2] + (d == 2 ? off : 0);
    return bx.goff[d] + (long long)x + (long long)n0 * (y + (long long)n1 * z);
}

// the box of a level owning cell c after the periodic wrap (-1: outside the domain or no box)
template <int DIM>
__device__ __forceinline__ int ap_owner(const FillDev& fd, int* c)
{
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const int n = fd.dhi[d] - fd.dlo[d] + 1;
        if (c[d] < fd.dlo[d] || c[d] > fd.dhi[d]) {
            if (!fd.periodic[d]) return -1;
            c[d] = (c[d] < fd.dlo[d]) ? c[d] + n : c[d] - n;
        }
    }
    int bi = 0;
    for (int d = DIM - 1; d >= 0; --d) bi = bi * fd.nb[d] + (c[d] - fd.dlo[d]) / fd.bs;
    return fd.owner[bi];
}

__device__ __forceinline__ int ap_floor_div(int a, int r) { return (a >= 0) ? a / r : -((-a + r - 1) / r); }

// (A x)_c of one cell of level l (-1 mask: the regular stencil over the filled ghosts of x; >= 0: its
// link list); diag: the sum of the transmissibilities instead
template <int DIM>
__device__ __forceinline__ double ap_cell_op(const CompArgs& a, int l, const BoxDev& ab, int i, int j, int k,
                                             int mask, bool diag)
{
    const CompLev& L = a.L[l];
    const long long c = ab.off + fab_index(ab, i, j, k);
    const double xc = diag ? 0.0 : L.v[c];
    double acc = 0.0;
    if (mask >= 0) {
        const CompCell cc = a.cells[mask];
        for (int q = 0; q < cc.nlink; ++q) {
            const CompLink lk = a.links[cc.link0 + q];
            const double T = a.L[lk.flev].bf[lk.face] * lk.geo;
            acc += diag ? T : T * (xc - a.L[lk.lev].v[lk.idx]);
        }
        return acc;
    }
    const int co[3] = {i, j, k};
    const long long st[3] = {1, ab.sj, ab.sk};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const bool lo_dom = !L.fd.periodic[d] && co[d] == L.fd.dlo[d];
        const bool hi_dom = !L.fd.periodic[d] && co[d] == L.fd.dhi[d];
        const double g = L.area[d] / L.dx[d];
        if (!lo_dom) {
            const double T = L.bf[ap_face(ab, d, i, j, k, 0)] * g;
            acc += diag ? T : T * (xc - L.v[c - st[d]]);
        }
        if (!hi_dom) {
            const double T = L.bf[ap_face(ab, d, i, j, k, 1)] * g;
            acc += diag ? T : T * (xc - L.v[c + st[d]]);
        }
    }
    return acc;
}

}  // namespace

// AverageDownFaces (amrex::average_down_faces): every coarse face of the coarsened fine box `b`
// (blockIdx.y) of direction d takes the mean of the fine faces on it (the lower transverse index
// fastest, times 1/ratio^(DIM-1)); it is written into every coarse box holding it (the boxes of its
// two cells, periodic images included).
template <int DIM>
__global__ __launch_bounds__(NT) void k_avgdown_faces(const BoxDev* __restrict__ fboxes, FillDev cfd,
                                                      const double* __restrict__ Ff, double* Fc, int ratio, int d)
{
    const BoxDev fb = fboxes[blockIdx.y];
    int clo[3] = {0, 0, 0}, ce[3] = {1, 1, 1};
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
        clo[e] = fb.lo[e] / ratio;
        ce[e] = (fb.hi[e] + 1) / ratio - clo[e] + (e == d ? 1 : 0);
    }
    const long long q = (long long)blockIdx.x * NT + threadIdx.x;
    if (q >= (long long)ce[0] * ce[1] * ce[2]) return;
    int C[3];
    pos3(q, ce[0], ce[1], C[0], C[1], C[2]);
    C[0] += clo[0];
    C[1] += clo[1];
    C[2] += clo[2];
    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    const int r1 = ratio, r2 = (DIM == 3) ? ratio : 1;
    double s = 0.0;
    for (int o2 = 0; o2 < r2; ++o2)
        for (int o1 = 0; o1 < r1; ++o1) {
            int f[3] = {C[0] * ratio, C[1] * ratio, C[2] * ratio};
            f[t1] += o1;
            if (DIM == 3) f[t2] += o2;
            s += Ff[ap_face(fb, d, f[0], f[1], f[2], 0)];
        }
    int rt = 1;
#pragma unroll
    for (int e = 1; e < DIM; ++e) rt *= ratio;
    const double v = s * (1.0 / (double)rt);
    // the face is the low face of cell C and the high face of cell C - e_d
    for (int side = 0; side < 2; ++side) {
        int c[3] = {C[0], C[1], C[2]};
        if (side) c[d] -= 1;
        const int own = ap_owner<DIM>(cfd, c);
        if (own < 0) continue;
        const BoxDev cb = cfd.boxes[own];
        Fc[ap_face(cb, d, c[0], c[1], c[2], side)] = v;
    }
}

namespace {

// y += a x (n elements)
__global__ __launch_bounds__(NT) void k_axpy(long long n, double a, const double* __restrict__ x, double* y)
{
    for (long long q = (long long)blockIdx.x * NT + threadIdx.x; q < n; q += (long long)gridDim.x * NT)
        y[q] = y[q] + a * x[q];
}

// max |x| over n elements into *bits (the bit pattern of a non-negative double; NaN wins)
__global__ __launch_bounds__(NT) void k_maxabs(long long n, const double* __restrict__ x, unsigned long long* bits)
{
    double m = 0.0;
    for (long long q = (long long)blockIdx.x * NT + threadIdx.x; q < n; q += (long long)gridDim.x * NT) {
        const double v = fabs(x[q]);
        m = (v > m || __builtin_isnan(v)) ? v : m;
    }
    __shared__ double red[NT / 64];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = (o > m || __builtin_isnan(o)) ? o : m;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = red[0];
        for (int w = 1; w < NT / 64; ++w) v = (red[w] > v || __builtin_isnan(red[w])) ? red[w] : v;
        atomicMax(bits, (unsigned long long)__double_as_longlong(v));
    }
}

}  // namespace

// the composite operator on level l's valid cells: ap = A p (0 on covered cells), partials of p.ap;
// diag mode (ap_out = inverse diagonal): 1/sum of transmissibilities
template <int DIM>
__global__ __launch_bounds__(NT) void k_comp_ap(CompArgs a, int l, const double* __restrict__ sc, double* out,
                                                double* part, int diag)
{
    if (!diag && sc[CG_DONE] != 0.0) return;
    const CompLev& L = a.L[l];
    long long q_;
    const BoxDev ab = L.abox[cell_block(L.cmap, q_)];
    int i, j, k;
    double v = 0.0;
    if (ap_cell<DIM>(ab, q_, i, j, k)) {
        const long long c = ab.off + fab_index(ab, i, j, k);
        const int mask = L.mask[c];
        if (mask == COMP_COVERED) {
            out[c] = 0.0;
        } else {
            const double y = ap_cell_op<DIM>(a, l, ab, i, j, k, mask, diag != 0);
            if (diag) {
                out[c] = (y > 0.0) ? 1.0 / y : 0.0;
            } else {
                out[c] = y;
                v = L.v[c] * y;
            }
        }
    }
    if (!diag) ap_block_sum(v, part, (long long)blockIdx.y * gridDim.x + blockIdx.x);
}

// r = -(rhs - mean) V on the composite cells (0 covered), x = 0, and the partial sums of V rhs
// (first pass, mean_sum null) or of r.r (second pass)
template <int DIM>
__global__ __launch_bounds__(NT) void k_comp_init(CompArgs a, int l, const double* __restrict__ rhs,
                                                  const double* __restrict__ mean_sum, double vol_total, double* X,
                                                  double* R, double* part)
{
    const CompLev& L = a.L[l];
    long long q_;
    const BoxDev ab = L.abox[cell_block(L.cmap, q_)];
    int i, j, k;
    double v = 0.0;
    if (ap_cell<DIM>(ab, q_, i, j, k)) {
        const long long c = ab.off + fab_index(ab, i, j, k);
        const bool cov = L.mask[c] == COMP_COVERED;
        if (!mean_sum) {
            v = cov ? 0.0 : L.vol * rhs[c];
        } else {
            const double mean = mean_sum[0] / vol_total;
            const double r = cov ? 0.0 : -(rhs[c] - mean) * L.vol;
            X[c] = 0.0;
            R[c] = r;
            v = r * r;
        }
    }
    ap_block_sum(v, part, (long long)blockIdx.y * gridDim.x + blockIdx.x);
}

// S = R / V on every valid cell (the covered cells take the mean of their children afterwards)
template <int DIM>
__global__ __launch_bounds__(NT) void k_comp_s(CompArgs a, int l, const double* __restrict__ sc,
                                               const double* __restrict__ R, double* S)
{
    if (sc && sc[CG_DONE] != 0.0) return;
    const CompLev& L = a.L[l];
    long long q_;
    const BoxDev ab = L.abox[cell_block(L.cmap, q_)];
    int i, j, k;
    if (!ap_cell<DIM>(ab, q_, i, j, k)) return;
    const long long c = ab.off + fab_index(ab, i, j, k);
    S[c] = R[c] * (1.0 / L.vol);
}

// z = jacobi R ID (levels above 0) + the level-0 multigrid correction of the cell's level-0 ancestor
// (dense x0 over the level-0 domain, grid G0) on the composite cells; 0 on covered cells
template <int DIM>
__global__ __launch_bounds__(NT) void k_comp_z(CompArgs a, int l, const double* __restrict__ sc,
                                               const double* __restrict__ R, const double* __restrict__ ID,
                                               const double* __restrict__ x0, MgGrid G0, int dlo0, int dlo1,
                                               int dlo2, double* Z)
{
    if (sc && sc[CG_DONE] != 0.0) return;
    const CompLev& L = a.L[l];
    long long q_;
    const BoxDev ab = L.abox[cell_block(L.cmap, q_)];
    int i, j, k;
    if (!ap_cell<DIM>(ab, q_, i, j, k)) return;
    const long long c = ab.off + fab_index(ab, i, j, k);
    if (L.mask[c] == COMP_COVERED) { Z[c] = 0.0; return; }
    int f = 1;
    for (int m = 0; m < l; ++m) f *= a.ratio;
    const int a0 = ap_floor_div(i, f) - dlo0, a1 = ap_floor_div(j, f) - dlo1;
    const int a2 = (DIM == 3) ? ap_floor_div(k, f) - dlo2 : 0;
    // without a multigrid level 0 (x0 null) every level takes the Jacobi term alone
    const double z0 = x0 ? x0[(long long)a0 + (long long)G0.n[0] * (a1 + (long long)G0.n[1] * a2)] : 0.0;
    // every level between 0 and l: its averaged-down residual per volume over its own diagonal at the
    // cell's ancestor there (an additive multilevel term per level: symmetric, as the sum of P D^-1 P^T)
    double zm = 0.0;
    int fm = 1;
    for (int m = l - 1; m >= 1; --m) {
        fm *= a.ratio;
        const CompLev& M = a.L[m];
        int q[3] = {ap_floor_div(i, fm), ap_floor_div(j, fm), (DIM == 3) ? ap_floor_div(k, fm) : 0};
        const int own = ap_owner<DIM>(M.fd, q);
        if (own < 0) continue;
        const BoxDev mb = M.abox[own];
        const long long mc = mb.off + fab_index(mb, q[0], q[1], q[2]);
        zm += M.s[mc] * M.idl[mc];
    }
    Z[c] = ((l > 0 || !x0) ? R[c] * ID[c] : 0.0) + zm + z0;
}

// Grad_Q on every face of direction D of level l from the composite solution X (covered cells
// averaged down): (x_c - x_m)/dx between two cells of the level, the two-point difference through a
// coarse-fine face, 0 on Neumann domain faces
template <int DIM, int D>
__global__ __launch_bounds__(NT) void k_comp_grad(CompArgs a, int l, double* gq)
{
    const CompLev& L = a.L[l];
    const BoxDev ab = L.abox[blockIdx.y];
    const int n0 = ab.hi[0] - ab.lo[0] + 1 + (D == 0), n1 = ab.hi[1] - ab.lo[1] + 1 + (D == 1);
    const int n2 = (DIM == 3) ? ab.hi[2] - ab.lo[2] + 1 + (D == 2) : 1;
    const long long q = (long long)blockIdx.x * NT + threadIdx.x;
    if (q >= (long long)n0 * n1 * n2) return;
    int c[3];
    pos3(q, n0, n1, c[0], c[1], c[2]);
    c[0] += ab.lo[0];
    c[1] += ab.lo[1];
    c[2] += ab.lo[2];
    if (DIM == 2) c[2] = ab.lo[2];
    const long long fi = ab.goff[D] + q;
    if (!L.fd.periodic[D] && (c[D] == L.fd.dlo[D] || c[D] == L.fd.dhi[D] + 1)) { gq[fi] = 0.0; return; }
    int m[3] = {c[0], c[1], c[2]};
    m[D] -= 1;
    const bool c_in = c[D] <= ab.hi[D], m_in = m[D] >= ab.lo[D];
    double xc, xm;
    double h = L.dx[D];
    if (c_in && m_in) {
        xc = L.v[ab.off + fab_index(ab, c[0], c[1], c[2])];
        xm = L.v[ab.off + fab_index(ab, m[0], m[1], m[2])];
    } else {
        int o[3] = {c_in ? m[0] : c[0], c_in ? m[1] : c[1], c_in ? m[2] : c[2]};
        const double xi = c_in ? L.v[ab.off + fab_index(ab, c[0], c[1], c[2])]
                               : L.v[ab.off + fab_index(ab, m[0], m[1], m[2])];
        double xo;
        const int own = ap_owner<DIM>(L.fd, o);
        if (own >= 0) {
            const BoxDev ob = L.abox[own];
            xo = L.v[ob.off + fab_index(ob, o[0], o[1], o[2])];
        } else {
            // a coarse-fine face: the coarse cell beyond it
            const CompLev& C = a.L[l - 1];
            int p[3] = {0, 0, 0};
#pragma unroll
            for (int e = 0; e < DIM; ++e) p[e] = ap_floor_div(o[e], a.ratio);
            const int po = ap_owner<DIM>(C.fd, p);
            const BoxDev pb = C.abox[po < 0 ? 0 : po];
            xo = (po < 0) ? 0.0 : C.v[pb.off + fab_index(pb, p[0], p[1], p[2])];
            h = 0.5 * (L.dx[D] + C.dx[D]);
        }
        xc = c_in ? xi : xo;
        xm = c_in ? xo : xi;
    }
    gq[fi] = (xc - xm) / h;
}

// ---- launchers -------------------------------------------------------------------------------------

namespace {
dim3 box_grid2(long long maxn, int nbox)
{
    long long b = (maxn + NT - 1) / NT;
    if (b < 1) b = 1;
    return dim3((unsigned)b, (unsigned)nbox);
}
// a launch over the cells of level p.lev of the composite solve (its compact block list)
dim3 comp_grid(const AmrPfArgs& p)
{
    const CompLev& L = p.ca->L[p.lev];
    return cell_grid(L.cmap, L.ncmap, p.maxcells, p.nbox);
}
dim3 flat_grid(long long n)
{
    long long b = (n + NT - 1) / NT;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}
}  // namespace

template <int DIM>
hipError_t launch_amrpf_dim(int op, const AmrPfArgs& p, hipStream_t s)
{
    switch (op) {
    case AP_AVGDOWN_FACES:
        for (int d = 0; d < DIM; ++d)
            hipLaunchKernelGGL((k_avgdown_faces<DIM>), box_grid2(p.maxfaces[d], p.nbox), dim3(NT), 0, s, p.fboxes,
                               p.cfd, p.src, p.dst, p.ratio, d);
        break;
    case AP_AXPY:
        hipLaunchKernelGGL(k_axpy, flat_grid(p.n), dim3(NT), 0, s, p.n, p.a, p.src, p.dst);
        break;
    case AP_MAXABS:
        hipLaunchKernelGGL(k_maxabs, flat_grid(p.n), dim3(NT), 0, s, p.n, p.src, p.bits);
        break;
    case AP_COMP_AP:
        hipLaunchKernelGGL((k_comp_ap<DIM>), comp_grid(p), dim3(NT), 0, s, *p.ca, p.lev, p.sc,
                           p.dst, p.part, 0);
        break;
    case AP_COMP_DIAG:
        hipLaunchKernelGGL((k_comp_ap<DIM>), comp_grid(p), dim3(NT), 0, s, *p.ca, p.lev, p.sc,
                           p.dst, p.part, 1);
        break;
    case AP_COMP_INIT:
        hipLaunchKernelGGL((k_comp_init<DIM>), comp_grid(p), dim3(NT), 0, s, *p.ca, p.lev, p.src,
                           p.mean_sum, p.a, p.dst, p.dst2, p.part);
        break;
    case AP_COMP_S:
        hipLaunchKernelGGL((k_comp_s<DIM>), comp_grid(p), dim3(NT), 0, s, *p.ca, p.lev, p.sc,
                           p.src, p.dst);
        break;
    case AP_COMP_Z:
        hipLaunchKernelGGL((k_comp_z<DIM>), comp_grid(p), dim3(NT), 0, s, *p.ca, p.lev, p.sc,
                           p.src, p.src2, p.src3, p.G0, p.dlo0[0], p.dlo0[1], p.dlo0[2], p.dst);
        break;
    case AP_COMP_GRAD:
        hipLaunchKernelGGL((k_comp_grad<DIM, 0>), box_grid2(p.maxfaces[0], p.nbox), dim3(NT), 0, s, *p.ca, p.lev,
                           p.dst);
        hipLaunchKernelGGL((k_comp_grad<DIM, 1>), box_grid2(p.maxfaces[1], p.nbox), dim3(NT), 0, s, *p.ca, p.lev,
                           p.dst);
        if (DIM == 3)
            hipLaunchKernelGGL((k_comp_grad<DIM, (DIM == 3 ? 2 : 1)>), box_grid2(p.maxfaces[2], p.nbox), dim3(NT), 0,
                               s, *p.ca, p.lev, p.dst);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template hipError_t launch_amrpf_dim<CMP_KERNEL_DIM>(int, const AmrPfArgs&, hipStream_t);

}  // namespace cmp
