// cmp_pf.hip - gfx950 kernels of the Phase-Field step's mechanism (ACDI, CAC-Adv) and Grad_Q (explicit
// RC, or the implicit RC's variable-coefficient Poisson solve by conjugate gradients).
//
// Built twice, with CMP_KERNEL_DIM=2 and CMP_KERNEL_DIM=3. Every kernel runs over all boxes of a
// level at once: blockIdx.y is the box, blockIdx.x * 256 + threadIdx.x the cell or face inside it.
// The fields are one-component level buffers with one ghost cell (alpha1), DIM components with one
// ghost cell (cell normals, offsets DIM times alpha1's) and nodal face buffers (J, Jhat, Grad_Q).
// These are streaming passes over single fields (a few bytes of HBM traffic per cell per pass);
// the Phase-Field right-hand side itself runs in the fused stage kernel of cmp_kernels.hip.
#include "cmp_kernels.hpp"

#ifndef CMP_KERNEL_DIM
#error "CMP_KERNEL_DIM must be 2 or 3"
#endif

namespace cmp {

namespace {

// Valid-cell coordinates of linear index q of box bx (i fastest); false past the box's cells.
template <int DIM>
__device__ __forceinline__ bool cell_of(const BoxDev& bx, long long q, int& i, int& j, int& k)
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

// Face coordinates (of the face's high cell) of linear index q of the direction-D nodal faces.
template <int DIM, int D>
__device__ __forceinline__ bool face_of(const BoxDev& bx, long long q, int& i, int& j, int& k)
{
    const int n0 = bx.hi[0] - bx.lo[0] + 1 + (D == 0), n1 = bx.hi[1] - bx.lo[1] + 1 + (D == 1);
    const int n2 = (DIM == 3) ? bx.hi[2] - bx.lo[2] + 1 + (D == 2) : 1;
    if (q >= (long long)n0 * n1 * n2) return false;
    pos3(q, n0, n1, i, j, k);
    i += bx.lo[0];
    j += bx.lo[1];
    k += bx.lo[2];
    return true;
}

// Face index of the face at cell (i,j,k)'s low side (off = 0) or high side (off = 1) along D.
template <int DIM, int D>
__device__ __forceinline__ long long face_index(const BoxDev& bx, int i, int j, int k, int off)
{
    const int n0 = bx.hi[0] - bx.lo[0] + 1 + (D == 0), n1 = bx.hi[1] - bx.lo[1] + 1 + (D == 1);
    const int x = i - bx.lo[0] + (D == 0 ? off : 0);
    const int y = j - bx.lo[1] + (D == 1 ? off : 0);
    const int z = k - bx.lo[2] + (D == 2 ? off : 0);
    return bx.goff[D] + (long long)x + (long long)n0 * (y + (long long)n1 * z);
}

// Block max of non-negative v into *bits (the bit pattern of a non-negative double orders like the
// double). The atomic is issued only when the block's value exceeds what *bits already holds, so
// that a large grid does not queue one atomic per block on one address.
__device__ __forceinline__ void block_max_to(double v, unsigned long long* bits)
{
    __shared__ double red[NT / 64];
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
#pragma unroll
        for (int w = 1; w < NT / 64; ++w) m = smax(m, red[w]);
        const unsigned long long mb = (unsigned long long)__double_as_longlong(m);
        if (mb > __hip_atomic_load(bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(bits, mb);
    }
}

// Face index (box face buffer of direction d, runtime d) of cell (i,j,k)'s low (off 0) or high
// (off 1) face.
__device__ __forceinline__ long long face_index_rt(const BoxDev& bx, int d, int i, int j, int k, int off)
{
    const int n0 = bx.hi[0] - bx.lo[0] + 1 + (d == 0), n1 = bx.hi[1] - bx.lo[1] + 1 + (d == 1);
    const int x = i - bx.lo[0] + (d == 0 ? off : 0);
    const int y = j - bx.lo[1] + (d == 1 ? off : 0);
    const int z = k - bx.lo[2] + (d == 2 ? off : 0);
    return bx.goff[d] + (long long)x + (long long)n0 * (y + (long long)n1 * z);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block sum of v into part[idx], in a fixed order (the same inputs give the same bits). Every thread
// of the block must call it.
__device__ __forceinline__ void block_sum_to(double v, double* part, long long idx)
{
    __shared__ double red[NT / 64];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = red[0];
#pragma unroll
        for (int w = 1; w < NT / 64; ++w) s += red[w];
        part[idx] = s;
    }
    __syncthreads();
}

}  // namespace

// One volume fraction (state component comp) of the flow state's valid cells -> the alpha buffer
// (and into A2 when given: the mechanism's working copy)
template <int DIM>
__global__ __launch_bounds__(NT) void k_pf_extract(PfGeom g, const double* __restrict__ U, double* A, double* A2,
                                                   int comp)
{
    const BoxDev hb = g.hbox[blockIdx.y], ab = g.abox[blockIdx.y];
    int i, j, k;
    if (!cell_of<DIM>(ab, (long long)blockIdx.x * NT + threadIdx.x, i, j, k)) return;
    const long long a = ab.off + fab_index(ab, i, j, k);
    const double v = U[hb.off + fab_index(hb, i, j, k) + comp * hb.sn];
    A[a] = v;
    if (A2) A2[a] = v;
}

// One phase's Grad_Q (a one-component face buffer, abox offsets) -> component comp of the level's
// Grad_Q face buffer (hbox offsets, ncomp components per box and direction)
template <int DIM>
__global__ __launch_bounds__(NT) void k_pf_scatter(PfGeom g, const double* __restrict__ src, double* dst,
                                                   int comp)
{
    const BoxDev hb = g.hbox[blockIdx.y], ab = g.abox[blockIdx.y];
    const long long q = (long long)blockIdx.x * NT + threadIdx.x;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        long long nf = 1;
#pragma unroll
        for (int e = 0; e < DIM; ++e) nf *= ab.hi[e] - ab.lo[e] + 1 + (e == d ? 1 : 0);
        if (q < nf) dst[hb.goff[d] + comp * nf + q] = src[ab.goff[d] + q];
    }
}

// Pointwise maps of a whole alpha-layout buffer (valid and ghost cells): mode 0 psi =
// Mapping_Backward(alpha, eta) (the LS field of PhaseField_ACDI, whose ghosts the reference fills
// from the same cells the alpha ghosts come from); mode 1 the re-sharpened alpha of Grad_Q,
// Mapping_Forward(Mapping_Backward(alpha, eta), eta2).
#if CMP_KERNEL_DIM == 2
__global__ __launch_bounds__(NT) void k_pf_map(const double* __restrict__ A, double* out, long long n,
                                               double eta, double eta2, int mode)
{
    for (long long q = (long long)blockIdx.x * NT + threadIdx.x; q < n; q += (long long)gridDim.x * NT) {
        const double psi = pf_map_backward(A[q], eta);
        out[q] = mode ? pf_map_forward(psi, eta2) : psi;
    }
}

hipError_t launch_pf_map(const double* A, double* out, long long n, double eta, double eta2, int mode,
                         hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    long long blocks = (n + NT - 1) / NT;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_pf_map, dim3((unsigned)blocks), dim3(NT), 0, s, A, out, n, eta, eta2, mode);
    return hipGetLastError();
}
#endif

// Cell normals of psi on valid cells (PhaseField_Normal with FDM_Normal_CenterFromCenter_D,
// include/Operators_FD/FDM_Normal_CenterFromCenter_D.H:12-63).
template <int DIM>
__global__ __launch_bounds__(NT) void k_pf_normals(PfGeom g, const double* __restrict__ Psi,
                                                   double* Nrm)
{
    const BoxDev ab = g.abox[blockIdx.y];
    int i, j, k;
    if (!cell_of<DIM>(ab, (long long)blockIdx.x * NT + threadIdx.x, i, j, k)) return;
    const double* Pb = Psi + ab.off;
    const long long c = fab_index(ab, i, j, k);
    const double fc = Pb[c];
    const long long st[3] = {1, ab.sj, ab.sk};
    double gr[DIM], sum = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const double fm = Pb[c - st[d]];
        const double fp = Pb[c + st[d]];
        gr[d] = 0.5 * ((fc - fm) / g.dx[d]) + 0.5 * ((fp - fc) / g.dx[d]);
        sum += gr[d] * gr[d];
    }
    double* Nb = Nrm + (long long)DIM * ab.off;
    const double r = (sum == 0.0) ? 0.0 : sqrt(sum);
#pragma unroll
    for (int d = 0; d < DIM; ++d) Nb[c + d * ab.sn] = (sum == 0.0) ? 0.0 : gr[d] / r;
}

// ACDI flux on every face of direction D: J stored, Jhat += dt_m J (explicit RC)
// (PhaseField_ACDI, source/Class_Compressible_PhaseField.cpp:1609-1640; Saxpy of :679-683).
template <int DIM, int D>
__global__ __launch_bounds__(NT) void k_pf_jfaces(PfGeom g, const double* __restrict__ A,
                                                  const double* __restrict__ Psi,
                                                  const double* __restrict__ Nrm, double* J, double* H,
                                                  double mob_face, double eta, double dt_m)
{
    const BoxDev ab = g.abox[blockIdx.y];
    const long long q = (long long)blockIdx.x * NT + threadIdx.x;
    int i, j, k;
    if (!face_of<DIM, D>(ab, q, i, j, k)) return;
    const long long c = fab_index(ab, i, j, k);
    const long long m = c - ((D == 0) ? 1 : ((D == 1) ? ab.sj : ab.sk));
    const double* Ab = A + ab.off;
    const double* Pb = Psi + ab.off;
    const double* Nb = Nrm + (long long)DIM * ab.off + D * ab.sn;
    const double Jf = pf_acdi_flux(Ab[m], Ab[c], Pb[m], Pb[c], Nb[m], Nb[c], mob_face, eta, g.dx[D]);
    const long long f = ab.goff[D] + q;
    J[f] = Jf;
    if (H) H[f] = H[f] + dt_m * Jf;   // H is null when Jhat is not needed (implicit RC)
}

// alpha1 += dt_m div J on valid cells (FVM_Divergence_K + LinComb, :685-688).
template <int DIM>
__global__ __launch_bounds__(NT) void k_pf_update(PfGeom g, double* A, const double* __restrict__ J,
                                                  double dt_m)
{
    const BoxDev ab = g.abox[blockIdx.y];
    int i, j, k;
    if (!cell_of<DIM>(ab, (long long)blockIdx.x * NT + threadIdx.x, i, j, k)) return;
    double div = (J[face_index<DIM, 0>(ab, i, j, k, 1)] - J[face_index<DIM, 0>(ab, i, j, k, 0)]) / g.dx[0];
    div = div + (J[face_index<DIM, 1>(ab, i, j, k, 1)] - J[face_index<DIM, 1>(ab, i, j, k, 0)]) / g.dx[1];
    if (DIM == 3)
        div = div + (J[face_index<DIM, (DIM == 3 ? 2 : 1)>(ab, i, j, k, 1)]
                     - J[face_index<DIM, (DIM == 3 ? 2 : 1)>(ab, i, j, k, 0)]) / g.dx[2];
    const long long c = ab.off + fab_index(ab, i, j, k);
    A[c] = 1.0 * A[c] + dt_m * div;
}

// Explicit Grad_Q on every face of direction D from the step's re-sharpened alpha1 S (k_pf_map
// mode 1) and Jhat, with max |Grad_Q|.
template <int DIM, int D>
__global__ __launch_bounds__(NT) void k_pf_gradq(PfGeom g, const double* __restrict__ S,
                                                 const double* __restrict__ H, double hscale, double* gq,
                                                 unsigned long long* gmax_bits, int* boxflag)
{
    const BoxDev ab = g.abox[blockIdx.y];
    const long long q = (long long)blockIdx.x * NT + threadIdx.x;
    int i, j, k;
    double v = 0.0;
    if (face_of<DIM, D>(ab, q, i, j, k)) {
        const long long c = fab_index(ab, i, j, k);
        const long long m = c - ((D == 0) ? 1 : ((D == 1) ? ab.sj : ab.sk));
        const double* Sb = S + ab.off;
        const long long f = ab.goff[D] + q;
        const double r = pf_gradq_sharp(Sb[m], Sb[c], H[f] * hscale);
        gq[f] = r;
        v = fabs(r);
        if (__builtin_isnan(r)) v = __longlong_as_double(0x7ff8000000000000LL);
        if (boxflag && r != 0.0) boxflag[blockIdx.y] = 1;   // every writer stores the same value
    }
    block_max_to(v, gmax_bits);
}

// max over valid cells of |u| (MaxVelocity, :3541-3579; Conservative2Velocity_Compressible5Eq.H:15-40)
template <int DIM>
// nmass partial densities come first (2, or the N-phase model's NP), then the momenta
__global__ __launch_bounds__(NT) void k_pf_maxvel(PfGeom g, const double* __restrict__ U,
                                                  unsigned long long* bits, int nmass)
{
    const BoxDev hb = g.hbox[blockIdx.y];
    int i, j, k;
    double v = 0.0;
    if (cell_of<DIM>(hb, (long long)blockIdx.x * NT + threadIdx.x, i, j, k)) {
        const double* Ub = U + hb.off;
        const long long c = fab_index(hb, i, j, k);
        double rho = Ub[c];
        for (int q = 1; q < nmass; ++q) rho += Ub[c + q * hb.sn];
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const double u = Ub[c + (nmass + d) * hb.sn] / rho;
            s = (d == 0) ? u * u : s + u * u;
        }
        v = sqrt(s);
        if (__builtin_isnan(v)) v = __longlong_as_double(0x7ff8000000000000LL);
    }
    block_max_to(v, bits);
}

static dim3 grid_of_cells(long long per_box, int nbox)
{
    return dim3((unsigned)((per_box + NT - 1) / NT), (unsigned)nbox);
}

// ---- CAC mechanism (PhaseField_Mechanism, source/Class_Compressible_PhaseField.cpp:1007-1152) -----

// mech and W = alpha (1 - alpha) of every valid cell, with block partial sums of both
template <int DIM>
__global__ __launch_bounds__(NT) void k_pf_cac_mech(PfGeom g, const double* __restrict__ A,
                                                    const double* __restrict__ Nrm, double* Mk, double* Wk,
                                                    double mob, double lambda, double eta, double* part)
{
    long long q_;
    const BoxDev ab = g.abox[cell_block(g.cmap, q_)];
    int i, j, k;
    double mk = 0.0, wk = 0.0;
    if (cell_of<DIM>(ab, q_, i, j, k)) {
        const double* Ab = A + ab.off;
        const double* Nb = Nrm + (long long)DIM * ab.off;
        const long long c = fab_index(ab, i, j, k);
        const long long st[3] = {1, ab.sj, ab.sk};
        double fm[DIM], fp[DIM], nc[DIM], nm[DIM], np[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            fm[d] = Ab[c - st[d]];
            fp[d] = Ab[c + st[d]];
            nc[d] = Nb[c + d * ab.sn];
            nm[d] = Nb[c - st[d] + d * ab.sn];
            np[d] = Nb[c + st[d] + d * ab.sn];
        }
        const double f = Ab[c];
        mk = pf_cac_mech<DIM>(g.dx, f, fm, fp, nc, nm, np, mob, lambda, eta);
        wk = f * (1.0 - f);
        Mk[ab.off + c] = mk;
        Wk[ab.off + c] = wk;
    }
    const long long blk = (long long)blockIdx.y * gridDim.x + blockIdx.x;
    const long long nblk = (long long)gridDim.x * gridDim.y;
    block_sum_to(mk, part, blk);
    block_sum_to(wk, part, nblk + blk);
}

// alpha += dt_m (mech + W L), L = -sum mech / sum W (0 when sum W = 0, PhaseFieldMechanism_CAC.H)
template <int DIM>
__global__ __launch_bounds__(NT) void k_pf_cac_update(PfGeom g, double* A, const double* __restrict__ Mk,
                                                      const double* __restrict__ Wk,
                                                      const double* __restrict__ sums, double dt_m)
{
    long long q_;
    const BoxDev ab = g.abox[cell_block(g.cmap, q_)];
    int i, j, k;
    if (!cell_of<DIM>(ab, q_, i, j, k)) return;
    const double sw = sums[1];
    const double L = (sw == 0.0) ? 0.0 : -sums[0] / sw;
    const long long c = ab.off + fab_index(ab, i, j, k);
    const double divj = Mk[c] + Wk[c] * L;
    A[c] = 1.0 * A[c] + dt_m * divj;
}

// Div_Jhat = (1/dt) alpha_new + (-1/dt) alpha (LinComb of PhaseFieldMechanism_Evolve, :703)
template <int DIM>
__global__ __launch_bounds__(NT) void k_pf_divjhat(PfGeom g, const double* __restrict__ A,
                                                   const double* __restrict__ A0, double* DJ, double a,
                                                   double b)
{
    long long q_;
    const BoxDev ab = g.abox[cell_block(g.cmap, q_)];
    int i, j, k;
    if (!cell_of<DIM>(ab, q_, i, j, k)) return;
    const long long c = ab.off + fab_index(ab, i, j, k);
    DJ[c] = a * A[c] + b * A0[c];
}

// ---- implicit Grad_Q (PhaseField_GradQPF with Div_J, :2133-2310) ---------------------------------

// b = W(phi_f) + 1e-3 on every face of direction D, phi_f the face mean of the re-sharpened alpha
template <int DIM, int D>
__global__ __launch_bounds__(NT) void k_pf_bcoef(PfGeom g, const double* __restrict__ S, double* Bf)
{
    const BoxDev ab = g.abox[blockIdx.y];
    const long long q = (long long)blockIdx.x * NT + threadIdx.x;
    int i, j, k;
    if (!face_of<DIM, D>(ab, q, i, j, k)) return;
    const long long c = fab_index(ab, i, j, k);
    const long long m = c - ((D == 0) ? 1 : ((D == 1) ? ab.sj : ab.sk));
    const double* Sb = S + ab.off;
    const double pf = 0.5 * Sb[m] + 0.5 * Sb[c];
    Bf[ab.goff[D] + q] = pf * (1.0 - pf) + 1.0e-3;
}

// 1/diag of -div(b grad .): sum over the cell's faces of b_f/dx_d^2, Neumann domain faces excluded
template <int DIM>
__global__ __launch_bounds__(NT) void k_cg_idiag(PfGeom g, const double* __restrict__ Bf, double* ID)
{
    long long q_;
    const BoxDev ab = g.abox[cell_block(g.cmap, q_)];
    int i, j, k;
    if (!cell_of<DIM>(ab, q_, i, j, k)) return;
    const int co[3] = {i, j, k};
    double dsum = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const bool lo_dom = !g.periodic[d] && co[d] == g.dlo[d];
        const bool hi_dom = !g.periodic[d] && co[d] == g.dhi[d];
        const double h2 = g.dx[d] * g.dx[d];
        if (!hi_dom) dsum += Bf[face_index_rt(ab, d, i, j, k, 1)] / h2;
        if (!lo_dom) dsum += Bf[face_index_rt(ab, d, i, j, k, 0)] / h2;
    }
    ID[ab.off + fab_index(ab, i, j, k)] = (dsum > 0.0) ? 1.0 / dsum : 0.0;
}

template <int DIM>
__global__ __launch_bounds__(NT) void k_cg_init(PfGeom g, const double* __restrict__ rhs,
                                                const double* __restrict__ mean_sum, double ncell,
                                                const double* __restrict__ ID, double* X, double* R,
                                                double* Z, double* P, double* part)
{
    long long q_;
    const BoxDev ab = g.abox[cell_block(g.cmap, q_)];
    int i, j, k;
    double v = 0.0, w = 0.0;
    if (cell_of<DIM>(ab, q_, i, j, k)) {
        const long long c = ab.off + fab_index(ab, i, j, k);
        const double mean = mean_sum[0] / ncell;
        const double r = -(rhs[c] - mean);
        const double z = r * ID[c];
        X[c] = 0.0;
        R[c] = r;
        Z[c] = z;
        P[c] = z;
        v = r * r;
        w = r * z;
    }
    const long long blk = (long long)blockIdx.y * gridDim.x + blockIdx.x;
    block_sum_to(v, part, blk);
    block_sum_to(w, part, (long long)gridDim.x * gridDim.y + blk);
}

// the level sum of rhs over valid cells (for its mean)
template <int DIM>
__global__ __launch_bounds__(NT) void k_cell_sum(PfGeom g, const double* __restrict__ F, double* part)
{
    long long q_;
    const BoxDev ab = g.abox[cell_block(g.cmap, q_)];
    int i, j, k;
    double v = 0.0;
    if (cell_of<DIM>(ab, q_, i, j, k)) v = F[ab.off + fab_index(ab, i, j, k)];
    block_sum_to(v, part, (long long)blockIdx.y * gridDim.x + blockIdx.x);
}

template <int DIM>
__global__ __launch_bounds__(NT) void k_cg_ap(PfGeom g, const double* __restrict__ sc,
                                              const double* __restrict__ Bf, const double* __restrict__ P,
                                              double* AP, double* part)
{
    if (sc[CG_DONE] != 0.0) return;   // uniform over the grid
    long long q_;
    const BoxDev ab = g.abox[cell_block(g.cmap, q_)];
    int i, j, k;
    double v = 0.0;
    if (cell_of<DIM>(ab, q_, i, j, k)) {
        const double* Pb = P + ab.off;
        const long long c = fab_index(ab, i, j, k);
        const long long st[3] = {1, ab.sj, ab.sk};
        const int co[3] = {i, j, k};
        const double pc = Pb[c];
        double acc = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const bool lo_dom = !g.periodic[d] && co[d] == g.dlo[d];
            const bool hi_dom = !g.periodic[d] && co[d] == g.dhi[d];
            const double fp_ = hi_dom ? 0.0 : Bf[face_index_rt(ab, d, i, j, k, 1)] * (Pb[c + st[d]] - pc) / g.dx[d];
            const double fm_ = lo_dom ? 0.0 : Bf[face_index_rt(ab, d, i, j, k, 0)] * (pc - Pb[c - st[d]]) / g.dx[d];
            const double t = (fp_ - fm_) / g.dx[d];
            acc = (d == 0) ? t : acc + t;
        }
        AP[ab.off + c] = -acc;
        v = pc * (-acc);
    }
    block_sum_to(v, part, (long long)blockIdx.y * gridDim.x + blockIdx.x);
}

template <int DIM>
__global__ __launch_bounds__(NT) void k_cg_xr(PfGeom g, const double* __restrict__ sc,
                                              const double* __restrict__ ID, double* X, double* R, double* Z,
                                              const double* __restrict__ P, const double* __restrict__ AP,
                                              double* part)
{
    if (sc[CG_DONE] != 0.0 || !(sc[CG_PAP] > 0.0)) return;
    long long q_;
    const BoxDev ab = g.abox[cell_block(g.cmap, q_)];
    int i, j, k;
    double v = 0.0, w = 0.0;
    if (cell_of<DIM>(ab, q_, i, j, k)) {
        const long long c = ab.off + fab_index(ab, i, j, k);
        const double al = sc[CG_RZ] / sc[CG_PAP];
        X[c] += al * P[c];
        const double r = R[c] - al * AP[c];
        R[c] = r;
        v = r * r;
        // ID null: another preconditioner computes z and r.z after this kernel
        if (ID) {
            const double z = r * ID[c];
            Z[c] = z;
            w = r * z;
        }
    }
    const long long blk = (long long)blockIdx.y * gridDim.x + blockIdx.x;
    block_sum_to(v, part, blk);
    block_sum_to(w, part, (long long)gridDim.x * gridDim.y + blk);
}

// With inv_n > 0 the search directions stay orthogonal to the constant vector, the null space of a
// Neumann / periodic operator: a multigrid preconditioner's coarsest grids leave the constant part of
// a residual's rounding unreduced and return it amplified, and once the residual is at the rounding
// floor that part would dominate p, p.Ap would fall to rounding and the step alpha = r.z / p.Ap
// would throw the iterate off.
template <int DIM>
__global__ __launch_bounds__(NT) void k_cg_p(PfGeom g, const double* __restrict__ sc,
                                             const double* __restrict__ Z, double* P, double inv_n,
                                             const int* __restrict__ mask, int first)
{
    if (sc[CG_DONE] != 0.0 || (!first && !(sc[CG_PAP] > 0.0))) return;
    long long q_;
    const BoxDev ab = g.abox[cell_block(g.cmap, q_)];
    int i, j, k;
    if (!cell_of<DIM>(ab, q_, i, j, k)) return;
    const long long c = ab.off + fab_index(ab, i, j, k);
    const bool covered = mask && mask[c] == COMP_COVERED;
    const double zc = (inv_n > 0.0 && !covered) ? Z[c] - sc[CG_ZS] * inv_n : Z[c];
    if (first) {
        P[c] = zc;
    } else {
        const double be = cg_rz_projected(sc, CG_RZNEW, inv_n) / sc[CG_RZ];
        P[c] = zc + be * P[c];
    }
}

// Grad_Q = (x_c - x_m)/dx on every face of direction D; zero on Neumann domain faces
template <int DIM, int D>
__global__ __launch_bounds__(NT) void k_pf_gradx(PfGeom g, const double* __restrict__ X, double* gq,
                                                 unsigned long long* gmax_bits, int* boxflag)
{
    const BoxDev ab = g.abox[blockIdx.y];
    const long long q = (long long)blockIdx.x * NT + threadIdx.x;
    int i, j, k;
    double v = 0.0;
    if (face_of<DIM, D>(ab, q, i, j, k)) {
        const int co = (D == 0) ? i : ((D == 1) ? j : k);
        const bool on_dom = !g.periodic[D] && (co == g.dlo[D] || co == g.dhi[D] + 1);
        const long long c = fab_index(ab, i, j, k);
        const long long m = c - ((D == 0) ? 1 : ((D == 1) ? ab.sj : ab.sk));
        const double* Xb = X + ab.off;
        const double r = on_dom ? 0.0 : (Xb[c] - Xb[m]) / g.dx[D];
        gq[ab.goff[D] + q] = r;
        v = fabs(r);
        if (__builtin_isnan(r)) v = __longlong_as_double(0x7ff8000000000000LL);
        if (boxflag && r != 0.0) boxflag[blockIdx.y] = 1;
    }
    block_max_to(v, gmax_bits);
}

// ---- launchers -----------------------------------------------------------------------------------

template <int DIM>
hipError_t launch_pf_cac_mech_dim(const PfGeom& g, long long maxcells, const double* A, const double* Nrm,
                                  double* Mk, double* Wk, double mob, double lambda, double eta,
                                  double* part, hipStream_t s)
{
    hipLaunchKernelGGL((k_pf_cac_mech<DIM>), cell_grid(g.cmap, g.ncmap, maxcells, g.nbox), dim3(NT), 0, s, g, A, Nrm, Mk,
                       Wk, mob, lambda, eta, part);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_pf_cac_update_dim(const PfGeom& g, long long maxcells, double* A, const double* Mk,
                                    const double* Wk, const double* sums, double dt_m, hipStream_t s)
{
    hipLaunchKernelGGL((k_pf_cac_update<DIM>), cell_grid(g.cmap, g.ncmap, maxcells, g.nbox), dim3(NT), 0, s, g, A, Mk, Wk,
                       sums, dt_m);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_pf_divjhat_dim(const PfGeom& g, long long maxcells, const double* A, const double* A0,
                                 double* DJ, double dt, hipStream_t s)
{
    hipLaunchKernelGGL((k_pf_divjhat<DIM>), cell_grid(g.cmap, g.ncmap, maxcells, g.nbox), dim3(NT), 0, s, g, A, A0, DJ,
                       1.0 / dt, -1.0 / dt);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_pf_bcoef_dim(const PfGeom& g, const long long* maxfaces, const double* S, double* Bf,
                               hipStream_t s)
{
    hipLaunchKernelGGL((k_pf_bcoef<DIM, 0>), grid_of_cells(maxfaces[0], g.nbox), dim3(NT), 0, s, g, S, Bf);
    hipLaunchKernelGGL((k_pf_bcoef<DIM, 1>), grid_of_cells(maxfaces[1], g.nbox), dim3(NT), 0, s, g, S, Bf);
    if (DIM == 3)
        hipLaunchKernelGGL((k_pf_bcoef<DIM, (DIM == 3 ? 2 : 1)>), grid_of_cells(maxfaces[2], g.nbox), dim3(NT),
                           0, s, g, S, Bf);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_cell_sum_dim(const PfGeom& g, long long maxcells, const double* F, double* part, hipStream_t s)
{
    hipLaunchKernelGGL((k_cell_sum<DIM>), cell_grid(g.cmap, g.ncmap, maxcells, g.nbox), dim3(NT), 0, s, g, F, part);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_cg_idiag_dim(const PfGeom& g, long long maxcells, const double* Bf, double* ID, hipStream_t s)
{
    hipLaunchKernelGGL((k_cg_idiag<DIM>), cell_grid(g.cmap, g.ncmap, maxcells, g.nbox), dim3(NT), 0, s, g, Bf, ID);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_cg_init_dim(const PfGeom& g, long long maxcells, const double* rhs, const double* mean_sum,
                              double ncell, const double* ID, double* X, double* R, double* Z, double* P,
                              double* part, hipStream_t s)
{
    hipLaunchKernelGGL((k_cg_init<DIM>), cell_grid(g.cmap, g.ncmap, maxcells, g.nbox), dim3(NT), 0, s, g, rhs, mean_sum,
                       ncell, ID, X, R, Z, P, part);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_cg_ap_dim(const PfGeom& g, long long maxcells, const double* sc, const double* Bf,
                            const double* P, double* AP, double* part, hipStream_t s)
{
    hipLaunchKernelGGL((k_cg_ap<DIM>), cell_grid(g.cmap, g.ncmap, maxcells, g.nbox), dim3(NT), 0, s, g, sc, Bf, P, AP, part);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_cg_xr_dim(const PfGeom& g, long long maxcells, const double* sc, const double* ID, double* X,
                            double* R, double* Z, const double* P, const double* AP, double* part, hipStream_t s)
{
    hipLaunchKernelGGL((k_cg_xr<DIM>), cell_grid(g.cmap, g.ncmap, maxcells, g.nbox), dim3(NT), 0, s, g, sc, ID, X, R, Z, P, AP,
                       part);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_cg_p_dim(const PfGeom& g, long long maxcells, const double* sc, const double* Z, double* P,
                           double inv_n, const int* mask, int first, hipStream_t s)
{
    hipLaunchKernelGGL((k_cg_p<DIM>), cell_grid(g.cmap, g.ncmap, maxcells, g.nbox), dim3(NT), 0, s, g, sc, Z, P, inv_n, mask,
                       first);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_pf_gradx_dim(const PfGeom& g, const long long* maxfaces, const double* X, double* gq,
                               unsigned long long* bits, int* boxflag, hipStream_t s)
{
    hipLaunchKernelGGL((k_pf_gradx<DIM, 0>), grid_of_cells(maxfaces[0], g.nbox), dim3(NT), 0, s, g, X, gq, bits,
                       boxflag);
    hipLaunchKernelGGL((k_pf_gradx<DIM, 1>), grid_of_cells(maxfaces[1], g.nbox), dim3(NT), 0, s, g, X, gq, bits,
                       boxflag);
    if (DIM == 3)
        hipLaunchKernelGGL((k_pf_gradx<DIM, (DIM == 3 ? 2 : 1)>), grid_of_cells(maxfaces[2], g.nbox), dim3(NT), 0,
                           s, g, X, gq, bits, boxflag);
    return hipGetLastError();
}

static dim3 grid_of(long long per_box, int nbox)
{
    return dim3((unsigned)((per_box + NT - 1) / NT), (unsigned)nbox);
}

template <int DIM>
hipError_t launch_pf_extract_dim(const PfGeom& g, long long maxcells, const double* U, double* A, double* A2,
                                 int comp, hipStream_t s)
{
    hipLaunchKernelGGL((k_pf_extract<DIM>), grid_of(maxcells, g.nbox), dim3(NT), 0, s, g, U, A, A2, comp);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_pf_scatter_dim(const PfGeom& g, const long long* maxfaces, const double* src, double* dst,
                                 int comp, hipStream_t s)
{
    long long m = 0;
    for (int d = 0; d < DIM; ++d) m = (maxfaces[d] > m) ? maxfaces[d] : m;
    hipLaunchKernelGGL((k_pf_scatter<DIM>), grid_of(m, g.nbox), dim3(NT), 0, s, g, src, dst, comp);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_pf_normals_dim(const PfGeom& g, long long maxcells, const double* Psi, double* Nrm,
                                 hipStream_t s)
{
    hipLaunchKernelGGL((k_pf_normals<DIM>), grid_of(maxcells, g.nbox), dim3(NT), 0, s, g, Psi, Nrm);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_pf_jfaces_dim(const PfGeom& g, const long long* maxfaces, const double* A,
                                const double* Psi, const double* Nrm, double* J, double* H,
                                double mob_face, double eta, double dt_m, hipStream_t s)
{
    hipLaunchKernelGGL((k_pf_jfaces<DIM, 0>), grid_of(maxfaces[0], g.nbox), dim3(NT), 0, s, g, A, Psi,
                       Nrm, J, H, mob_face, eta, dt_m);
    hipLaunchKernelGGL((k_pf_jfaces<DIM, 1>), grid_of(maxfaces[1], g.nbox), dim3(NT), 0, s, g, A, Psi,
                       Nrm, J, H, mob_face, eta, dt_m);
    if (DIM == 3)
        hipLaunchKernelGGL((k_pf_jfaces<DIM, (DIM == 3 ? 2 : 1)>), grid_of(maxfaces[2], g.nbox), dim3(NT),
                           0, s, g, A, Psi, Nrm, J, H, mob_face, eta, dt_m);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_pf_update_dim(const PfGeom& g, long long maxcells, double* A, const double* J,
                                double dt_m, hipStream_t s)
{
    hipLaunchKernelGGL((k_pf_update<DIM>), grid_of(maxcells, g.nbox), dim3(NT), 0, s, g, A, J, dt_m);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_pf_gradq_dim(const PfGeom& g, const long long* maxfaces, const double* S,
                               const double* H, double hscale, double* gq, unsigned long long* gmax_bits,
                               int* boxflag, hipStream_t s)
{
    hipLaunchKernelGGL((k_pf_gradq<DIM, 0>), grid_of(maxfaces[0], g.nbox), dim3(NT), 0, s, g, S, H, hscale, gq,
                       gmax_bits, boxflag);
    hipLaunchKernelGGL((k_pf_gradq<DIM, 1>), grid_of(maxfaces[1], g.nbox), dim3(NT), 0, s, g, S, H, hscale, gq,
                       gmax_bits, boxflag);
    if (DIM == 3)
        hipLaunchKernelGGL((k_pf_gradq<DIM, (DIM == 3 ? 2 : 1)>), grid_of(maxfaces[2], g.nbox), dim3(NT),
                           0, s, g, S, H, hscale, gq, gmax_bits, boxflag);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_pf_maxvel_dim(const PfGeom& g, long long maxcells, const double* U,
                                unsigned long long* bits, int nmass, hipStream_t s)
{
    hipLaunchKernelGGL((k_pf_maxvel<DIM>), grid_of(maxcells, g.nbox), dim3(NT), 0, s, g, U, bits, nmass);
    return hipGetLastError();
}

#if CMP_KERNEL_DIM == 2
__global__ __launch_bounds__(NT) void k_scale(double* p, long long n, double v)
{
    for (long long q = (long long)blockIdx.x * NT + threadIdx.x; q < n; q += (long long)gridDim.x * NT)
        p[q] *= v;
}

__global__ void k_setd(double* p, double v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ void k_zero_u64(unsigned long long* p, int n)
{
    if (blockIdx.x == 0 && (int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

hipError_t launch_scale(double* p, long long n, double v, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    long long blocks = (n + NT - 1) / NT;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_scale, dim3((unsigned)blocks), dim3(NT), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_setd(double* p, double v, hipStream_t s)
{
    hipLaunchKernelGGL(k_setd, dim3(1), dim3(64), 0, s, p, v);
    return hipGetLastError();
}

hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t s)
{
    if (n <= 0 || n > 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, p, n);
    return hipGetLastError();
}

// sums[v] = sum_b part[v * nblocks + b], one workgroup of 1024 threads, a fixed order of additions
// (each thread four interleaved running sums, then the wave and workgroup sums in index order).
// sc non-null: nothing once the solve's stop flag is set (the queued iterations past the stop).
constexpr int NRED = 1024;
// The fixed-order sum of nblocks partials p[0..nblocks) by the NRED threads of one workgroup: thread
// t adds p[t], p[t + NRED], ... in four interleaved accumulators, then the waves and the workgroup
// add their values in a fixed tree; the result is returned to thread 0.
__device__ __forceinline__ double reduce_block_parts(const double* __restrict__ p, int nblocks, double* red)
{
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int b = threadIdx.x;
    for (; b + 3 * NRED < nblocks; b += 4 * NRED) {
        a0 += p[b];
        a1 += p[b + NRED];
        a2 += p[b + 2 * NRED];
        a3 += p[b + 3 * NRED];
    }
    for (; b < nblocks; b += NRED) a0 += p[b];
    double t = (a0 + a1) + (a2 + a3);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) t += __shfl_xor(t, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
    __syncthreads();
    double u = 0.0;
    if (threadIdx.x == 0) {
        u = red[0];
#pragma unroll
        for (int w = 1; w < NRED / 64; ++w) u += red[w];
    }
    __syncthreads();
    return u;
}

__global__ __launch_bounds__(NRED) void k_reduce_sums(const double* __restrict__ part, int nblocks, int nvals,
                                                      double* sums, const double* __restrict__ sc)
{
    if (sc && sc[CG_DONE] != 0.0) return;
    __shared__ double red[NRED / 64];
    for (int v = 0; v < nvals; ++v) {
        const double u = reduce_block_parts(part + (long long)v * nblocks, nblocks, red);
        if (threadIdx.x == 0) sums[v] = u;
    }
}

// every level's sums (per[l * nvals + v], as k_reduce_sums leaves them) and their sum in level order
__global__ __launch_bounds__(NRED) void k_reduce_levels(LevelParts lp, int nvals, double* per, double* sums,
                                                        const double* __restrict__ sc)
{
    if (sc && sc[CG_DONE] != 0.0) return;
    __shared__ double red[NRED / 64];
    for (int l = 0; l < lp.nlev; ++l)
        for (int v = 0; v < nvals; ++v) {
            const double u = reduce_block_parts(lp.part[l] + (long long)v * lp.nblocks[l], lp.nblocks[l], red);
            if (threadIdx.x == 0) per[l * nvals + v] = u;
        }
    if (threadIdx.x == 0)
        for (int v = 0; v < nvals; ++v) {
            double t = per[v];
            for (int l = 1; l < lp.nlev; ++l) t += per[l * nvals + v];
            sums[v] = t;
        }
}

// CG bookkeeping on the device, so that the iteration stops where the host loop of the solve's
// restatement stops without a host round trip per iteration: start after r.r of the initial
// residual landed in sc[CG_RR]; end after each iteration (pAp <= 0 stops without counting it;
// r.r <= 1e-28 b.b, 400 iterations without a new minimum of r.r (CG's residual norm is not monotone
// and can plateau for tens of iterations before it falls again) or maxit stop after it).
__global__ void k_cg_start(double* sc, double inv_n)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    sc[CG_RZ] = cg_rz_projected(sc, CG_RZ, inv_n);
    const double rr = sc[CG_RR];
    sc[CG_B2] = rr;
    sc[CG_BEST] = rr;
    sc[CG_STALL] = 0.0;
    sc[CG_IT] = 0.0;
    sc[CG_DONE] = (rr > 0.0 && sc[CG_RZ] > 0.0) ? 0.0 : 1.0;   // a zero (or NaN) right-hand side: x = 0
}

__global__ void k_cg_end(double* sc, double maxit, double inv_n)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (sc[CG_DONE] != 0.0) return;
    if (!(sc[CG_PAP] > 0.0)) {
        sc[CG_DONE] = 1.0;
        return;
    }
    const double rr = sc[CG_RRNEW];
    sc[CG_RR] = rr;
    sc[CG_RZ] = cg_rz_projected(sc, CG_RZNEW, inv_n);
    const double it = sc[CG_IT] + 1.0;
    sc[CG_IT] = it;
    bool done = false;
    if (rr < sc[CG_BEST]) {
        sc[CG_BEST] = rr;
        sc[CG_STALL] = 0.0;
    } else {
        const double st = sc[CG_STALL] + 1.0;
        sc[CG_STALL] = st;
        done = st > 400.0;
    }
    if (!(rr > 1.0e-28 * sc[CG_B2]) || !(it < maxit)) done = true;
    if (done) sc[CG_DONE] = 1.0;
}

hipError_t launch_reduce_levels(const LevelParts& lp, int nvals, double* per, double* sums, hipStream_t s,
                                const double* sc)
{
    if (lp.nlev < 1 || lp.nlev > AMR_MAX_LEVELS) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_reduce_levels, dim3(1), dim3(NRED), 0, s, lp, nvals, per, sums, sc);
    return hipGetLastError();
}

hipError_t launch_reduce_sums(const double* part, int nblocks, int nvals, double* sums, hipStream_t s,
                              const double* sc)
{
    hipLaunchKernelGGL(k_reduce_sums, dim3(1), dim3(NRED), 0, s, part, nblocks, nvals, sums, sc);
    return hipGetLastError();
}

hipError_t launch_cg_start(double* sc, hipStream_t s, double inv_n)
{
    hipLaunchKernelGGL(k_cg_start, dim3(1), dim3(64), 0, s, sc, inv_n);
    return hipGetLastError();
}

hipError_t launch_cg_end(double* sc, int maxit, hipStream_t s, double inv_n)
{
    hipLaunchKernelGGL(k_cg_end, dim3(1), dim3(64), 0, s, sc, (double)maxit, inv_n);
    return hipGetLastError();
}
#endif

template hipError_t launch_pf_extract_dim<CMP_KERNEL_DIM>(const PfGeom&, long long, const double*,
                                                          double*, double*, int, hipStream_t);
template hipError_t launch_pf_scatter_dim<CMP_KERNEL_DIM>(const PfGeom&, const long long*, const double*,
                                                          double*, int, hipStream_t);
template hipError_t launch_pf_normals_dim<CMP_KERNEL_DIM>(const PfGeom&, long long, const double*,
                                                          double*, hipStream_t);
template hipError_t launch_pf_jfaces_dim<CMP_KERNEL_DIM>(const PfGeom&, const long long*, const double*,
                                                         const double*, const double*, double*, double*,
                                                         double, double, double, hipStream_t);
template hipError_t launch_pf_update_dim<CMP_KERNEL_DIM>(const PfGeom&, long long, double*, const double*,
                                                         double, hipStream_t);
template hipError_t launch_pf_gradq_dim<CMP_KERNEL_DIM>(const PfGeom&, const long long*, const double*,
                                                        const double*, double, double*, unsigned long long*,
                                                        int*, hipStream_t);
template hipError_t launch_pf_maxvel_dim<CMP_KERNEL_DIM>(const PfGeom&, long long, const double*,
                                                         unsigned long long*, int, hipStream_t);
template hipError_t launch_pf_cac_mech_dim<CMP_KERNEL_DIM>(const PfGeom&, long long, const double*,
                                                           const double*, double*, double*, double, double,
                                                           double, double*, hipStream_t);
template hipError_t launch_pf_cac_update_dim<CMP_KERNEL_DIM>(const PfGeom&, long long, double*, const double*,
                                                             const double*, const double*, double, hipStream_t);
template hipError_t launch_pf_divjhat_dim<CMP_KERNEL_DIM>(const PfGeom&, long long, const double*,
                                                          const double*, double*, double, hipStream_t);
template hipError_t launch_pf_bcoef_dim<CMP_KERNEL_DIM>(const PfGeom&, const long long*, const double*,
                                                        double*, hipStream_t);
template hipError_t launch_cell_sum_dim<CMP_KERNEL_DIM>(const PfGeom&, long long, const double*, double*,
                                                        hipStream_t);
template hipError_t launch_cg_idiag_dim<CMP_KERNEL_DIM>(const PfGeom&, long long, const double*, double*,
                                                        hipStream_t);
template hipError_t launch_cg_init_dim<CMP_KERNEL_DIM>(const PfGeom&, long long, const double*, const double*,
                                                       double, const double*, double*, double*, double*, double*,
                                                       double*, hipStream_t);
template hipError_t launch_cg_ap_dim<CMP_KERNEL_DIM>(const PfGeom&, long long, const double*, const double*,
                                                     const double*, double*, double*, hipStream_t);
template hipError_t launch_cg_xr_dim<CMP_KERNEL_DIM>(const PfGeom&, long long, const double*, const double*,
                                                     double*, double*, double*, const double*, const double*,
                                                     double*, hipStream_t);
template hipError_t launch_cg_p_dim<CMP_KERNEL_DIM>(const PfGeom&, long long, const double*, const double*,
                                                   double*, double, const int*, int, hipStream_t);
template hipError_t launch_pf_gradx_dim<CMP_KERNEL_DIM>(const PfGeom&, const long long*, const double*, double*,
                                                        unsigned long long*, int*, hipStream_t);

}  // namespace cmp
