// cmp_pf.hip - gfx950 kernels of the Phase-Field step's mechanism and Grad_Q (ACDI, explicit RC).
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
    i = bx.lo[0] + (int)(q % n0);
    j = bx.lo[1] + (int)((q / n0) % n1);
    k = bx.lo[2] + (int)(q / ((long long)n0 * n1));
    return true;
}

// Face coordinates (of the face's high cell) of linear index q of the direction-D nodal faces.
template <int DIM, int D>
__device__ __forceinline__ bool face_of(const BoxDev& bx, long long q, int& i, int& j, int& k)
{
    const int n0 = bx.hi[0] - bx.lo[0] + 1 + (D == 0), n1 = bx.hi[1] - bx.lo[1] + 1 + (D == 1);
    const int n2 = (DIM == 3) ? bx.hi[2] - bx.lo[2] + 1 + (D == 2) : 1;
    if (q >= (long long)n0 * n1 * n2) return false;
    i = bx.lo[0] + (int)(q % n0);
    j = bx.lo[1] + (int)((q / n0) % n1);
    k = bx.lo[2] + (int)(q / ((long long)n0 * n1));
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

}  // namespace

// alpha1 of the flow state's valid cells -> the alpha buffer
template <int DIM>
__global__ __launch_bounds__(NT) void k_pf_extract(PfGeom g, const double* __restrict__ U, double* A)
{
    const BoxDev hb = g.hbox[blockIdx.y], ab = g.abox[blockIdx.y];
    int i, j, k;
    if (!cell_of<DIM>(ab, (long long)blockIdx.x * NT + threadIdx.x, i, j, k)) return;
    A[ab.off + fab_index(ab, i, j, k)] = U[hb.off + fab_index(hb, i, j, k) + (3 + DIM) * hb.sn];
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

// ACDI flux on every face of direction D: J stored, Jhat += dt_m J
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
    H[f] = H[f] + dt_m * Jf;
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
                                                 const double* __restrict__ H, double* gq,
                                                 unsigned long long* gmax_bits)
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
        const double r = pf_gradq_sharp(Sb[m], Sb[c], H[f]);
        gq[f] = r;
        v = fabs(r);
        if (__builtin_isnan(r)) v = __longlong_as_double(0x7ff8000000000000LL);
    }
    block_max_to(v, gmax_bits);
}

// max over valid cells of |u| (MaxVelocity, :3541-3579; Conservative2Velocity_Compressible5Eq.H:15-40)
template <int DIM>
__global__ __launch_bounds__(NT) void k_pf_maxvel(PfGeom g, const double* __restrict__ U,
                                                  unsigned long long* bits)
{
    const BoxDev hb = g.hbox[blockIdx.y];
    int i, j, k;
    double v = 0.0;
    if (cell_of<DIM>(hb, (long long)blockIdx.x * NT + threadIdx.x, i, j, k)) {
        const double* Ub = U + hb.off;
        const long long c = fab_index(hb, i, j, k);
        const double rho = Ub[c] + Ub[c + hb.sn];
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const double u = Ub[c + (2 + d) * hb.sn] / rho;
            s = (d == 0) ? u * u : s + u * u;
        }
        v = sqrt(s);
        if (__builtin_isnan(v)) v = __longlong_as_double(0x7ff8000000000000LL);
    }
    block_max_to(v, bits);
}

// ---- launchers -----------------------------------------------------------------------------------

static dim3 grid_of(long long per_box, int nbox)
{
    return dim3((unsigned)((per_box + NT - 1) / NT), (unsigned)nbox);
}

template <int DIM>
hipError_t launch_pf_extract_dim(const PfGeom& g, long long maxcells, const double* U, double* A,
                                 hipStream_t s)
{
    hipLaunchKernelGGL((k_pf_extract<DIM>), grid_of(maxcells, g.nbox), dim3(NT), 0, s, g, U, A);
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
                               const double* H, double* gq, unsigned long long* gmax_bits, hipStream_t s)
{
    hipLaunchKernelGGL((k_pf_gradq<DIM, 0>), grid_of(maxfaces[0], g.nbox), dim3(NT), 0, s, g, S, H, gq,
                       gmax_bits);
    hipLaunchKernelGGL((k_pf_gradq<DIM, 1>), grid_of(maxfaces[1], g.nbox), dim3(NT), 0, s, g, S, H, gq,
                       gmax_bits);
    if (DIM == 3)
        hipLaunchKernelGGL((k_pf_gradq<DIM, (DIM == 3 ? 2 : 1)>), grid_of(maxfaces[2], g.nbox), dim3(NT),
                           0, s, g, S, H, gq, gmax_bits);
    return hipGetLastError();
}

template <int DIM>
hipError_t launch_pf_maxvel_dim(const PfGeom& g, long long maxcells, const double* U,
                                unsigned long long* bits, hipStream_t s)
{
    hipLaunchKernelGGL((k_pf_maxvel<DIM>), grid_of(maxcells, g.nbox), dim3(NT), 0, s, g, U, bits);
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
#endif

template hipError_t launch_pf_extract_dim<CMP_KERNEL_DIM>(const PfGeom&, long long, const double*,
                                                          double*, hipStream_t);
template hipError_t launch_pf_normals_dim<CMP_KERNEL_DIM>(const PfGeom&, long long, const double*,
                                                          double*, hipStream_t);
template hipError_t launch_pf_jfaces_dim<CMP_KERNEL_DIM>(const PfGeom&, const long long*, const double*,
                                                         const double*, const double*, double*, double*,
                                                         double, double, double, hipStream_t);
template hipError_t launch_pf_update_dim<CMP_KERNEL_DIM>(const PfGeom&, long long, double*, const double*,
                                                         double, hipStream_t);
template hipError_t launch_pf_gradq_dim<CMP_KERNEL_DIM>(const PfGeom&, const long long*, const double*,
                                                        const double*, double*, unsigned long long*,
                                                        hipStream_t);
template hipError_t launch_pf_maxvel_dim<CMP_KERNEL_DIM>(const PfGeom&, long long, const double*,
                                                         unsigned long long*, hipStream_t);

}  // namespace cmp
