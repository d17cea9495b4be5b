// cmp_kernels.hpp - host-visible launch interface of the gfx950 kernels (internal to the library).
#pragma once

#include <hip/hip_runtime.h>
#include "cmp_device.hpp"

namespace cmp {

// One box of a level as the kernels see it: valid range, fab placement and strides.
struct BoxDev {
    int lo[3], hi[3];
    int flo[3];          // lowest stored index of the fab per dim
    int pad;
    long long off;       // element offset of the fab in the level buffer
    long long sj, sk, sn;
    // output layout (cmp_rhs_box writes dUdt over the valid box only)
    long long ooff;
    int olo[3];
    int pad2;
    long long osj, osk, osn;
    // offsets of this box's face arrays in a level face buffer (cmp_face_layout, 1 comp)
    long long goff[3];
};

struct TileDev {
    int box;
    int lo[3];
};

constexpr int NT = 256;   // threads per workgroup of every kernel here

__device__ __forceinline__ long long fab_index(const BoxDev& bx, int i, int j, int k)
{
    return (long long)(i - bx.flo[0]) + bx.sj * (j - bx.flo[1]) + bx.sk * (k - bx.flo[2]);
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = smax(v, __shfl_xor(v, off, 64));
    return v;
}

// Ghost-fill tables of a level.
struct FillDev {
    const BoxDev* boxes;
    const long long* gprefix;   // nbox+1 prefix sums of ghost-cell counts
    const int* owner;           // box owning each bs^d block of the domain (-1 if none)
    int nbox;
    int ng;
    int bs;
    int nb[3];
    int dlo[3], dhi[3];
    int periodic[3];
    int lo_bc[3], hi_bc[3];
};

enum StageMode { MODE_RK = 0, MODE_RHS = 1, MODE_CMAX = 2 };

// Scheme selector: 0 Godunov, 1 MUSCL (limiter from id_limiter), 5 WENO5, 6 WENO5B; THINC on top
// of WENO5 or WENO5B. riemann: 1 HLL, 2 HLLC for the flow's fluxes; RIEMANN_PF_MIX / RIEMANN_PF_PHASIC
// select the Phase-Field flux (UD, mixture or phasic formulation), read with a Grad_Q face buffer.
enum { RIEMANN_PF_MIX = 10, RIEMANN_PF_PHASIC = 11 };
struct LaunchCfg {
    int dim;
    int scheme;
    int limiter;
    int riemann;
    int thinc;
};

hipError_t launch_stage(const LaunchCfg& lc, const KParams& kp, int mode, int stage,
                        const TileDev* tiles, int ntiles, const BoxDev* boxes,
                        const BoxDev* single_box, int single_ntiles,
                        const double* Us, const double* Un, double* Uout, const double* dt,
                        unsigned long long* cmax_bits, int* err, hipStream_t s,
                        const double* gradq = nullptr);

// N = 4 + dim: the flow state (reflecting BCs flip momenta); N = 1 or dim: a Phase-Field field
// (alpha1 or the normals), filled without sign changes
hipError_t launch_fill(int dim, int N, const FillDev& fd, long long nghost, double* U, int* err,
                       hipStream_t s);

// ---- Phase-Field mechanism kernels (grid.y = box) ----
// PfGeom: per-level constants of the mechanism kernels.
struct PfGeom {
    const BoxDev* hbox;    // flow-state layout (ng ghosts, nstate comps)
    const BoxDev* abox;    // alpha layout (1 ghost, 1 comp); normals use DIM * off with the same strides
    int nbox;
    double dx[3];
};

hipError_t launch_pf_extract(int dim, const PfGeom& g, long long maxcells, const double* U, double* A,
                             hipStream_t s);
// psi (mode 0) or the re-sharpened alpha1 (mode 1) of every element of an alpha-layout buffer
hipError_t launch_pf_map(const double* A, double* out, long long n, double eta, double eta2, int mode,
                         hipStream_t s);
hipError_t launch_pf_normals(int dim, const PfGeom& g, long long maxcells, const double* Psi,
                             double* Nrm, hipStream_t s);
hipError_t launch_pf_jfaces(int dim, const PfGeom& g, const long long* maxfaces, const double* A,
                            const double* Psi, const double* Nrm, double* J, double* H, double mob_face,
                            double eta, double dt_m, hipStream_t s);
hipError_t launch_pf_update(int dim, const PfGeom& g, long long maxcells, double* A, const double* J,
                            double dt_m, hipStream_t s);
hipError_t launch_pf_gradq(int dim, const PfGeom& g, const long long* maxfaces, const double* S,
                           const double* H, double* gq, unsigned long long* gmax_bits, hipStream_t s);
hipError_t launch_pf_maxvel(int dim, const PfGeom& g, long long maxcells, const double* U,
                            unsigned long long* bits, hipStream_t s);
hipError_t launch_scale(double* p, long long n, double s_, hipStream_t s);
hipError_t launch_setd(double* p, double v, hipStream_t s);
hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t s);

hipError_t launch_compute_dt(int dim, const double* dx, double* cmax, double* dt, double cfl,
                             int limiter, double* hist, int* step, int* err, hipStream_t s);

hipError_t launch_faces(const LaunchCfg& lc, const KParams& kp, const BoxDev& box, int dir,
                        const double* U, double* flux, double* fluxnc, hipStream_t s);

hipError_t launch_set2(int* p, int a, int b, hipStream_t s);

int tile_dims(int dim, int* tx, int* ty, int* tz);
int scheme_ghosts(int scheme);

}  // namespace cmp
