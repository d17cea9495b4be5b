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
};

struct TileDev {
    int box;
    int lo[3];
};

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
// of WENO5 or WENO5B.
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
                        unsigned long long* cmax_bits, int* err, hipStream_t s);

hipError_t launch_fill(int dim, int N, const FillDev& fd, long long nghost, double* U, int* err,
                       hipStream_t s);

hipError_t launch_compute_dt(int dim, const double* dx, double* cmax, double* dt, double cfl,
                             int limiter, double* hist, int* step, int* err, hipStream_t s);

hipError_t launch_faces(const LaunchCfg& lc, const KParams& kp, const BoxDev& box, int dir,
                        const double* U, double* flux, double* fluxnc, hipStream_t s);

hipError_t launch_set2(int* p, int a, int b, hipStream_t s);

int tile_dims(int dim, int* tx, int* ty, int* tz);
int scheme_ghosts(int scheme);

}  // namespace cmp
