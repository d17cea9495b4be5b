"""Initial states of the benchmark and test cases (synthetic, seeded, FP64).

Cell-centred evaluation (the reference's run.QuadratureIC = 0, include/InitData.H:66-98). The
primitive->conservative map used here is the EOS of Parm::Initialize (source/Parm.cpp:124-156),
rho e = A p + B + rho D with A, B mixed by volume fraction; CPU and GPU start from the same arrays.

Conventions: the state's a1 component is the volume fraction of phase 1; in the advected-interface
case phase 1 (gamma 1.4, rho 1) is inside the circle and phase 2 (gamma 1.6, rho 10) outside; in the
water-air case phase 1 is water (the droplet) and phase 2 air.
"""
import math

import numpy as np

from . import make_geometry, make_params, split_domain


def prim_to_cons(params, W):
    """W = [a1r1, a2r2, u.., p, a1] (leading axis) -> conservative state (same shape)."""
    dim = params.dim
    a1r1, a2r2 = W[0], W[1]
    vel = [W[2 + m] for m in range(dim)]
    p, a1 = W[2 + dim], W[3 + dim]
    a2 = 1.0 - a1
    A, B, C, D = params.eos_A, params.eos_B, params.eos_C, params.eos_D
    phi1 = np.where(C[0] > 0.0, a1 - C[0] * a1r1, a1)
    phi2 = np.where(C[1] > 0.0, a2 - C[1] * a2r2, a2)
    AP = phi1 * A[0] + phi2 * A[1]
    BP = phi1 * B[0] + phi2 * B[1]
    rho = a1r1 + a2r2
    rho_e = AP * p + BP + (a1r1 * D[0] + a2r2 * D[1])
    ke = sum(v * v for v in vel)
    U = np.empty_like(W)
    U[0], U[1] = a1r1, a2r2
    for m in range(dim):
        U[2 + m] = rho * vel[m]
    U[2 + dim] = rho_e + 0.5 * rho * ke
    U[3 + dim] = a1
    return U


def _centres(geom, shape):
    dim = geom.dim
    axes = []
    for d in range(dim):
        n = shape[dim - 1 - d]
        axes.append(geom.prob_lo[d] + (np.arange(n) + 0.5) * geom.dx[d])
    grids = np.meshgrid(*reversed(axes), indexing="ij")   # [(k,) j, i] ordering
    return list(reversed(grids))                          # x, y, (z)


def advected_interface_2d(n=256, max_grid_size=None, scheme="WENO5", riemann="HLLC"):
    """2D two-phase advected interface: [0,1]^2 periodic, u = v = 1, p = 1, circle r = 0.25 at
    (0.5, 0.5), alpha = 0.5 (1 + tanh((0.25 - r)/(2 dx))) clipped to [1e-8, 1 - 1e-8]."""
    params = make_params(2, gamma=(1.4, 1.6), scheme=scheme, riemann=riemann, source_term=0)
    geom = make_geometry((n, n), (0.0, 0.0), (1.0, 1.0), (1, 1))
    x, y = _centres(geom, (n, n))
    r = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2)
    a1 = np.clip(0.5 * (1.0 + np.tanh((0.25 - r) / (2.0 * geom.dx[0]))), 1e-8, 1.0 - 1e-8)
    W = np.empty((6, n, n))
    W[0] = a1 * 1.0
    W[1] = (1.0 - a1) * 10.0
    W[2] = 1.0
    W[3] = 1.0
    W[4] = 1.0
    W[5] = a1
    U = prim_to_cons(params, W)
    boxes = split_domain((n, n), max_grid_size or n)
    return dict(params=params, geom=geom, U=U, boxes=boxes, shape=(n, n), cfl=0.5,
                name="advected-interface-2d")


def rankine_hugoniot(gamma, p1, rho1, mach):
    """Post-shock (p2, rho2, u2) of a shock moving at Mach `mach` into gas at rest
    (ideal gas; the relations of include/PhysicsFunctions.H:15-34)."""
    g = gamma
    M2 = mach * mach
    p2 = p1 * (1.0 + 2.0 * g / (g + 1.0) * (M2 - 1.0))
    rho2 = rho1 * (g + 1.0) * M2 / ((g - 1.0) * M2 + 2.0)
    c1 = math.sqrt(g * p1 / rho1)
    u2 = mach * c1 * (1.0 - rho1 / rho2)
    return p2, rho2, u2


def water_air_droplet_3d(n=256, box=32, scheme="WENO5", riemann="HLLC", thinc=1):
    """3D water-air shock-droplet: unit cube, outflow BCs, water (stiffened gas gamma 4.4,
    p_inf 6e8 Pa, rho 1000) droplet r = 0.1 at the centre, air (gamma 1.4, rho 1.2), p = 1e5 Pa,
    Mach 2.4 shock in air at x = 0.25 moving in +x. Interface: tanh profile of width dx. THINC on
    a WENO5 base scheme (FiniteVolume.THINC = 1 with its default parameters)."""
    params = make_params(3, gamma=(4.4, 1.4), pinf=(6.0e8, 0.0), scheme=scheme, riemann=riemann,
                         source_term=0, thinc=thinc)
    geom = make_geometry((n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (0, 0, 0))
    x, y, z = _centres(geom, (n, n, n))
    r = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2)
    a1 = np.clip(0.5 * (1.0 + np.tanh((0.1 - r) / (2.0 * geom.dx[0]))), 1e-8, 1.0 - 1e-8)
    p0, rho_w, rho_a = 1.0e5, 1000.0, 1.2
    p2, rho2, u2 = rankine_hugoniot(1.4, p0, rho_a, 2.4)
    post = x < 0.25
    W = np.empty((7, n, n, n))
    W[0] = a1 * rho_w
    W[1] = (1.0 - a1) * np.where(post, rho2, rho_a)
    W[2] = np.where(post, u2, 0.0)
    W[3] = 0.0
    W[4] = 0.0
    W[5] = np.where(post, p2, p0)
    W[6] = a1
    U = prim_to_cons(params, W)
    boxes = split_domain((n, n, n), box)
    return dict(params=params, geom=geom, U=U, boxes=boxes, shape=(n, n, n), cfl=0.4,
                name="water-air-droplet-3d")


def random_smooth(dim, n, seed, periodic=True, max_grid_size=None, scheme="WENO5", riemann="HLLC",
                  gamma=(1.4, 1.6), pinf=(0.0, 0.0), source_term=0, amp=0.1, thinc=0):
    """Seeded smooth two-phase state for parity tests: a few random Fourier modes on top of a
    uniform state, with a tanh interface across the middle."""
    rng = np.random.default_rng(seed)
    params = make_params(dim, gamma=gamma, pinf=pinf, scheme=scheme, riemann=riemann,
                         source_term=source_term, thinc=thinc)
    shape = (n,) * dim
    geom = make_geometry(shape, (0.0,) * dim, (1.0,) * dim, (int(periodic),) * dim)
    X = _centres(geom, tuple(reversed(shape)))

    def modes():
        f = np.zeros_like(X[0])
        for _ in range(3):
            k = rng.integers(1, 3, size=dim)
            ph = rng.uniform(0, 2 * np.pi, size=dim)
            term = np.ones_like(X[0])
            for d in range(dim):
                term = term * np.sin(2 * np.pi * k[d] * X[d] + ph[d])
            f = f + rng.uniform(-1, 1) * term
        return f / 3.0

    xi = X[0] - 0.5 + 0.05 * modes()
    a1 = np.clip(0.5 * (1.0 + np.tanh(np.sin(2 * np.pi * xi) / (4.0 * geom.dx[0]))), 1e-8, 1 - 1e-8)
    pscale = 1.0 + max(pinf)
    W = np.empty((4 + dim,) + tuple(reversed(shape)))
    W[0] = a1 * (1.0 + amp * modes())
    W[1] = (1.0 - a1) * (5.0 + amp * 5.0 * modes())
    for m in range(dim):
        W[2 + m] = 0.5 + amp * modes()
    W[2 + dim] = pscale * (1.0 + amp * modes())
    W[3 + dim] = a1
    U = prim_to_cons(params, W)
    boxes = split_domain(shape, max_grid_size or n)
    return dict(params=params, geom=geom, U=U, boxes=boxes, shape=tuple(reversed(shape)), cfl=0.4,
                name="random-smooth-%dd" % dim)


def two_level_smooth(dim, nc, seed, fine_boxes, ratio=2, periodic=True, nphase=0, **kw):
    """A two-level hierarchy of the seeded smooth state: the fine level's valid cells from
    random_smooth (nphase_smooth for nphase > 0) at nc * ratio cells a side, the coarse level (nc a
    side) the mean of the fine field over each coarse cell. fine_boxes in fine cells. Returns
    dict(params, geoms, boxes, fields, shapes, cfl) with the coarse level first."""
    if nphase:
        f = nphase_smooth(dim, nc * ratio, seed, nphase=nphase, periodic=periodic, **kw)
    else:
        f = random_smooth(dim, nc * ratio, seed, periodic=periodic, **kw)
    Uf = f["U"]
    shp = Uf.shape[1:]
    Uc = Uf.reshape((Uf.shape[0],) + tuple(x for n in shp for x in (n // ratio, ratio)))
    Uc = Uc.mean(axis=tuple(2 + 2 * a for a in range(dim)))
    cg = make_geometry((nc,) * dim, (0.0,) * dim, (1.0,) * dim, (int(periodic),) * dim)
    return dict(params=f["params"], geoms=[cg, f["geom"]], boxes=[split_domain((nc,) * dim, max(8, nc // 2)),
                                                                  [(list(lo), list(hi)) for lo, hi in fine_boxes]],
                fields=[np.ascontiguousarray(Uc), Uf], shapes=[Uc.shape[1:], shp], cfl=f["cfl"])


# ---- the N-phase model (cmp_params.nphase = 2 or 3) --------------------------------------------------

def np_prim_to_cons(params, W):
    """N-phase primitives [alpha_k rho_k (NP), u.., p, alpha_k (NP)] -> conservative (same shape)."""
    dim, np_ = params.dim, params.nphase
    A, B, C, D = params.eos_A, params.eos_B, params.eos_C, params.eos_D
    arho = [W[k] for k in range(np_)]
    alpha = [W[np_ + dim + 1 + k] for k in range(np_)]
    vel = [W[np_ + m] for m in range(dim)]
    p = W[np_ + dim]
    AP = sum(np.where(C[k] > 0.0, alpha[k] - C[k] * arho[k], alpha[k]) * A[k] for k in range(np_))
    BP = sum(np.where(C[k] > 0.0, alpha[k] - C[k] * arho[k], alpha[k]) * B[k] for k in range(np_))
    rho = sum(arho)
    rho_e = AP * p + BP + sum(arho[k] * D[k] for k in range(np_))
    U = np.empty_like(W)
    for k in range(np_):
        U[k] = arho[k]
        U[np_ + dim + 1 + k] = alpha[k]
    for m in range(dim):
        U[np_ + m] = rho * vel[m]
    U[np_ + dim] = rho_e + 0.5 * rho * sum(v * v for v in vel)
    return U


def nphase_from_two_phase(c, nphase=3, split=0.35):
    """The same flow in the N-phase model: nphase 2 keeps both phases (alpha2 = 1 - alpha1 stored);
    nphase 3 splits phase 2 into two phases with phase 2's EOS, a fraction `split` of its volume and
    mass in the second one (identical densities: the mixture is the same material)."""
    P2 = c["params"]
    dim = P2.dim
    U2 = c["U"]
    W2 = _cons_to_prim_two(P2, U2)
    a1r1, a2r2 = W2[0], W2[1]
    a1 = W2[3 + dim]
    gam = [1.0 + 1.0 / P2.eos_A[k] for k in range(2)]
    pinf = [P2.eos_B[k] / (P2.eos_A[k] + 1.0) for k in range(2)]
    if nphase == 2:
        g, pi = gam, pinf
        arho = [a1r1, a2r2]
        alpha = [a1, 1.0 - a1]
    else:
        g, pi = [gam[0], gam[1], gam[1]], [pinf[0], pinf[1], pinf[1]]
        arho = [a1r1, (1.0 - split) * a2r2, split * a2r2]
        alpha = [a1, (1.0 - split) * (1.0 - a1), split * (1.0 - a1)]
    scheme = [k for k, v in SCHEME_IDS.items() if v == (P2.id_weno, P2.id_limiter)][0]
    riemann = "HLLC" if P2.id_riemann == 2 else "HLL"
    P = make_params(dim, gamma=tuple(g), pinf=tuple(pi), scheme=scheme, riemann=riemann,
                    source_term=P2.source_term, nphase=nphase)
    W = np.empty((P.nstate,) + U2.shape[1:])
    for k in range(nphase):
        W[k] = arho[k]
        W[nphase + dim + 1 + k] = alpha[k]
    for m in range(dim):
        W[nphase + m] = W2[2 + m]
    W[nphase + dim] = W2[2 + dim]
    out = dict(c)
    out["params"] = P
    out["U"] = np_prim_to_cons(P, W)
    return out


def _cons_to_prim_two(P, U):
    dim = P.dim
    a1r1, a2r2 = U[0], U[1]
    rho = a1r1 + a2r2
    vel = [U[2 + m] / rho for m in range(dim)]
    a1 = U[3 + dim]
    a2 = 1.0 - a1
    A, B, C, D = P.eos_A, P.eos_B, P.eos_C, P.eos_D
    phi1 = np.where(C[0] > 0.0, a1 - C[0] * a1r1, a1)
    phi2 = np.where(C[1] > 0.0, a2 - C[1] * a2r2, a2)
    AP = phi1 * A[0] + phi2 * A[1]
    BP = phi1 * B[0] + phi2 * B[1]
    rho_e = U[2 + dim] - 0.5 * rho * sum(v * v for v in vel)
    W = np.empty_like(U)
    W[0], W[1] = a1r1, a2r2
    for m in range(dim):
        W[2 + m] = vel[m]
    W[2 + dim] = (rho_e - (a1r1 * D[0] + a2r2 * D[1]) - BP) / AP
    W[3 + dim] = a1
    return W


SCHEME_IDS = {"WENO5": (1, 5), "WENO5B": (1, 6), "MUSCL": (0, 1), "GODUNOV": (0, 0)}


def nphase_smooth(dim, n, seed, nphase=3, periodic=True, max_grid_size=None, scheme="WENO5",
                  riemann="HLLC", gamma=(1.4, 1.67, 1.09), pinf=(0.0, 0.0, 0.0), source_term=0,
                  amp=0.1):
    """Seeded smooth N-phase state for parity tests: volume fractions from smooth positive weights
    (normalised to sum 1), phase densities, velocity and pressure with a few random Fourier modes."""
    rng = np.random.default_rng(seed)
    params = make_params(dim, gamma=tuple(gamma[:nphase]), pinf=tuple(pinf[:nphase]), scheme=scheme,
                         riemann=riemann, source_term=source_term, nphase=nphase)
    shape = (n,) * dim
    geom = make_geometry(shape, (0.0,) * dim, (1.0,) * dim, (int(periodic),) * dim)
    X = _centres(geom, tuple(reversed(shape)))

    def modes():
        f = np.zeros_like(X[0])
        for _ in range(3):
            k = rng.integers(1, 3, size=dim)
            ph = rng.uniform(0, 2 * np.pi, size=dim)
            term = np.ones_like(X[0])
            for d in range(dim):
                term = term * np.sin(2 * np.pi * k[d] * X[d] + ph[d])
            f = f + rng.uniform(-1, 1) * term
        return f / 3.0

    wts = [np.exp(3.0 * modes()) for _ in range(nphase)]
    tot = sum(wts)
    alpha = [w / tot for w in wts]
    pscale = 1.0 + max(pinf[:nphase])
    W = np.empty((2 * nphase + dim + 1,) + tuple(reversed(shape)))
    for k in range(nphase):
        W[k] = alpha[k] * ((1.0 + 2.0 * k) * (1.0 + amp * modes()))
        W[nphase + dim + 1 + k] = alpha[k]
    for m in range(dim):
        W[nphase + m] = 0.5 + amp * modes()
    W[nphase + dim] = pscale * (1.0 + amp * modes())
    U = np_prim_to_cons(params, W)
    boxes = split_domain(shape, max_grid_size or n)
    return dict(params=params, geom=geom, U=U, boxes=boxes, shape=shape, cfl=0.4,
                name="nphase-smooth")


# ---- the served stream: seeded 2D two-phase shock-bubble cases -----------------------------------

SERVED_GRIDS = [(128, 128), (256, 256), (512, 128), (512, 512), (1024, 256)]


def shock_bubble_2d(nx, ny, mach, rho_ratio, gamma_bubble, gamma_ambient, radius, max_grid_size=64,
                    scheme="WENO5", riemann="HLLC", xb=None, xs=None):
    """2D two-phase shock-bubble: domain [0, nx/ny] x [0, 1] (dx = dy = 1/ny), ambient ideal gas
    (phase 2: gamma_ambient, rho 1, p 1) at rest, a bubble of phase 1 (gamma_bubble, rho_ratio) of
    radius `radius` at (x_b, 0.5) with a tanh profile of width dx clipped to [1e-8, 1 - 1e-8], and a
    planar shock of Mach `mach` at x_b - radius - 0.1 moving in +x with the Rankine-Hugoniot
    post-shock state behind it (include/PhysicsFunctions.H:15-34). Outflow (foextrap) on every
    side: the post-shock inflow side holds a uniform state. xb / xs place the bubble centre and the
    shock explicitly."""
    width = nx / ny
    params = make_params(2, gamma=(gamma_bubble, gamma_ambient), scheme=scheme, riemann=riemann, source_term=0)
    lo_bc, hi_bc = [2, 2], [2, 2]
    geom = make_geometry((nx, ny), (0.0, 0.0), (width, 1.0), (0, 0), lo_bc=lo_bc, hi_bc=hi_bc)
    x, y = _centres(geom, (ny, nx))
    if xb is None:
        xb = max(radius + 0.25, 0.45 * width)
    if xs is None:
        xs = max(xb - radius - 0.1, 0.05)
    r = np.sqrt((x - xb) ** 2 + (y - 0.5) ** 2)
    a1 = np.clip(0.5 * (1.0 + np.tanh((radius - r) / (2.0 * geom.dx[0]))), 1e-8, 1.0 - 1e-8)
    p2, rho2, u2 = rankine_hugoniot(gamma_ambient, 1.0, 1.0, mach)
    post = x < xs
    W = np.empty((6, ny, nx))
    W[0] = a1 * rho_ratio
    W[1] = (1.0 - a1) * np.where(post, rho2, 1.0)
    W[2] = np.where(post, u2, 0.0)
    W[3] = 0.0
    W[4] = np.where(post, p2, 1.0)
    W[5] = a1
    U = prim_to_cons(params, W)
    boxes = split_domain((nx, ny), max_grid_size)
    return dict(params=params, geom=geom, U=U, boxes=boxes, shape=(ny, nx), cfl=0.4,
                name="shock-bubble-2d")


def served_case(seed, max_grid_size=64, grids=None):
    """One case of the served stream, drawn from its seed: grid from SERVED_GRIDS, shock Mach
    uniform in [1.1, 3.0], bubble/ambient density ratio log-uniform in [0.1, 10], both gammas
    uniform in [1.1, 1.7], bubble radius uniform in [0.1, 0.3] of the domain height."""
    rng = np.random.default_rng(seed)
    grids = grids or SERVED_GRIDS
    nx, ny = grids[int(rng.integers(len(grids)))]
    mach = rng.uniform(1.1, 3.0)
    ratio = float(np.exp(rng.uniform(np.log(0.1), np.log(10.0))))
    gb, ga = rng.uniform(1.1, 1.7), rng.uniform(1.1, 1.7)
    radius = rng.uniform(0.1, 0.3)
    c = shock_bubble_2d(nx, ny, mach, ratio, gb, ga, radius, max_grid_size=max_grid_size)
    c["seed"] = seed
    c["draw"] = dict(grid=(nx, ny), mach=mach, rho_ratio=ratio, gamma_bubble=gb, gamma_ambient=ga,
                     radius=radius)
    return c


# ---- the AMR configuration's base level: three-phase converging rings on uneven boxes -------------

def uneven_boxes(n_cell, lo_size=8, hi_size=64, seed=0):
    """A seeded split of the domain into boxes of lo_size..hi_size cells a side (multiples of
    lo_size), the shape a regrid's box list has: recursive cuts at random multiples of lo_size until
    every box is at most hi_size a side."""
    rng = np.random.default_rng(seed)
    dim = len(n_cell)
    out = []
    stack = [([0] * dim, [n - 1 for n in n_cell])]
    while stack:
        lo, hi = stack.pop()
        ext = [hi[d] - lo[d] + 1 for d in range(dim)]
        d = int(np.argmax(ext))
        if ext[d] <= hi_size and rng.uniform() < 0.5 or ext[d] < 2 * lo_size:
            if max(ext) <= hi_size:
                out.append((list(lo), list(hi)))
                continue
            d = int(np.argmax(ext))
        blocks = ext[d] // lo_size
        cut = lo[d] + lo_size * int(rng.integers(1, blocks))
        a_hi = list(hi)
        a_hi[d] = cut - 1
        b_lo = list(lo)
        b_lo[d] = cut
        stack.append((list(lo), a_hi))
        stack.append((b_lo, list(hi)))
    return out


def nphase_rings_2d(n=512, seed=0, boxes=None, scheme="WENO5", riemann="HLLC"):
    """The AMR configuration's initial state on its base grid: three ideal gases, (gamma, rho) =
    (1.4, 1), (1.67, 0.14), (1.09, 5.0), in [0,1]^2 with two concentric annular interfaces at r = 0.3
    and r = 0.35 around the centre (phase 1 inside, phase 2 in the annulus, phase 3 outside), each
    with a seeded 8-mode sinusoidal perturbation of amplitude 0.005 and a tanh profile of width dx,
    and an inward-moving Mach 1.5 shock at r = 0.45 (post-shock state of phase 3 outside it, moving
    towards the centre). Outflow sides. Boxes: uneven_boxes (8..64 cells a side) unless given."""
    rng = np.random.default_rng(seed)
    params = make_params(2, gamma=(1.4, 1.67, 1.09), scheme=scheme, riemann=riemann, nphase=3)
    geom = make_geometry((n, n), (0.0, 0.0), (1.0, 1.0), (0, 0), lo_bc=[2, 2], hi_bc=[2, 2])
    x, y = _centres(geom, (n, n))
    r = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2)
    th = np.arctan2(y - 0.5, x - 0.5)

    def perturbed(r0):
        s = np.zeros_like(th)
        for m in range(1, 9):
            s += rng.uniform(-1, 1) * np.cos(m * th + rng.uniform(0, 2 * np.pi))
        return r0 + 0.005 * s / 8.0

    w = 2.0 * geom.dx[0]
    h1 = 0.5 * (1.0 + np.tanh((perturbed(0.30) - r) / w))
    h2 = 0.5 * (1.0 + np.tanh((perturbed(0.35) - r) / w))
    a = [np.clip(h1, 1e-8, 1.0), np.clip(h2 - h1, 1e-8, 1.0), np.clip(1.0 - h2, 1e-8, 1.0)]
    tot = sum(a)
    a = [v / tot for v in a]
    p2, rho2, u2 = rankine_hugoniot(1.09, 1.0, 5.0, 1.5)
    post = r > 0.45
    rho = [np.full_like(r, 1.0), np.full_like(r, 0.14), np.where(post, rho2, 5.0)]
    W = np.empty((9, n, n))
    for k in range(3):
        W[k] = a[k] * rho[k]
        W[6 + k] = a[k]
    ur = np.where(post, -u2, 0.0)
    W[3] = ur * np.cos(th)
    W[4] = ur * np.sin(th)
    W[5] = np.where(post, p2, 1.0)
    U = np_prim_to_cons(params, W)
    boxes = boxes if boxes is not None else uneven_boxes((n, n), seed=seed)
    return dict(params=params, geom=geom, U=U, boxes=boxes, shape=(n, n), cfl=0.4, name="nphase-rings-2d")


def interface_hierarchy(dim, nc, fine_boxes, ratio=2, periodic=True, width=2.0, eos=None, nphase=0,
                        velocity=(1.0, 0.5, -0.25), max_grid_size=None, rings=False):
    """A circle (sphere) of phase 1 of radius 0.25 about the domain centre in phase 2, uniform u and p,
    tanh profile of width*dx at the finest level, on a hierarchy: level 0 nc cells a side, every
    finer level the refinement of the one before by ratio with the boxes fine_boxes[l - 1] (in that
    level's cells). The finest level's field is evaluated at its cell centres and every coarser
    level's is the mean of it over each cell (the state AverageDown leaves). eos: dict(gamma, pinf);
    nphase 3: phase 2 split 65/35 into two phases of its material (every phase fraction of the outer
    region lies inside (0, 1), where the sharpening normal is set by rounding), or with rings=True
    three ideal gases with separate interfaces: a disk of radius 0.2 (phase 1) inside a ring to
    radius 0.32 (phase 3) inside phase 2. Returns dict(params, geoms, boxes, fields, shapes) with the
    coarsest level first."""
    eos = eos or dict(gamma=(1.4, 1.6), pinf=(0.0, 0.0))
    nlev = 1 + len(fine_boxes)
    nf = nc * ratio ** (nlev - 1)
    P = make_params(dim, **eos)
    shape_f = (nf,) * dim
    Gf = make_geometry(shape_f, (0.0,) * dim, (1.0,) * dim, (int(periodic),) * dim)
    X = _centres(Gf, shape_f)
    r = np.sqrt(sum((X[d] - 0.5) ** 2 for d in range(dim)))
    a1 = np.clip(0.5 * (1.0 + np.tanh((0.25 - r) / (width * Gf.dx[0]))), 1e-8, 1 - 1e-8)
    sg = eos["pinf"][0] > 0
    W = np.empty((P.nstate,) + shape_f)
    W[0] = a1 * (1000.0 if sg else 1.0)
    W[1] = (1 - a1) * (1.2 if sg else 10.0)
    for m in range(dim):
        W[2 + m] = velocity[m]
    W[2 + dim] = 1.0e5 if sg else 1.0
    W[3 + dim] = a1
    U = prim_to_cons(P, W)
    c = dict(params=P, U=U)
    if nphase == 3 and rings:
        P = make_params(dim, gamma=(1.4, 1.67, 1.09), pinf=(0.0, 0.0, 0.0), nphase=3)
        h1 = 0.5 * (1 + np.tanh((0.2 - r) / (width * Gf.dx[0])))
        h2 = 0.5 * (1 + np.tanh((0.32 - r) / (width * Gf.dx[0])))
        a = [np.clip(v, 1e-8, 1.0) for v in (h1, 1.0 - h2, h2 - h1)]
        tot = sum(a)
        W = np.empty((P.nstate,) + shape_f)
        for k, rho in enumerate((1.0, 0.14, 5.0)):
            W[k] = a[k] / tot * rho
            W[4 + dim + k] = a[k] / tot
        for m in range(dim):
            W[3 + m] = velocity[m]
        W[3 + dim] = 1.0
        c = dict(params=P, U=np_prim_to_cons(P, W))
    elif nphase:
        c = nphase_from_two_phase(c, nphase=nphase)
    P, U = c["params"], c["U"]
    fields, shapes, geoms, boxes = [], [], [], []
    for l in range(nlev):
        n = nc * ratio ** l
        f = ratio ** (nlev - 1 - l)
        Ul = U.reshape((U.shape[0],) + tuple(x for s in shape_f for x in (s // f, f)))
        Ul = Ul.mean(axis=tuple(2 + 2 * a for a in range(dim))) if f > 1 else U
        fields.append(np.ascontiguousarray(Ul))
        shapes.append(Ul.shape[1:])
        geoms.append(make_geometry((n,) * dim, (0.0,) * dim, (1.0,) * dim, (int(periodic),) * dim))
        if l == 0:
            boxes.append(split_domain((nc,) * dim, max_grid_size or max(8, nc // 2)))
        else:
            boxes.append([(list(lo), list(hi)) for lo, hi in fine_boxes[l - 1]])
    return dict(params=P, geoms=geoms, boxes=boxes, fields=fields, shapes=shapes)
