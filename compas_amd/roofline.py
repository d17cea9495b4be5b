"""Algorithmic work per cell of the five-equation RK3 step (the roofline's numerator).

Unit: one cell-update = one cell advanced by one full TVD-RK3 step (three stages).

Bytes (minimum for a fused stage kernel with the state in HBM): stage 1 reads U^n and writes U^1,
stage 2 reads U^1 and U^n and writes U^2, stage 3 reads U^2 and U^n and writes U^{n+1}: 8 N doubles
= 64 N bytes per cell-update.

Flops: +, -, *, /, sqrt count 1 each (an FMA is a multiply and an add, 2); min, max, compare,
select, negation and copies count 0. Counted by running the minimal algorithm below on an
op-counting scalar: per stage and cell one cons->prim; per face WENO5 (or MUSCL) of every
primitive on both sides, one prim->cons per side, one HLLC (upwind star state only) or HLL;
the flux differences, the alpha1 div(u*) term and the RK3 combination. Divisions and square roots
also cost extra instructions on the GPU; they are reported separately.
"""
import math

_COUNT = {"add": 0, "mul": 0, "div": 0, "sqrt": 0}


class Op:
    __slots__ = ("v",)

    def __init__(self, v):
        self.v = float(v)

    @staticmethod
    def _v(x):
        return x.v if isinstance(x, Op) else float(x)

    def __add__(self, o):
        _COUNT["add"] += 1
        return Op(self.v + Op._v(o))

    __radd__ = __add__

    def __sub__(self, o):
        _COUNT["add"] += 1
        return Op(self.v - Op._v(o))

    def __rsub__(self, o):
        _COUNT["add"] += 1
        return Op(Op._v(o) - self.v)

    def __mul__(self, o):
        _COUNT["mul"] += 1
        return Op(self.v * Op._v(o))

    __rmul__ = __mul__

    def __truediv__(self, o):
        _COUNT["div"] += 1
        return Op(self.v / Op._v(o))

    def __rtruediv__(self, o):
        _COUNT["div"] += 1
        return Op(Op._v(o) / self.v)

    def __neg__(self):
        return Op(-self.v)

    def __lt__(self, o):
        return self.v < Op._v(o)

    def __gt__(self, o):
        return self.v > Op._v(o)


def _sqrt(x):
    _COUNT["sqrt"] += 1
    return Op(math.sqrt(x.v))


def _min(a, b):
    return b if Op._v(b) < Op._v(a) else a


def _max(a, b):
    return b if Op._v(a) < Op._v(b) else a


class _Eos:
    def __init__(self, A, B):
        self.A, self.B = A, B


def _mix(e, a1):
    a2 = 1.0 - a1
    AP = a1 * e.A[0] + a2 * e.A[1]
    BP = a1 * e.B[0] + a2 * e.B[1]
    return AP, BP


def _cons2prim(e, U, dim):
    rho = U[0] + U[1]
    vel = [U[2 + m] / rho for m in range(dim)]
    ke = vel[0] * vel[0]
    for m in range(1, dim):
        ke = ke + vel[m] * vel[m]
    rho_e = U[2 + dim] - (0.5 * rho) * ke
    AP, BP = _mix(e, U[3 + dim])
    p = (rho_e - BP) / AP
    return [U[0], U[1]] + vel + [p, U[3 + dim]]


def _prim2cons(e, W, dim):
    rho = W[0] + W[1]
    mom = [rho * W[2 + m] for m in range(dim)]
    ke = W[2] * W[2]
    for m in range(1, dim):
        ke = ke + W[2 + m] * W[2 + m]
    AP, BP = _mix(e, W[3 + dim])
    rho_e = AP * W[2 + dim] + BP
    return [W[0], W[1]] + mom + [rho_e + (0.5 * rho) * ke, W[3 + dim]]


def _weno5(fmm, fm, fc, fp, fpp):
    f0 = fc + (2.0 / 3.0) * (fp - fc) - (1.0 / 6.0) * (fpp - fp)
    f1 = fc + (1.0 / 6.0) * (fc - fm) + (1.0 / 3.0) * (fp - fc)
    f2 = fc - (1.0 / 3.0) * (fm - fmm) + (5.0 / 6.0) * (fc - fm)
    d0, d1, d2 = 0.3, 0.6, 0.1
    a, b, c, ee = fp - fc, fpp - fp, fc - fm, fm - fmm
    IS0 = (10.0 / 3.0) * a * a - (11.0 / 3.0) * a * b + (4.0 / 3.0) * b * b
    IS1 = -(5.0 / 3.0) * c * a + (4.0 / 3.0) * c * c + (4.0 / 3.0) * a * a
    IS2 = (10.0 / 3.0) * c * c - (11.0 / 3.0) * ee * c + (4.0 / 3.0) * ee * ee
    IS01, IS02, IS12 = IS0 / IS1, IS0 / IS2, IS1 / IS2
    IS10, IS20, IS21 = 1.0 / IS01, 1.0 / IS02, 1.0 / IS12
    w0 = d0 / (d0 + (IS01 * IS01) * d1 + (IS02 * IS02) * d2)
    w1 = d1 / ((IS10 * IS10) * d0 + d1 + (IS12 * IS12) * d2)
    w2 = d2 / ((IS20 * IS20) * d0 + (IS21 * IS21) * d1 + d2)
    return w0 * f0 + w1 * f1 + w2 * f2


def _muscl(fm, fc, fp, dx, face):
    sm = (fc - fm) / dx
    sp = (fp - fc) / dx
    s = _min(_max(sm, 0.0), _max(sp, 0.0)) + _max(_min(sm, 0.0), _min(sp, 0.0))
    return fc + face * 0.5 * s * dx


def _sound(e, p, a1, rho):
    r1 = ((e.A[0] + 1.0) * p + e.B[0]) / e.A[0]
    r2 = ((e.A[1] + 1.0) * p + e.B[1]) / e.A[1]
    a2 = 1.0 - a1
    num = a1 * e.A[0] * r1 + a2 * e.A[1] * r2
    AP = a1 * e.A[0] + a2 * e.A[1]
    return _sqrt((num / AP) / rho)


def _hllc(e, UL, UR, WL, WR, dim, riemann):
    N = 4 + dim
    pL, pR = WL[2 + dim], WR[2 + dim]
    rhoL, rhoR = WL[0] + WL[1], WR[0] + WR[1]
    cL, cR = _sound(e, pL, WL[3 + dim], rhoL), _sound(e, pR, WR[3 + dim], rhoR)
    qL = WL[2] * WL[2]
    qR = WR[2] * WR[2]
    for m in range(1, dim):
        qL = qL + WL[2 + m] * WL[2 + m]
        qR = qR + WR[2 + m] * WR[2 + m]
    _max(cL + _sqrt(qL), cR + _sqrt(qR))
    VL, VR = WL[2], WR[2]
    if riemann == 2:
        V_avg, c_avg = 0.5 * (VL + VR), 0.5 * (cL + cR)
        SL = _min(V_avg - c_avg, VL - cL)
        SR = _max(V_avg + c_avg, VR + cR)
        S = (pR - pL + rhoL * VL * (SL - VL) - rhoR * VR * (SR - VR)) / (rhoL * (SL - VL) - rhoR * (SR - VR))
        Sb = _min(0.0, SL)
        q = (SL - VL) / (SL - S)
        F = [WL[0] * VL + Sb * (q * UL[0] - UL[0]), WL[1] * VL + Sb * (q * UL[1] - UL[1])]
        for m in range(dim):
            Fm = UL[2 + m] * VL
            Cs = q * UL[2 + m]
            if m == 0:
                Fm = Fm + pL
                Cs = Cs + q * rhoL * (S - VL)
            F.append(Fm + Sb * (Cs - UL[2 + m]))
        CsE = q * UL[2 + dim] + q * (S - VL) * (rhoL * S + pL / (SL - VL))
        F.append((UL[2 + dim] + pL) * VL + Sb * (CsE - UL[2 + dim]))
        F.append(WL[3 + dim] * VL + Sb * (q * UL[3 + dim] - UL[3 + dim]))
        un = VL + Sb * (q - 1.0)
        return F, un
    Sp = _max(_max(0.0, VL + cL), VR + cR)
    Sm = _min(_min(0.0, VL - cL), VR - cR)
    inv = 1.0 / (Sp - Sm)
    FL = [WL[0] * VL, WL[1] * VL] + [UL[2 + m] * VL for m in range(dim)] + [(UL[2 + dim] + pL) * VL, WL[3 + dim] * VL]
    FR = [WR[0] * VR, WR[1] * VR] + [UR[2 + m] * VR for m in range(dim)] + [(UR[2 + dim] + pR) * VR, WR[3 + dim] * VR]
    FL[2] = FL[2] + pL
    FR[2] = FR[2] + pR
    F = [(Sp * FL[n] - Sm * FR[n] + Sp * Sm * (UR[n] - UL[n])) * inv for n in range(N)]
    return F, (Sp * VL - Sm * VR) * inv


def count_cell_stage(dim, scheme="WENO5", riemann=2, stage=1):
    """Ops of one cell in one RK stage of the minimal fused algorithm (dict of add/mul/div/sqrt)."""
    for k in _COUNT:
        _COUNT[k] = 0
    e = _Eos([2.5, 1.6666], [0.0, 0.0])
    N = 4 + dim
    # a generic, non-degenerate state (no 0/0 branches, S* > 0)
    W = [Op(0.4), Op(3.0)] + [Op(0.3 + 0.1 * m) for m in range(dim)] + [Op(1.2), Op(0.4)]
    U = _prim2cons(e, W, dim)
    base = dict(_COUNT)
    for k in _COUNT:
        _COUNT[k] = 0
    _cons2prim(e, U, dim)                                          # once per cell
    for _ in range(dim):                                           # one face per direction
        WLf, WRf = [], []
        for n in range(N):
            v = [W[n] * Op(1.0 + 0.01 * s * s) for s in range(6)]
            for k in _COUNT:
                _COUNT[k] -= 0
            if scheme == "WENO5":
                WLf.append(_weno5(*v[0:5]))
                WRf.append(_weno5(*v[1:6]))
            else:
                WLf.append(_muscl(v[1], v[2], v[3], 0.1, 1.0))
                WRf.append(_muscl(v[2], v[3], v[4], 0.1, -1.0))
        UL, UR = _prim2cons(e, WLf, dim), _prim2cons(e, WRf, dim)
        _hllc(e, UL, UR, WLf, WRf, dim, riemann)
    # the stencil scaling above is test-data setup, not algorithm: remove its multiplies
    _COUNT["mul"] -= dim * N * 6
    for _ in range(N):                                             # flux differences
        acc = None
        for d in range(dim):
            t = (Op(1.0) - Op(0.5)) / Op(0.1)
            acc = t if acc is None else acc + t
    divu = None
    for d in range(dim):
        t = (Op(1.0) - Op(0.5)) / Op(0.1)
        divu = t if divu is None else divu + t
    U[3 + dim] * (1.0 + 0.0) * divu + Op(0.0)                     # alpha1 (1+K) div u*
    for n in range(N):                                             # RK combination
        if stage == 0:
            U[n] + Op(0.1) * Op(0.2)
        else:
            (Op(0.25) * U[n] + Op(0.75) * U[n]) + Op(0.1) * Op(0.2)
    out = dict(_COUNT)
    del base
    return out


def flops_per_cell_stage(params, scheme=None):
    if scheme is None:
        # WENO5B and THINC faces are counted as WENO5 faces (a lower bound of their work)
        scheme = "WENO5" if params.id_weno == 1 else "MUSCL"
    c = count_cell_stage(params.dim, scheme, params.id_riemann, stage=1)
    return c["add"] + c["mul"] + c["div"] + c["sqrt"]


def flops_per_cell_update(params, scheme=None):
    """Three stages (stage 1's RK combination has 3N fewer ops than stages 2 and 3)."""
    if scheme is None:
        # WENO5B and THINC faces are counted as WENO5 faces (a lower bound of their work)
        scheme = "WENO5" if params.id_weno == 1 else "MUSCL"
    tot = 0
    for s in range(3):
        c = count_cell_stage(params.dim, scheme, params.id_riemann, stage=s)
        tot += c["add"] + c["mul"] + c["div"] + c["sqrt"]
    return tot


def bytes_per_cell_update(params):
    return 64 * params.nstate


if __name__ == "__main__":
    for dim in (2, 3):
        for sch in ("WENO5", "MUSCL"):
            for r in (2, 1):
                c = count_cell_stage(dim, sch, r)
                print(dim, sch, "HLLC" if r == 2 else "HLL", c, "total/stage", sum(c.values()))
