"""TEST INFRASTRUCTURE ONLY -- scalar restatement of the reference pair kernel.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may
import this module.

This is a second, independent restatement of
``solvers/NBody/extra/hydrodynamicKernels.cuh`` that follows the reference's
*vector-application* structure (``dotProduct_*`` applying coefficients to the
source vector), whereas ``rpy_oracle.c`` builds explicit 3x3 matrices.  The two
are checked against each other on random pairs (tests/test_oracle_golden.py),
which pins both.

The functions are generic over the scalar type, so running them with
:class:`Flop` counts the arithmetic of the reference formulas (the frozen flop
model: +, -, *, /, sqrt each count 1, an FMA therefore counts 2; selects and
branches are not counted).  Positions are pre-scaled by the hydrodynamic
radius and the mobility constants t0/rt0/r0 are applied once per target, as
both are O(N) work and not part of the per-pair cost.
"""
import math


class Flop:
    """Float that counts arithmetic operations into a shared tally."""

    tally = {"ops": 0}

    __slots__ = ("v",)

    def __init__(self, v):
        self.v = float(v.v if isinstance(v, Flop) else v)

    @staticmethod
    def _val(x):
        return x.v if isinstance(x, Flop) else float(x)

    def _op(self, val):
        Flop.tally["ops"] += 1
        return Flop(val)

    def __add__(self, o): return self._op(self.v + Flop._val(o))
    def __radd__(self, o): return self._op(Flop._val(o) + self.v)
    def __sub__(self, o): return self._op(self.v - Flop._val(o))
    def __rsub__(self, o): return self._op(Flop._val(o) - self.v)
    def __mul__(self, o): return self._op(self.v * Flop._val(o))
    def __rmul__(self, o): return self._op(Flop._val(o) * self.v)
    def __truediv__(self, o): return self._op(self.v / Flop._val(o))
    def __rtruediv__(self, o): return self._op(Flop._val(o) / self.v)
    def __neg__(self): return Flop(-self.v)
    def __lt__(self, o): return self.v < Flop._val(o)
    def __le__(self, o): return self.v <= Flop._val(o)
    def __gt__(self, o): return self.v > Flop._val(o)
    def __ge__(self, o): return self.v >= Flop._val(o)
    def __eq__(self, o): return self.v == Flop._val(o)
    def __hash__(self): return hash(self.v)
    def __float__(self): return self.v


def _sqrt(x):
    if isinstance(x, Flop):
        return x._op(math.sqrt(x.v))
    return math.sqrt(x)


def _max(x, y):
    return x if x >= y else y


def _min(x, y):
    return x if x <= y else y


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _zero():
    return [0.0, 0.0, 0.0]


# --- open boundary, hydrodynamicKernels.cuh:41-118 (scaled units) ---------

def rpy_uf(r):
    """(f, g/r^2), hydrodynamicKernels.cuh:41-57."""
    if r >= 2.0:
        invr = 1.0 / r
        invr2 = invr * invr
        f = (0.75 + 0.5 * invr2) * invr
        g = (0.75 - 1.5 * invr2) * invr * invr2
        return f, g
    f = 1.0 - 0.28125 * r
    g = 0.09375 / r if r > 0.0 else 0.0
    return f, g


def rpy_wt(r):
    """(f, g/r^2), hydrodynamicKernels.cuh:63-86."""
    if r >= 2.0:
        invr = 1.0 / r
        invr2 = invr * invr
        invr3 = invr * invr2
        return -0.5 * invr3, 1.5 * invr2 * invr3
    r3 = r * r * r
    f = 1.0 - 0.84375 * r + 0.078125 * r3
    g = (0.28125 / r - 0.046875 * r) if r > 0.0 else 0.0
    return f, g


def rpy_rt(d, r):
    """(M_xy, M_xz, M_yz), hydrodynamicKernels.cuh:89-118."""
    if r >= 2.0:
        c = 1.0 / (r * r * r)
    else:
        c = 0.5 * (1.0 - 0.375 * r)
    return [d[2] * c, -(d[1] * c), d[0] * c]


def _apply_sym(coef, d, r, v):
    f, g = coef
    gv = g * _dot(d, v)
    out = [f * v[0], f * v[1], f * v[2]]
    if r > 0.0:
        out = [out[k] + gv * d[k] for k in range(3)]
    return out


def _apply_rt(m, v):
    """hydrodynamicKernels.cuh:163-164 / :170-171."""
    return [m[0] * v[1] + m[1] * v[2], -(m[0] * v[0]) + m[2] * v[2], -(m[1] * v[0]) - m[2] * v[1]]


# --- bottom wall corrections, hydrodynamicKernels.cuh:187-359 -------------

def wall_uf(R, self_, h, v):
    c = _zero()
    if self_:
        iz = 1.0 / h
        iz3 = iz * iz * iz
        iz5 = iz3 * iz * iz
        c[0] = -(v[0] * (9.0 * iz - 2.0 * iz3 + iz5) * 0.0625)
        c[1] = -(v[1] * (9.0 * iz - 2.0 * iz3 + iz5) * 0.0625)
        c[2] = -(v[2] * (9.0 * iz - 4.0 * iz3 + iz5) * 0.125)
        return c
    hh = h / R[2]
    invR = 1.0 / _sqrt(_dot(R, R))
    e = [R[0] * invR, R[1] * invR, R[2] * invR]
    invR3 = invR * invR * invR
    invR5 = invR3 * invR * invR
    f1 = -((3.0 * (1.0 + 2.0 * hh * (1.0 - hh) * e[2] * e[2]) * invR
            + 2.0 * (1.0 - 3.0 * e[2] * e[2]) * invR3
            - 2.0 * (1.0 - 5.0 * e[2] * e[2]) * invR5) * 0.25)
    f2 = -((3.0 * (1.0 - 6.0 * hh * (1.0 - hh) * e[2] * e[2]) * invR
            - 6.0 * (1.0 - 5.0 * e[2] * e[2]) * invR3
            + 10.0 * (1.0 - 7.0 * e[2] * e[2]) * invR5) * 0.25)
    f3 = e[2] * (3.0 * hh * (1.0 - 6.0 * (1.0 - hh) * e[2] * e[2]) * invR
                 - 6.0 * (1.0 - 5.0 * e[2] * e[2]) * invR3
                 + 10.0 * (2.0 - 7.0 * e[2] * e[2]) * invR5) * 0.5
    f4 = e[2] * (3.0 * hh * invR - 10.0 * invR5) * 0.5
    f5 = -(3.0 * hh * hh * e[2] * e[2] * invR + 3.0 * e[2] * e[2] * invR3
           + (2.0 - 15.0 * e[2] * e[2]) * invR5)
    c[0] = (f1 + f2 * e[0] * e[0]) * v[0] + (f2 * e[0] * e[1]) * v[1] + (f2 * e[0] * e[2] + f3 * e[0]) * v[2]
    c[1] = (f2 * e[1] * e[0]) * v[0] + (f1 + f2 * e[1] * e[1]) * v[1] + (f2 * e[1] * e[2] + f3 * e[1]) * v[2]
    c[2] = ((f2 * e[2] * e[0] + f4 * e[0]) * v[0] + (f2 * e[2] * e[1] + f4 * e[1]) * v[1]
            + (f1 + f2 * e[2] * e[2] + f3 * e[2] + f4 * e[2] + f5) * v[2])
    return c


def wall_wt(R, self_, h, v):
    c = _zero()
    if self_:
        iz = 1.0 / h
        iz3 = iz * iz * iz
        return [v[0] * -(iz3 * 0.3125), v[1] * -(iz3 * 0.3125), v[2] * -(iz3 * 0.125)]
    invR = 1.0 / _sqrt(_dot(R, R))
    invR3 = invR * invR * invR
    e = [R[0] * invR, R[1] * invR, R[2] * invR]
    f1 = (1.0 - 6.0 * e[2] * e[2]) * invR3 * 0.5
    f2 = -(invR3 * 1.5)
    f3 = 3.0 * invR3 * e[2]
    f4 = 3.0 * invR3
    c[0] = (f1 + f2 * e[0] * e[0] + f4 * e[1] * e[1]) * v[0] + ((f2 - f4) * e[0] * e[1]) * v[1] + (f2 * e[0] * e[2]) * v[2]
    c[1] = ((f2 - f4) * e[0] * e[1]) * v[0] + (f1 + f2 * e[1] * e[1] + f4 * e[0] * e[0]) * v[1] + (f2 * e[1] * e[2]) * v[2]
    c[2] = (f2 * e[2] * e[0] + f3 * e[0]) * v[0] + (f2 * e[2] * e[1] + f3 * e[1]) * v[1] + (f1 + f2 * e[2] * e[2] + f3 * e[2]) * v[2]
    return c


def _rt_factors(R, h):
    hh = h / R[2]
    invR = 1.0 / _sqrt(_dot(R, R))
    invR2 = invR * invR
    invR4 = invR2 * invR2
    e = [R[0] * invR, R[1] * invR, R[2] * invR]
    f1 = invR2
    f2 = (6.0 * hh * e[2] * e[2] * invR2 + (1.0 - 10.0 * e[2] * e[2]) * invR4) * 2.0
    f3 = -(e[2] * (3.0 * hh * invR2 - 5.0 * invR4) * 2.0)
    f4 = -(e[2] * (hh * invR2 - invR4) * 2.0)
    return e, f1, f2, f3, f4


def wall_ut(R, self_, h, v):
    c = _zero()
    if self_:
        iz = 1.0 / h
        iz4 = iz * iz * iz * iz
        return [(iz4 * 0.125) * v[1], -(iz4 * 0.125) * v[0], 0.0]
    e, f1, f2, f3, f4 = _rt_factors(R, h)
    c[0] = 0.0 - ((-(f3 * e[0] * e[1])) * v[0] + (-(f1 * e[2]) + f3 * e[0] * e[0] - f4) * v[1] + (f1 * e[1]) * v[2])
    c[1] = 0.0 - ((f1 * e[2] - f3 * e[1] * e[1] + f4) * v[0] + (f3 * e[0] * e[1]) * v[1] + (-(f1 * e[0])) * v[2])
    c[2] = 0.0 - ((-(f1 * e[1]) - f2 * e[1] - f3 * e[1] * e[2]) * v[0] + (f1 * e[0] + f2 * e[0] + f3 * e[0] * e[2]) * v[1])
    return c


def wall_wf(R, self_, h, v):
    c = _zero()
    if self_:
        iz = 1.0 / h
        iz4 = iz * iz * iz * iz
        return [-(iz4 * 0.125) * v[1], (iz4 * 0.125) * v[0], 0.0]
    e, f1, f2, f3, f4 = _rt_factors(R, h)
    c[0] = 0.0 - ((-(f3 * e[0] * e[1])) * v[0] + (f1 * e[2] - f3 * e[1] * e[1] + f4) * v[1]
                  + (-(f1 * e[1]) - f2 * e[1] - f3 * e[1] * e[2]) * v[2])
    c[1] = 0.0 - ((-(f1 * e[2]) + f3 * e[0] * e[0] - f4) * v[0] + (f3 * e[0] * e[1]) * v[1]
                  + (f1 * e[0] + f2 * e[0] + f3 * e[0] * e[2]) * v[2])
    c[2] = 0.0 - ((f1 * e[1]) * v[0] + (-(f1 * e[0])) * v[1])
    return c


def pair_apply(wall, pi, pj, fj, tj, blocks):
    """Dimensionless contributions of source j to target i.

    Follows OpenBoundary::dotProduct (:124-140) or BottomWall::dotProduct
    (:362-390); positions already divided by a.  ``blocks`` is a subset of
    {"UF", "UT", "WT", "WF"}.  Returns dict block -> 3-vector (before the
    t0 / rt0 / r0 factors).
    """
    pi = list(pi)
    pj = list(pj)
    bij = 1.0
    if wall:
        bi = _max(_min(pi[2], 1.0), 0.0)
        bj = _max(_min(pj[2], 1.0), 0.0)
        bij = bi * bj
        pi[2] = _max(pi[2], 1.0)
        pj[2] = _max(pj[2], 1.0)
    d = [pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]]
    r = _sqrt(_dot(d, d))
    out = {}
    if "UF" in blocks:
        u = _apply_sym(rpy_uf(r), d, r, fj)
        if wall:
            R = [d[0], d[1], d[2] + 2.0 * pj[2]]
            cc = wall_uf(R, r == 0.0, pj[2], fj)
            u = [u[k] + cc[k] for k in range(3)]
        out["UF"] = u
    if "UT" in blocks:
        u = _apply_rt(rpy_rt(d, r), tj)
        if wall:
            R = [-d[0], -d[1], -d[2] + 2.0 * pi[2]]
            cc = wall_ut(R, r == 0.0, pi[2], tj)
            u = [u[k] + cc[k] for k in range(3)]
        out["UT"] = u
    if "WT" in blocks:
        w = _apply_sym(rpy_wt(r), d, r, tj)
        if wall:
            R = [d[0], d[1], d[2] + 2.0 * pj[2]]
            cc = wall_wt(R, r == 0.0, pj[2], tj)
            w = [w[k] + cc[k] for k in range(3)]
        out["WT"] = w
    if "WF" in blocks:
        w = _apply_rt(rpy_rt(d, r), fj)
        if wall:
            R = [d[0], d[1], d[2] + 2.0 * pj[2]]
            cc = wall_wf(R, r == 0.0, pj[2], fj)
            w = [w[k] + cc[k] for k in range(3)]
        out["WF"] = w
    if wall:
        out = {k: [bij * x for x in v] for k, v in out.items()}
    return out


def count_pair_flops(wall, blocks):
    """Flops of one far-field (r >= 2a), non-self pair evaluation plus the
    accumulation of the result into the per-target sums."""
    pi = [Flop(0.3), Flop(0.1), Flop(6.0)]
    pj = [Flop(3.1), Flop(-0.4), Flop(4.5)]
    fj = [Flop(0.2), Flop(-1.0), Flop(0.7)]
    tj = [Flop(-0.3), Flop(0.5), Flop(0.9)]
    acc = {k: [Flop(0.0)] * 3 for k in blocks}
    Flop.tally["ops"] = 0
    res = pair_apply(wall, pi, pj, fj, tj, blocks)
    for k in blocks:
        acc[k] = [acc[k][c] + res[k][c] for c in range(3)]
    return Flop.tally["ops"]
