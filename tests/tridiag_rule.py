"""numpy restatement of the device's tridiagonal square root, for tests.

The Lanczos solver needs s = T^{1/2} e1 for its m x m tridiagonal T
(the reference calls host LAPACKE ?steqr + cblas ?gemv,
include/lanczos/LanczosAlgorithm.cu:48-85).  The HIP kernel
(libmobility_amd/csrc/lanczos_kernels.hip, tridiag_sqrt_body) evaluates
    T^{1/2} e1 = (2/pi) Int_R e^u T (e^{2u} I + T)^{-1} e1 du
on sinh-mapped trapezoid nodes; this module follows it step for step so that
its accuracy over condition numbers can be measured on the CPU:
  * spectrum bounds: Gershgorin above; below, one round of Sturm counts at
    511 points log-spaced from 1e-16 of the largest up to it, which brackets
    the smallest eigenvalue within 7.5% (or puts it on that floor);
  * nodes u_k = c + s sinh(t_k), t_k = (k - half) h, |t_k| <= 4, with
    s = 1.2 + 0.2 log10(kappa), h = 0.18 / (1 + 0.3 log10(kappa)),
    c = (ln lmin + ln lmax) / 4, weights (2h/pi) e^u s cosh(t);
  * per node, (sig I + T) x = e1 by UL elimination and T x formed as
    e1 - sig x below sqrt(lmin lmax), as the tridiagonal product above.
"""
import numpy as np

NODES = 512


def _sturm(a, b2, x, pivmin):
    cnt = 0
    q = a[0] - x
    if abs(q) < pivmin:
        q = -pivmin
    cnt += q < 0
    for i in range(1, len(a)):
        q = (a[i] - x) - b2[i - 1] / q
        if abs(q) < pivmin:
            q = -pivmin
        cnt += q < 0
    return cnt


def bounds(a, b):
    """(lmin, lmax) as the kernel brackets them: Gershgorin above; one round of
    Sturm counts at x_t = lmax 1e-16^{1 - t/511}, t = 1..511 (log-spaced 7.5%
    apart); lmin = the last point below the smallest eigenvalue, or the floor
    1e-16 lmax."""
    m = len(a)
    bb = np.zeros(m)
    bb[:m - 1] = b
    b2 = bb * bb
    r = np.abs(np.concatenate([[0.0], bb[:m - 1]])) + np.abs(bb)
    hi = np.max(a + r)
    hi = hi + 1e-12 * abs(hi) + 1e-300
    pivmin = 1e-290 * max(1.0, b2.max())
    floor = 1e-16 * hi
    f = NODES
    for t in range(1, NODES):
        x = hi * np.exp(np.log(1e-16) * (1.0 - t / (NODES - 1)))
        if _sturm(a, b2, x, pivmin) >= 1:
            f = t
            break
    if f <= 1 or f >= NODES:
        return floor, hi
    return max(floor, hi * np.exp(np.log(1e-16) * (1.0 - (f - 1) / (NODES - 1)))), hi


def nodes(lmin, lmax):
    """Shifts sig_k = e^{2 u_k} and weights of the quadrature."""
    lk = np.log10(max(lmax / lmin, 1.0))
    s = 1.2 + 0.2 * lk
    h = 0.18 / (1.0 + 0.3 * lk)
    half = int(np.floor(4.0 / h))
    if 2 * half + 1 > NODES:
        half = (NODES - 1) // 2
        h = 4.0 / half
    tk = (np.arange(2 * half + 1) - half) * h
    u = 0.25 * (np.log(lmin) + np.log(lmax)) + s * np.sinh(tk)
    return np.exp(2 * u), (2 * h / np.pi) * np.exp(u) * s * np.cosh(tk)


def sqrt_e1(a, b):
    """T^{1/2} e1 for T = tridiag(b, a, b) as the kernel computes it."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    m = len(a)
    lmin, lmax = bounds(a, b)
    sig, w = nodes(lmin, lmax)
    tau = np.sqrt(lmin * lmax)
    out = np.zeros(m)
    for sg, wg in zip(sig, w):
        q = np.zeros(m)
        rg = 1.0 / (sg + a[m - 1])
        if m > 1:
            q[m - 1] = -b[m - 2] * rg
        for i in range(m - 2, -1, -1):
            rg = 1.0 / (sg + a[i] - b[i] * b[i] * rg)
            if i >= 1:
                q[i] = -b[i - 1] * rg
        x = np.empty(m)
        x[0] = rg
        for i in range(1, m):
            x[i] = x[i - 1] * q[i]
        if sg < tau:
            y = -sg * x
            y[0] += 1.0
        else:
            y = a * x
            y[1:] += b * x[:-1]
            y[:-1] += b * x[1:]
        out += wg * y
    return out
