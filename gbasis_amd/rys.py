"""Rys quadrature nodes and weights for the high angular momentum ERI kernel.

For T >= 0 and n = 1 .. RYS_NMAX the n-point Rys rule gives nodes x_r = t_r^2 in
(0, 1) and weights W_r with

    sum_r W_r x_r^m = F_m(T) = int_0^1 t^(2m) exp(-T t^2) dt,   m = 0 .. 2n - 1,

so that a primitive [e0|f0] of total angular momentum L <= 2n - 1 is a sum over n
nodes of products of three one-dimensional integrals (the Boys-function form of
gbasis/integrals/_two_elec_int_improved.py:87-196 with F_m replaced by the rule).

The device reads a table built once, offline, by ``build_table`` (run as
``python -m gbasis_amd.rys``), written to ``rys_table.npz`` next to this file:

* 0 <= T < RYS_TMAX: piecewise Chebyshev interpolants of degree RYS_DEG on
  intervals of width 1 in T, of every node and weight: the interpolants through the
  Chebyshev points, with values and coefficients computed in 80-digit arithmetic
  (mpmath) from the moments F_m(T) by the Chebyshev algorithm and the
  eigen-decomposition of the Jacobi matrix (Golub-Welsch), rounded to float64 last;
* T >= RYS_TMAX: the rule of the half-range Hermite weight, x_r = s_r^2 / T,
  W_r = w_r / sqrt(T), with s_r, w_r the positive nodes and weights of 2n-point
  Gauss-Hermite quadrature.  The neglected part of the integral beyond t = 1 is
  below 1e-20 relative for n <= 9 at T = 100.

Table layout (float64, flat): for n = 1 .. RYS_NMAX at offset RYS_OFF[n],
RYS_NSEG intervals of RYS_DEG + 1 Chebyshev coefficients of the n nodes followed by
those of the n weights ([seg][coef][node 0..n-1, weight 0..n-1]), then, at
RYS_HOFF[n], the n Hermite nodes s_r^2 and the n weights w_r.
"""

import os

import numpy as np

RYS_NMAX = 9          # nodes for L <= 16 (up to (gg|gg))
RYS_TMAX = 100.0
RYS_NSEG = 100        # intervals [k, k + 1) of T
RYS_DEG = 13
_NCOEF = RYS_DEG + 1

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rys_table.npz")


def _offsets():
    off, hoff = {}, {}
    pos = 0
    for n in range(1, RYS_NMAX + 1):
        off[n] = pos
        pos += RYS_NSEG * _NCOEF * 2 * n
    for n in range(1, RYS_NMAX + 1):
        hoff[n] = pos
        pos += 2 * n
    return off, hoff, pos


RYS_OFF, RYS_HOFF, RYS_SIZE = _offsets()


# ---------------------------------------------------------------- generation (mpmath)
def _moments_mp(T, nm):
    import mpmath as mp

    T = mp.mpf(T)
    half = mp.mpf(1) / 2
    if T == 0:
        return [mp.mpf(1) / (2 * m + 1) for m in range(nm)]
    return [mp.gammainc(m + half, 0, T) / (2 * T ** (m + half)) for m in range(nm)]


def rys_rule_mp(n, T, dps=80):
    """Nodes x_r and weights W_r (increasing x) of the n-point rule at T, as mpmath
    numbers: ordinary-moment Chebyshev algorithm for the three-term recurrence of the
    monic orthogonal polynomials in x, then the Jacobi matrix's eigen-decomposition."""
    import mpmath as mp

    with mp.workdps(dps):
        mu = _moments_mp(T, 2 * n)
        a = [mu[1] / mu[0]]
        b = [mu[0]]
        sig_m1 = [mp.mpf(0)] * (2 * n)
        sig = list(mu)
        for k in range(1, n):
            new = [mp.mpf(0)] * (2 * n)
            for l in range(k, 2 * n - k):
                new[l] = sig[l + 1] - a[k - 1] * sig[l] - b[k - 1] * sig_m1[l]
            a.append(new[k + 1] / new[k] - sig[k] / sig[k - 1])
            b.append(new[k] / sig[k - 1])
            sig_m1, sig = sig, new
        J = mp.matrix(n, n)
        for i in range(n):
            J[i, i] = a[i]
            if i + 1 < n:
                J[i, i + 1] = J[i + 1, i] = mp.sqrt(b[i + 1])
        E, Q = mp.eigsy(J)
        pairs = sorted((E[i], mu[0] * Q[0, i] ** 2) for i in range(n))
        return [p[0] for p in pairs], [p[1] for p in pairs]


def _hermite_half(n, dps=80):
    """Positive nodes s_r (as s_r^2) and weights w_r of 2n-point Gauss-Hermite."""
    import mpmath as mp

    with mp.workdps(dps):
        s0, _ = np.polynomial.hermite.hermgauss(2 * n)
        out = []
        for x0 in s0[s0 > 0]:
            x = mp.mpf(float(x0))
            for _ in range(60):
                x -= mp.hermite(2 * n, x) / (4 * n * mp.hermite(2 * n - 1, x))
            w = (mp.mpf(2) ** (2 * n - 1) * mp.factorial(2 * n) * mp.sqrt(mp.pi)
                 / ((2 * n) ** 2 * mp.hermite(2 * n - 1, x) ** 2))
            out.append((x * x, w))
        out.sort()
        return [float(p[0]) for p in out], [float(p[1]) for p in out]


def _fit_segment(args):
    """Chebyshev interpolant of degree RYS_DEG of the nodes and weights on [k, k + 1],
    its coefficients computed in 80-digit arithmetic from the values at the Chebyshev
    points (discrete cosine transform) and only then rounded to float64."""
    import mpmath as mp

    n, k = args
    N = _NCOEF
    with mp.workdps(80):
        theta = [mp.pi * (i + mp.mpf(1) / 2) / N for i in range(N)]
        vals = []
        for th in theta:
            x, w = rys_rule_mp(n, mp.mpf(k) + (1 + mp.cos(th)) / 2)
            vals.append(list(x) + list(w))
        coef = np.zeros((N, 2 * n))
        for j in range(N):
            cj = [mp.fsum(vals[i][r] * mp.cos(j * theta[i]) for i in range(N)) * 2 / N
                  for r in range(2 * n)]
            if j == 0:
                cj = [v / 2 for v in cj]
            coef[j] = [float(v) for v in cj]
    return n, k, coef   # (ncoef, 2n), in s = 2 (T - k) - 1


def build_table(processes=None):
    """The flat table described in the module docstring (takes a few minutes)."""
    from multiprocessing import Pool

    table = np.zeros(RYS_SIZE)
    jobs = [(n, k) for n in range(1, RYS_NMAX + 1) for k in range(RYS_NSEG)]
    with Pool(processes) as pool:
        for n, k, c in pool.imap_unordered(_fit_segment, jobs, chunksize=8):
            base = RYS_OFF[n] + k * _NCOEF * 2 * n
            table[base:base + _NCOEF * 2 * n] = c.reshape(-1)
    for n in range(1, RYS_NMAX + 1):
        s2, w = _hermite_half(n)
        table[RYS_HOFF[n]:RYS_HOFF[n] + n] = s2
        table[RYS_HOFF[n] + n:RYS_HOFF[n] + 2 * n] = w
    return table


# ---------------------------------------------------------------- use
_TABLE = None


def rys_table():
    """The flat float64 table (loaded once from rys_table.npz)."""
    global _TABLE
    if _TABLE is None:
        if not os.path.exists(TABLE_PATH):
            raise RuntimeError(f"{TABLE_PATH} missing: build it with python -m gbasis_amd.rys")
        with np.load(TABLE_PATH, allow_pickle=False) as z:
            t = np.ascontiguousarray(z["table"], dtype=np.float64)
            meta = z["meta"].tolist()
        if meta != [RYS_NMAX, RYS_NSEG, RYS_DEG, RYS_TMAX] or t.size != RYS_SIZE:
            raise RuntimeError(f"{TABLE_PATH} does not match this module's layout")
        _TABLE = t
    return _TABLE


def rys_rule(n, T):
    """Nodes and weights, each (len(T), n), of the n-point rule at the T values, from
    the table (the device kernel's evaluation, in NumPy)."""
    tab = rys_table()
    T = np.atleast_1d(np.asarray(T, dtype=np.float64))
    x = np.zeros((T.size, n))
    w = np.zeros((T.size, n))
    far = T >= RYS_TMAX
    if far.any():
        s2 = tab[RYS_HOFF[n]:RYS_HOFF[n] + n]
        hw = tab[RYS_HOFF[n] + n:RYS_HOFF[n] + 2 * n]
        x[far] = s2[None, :] / T[far, None]
        w[far] = hw[None, :] / np.sqrt(T[far, None])
    near = ~far
    if near.any():
        k = np.minimum(np.floor(T[near]).astype(np.int64), RYS_NSEG - 1)
        s = 2.0 * (T[near] - k) - 1.0
        c = tab[RYS_OFF[n]:RYS_OFF[n] + RYS_NSEG * _NCOEF * 2 * n].reshape(RYS_NSEG, _NCOEF,
                                                                           2 * n)[k]
        # Clenshaw recurrence, as the kernel evaluates it
        s = s[:, None]
        b1 = np.zeros((s.shape[0], 2 * n))
        b2 = np.zeros_like(b1)
        for j in range(RYS_DEG, 0, -1):
            b1, b2 = c[:, j, :] + 2.0 * s * b1 - b2, b1
        v = c[:, 0, :] + s * b1 - b2
        x[near] = v[:, :n]
        w[near] = v[:, n:]
    return x, w


if __name__ == "__main__":
    import time

    t0 = time.time()
    tab = build_table()
    np.savez_compressed(TABLE_PATH, table=tab,
                        meta=np.array([RYS_NMAX, RYS_NSEG, RYS_DEG, RYS_TMAX], dtype=np.float64))
    print(f"wrote {TABLE_PATH} ({tab.size} values) in {time.time() - t0:.0f} s")
