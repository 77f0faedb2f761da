"""Rys quadrature rule of the high angular momentum kernel (CPU): the shipped table
against an independent 40-digit evaluation of the Boys moments, and the Rys form of
the primitive part, replayed in NumPy exactly as the kernel evaluates it, against the
binary128 oracle on shells with exponents 0.02 .. 1e5."""
import numpy as np
import pytest

from gbasis_amd.programs import comps
from gbasis_amd.rys import RYS_NMAX, RYS_TMAX, rys_rule
from oracle.hp import hp_block_raw
from tests.test_subpairs_cpu import _wide_basis, replay_quartet


def _boys_mp(m, T):
    import mpmath as mp

    with mp.workdps(40):
        T = mp.mpf(float(T))
        if T == 0:
            return float(mp.mpf(1) / (2 * m + 1))
        return float(mp.gammainc(m + mp.mpf(1) / 2, 0, T) / (2 * T ** (m + mp.mpf(1) / 2)))


def test_rule_reproduces_boys_moments():
    """sum_r W_r x_r^m = F_m(T) for m < 2n, n = 1..9, over the fitted range, its ends
    and the asymptotic range (T up to 1e5): within 4e-15 relative (the m-th power of a
    node carries m times its rounding)."""
    rng = np.random.default_rng(3)
    Ts = np.concatenate([[0.0, 1e-13, 0.999999, 1.0, 37.5, RYS_TMAX - 1e-9, RYS_TMAX,
                          RYS_TMAX + 0.3, 400.0, 1e5], rng.uniform(0, RYS_TMAX, 12)])
    worst = 0.0
    for n in range(1, RYS_NMAX + 1):
        x, w = rys_rule(n, Ts)
        assert np.all((x > 0) & (x < 1)) and np.all(w > 0)
        assert np.all(np.diff(x, axis=1) > 0)
        for i, T in enumerate(Ts):
            for m in range(2 * n):
                worst = max(worst, abs(np.sum(w[i] * x[i] ** m) / _boys_mp(m, T) - 1))
    assert worst <= 4e-15, worst


def rys_prim_acc(prog, rb, rk):
    """Contracted [e0|f0] of one entry quartet from its primitive records by Rys
    quadrature, in the class program's accumulator order: per node u (x = t^2) the
    two-index integrals of each Cartesian direction,
        I(e+1, 0) = C00 I(e, 0) + e B10 I(e-1, 0),
        I(e, f+1) = D00 I(e, f) + f B01 I(e, f-1) + e B00 I(e-1, f),  I(0, 0) = 1,
    C00 = PA + u WP, D00 = QC + u WQ, B10 = (1 - u rho/zeta)/(2 zeta),
    B01 = (1 - u rho/eta)/(2 eta), B00 = u/(2(zeta + eta)), then
    [e0|f0] = sum_prim pref sum_r W_r Ix Iy Iz."""
    la, lb, lc, ld = prog.cls
    Lab, Lcd = la + lb, lc + ld
    n = prog.L // 2 + 1
    acc = np.zeros(prog.nacc)
    if rb.shape[0] == 0:
        return acc
    zeta, eta = rb[:, 0], rk[:, 0]
    inv = 1 / (zeta + eta)
    PQ = rb[:, 2:5] - rk[:, 2:5]
    T = zeta * eta * inv * np.sum(PQ * PQ, axis=1)
    pref = rb[:, 8] * rk[:, 8] * np.sqrt(inv)
    roz, roe = eta * inv, zeta * inv
    WP, WQ = -roz[:, None] * PQ, roe[:, None] * PQ
    x, w = rys_rule(n, T)                                   # (P, n)
    u = x[:, :, None]                                       # (P, n, 1) per direction
    C00 = rb[:, None, 5:8] + u * WP[:, None, :]
    D00 = rk[:, None, 5:8] + u * WQ[:, None, :]
    B10 = (rb[:, 1, None] * (1 - x * roz[:, None]))[:, :, None]
    B01 = (rk[:, 1, None] * (1 - x * roe[:, None]))[:, :, None]
    B00 = (x * (0.5 * inv)[:, None])[:, :, None]
    I = np.zeros((Lab + 1, Lcd + 1) + C00.shape)
    I[0, 0] = 1.0
    for e in range(Lab):
        I[e + 1, 0] = C00 * I[e, 0] + (e * B10 * I[e - 1, 0] if e else 0.0)
    for f in range(Lcd):
        for e in range(Lab + 1):
            v = D00 * I[e, f]
            if f:
                v = v + f * B01 * I[e, f - 1]
            if e:
                v = v + e * B00 * I[e - 1, f]
            I[e, f + 1] = v
    wt = w * pref[:, None]
    f = 0
    for nc in range(lc, Lcd + 1):
        for c in comps(nc):
            for na in range(la, Lab + 1):
                for a in comps(na):
                    acc[f] = np.sum(wt * I[a[0], c[0], :, :, 0] * I[a[1], c[1], :, :, 1]
                                    * I[a[2], c[2], :, :, 2])
                    f += 1
    assert f == prog.nacc
    return acc


@pytest.mark.parametrize("ls", [[1, 1, 0, 0], [2, 1, 2, 1], [3, 2, 1, 0], [3, 3, 2, 2],
                                [4, 2, 3, 1], [4, 4, 2, 1]])
def test_rys_primitive_part_vs_binary128(ls):
    """Shell quartets with exponents 0.02 .. 1e5 evaluated from the split device tables
    with the Rys primitive part and the kernels' horizontal recursion: within 5e-13 of
    the largest value of the binary128 oracle's block, a fifth of the 1e-12 bar (the
    worst case, (ff|dd), reaches 2.1e-13 where the Obara-Saika replay reaches 4.5e-14:
    a node's one-ulp rounding enters its powers up to L and the transfer of the
    horizontal recursion)."""
    from gbasis_amd.tables import BatchTables, MoleculeTables

    basis = _wide_basis(17 + sum(ls), ls)
    m = MoleculeTables(basis)
    b = BatchTables([m])
    bra, ket = m.quartets()
    worst = 0.0
    for p, q in list(zip(bra.tolist(), ket.tolist()))[::5]:
        got = replay_quartet(b, p, q, prim_acc=rys_prim_acc)
        sh = [m.shells[int(i)] for i in (*b.pair_info[p, :2], *b.pair_info[q, :2])]
        ref = hp_block_raw([s.l for s in sh], [s.coord for s in sh], [s.exps for s in sh],
                           [s.coeffs for s in sh], "quad", 0)
        worst = max(worst, np.abs(got - ref).max() / max(1.0, np.abs(ref).max()))
    assert worst <= 5e-13, worst


@pytest.mark.parametrize("cls", [(4, 4, 4, 4), (4, 4, 4, 3), (3, 3, 2, 2), (4, 0, 4, 0),
                                 (2, 4, 3, 3), (4, 1, 3, 0)])
def test_rys_lane_groups(cls):
    """Lane slots of the Rys pass: every accumulator of the class program exactly once, in
    a group whose members share (a_x, a_y, c_x, c_y); the groups of a wavefront have one
    shape and its first lane is occupied; whole 512-lane blocks."""
    from types import SimpleNamespace

    from gbasis_amd.engine import EriEngine
    from gbasis_amd.programs import class_program

    stub = SimpleNamespace(_rys_idx={}, device="cpu")
    arr, n = EriEngine._rys_groups(stub, cls)
    a = arr.numpy()
    la, lb, lc, ld = cls
    W, MA, MC = lc + ld + 1, lb + 1, ld + 1
    desc, rows = a[:n], a[n:].reshape(n, MA * MC)
    assert n % 512 == 0
    prog = class_program(*cls)
    final = [(c, x) for nc in range(lc, lc + ld + 1) for c in comps(nc)
             for na in range(la, la + lb + 1) for x in comps(na)]
    seen = []
    for w in range(0, n, 64):
        d = desc[w:w + 64]
        ok = d >= 0
        if ok.any():
            assert d[0] >= 0
            assert len({(int(x) >> 18) & 15 for x in d[ok]}) == 1
            assert len({(int(x) >> 26) & 15 for x in d[ok]}) == 1
    for g, r in zip(desc.tolist(), rows.tolist()):
        if g < 0:
            assert all(f < 0 for f in r)
            continue
        ox, oy = g & 127, (g >> 7) & 127
        az0, alen, cz0, clen = (g >> 14) & 15, (g >> 18) & 15, (g >> 22) & 15, (g >> 26) & 15
        for i in range(MA):
            for j in range(MC):
                f = r[i * MC + j]
                if i < alen and j < clen:
                    c, x = final[f]
                    assert (x[0] * W + c[0], x[1] * W + c[1]) == (ox, oy)
                    assert (x[2], c[2]) == (az0 + i, cz0 + j)
                    seen.append(f)
                else:
                    assert f < 0
    assert sorted(seen) == list(range(prog.nacc))
