"""NumPy restatement of the reference OS+HGP engine.  TEST INFRASTRUCTURE ONLY.

Follows gbasis/integrals/_two_elec_int_improved.py:591-732
(``compute_two_electron_integrals_os_hgp``) stage by stage:

 1. primitive-quartet geometry and prefactor          (:651-690)
 2. Boys F_m(T) * prefactor, m = 0..L                  (:696-697)
 3. vertical recursion on centre A                     (:87-196)
 4. electron transfer from A to C                      (:199-361)
 5. contraction with (2a/pi)^(3/4) (4a)^(l/2) c_p      (:44-84)
 6. horizontal recursion on D, then on B               (:364-570)
 7. factorial2 component norms                         (:572-586, :25-41)

Intermediates are held per Cartesian component (dicts keyed by exponent
tuples) with the primitive quartets flattened on the last axis, instead of the
reference's over-allocated cubes; the recursions, their direction choices
(z before y before x, matching the reference's x-, y-, z-sweep order) and the
arithmetic are the same.
"""

import numpy as np

from .boys import boys_hyp1f1

TWO_PI_2P5 = 2.0 * np.pi ** 2.5


def _comps(l):
    return [(x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1)]


def _dfact(n):
    out = 1
    while n > 1:
        out *= n
        n -= 2
    return out


def _step(comp, lvl_dir=None):
    """Recursion direction for a component: z if a_z>0, elif y if a_y>0, else x."""
    for i in (2, 1, 0):
        if comp[i] > 0:
            return i
    return None


def _minus(comp, i, k=1):
    c = list(comp)
    c[i] -= k
    return tuple(c)


def _plus(comp, i):
    c = list(comp)
    c[i] += 1
    return tuple(c)


def boys_potential(boys, orders, T, rho, potential="coulomb", omega=None, dtype=np.float64):
    """Boys values of the operator: F_m(T) for 1/r; for erf(omega r)/r
    G_m = s^(m+1/2) F_m(s T) with s = omega^2 / (omega^2 + rho), and F_m(T) - G_m for
    erfc(omega r)/r (gbasis/integrals/boys_functions.py:145-223)."""
    if potential == "coulomb":
        return boys(orders, T).astype(dtype)
    if potential not in ("erf", "erfc") or omega is None:
        raise ValueError("potential must be 'coulomb', or 'erf' / 'erfc' with omega")
    w2 = dtype(omega) * dtype(omega)
    s = w2 / (w2 + rho)
    g = s ** (np.asarray(orders, dtype=dtype) + dtype(0.5)) * boys(orders, s * T).astype(dtype)
    return g if potential == "erf" else boys(orders, T).astype(dtype) - g


def os_hgp_quartet(coord_a, la, comps_a, exps_a, coeffs_a,
                   coord_b, lb, comps_b, exps_b, coeffs_b,
                   coord_c, lc, comps_c, exps_c, coeffs_c,
                   coord_d, ld, comps_d, exps_d, coeffs_d,
                   boys=boys_hyp1f1, dtype=np.float64, ket="etr", potential="coulomb",
                   omega=None):
    """Contracted Cartesian ERI block in chemists' notation.

    Returns an array of shape ``(n_a, n_b, n_c, n_d, M_a, M_b, M_c, M_d)`` with
    the components in the orders given by ``comps_*`` (as the reference engine,
    ``_two_elec_int_improved.py:591-732``); no ``norm_cont`` is applied.
    """
    A, B, C, D = (np.asarray(x, dtype=dtype) for x in (coord_a, coord_b, coord_c, coord_d))
    ea, eb, ec, ed = (np.asarray(x, dtype=dtype) for x in (exps_a, exps_b, exps_c, exps_d))
    ca, cb, cc, cd = (np.asarray(x, dtype=dtype) for x in (coeffs_a, coeffs_b, coeffs_c,
                                                           coeffs_d))
    pi = dtype(np.pi) if dtype is np.float64 else dtype("3.14159265358979323846264338327950288")
    ca, cb, cc, cd = (x[:, None] if x.ndim == 1 else x for x in (ca, cb, cc, cd))
    L = la + lb + lc + ld
    Lab, Lcd = la + lb, lc + ld

    # primitive quartets on axes (Ka, Kb, Kc, Kd), flattened
    a4 = ea[:, None, None, None]
    b4 = eb[None, :, None, None]
    c4 = ec[None, None, :, None]
    d4 = ed[None, None, None, :]
    zeta = a4 + b4
    eta = c4 + d4
    shape = np.broadcast(a4, b4, c4, d4).shape
    zeta = np.broadcast_to(zeta, shape).ravel()
    eta = np.broadcast_to(eta, shape).ravel()
    aa = np.broadcast_to(a4, shape).ravel()
    bb = np.broadcast_to(b4, shape).ravel()
    cc4 = np.broadcast_to(c4, shape).ravel()
    dd4 = np.broadcast_to(d4, shape).ravel()
    rho = zeta * eta / (zeta + eta)
    P = (aa[:, None] * A + bb[:, None] * B) / zeta[:, None]
    Q = (cc4[:, None] * C + dd4[:, None] * D) / eta[:, None]
    PQ = P - Q
    PA = P - A
    QC = Q - C
    T = rho * np.sum(PQ * PQ, axis=1)
    pref = (2 * pi ** dtype(2.5) / (zeta * eta * np.sqrt(zeta + eta))
            * np.exp(-(aa * bb / zeta) * np.sum((A - B) ** 2))
            * np.exp(-(cc4 * dd4 / eta) * np.sum((C - D) ** 2)))
    roz = rho / zeta
    h2z = 0.5 / zeta
    zoe = zeta / eta
    h2e = 0.5 / eta

    # 2-3. Boys * prefactor and VRR: V[a] has shape (L - |a| + 1, nprim)
    fm = boys_potential(boys, np.arange(L + 1)[:, None], T[None, :], rho[None, :], potential,
                        omega, dtype) * pref[None, :]
    V = {(0, 0, 0): fm}
    for n in range(1, L + 1):
        for a in _comps(n):
            i = _step(a)
            a1 = _minus(a, i)
            prev = V[a1]
            mmax = L - n
            val = PA[:, i] * prev[:mmax + 1] - roz * PQ[:, i] * prev[1:mmax + 2]
            if a[i] > 1:
                a2 = _minus(a, i, 2)
                p2 = V[a2]
                val = val + (a[i] - 1) * h2z * (p2[:mmax + 1] - roz * p2[1:mmax + 2])
            V[a] = val

    # 4. ETR: E[c][a], m = 0 only, |a| <= L - |c|
    E = {(0, 0, 0): {a: V[a][0] for n in range(L + 1) for a in _comps(n)}}
    coef = QC + zoe[:, None] * PA
    if ket == "vrr":
        # Obara-Saika two-centre build of C (bounded coefficients), for the
        # numerical-stability checks: [a|c+1_i]^(m) = QC_i [a|c]^(m) + WQ_i [a|c]^(m+1)
        #   + c_i/(2eta)([a|c-1_i]^(m) - rho/eta [a|c-1_i]^(m+1)) + a_i/(2(zeta+eta)) [a-1_i|c]^(m+1)
        roe = rho / eta
        WQ = roe[:, None] * PQ
        h2s = 0.5 / (zeta + eta)
        K = {(0, 0, 0): {a: V[a] for n in range(L + 1) for a in _comps(n)}}
        for n in range(1, Lcd + 1):
            for c in _comps(n):
                i = _step(c)
                c1 = _minus(c, i)
                k1 = K[c1]
                k2 = K[_minus(c, i, 2)] if c[i] > 1 else None
                row = {}
                for na in range(0, L - n + 1):
                    for a in _comps(na):
                        mm = L - n - na + 1
                        val = QC[:, i] * k1[a][:mm] + WQ[:, i] * k1[a][1:mm + 1]
                        if k2 is not None:
                            val = val + (c[i] - 1) * h2e * (k2[a][:mm] - roe * k2[a][1:mm + 1])
                        if a[i] > 0:
                            val = val + a[i] * h2s * k1[_minus(a, i)][1:mm + 1]
                        row[a] = val
                K[c] = row
        E = {c: {a: v[0] for a, v in row.items()} for c, row in K.items()}
    for n in range(1, Lcd + 1 if ket == "etr" else 1):
        for c in _comps(n):
            i = _step(c)
            c1 = _minus(c, i)
            e1 = E[c1]
            e2 = E[_minus(c, i, 2)] if c[i] > 1 else None
            row = {}
            for na in range(0, L - n + 1):
                for a in _comps(na):
                    val = coef[:, i] * e1[a] - zoe * e1[_plus(a, i)]
                    if e2 is not None:
                        val = val + (c[i] - 1) * h2e * e2[a]
                    if a[i] > 0:
                        val = val + a[i] * h2e * e1[_minus(a, i)]
                    row[a] = val
            E[c] = row

    # 5. contraction; primitive norms folded into the coefficients (:44-84)
    def normed(e, c, l):
        return c * (((2 / pi) * e) ** dtype(0.75) * (4 * e) ** (dtype(l) / 2))[:, None]

    na_, nb_, nc_, nd_ = normed(ea, ca, la), normed(eb, cb, lb), normed(ec, cc, lc), normed(ed, cd,
                                                                                              ld)
    Ka, Kb, Kc, Kd = ea.size, eb.size, ec.size, ed.size

    def contract(x):
        x = x.reshape(Ka, Kb, Kc, Kd)
        return np.einsum("abcd,aA,bB,cC,dD->ABCD", x, na_, nb_, nc_, nd_, optimize=True)

    G = {}
    for n in range(lc, Lcd + 1):
        for c in _comps(n):
            for m in range(la, Lab + 1):
                for a in _comps(m):
                    G[(a, (0, 0, 0), c, (0, 0, 0))] = contract(E[c][a])

    # 6. HRR on D: [a0|c d+1_i] = [a0|c+1_i d] + CD_i [a0|c d]
    CD = C - D
    AB = A - B
    for nd in range(1, ld + 1):
        for d in _comps(nd):
            i = _step(d)
            d1 = _minus(d, i)
            for n in range(lc, Lcd - nd + 1):
                for c in _comps(n):
                    for m in range(la, Lab + 1):
                        for a in _comps(m):
                            G[(a, (0, 0, 0), c, d)] = (G[(a, (0, 0, 0), _plus(c, i), d1)]
                                                       + CD[i] * G[(a, (0, 0, 0), c, d1)])
    # HRR on B: [a b+1_i|cd] = [a+1_i b|cd] + AB_i [a b|cd]
    for nb in range(1, lb + 1):
        for b in _comps(nb):
            i = _step(b)
            b1 = _minus(b, i)
            for m in range(la, Lab - nb + 1):
                for a in _comps(m):
                    for c in _comps(lc):
                        for d in _comps(ld):
                            G[(a, b, c, d)] = G[(_plus(a, i), b1, c, d)] + AB[i] * G[(a, b1, c, d)]

    # 7. factorial2 component norms, output in the requested component orders
    def f2norm(comp):
        return 1 / np.sqrt(dtype(_dfact(2 * comp[0] - 1) * _dfact(2 * comp[1] - 1)
                                 * _dfact(2 * comp[2] - 1)))

    Ma, Mb, Mc, Md = ca.shape[1], cb.shape[1], cc.shape[1], cd.shape[1]
    out = np.empty((len(comps_a), len(comps_b), len(comps_c), len(comps_d), Ma, Mb, Mc, Md),
                   dtype=dtype)
    for ia, a in enumerate(map(tuple, np.asarray(comps_a).tolist())):
        for ib, b in enumerate(map(tuple, np.asarray(comps_b).tolist())):
            for ic, c in enumerate(map(tuple, np.asarray(comps_c).tolist())):
                for id_, d in enumerate(map(tuple, np.asarray(comps_d).tolist())):
                    out[ia, ib, ic, id_] = (G[(a, b, c, d)] * f2norm(a) * f2norm(b) * f2norm(c)
                                            * f2norm(d))
    return out
