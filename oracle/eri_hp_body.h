/* Body of the high-precision ERI oracle, included by eri_hp.c once per arithmetic type.
 * TEST INFRASTRUCTURE ONLY.  Expects the macros
 *   R            the arithmetic type (__float128 or double)
 *   FN(name)     the exported name of a function for this type
 *   RSQRT, REXP, RERF, RPI   sqrt, exp, erf and pi in that type.
 *
 * One contracted shell quartet (la lb | lc ld), one segmented contraction per shell,
 * by the Obara-Saika / Head-Gordon-Pople scheme of
 * gbasis/integrals/_two_elec_int_improved.py:591-732 (stage order :651-732):
 *   Boys F_m(T) times the primitive prefactor                     (:684-697)
 *   vertical recursion building angular momentum on one bra centre (:87-196)
 *   Obara-Saika two-centre build of the ket centre (an exact replacement of the
 *     electron transfer of :199-361 with bounded coefficients)
 *   contraction with c (2a/pi)^(3/4) (4a)^(l/2)                    (:44-84)
 *   horizontal recursion on the ket, then on the bra               (:364-570)
 *   factorial2 component norms                                     (:572-586)
 * with one difference of organisation: the primitive pairs of a shell pair may be
 * split into two groups, each built on its own centre (A or B for the bra, C or D
 * for the ket) and transferred to the other by its own horizontal recursion; the
 * integral is the sum over the groups (exact, by linearity).  See eri_hp.c.
 */

/* F_m(T), m = 0..L, into F[].  T < 100: the series
 *   F_L(T) = e^-T sum_i (2T)^i / ((2L+1)(2L+3)...(2L+2i+1))
 * (all terms positive), then the downward recurrence F_m = (2T F_{m+1} + e^-T)/(2m+1);
 * T >= 100: F_0 = sqrt(pi/T)/2 erf(sqrt T), then the upward recurrence
 * F_{m+1} = ((2m+1) F_m - e^-T) / (2T), stable for T > m. */
static void FN(boys)(R T, int L, R* F) {
  const R emT = REXP(-T);
  if (T < (R)100) {
    R term = (R)1 / (R)(2 * L + 1), sum = term;
    for (int i = 1; i < 100000; ++i) {
      term *= ((R)2 * T) / (R)(2 * L + 2 * i + 1);
      sum += term;
      if (term < sum * (R)1e-40) break;
    }
    F[L] = emT * sum;
    for (int m = L - 1; m >= 0; --m) F[m] = ((R)2 * T * F[m + 1] + emT) / (R)(2 * m + 1);
  } else {
    const R st = RSQRT(T);
    F[0] = RSQRT(RPI) / (R)2 * RERF(st) / st;
    for (int m = 0; m < L; ++m) F[m + 1] = ((R)(2 * m + 1) * F[m] - emT) / ((R)2 * T);
  }
}

/* Primitive normalisation (2a/pi)^(3/4) (4a)^(l/2) in R. */
static R FN(prim_norm)(double a, int l) {
  const R x = (R)2 * (R)a / RPI;
  R n = RSQRT(x) * RSQRT(RSQRT(x));
  const R s = RSQRT((R)4 * (R)a);
  for (int i = 0; i < l; ++i) n *= s;
  return n;
}

typedef struct {
  int l;
  double c[3];
  int K;
  const double* e;
  const double* w;  /* raw contraction coefficients */
} FN(shell_t);

/* One orientation group.  The bra is built on shell X (the other bra shell is Y),
 * the ket on Z (other W); bp / kp list the primitive pairs of the group as (ix, iy)
 * and (iz, iw).  Adds the group's Cartesian block, in canonical component order of
 * (X, Y, Z, W), times the factorial2 norms, into outg[x][y][z][w]. */
static void FN(group)(const FN(shell_t) * X, const FN(shell_t) * Y, const int* bp, int nbp,
                      const FN(shell_t) * Z, const FN(shell_t) * W, const int* kp, int nkp,
                      R* outg) {
  const int lx = X->l, ly = Y->l, lz = Z->l, lw = W->l;
  const int Lxy = lx + ly, Lzw = lz + lw, L = Lxy + Lzw;
  const int NE = ncle(Lxy), NF = ncle(Lzw);
  const int M = Lzw + 1;               /* orders kept in the ket build */
  R* V = (R*)calloc((size_t)NE * (L + 1), sizeof(R));
  R* K = (R*)calloc((size_t)NF * NE * M, sizeof(R));
  R* acc = (R*)calloc((size_t)NE * NF, sizeof(R));
  R F[MAXLT + 1];
  R XY[3], ZW[3];
  for (int i = 0; i < 3; ++i) {
    XY[i] = (R)X->c[i] - (R)Y->c[i];
    ZW[i] = (R)Z->c[i] - (R)W->c[i];
  }
  const R XY2 = XY[0] * XY[0] + XY[1] * XY[1] + XY[2] * XY[2];
  const R ZW2 = ZW[0] * ZW[0] + ZW[1] * ZW[1] + ZW[2] * ZW[2];
  const R two_pi_2p5 = (R)2 * RPI * RPI * RSQRT(RPI);

  for (int ib = 0; ib < nbp; ++ib) {
    const int ix = bp[2 * ib], iy = bp[2 * ib + 1];
    const R ax = (R)X->e[ix], ay = (R)Y->e[iy];
    const R zeta = ax + ay;
    R P[3], PX[3];
    for (int i = 0; i < 3; ++i) {
      P[i] = (ax * (R)X->c[i] + ay * (R)Y->c[i]) / zeta;
      PX[i] = P[i] - (R)X->c[i];
    }
    const R cb = (R)X->w[ix] * FN(prim_norm)(X->e[ix], lx) * (R)Y->w[iy] *
                 FN(prim_norm)(Y->e[iy], ly) * REXP(-(ax * ay / zeta) * XY2);
    for (int ik = 0; ik < nkp; ++ik) {
      const int iz = kp[2 * ik], iw = kp[2 * ik + 1];
      const R az = (R)Z->e[iz], aw = (R)W->e[iw];
      const R eta = az + aw;
      R Q[3], QZ[3], PQ[3];
      for (int i = 0; i < 3; ++i) {
        Q[i] = (az * (R)Z->c[i] + aw * (R)W->c[i]) / eta;
        QZ[i] = Q[i] - (R)Z->c[i];
        PQ[i] = P[i] - Q[i];
      }
      const R ck = (R)Z->w[iz] * FN(prim_norm)(Z->e[iz], lz) * (R)W->w[iw] *
                   FN(prim_norm)(W->e[iw], lw) * REXP(-(az * aw / eta) * ZW2);
      const R s = zeta + eta;
      const R rho = zeta * eta / s;
      const R T = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
      const R pref = two_pi_2p5 / (zeta * eta * RSQRT(s)) * cb * ck;
      const R roz = rho / zeta, roe = rho / eta;
      const R h2z = (R)1 / ((R)2 * zeta), h2e = (R)1 / ((R)2 * eta), h2s = (R)1 / ((R)2 * s);
      R WP[3], WQ[3];
      for (int i = 0; i < 3; ++i) {
        WP[i] = -roz * PQ[i];   /* W - P */
        WQ[i] = roe * PQ[i];    /* W - Q */
      }
      FN(boys)(T, L, F);
      /* vertical recursion on X: V[a][m], |a| <= Lxy, m <= L - |a| */
      for (int m = 0; m <= L; ++m) V[m] = pref * F[m];
      for (int g = 1; g < NE; ++g) {
        const int n = cdeg[g], i = cdir[g], g1 = cm1[g], g2 = cm2[g];
        const R* v1 = V + (size_t)g1 * (L + 1);
        R* v = V + (size_t)g * (L + 1);
        for (int m = 0; m <= L - n; ++m) {
          R val = PX[i] * v1[m] + WP[i] * v1[m + 1];
          if (g2 >= 0) {
            const R* v2 = V + (size_t)g2 * (L + 1);
            val += (R)(cval[g] - 1) * h2z * (v2[m] - roz * v2[m + 1]);
          }
          v[m] = val;
        }
      }
      /* ket build on Z: K[c][a][m], |c| <= Lzw, |a| <= Lxy, m <= Lzw - |c| */
      for (int a = 0; a < NE; ++a)
        for (int m = 0; m < M; ++m) K[(size_t)a * M + m] = V[(size_t)a * (L + 1) + m];
      for (int c = 1; c < NF; ++c) {
        const int n = cdeg[c], i = cdir[c], c1 = cm1[c], c2 = cm2[c];
        for (int a = 0; a < NE; ++a) {
          const R* k1 = K + ((size_t)c1 * NE + a) * M;
          const R* k2 = c2 >= 0 ? K + ((size_t)c2 * NE + a) * M : 0;
          const int am = csub[a][i];
          const R* ka = am >= 0 ? K + ((size_t)c1 * NE + am) * M : 0;
          R* k = K + ((size_t)c * NE + a) * M;
          for (int m = 0; m <= Lzw - n; ++m) {
            R val = QZ[i] * k1[m] + WQ[i] * k1[m + 1];
            if (k2) val += (R)(cval[c] - 1) * h2e * (k2[m] - roe * k2[m + 1]);
            if (ka) val += (R)ccomp[a][i] * h2s * ka[m + 1];
            k[m] = val;
          }
        }
      }
      for (int f = ncle(lz - 1); f < NF; ++f)
        for (int e = ncle(lx - 1); e < NE; ++e)
          acc[(size_t)e * NF + f] += K[((size_t)f * NE + e) * M];
    }
  }

  /* horizontal recursion on the ket: H[e][c][d] = [e 0 | c d] */
  const int ND = ncle(lw);
  R* H = (R*)calloc((size_t)NE * NF * ND, sizeof(R));
  for (int e = 0; e < NE; ++e)
    for (int f = 0; f < NF; ++f) H[((size_t)e * NF + f) * ND] = acc[(size_t)e * NF + f];
  for (int d = 1; d < ND; ++d) {
    const int k = cdeg[d], i = cdir[d], dm = cm1[d];
    for (int c = ncle(lz - 1); c < ncle(Lzw - k); ++c) {
      const int cp = cadd[c][i];
      for (int e = ncle(lx - 1); e < NE; ++e) {
        R* h = H + (size_t)e * NF * ND;
        h[(size_t)c * ND + d] = h[(size_t)cp * ND + dm] + ZW[i] * h[(size_t)c * ND + dm];
      }
    }
  }
  /* horizontal recursion on the bra: B[a][b][cd], |c| = lz, |d| = lw */
  const int NB = ncle(ly), nz = ncart(lz), nw = ncart(lw), ncd = nz * nw;
  R* Bv = (R*)calloc((size_t)NE * NB * ncd, sizeof(R));
  const int z0 = ncle(lz - 1), w0 = ncle(lw - 1);
  for (int a = ncle(lx - 1); a < NE; ++a)
    for (int c = 0; c < nz; ++c)
      for (int d = 0; d < nw; ++d)
        Bv[((size_t)a * NB) * ncd + c * nw + d] = H[((size_t)a * NF + z0 + c) * ND + w0 + d];
  for (int b = 1; b < NB; ++b) {
    const int k = cdeg[b], i = cdir[b], bm = cm1[b];
    for (int a = ncle(lx - 1); a < ncle(Lxy - k); ++a) {
      const int ap = cadd[a][i];
      for (int t = 0; t < ncd; ++t)
        Bv[((size_t)a * NB + b) * ncd + t] =
            Bv[((size_t)ap * NB + bm) * ncd + t] + XY[i] * Bv[((size_t)a * NB + bm) * ncd + t];
    }
  }
  const int nx = ncart(lx), ny = ncart(ly), x0 = ncle(lx - 1), y0 = ncle(ly - 1);
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int c = 0; c < nz; ++c)
        for (int d = 0; d < nw; ++d) {
          const R nrm = f2n[x0 + x] * f2n[y0 + y] * f2n[z0 + c] * f2n[w0 + d];
          outg[((x * ny + y) * nz + c) * nw + d] +=
              (R)Bv[((size_t)(x0 + x) * NB + y0 + y) * ncd + c * nw + d] * (R)nrm;
        }
  free(V);
  free(K);
  free(acc);
  free(H);
  free(Bv);
}

/* Build-centre choice of the primitive pairs of one shell pair (first, second):
 * group[r] = 1 when pair r = (i, j) is built on the second shell.  mode 0: never
 * (the reference's orientation); mode 1: when the second shell has l > 0 and the
 * larger exponent, and the horizontal recursion from the first centre would
 * multiply rounding by more than tau per unit of angular momentum, estimated as
 * r = (|AB| + max(|PA|, s)) / max(|PB|, s) with s = 1/sqrt(2 zeta) the width of the
 * product Gaussian (tau = 0: every such pair); mode 2: the same with r^l_second >
 * tau (the whole amplification of the second shell's transfer). */
static void FN(orient)(const FN(shell_t) * S1, const FN(shell_t) * S2, int mode, double tau,
                       int* group) {
  double ab2 = 0;
  for (int i = 0; i < 3; ++i) ab2 += (S1->c[i] - S2->c[i]) * (S1->c[i] - S2->c[i]);
  const double ab = sqrt(ab2);
  for (int i = 0; i < S1->K; ++i)
    for (int j = 0; j < S2->K; ++j) {
      int g = 0;
      if ((mode == 1 || mode == 2) && S2->l > 0 && S2->e[j] > S1->e[i]) {
        const double z = S1->e[i] + S2->e[j];
        const double s = 1.0 / sqrt(2.0 * z);
        const double pa = S2->e[j] / z * ab, pb = S1->e[i] / z * ab;
        const double ratio = (ab + fmax(pa, s)) / fmax(pb, s);
        g = mode == 1 ? ratio > tau : pow(ratio, S2->l) > tau;
      }
      group[i * S2->K + j] = g;
    }
}

/* Cartesian block of (a b | c d), canonical component order, factorial2 norms applied,
 * no norm_cont.  l[4], centre[12], K[4]; exps / coefs concatenated over the four
 * shells.  mode / tau: build-centre choice (see orient).  out_hi[na nb nc nd] gets
 * the value rounded to double, out_lo (optional) the remainder.  Returns 0, or -1 on
 * unsupported input. */
int FN(eri_block)(const int* l, const double* centre, const int* K, const double* exps,
                  const double* coefs, int mode, double tau, double* out_hi, double* out_lo) {
  if (!tables_ready) init_tables();
  FN(shell_t) sh[4];
  int off = 0;
  for (int s = 0; s < 4; ++s) {
    if (l[s] < 0 || l[s] > MAXL || K[s] < 1) return -1;
    sh[s].l = l[s];
    for (int i = 0; i < 3; ++i) sh[s].c[i] = centre[3 * s + i];
    sh[s].K = K[s];
    sh[s].e = exps + off;
    sh[s].w = coefs + off;
    off += K[s];
  }
  const int na = ncart(l[0]), nb = ncart(l[1]), nc = ncart(l[2]), nd = ncart(l[3]);
  const int nout = na * nb * nc * nd;
  R* out = (R*)calloc((size_t)nout, sizeof(R));
  /* primitive-pair groups of the bra (a, b) and ket (c, d) */
  int* gb = (int*)malloc(sizeof(int) * K[0] * K[1]);
  int* gk = (int*)malloc(sizeof(int) * K[2] * K[3]);
  FN(orient)(&sh[0], &sh[1], mode, tau, gb);
  FN(orient)(&sh[2], &sh[3], mode, tau, gk);
  int* bp[2];
  int nbp[2] = {0, 0}, nkp[2] = {0, 0};
  int* kp[2];
  for (int g = 0; g < 2; ++g) {
    bp[g] = (int*)malloc(sizeof(int) * 2 * K[0] * K[1]);
    kp[g] = (int*)malloc(sizeof(int) * 2 * K[2] * K[3]);
  }
  for (int i = 0; i < K[0]; ++i)
    for (int j = 0; j < K[1]; ++j) {
      const int g = gb[i * K[1] + j];
      /* (index on the build centre, index on the other) */
      bp[g][2 * nbp[g]] = g ? j : i;
      bp[g][2 * nbp[g] + 1] = g ? i : j;
      ++nbp[g];
    }
  for (int i = 0; i < K[2]; ++i)
    for (int j = 0; j < K[3]; ++j) {
      const int g = gk[i * K[3] + j];
      kp[g][2 * nkp[g]] = g ? j : i;
      kp[g][2 * nkp[g] + 1] = g ? i : j;
      ++nkp[g];
    }
  R* outg = (R*)calloc((size_t)nout, sizeof(R));
  for (int g1 = 0; g1 < 2; ++g1)
    for (int g2 = 0; g2 < 2; ++g2) {
      if (!nbp[g1] || !nkp[g2]) continue;
      const FN(shell_t)* X = g1 ? &sh[1] : &sh[0];
      const FN(shell_t)* Y = g1 ? &sh[0] : &sh[1];
      const FN(shell_t)* Z = g2 ? &sh[3] : &sh[2];
      const FN(shell_t)* W = g2 ? &sh[2] : &sh[3];
      memset(outg, 0, sizeof(R) * nout);
      FN(group)(X, Y, bp[g1], nbp[g1], Z, W, kp[g2], nkp[g2], outg);
      const int nx = ncart(X->l), ny = ncart(Y->l), nz = ncart(Z->l), nw = ncart(W->l);
      for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
          for (int z = 0; z < nz; ++z)
            for (int w = 0; w < nw; ++w) {
              const int ia = g1 ? y : x, ib = g1 ? x : y;
              const int ic = g2 ? w : z, id = g2 ? z : w;
              out[((ia * nb + ib) * nc + ic) * nd + id] += outg[((x * ny + y) * nz + z) * nw + w];
            }
    }
  for (int o = 0; o < nout; ++o) {
    const double hi = (double)out[o];
    out_hi[o] = hi;
    if (out_lo) out_lo[o] = (double)(out[o] - (R)hi);
  }
  free(outg);
  free(out);
  free(gb);
  free(gk);
  for (int g = 0; g < 2; ++g) {
    free(bp[g]);
    free(kp[g]);
  }
  return 0;
}

/* F_m(T), m = 0..L, rounded to double (hi) with the remainder (lo). */
int FN(boys_values)(double T, int L, double* hi, double* lo) {
  if (L < 0 || L > MAXLT || !(T >= 0)) return -1;
  R F[MAXLT + 1];
  FN(boys)((R)T, L, F);
  for (int m = 0; m <= L; ++m) {
    hi[m] = (double)F[m];
    if (lo) lo[m] = (double)(F[m] - (R)hi[m]);
  }
  return 0;
}
