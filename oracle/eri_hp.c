/* High-precision ERI oracle.  TEST INFRASTRUCTURE ONLY: only tests/, the fixture
 * scripts under tests/golden/ and the cpu_baseline leg of bench.py may load it, as the
 * checker.  Built by oracle/build_hp.sh into oracle/liboracle_hp.so.
 *
 * A C restatement of the Obara-Saika / Head-Gordon-Pople engine of
 * gbasis/integrals/_two_elec_int_improved.py:591-732 (see eri_hp_body.h for the
 * stages), compiled twice from one body:
 *   *_q  in IEEE binary128 (__float128, 113-bit significand, libquadmath), the
 *        accuracy reference for FP64 parity: the horizontal recursion after
 *        contraction amplifies the rounding of its inputs by up to
 *        (|AB| sqrt(2 zeta))^l_b when a primitive pair's product Gaussian sits on
 *        the second centre (tight exponent there, diffuse on the first); 113 bits
 *        leave margin for that in every benchmark workload, and the build-centre
 *        choice of mode 1 removes the amplification altogether;
 *   *_d  in binary64 with the same operation order, to measure what plain FP64
 *        reaches with a given build-centre rule.
 * How the high-precision values are pinned: the HORTON primitive integrals of the
 * reference's test suite (tests/test_two_elec_int.py:1153-1365), an independent
 * McMurchie-Davidson evaluation in mpmath (up to g functions, exponents up to 1e5),
 * agreement of the two build-centre modes where both are well conditioned, the
 * (ss|ss) closed form, and the permutational and rotational symmetries
 * (tests/test_oracle_hp.py).
 */
#include <math.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>

#define MAXL 8           /* angular momentum per shell */
#define MAXLT 32         /* total angular momentum of a quartet */
#define MAXDEG 16        /* degree of the component tables (Lab <= 16) */

static int ncle(int n) { return n < 0 ? 0 : (n + 1) * (n + 2) * (n + 3) / 6; }
static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

/* Cartesian components of degree <= MAXDEG in the canonical order
 * (x descending, then y descending), indexed globally: degree-n components start
 * at ncle(n-1). */
#define NCOMP 969        /* ncle(MAXDEG) */
static int tables_ready = 0;
static int ccomp[NCOMP][3], cdeg[NCOMP], cdir[NCOMP], cval[NCOMP], cm1[NCOMP], cm2[NCOMP];
static int csub[NCOMP][3], cadd[NCOMP][3];
static long double f2n_ld[NCOMP];
static __float128 f2n_q[NCOMP];
static double f2n_d[NCOMP];

static int cidx(int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0) return -1;
  const int n = x + y + z;
  if (n > MAXDEG) return -1;
  const int r = n - x;
  return ncle(n - 1) + r * (r + 1) / 2 + z;
}

static double dfact(int n) {
  double r = 1;
  while (n > 1) {
    r *= n;
    n -= 2;
  }
  return r;
}

static void init_tables(void) {
  for (int n = 0; n <= MAXDEG; ++n)
    for (int x = n; x >= 0; --x)
      for (int y = n - x; y >= 0; --y) {
        const int z = n - x - y, g = cidx(x, y, z);
        ccomp[g][0] = x;
        ccomp[g][1] = y;
        ccomp[g][2] = z;
        cdeg[g] = n;
        /* recursion direction: z if a_z > 0, else y if a_y > 0, else x (the
         * reference's sweep order) */
        const int i = z > 0 ? 2 : (y > 0 ? 1 : 0);
        cdir[g] = n > 0 ? i : -1;
        cval[g] = ccomp[g][i];
        int c[3] = {x, y, z};
        c[i] -= 1;
        cm1[g] = n > 0 ? cidx(c[0], c[1], c[2]) : -1;
        c[i] -= 1;
        cm2[g] = (n > 0 && c[i] >= 0) ? cidx(c[0], c[1], c[2]) : -1;
        for (int j = 0; j < 3; ++j) {
          int d[3] = {x, y, z};
          d[j] -= 1;
          csub[g][j] = cidx(d[0], d[1], d[2]);
          d[j] += 2;
          cadd[g][j] = cidx(d[0], d[1], d[2]);
        }
        const double p = dfact(2 * x - 1) * dfact(2 * y - 1) * dfact(2 * z - 1);
        f2n_q[g] = 1 / sqrtq((__float128)p);
        f2n_d[g] = 1 / sqrt(p);
        f2n_ld[g] = 1 / sqrtl((long double)p);
      }
  tables_ready = 1;
}

/* ---- binary128 ---- */
#define R __float128
#define FN(name) name##_q
#define RSQRT sqrtq
#define REXP expq
#define RERF erfq
#define RPI M_PIq
#define f2n f2n_q
#include "eri_hp_body.h"
#undef R
#undef FN
#undef RSQRT
#undef REXP
#undef RERF
#undef RPI
#undef f2n

/* ---- binary64 ---- */
#define R double
#define FN(name) name##_d
#define RSQRT sqrt
#define REXP exp
#define RERF erf
#define RPI M_PI
#define f2n f2n_d
#include "eri_hp_body.h"
