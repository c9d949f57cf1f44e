/*
 * Wigner-3j family recurrence, included twice by wigner.c with
 *   REAL  = double       / long double
 *   SFX   = _d           / _ld
 *   RSQRT = sqrt         / sqrtl
 *   RABS  = fabs         / fabsl
 *
 * TEST INFRASTRUCTURE (oracle): restates the algorithm behind the external
 * WignerFamilies.jl calls of the reference (`WignerF(T, l1, l2, m1, m2)` +
 * `wigner3j_f!`, call sites src/modecoupling.jl:70,73 and
 * src/covariance.jl:105-108).  WignerFamilies.jl (compat "1",
 * Project.toml:25,43) is not vendored; its published method is the
 * Schulten & Gordon (J. Math. Phys. 16, 1961 (1975)) three-term recurrence in
 * the first angular momentum with the two-sided, sum-rule normalised scheme of
 * Luscombe & Luban (Phys. Rev. E 57, 7274 (1998)).
 *
 * Family: f(j) = ( j  l1  l2 ; -m1-m2  m1  m2 ),  j in [nmin, nmax],
 *   nmin = max(|l1-l2|, |m1+m2|), nmax = l1+l2.
 * In Schulten-Gordon notation (j1 j2 j3; M1 M2 M3) = (j l1 l2; -m1-m2 m1 m2):
 *   X(j) f(j+1) + Y(j) f(j) + Z(j) f(j-1) = 0
 *   A(j) = sqrt[(j^2-(j2-j3)^2)((j2+j3+1)^2-j^2)(j^2-M1^2)]
 *   X(j) = j A(j+1),  Z(j) = (j+1) A(j)
 *   Y(j) = -(2j+1)[ j2(j2+1) M1 - j3(j3+1) M1 - j(j+1)(M3-M2) ]
 * Normalisation  sum_j (2j+1) f(j)^2 = 1,  sign f(nmax) = (-1)^(j2-j3+M1).
 */

static REAL SG_A(SFX)(long j, long j2, long j3, long M1) {
  REAL jj = (REAL)j;
  REAL d = (REAL)(j2 - j3), s = (REAL)(j2 + j3 + 1), m = (REAL)M1;
  REAL v = (jj * jj - d * d) * (s * s - jj * jj) * (jj * jj - m * m);
  return v > 0 ? RSQRT(v) : (REAL)0;
}

static REAL SG_Y(SFX)(long j, long j2, long j3, long M1, long M2, long M3) {
  REAL jj = (REAL)j;
  return -((REAL)(2 * j + 1)) *
         ((REAL)(j2 * (j2 + 1) * M1) - (REAL)(j3 * (j3 + 1) * M1) -
          jj * (jj + 1) * (REAL)(M3 - M2));
}

/* Fill f[0 .. nmax-nmin] and return the count; *nmin_out receives nmin.
 * Families with |m1| > l1 or |m2| > l2 (no valid 3j symbol) are returned as
 * zeros: the physical value.  What WignerFamilies.jl returns there is not
 * pinned by anything in the reference. */
static int wigner_family(SFX)(int l1, int l2, int m1, int m2, REAL* f,
                              int* nmin_out) {
  const long j2 = l1, j3 = l2, M2 = m1, M3 = m2, M1 = -(long)m1 - m2;
  long nmin = labs(j2 - j3);
  if (labs(M1) > nmin) nmin = labs(M1);
  const long nmax = j2 + j3;
  *nmin_out = (int)nmin;
  if (nmax < nmin) return 0;
  const long n = nmax - nmin + 1;
  if (labs(m1) > l1 || labs(m2) > l2) {
    for (long k = 0; k < n; ++k) f[k] = 0;
    return (int)n;
  }
  const REAL sgn_max = (((j2 - j3 + M1) % 2) == 0) ? (REAL)1 : (REAL)-1;
  if (n == 1) {
    f[0] = sgn_max / RSQRT((REAL)(2 * nmin + 1));
    return 1;
  }
  const REAL big = (REAL)1e150, tiny = (REAL)1e-150;

  if (M1 == 0 && M2 == 0 && M3 == 0) {
    /* Y == 0: two-term recurrence f(j+1) = -Z(j)/X(j) f(j-1); entries with
     * odd j+l1+l2 vanish.  Forward from nmin (X(j) != 0 for j >= 1). */
    for (long k = 0; k < n; ++k) f[k] = 0;
    f[0] = 1;
    for (long j = nmin + 1; j + 1 <= nmax; j += 2) {
      REAL X = (REAL)j * SG_A(SFX)(j + 1, j2, j3, M1);
      REAL Z = (REAL)(j + 1) * SG_A(SFX)(j, j2, j3, M1);
      f[j + 1 - nmin] = -Z / X * f[j - 1 - nmin];
      if (RABS(f[j + 1 - nmin]) > big)
        for (long k = 0; k <= j + 1 - nmin; ++k) f[k] *= tiny;
    }
  } else {
    /* Forward three-term recurrence from nmin through the lower classically
     * forbidden region, where the wanted solution grows with j and forward
     * recursion is stable.  The sweep stops at the first j where |f| would
     * decrease; jf is then the position of the largest forward value. */
    long jf = nmin;
    f[0] = 1;
    if (nmin + 1 <= nmax) {
      REAL f1;
      if (nmin == 0) {
        /* j = 0 (l1 == l2, M1 == 0): X(0) = Z(0) = Y(0) = 0.  Dividing the
         * recurrence by j and letting j -> 0 gives A(1) f(1) + (M3-M2) f(0) = 0. */
        f1 = -(REAL)(M3 - M2) / SG_A(SFX)(1, j2, j3, M1);
      } else {
        REAL X = (REAL)nmin * SG_A(SFX)(nmin + 1, j2, j3, M1);
        f1 = -SG_Y(SFX)(nmin, j2, j3, M1, M2, M3) / X;
      }
      if (RABS(f1) > RABS(f[0])) {
        f[1] = f1;
        jf = nmin + 1;
        while (jf + 1 <= nmax) {
          long j = jf;
          REAL X = (REAL)j * SG_A(SFX)(j + 1, j2, j3, M1);
          REAL Z = (REAL)(j + 1) * SG_A(SFX)(j, j2, j3, M1);
          REAL Y = SG_Y(SFX)(j, j2, j3, M1, M2, M3);
          REAL nxt = -(Y * f[j - nmin] + Z * f[j - 1 - nmin]) / X;
          if (!(RABS(nxt) > RABS(f[j - nmin]))) break;
          f[j + 1 - nmin] = nxt;
          jf = j + 1;
          if (RABS(nxt) > big)
            for (long k = 0; k <= jf - nmin; ++k) f[k] *= tiny;
        }
      }
    }
    if (jf < nmax) {
      /* Backward three-term recurrence from nmax down to jf: the wanted
       * solution grows away from nmax through the upper forbidden region and
       * both solutions are of comparable size in the classical region.
       * X(nmax) = 0, so the start is g(nmax-1) = -Y(nmax) g(nmax) / Z(nmax). */
      const REAL fj = f[jf - nmin];
      REAL gj;
      REAL gnm1 = -SG_Y(SFX)(nmax, j2, j3, M1, M2, M3) /
                  ((REAL)(nmax + 1) * SG_A(SFX)(nmax, j2, j3, M1));
      f[nmax - nmin] = 1;
      if (nmax - 1 == jf) {
        gj = gnm1;
      } else {
        f[nmax - 1 - nmin] = gnm1;
        gj = 0;
        for (long j = nmax - 1; j - 1 >= jf; --j) {
          REAL X = (REAL)j * SG_A(SFX)(j + 1, j2, j3, M1);
          REAL Z = (REAL)(j + 1) * SG_A(SFX)(j, j2, j3, M1);
          REAL Y = SG_Y(SFX)(j, j2, j3, M1, M2, M3);
          REAL prv = -(X * f[j + 1 - nmin] + Y * f[j - nmin]) / Z;
          if (j - 1 == jf) {
            gj = prv;
            break;
          }
          f[j - 1 - nmin] = prv;
          if (RABS(prv) > big)
            for (long k = j - 1; k <= nmax; ++k) f[k - nmin] *= tiny;
        }
      }
      const REAL scale = fj / gj;
      for (long k = jf + 1; k <= nmax; ++k) f[k - nmin] *= scale;
    }
  }
  /* sum-rule normalisation and sign convention */
  REAL s = 0;
  for (long k = 0; k < n; ++k) s += (REAL)(2 * (nmin + k) + 1) * f[k] * f[k];
  REAL norm = (REAL)1 / RSQRT(s);
  if ((f[n - 1] < 0) != (sgn_max < 0)) norm = -norm;
  for (long k = 0; k < n; ++k) f[k] *= norm;
  return (int)n;
}
