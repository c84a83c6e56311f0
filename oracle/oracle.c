/*
 * oracle.c -- CPU restatement of exciting's LAPW H/O setup hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the HIP path and
 * the timed CPU baseline, never part of the product.
 *
 * Complex arithmetic is written out component-wise so the rounding of each
 * step is the one the Fortran source implies (Smith's division for complex
 * quotients, as gfortran/flang emit it).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static int g_fault = 0;
void orc_set_fault(int code) { g_fault = code; }

static inline double zre(orc_z a) { return creal(a); }
static inline double zim(orc_z a) { return cimag(a); }
static inline orc_z zmk(double r, double i) {
  orc_z z;
  __real__ z = r;
  __imag__ z = i;
  return z;
}
static inline orc_z zmul(orc_z a, orc_z b) {
  return zmk(zre(a) * zre(b) - zim(a) * zim(b),
             zre(a) * zim(b) + zim(a) * zre(b));
}
static inline orc_z zconj(orc_z a) { return zmk(zre(a), -zim(a)); }
static inline orc_z zadd(orc_z a, orc_z b) {
  return zmk(zre(a) + zre(b), zim(a) + zim(b));
}
static inline orc_z zsub(orc_z a, orc_z b) {
  return zmk(zre(a) - zre(b), zim(a) - zim(b));
}
/* Smith's algorithm, the complex quotient gfortran and flang generate. */
static inline orc_z zdiv(orc_z a, orc_z b) {
  double ar = zre(a), ai = zim(a), br = zre(b), bi = zim(b);
  if (fabs(bi) <= fabs(br)) {
    double r = bi / br, den = br + bi * r;
    return zmk((ar + ai * r) / den, (ai - ar * r) / den);
  } else {
    double r = br / bi, den = bi + br * r;
    return zmk((ar * r + ai) / den, (ai * r - ar) / den);
  }
}

/* ------------------------------------------------------------------ */
/* src/src_lib/factnm.f90, factr.f90                                    */
/* ------------------------------------------------------------------ */

/* factnm.f90:13-75: multifactorial n!^(m); m=1 and m=2 use the tabulated
 * exact values (n! for n<=24, n!! for n<=38) as decimal literals. */
static const double k_fact1[24] = {
    1.0, 2.0, 6.0, 24.0, 120.0, 720.0, 5040.0, 40320.0, 362880.0, 3628800.0,
    39916800.0, 479001600.0, 6227020800.0, 87178291200.0, 1307674368000.0,
    20922789888000.0, 355687428096000.0, 6402373705728000.0,
    121645100408832000.0, 2432902008176640000.0, 51090942171709440000.0,
    1124000727777607680000.0, 25852016738884976640000.0,
    620448401733239439360000.0};
static const double k_fact2[38] = {
    1.0, 2.0, 3.0, 8.0, 15.0, 48.0, 105.0, 384.0, 945.0, 3840.0, 10395.0,
    46080.0, 135135.0, 645120.0, 2027025.0, 10321920.0, 34459425.0,
    185794560.0, 654729075.0, 3715891200.0, 13749310575.0, 81749606400.0,
    316234143225.0, 1961990553600.0, 7905853580625.0, 51011754393600.0,
    213458046676875.0, 1428329123020800.0, 6190283353629375.0,
    42849873690624000.0, 191898783962510625.0, 1371195958099968000.0,
    6332659870762850625.0, 46620662575398912000.0, 221643095476699771875.0,
    1678343852714360832000.0, 8200794532637891559375.0,
    63777066403145711616000.0};

double orc_factnm(int n, int m) {
  if (n == 0) return 1.0;
  if (m == 1 && n >= 1 && n <= 24) return k_fact1[n - 1];
  if (m == 2 && n >= 1 && n <= 38) return k_fact2[n - 1];
  if (n < 0 || m <= 0 || n > 150) return NAN;
  double f;
  if (m == 1) {
    f = k_fact1[23];
    for (int i = 25; i <= n; i++) f *= (double)i;
  } else {
    int j = n / m;
    if (n % m == 0) j--;
    f = (double)n;
    for (int i = 1; i <= j; i++) f *= (double)(n - i * m);
  }
  return f;
}

/* factr.f90:13-52: n!/d! */
double orc_factr(int n, int d) {
  if (n < 0 || d < 0) return NAN;
  double f;
  if (n < d) {
    f = 1.0 / (double)(n + 1);
    for (int i = n + 2; i <= d; i++) f /= (double)i;
  } else if (n == d) {
    f = 1.0;
  } else {
    f = (double)(d + 1);
    for (int i = d + 2; i <= n; i++) f *= (double)i;
  }
  return f;
}

/* ------------------------------------------------------------------ */
/* src/src_lib/sbessel.f90:13-146                                       */
/* ------------------------------------------------------------------ */
int orc_sbessel(int lmax, double x, double *jl) {
  const int lst = 25;           /* sbessel.f90:40 */
  const double rsc = 1.0e150;   /* sbessel.f90:43 */
  const double rsci = 1.0 / rsc;
  if (lmax < 0 || lmax > 50) return -1;
  if (x < 0.0 || x > 1.0e5) return -2;
  if (x < 1.0e-8) { /* :61-71 */
    double t1 = 1.0, t2 = 1.0;
    jl[0] = 1.0;
    for (int l = 1; l <= lmax; l++) {
      t1 = t1 / (double)(2 * l + 1);
      t2 = t2 * x;
      jl[l] = t2 * t1;
    }
    return 0;
  }
  double xi = 1.0 / x;
  if (x < (double)lmax) { /* :74-118 downward recursion */
    double j0 = 0.6370354636449841609 * rsci;
    double j1 = 0.3532702964695481204 * rsci;
    double jt;
    for (int l = lmax + lst; l >= lmax + 1; l--) {
      jt = j0 * (double)(2 * l + 1) * xi - j1;
      j1 = j0;
      j0 = jt;
      if (fabs(j0) > rsc) {
        jt *= rsci;
        j1 *= rsci;
        j0 *= rsci;
      }
    }
    for (int l = lmax; l >= 0; l--) {
      jt = j0 * (double)(2 * l + 1) * xi - j1;
      j1 = j0;
      j0 = jt;
      if (fabs(j0) > rsc) {
        jt *= rsci;
        j1 *= rsci;
        j0 *= rsci;
        for (int k = l + 1; k <= lmax; k++) jl[k] *= rsci;
      }
      jl[l] = j1;
    }
    double t1 = 1.0 / ((jl[0] - x * jl[1]) * cos(x) + x * jl[0] * sin(x));
    for (int l = 0; l <= lmax; l++) jl[l] = t1 * jl[l];
    return 0;
  }
  /* :119-145 upward recursion */
  jl[0] = sin(x) * xi;
  if (lmax == 0) return 0;
  jl[1] = (jl[0] - cos(x)) * xi;
  if (lmax == 1) return 0;
  double j0 = jl[0], j1 = jl[1], jt;
  for (int l = 2; l <= lmax; l++) {
    jt = (double)(2 * l - 1) * j1 * xi - j0;
    j0 = j1;
    j1 = jt;
    jl[l] = j1;
  }
  if (g_fault == 1) jl[lmax] *= 1.0 + 1e-9; /* planted mistake */
  return 0;
}

/* integer power by repeated multiplication (Fortran x**j, j>=0) */
static double ipow(double x, int j) {
  double r = 1.0;
  for (int i = 0; i < j; i++) r *= x;
  return r;
}

/* ------------------------------------------------------------------ */
/* src/src_lib/sbesseldm.f90:13-133                                     */
/* ------------------------------------------------------------------ */
int orc_sbesseldm(int m, int lmax, double x, double *djl) {
  const int maxm = 6, maxns = 20; /* :44-45 */
  if (m < 0 || m > maxm) return -1;
  if (lmax < 0 || lmax > 30) return -2;
  if (x < 0.0 || x > 1.0e5) return -3;
  if (m == 0) return orc_sbessel(lmax, x, djl);
  if (x > 1.0) { /* :79-106 */
    double jl[32];
    orc_sbessel(lmax + 1, x, jl);
    for (int l = 0; l <= lmax; l++) {
      int a[8] = {0}, a1[8] = {0}, b[8] = {0}, b1[8] = {0};
      a[0] = 1;
      for (int i = 1; i <= m; i++) {
        b[0] = 0;
        b1[0] = 0;
        for (int j = 0; j <= i; j++) {
          b[j + 1] = a[j] * (l - j);
          b1[j + 1] = -a1[j] * (j + l + 2);
        }
        for (int j = 0; j <= i; j++) {
          b1[j] = b1[j] - a[j];
          b[j] = b[j] + a1[j];
        }
        for (int j = 0; j <= i + 1; j++) {
          a[j] = b[j];
          a1[j] = b1[j];
        }
      }
      double t1 = 1.0;
      double sum = (double)a[0] * jl[l] + (double)a1[0] * jl[l + 1];
      for (int i = 1; i <= m + 1; i++) {
        t1 = t1 * x;
        sum = sum + ((double)a[i] * jl[l] + (double)a1[i] * jl[l + 1]) / t1;
      }
      djl[l] = sum;
    }
    return 0;
  }
  /* :107-130 series */
  double x2 = x * x;
  for (int l = 0; l <= lmax; l++) {
    int i0 = (m - l + 1) / 2; /* Fortran integer division truncates */
    if (i0 < 0) i0 = 0;
    int j = 2 * i0 + l - m;
    double t1 = (j == 0) ? 1.0 : ipow(x, j);
    double pw = 1.0;
    for (int k = 0; k < i0; k++) pw *= -2.0;
    t1 = orc_factr(j + m, j) * t1 /
         (orc_factnm(i0, 1) * orc_factnm(j + l + m + 1, 2) * pw);
    double sum = t1;
    for (int i = i0 + 1; i <= maxns; i++) {
      j = 2 * i + l;
      t1 = -t1 * (double)((j - 1) * j) * x2 /
           (double)((j - l) * (j - m - 1) * (j - m) * (j + l + 1));
      if (fabs(t1) <= 1.0e-40) break;
      sum = sum + t1;
    }
    djl[l] = sum;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* src/src_lib/genylm.f90:13-94 (Masters & Richards-Dinger recursion)   */
/* ------------------------------------------------------------------ */
int orc_genylm(int lmax, const double *tp, orc_z *ylm) {
  const double fourpi = 12.566370614359172954;
  if (lmax < 0 || lmax > 50) return -1;
  ylm[0] = zmk(0.28209479177387814347, 0.0);
  if (lmax == 0) return 0;
  double sn = sin(tp[0]), cs = cos(tp[0]);
  double x[51];
  orc_z z[51];
  for (int m = 1; m <= lmax; m++) {
    double t1 = (double)m * tp[1];
    z[m] = zmk(cos(t1), sin(t1));
  }
  for (int l = 1; l <= lmax; l++) {
    x[l] = (l % 2 == 0) ? 1.0 : -1.0;
    double dx = 0.0;
    for (int m = l; m >= 1; m--) {
      double t1 = sqrt((double)((l + m) * (l - m + 1)));
      x[m - 1] = -(sn * dx + (double)(2 * m) * cs * x[m]) / t1;
      dx = sn * x[m] * t1;
    }
    double t1 = sn, sum = 0.0;
    for (int m = 1; m <= l; m++) {
      x[m] = t1 * x[m];
      sum = sum + x[m] * x[m];
      t1 = t1 * sn;
    }
    sum = 2.0 * sum + x[0] * x[0];
    t1 = sqrt((double)(2 * l + 1) / (fourpi * sum));
    int lm1 = l * (l + 1), lm2 = lm1; /* 0-based l(l+1)+m */
    ylm[lm1] = zmk(t1 * x[0], 0.0);
    for (int m = 1; m <= l; m++) {
      lm1++;
      lm2--;
      /* t1*x(m)*z(m): real*real then real*complex */
      double s = t1 * x[m];
      ylm[lm1] = zmk(s * zre(z[m]), s * zim(z[m]));
      orc_z c = zconj(ylm[lm1]);
      if (m % 2 != 0) c = zmk(-zre(c), -zim(c));
      ylm[lm2] = c;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* src/src_lib/polynom.f90:11-188                                       */
/* ------------------------------------------------------------------ */
double orc_polynom(int m, int np, const double *xa, const double *ya,
                   double *c, double x) {
  double x0, x1, x2, x3, y1, y2, y3, t1, t2, t3, t4, t5, t6, t7, sum;
  switch (np) {
    case 1:
      if (m < 0) return ya[0] * (x - xa[0]);
      if (m == 0) return ya[0];
      return 0.0;
    case 2:
      c[1] = (ya[1] - ya[0]) / (xa[1] - xa[0]);
      t1 = x - xa[0];
      if (m < 0) return t1 * (ya[0] + 0.5 * c[1] * t1);
      if (m == 0) return c[1] * t1 + ya[0];
      if (m == 1) return c[1];
      return 0.0;
    case 3:
      x1 = xa[1] - xa[0];
      x2 = xa[2] - xa[0];
      y1 = ya[1] - ya[0];
      y2 = ya[2] - ya[0];
      t1 = 1.0 / (x1 * x2 * (x2 - x1));
      t2 = x1 * y2;
      t3 = x2 * y1;
      c[1] = t1 * (x2 * t3 - x1 * t2);
      c[2] = t1 * (t2 - t3);
      t1 = x - xa[0];
      if (m < 0)
        return t1 * (ya[0] + t1 * (0.5 * c[1] + 0.3333333333333333333 * c[2] * t1));
      if (m == 0) return ya[0] + t1 * (c[1] + c[2] * t1);
      if (m == 1) return 2.0 * c[2] * t1 + c[1];
      if (m == 2) return 2.0 * c[2];
      return 0.0;
    case 4:
      x1 = xa[1] - xa[0];
      x2 = xa[2] - xa[0];
      x3 = xa[3] - xa[0];
      y1 = ya[1] - ya[0];
      y2 = ya[2] - ya[0];
      y3 = ya[3] - ya[0];
      t1 = 1.0 / (x1 * x2 * x3 * (x1 - x2) * (x1 - x3) * (x2 - x3));
      t2 = x1 * x2 * y3;
      t3 = x2 * x3 * y1;
      t4 = x3 * x1 * y2;
      t5 = x1 * x1;
      t6 = x2 * x2;
      t7 = x3 * x3;
      c[1] = t1 * (t3 * (x3 * t6 - x2 * t7) + t4 * (x1 * t7 - x3 * t5) +
                   t2 * (x2 * t5 - x1 * t6));
      c[2] = t1 * (t3 * (t7 - t6) + t4 * (t5 - t7) + t2 * (t6 - t5));
      c[3] = t1 * (t3 * (x2 - x3) + t4 * (x3 - x1) + t2 * (x1 - x2));
      t1 = x - xa[0];
      if (m < 0)
        return t1 * (ya[0] + t1 * (0.5 * c[1] +
                                   t1 * (0.3333333333333333333 * c[2] +
                                         0.25 * c[3] * t1)));
      if (m == 0) return ya[0] + t1 * (c[1] + t1 * (c[2] + c[3] * t1));
      if (m == 1) return c[1] + t1 * (2.0 * c[2] + 3.0 * c[3] * t1);
      if (m == 2) return 6.0 * c[3] * t1 + 2.0 * c[2];
      if (m == 3) return 6.0 * c[3];
      return 0.0;
    default:
      break;
  }
  if (np <= 0) return NAN;
  if (m >= np) return 0.0;
  /* divided differences (1-based loops transcribed) */
  for (int i = 0; i < np; i++) c[i] = ya[i];
  for (int i = 2; i <= np; i++)
    for (int j = np; j >= i; j--)
      c[j - 1] = (c[j - 1] - c[j - 2]) / (xa[j - 1] - xa[j - i]);
  if (m == 0) {
    sum = c[0];
    t1 = 1.0;
    for (int i = 2; i <= np; i++) {
      t1 = t1 * (x - xa[i - 2]);
      sum = sum + c[i - 1] * t1;
    }
    return sum;
  }
  x0 = xa[0];
  for (int j = 1; j <= np - 1; j++)
    for (int i = 1; i <= np - j; i++) {
      int k = np - i;
      c[k - 1] = c[k - 1] + (x0 - xa[k - j]) * c[k];
    }
  if (m > 0) {
    for (int j = 1; j <= m; j++)
      for (int i = m + 1; i <= np; i++) c[i - 1] = c[i - 1] * (double)(i - j);
    t1 = c[np - 1];
    t2 = x - x0;
    for (int i = np - 1; i >= m + 1; i--) t1 = t1 * t2 + c[i - 1];
    return t1;
  }
  t1 = c[np - 1] / (double)np;
  t2 = x - x0;
  for (int i = np - 1; i >= 1; i--) t1 = t1 * t2 + c[i - 1] / (double)i;
  return t1 * t2;
}

/* src/src_lib/sphcrd.f90:13-50 */
void orc_sphcrd(const double *v, double *r, double *tp) {
  const double twopi = 6.2831853071795864769, eps = 1.0e-14;
  *r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (*r > eps) {
    double t1 = v[2] / *r;
    if (t1 > 0.0) t1 = fmin(t1, 1.0);
    else t1 = fmax(t1, -1.0);
    tp[0] = acos(t1);
    if (fabs(v[0]) > eps || fabs(v[1]) > eps) {
      tp[1] = atan2(v[1], v[0]);
      if (tp[1] < 0.0) tp[1] += twopi;
    } else {
      tp[1] = 0.0;
    }
  } else {
    tp[0] = 0.0;
    tp[1] = 0.0;
  }
}

/* ------------------------------------------------------------------ */
/* LAPACK 3.4.0 zgesv = zgetrf (unblocked zgetf2 for n < 64) + zgetrs   */
/* (external/LAPACK/zgesv.f, zgetf2.f, zgetrs.f, ztrsm.f)               */
/* ------------------------------------------------------------------ */
int orc_zgesv(int n, int nrhs, orc_z *a, int lda, int *ipiv, orc_z *b,
              int ldb) {
  const double sfmin = 2.2250738585072014e-308; /* dlamch('S') */
  int info = 0;
#define A_(i, j) a[(i) + (size_t)(j) * lda]
#define B_(i, j) b[(i) + (size_t)(j) * ldb]
  for (int j = 0; j < n; j++) {
    /* izamax over |re|+|im|, first maximum */
    int jp = j;
    double best = fabs(zre(A_(j, j))) + fabs(zim(A_(j, j)));
    for (int i = j + 1; i < n; i++) {
      double v = fabs(zre(A_(i, j))) + fabs(zim(A_(i, j)));
      if (v > best) {
        best = v;
        jp = i;
      }
    }
    ipiv[j] = jp + 1;
    if (zre(A_(jp, j)) != 0.0 || zim(A_(jp, j)) != 0.0) {
      if (jp != j)
        for (int k = 0; k < n; k++) {
          orc_z t = A_(j, k);
          A_(j, k) = A_(jp, k);
          A_(jp, k) = t;
        }
      if (j < n - 1) {
        if (cabs(A_(j, j)) >= sfmin) {
          orc_z r = zdiv(zmk(1.0, 0.0), A_(j, j));
          for (int i = j + 1; i < n; i++) A_(i, j) = zmul(r, A_(i, j));
        } else {
          for (int i = j + 1; i < n; i++) A_(i, j) = zdiv(A_(i, j), A_(j, j));
        }
      }
    } else if (info == 0) {
      info = j + 1;
    }
    if (j < n - 1) /* zgeru rank-1 update, alpha = -1 */
      for (int k = j + 1; k < n; k++) {
        orc_z y = A_(j, k);
        if (zre(y) != 0.0 || zim(y) != 0.0) {
          orc_z t = zmul(zmk(-1.0, 0.0), y);
          for (int i = j + 1; i < n; i++)
            A_(i, k) = zadd(A_(i, k), zmul(A_(i, j), t));
        }
      }
  }
  if (info != 0) return info;
  for (int r = 0; r < nrhs; r++) {
    for (int i = 0; i < n; i++) { /* zlaswp */
      int ip = ipiv[i] - 1;
      if (ip != i) {
        orc_z t = B_(i, r);
        B_(i, r) = B_(ip, r);
        B_(ip, r) = t;
      }
    }
    for (int k = 0; k < n; k++) /* ztrsm L, lower, N, unit */
      if (zre(B_(k, r)) != 0.0 || zim(B_(k, r)) != 0.0)
        for (int i = k + 1; i < n; i++)
          B_(i, r) = zsub(B_(i, r), zmul(B_(k, r), A_(i, k)));
    for (int k = n - 1; k >= 0; k--) /* ztrsm L, upper, N, non-unit */
      if (zre(B_(k, r)) != 0.0 || zim(B_(k, r)) != 0.0) {
        B_(k, r) = zdiv(B_(k, r), A_(k, k));
        for (int i = 0; i < k; i++)
          B_(i, r) = zsub(B_(i, r), zmul(B_(k, r), A_(i, k)));
      }
  }
#undef A_
#undef B_
  return 0;
}

/* ------------------------------------------------------------------ */
/* Index helpers: idxlm (init0.f90:93-100), genidxlo.f90:27-45,          */
/* MTInitAll losize (mod_eigensystem.F90:206-230), MaxAPWs (:151-184)   */
/* ------------------------------------------------------------------ */
static inline int idxlm(int l, int m) { return l * (l + 1) + m; } /* 0-based */
static inline int apword_of(const orc_system *s, int l, int is) {
  return s->apword[l + (s->lmaxapw + 1) * is];
}
static inline int lorbl_of(const orc_system *s, int ilo, int is) {
  return s->lorbl[ilo + s->nlomax * is];
}
static int species_of(const orc_system *s, int ias, int *ia_out) {
  int base = 0;
  for (int is = 0; is < s->nspecies; is++) {
    if (ias < base + s->natoms[is]) {
      if (ia_out) *ia_out = ias - base;
      return is;
    }
    base += s->natoms[is];
  }
  return -1;
}

int orc_losize(const orc_system *s, int is) {
  int n = 0;
  for (int ilo = 0; ilo < s->nlorb[is]; ilo++) n += 2 * lorbl_of(s, ilo, is) + 1;
  return n;
}

int orc_nlotot(const orc_system *s) {
  int n = 0;
  for (int is = 0; is < s->nspecies; is++) n += s->natoms[is] * orc_losize(s, is);
  return n;
}

int orc_maxapw(const orc_system *s) {
  int mx = 0;
  for (int is = 0; is < s->nspecies; is++) {
    int n = 0;
    for (int l = 0; l <= s->lmaxmat; l++) n += (2 * l + 1) * apword_of(s, l, is);
    if (n > mx) mx = n;
  }
  return mx;
}

/* 0-based first LO row of atom ias (idxlo(lm(l1,-l1),1,ias)-1) */
static int lo_base(const orc_system *s, int ias) {
  int base = 0;
  for (int j = 0; j < ias; j++) base += orc_losize(s, species_of(s, j, NULL));
  return base;
}
/* 0-based offset of (ilo, m=-l) within the atom's LO rows */
static int lo_offset(const orc_system *s, int is, int ilo) {
  int off = 0;
  for (int k = 0; k < ilo; k++) off += 2 * lorbl_of(s, k, is) + 1;
  return off;
}

/* ------------------------------------------------------------------ */
/* match.f90:13-194                                                     */
/* ------------------------------------------------------------------ */
int orc_match(const orc_system *s, int ngp, const double *gpc,
              const double *tpgpc, const orc_z *sfacgp, orc_z *apwalm) {
  const double fourpi = 12.566370614359172954;
  const int lmax = s->lmaxapw, lmmax = (lmax + 1) * (lmax + 1);
  const int pmax = s->apwordmax, ld = s->ngkmax;
  const double fpso = fourpi / sqrt(s->omega); /* :83 */
  const int np = (pmax + 1 > 4) ? pmax + 1 : 4; /* :85 */
  int info_all = 0;
  orc_z *ylm = (orc_z *)malloc(sizeof(orc_z) * (size_t)lmmax * (ngp > 0 ? ngp : 1));
  double *djl = (double *)malloc(sizeof(double) * (size_t)(lmax + 1) * pmax * (ngp > 0 ? ngp : 1));
  orc_z *zb = (orc_z *)malloc(sizeof(orc_z) * (size_t)pmax * (ngp > 0 ? ngp : 1) * (2 * lmax + 1));
  orc_z zd[16];
  int ipiv[8];
  double cwork[16];
  orc_z zil[4] = {zmk(1, 0), zmk(0, 1), zmk(-1, 0), zmk(0, -1)};
  for (int igp = 0; igp < ngp; igp++) /* :94-104 */
    orc_genylm(lmax, tpgpc + 2 * igp, ylm + (size_t)lmmax * igp);
  int ias0 = 0;
  for (int is = 0; is < s->nspecies; is++) {
    int omax = 0;
    for (int l = 0; l <= lmax; l++)
      if (apword_of(s, l, is) > omax) omax = apword_of(s, l, is);
    for (int igp = 0; igp < ngp; igp++) { /* :114-131 */
      double *d = djl + (size_t)(lmax + 1) * pmax * igp;
      double t1 = gpc[igp] * s->rmt[is];
      for (int io1 = 1; io1 <= omax; io1++)
        orc_sbesseldm(io1 - 1, lmax, t1, d + (lmax + 1) * (io1 - 1));
      t1 = 1.0;
      for (int io1 = 2; io1 <= omax; io1++) {
        t1 = t1 * gpc[igp];
        for (int l = 0; l <= lmax; l++) d[l + (lmax + 1) * (io1 - 1)] *= t1;
      }
    }
    for (int ia = 0; ia < s->natoms[is]; ia++) {
      int ias = ias0 + ia;
      for (int l = 0; l <= lmax; l++) {
        orc_z zt1 = zmul(zmk(fpso, 0.0), zil[l % 4]);
        int p = apword_of(s, l, is);
        int ir = s->nrmt[is] - np; /* 0-based nrmt-np+1 */
        for (int io2 = 0; io2 < p; io2++)
          for (int io1 = 0; io1 < p; io1++) {
            const double *ya =
                s->apwfr + ir +
                (size_t)s->nrmtmax *
                    (0 + 2 * (io2 + (size_t)pmax * (l + (size_t)(lmax + 1) * ias)));
            double v = orc_polynom(io1, np, s->spr + ir + (size_t)s->nrmtmax * is,
                                   ya, cwork, s->rmt[is]);
            zd[io1 + pmax * io2] = zmk(v, 0.0);
          }
        int i = 0;
        for (int igp = 0; igp < ngp; igp++) {
          orc_z zt2 = zmul(zt1, sfacgp[igp + (size_t)ld * ias]);
          const double *d = djl + (size_t)(lmax + 1) * pmax * igp;
          for (int m = -l; m <= l; m++) {
            orc_z yc = zconj(ylm[(size_t)lmmax * igp + idxlm(l, m)]);
            for (int io1 = 0; io1 < p; io1++) {
              orc_z t = zmul(zmk(d[l + (lmax + 1) * io1], 0.0), zt2);
              zb[io1 + (size_t)pmax * i] = zmul(t, yc);
            }
            i++;
          }
        }
        int info = orc_zgesv(p, i, zd, pmax, ipiv, zb, pmax); /* :163 */
        if (info != 0 && info_all == 0) info_all = info; /* reference prints, continues */
        i = 0;
        for (int igp = 0; igp < ngp; igp++)
          for (int m = -l; m <= l; m++) {
            int lm = idxlm(l, m);
            for (int io1 = 0; io1 < p; io1++)
              apwalm[igp + (size_t)ld * (io1 + (size_t)pmax * (lm + (size_t)lmmax * ias))] =
                  zb[io1 + (size_t)pmax * i];
            i++;
          }
      }
    }
    ias0 += s->natoms[is];
  }
  if (g_fault == 2 && ngp > 0) /* planted mistake: drop the i^l phase for l=1 */
    for (int ias = 0; ias < s->natmtot; ias++)
      for (int m = -1; m <= 1 && lmax >= 1; m++)
        for (int io = 0; io < pmax; io++)
          for (int igp = 0; igp < ngp; igp++) {
            orc_z *z = &apwalm[igp + (size_t)ld * (io + (size_t)pmax * (idxlm(1, m) + (size_t)lmmax * ias))];
            *z = zmk(zim(*z), -zre(*z));
          }
  free(ylm);
  free(djl);
  free(zb);
  return info_all;
}

/* ------------------------------------------------------------------ */
/* zgemm with netlib loop order (external/BLAS/zgemm.f), alpha = beta = 1 */
/* for the three forms the setup uses: NN, CN, CC.                      */
/* ------------------------------------------------------------------ */
static void own_zgemm(char ta, char tb, int m, int n, int k, const orc_z *a,
                      int lda, const orc_z *b, int ldb, orc_z *c, int ldc) {
#define A_(i, j) a[(i) + (size_t)(j) * lda]
#define B_(i, j) b[(i) + (size_t)(j) * ldb]
#define C_(i, j) c[(i) + (size_t)(j) * ldc]
  if (ta == 'N' && tb == 'N') {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++)
      for (int l = 0; l < k; l++) {
        orc_z t = B_(l, j);
        if (zre(t) != 0.0 || zim(t) != 0.0)
          for (int i = 0; i < m; i++) C_(i, j) = zadd(C_(i, j), zmul(t, A_(i, l)));
      }
  } else if (ta == 'C' && tb == 'N') {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < m; i++) {
        double tr = 0.0, ti = 0.0;
        for (int l = 0; l < k; l++) {
          orc_z p = zmul(zconj(A_(l, i)), B_(l, j));
          tr += zre(p);
          ti += zim(p);
        }
        C_(i, j) = zadd(zmk(tr, ti), C_(i, j));
      }
  } else { /* C, C */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < m; i++) {
        double tr = 0.0, ti = 0.0;
        for (int l = 0; l < k; l++) {
          orc_z p = zmul(zconj(A_(l, i)), zconj(B_(j, l)));
          tr += zre(p);
          ti += zim(p);
        }
        C_(i, j) = zadd(zmk(tr, ti), C_(i, j));
      }
  }
#undef A_
#undef B_
#undef C_
}

static void do_zgemm(orc_zgemm_t zg, char ta, char tb, int m, int n, int k,
                     const orc_z *a, int lda, const orc_z *b, int ldb, orc_z *c,
                     int ldc) {
  if (m <= 0 || n <= 0) return;
  if (zg) {
    orc_z one = zmk(1.0, 0.0);
    zg(&ta, &tb, &m, &n, &k, &one, a, &lda, b, &ldb, &one, c, &ldc);
  } else {
    own_zgemm(ta, tb, m, n, k, a, lda, b, ldb, c, ldc);
  }
}

/* Build apwi(maxaa, ngp): apwi(if3,:) = apwalm(1:ngp, io2, lm3, ias) with
 * if3 running over (l3 <= lmaxmat, m3, io2) io2 fastest
 * (hamiltonsetup.f90:58-71). */
static void build_apwi(const orc_system *s, int ngp, const orc_z *apwalm,
                       int ias, int is, orc_z *apwi) {
  const int lmmax = (s->lmaxapw + 1) * (s->lmaxapw + 1);
  memset(apwi, 0, sizeof(orc_z) * (size_t)s->maxaa * ngp);
  int if3 = 0;
  for (int l3 = 0; l3 <= s->lmaxmat; l3++)
    for (int m3 = -l3; m3 <= l3; m3++) {
      int lm3 = idxlm(l3, m3);
      for (int io2 = 0; io2 < apword_of(s, l3, is); io2++) {
        const orc_z *col = apwalm + (size_t)s->ngkmax *
                                        (io2 + (size_t)s->apwordmax * (lm3 + (size_t)lmmax * ias));
        for (int g = 0; g < ngp; g++) apwi[if3 + (size_t)s->maxaa * g] = col[g];
        if3++;
      }
    }
}

/* ------------------------------------------------------------------ */
/* hmlistln.f90:13-92                                                   */
/* ------------------------------------------------------------------ */
static int ivgig_lookup(const orc_system *s, const int *iv) {
  const int *lo = s->intgv, *hi = s->intgv + 3;
  for (int d = 0; d < 3; d++)
    if (iv[d] < lo[d] || iv[d] > hi[d]) return 0; /* outside the box: no term */
  int n1 = hi[0] - lo[0] + 1, n2 = hi[1] - lo[1] + 1;
  size_t idx = (size_t)(iv[0] - lo[0]) +
               (size_t)n1 * ((size_t)(iv[1] - lo[1]) + (size_t)n2 * (iv[2] - lo[2]));
  return s->ivgig[idx];
}

void orc_hmlistln(const orc_system *s, int ngp, const int *igpig,
                  const double *vgpc, int rank, orc_z *h) {
  const orc_z *kin = s->valence_rel_none ? s->cfunig : s->meffig;
#pragma omp parallel for schedule(dynamic, 16)
  for (int j = 0; j < ngp; j++)
    for (int i = 0; i <= j; i++) {
      int iv[3];
      const int *gi = s->ivg + 3 * (size_t)(igpig[i] - 1);
      const int *gj = s->ivg + 3 * (size_t)(igpig[j] - 1);
      for (int d = 0; d < 3; d++) iv[d] = gi[d] - gj[d];
      int ig = ivgig_lookup(s, iv);
      if (ig > 0 && ig <= s->ngvec) {
        const double *a = vgpc + 3 * (size_t)i, *b = vgpc + 3 * (size_t)j;
        double t1 = 0.5 * (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]);
        if (g_fault == 3) t1 = 0.5 * t1; /* planted mistake */
        orc_z zt = zadd(s->veffig[ig - 1], zmul(zmk(t1, 0.0), kin[ig - 1]));
        h[i + (size_t)rank * j] = zadd(h[i + (size_t)rank * j], zt);
      }
    }
}

/* olpistln.f90:13-85 */
void orc_olpistln(const orc_system *s, int ngp, const int *igpig,
                  const double *vgpc, int rank, orc_z *o) {
  const double alpha = 1.0 / 137.03599911, a2 = 0.5 * alpha * alpha;
#pragma omp parallel for schedule(dynamic, 16)
  for (int j = 0; j < ngp; j++)
    for (int i = 0; i <= j; i++) {
      int iv[3];
      const int *gi = s->ivg + 3 * (size_t)(igpig[i] - 1);
      const int *gj = s->ivg + 3 * (size_t)(igpig[j] - 1);
      for (int d = 0; d < 3; d++) iv[d] = gi[d] - gj[d];
      int ig = ivgig_lookup(s, iv);
      if (ig > 0 && ig <= s->ngvec) {
        orc_z zt;
        if (s->h1on) {
          const double *a = vgpc + 3 * (size_t)i, *b = vgpc + 3 * (size_t)j;
          double t1 = 0.5 * a2 * (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]);
          zt = zadd(zmul(zmk(t1, 0.0), s->m2effig[ig - 1]), s->cfunig[ig - 1]);
        } else {
          zt = s->cfunig[ig - 1];
        }
        o[i + (size_t)rank * j] = zadd(o[i + (size_t)rank * j], zt);
      }
    }
}

/* hamiltonoverlapocopy_UL.f90:6-30: lower := conj(upper) */
void orc_copy_ul(int rank, orc_z *h, orc_z *o) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < rank - 1; i++)
    for (int j = i + 1; j < rank; j++) {
      h[j + (size_t)rank * i] = zconj(h[i + (size_t)rank * j]);
      o[j + (size_t)rank * i] = zconj(o[i + (size_t)rank * j]);
    }
}

/* ------------------------------------------------------------------ */
/* hamiltonsetup.f90:8-193                                              */
/* ------------------------------------------------------------------ */
int orc_hamiltonsetup(const orc_system *s, int ngp, const orc_z *apwalm,
                      const int *igpig, const double *vgpc, int rank, orc_z *h,
                      orc_zgemm_t zgemm) {
  const int maxaa = s->maxaa, maxnlo = s->maxnlo;
  if (ngp < 0 || rank < ngp) return -1;
  orc_z *apwi = (orc_z *)malloc(sizeof(orc_z) * (size_t)maxaa * (ngp > 0 ? ngp : 1));
  orc_z *zm = (orc_z *)malloc(sizeof(orc_z) * (size_t)maxaa * (ngp > 0 ? ngp : 1));
  for (int ias = 0; ias < s->natmtot; ias++) {
    int ia, is = species_of(s, ias, &ia);
    build_apwi(s, ngp, apwalm, ias, is, apwi);
    memset(zm, 0, sizeof(orc_z) * (size_t)maxaa * ngp);
    /* :75-88 zm = aa * apwi */
    do_zgemm(zgemm, 'N', 'N', maxaa, ngp, maxaa,
             s->aa + (size_t)maxaa * maxaa * ias, maxaa, apwi, maxaa, zm, maxaa);
    /* :90-103 H += apwi^H * zm (full ngp x ngp) */
    if (g_fault != 4)
      do_zgemm(zgemm, 'C', 'N', ngp, ngp, maxaa, apwi, maxaa, zm, maxaa, h, rank);
    if (s->nlorb[is] != 0) {
      int los = orc_losize(s, is);
      int r0 = ngp + lo_base(s, ias); /* 0-based row ngp+idxlo(lm3,1,ias) */
      /* :131-145 H(r0:r0+los-1, 1:ngp) += loa * apwi */
      do_zgemm(zgemm, 'N', 'N', los, ngp, maxaa,
               s->loa + (size_t)maxnlo * maxaa * ias, maxnlo, apwi, maxaa,
               h + r0, rank);
      /* :147-149 mirror */
      for (int i = r0; i < r0 + los; i++)
        for (int g = 0; g < ngp; g++)
          h[g + (size_t)rank * i] = zconj(h[i + (size_t)rank * g]);
      /* :153-161 lo-lo block */
      const orc_z *ll = s->lolo + (size_t)maxnlo * maxnlo * ias;
      for (int jj = 0; jj < los; jj++)
        for (int ii = 0; ii < los; ii++) {
          size_t e = (size_t)(r0 + ii) + (size_t)rank * (r0 + jj);
          h[e] = zadd(h[e], ll[ii + (size_t)maxnlo * jj]);
        }
    }
  }
  free(apwi);
  free(zm);
  orc_hmlistln(s, ngp, igpig, vgpc, rank, h); /* :174 */
  return 0;
}

/* ------------------------------------------------------------------ */
/* overlapsetup.f90:8-183                                               */
/* ------------------------------------------------------------------ */
int orc_overlapsetup(const orc_system *s, int ngp, const orc_z *apwalm,
                     const int *igpig, const double *vgpc, int rank, orc_z *o,
                     orc_zgemm_t zgemm) {
  const int maxaa = s->maxaa, pmax = s->apwordmax, lmax = s->lmaxapw;
  const int lmmax = (lmax + 1) * (lmax + 1);
  if (ngp < 0 || rank < ngp) return -1;
  orc_z *apwi = (orc_z *)malloc(sizeof(orc_z) * (size_t)maxaa * (ngp > 0 ? ngp : 1));
  orc_z *zm = (orc_z *)malloc(sizeof(orc_z) * (size_t)maxaa * (ngp > 0 ? ngp : 1));
  for (int ias = 0; ias < s->natmtot; ias++) {
    int ia, is = species_of(s, ias, &ia);
    build_apwi(s, ngp, apwalm, ias, is, apwi);
    /* :82-94 zm(:, f) = conj(apwi(f,:)) + sum_io1 h1aa(io1,io2,l,ias) conj(apwi(f1,:)) */
    memset(zm, 0, sizeof(orc_z) * (size_t)maxaa * ngp);
    int if3 = 0;
    for (int l3 = 0; l3 <= s->lmaxmat; l3++)
      for (int m3 = -l3; m3 <= l3; m3++) {
        int p = apword_of(s, l3, is);
        for (int io2 = 0; io2 < p; io2++) {
          orc_z *zc = zm + (size_t)ngp * (if3 + io2);
          if (s->h1aa)
            for (int io1 = 0; io1 < p; io1++) {
              double hv = s->h1aa[io1 + pmax * (io2 + (size_t)pmax * (l3 + (size_t)(lmax + 1) * ias))];
              for (int g = 0; g < ngp; g++)
                zc[g] = zadd(zc[g], zmul(zmk(hv, 0.0),
                                         zconj(apwi[if3 + io1 + (size_t)maxaa * g])));
            }
          for (int g = 0; g < ngp; g++)
            zc[g] = zadd(zc[g], zconj(apwi[if3 + io2 + (size_t)maxaa * g]));
        }
        if3 += p;
      }
    /* :95-108 O += apwi^H * zm^H */
    do_zgemm(zgemm, 'C', 'C', ngp, ngp, maxaa, apwi, maxaa, zm, ngp, o, rank);
    if (s->nlorb[is] != 0) {
      int r0 = ngp + lo_base(s, ias);
      /* :116-133 APW-lo */
      for (int ilo = 0; ilo < s->nlorb[is]; ilo++) {
        int l = lorbl_of(s, ilo, is);
        int j1 = r0 + lo_offset(s, is, ilo), j2 = j1 + 2 * l;
        for (int io = 0; io < apword_of(s, l, is); io++) {
          size_t k = io + (size_t)pmax * (ilo + (size_t)s->nlomax * ias);
          double f = s->oalo[k] + (s->h1loa ? s->h1loa[k] : 0.0);
          for (int j = j1; j <= j2; j++) {
            int lm = idxlm(l, -l) + (j - j1);
            const orc_z *col = apwalm + (size_t)s->ngkmax * (io + (size_t)pmax * (lm + (size_t)lmmax * ias));
            for (int g = 0; g < ngp; g++) {
              size_t e = g + (size_t)rank * j;
              o[e] = zadd(o[e], zconj(zmul(col[g], zmk(f, 0.0))));
            }
          }
        }
        for (int j = j1; j <= j2; j++)
          for (int g = 0; g < ngp; g++)
            o[j + (size_t)rank * g] = zconj(o[g + (size_t)rank * j]);
      }
      /* :134-151 lo-lo, diagonal in m */
      for (int ilo1 = 0; ilo1 < s->nlorb[is]; ilo1++) {
        int l = lorbl_of(s, ilo1, is);
        for (int ilo2 = 0; ilo2 < s->nlorb[is]; ilo2++) {
          if (lorbl_of(s, ilo2, is) != l) continue;
          int j1 = r0 + lo_offset(s, is, ilo1), j2 = r0 + lo_offset(s, is, ilo2);
          size_t k = ilo1 + (size_t)s->nlomax * (ilo2 + (size_t)s->nlomax * ias);
          double f = s->ololo[k] + (s->h1lolo ? s->h1lolo[k] : 0.0);
          for (int d = 0; d <= 2 * l; d++) {
            size_t e = (size_t)(j1 + d) + (size_t)rank * (j2 + d);
            o[e] = zadd(o[e], zmk(f, 0.0));
          }
        }
      }
    }
  }
  free(apwi);
  free(zm);
  orc_olpistln(s, ngp, igpig, vgpc, rank, o); /* :160 */
  return 0;
}

int orc_hamiltonandoverlapsetup(const orc_system *s, int ngp,
                                const orc_z *apwalm, const int *igpig,
                                const double *vgpc, int rank, orc_z *h,
                                orc_z *o, orc_zgemm_t zgemm) {
  int rc = orc_hamiltonsetup(s, ngp, apwalm, igpig, vgpc, rank, h, zgemm);
  if (rc) return rc;
  rc = orc_overlapsetup(s, ngp, apwalm, igpig, vgpc, rank, o, zgemm);
  if (rc) return rc;
  orc_copy_ul(rank, h, o); /* overlapsetup.f90:164-166 */
  return 0;
}

/* gencfun.f90:96-159.  The G vector and its length follow gengvec.f90:61-70
 * and :99-100 (v = i1*b1 + i2*b2 + i3*b3 summed left to right); the loop
 * order per G (species, then atoms) is the reference's. */
void orc_gencfunig(int ng, const int *ivg, const double *bvec, double omega,
                   int nspecies, const int *natoms, const double *rmt,
                   const double *atposc, double epslat, orc_z *cfunig) {
  const double fourpi = 12.566370614359172954;
  const double t1 = fourpi / omega; /* :120 */
#pragma omp parallel for schedule(static)
  for (int ig = 0; ig < ng; ig++) {
    const double i1 = ivg[3 * ig], i2 = ivg[3 * ig + 1], i3 = ivg[3 * ig + 2];
    double v[3];
    for (int d = 0; d < 3; d++) v[d] = i1 * bvec[d] + i2 * bvec[3 + d] + i3 * bvec[6 + d];
    const double gc = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const int finite = gc > epslat; /* :123-124 */
    orc_z acc = finite ? 0.0 : 1.0; /* :126-127 */
    int ias = 0;
    for (int is = 0; is < nspecies; is++) {
      double ffacg;
      if (finite) { /* :133-137 */
        const double t2 = gc * rmt[is];
        ffacg = g_fault == 5 ? t1 * (sin(t2) + t2 * cos(t2)) / (gc * gc * gc) /* planted */
                             : t1 * (sin(t2) - t2 * cos(t2)) / (gc * gc * gc);
      } else { /* :140-143 */
        ffacg = t1 / 3.0 * rmt[is] * rmt[is] * rmt[is];
      }
      for (int ia = 0; ia < natoms[is]; ia++, ias++) { /* :147-155 */
        const double *tau = atposc + 3 * ias;
        double t2 = -(v[0] * tau[0] + v[1] * tau[1] + v[2] * tau[2]);
        if (g_fault == 6) t2 = -t2; /* planted mistake: phase sign */
        acc -= ffacg * (cos(t2) + I * sin(t2));
      }
    }
    cfunig[ig] = acc;
  }
}

/* ------------------------------------------------------------------ */
/* src/src_lib/spline.f90:13-96 (ld = 1)                                */
/* ------------------------------------------------------------------ */
void orc_spline(int n, const double *x, const double *f, double *cf) {
#define CF(k, i) cf[(k - 1) + 3 * (size_t)((i) - 1)]
#define X(i) x[(i) - 1]
#define F(i) f[(i) - 1]
  double *w = (double *)malloc(sizeof(double) * (size_t)n);
  double t1, t2, t3, t4;
  for (int i = 1; i <= n - 1; i++) { /* :31-34 */
    w[i - 1] = 1.0 / (X(i + 1) - X(i));
    CF(1, i) = w[i - 1] * (F(i + 1) - F(i));
  }
  CF(2, 1) = 1.0; /* :35-43 */
  CF(3, 1) = (CF(1, 2) - CF(1, 1)) / (X(3) - X(1));
  t1 = (X(2) - X(1)) * w[1];
  t2 = t1 * CF(2, 1);
  t3 = 1.0 / (2.0 * (t1 + 1.0));
  CF(2, 2) = t3;
  t4 = 3.0 * (CF(1, 2) - CF(1, 1)) * w[1] - t2 * CF(3, 1);
  CF(3, 2) = t4;
  for (int i = 3; i <= n - 1; i++) { /* :44-51 */
    t1 = (X(i) - X(i - 1)) * w[i - 1];
    t2 = t1 * t3;
    t3 = 1.0 / (2.0 * t1 + 2.0 - t2);
    CF(2, i) = t3;
    t4 = 3.0 * (CF(1, i) - CF(1, i - 1)) * w[i - 1] - t2 * t4;
    CF(3, i) = t4;
  }
  CF(2, n) = 1.0; /* :52-59 */
  t1 = (CF(1, n - 1) - CF(1, n - 2)) / (X(n) - X(n - 2));
  CF(3, n) = t1;
  for (int i = n - 1; i >= 2; i--) {
    t1 = CF(3, i) - t1 * CF(2, i + 1);
    CF(3, i) = t1;
  }
  for (int i = 1; i <= n; i++) CF(2, i) = CF(2, i) * CF(3, i); /* :60-62 */
  for (int i = 1; i <= n - 1; i++) { /* :63-68 */
    t1 = 0.3333333333333333333 * (CF(2, i + 1) - CF(2, i)) * w[i - 1];
    CF(3, i) = t1;
    t2 = X(i + 1) - X(i);
    CF(1, i) = CF(1, i) - (CF(2, i) + t1 * t2) * t2;
  }
  t1 = X(n) - X(n - 1); /* :69-72 */
  CF(1, n) = CF(1, n - 1) + (2.0 * CF(2, n - 1) + 3.0 * CF(3, n - 1) * t1) * t1;
  CF(3, n) = CF(3, n - 1);
  free(w);
#undef CF
#undef X
#undef F
}

/* src/src_lib/fderiv.f90:13-80, the m <= -1 branch (:64-70). */
void orc_fderiv_int(int n, const double *x, const double *f, double *g) {
  double *cf = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  orc_spline(n, x, f, cf);
  g[0] = 0.0;
  for (int i = 0; i < n - 1; i++) {
    const double dx = x[i + 1] - x[i];
    g[i + 1] = g[i] + (((0.25 * cf[3 * i + 2] * dx + 0.3333333333333333333 * cf[3 * i + 1]) * dx +
                        0.5 * cf[3 * i]) * dx + f[i]) * dx;
  }
  free(cf);
}

/* ------------------------------------------------------------------ */
/* src/math/wigner3j_symbol.f90                                         */
/* ------------------------------------------------------------------ */
double orc_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) { /* :31-80 */
  if (j1 == 0 && j2 == 0 && j3 == 0) return 1.0;
  const int l1 = j2 - j1 + j3, l2 = j1 - j2 + j3, l3 = j1 + j2 - j3;
  if (m1 + m2 + m3 != 0 || l1 < 0 || l2 < 0 || l3 < 0) return 0.0;
  const int n1 = j1 - m1, n2 = j2 + m2;
  int k1 = 0;
  if (j2 - j3 - m1 > k1) k1 = j2 - j3 - m1;
  if (j1 - j3 + m2 > k1) k1 = j1 - j3 + m2;
  int k2 = l3;
  if (n1 < k2) k2 = n1;
  if (n2 < k2) k2 = n2;
  double sgn = ((k1 + j1 + j2 + m3) % 2 == 0) ? 1.0 : -1.0;
  double add = 0.0;
  for (int k = k1; k <= k2; k++) {
    const double t1 = sgn * orc_factr(l1, l1 - n2 + k) * orc_factr(l2, l2 - n1 + k) *
                      orc_factr(l3, l3 - k);
    add = add + t1 / (orc_factnm(k, 1) * orc_factnm(n1 - k, 1) * orc_factnm(n2 - k, 1));
    sgn = -sgn;
  }
  const double t1 = orc_factr(j1 + m1, l1) * orc_factr(j2 + m2, l2) * orc_factr(j3 + m3, l3) *
                    orc_factr(j3 - m3, 1 + j1 + j2 + j3) * orc_factnm(j1 - m1, 1) *
                    orc_factnm(j2 - m2, 1);
  return add * sqrt(t1);
}

double orc_gaunt_yyy(int l1, int l2, int l3, int m1, int m2, int m3) { /* :115-152 */
  const double isqrtfourpi = 1.0 / sqrt(12.566370614359172954);
  if (m1 - m2 - m3 != 0) return 0.0;
  const int j1 = l2 - l1 + l3, j2 = l1 - l2 + l3, j3 = l1 + l2 - l3;
  if (j1 < 0 || j2 < 0 || j3 < 0) return 0.0;
  const int j = l1 + l2 + l3;
  if (j % 2 != 0) return 0.0;
  const int jh = j / 2;
  double t1 = sqrt((double)((2 * l1 + 1) * (2 * l2 + 1) * (2 * l3 + 1)) * orc_factr(j1, j + 1) *
                   orc_factnm(j2, 1) * orc_factnm(j3, 1));
  t1 = t1 * orc_factr(jh, jh - l1) / (orc_factnm(jh - l2, 1) * orc_factnm(jh - l3, 1));
  double g = t1 * isqrtfourpi * orc_wigner3j(l1, l2, l3, -m1, m2, m3);
  if (((m1 + jh) % 2 + 2) % 2 != 0) g = -g;
  return g;
}

void orc_gaunt_yry(int l1, int l2, int l3, int m1, int m2, int m3, double *out) { /* :166-194 */
  const double sqrthalf = 1.4142135623730950488 / 2.0;
  if (m2 > 0) {
    const double t1 = (m2 % 2 == 0) ? 1.0 : -1.0;
    out[0] = sqrthalf * (orc_gaunt_yyy(l1, l2, l3, m1, m2, m3) +
                         t1 * orc_gaunt_yyy(l1, l2, l3, m1, -m2, m3));
    out[1] = 0.0;
  } else if (m2 < 0) {
    const double t1 = ((-m2) % 2 == 0) ? 1.0 : -1.0;
    const double t2 = sqrthalf * (orc_gaunt_yyy(l1, l2, l3, m1, m2, m3) -
                                  t1 * orc_gaunt_yyy(l1, l2, l3, m1, -m2, m3));
    out[0] = 0.0;
    out[1] = g_fault == 7 ? t2 : -t2; /* planted mistake: sign of the imaginary part */
  } else {
    out[0] = orc_gaunt_yyy(l1, l2, l3, m1, m2, m3);
    out[1] = 0.0;
  }
}

/* ------------------------------------------------------------------ */
/* hmlint.f90:53-55: mt_kin then mt_pot for the spinless Hamiltonian    */
/* ------------------------------------------------------------------ */
int orc_hmlint(const orc_system *s, int lmaxvr, const double *veffmt, const double *lofr,
               int valence_rel_none, orc_z *aa, orc_z *loa, orc_z *lolo) {
  return orc_hmlint_parts(s, lmaxvr, veffmt, lofr, valence_rel_none, aa, loa, lolo, 3);
}

/* parts: 1 = mt_kin, 2 = mt_pot (with pot = veffmt), 3 = both (hmlint). */
int orc_hmlint_parts(const orc_system *s, int lmaxvr, const double *veffmt, const double *lofr,
                     int valence_rel_none, orc_z *aa, orc_z *loa, orc_z *lolo, int parts) {
  const int want_kin = parts & 1, want_pot = parts & 2;
  const double y00 = 0.28209479177387814347; /* constants.f90:90 */
  const double alpha = 1.0 / 137.03599911;
  const int lmmaxvr = (lmaxvr + 1) * (lmaxvr + 1);
  const int lmaxmat = s->lmaxmat, lmmaxmat = (lmaxmat + 1) * (lmaxmat + 1);
  const int nrmtmax = s->nrmtmax, p = s->apwordmax, nlomax = s->nlomax;
  const int maxaa = s->maxaa, mnlo = s->maxnlo > 1 ? s->maxnlo : 1;
  const double a = valence_rel_none ? 0.0 : 0.5 * alpha * alpha; /* mt_kin.f90:55-59 */
  /* Gaunt list gntnonz (init1.f90:563-609): for lm1 then lm3 (both <= lmmaxmat),
   * the lm2 <= lmmaxvr with |gntryy(lm2,lm3,lm1)| > 1e-20, in order. */
  const int npairs = lmmaxmat * lmmaxmat;
  int *start = (int *)calloc((size_t)npairs + 1, sizeof(int));
  int cap = 1024, nnz = 0;
  int *glm2 = (int *)malloc(sizeof(int) * cap);
  orc_z *gval = (orc_z *)malloc(sizeof(orc_z) * cap);
  for (int l1 = 0; l1 <= lmaxmat; l1++)
    for (int m1 = -l1; m1 <= l1; m1++)
      for (int l3 = 0; l3 <= lmaxmat; l3++)
        for (int m3 = -l3; m3 <= l3; m3++) {
          const int lm1 = idxlm(l1, m1), lm3 = idxlm(l3, m3);
          start[lm1 * lmmaxmat + lm3] = nnz;
          for (int l2 = 0; l2 <= lmaxvr; l2++)
            for (int m2 = -l2; m2 <= l2; m2++) {
              double g[2];
              orc_gaunt_yry(l1, l2, l3, m1, m2, m3, g); /* gntryy(lm2,lm3,lm1), :577-589 */
              if (cabs(zmk(g[0], g[1])) > 1e-20) {
                if (nnz == cap) {
                  cap *= 2;
                  glm2 = (int *)realloc(glm2, sizeof(int) * cap);
                  gval = (orc_z *)realloc(gval, sizeof(orc_z) * cap);
                }
                glm2[nnz] = idxlm(l2, m2);
                gval[nnz] = zmk(g[0], g[1]);
                nnz++;
              }
            }
        }
  start[npairs] = nnz;
  /* gather the pair list end of each (lm1, lm3) */
  int *endp = (int *)malloc(sizeof(int) * (size_t)npairs);
  for (int q = 0; q < npairs; q++) endp[q] = q + 1 < npairs ? start[q + 1] : nnz;

  double *r2 = (double *)malloc(sizeof(double) * nrmtmax);
  double *r2inv = (double *)malloc(sizeof(double) * nrmtmax);
  double *rmtable = (double *)malloc(sizeof(double) * nrmtmax);
  double *fr = (double *)malloc(sizeof(double) * nrmtmax);
  double *gr = (double *)malloc(sizeof(double) * nrmtmax);
  /* radial functions of one atom: APW (l, io) for l <= lmaxmat, then LOs */
  const int nfa = (lmaxmat + 1) * p, nf = nfa + nlomax;
  double *Ipot = (double *)malloc(sizeof(double) * (size_t)nf * nf * lmmaxvr);
  double *Ikin = (double *)malloc(sizeof(double) * (size_t)nf * nf);
#define APWFR(ir, k, io, l, ias) \
  s->apwfr[(ir) + (size_t)nrmtmax * ((k) + 2 * ((io) + (size_t)p * ((l) + (size_t)(s->lmaxapw + 1) * (ias))))]
#define LOFR(ir, k, ilo, ias) lofr[(ir) + (size_t)nrmtmax * ((k) + 2 * ((ilo) + (size_t)nlomax * (ias)))]
#define VEFF(lm, ir, ias) veffmt[(lm) + (size_t)lmmaxvr * ((ir) + (size_t)nrmtmax * (ias))]
  int ias = 0;
  for (int is = 0; is < s->nspecies; is++) {
    const int nr = s->nrmt[is];
    const double *spr = s->spr + (size_t)nrmtmax * is;
    for (int ir = 0; ir < nr; ir++) { /* mt_kin.f90:101-104 */
      r2[ir] = spr[ir] * spr[ir];
      r2inv[ir] = 1.0 / r2[ir];
    }
    for (int ia = 0; ia < s->natoms[is]; ia++, ias++) {
      for (int ir = 0; ir < nr; ir++) rmtable[ir] = 1.0 / (1.0 - a * VEFF(0, ir, ias) * y00);
      /* radial function value (k = 0) or derivative (k = 1) of function fi */
#define RF(fi, ir, k) ((fi) < nfa ? APWFR(ir, k, (fi) % p, (fi) / p, ias) : LOFR(ir, k, (fi) - nfa, ias))
#define LOF(fi) ((fi) < nfa ? (fi) / p : lorbl_of(s, (fi) - nfa, is))
#define VALID(fi) ((fi) < nfa ? ((fi) % p) < apword_of(s, (fi) / p, is) : ((fi) - nfa) < s->nlorb[is])
      /* kinetic integrals, equal l only (mt_kin.f90:117-170) */
      for (int f1 = 0; f1 < nf; f1++)
        for (int f2 = 0; f2 < nf; f2++) {
          Ikin[f1 * nf + f2] = 0.0;
          if (!VALID(f1) || !VALID(f2) || LOF(f1) != LOF(f2)) continue;
          const int l = LOF(f1);
          const double angular = (double)(l * (l + 1));
          for (int ir = 0; ir < nr; ir++) {
            const double t1 = RF(f1, ir, 0) * RF(f2, ir, 0);
            const double t2 = RF(f1, ir, 1) * RF(f2, ir, 1);
            fr[ir] = (0.5 * t2 * rmtable[ir] + 0.5 * angular * t1 * rmtable[ir] * r2inv[ir]) * r2[ir];
          }
          orc_fderiv_int(nr, spr, fr, gr);
          Ikin[f1 * nf + f2] = gr[nr - 1] / y00;
        }
      /* potential integrals (mt_pot.f90:131-218) */
      for (int f1 = 0; f1 < nf; f1++)
        for (int f2 = 0; f2 < nf; f2++) {
          if (!VALID(f1) || !VALID(f2)) continue;
          for (int lm2 = 0; lm2 < lmmaxvr; lm2++) {
            for (int ir = 0; ir < nr; ir++) {
              const double t1 = RF(f1, ir, 0) * RF(f2, ir, 0) * r2[ir];
              fr[ir] = t1 * VEFF(lm2, ir, ias);
            }
            orc_fderiv_int(nr, spr, fr, gr);
            Ipot[((size_t)f1 * nf + f2) * lmmaxvr + lm2] = gr[nr - 1];
          }
        }
      /* index of the APW function f = (l, io) and its row if = (l, m, io) */
      int if1 = 0;
      /* mt_kin.f90:173-186 then mt_pot.f90:227-258: aa */
      for (int l1 = 0; l1 <= lmaxmat; l1++)
        for (int m1 = -l1; m1 <= l1; m1++)
          for (int io1 = 0; io1 < apword_of(s, l1, is); io1++, if1++) {
            int if3 = 0;
            for (int l3 = 0; l3 <= lmaxmat; l3++)
              for (int m3 = -l3; m3 <= l3; m3++)
                for (int io2 = 0; io2 < apword_of(s, l3, is); io2++, if3++) {
                  const int fa = l1 * p + io1, fb = l3 * p + io2;
                  orc_z *e = aa + if1 + (size_t)maxaa * (if3 + (size_t)maxaa * ias);
                  if (want_kin && l1 == l3 && m1 == m3) /* kinetic: haaintegrals(1,io2,l1,io1,l1)*y00 */
                    *e = zadd(*e, zmk(Ikin[fb * nf + fa] * y00, 0.0));
                }
          }
      if1 = 0;
      for (int l1 = 0; l1 <= lmaxmat; l1++)
        for (int m1 = -l1; m1 <= l1; m1++)
          for (int io1 = 0; io1 < apword_of(s, l1, is); io1++, if1++) {
            int if3 = 0;
            for (int l3 = 0; l3 <= lmaxmat; l3++)
              for (int m3 = -l3; m3 <= l3; m3++)
                for (int io2 = 0; io2 < apword_of(s, l3, is); io2++, if3++) {
                  const int q = idxlm(l1, m1) * lmmaxmat + idxlm(l3, m3);
                  const int fa = l1 * p + io1, fb = l3 * p + io2;
                  orc_z zsum = 0.0;
                  for (int t = start[q]; t < endp[q]; t++) /* haaintegrals(lm2,io2,l3,io1,l1) */
                    zsum = zadd(zsum, zmul(gval[t], zmk(Ipot[((size_t)fb * nf + fa) * lmmaxvr + glm2[t]], 0.0)));
                  orc_z *e = aa + if1 + (size_t)maxaa * (if3 + (size_t)maxaa * ias);
                  if (want_pot) *e = zadd(*e, zsum);
                }
          }
      /* loa (mt_kin.f90:189-206, mt_pot.f90:260-287) and lolo (mt_kin.f90:208-226,
       * mt_pot.f90:289-311) */
      if (s->maxnlo > 0) {
        int r1 = 0;
        for (int ilo = 0; ilo < s->nlorb[is]; ilo++) {
          const int l1 = lorbl_of(s, ilo, is);
          for (int m1 = -l1; m1 <= l1; m1++, r1++) {
            int if3 = 0;
            for (int l3 = 0; l3 <= lmaxmat; l3++)
              for (int m3 = -l3; m3 <= l3; m3++)
                for (int io = 0; io < apword_of(s, l3, is); io++, if3++) {
                  const int fl = nfa + ilo, fb = l3 * p + io;
                  orc_z *e = loa + r1 + (size_t)mnlo * (if3 + (size_t)maxaa * ias);
                  if (want_kin && idxlm(l1, m1) == idxlm(l3, m3))
                    *e = zadd(*e, zmk(Ikin[fb * nf + fl] * y00, 0.0));
                }
          }
        }
        r1 = 0;
        for (int ilo = 0; ilo < s->nlorb[is]; ilo++) {
          const int l1 = lorbl_of(s, ilo, is);
          for (int m1 = -l1; m1 <= l1; m1++, r1++) {
            int if3 = 0;
            for (int l3 = 0; l3 <= lmaxmat; l3++)
              for (int m3 = -l3; m3 <= l3; m3++)
                for (int io = 0; io < apword_of(s, l3, is); io++, if3++) {
                  const int q = idxlm(l1, m1) * lmmaxmat + idxlm(l3, m3);
                  const int fl = nfa + ilo, fb = l3 * p + io;
                  orc_z zsum = 0.0;
                  for (int t = start[q]; t < endp[q]; t++) /* halointegrals(lm2,io,l3,ilo) */
                    zsum = zadd(zsum, zmul(gval[t], zmk(Ipot[((size_t)fl * nf + fb) * lmmaxvr + glm2[t]], 0.0)));
                  orc_z *e = loa + r1 + (size_t)mnlo * (if3 + (size_t)maxaa * ias);
                  if (want_pot) *e = zadd(*e, zsum);
                }
          }
        }
        r1 = 0;
        for (int ilo1 = 0; ilo1 < s->nlorb[is]; ilo1++) {
          const int l1 = lorbl_of(s, ilo1, is);
          for (int m1 = -l1; m1 <= l1; m1++, r1++) {
            int r3 = 0;
            for (int ilo2 = 0; ilo2 < s->nlorb[is]; ilo2++) {
              const int l3 = lorbl_of(s, ilo2, is);
              for (int m3 = -l3; m3 <= l3; m3++, r3++) {
                orc_z *e = lolo + r1 + (size_t)mnlo * (r3 + (size_t)mnlo * ias);
                if (want_kin && idxlm(l1, m1) == idxlm(l3, m3))
                  *e = zadd(*e, zmk(Ikin[(nfa + ilo1) * nf + nfa + ilo2] * y00, 0.0));
              }
            }
          }
        }
        r1 = 0;
        for (int ilo1 = 0; ilo1 < s->nlorb[is]; ilo1++) {
          const int l1 = lorbl_of(s, ilo1, is);
          for (int m1 = -l1; m1 <= l1; m1++, r1++) {
            int r3 = 0;
            for (int ilo2 = 0; ilo2 < s->nlorb[is]; ilo2++) {
              const int l3 = lorbl_of(s, ilo2, is);
              for (int m3 = -l3; m3 <= l3; m3++, r3++) {
                const int q = idxlm(l1, m1) * lmmaxmat + idxlm(l3, m3);
                orc_z zsum = 0.0;
                for (int t = start[q]; t < endp[q]; t++) /* hlolointegrals(lm2,ilo1,ilo2) */
                  zsum = zadd(zsum, zmul(gval[t], zmk(Ipot[((size_t)(nfa + ilo1) * nf + nfa + ilo2) * lmmaxvr + glm2[t]], 0.0)));
                orc_z *e = lolo + r1 + (size_t)mnlo * (r3 + (size_t)mnlo * ias);
                if (want_pot) *e = zadd(*e, zsum);
              }
            }
          }
        }
      }
#undef RF
#undef LOF
#undef VALID
    }
  }
#undef APWFR
#undef LOFR
#undef VEFF
  free(start); free(endp); free(glm2); free(gval); free(r2); free(r2inv); free(rmtable);
  free(fr); free(gr); free(Ipot); free(Ikin);
  return 0;
}

/* ------------------------------------------------------------------ */
/* mt_so.f90:11-397 (the branch with InterstitialSO false, the only one */
/* the reference runs): the spin-orbit terms of the alpha, beta and ab  */
/* blocks, accumulated.                                                 */
/* ------------------------------------------------------------------ */
int orc_mt_so(const orc_system *s, int lmaxvr, const double *pot, const double *lofr,
              int rel_level, orc_z *const *t) {
  const double y00 = 0.28209479177387814347;
  const double alpha = 1.0 / 137.03599911, a = 0.5 * alpha * alpha; /* mt_so.f90:38,75 */
  const int lmmaxvr = (lmaxvr + 1) * (lmaxvr + 1);
  const int lmaxmat = s->lmaxmat, nrmtmax = s->nrmtmax, p = s->apwordmax;
  const int nlomax = s->nlomax, maxaa = s->maxaa, mnlo = s->maxnlo > 1 ? s->maxnlo : 1;
  if (lmaxmat > 62) return 1;
  orc_z *al_aa = t[0], *al_loa = t[1], *al_lolo = t[2], *al_alo = t[3];
  orc_z *be_aa = t[4], *be_loa = t[5], *be_lolo = t[6], *be_alo = t[7];
  orc_z *ab_aa = t[8], *ab_loa = t[9], *ab_lolo = t[10], *ab_alo = t[11];
  double *rmtable = (double *)malloc(sizeof(double) * nrmtmax);
  double *rm2 = (double *)malloc(sizeof(double) * nrmtmax);
  double *fr = (double *)malloc(sizeof(double) * nrmtmax);
  double *gr = (double *)malloc(sizeof(double) * nrmtmax);
  double *cf = (double *)malloc(sizeof(double) * 3 * nrmtmax);
#define APWFR(ir, k, io, l, ias) \
  s->apwfr[(ir) + (size_t)nrmtmax * ((k) + 2 * ((io) + (size_t)p * ((l) + (size_t)(s->lmaxapw + 1) * (ias))))]
#define LOFR(ir, ilo, ias) lofr[(ir) + (size_t)nrmtmax * (2 * ((ilo) + (size_t)nlomax * (ias)))]
#define POT00(ir, ias) pot[(size_t)lmmaxvr * ((ir) + (size_t)nrmtmax * (ias))]
#define AA(x, i, j) (x)[(i) + (size_t)maxaa * ((j) + (size_t)maxaa * ias)]
#define LOA(x, i, j) (x)[(i) + (size_t)mnlo * ((j) + (size_t)maxaa * ias)]
#define LOLO(x, i, j) (x)[(i) + (size_t)mnlo * ((j) + (size_t)mnlo * ias)]
#define ALO(x, i, j) (x)[(i) + (size_t)maxaa * ((j) + (size_t)mnlo * ias)]
  int ias = 0;
  for (int is = 0; is < s->nspecies; is++) {
    const int nr = s->nrmt[is];
    const double *spr = s->spr + (size_t)nrmtmax * is;
    for (int ia = 0; ia < s->natoms[is]; ia++, ias++) {
      for (int ir = 0; ir < nr; ir++) { /* :81-93 */
        const double v = POT00(ir, ias) * y00;
        if (rel_level == 0) rmtable[ir] = a * v;
        else if (rel_level == 1) rmtable[ir] = 1.0 / (1.0 - a * v);
        else rmtable[ir] = a / ((1.0 - a * v) * (1.0 - a * v));
      }
      /* :97-101: (1/r) d/dr of rmtable, the derivative being fderiv m = 1,
       * i.e. the first spline coefficient */
      orc_spline(nr, spr, rmtable, cf);
      for (int ir = 0; ir < nr; ir++) rm2[ir] = cf[3 * ir] / spr[ir];
      /* APW row offset of l: sum over l' < l of apword(l') (2l'+1) */
      int off[64];
      off[0] = 0;
      for (int l = 1; l <= lmaxmat + 1; l++)
        off[l] = off[l - 1] + apword_of(s, l - 1, is) * (2 * l - 1);
#define RADINT(u1, u2) \
  (__extension__({ \
    for (int ir = 0; ir < nr; ir++) fr[ir] = (u1) * (u2) * rm2[ir] * spr[ir] * spr[ir]; \
    orc_fderiv_int(nr, spr, fr, gr); \
    gr[nr - 1]; \
  }))
      /* APW-APW (:108-165) */
      for (int l1 = 1; l1 <= lmaxmat; l1++) {
        const int na = apword_of(s, l1, is);
        for (int io1 = 0; io1 < na; io1++)
          for (int io2 = 0; io2 < na; io2++) {
            const double R = RADINT(APWFR(ir, 0, io1, l1, ias), APWFR(ir, 0, io2, l1, ias));
            for (int m1 = -l1; m1 <= l1; m1++) {
              const int r = off[l1] + (m1 + l1) * na + io1;
              if (m1 != l1) {
                int c = off[l1] + (m1 + l1 + 1) * na + io2, rr = r;
                if (g_fault == 8) { rr = c; c = r; } /* planted mistake: L+ in place of L- */
                AA(ab_aa, rr, c) = zadd(AA(ab_aa, rr, c),
                                        zmk(0.5 * sqrt((double)(l1 * (l1 + 1) - m1 * (m1 + 1))) * R, 0.0));
              }
              const int c = off[l1] + (m1 + l1) * na + io2;
              AA(al_aa, r, c) = zadd(AA(al_aa, r, c), zmk(0.5 * m1 * R, 0.0));
              AA(be_aa, r, c) = zadd(AA(be_aa, r, c), zmk(-0.5 * m1 * R, 0.0));
            }
          }
      }
      if (s->maxnlo > 0) {
        /* APW row, LO column: ab%alo (:174-221) */
        int lo0 = 0;
        for (int ilo = 0; ilo < s->nlorb[is]; ilo++) {
          const int l1 = lorbl_of(s, ilo, is);
          if (l1 != 0) {
            const int na = apword_of(s, l1, is);
            for (int io = 0; io < na; io++) {
              const double R = RADINT(APWFR(ir, 0, io, l1, ias), LOFR(ir, ilo, ias));
              for (int m1 = -l1; m1 < l1; m1++) {
                const int r = off[l1] + (m1 + l1) * na + io, c = lo0 + m1 + l1 + 1;
                ALO(ab_alo, r, c) = zadd(ALO(ab_alo, r, c),
                                         zmk(0.5 * sqrt((double)(l1 * (l1 + 1) - m1 * (m1 + 1))) * R, 0.0));
              }
            }
          }
          lo0 += 2 * l1 + 1;
        }
        /* LO row, APW column: ab%loa, alpha%loa, beta%loa (:223-284) */
        lo0 = 0;
        for (int ilo = 0; ilo < s->nlorb[is]; ilo++) {
          const int l1 = lorbl_of(s, ilo, is);
          if (l1 != 0) {
            const int na = apword_of(s, l1, is);
            for (int io = 0; io < na; io++) {
              const double R = RADINT(LOFR(ir, ilo, ias), APWFR(ir, 0, io, l1, ias));
              for (int m1 = -l1; m1 <= l1; m1++) {
                const int r = lo0 + m1 + l1;
                if (m1 != l1) {
                  const int c = off[l1] + (m1 + l1 + 1) * na + io;
                  LOA(ab_loa, r, c) = zadd(LOA(ab_loa, r, c),
                                           zmk(0.5 * sqrt((double)(l1 * (l1 + 1) - m1 * (m1 + 1))) * R, 0.0));
                }
                const int c = off[l1] + (m1 + l1) * na + io;
                LOA(al_loa, r, c) = zadd(LOA(al_loa, r, c), zmk(0.5 * m1 * R, 0.0));
                LOA(be_loa, r, c) = zadd(LOA(be_loa, r, c), zmk(-0.5 * m1 * R, 0.0));
              }
            }
          }
          lo0 += 2 * l1 + 1;
        }
        /* LO-LO (:291-357) */
        int r0 = 0;
        for (int ilo1 = 0; ilo1 < s->nlorb[is]; ilo1++) {
          const int l1 = lorbl_of(s, ilo1, is);
          int c0 = 0;
          for (int ilo2 = 0; ilo2 < s->nlorb[is]; ilo2++) {
            const int l3 = lorbl_of(s, ilo2, is);
            if (l1 == l3 && l1 != 0) {
              const double R = RADINT(LOFR(ir, ilo1, ias), LOFR(ir, ilo2, ias));
              for (int m1 = -l1; m1 <= l1; m1++) {
                const int r = r0 + m1 + l1;
                if (m1 != l1) {
                  const int c = c0 + m1 + l1 + 1;
                  LOLO(ab_lolo, r, c) = zadd(LOLO(ab_lolo, r, c),
                                             zmk(0.5 * sqrt((double)(l1 * (l1 + 1) - m1 * (m1 + 1))) * R, 0.0));
                }
                const int c = c0 + m1 + l1;
                LOLO(al_lolo, r, c) = zadd(LOLO(al_lolo, r, c), zmk(0.5 * m1 * R, 0.0));
                LOLO(be_lolo, r, c) = zadd(LOLO(be_lolo, r, c), zmk(-0.5 * m1 * R, 0.0));
              }
            }
            c0 += 2 * l3 + 1;
          }
          r0 += 2 * l1 + 1;
        }
        /* :361-366: alo = conjg(transpose(loa)) for alpha and beta, whole table */
        for (int i = 0; i < maxaa; i++)
          for (int j = 0; j < s->maxnlo; j++) {
            ALO(al_alo, i, j) = conj(LOA(al_loa, j, i));
            ALO(be_alo, i, j) = conj(LOA(be_loa, j, i));
          }
      }
#undef RADINT
    }
  }
#undef APWFR
#undef LOFR
#undef POT00
#undef AA
#undef LOA
#undef LOLO
#undef ALO
  free(rmtable); free(rm2); free(fr); free(gr); free(cf);
  return 0;
}
