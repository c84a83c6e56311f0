/* C entry points for functions of the reference's Fortran module
 * wigner3j_symbol (src/math/wigner3j_symbol.f90), compiled by flang into
 * oracle/_ref.  Test infrastructure only: the module functions take their
 * integer arguments by reference and return real(dp) / complex(dp) values;
 * these wrappers store the result through a pointer so that ctypes can read
 * it. */
#include <complex.h>

double _QMwigner3j_symbolPwigner3j(const int *, const int *, const int *, const int *,
                                   const int *, const int *);
double _Complex _QMwigner3j_symbolPgaunt_yry(const int *, const int *, const int *,
                                             const int *, const int *, const int *);
double _QMwigner3j_symbolPgaunt_yyy(const int *, const int *, const int *, const int *,
                                    const int *, const int *);

void ref_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3, double *out) {
  *out = _QMwigner3j_symbolPwigner3j(&j1, &j2, &j3, &m1, &m2, &m3);
}

void ref_gaunt_yyy(int l1, int l2, int l3, int m1, int m2, int m3, double *out) {
  *out = _QMwigner3j_symbolPgaunt_yyy(&l1, &l2, &l3, &m1, &m2, &m3);
}

void ref_gaunt_yry(int l1, int l2, int l3, int m1, int m2, int m3, double *out) {
  double _Complex z = _QMwigner3j_symbolPgaunt_yry(&l1, &l2, &l3, &m1, &m2, &m3);
  out[0] = creal(z);
  out[1] = cimag(z);
}
