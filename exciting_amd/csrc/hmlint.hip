// Muffin-tin Hamiltonian tables on the device: hmlint.f90:53-55 for the
// spinless main Hamiltonian, i.e. mt_kin.f90:14-237 (kinetic, diagonal in
// (l, m)) followed by mt_pot.f90:14-323 (potential, coupled through the
// complex Gaunt coefficients of init1.f90:563-609), producing
// mt_hscf%main%aa(maxaa,maxaa,natmtot), %loa(maxnlo,maxaa,natmtot) and
// %lolo(maxnlo,maxnlo,natmtot).  exh_set_mt_device installs such tables
// (and olprad's oalo/ololo, iora*'s h1 terms) from device memory.
//
// Two kernels:
//   radial    every radial integral of one atom: for each pair of radial
//             functions (APW (l, io) for l <= lmaxmat, then the LOs) and each
//             lm of the potential, the spline integral of f1 f2 r^2 V_lm;
//             plus the kinetic integral of equal-l pairs.  The spline
//             integral of fderiv.f90 (m = -1) with spline.f90 is linear in
//             the integrand, so it is a weighted sum with per-species
//             weights computed once on the host by applying that rule to unit
//             vectors.  Radial functions and a slab of potential channels are
//             staged in LDS per chunk of grid points.
//   contract  one table element per thread: the sum over the Gaunt list of
//             (lm1, lm3) of coefficient x radial integral, plus the kinetic
//             integral when lm1 = lm3, in the reference's order.
#include <cmath>
#include <cstring>
#include <vector>

#include "exh_device.h"

namespace exh {
namespace {

constexpr int kRC = 32;        // grid points per LDS chunk
constexpr int kCH = 8;         // potential channels (lm) per workgroup
constexpr int kNfMax = 63;     // radial functions per atom (APW + LO)
constexpr int kThreads = 256;    // contraction kernel
constexpr int kRadThreads = 384; // radial kernel: one pair per thread up to 27 functions

struct HmlArgs {
  int nf, nfa, p, lmaxapw, nrmtmax, nlomax, lmmaxvr, nchunk;
  double zora;   // alpha^2/2 when a mass factor is applied (mt_kin.f90:55-59)
  int kmass;     // kinetic mass factor on APW-APW pairs (1) / pairs with an LO (2)
};

__device__ __forceinline__ double radial_value(const HmlArgs &a, const double *__restrict__ apwfr,
                                               const double *__restrict__ lofr, int ias, int f,
                                               int ir, int k) {
  if (f < a.nfa) {
    const int l = f / a.p, io = f - l * a.p;
    return apwfr[ir + (size_t)a.nrmtmax *
                          (k + 2 * (io + (size_t)a.p * (l + (size_t)(a.lmaxapw + 1) * ias)))];
  }
  return lofr[ir + (size_t)a.nrmtmax * (k + 2 * ((f - a.nfa) + (size_t)a.nlomax * ias))];
}

// pair t of the upper triangle (f1 <= f2) of nf functions
__device__ __forceinline__ void pair_of(int t, int &f1, int &f2) {
  int b = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((b + 1) * (b + 2) / 2 <= t) b++;
  while (b * (b + 1) / 2 > t) b--;
  f2 = b;
  f1 = t - b * (b + 1) / 2;
}

// blockIdx.x = atom, blockIdx.y = potential slab (lm0 = y * kCH) or, for
// y == nchunk, the kinetic integrals.  ipot[ias][f1][f2][lm], ikin[ias][f1][f2],
// both written for (f1,f2) and (f2,f1).
template <int PPT>
__global__ __launch_bounds__(kRadThreads) void hml_radial_kernel(
    HmlArgs a, const int *__restrict__ atom_species, const int *__restrict__ nrmt,
    const double *__restrict__ wts, const double *__restrict__ spr,
    const int *__restrict__ fl, const double *__restrict__ apwfr,
    const double *__restrict__ lofr, const double *__restrict__ veffmt,
    double *__restrict__ ipot, double *__restrict__ ikin) {
  __shared__ double sF[kRC][kNfMax + 1];
  __shared__ double sD[kRC][kNfMax + 1];
  __shared__ double sV[kRC][kCH];
  __shared__ double sW[kRC], sR2[kRC], sRM[kRC];
  const int ias = blockIdx.x, is = atom_species[ias], nr = nrmt[is];
  const bool kin = (int)blockIdx.y == a.nchunk;
  const int lm0 = blockIdx.y * kCH;
  const int nf = a.nf, npair = nf * (nf + 1) / 2, tid = threadIdx.x;
  const double y00 = 0.28209479177387814347;
  double acc[PPT][kCH];
#pragma unroll
  for (int q = 0; q < PPT; q++)
#pragma unroll
    for (int k = 0; k < kCH; k++) acc[q][k] = 0.0;
  int pf1[PPT], pf2[PPT];
  double lfac[PPT];
#pragma unroll
  for (int q = 0; q < PPT; q++) {
    const int t = tid + q * kRadThreads;
    pf1[q] = pf2[q] = 0;
    lfac[q] = 0.0;
    if (t < npair) {
      pair_of(t, pf1[q], pf2[q]);
      const int l = fl[is * kNfMax + pf1[q]];
      lfac[q] = (double)(l * (l + 1));
    }
  }
  for (int r0 = 0; r0 < nr; r0 += kRC) {
    for (int i = tid; i < kRC * nf; i += kRadThreads) {
      const int r = i / nf, f = i - r * nf, ir = r0 + r;
      sF[r][f] = ir < nr ? radial_value(a, apwfr, lofr, ias, f, ir, 0) : 0.0;
      if (kin) sD[r][f] = ir < nr ? radial_value(a, apwfr, lofr, ias, f, ir, 1) : 0.0;
    }
    if (kin) {
      for (int r = tid; r < kRC; r += kRadThreads) {
        const int ir = r0 + r;
        double w = 0.0, r2 = 0.0, rm = 1.0;
        if (ir < nr) {
          const double x = spr[ir + (size_t)a.nrmtmax * is];
          r2 = x * x;
          w = 0.5 * wts[ir + (size_t)a.nrmtmax * is];
          if (a.kmass) {
            const double v00 = veffmt[(size_t)a.lmmaxvr * (ir + (size_t)a.nrmtmax * ias)];
            rm = 1.0 / (1.0 - a.zora * v00 * y00);
          }
        }
        sW[r] = w;
        sR2[r] = r2;
        sRM[r] = rm;
      }
    } else {
      for (int i = tid; i < kRC * kCH; i += kRadThreads) {
        const int r = i / kCH, k = i - r * kCH, ir = r0 + r, lm = lm0 + k;
        sV[r][k] = (ir < nr && lm < a.lmmaxvr)
                       ? veffmt[lm + (size_t)a.lmmaxvr * (ir + (size_t)a.nrmtmax * ias)]
                       : 0.0;
      }
      for (int r = tid; r < kRC; r += kRadThreads) {
        const int ir = r0 + r;
        double w = 0.0;
        if (ir < nr) {
          const double x = spr[ir + (size_t)a.nrmtmax * is];
          w = wts[ir + (size_t)a.nrmtmax * is] * (x * x);
        }
        sW[r] = w;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PPT; q++) {
      if (tid + q * kRadThreads >= npair) continue;
      const int f1 = pf1[q], f2 = pf2[q];
      if (kin) {  // acc[q][0]: plain, acc[q][1]: with the mass factor
        double s0 = acc[q][0], s1 = acc[q][1];
        for (int r = 0; r < kRC; r++) {
          const double t =
              sW[r] * (sD[r][f1] * sD[r][f2] * sR2[r] + lfac[q] * (sF[r][f1] * sF[r][f2]));
          s0 += t;
          s1 += t * sRM[r];
        }
        acc[q][0] = s0;
        acc[q][1] = s1;
      } else {
        for (int r = 0; r < kRC; r++) {
          const double pr = sF[r][f1] * sF[r][f2] * sW[r];
#pragma unroll
          for (int k = 0; k < kCH; k++) acc[q][k] += pr * sV[r][k];
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < PPT; q++) {
    if (tid + q * kRadThreads >= npair) continue;
    const int f1 = pf1[q], f2 = pf2[q];
    if (kin) {
      const bool same = fl[is * kNfMax + f1] == fl[is * kNfMax + f2];
      const bool lo = f1 >= a.nfa || f2 >= a.nfa;
      const bool mass = a.kmass & (lo ? 2 : 1);
      const double v = same ? (mass ? acc[q][1] : acc[q][0]) : 0.0;
      ikin[((size_t)ias * nf + f1) * nf + f2] = v;
      ikin[((size_t)ias * nf + f2) * nf + f1] = v;
    } else {
#pragma unroll
      for (int k = 0; k < kCH; k++) {
        const int lm = lm0 + k;
        if (lm >= a.lmmaxvr) break;
        ipot[(((size_t)ias * nf + f1) * nf + f2) * a.lmmaxvr + lm] = acc[q][k];
        ipot[(((size_t)ias * nf + f2) * nf + f1) * a.lmmaxvr + lm] = acc[q][k];
      }
    }
  }
}

// Rows of one table kind per species: int2 {lm (0-based), radial function}.
struct RowTable {
  const int2 *rows;    // [nspecies][ld]
  const int *count;    // [nspecies]
  int ld;
};

// out(i, j, ias) (+)= [lm_i == lm_j] kin(f_i, f_j) + sum_t g_t ipot(f_i, f_j, lm2_t)
// over the Gaunt list of (lm_i, lm_j): aa (APW x APW), loa (LO x APW) and
// lolo (LO x LO) of mt_kin.f90:173-226 and mt_pot.f90:227-311.
__global__ __launch_bounds__(kThreads) void hml_contract_kernel(
    int nf, int lmmaxvr, int lmmaxmat, const int *__restrict__ atom_species, RowTable R,
    RowTable C, int ldo, int nco, const int *__restrict__ gstart, const int *__restrict__ glm2,
    const double2 *__restrict__ gval, const double *__restrict__ ipot,
    const double *__restrict__ ikin, double2 *__restrict__ out, int accumulate, int with_kin) {
  const int ias = blockIdx.y;
  const long long e = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (e >= (long long)ldo * nco) return;
  const int i = (int)(e % ldo), j = (int)(e / ldo);
  const int is = atom_species[ias];
  double2 *o = out + (size_t)i + (size_t)ldo * (j + (size_t)nco * ias);
  double re = 0.0, im = 0.0;
  if (accumulate) {
    re = o->x;
    im = o->y;
  }
  if (i < R.count[is] && j < C.count[is]) {
    const int2 ri = R.rows[is * R.ld + i], cj = C.rows[is * C.ld + j];
    const size_t base = ((size_t)ias * nf + ri.y) * nf + cj.y;
    if (with_kin && ri.x == cj.x) re += ikin[base];
    double zr = 0.0, zi = 0.0;
    const int q = ri.x * lmmaxmat + cj.x;
    const double *ip = ipot + base * lmmaxvr;
    for (int t = gstart[q]; t < gstart[q + 1]; t++) {
      const double v = ip[glm2[t]];
      const double2 g = gval[t];
      zr += g.x * v;
      zi += g.y * v;
    }
    re += zr;
    im += zi;
  }
  *o = make_double2(re, im);
}

// ---- host-side tables --------------------------------------------------

// Cubic spline coefficients cf(3, n) of f on the grid x, with the steps of
// spline.f90 (natural-like end conditions, cf(1,n) and cf(3,n) extrapolated
// from the last interval): the same rule has to give the same numbers.  iw
// holds n scratch values.
void spline_coefficients(int n, const double *x, const double *f, double *cf, double *iw) {
  for (int i = 0; i + 1 < n; i++) {
    iw[i] = 1.0 / (x[i + 1] - x[i]);
    cf[3 * i] = iw[i] * (f[i + 1] - f[i]);
  }
  cf[1] = 1.0;
  cf[2] = (cf[3] - cf[0]) / (x[2] - x[0]);
  double t1 = (x[1] - x[0]) * iw[1];
  double t2 = t1 * cf[1];
  double t3 = 1.0 / (2.0 * (t1 + 1.0));
  cf[3 + 1] = t3;
  double t4 = 3.0 * (cf[3] - cf[0]) * iw[1] - t2 * cf[2];
  cf[3 + 2] = t4;
  for (int i = 2; i + 1 < n; i++) {
    t1 = (x[i] - x[i - 1]) * iw[i];
    t2 = t1 * t3;
    t3 = 1.0 / (2.0 * t1 + 2.0 - t2);
    cf[3 * i + 1] = t3;
    t4 = 3.0 * (cf[3 * i] - cf[3 * (i - 1)]) * iw[i] - t2 * t4;
    cf[3 * i + 2] = t4;
  }
  cf[3 * (n - 1) + 1] = 1.0;
  t1 = (cf[3 * (n - 2)] - cf[3 * (n - 3)]) / (x[n - 1] - x[n - 3]);
  cf[3 * (n - 1) + 2] = t1;
  for (int i = n - 2; i >= 1; i--) {
    t1 = cf[3 * i + 2] - t1 * cf[3 * (i + 1) + 1];
    cf[3 * i + 2] = t1;
  }
  for (int i = 0; i < n; i++) cf[3 * i + 1] *= cf[3 * i + 2];
  for (int i = 0; i + 1 < n; i++) {
    t1 = (1.0 / 3.0) * (cf[3 * (i + 1) + 1] - cf[3 * i + 1]) * iw[i];
    cf[3 * i + 2] = t1;
    t2 = x[i + 1] - x[i];
    cf[3 * i] = cf[3 * i] - (cf[3 * i + 1] + t1 * t2) * t2;
  }
  t1 = x[n - 1] - x[n - 2];
  cf[3 * (n - 1)] = cf[3 * (n - 2)] + (2.0 * cf[3 * (n - 2) + 1] + 3.0 * cf[3 * (n - 2) + 2] * t1) * t1;
  cf[3 * (n - 1) + 2] = cf[3 * (n - 2) + 2];
}

// Weights w with sum_i w_i f_i equal to the integral over the whole grid that
// fderiv.f90 (m = -1) returns: its spline (spline.f90) and the per-interval
// cubic integral are linear in f, so w_j is that rule applied to the j-th
// unit vector.
void spline_integral_weights(int n, const double *x, double *w) {
  std::vector<double> cf(3 * (size_t)n), iw(n), f(n);
  for (int j = 0; j < n; j++) {
    std::fill(f.begin(), f.end(), 0.0);
    f[j] = 1.0;
    spline_coefficients(n, x, f.data(), cf.data(), iw.data());
    double g = 0.0;
    for (int i = 0; i + 1 < n; i++) {
      const double dx = x[i + 1] - x[i];
      g += (((0.25 * cf[3 * i + 2] * dx + (1.0 / 3.0) * cf[3 * i + 1]) * dx + 0.5 * cf[3 * i]) *
                dx + f[i]) * dx;
    }
    w[j] = g;
  }
}

// The first derivative at the grid points that fderiv.f90 (m = 1) returns,
// cf(1, :) of the spline, as a matrix: dt[j * ld + i] is the derivative at
// point i of the j-th unit vector (stored so that a workgroup reading one j
// reads consecutive i).
void spline_derivative_matrix(int n, const double *x, int ld, double *dt) {
  std::vector<double> cf(3 * (size_t)n), iw(n), f(n);
  for (int j = 0; j < n; j++) {
    std::fill(f.begin(), f.end(), 0.0);
    f[j] = 1.0;
    spline_coefficients(n, x, f.data(), cf.data(), iw.data());
    for (int i = 0; i < n; i++) dt[(size_t)j * ld + i] = cf[3 * i];
  }
}

// Wigner 3j symbol by Racah's formula, in long double.
long double fact(int n) {
  long double r = 1.0L;
  for (int i = 2; i <= n; i++) r *= i;
  return r;
}

long double w3j(int j1, int j2, int j3, int m1, int m2, int m3) {
  if (m1 + m2 + m3 != 0) return 0.0L;
  if (j3 < std::abs(j1 - j2) || j3 > j1 + j2) return 0.0L;
  if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0.0L;
  const long double tri = fact(j1 + j2 - j3) * fact(j1 - j2 + j3) * fact(-j1 + j2 + j3) /
                          fact(j1 + j2 + j3 + 1);
  const long double pre = std::sqrt(tri * fact(j1 + m1) * fact(j1 - m1) * fact(j2 + m2) *
                                    fact(j2 - m2) * fact(j3 + m3) * fact(j3 - m3));
  const int kmin = std::max(0, std::max(j2 - j3 - m1, j1 - j3 + m2));
  const int kmax = std::min(j1 + j2 - j3, std::min(j1 - m1, j2 + m2));
  long double sum = 0.0L;
  for (int k = kmin; k <= kmax; k++) {
    const long double d = fact(k) * fact(j3 - j2 + k + m1) * fact(j3 - j1 + k - m2) *
                          fact(j1 + j2 - j3 - k) * fact(j1 - k - m1) * fact(j2 - k + m2);
    sum += ((k % 2) ? -1.0L : 1.0L) / d;
  }
  const int ph = j1 - j2 - m3;
  return ((ph % 2 + 2) % 2 ? -1.0L : 1.0L) * pre * sum;
}

// Integral of conj(Y_l1m1) Y_l2m2 Y_l3m3 over the sphere.
long double gaunt_complex(int l1, int l2, int l3, int m1, int m2, int m3) {
  if (m1 != m2 + m3) return 0.0L;
  const long double pi = 3.14159265358979323846264338327950288L;
  const long double f =
      std::sqrt((long double)((2 * l1 + 1) * (2 * l2 + 1) * (2 * l3 + 1)) / (4.0L * pi));
  return ((m1 % 2 + 2) % 2 ? -1.0L : 1.0L) * f * w3j(l1, l2, l3, 0, 0, 0) *
         w3j(l1, l2, l3, -m1, m2, m3);
}

// Same with the real harmonic R_l2m2 of genrlm.f90 (sqrt(2) Re Y_lm for m > 0,
// Re Y_l0, sqrt(2) Im Y_lm for m < 0) in the middle.
void gaunt_real_middle(int l1, int l2, int l3, int m1, int m2, int m3, double *re, double *im) {
  const long double h = std::sqrt(0.5L);
  const long double s = (std::abs(m2) % 2) ? -1.0L : 1.0L;
  *re = *im = 0.0;
  if (m2 > 0) {
    *re = (double)(h * (gaunt_complex(l1, l2, l3, m1, m2, m3) +
                        s * gaunt_complex(l1, l2, l3, m1, -m2, m3)));
  } else if (m2 < 0) {
    *im = (double)(-h * (gaunt_complex(l1, l2, l3, m1, m2, m3) -
                         s * gaunt_complex(l1, l2, l3, m1, -m2, m3)));
  } else {
    *re = (double)gaunt_complex(l1, l2, l3, m1, 0, m3);
  }
}

int cdiv_ll(long long a, int b) { return (int)((a + b - 1) / b); }

// Species tables, Gaunt list and spline weights for (lmaxvr, the basis, the
// radial grid): rebuilt when any of them changed.
int prepare(exh_context *c, int lmaxvr) {
  const int p = c->apwordmax, lmaxmat = c->lmaxmat;
  const int nfa = (lmaxmat + 1) * p, nf = nfa + c->nlomax;
  if (nf > kNfMax) return fail(c, EXH_ERANGE, "hmlint: more than 63 radial functions per atom");
  for (int is = 0; is < c->nspecies; is++)
    for (int ilo = 0; ilo < c->nlorb[is]; ilo++)
      if (c->lorbl[ilo + c->nlomax * is] > lmaxmat)
        return fail(c, EXH_ERANGE, "hmlint: a local orbital has l > lmaxmat");
  if (c->hml_lmaxvr == lmaxvr && c->hml_radial_epoch == c->radial_epoch &&
      c->hml_basis_epoch == c->basis_epoch)
    return 0;
  const int ns = c->nspecies, lmmaxmat = (lmaxmat + 1) * (lmaxmat + 1);
  // l of each radial function per species (-1 = not a function of the species)
  std::vector<int> fl((size_t)ns * kNfMax, -1);
  std::vector<int2> arow((size_t)ns * c->maxaa), lrow((size_t)ns * std::max(c->maxnlo, 1));
  std::vector<int> acount(ns), lcount(ns);
  for (int is = 0; is < ns; is++) {
    for (int l = 0; l <= lmaxmat; l++)
      for (int io = 0; io < c->apword[l + (c->lmaxapw + 1) * is]; io++) fl[is * kNfMax + l * p + io] = l;
    int na = 0;
    for (int l = 0; l <= lmaxmat; l++)
      for (int m = -l; m <= l; m++)
        for (int io = 0; io < c->apword[l + (c->lmaxapw + 1) * is]; io++)
          arow[(size_t)is * c->maxaa + na++] = make_int2(l * (l + 1) + m, l * p + io);
    acount[is] = na;
    int nl = 0;
    for (int ilo = 0; ilo < c->nlorb[is]; ilo++) {
      const int l = c->lorbl[ilo + c->nlomax * is];
      fl[is * kNfMax + nfa + ilo] = l;
      for (int m = -l; m <= l; m++)
        lrow[(size_t)is * std::max(c->maxnlo, 1) + nl++] = make_int2(l * (l + 1) + m, nfa + ilo);
    }
    lcount[is] = nl;
  }
  // Gaunt list: for lm1, then lm3 (both <= lmmaxmat), the lm2 <= lmmaxvr with a
  // nonzero coefficient, in increasing lm2 (the order of init1.f90:592-609)
  std::vector<int> gstart((size_t)lmmaxmat * lmmaxmat + 1), glm2;
  std::vector<double2> gval;
  for (int l1 = 0; l1 <= lmaxmat; l1++)
    for (int m1 = -l1; m1 <= l1; m1++)
      for (int l3 = 0; l3 <= lmaxmat; l3++)
        for (int m3 = -l3; m3 <= l3; m3++) {
          gstart[(size_t)(l1 * (l1 + 1) + m1) * lmmaxmat + l3 * (l3 + 1) + m3] = (int)glm2.size();
          for (int l2 = 0; l2 <= lmaxvr; l2++)
            for (int m2 = -l2; m2 <= l2; m2++) {
              double re, im;
              gaunt_real_middle(l1, l2, l3, m1, m2, m3, &re, &im);
              if (std::sqrt(re * re + im * im) > 1e-20) {
                glm2.push_back(l2 * (l2 + 1) + m2);
                gval.push_back(make_double2(re, im));
              }
            }
        }
  gstart.back() = (int)glm2.size();
  // spline weights per species
  std::vector<double> w((size_t)c->nrmtmax * ns, 0.0);
  for (int is = 0; is < ns; is++) {
    if (c->nrmt_h[is] < 3) return fail(c, EXH_EINVAL, "hmlint: fewer than 3 radial points");
    spline_integral_weights(c->nrmt_h[is], c->spr_h.data() + (size_t)c->nrmtmax * is,
                            w.data() + (size_t)c->nrmtmax * is);
  }
  std::vector<int> sp_int(c->nrmt_h.begin(), c->nrmt_h.end());
  auto put = [&](DevBuf &b, const void *src, size_t bytes) {
    int rc = ensure(c, b, bytes);
    if (rc) return rc;
    hipError_t e = hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice);
    return e == hipSuccess ? 0 : hip_fail(c, e, "hmlint tables");
  };
  int rc = put(c->hml_fl, fl.data(), fl.size() * sizeof(int));
  if (!rc) rc = put(c->hml_arow, arow.data(), arow.size() * sizeof(int2));
  if (!rc) rc = put(c->hml_lrow, lrow.data(), lrow.size() * sizeof(int2));
  std::vector<int> counts(acount);
  counts.insert(counts.end(), lcount.begin(), lcount.end());
  counts.insert(counts.end(), sp_int.begin(), sp_int.end());
  if (!rc) rc = put(c->hml_count, counts.data(), counts.size() * sizeof(int));
  if (!rc) rc = put(c->hml_gstart, gstart.data(), gstart.size() * sizeof(int));
  if (!rc) rc = put(c->hml_glm2, glm2.data(), std::max<size_t>(glm2.size(), 1) * sizeof(int));
  if (!rc) rc = put(c->hml_gval, gval.data(), std::max<size_t>(gval.size(), 1) * sizeof(double2));
  if (!rc) rc = put(c->hml_w, w.data(), w.size() * sizeof(double));
  std::vector<double> spr(c->spr_h);
  if (!rc) rc = put(c->hml_spr, spr.data(), spr.size() * sizeof(double));
  if (rc) return rc;
  c->hml_lmaxvr = lmaxvr;
  c->hml_radial_epoch = c->radial_epoch;
  c->hml_basis_epoch = c->basis_epoch;
  return 0;
}

template <typename T>
int copy_in(exh_context *c, DevBuf &b, const T *src, size_t n) {
  int rc = ensure(c, b, n * sizeof(T));
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? 0 : hip_fail(c, e, "upload");
}

int copy_out(exh_context *c, void *dst, const DevBuf &b, size_t bytes) {
  hipError_t e = hipMemcpyAsync(dst, b.p, bytes, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? 0 : hip_fail(c, e, "download");
}

__global__ void add_tables_kernel(size_t n, const double *__restrict__ a,
                                  const double *__restrict__ b, double *__restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a ? a[i] + (b ? b[i] : 0.0) : 0.0;
}

// ---- spin-orbit terms (mt_so.f90) ----------------------------------------

// The twelve tables of one call: alpha (aa, loa, lolo, alo), beta, ab.
struct SoTables {
  double2 *t[12];
};

struct SoArgs {
  int nfa, p, lmaxapw, nrmtmax, nlomax, lmmaxvr, maxaa, mnlo, rel;
  double a;   // alpha^2/2 (mt_so.f90:38,75)
};

constexpr int kSoThreads = 256;

// One workgroup per atom.  LDS: rm (the relativistic factor of mt_so.f90:81-93
// on the grid) and q = w r^2 (1/r) d rm/dr, w the spline integration weights,
// so that the radial integral of a pair is sum_r q u1 u2.  Each thread then
// takes same-l pairs of radial functions (f1, f2, l >= 1) and adds, for every
// m, the L.S elements: alpha += m R/2, beta -= m R/2 on (m, m), ab +=
// sqrt(l(l+1) - m(m+1)) R/2 on (m, m+1).  Every table element is written by
// one (pair, m) only.
__global__ __launch_bounds__(kSoThreads) void so_kernel(
    SoArgs A, const int *__restrict__ atom_species, const int *__restrict__ nrmt,
    const double *__restrict__ w, const double *__restrict__ spr, const double *__restrict__ dt,
    const int4 *__restrict__ pairs, const int *__restrict__ pair_start,
    const int2 *__restrict__ rowb, const double *__restrict__ apwfr,
    const double *__restrict__ lofr, const double *__restrict__ pot, SoTables T) {
  extern __shared__ double so_lds[];
  double *rm = so_lds, *q = so_lds + A.nrmtmax;
  const int ias = blockIdx.x, is = atom_species[ias], nr = nrmt[is];
  const double *x = spr + (size_t)A.nrmtmax * is;
  constexpr double kY00 = 0.28209479177387814347;
  for (int ir = threadIdx.x; ir < nr; ir += blockDim.x) {
    const double v = A.a * pot[(size_t)A.lmmaxvr * (ir + (size_t)A.nrmtmax * ias)] * kY00;
    rm[ir] = A.rel == 0 ? v : A.rel == 1 ? 1.0 / (1.0 - v) : A.a / ((1.0 - v) * (1.0 - v));
  }
  __syncthreads();
  const double *D = dt + (size_t)A.nrmtmax * A.nrmtmax * is;
  const double *ws = w + (size_t)A.nrmtmax * is;
  // The spline maps a constant to a zero derivative, so D is applied to
  // rm - rm(nr): rm is near 1 for zora and the entries of D reach 1/dx, and
  // without the shift the sum would cancel to a few digits near the origin.
  const double c0 = rm[nr - 1];
  for (int ir = threadIdx.x; ir < nr; ir += blockDim.x) {
    double d = 0.0;
    for (int j = 0; j < nr; j++) d += D[(size_t)j * A.nrmtmax + ir] * (rm[j] - c0);
    q[ir] = ws[ir] * (d / x[ir]) * x[ir] * x[ir];
  }
  __syncthreads();
  for (int t = pair_start[is] + threadIdx.x; t < pair_start[is + 1]; t += blockDim.x) {
    const int4 P = pairs[t];   // f1, f2, l, kind (1: f1 is an LO, 2: f2 is an LO)
    double R = 0.0;
    for (int ir = 0; ir < nr; ir++) {
      const double u1 = P.x < A.nfa
          ? apwfr[ir + (size_t)A.nrmtmax * 2 * (P.x % A.p + (size_t)A.p * (P.x / A.p + (size_t)(A.lmaxapw + 1) * ias))]
          : lofr[ir + (size_t)A.nrmtmax * 2 * ((P.x - A.nfa) + (size_t)A.nlomax * ias)];
      const double u2 = P.y < A.nfa
          ? apwfr[ir + (size_t)A.nrmtmax * 2 * (P.y % A.p + (size_t)A.p * (P.y / A.p + (size_t)(A.lmaxapw + 1) * ias))]
          : lofr[ir + (size_t)A.nrmtmax * 2 * ((P.y - A.nfa) + (size_t)A.nlomax * ias)];
      R += q[ir] * u1 * u2;
    }
    const int l = P.z, kind = P.w;
    const int2 b1 = rowb[is * kNfMax + P.x], b2 = rowb[is * kNfMax + P.y];
    const int ld = (kind & 1) ? A.mnlo : A.maxaa, nc = (kind & 2) ? A.mnlo : A.maxaa;
    // alpha/beta/ab table of this kind: aa (0), loa (1), lolo (3); ab also alo (2)
    const int slot = kind == 0 ? 0 : kind == 1 ? 1 : kind == 3 ? 2 : 3;
    double2 *al = slot < 3 ? T.t[slot] : nullptr, *be = slot < 3 ? T.t[4 + slot] : nullptr;
    double2 *ab = T.t[8 + slot];
    for (int m = -l; m <= l; m++) {
      const size_t r = b1.x + (size_t)(m + l) * b1.y;
      if (al) {
        const size_t e = r + (size_t)ld * (b2.x + (size_t)(m + l) * b2.y + (size_t)nc * ias);
        al[e].x += 0.5 * m * R;
        be[e].x -= 0.5 * m * R;
      }
      if (m < l) {
        const size_t e = r + (size_t)ld * (b2.x + (size_t)(m + l + 1) * b2.y + (size_t)nc * ias);
        ab[e].x += 0.5 * sqrt((double)(l * (l + 1) - m * (m + 1))) * R;
      }
    }
  }
}

// alo(i, j, ias) = conjg(loa(j, i, ias)) for alpha and beta (mt_so.f90:361-366).
__global__ void so_alo_kernel(int maxaa, int maxnlo, int mnlo, int natm, SoTables T) {
  const long long n = (long long)maxaa * maxnlo * natm;
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = (int)(k % maxaa), j = (int)((k / maxaa) % maxnlo), ias = (int)(k / ((long long)maxaa * maxnlo));
  const size_t alo = i + (size_t)maxaa * (j + (size_t)mnlo * ias);
  const size_t loa = j + (size_t)mnlo * (i + (size_t)maxaa * ias);
  const double2 a = T.t[1][loa], b = T.t[5][loa];
  T.t[3][alo] = make_double2(a.x, -a.y);
  T.t[7][alo] = make_double2(b.x, -b.y);
}

__global__ void zero_kernel(size_t n, double2 *__restrict__ p) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_double2(0.0, 0.0);
}

// Per species: the derivative matrices, the same-l pair list and the row base
// and m stride of every radial function (APW (l, io): off(l) + io, stride
// apword(l); LO ilo: its first row, stride 1).  Rebuilt when the basis or
// the radial grid changed.
int so_prepare(exh_context *c) {
  if (int rc = prepare(c, c->hml_lmaxvr >= 0 ? c->hml_lmaxvr : 0)) return rc;
  if (c->so_radial_epoch == c->radial_epoch && c->so_basis_epoch == c->basis_epoch) return 0;
  const int ns = c->nspecies, p = c->apwordmax, lmaxmat = c->lmaxmat, nrm = c->nrmtmax;
  const int nfa = (lmaxmat + 1) * p;
  std::vector<double> dt((size_t)ns * nrm * nrm, 0.0);
  for (int is = 0; is < ns; is++)
    spline_derivative_matrix(c->nrmt_h[is], c->spr_h.data() + (size_t)nrm * is, nrm,
                             dt.data() + (size_t)nrm * nrm * is);
  std::vector<int2> rowb((size_t)ns * kNfMax, make_int2(0, 0));
  std::vector<int> fl((size_t)ns * kNfMax, -1);
  std::vector<int4> pairs;
  std::vector<int> start(ns + 1, 0);
  for (int is = 0; is < ns; is++) {
    const int *aw = c->apword.data() + (c->lmaxapw + 1) * is;
    int off = 0;
    for (int l = 0; l <= lmaxmat; l++) {
      for (int io = 0; io < aw[l]; io++) {
        rowb[is * kNfMax + l * p + io] = make_int2(off + io, aw[l]);
        fl[is * kNfMax + l * p + io] = l;
      }
      off += aw[l] * (2 * l + 1);
    }
    int lo0 = 0;
    for (int ilo = 0; ilo < c->nlorb[is]; ilo++) {
      const int l = c->lorbl[ilo + c->nlomax * is];
      rowb[is * kNfMax + nfa + ilo] = make_int2(lo0, 1);
      fl[is * kNfMax + nfa + ilo] = l;
      lo0 += 2 * l + 1;
    }
    start[is] = (int)pairs.size();
    for (int f1 = 0; f1 < kNfMax; f1++)
      for (int f2 = 0; f2 < kNfMax; f2++) {
        const int l1 = fl[is * kNfMax + f1], l2 = fl[is * kNfMax + f2];
        if (l1 < 1 || l1 != l2) continue;
        pairs.push_back(make_int4(f1, f2, l1, (f1 >= nfa ? 1 : 0) | (f2 >= nfa ? 2 : 0)));
      }
  }
  start[ns] = (int)pairs.size();
  auto put = [&](DevBuf &b, const void *src, size_t bytes) {
    int rc = ensure(c, b, std::max<size_t>(bytes, 16));
    if (rc || bytes == 0) return rc;
    hipError_t e = hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice);
    return e == hipSuccess ? 0 : hip_fail(c, e, "mt_so tables");
  };
  int rc = put(c->so_dt, dt.data(), dt.size() * sizeof(double));
  if (!rc) rc = put(c->so_rowb, rowb.data(), rowb.size() * sizeof(int2));
  if (!rc) rc = put(c->so_pairs, pairs.data(), pairs.size() * sizeof(int4));
  if (!rc) rc = put(c->so_pstart, start.data(), start.size() * sizeof(int));
  if (rc) return rc;
  c->so_radial_epoch = c->radial_epoch;
  c->so_basis_epoch = c->basis_epoch;
  return 0;
}

}  // namespace
}  // namespace exh

using namespace exh;

extern "C" {

// hmlint (mt_kin + mt_pot) or, with with_kin false, mt_pot alone.
static int hmlint_impl(exh_context *c, void *stream, int lmaxvr, const double *veffmt,
                       const double *lofr, int valence_rel_none, exh_z *aa, exh_z *loa,
                       exh_z *lolo, int flags, bool with_kin) {
  if (!c) return EXH_EINVAL;
  if (int rc = wait_done(c)) return rc;
  if (!c->have_basis || !c->have_radial)
    return fail(c, EXH_ESTATE, "hmlint: basis and radial data not set");
  if (lmaxvr < 0 || lmaxvr > 50) return fail(c, EXH_EINVAL, "hmlint: lmaxvr outside 0..50");
  if (!veffmt || !aa || (c->nlomax > 0 && (!lofr || !loa || !lolo)))
    return fail(c, EXH_EINVAL, "hmlint: null array");
  if (flags & ~EXH_ACCUMULATE) return fail(c, EXH_EINVAL, "hmlint: unknown flags");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, EXH_ESTATE, "hipSetDevice");
  if (int rc = prepare(c, lmaxvr)) return rc;
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  const int p = c->apwordmax, nfa = (c->lmaxmat + 1) * p, nf = nfa + c->nlomax;
  const int lmmaxvr = (lmaxvr + 1) * (lmaxvr + 1), nchunk = (lmmaxvr + kCH - 1) / kCH;
  const int natm = c->natmtot, mnlo = std::max(c->maxnlo, 1);
  int rc = ensure(c, c->hml_ipot, (size_t)natm * nf * nf * lmmaxvr * sizeof(double));
  if (!rc) rc = ensure(c, c->hml_ikin, (size_t)natm * nf * nf * sizeof(double));
  if (rc) return rc;
  const double alpha = 1.0 / 137.03599911;
  HmlArgs a{nf, nfa, p, c->lmaxapw, c->nrmtmax, std::max(c->nlomax, 1), lmmaxvr, nchunk,
            0.5 * alpha * alpha, valence_rel_none ? 0 : 3};
  const int *counts = (const int *)c->hml_count.p;
  const int npair = nf * (nf + 1) / 2;
  const double *lf = lofr ? lofr : (const double *)c->hml_w.p;  // never read when nlomax = 0
  hipError_t e = timed(c, EXH_K_HMLRAD, s, [&] {
    dim3 grid(natm, nchunk + (with_kin ? 1 : 0));
#define LAUNCH(P)                                                                               \
  hipLaunchKernelGGL(hml_radial_kernel<P>, grid, dim3(kRadThreads), 0, s, a,                       \
                     (const int *)c->atom_species_d.p, counts + 2 * c->nspecies,                \
                     (const double *)c->hml_w.p, (const double *)c->hml_spr.p,                  \
                     (const int *)c->hml_fl.p, (const double *)c->apwfr_d.p, lf, veffmt,        \
                     (double *)c->hml_ipot.p, (double *)c->hml_ikin.p)
    if (npair <= kRadThreads) LAUNCH(1);
    else if (npair <= 2 * kRadThreads) LAUNCH(2);
    else if (npair <= 4 * kRadThreads) LAUNCH(4);
    else LAUNCH(8);
#undef LAUNCH
    return hipGetLastError();
  });
  if (e != hipSuccess) return hip_fail(c, e, "hmlint radial");
  const int lmmaxmat = (c->lmaxmat + 1) * (c->lmaxmat + 1);
  RowTable A{(const int2 *)c->hml_arow.p, counts, c->maxaa};
  RowTable L{(const int2 *)c->hml_lrow.p, counts + c->nspecies, mnlo};
  struct Part {
    RowTable r, col;
    int ldo, nco;
    exh_z *out;
  } parts[3] = {{A, A, c->maxaa, c->maxaa, aa}, {L, A, mnlo, c->maxaa, loa}, {L, L, mnlo, mnlo, lolo}};
  const int nparts = c->nlomax > 0 ? 3 : 1;
  e = timed(c, EXH_K_HMLCON, s, [&] {
    hipError_t le = hipSuccess;
    for (int k = 0; k < nparts && le == hipSuccess; k++) {
      const Part &P = parts[k];
      dim3 grid(cdiv_ll((long long)P.ldo * P.nco, kThreads), natm);
      hipLaunchKernelGGL(hml_contract_kernel, grid, dim3(kThreads), 0, s, nf, lmmaxvr, lmmaxmat,
                         (const int *)c->atom_species_d.p, P.r, P.col, P.ldo, P.nco,
                         (const int *)c->hml_gstart.p, (const int *)c->hml_glm2.p,
                         (const double2 *)c->hml_gval.p, (const double *)c->hml_ipot.p,
                         (const double *)c->hml_ikin.p, (double2 *)P.out,
                         (flags & EXH_ACCUMULATE) ? 1 : 0, with_kin ? 1 : 0);
      le = hipGetLastError();
    }
    return le;
  });
  if (e == hipSuccess) e = mark_done(c, s);
  return e == hipSuccess ? 0 : hip_fail(c, e, "hmlint contract");
}

int exh_hmlint_device(exh_context *c, void *stream, int lmaxvr, const double *veffmt,
                      const double *lofr, int valence_rel_none, exh_z *aa, exh_z *loa,
                      exh_z *lolo, int flags) {
  return hmlint_impl(c, stream, lmaxvr, veffmt, lofr, valence_rel_none, aa, loa, lolo, flags,
                     true);
}

int exh_mt_pot_device(exh_context *c, void *stream, int lmaxvr, const double *pot,
                      const double *lofr, exh_z *aa, exh_z *loa, exh_z *lolo, int flags) {
  return hmlint_impl(c, stream, lmaxvr, pot, lofr, 1, aa, loa, lolo, flags, false);
}

int exh_mt_kinetic_device(exh_context *c, void *stream, int lmaxvr, const double *veffmt,
                          const double *lofr, int mass_flags, double *tkin) {
  if (!c) return EXH_EINVAL;
  if (int rc = wait_done(c)) return rc;
  if (!c->have_basis || !c->have_radial)
    return fail(c, EXH_ESTATE, "mt_kinetic: basis and radial data not set");
  if (mass_flags & ~3) return fail(c, EXH_EINVAL, "mt_kinetic: unknown mass flags");
  if (!tkin || (c->nlomax > 0 && !lofr) || (mass_flags && !veffmt))
    return fail(c, EXH_EINVAL, "mt_kinetic: null array");
  if (mass_flags && (lmaxvr < 0 || lmaxvr > 50))
    return fail(c, EXH_EINVAL, "mt_kinetic: lmaxvr outside 0..50");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, EXH_ESTATE, "hipSetDevice");
  const int lmv = mass_flags ? lmaxvr : 0;
  if (int rc = prepare(c, c->hml_lmaxvr >= 0 ? c->hml_lmaxvr : lmv)) return rc;
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  const int p = c->apwordmax, nfa = (c->lmaxmat + 1) * p, nf = nfa + c->nlomax;
  const double alpha = 1.0 / 137.03599911;
  HmlArgs a{nf, nfa, p, c->lmaxapw, c->nrmtmax, std::max(c->nlomax, 1), (lmv + 1) * (lmv + 1), 0,
            0.5 * alpha * alpha, mass_flags};
  const int *counts = (const int *)c->hml_count.p;
  const int npair = nf * (nf + 1) / 2;
  const double *lf = lofr ? lofr : (const double *)c->hml_w.p;  // never read when nlomax = 0
  const double *vm = veffmt ? veffmt : (const double *)c->hml_w.p;  // read only with a mass flag
  hipError_t e = timed(c, EXH_K_HMLRAD, s, [&] {
    dim3 grid(c->natmtot, 1);
#define LAUNCH(P)                                                                               \
  hipLaunchKernelGGL(hml_radial_kernel<P>, grid, dim3(kRadThreads), 0, s, a,                    \
                     (const int *)c->atom_species_d.p, counts + 2 * c->nspecies,                \
                     (const double *)c->hml_w.p, (const double *)c->hml_spr.p,                  \
                     (const int *)c->hml_fl.p, (const double *)c->apwfr_d.p, lf, vm,            \
                     (double *)nullptr, tkin)
    if (npair <= kRadThreads) LAUNCH(1);
    else if (npair <= 2 * kRadThreads) LAUNCH(2);
    else if (npair <= 4 * kRadThreads) LAUNCH(4);
    else LAUNCH(8);
#undef LAUNCH
    return hipGetLastError();
  });
  if (e == hipSuccess) e = mark_done(c, s);
  return e == hipSuccess ? 0 : hip_fail(c, e, "mt_kinetic");
}

static int hmlint_host(exh_context *c, int lmaxvr, const double *veffmt, const double *lofr,
                       int valence_rel_none, exh_z *aa, exh_z *loa, exh_z *lolo, int flags,
                       bool with_kin) {
  if (!c) return EXH_EINVAL;
  if (!c->have_basis || !c->have_radial)
    return fail(c, EXH_ESTATE, "hmlint: basis and radial data not set");
  if (lmaxvr < 0 || lmaxvr > 50) return fail(c, EXH_EINVAL, "hmlint: lmaxvr outside 0..50");
  if (!veffmt || !aa || (c->nlomax > 0 && (!lofr || !loa || !lolo)))
    return fail(c, EXH_EINVAL, "hmlint: null array");
  if (int rc = wait_done(c)) return rc;
  const size_t na = c->natmtot, mnlo = std::max(c->maxnlo, 1);
  const size_t lmm = (size_t)(lmaxvr + 1) * (lmaxvr + 1);
  const size_t naa = (size_t)c->maxaa * c->maxaa * na, nloa = mnlo * c->maxaa * na,
               nlolo = mnlo * mnlo * na;
  int rc = copy_in(c, c->hml_vin, veffmt, lmm * c->nrmtmax * na);
  if (!rc && c->nlomax > 0)
    rc = copy_in(c, c->hml_loin, lofr, (size_t)c->nrmtmax * 2 * c->nlomax * na);
  const bool acc = flags & EXH_ACCUMULATE;
  if (!rc) rc = acc ? copy_in(c, c->hml_aa, (const double2 *)aa, naa)
                    : ensure(c, c->hml_aa, naa * sizeof(double2));
  if (!rc && c->nlomax > 0) {
    rc = acc ? copy_in(c, c->hml_loa, (const double2 *)loa, nloa)
             : ensure(c, c->hml_loa, nloa * sizeof(double2));
    if (!rc)
      rc = acc ? copy_in(c, c->hml_lolo, (const double2 *)lolo, nlolo)
               : ensure(c, c->hml_lolo, nlolo * sizeof(double2));
  }
  if (rc) return rc;
  rc = hmlint_impl(c, nullptr, lmaxvr, (const double *)c->hml_vin.p,
                   c->nlomax > 0 ? (const double *)c->hml_loin.p : nullptr, valence_rel_none,
                   (exh_z *)c->hml_aa.p, c->nlomax > 0 ? (exh_z *)c->hml_loa.p : nullptr,
                   c->nlomax > 0 ? (exh_z *)c->hml_lolo.p : nullptr, flags, with_kin);
  if (!rc) rc = copy_out(c, aa, c->hml_aa, naa * sizeof(double2));
  if (!rc && c->nlomax > 0) rc = copy_out(c, loa, c->hml_loa, nloa * sizeof(double2));
  if (!rc && c->nlomax > 0) rc = copy_out(c, lolo, c->hml_lolo, nlolo * sizeof(double2));
  return rc;
}

int exh_hmlint(exh_context *c, int lmaxvr, const double *veffmt, const double *lofr,
               int valence_rel_none, exh_z *aa, exh_z *loa, exh_z *lolo, int flags) {
  return hmlint_host(c, lmaxvr, veffmt, lofr, valence_rel_none, aa, loa, lolo, flags, true);
}

int exh_mt_pot(exh_context *c, int lmaxvr, const double *pot, const double *lofr, exh_z *aa,
               exh_z *loa, exh_z *lolo, int flags) {
  return hmlint_host(c, lmaxvr, pot, lofr, 1, aa, loa, lolo, flags, false);
}

int exh_set_mt_device(exh_context *c, void *stream, int maxaa, int maxnlo, const exh_z *aa,
                      const exh_z *loa, const exh_z *lolo, const double *oalo,
                      const double *ololo, const double *h1aa, const double *h1loa,
                      const double *h1lolo) {
  if (!c) return EXH_EINVAL;
  if (int rc = wait_done(c)) return rc;
  if (!c->have_basis) return fail(c, EXH_ESTATE, "set_mt: call exh_set_basis first");
  if (maxaa != c->maxaa || maxnlo != c->maxnlo)
    return fail(c, EXH_EINVAL, "set_mt: maxaa/maxnlo differ from the basis");
  if (!aa || (c->nlotot > 0 && (!loa || !lolo || !oalo || !ololo)))
    return fail(c, EXH_EINVAL, "set_mt: null table");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, EXH_ESTATE, "hipSetDevice");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  const size_t na = c->natmtot;
  const size_t mnlo = std::max(maxnlo, 1), nlom = std::max(c->nlomax, 1);
  const size_t naa = (size_t)maxaa * maxaa * na, nloa = mnlo * maxaa * na, nlolo = mnlo * mnlo * na;
  const size_t nsalo = (size_t)c->apwordmax * nlom * na, nolo = nlom * nlom * na;
  const size_t nh1 = (size_t)c->apwordmax * c->apwordmax * (c->lmaxapw + 1) * na;
  int rc = ensure(c, c->aa, naa * sizeof(double2));
  if (!rc) rc = ensure(c, c->loa, loa ? nloa * sizeof(double2) : 16);
  if (!rc) rc = ensure(c, c->lolo, lolo ? nlolo * sizeof(double2) : 16);
  if (!rc) rc = ensure(c, c->salo, nsalo * sizeof(double));
  if (!rc) rc = ensure(c, c->ololo_eff, nolo * sizeof(double));
  if (!rc && h1aa) rc = ensure(c, c->h1aa, nh1 * sizeof(double));
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(c->aa.p, aa, naa * sizeof(double2), hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess && loa)
    e = hipMemcpyAsync(c->loa.p, loa, nloa * sizeof(double2), hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess && lolo)
    e = hipMemcpyAsync(c->lolo.p, lolo, nlolo * sizeof(double2), hipMemcpyDeviceToDevice, s);
  // salo = oalo + h1loa, ololo_eff = ololo + h1lolo (zero without LOs)
  const bool lo = c->nlomax > 0;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(add_tables_kernel, dim3(cdiv_ll(nsalo, 256)), dim3(256), 0, s, nsalo,
                       lo ? oalo : nullptr, h1loa, (double *)c->salo.p);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(add_tables_kernel, dim3(cdiv_ll(nolo, 256)), dim3(256), 0, s, nolo,
                       lo ? ololo : nullptr, h1lolo, (double *)c->ololo_eff.p);
    e = hipGetLastError();
  }
  if (e == hipSuccess && h1aa)
    e = hipMemcpyAsync(c->h1aa.p, h1aa, nh1 * sizeof(double), hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(c, e, "set_mt_device");
  c->have_h1aa = h1aa != nullptr;
  c->have_mt = true;
  return 0;
}


int exh_mt_so_device(exh_context *c, void *stream, int lmaxvr, const double *pot,
                     const double *lofr, int rel_level, exh_z *alpha_aa, exh_z *alpha_loa,
                     exh_z *alpha_lolo, exh_z *alpha_alo, exh_z *beta_aa, exh_z *beta_loa,
                     exh_z *beta_lolo, exh_z *beta_alo, exh_z *ab_aa, exh_z *ab_loa,
                     exh_z *ab_lolo, exh_z *ab_alo, int flags) {
  if (!c) return EXH_EINVAL;
  if (int rc = wait_done(c)) return rc;
  if (!c->have_basis || !c->have_radial)
    return fail(c, EXH_ESTATE, "mt_so: basis and radial data not set");
  if (lmaxvr < 0 || lmaxvr > 50) return fail(c, EXH_EINVAL, "mt_so: lmaxvr outside 0..50");
  if (rel_level < 0 || rel_level > 2) return fail(c, EXH_EINVAL, "mt_so: rel_level outside 0..2");
  if (flags & ~EXH_ACCUMULATE) return fail(c, EXH_EINVAL, "mt_so: unknown flags");
  exh_z *tab[12] = {alpha_aa, alpha_loa, alpha_lolo, alpha_alo, beta_aa, beta_loa,
                    beta_lolo, beta_alo, ab_aa, ab_loa, ab_lolo, ab_alo};
  const bool lo = c->nlomax > 0;
  for (int k = 0; k < 12; k++)
    if (!tab[k] && (k % 4 == 0 || lo)) return fail(c, EXH_EINVAL, "mt_so: null array");
  if (!pot || (lo && !lofr)) return fail(c, EXH_EINVAL, "mt_so: null array");
  if ((size_t)2 * c->nrmtmax * sizeof(double) > 64 * 1024)
    return fail(c, EXH_ERANGE, "mt_so: nrmtmax above 4096");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, EXH_ESTATE, "hipSetDevice");
  if (int rc = so_prepare(c)) return rc;
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  const int natm = c->natmtot, maxaa = c->maxaa, mnlo = std::max(c->maxnlo, 1);
  const size_t sizes[4] = {(size_t)maxaa * maxaa * natm, (size_t)mnlo * maxaa * natm,
                           (size_t)mnlo * mnlo * natm, (size_t)maxaa * mnlo * natm};
  SoTables T;
  for (int k = 0; k < 12; k++) T.t[k] = (double2 *)tab[k];
  const double alpha = 1.0 / 137.03599911;
  SoArgs A{(c->lmaxmat + 1) * c->apwordmax, c->apwordmax, c->lmaxapw, c->nrmtmax,
           std::max(c->nlomax, 1), (lmaxvr + 1) * (lmaxvr + 1), maxaa, mnlo, rel_level,
           0.5 * alpha * alpha};
  const int *counts = (const int *)c->hml_count.p;
  const double *lf = lofr ? lofr : (const double *)c->hml_w.p;  // never read when nlomax = 0
  hipError_t e = timed(c, EXH_K_SO, s, [&] {
    hipError_t le = hipSuccess;
    if (!(flags & EXH_ACCUMULATE))
      for (int k = 0; k < 12 && le == hipSuccess; k++) {
        if (!tab[k]) continue;
        hipLaunchKernelGGL(zero_kernel, dim3(cdiv_ll((long long)sizes[k % 4], 256)), dim3(256), 0,
                           s, sizes[k % 4], T.t[k]);
        le = hipGetLastError();
      }
    if (le != hipSuccess) return le;
    hipLaunchKernelGGL(so_kernel, dim3(natm), dim3(kSoThreads),
                       2 * c->nrmtmax * sizeof(double), s, A, (const int *)c->atom_species_d.p,
                       counts + 2 * c->nspecies, (const double *)c->hml_w.p,
                       (const double *)c->hml_spr.p, (const double *)c->so_dt.p,
                       (const int4 *)c->so_pairs.p, (const int *)c->so_pstart.p,
                       (const int2 *)c->so_rowb.p, (const double *)c->apwfr_d.p, lf, pot, T);
    le = hipGetLastError();
    if (le == hipSuccess && lo && c->maxnlo > 0) {
      hipLaunchKernelGGL(so_alo_kernel,
                         dim3(cdiv_ll((long long)maxaa * c->maxnlo * natm, 256)), dim3(256), 0, s,
                         maxaa, c->maxnlo, mnlo, natm, T);
      le = hipGetLastError();
    }
    return le;
  });
  if (e == hipSuccess) e = mark_done(c, s);
  return e == hipSuccess ? 0 : hip_fail(c, e, "mt_so");
}

int exh_mt_so(exh_context *c, int lmaxvr, const double *pot, const double *lofr, int rel_level,
              exh_z *alpha_aa, exh_z *alpha_loa, exh_z *alpha_lolo, exh_z *alpha_alo,
              exh_z *beta_aa, exh_z *beta_loa, exh_z *beta_lolo, exh_z *beta_alo, exh_z *ab_aa,
              exh_z *ab_loa, exh_z *ab_lolo, exh_z *ab_alo, int flags) {
  if (!c) return EXH_EINVAL;
  if (int rc = wait_done(c)) return rc;
  if (!c->have_basis || !c->have_radial)
    return fail(c, EXH_ESTATE, "mt_so: basis and radial data not set");
  if (lmaxvr < 0 || lmaxvr > 50) return fail(c, EXH_EINVAL, "mt_so: lmaxvr outside 0..50");
  exh_z *tab[12] = {alpha_aa, alpha_loa, alpha_lolo, alpha_alo, beta_aa, beta_loa,
                    beta_lolo, beta_alo, ab_aa, ab_loa, ab_lolo, ab_alo};
  const bool lo = c->nlomax > 0;
  for (int k = 0; k < 12; k++)
    if (!tab[k] && (k % 4 == 0 || lo)) return fail(c, EXH_EINVAL, "mt_so: null array");
  if (!pot || (lo && !lofr)) return fail(c, EXH_EINVAL, "mt_so: null array");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, EXH_ESTATE, "hipSetDevice");
  const int natm = c->natmtot, maxaa = c->maxaa, mnlo = std::max(c->maxnlo, 1);
  const size_t sizes[4] = {(size_t)maxaa * maxaa * natm, (size_t)mnlo * maxaa * natm,
                           (size_t)mnlo * mnlo * natm, (size_t)maxaa * mnlo * natm};
  size_t off[13] = {0};
  for (int k = 0; k < 12; k++) off[k + 1] = off[k] + (tab[k] ? sizes[k % 4] : 0);
  const size_t lmm = (size_t)(lmaxvr + 1) * (lmaxvr + 1);
  int rc = copy_in(c, c->hml_vin, pot, lmm * c->nrmtmax * natm);
  if (!rc && lo) rc = copy_in(c, c->hml_loin, lofr, (size_t)c->nrmtmax * 2 * c->nlomax * natm);
  if (!rc) rc = ensure(c, c->so_host, off[12] * sizeof(double2));
  if (rc) return rc;
  double2 *base = (double2 *)c->so_host.p;
  exh_z *dev[12];
  hipError_t e = hipSuccess;
  for (int k = 0; k < 12; k++) {
    dev[k] = tab[k] ? (exh_z *)(base + off[k]) : nullptr;
    if (tab[k] && (flags & EXH_ACCUMULATE) && e == hipSuccess)
      e = hipMemcpyAsync(dev[k], tab[k], sizes[k % 4] * sizeof(double2), hipMemcpyHostToDevice,
                         c->stream);
  }
  if (e != hipSuccess) return hip_fail(c, e, "upload");
  rc = exh_mt_so_device(c, nullptr, lmaxvr, (const double *)c->hml_vin.p,
                        lo ? (const double *)c->hml_loin.p : nullptr, rel_level, dev[0], dev[1],
                        dev[2], dev[3], dev[4], dev[5], dev[6], dev[7], dev[8], dev[9], dev[10],
                        dev[11], flags);
  if (rc) return rc;
  for (int k = 0; k < 12; k++) {
    if (!tab[k]) continue;
    e = hipMemcpyAsync(tab[k], dev[k], sizes[k % 4] * sizeof(double2), hipMemcpyDeviceToHost,
                       c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "download");
  }
  e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? 0 : hip_fail(c, e, "download");
}

}  // extern "C"
