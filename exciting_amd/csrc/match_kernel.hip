// APW matching coefficients on the GPU (match.f90:13-194).
//
// One thread per (G+k vector, atom).  Each thread evaluates the spherical
// Bessel functions and their derivatives at |G+k| R (sbessel.f90,
// sbesseldm.f90), the complex spherical harmonics of (G+k)^ by the
// Masters & Richards-Dinger recursion (genylm.f90), forms
//   b_io = 4 pi i^l / sqrt(Omega) |G+k|^(io-1) j_l^(io-1)(|G+k|R) S_a(G+k) Y*_lm
// and solves D A = b with the LU factors of the boundary-derivative matrix D
// (zgetrf/zgetrs order; factors computed once per (l, atom) on the host).
// Writes apwalm(G, io, lm, atom): consecutive threads write consecutive G,
// so every store is a coalesced 1 KB row segment.  The kernel is bound by
// those writes (16 B x G x p x (lmax+1)^2 x atoms).
#include "exh_device.h"

namespace exh {
namespace {

constexpr int PCAP = 3;  // largest APW order supported (apwordmax)

__constant__ double k_fact2[38] = {
    1.0, 2.0, 3.0, 8.0, 15.0, 48.0, 105.0, 384.0, 945.0, 3840.0, 10395.0,
    46080.0, 135135.0, 645120.0, 2027025.0, 10321920.0, 34459425.0,
    185794560.0, 654729075.0, 3715891200.0, 13749310575.0, 81749606400.0,
    316234143225.0, 1961990553600.0, 7905853580625.0, 51011754393600.0,
    213458046676875.0, 1428329123020800.0, 6190283353629375.0,
    42849873690624000.0, 191898783962510625.0, 1371195958099968000.0,
    6332659870762850625.0, 46620662575398912000.0, 221643095476699771875.0,
    1678343852714360832000.0, 8200794532637891559375.0,
    63777066403145711616000.0};

// n! for small n (factnm(n,1) with n <= 24 is exact to the tabulated value)
__device__ __forceinline__ double d_fact(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; i++) f *= (double)i;
  return f;
}
// factnm(n, 2): n!! from the table
__device__ __forceinline__ double d_fact2(int n) { return n == 0 ? 1.0 : k_fact2[n - 1]; }
// factr(n, d) = n!/d! (factr.f90)
__device__ __forceinline__ double d_factr(int n, int d) {
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

// sbessel.f90:13-146 for compile-time lmax
template <int LM>
__device__ __forceinline__ void d_sbessel(double x, double (&jl)[LM + 1]) {
  const double rsc = 1.0e150, rsci = 1.0 / rsc;
  if (x < 1.0e-8) {
    double t1 = 1.0, t2 = 1.0;
    jl[0] = 1.0;
#pragma unroll
    for (int l = 1; l <= LM; l++) {
      t1 = t1 / (double)(2 * l + 1);
      t2 = t2 * x;
      jl[l] = t2 * t1;
    }
    return;
  }
  const double xi = 1.0 / x;
  if constexpr (LM == 0) {
    jl[0] = sin(x) * xi;  // x >= 1e-8 > lmax: upward branch
    return;
  } else {
  if (x < (double)LM) {
    double j0 = 0.6370354636449841609 * rsci, j1 = 0.3532702964695481204 * rsci, jt;
#pragma unroll
    for (int l = LM + 25; l >= LM + 1; l--) {
      jt = j0 * (double)(2 * l + 1) * xi - j1;
      j1 = j0;
      j0 = jt;
      if (fabs(j0) > rsc) {
        jt *= rsci;
        j1 *= rsci;
        j0 *= rsci;
      }
    }
#pragma unroll
    for (int l = LM; l >= 0; l--) {
      jt = j0 * (double)(2 * l + 1) * xi - j1;
      j1 = j0;
      j0 = jt;
      if (fabs(j0) > rsc) {
        jt *= rsci;
        j1 *= rsci;
        j0 *= rsci;
#pragma unroll
        for (int k = l + 1; k <= LM; k++) jl[k] *= rsci;
      }
      jl[l] = j1;
    }
    const double t1 = 1.0 / ((jl[0] - x * jl[1]) * cos(x) + x * jl[0] * sin(x));
#pragma unroll
    for (int l = 0; l <= LM; l++) jl[l] = t1 * jl[l];
    return;
  }
  jl[0] = sin(x) * xi;
  jl[1] = (jl[0] - cos(x)) * xi;
  double j0 = jl[0], j1 = jl[1], jt;
#pragma unroll
  for (int l = 2; l <= LM; l++) {
    jt = (double)(2 * l - 1) * j1 * xi - j0;
    j0 = j1;
    j1 = jt;
    jl[l] = j1;
  }
  }
}

// sbesseldm.f90:13-133 for m >= 1; jp1 = sbessel(LM+1, x), valid when x > 1
template <int LM>
__device__ __forceinline__ void d_sbesseldm(int m, double x,
                                            const double (&jp1)[LM + 2],
                                            double (&djl)[LM + 1]) {
  if (x > 1.0) {
#pragma unroll
    for (int l = 0; l <= LM; l++) {
      int a[PCAP + 2], a1[PCAP + 2], b[PCAP + 2], b1[PCAP + 2];
#pragma unroll
      for (int j = 0; j < PCAP + 2; j++) a[j] = a1[j] = b[j] = b1[j] = 0;
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
      double sum = (double)a[0] * jp1[l] + (double)a1[0] * jp1[l + 1];
      for (int i = 1; i <= m + 1; i++) {
        t1 = t1 * x;
        sum = sum + ((double)a[i] * jp1[l] + (double)a1[i] * jp1[l + 1]) / t1;
      }
      djl[l] = sum;
    }
    return;
  }
  const double x2 = x * x;
#pragma unroll
  for (int l = 0; l <= LM; l++) {
    int i0 = (m - l + 1) / 2;
    if (i0 < 0) i0 = 0;
    int j = 2 * i0 + l - m;
    double t1 = 1.0;
    for (int k = 0; k < j; k++) t1 *= x;
    double pw = 1.0;
    for (int k = 0; k < i0; k++) pw *= -2.0;
    t1 = d_factr(j + m, j) * t1 / (d_fact(i0) * d_fact2(j + l + m + 1) * pw);
    double sum = t1;
    for (int i = i0 + 1; i <= 20; i++) {
      j = 2 * i + l;
      t1 = -t1 * (double)((j - 1) * j) * x2 /
           (double)((j - l) * (j - m - 1) * (j - m) * (j + l + 1));
      if (fabs(t1) <= 1.0e-40) break;
      sum = sum + t1;
    }
    djl[l] = sum;
  }
}

template <int LM, typename PD, typename PS, typename PA>
__device__ __forceinline__ void match_body(
    int g, int ias, int ngp, int ld, int pmax, PD __restrict__ gpc, PD __restrict__ tpgpc,
    PS __restrict__ sfacgp, PA __restrict__ apwalm, const int *__restrict__ atom_species,
    const double *__restrict__ rmt, const int *__restrict__ apword,
    const int *__restrict__ omaxs, const double *__restrict__ lu,
    const int *__restrict__ piv, double fpso) {
  if (g >= ngp) return;
  const int is = atom_species[ias];
  const int omax = omaxs[is];
  constexpr int LMM = (LM + 1) * (LM + 1);
  const double q = gpc[g];
  const double x = q * rmt[is];

  // radial factors djl[io][l] = |G+k|^io j_l^(io)(x)      (match.f90:114-131)
  double djl[PCAP][LM + 1];
  {
    double jl[LM + 1];
    d_sbessel<LM>(x, jl);
#pragma unroll
    for (int l = 0; l <= LM; l++) djl[0][l] = jl[l];
  }
  if (omax > 1) {
    double jp1[LM + 2];
#pragma unroll
    for (int l = 0; l <= LM + 1; l++) jp1[l] = 0.0;
    if (x > 1.0) d_sbessel<LM + 1>(x, jp1);
    double t1 = 1.0;
#pragma unroll
    for (int io = 1; io < PCAP; io++) {
      if (io < omax) {
        double d[LM + 1];
        d_sbesseldm<LM>(io, x, jp1, d);
        t1 = t1 * q;
#pragma unroll
        for (int l = 0; l <= LM; l++) djl[io][l] = t1 * d[l];
      }
    }
  }

  // spherical harmonics ingredients (genylm.f90:56-62)
  const double sn = sin(tpgpc[2 * g]), cs = cos(tpgpc[2 * g]);
  double zr[LM + 1], zi[LM + 1];
#pragma unroll
  for (int m = 1; m <= LM; m++) {
    const double t1 = (double)m * tpgpc[2 * g + 1];
    zr[m] = cos(t1);
    zi[m] = sin(t1);
  }
  const double2 S = exh::ld(sfacgp + g + (size_t)ld * ias);
  const PA out = apwalm + g + (size_t)ld * pmax * LMM * ias;

#pragma unroll
  for (int l = 0; l <= LM; l++) {
    // x_m (genylm.f90:63-81) for this l
    double xm[LM + 1];
    double tn = 0.0;
    if (l == 0) {
      xm[0] = 1.0;
    } else {
      xm[l] = (l % 2 == 0) ? 1.0 : -1.0;
      double dx = 0.0;
#pragma unroll
      for (int m = l; m >= 1; m--) {
        const double t1 = sqrt((double)((l + m) * (l - m + 1)));
        xm[m - 1] = -(sn * dx + (double)(2 * m) * cs * xm[m]) / t1;
        dx = sn * xm[m] * t1;
      }
      double t1 = sn, sum = 0.0;
#pragma unroll
      for (int m = 1; m <= l; m++) {
        xm[m] = t1 * xm[m];
        sum = sum + xm[m] * xm[m];
        t1 = t1 * sn;
      }
      sum = 2.0 * sum + xm[0] * xm[0];
      tn = sqrt((double)(2 * l + 1) / (12.566370614359172954 * sum));
    }
    // zt1 = fpso * i^l ; zt2 = zt1 * S       (match.f90:139,152)
    double z1r = 0.0, z1i = 0.0;
    switch (l & 3) {
      case 0: z1r = fpso; break;
      case 1: z1i = fpso; break;
      case 2: z1r = -fpso; break;
      default: z1i = -fpso; break;
    }
    const double z2r = z1r * S.x - z1i * S.y, z2i = z1r * S.y + z1i * S.x;
    const int p = apword[l + (LM + 1) * is];
    const double *LU = lu + (size_t)PCAP * PCAP * (l + (LM + 1) * ias);
    const int *PV = piv + PCAP * (l + (LM + 1) * ias);
#pragma unroll
    for (int m = -l; m <= l; m++) {
      // Y_lm (genylm.f90:82-90), then conj
      double yr, yi;
      if (l == 0) {
        yr = 0.28209479177387814347;
        yi = 0.0;
      } else if (m == 0) {
        yr = tn * xm[0];
        yi = 0.0;
      } else {
        const int am = m < 0 ? -m : m;
        const double s = tn * xm[am];
        yr = s * zr[am];
        yi = s * zi[am];
        if (m < 0) {
          yi = -yi;
          if (am & 1) {
            yr = -yr;
            yi = -yi;
          }
        }
      }
      const double ycr = yr, yci = -yi;
      double br[PCAP], bi[PCAP];
#pragma unroll
      for (int io = 0; io < PCAP; io++) {
        if (io < p) {
          const double tr = djl[io][l] * z2r, ti = djl[io][l] * z2i;
          br[io] = tr * ycr - ti * yci;
          bi[io] = tr * yci + ti * ycr;
        }
      }
      // zgetrs: row interchanges, unit-lower then upper solve
#pragma unroll
      for (int i = 0; i < PCAP; i++)
        if (i < p) {
          const int ip = PV[i] - 1;
          if (ip != i) {
            double t = br[i];
            br[i] = br[ip];
            br[ip] = t;
            t = bi[i];
            bi[i] = bi[ip];
            bi[ip] = t;
          }
        }
#pragma unroll
      for (int k = 0; k < PCAP; k++)
#pragma unroll
        for (int i = k + 1; i < PCAP; i++)
          if (i < p) {
            const double lv = LU[i + PCAP * k];
            br[i] -= br[k] * lv;
            bi[i] -= bi[k] * lv;
          }
#pragma unroll
      for (int k = PCAP - 1; k >= 0; k--)
        if (k < p) {
          const double u = LU[k + PCAP * k];
          br[k] = br[k] / u;
          bi[k] = bi[k] / u;
#pragma unroll
          for (int i = 0; i < k; i++) {
            const double uv = LU[i + PCAP * k];
            br[i] -= br[k] * uv;
            bi[i] -= bi[k] * uv;
          }
        }
      const int lm = l * (l + 1) + m;
#pragma unroll
      for (int io = 0; io < PCAP; io++)
        if (io < p) {
          double2 v;
          v.x = br[io];
          v.y = bi[io];
          st(out + (size_t)ld * (io + (size_t)pmax * lm), v);
        }
    }
  }
}

template <int LM>
__global__ __launch_bounds__(64) void match_kernel(
    int ngp, int ld, int pmax, const double *__restrict__ gpc,
    const double *__restrict__ tpgpc, const double2 *__restrict__ sfacgp,
    double2 *__restrict__ apwalm, const int *__restrict__ atom_species,
    const double *__restrict__ rmt, const int *__restrict__ apword,
    const int *__restrict__ omaxs, const double *__restrict__ lu,
    const int *__restrict__ piv, double fpso) {
  match_body<LM>(blockIdx.x * 64 + threadIdx.x, blockIdx.y, ngp, ld, pmax, gpc, tpgpc, sfacgp,
                 apwalm, atom_species, rmt, apword, omaxs, lu, piv, fpso);
}

// Grouped over k-points: blockIdx.z selects the descriptor, blockIdx.y the
// atom.  One launch per lmaxapw present in the group; k-points of structures
// with another lmaxapw exit at once.
template <int LM>
__global__ __launch_bounds__(64) void match_grouped_kernel(const KDesc *__restrict__ kd) {
  const KDesc &d = kd[blockIdx.z];
  const SDesc &S = *d.sd;
  if (S.lmaxapw != LM || (int)blockIdx.y >= S.natmtot) return;
  match_body<LM>(blockIdx.x * 64 + threadIdx.x, blockIdx.y, d.ngp, S.ngkmax, S.pmax,
                 gptr(d.gpc), gptr(d.tpgpc), gptr(d.sfac), gptr(d.apwalm), S.atom_species,
                 S.rmt, S.apword, S.omax, S.lu, S.piv, S.fpso);
}

inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

// Largest lmaxapw with a compiled match kernel.
constexpr int kMatchLmaxCap = 12;

hipError_t launch_match(hipStream_t s, const exh_context *c, int ngp,
                        const double *gpc, const double *tpgpc,
                        const double2 *sfacgp, double2 *apwalm) {
  if (ngp <= 0 || c->natmtot <= 0) return hipSuccess;
  const double fpso = 12.566370614359172954 / sqrt(c->omega);
  dim3 grid(cdiv(ngp, 64), c->natmtot);
  const int *as = (const int *)c->atom_species_d.p;
  const int *apw = (const int *)c->apword_d.p;
  const int *om = (const int *)c->omax_d.p;
  const double *lu = (const double *)c->mt_match.lu.p;
  const int *pv = (const int *)c->mt_match.piv.p;
#define EXH_M(L)                                                                      \
  case L:                                                                             \
    hipLaunchKernelGGL(match_kernel<L>, grid, dim3(64), 0, s, ngp, c->ngkmax,         \
                       c->apwordmax, gpc, tpgpc, sfacgp, apwalm, as,                  \
                       (const double *)c->rmt_d.p, apw, om, lu, pv, fpso);            \
    break;
  switch (c->lmaxapw) {
    EXH_M(0) EXH_M(1) EXH_M(2) EXH_M(3) EXH_M(4) EXH_M(5) EXH_M(6)
    EXH_M(7) EXH_M(8) EXH_M(9) EXH_M(10) EXH_M(11) EXH_M(12)
    default:
      return hipErrorInvalidValue;
  }
#undef EXH_M
  return hipGetLastError();
}

hipError_t launch_match_grouped(hipStream_t s, const GroupDims &g, const KDesc *kd) {
  if (g.nk <= 0 || g.ngpmax <= 0 || g.natmax <= 0) return hipSuccess;
  dim3 grid(cdiv(g.ngpmax, 64), g.natmax, g.nk);
  if (g.lmax_mask >> (kMatchLmaxCap + 1)) return hipErrorInvalidValue;
#define EXH_M(L)                                                                  \
  if (g.lmax_mask & (1u << L)) {                                                  \
    hipLaunchKernelGGL(match_grouped_kernel<L>, grid, dim3(64), 0, s, kd);        \
    hipError_t e = hipGetLastError();                                             \
    if (e != hipSuccess) return e;                                                \
  }
  EXH_M(0) EXH_M(1) EXH_M(2) EXH_M(3) EXH_M(4) EXH_M(5) EXH_M(6)
  EXH_M(7) EXH_M(8) EXH_M(9) EXH_M(10) EXH_M(11) EXH_M(12)
#undef EXH_M
  return hipSuccess;
}

}  // namespace exh
