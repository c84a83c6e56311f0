// APW matching coefficients on the GPU (match.f90:13-194).
//
// The reference solves, per (l, atom), D A = b with
//   b_io = 4 pi i^l / sqrt(Omega) |G+k|^(io-1) j_l^(io-1)(|G+k|R) S_a(G+k) Y*_lm
// (spherical Bessel functions and derivatives from sbessel.f90 and
// sbesseldm.f90, complex spherical harmonics by the Masters &
// Richards-Dinger recursion of genylm.f90, zgesv on D).  b is a real radial
// vector times one complex factor, so the solve runs on the radial vector
// (zgetrs order with the LU factors of D, computed once per (l, atom) on the
// host) and the factor multiplies the solution.  The harmonics depend on
// G+k only and the radial factors on (G+k, species): a workgroup computes
// them once for its G+k vectors into LDS and reuses them for its atoms.
// The kernel is bound by the writes of apwalm(G, io, lm, atom)
// (16 B x G x p x (lmax+1)^2 x atoms), all coalesced along G.
#include <algorithm>

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

// ---------------------------------------------------------------------------
// One workgroup: MG = 256 consecutive G+k vectors (one per lane) x MA
// consecutive atoms, after match_prep has tabulated, per G+k,
// conj(Y_lm) * 4 pi i^l / sqrt(Omega) and, per (G+k, species),
// |G+k|^io j_l^(io)(|G+k| R).  For each l the workgroup solves the boundary
// system D v = (j_l^(io)) per (G+k, atom) with the LU factors of (l, atom)
// (zgetrs order on the real radial vector; the complex factor
// S_a(G+k) conj(Y_lm) fpso i^l multiplies the solution), then for each m
// reads the harmonics once and writes apwalm(G, io, lm, a) = v_io S_a yt_lm
// for every atom of the tile.  All four waves walk the same column sequence,
// so every column receives 4 KB contiguous per workgroup step: the widest
// row segments apwalm's leading dimension (the caller's ngkmax, not aligned
// to cache lines) allows.  The tables (a few MB per k-point) stay in cache;
// the loop holds few registers and no LDS, so many waves keep stores in
// flight.
// ---------------------------------------------------------------------------
constexpr int MG = 256;   // G+k vectors per workgroup
constexpr int MA = 4;     // atoms per workgroup
constexpr int MTH = 256;  // threads: one per G+k

// conj(Y_lm) * fpso * i^l for m = -L..L of one G+k direction (genylm.f90:56-90),
// into out[(L + m) * stride] (out points at lm = L*L).
template <int L, typename PO>
__device__ __forceinline__ void ylm_row(double sn, double cs, double phi, double fpso, PO out,
                                        size_t stride) {
  double xm[L + 1];
  double tn = 0.0;
  if constexpr (L == 0) {
    xm[0] = 1.0;
  } else {
    xm[L] = (L % 2 == 0) ? 1.0 : -1.0;
    double dx = 0.0;
#pragma unroll
    for (int m = L; m >= 1; m--) {
      const double t1 = sqrt((double)((L + m) * (L - m + 1)));
      xm[m - 1] = -(sn * dx + (double)(2 * m) * cs * xm[m]) / t1;
      dx = sn * xm[m] * t1;
    }
    double t1 = sn, sum = 0.0;
#pragma unroll
    for (int m = 1; m <= L; m++) {
      xm[m] = t1 * xm[m];
      sum = sum + xm[m] * xm[m];
      t1 = t1 * sn;
    }
    sum = 2.0 * sum + xm[0] * xm[0];
    tn = sqrt((double)(2 * L + 1) / (12.566370614359172954 * sum));
  }
  // zt1 = fpso * i^l (match.f90:139)
  double z1r = 0.0, z1i = 0.0;
  switch (L & 3) {
    case 0: z1r = fpso; break;
    case 1: z1i = fpso; break;
    case 2: z1r = -fpso; break;
    default: z1i = -fpso; break;
  }
#pragma unroll
  for (int m = -L; m <= L; m++) {
    double yr, yi;
    if (L == 0) {
      yr = 0.28209479177387814347;
      yi = 0.0;
    } else if (m == 0) {
      yr = tn * xm[0];
      yi = 0.0;
    } else {
      const int am = m < 0 ? -m : m;
      const double t1 = (double)am * phi;
      const double s = tn * xm[am];
      yr = s * cos(t1);
      yi = s * sin(t1);
      if (m < 0) {
        yi = -yi;
        if (am & 1) {
          yr = -yr;
          yi = -yi;
        }
      }
    }
    const double ycr = yr, yci = -yi;  // conj
    st(out + (size_t)(L + m) * stride, z2(fma(z1r, ycr, -(z1i * yci)), fma(z1r, yci, z1i * ycr)));
  }
}

template <int LM, int L = 0, typename PO>
__device__ __forceinline__ void ylm_all(double sn, double cs, double phi, double fpso, PO yt,
                                        size_t stride) {
  if constexpr (L <= LM) {
    ylm_row<L>(sn, cs, phi, fpso, yt + (size_t)L * L * stride, stride);
    ylm_all<LM, L + 1>(sn, cs, phi, fpso, yt, stride);
  }
}

// |G+k|^io j_l^(io)(|G+k| R), io < omax (match.f90:114-131)
template <int LM>
__device__ __forceinline__ void radial_row(double q, double rmt, int omax, double (&djl)[PCAP][LM + 1]) {
  const double x = q * rmt;
  {
    double jl[LM + 1];
    d_sbessel<LM>(x, jl);
#pragma unroll
    for (int l = 0; l <= LM; l++) djl[0][l] = jl[l];
  }
#pragma unroll
  for (int io = 1; io < PCAP; io++)
#pragma unroll
    for (int l = 0; l <= LM; l++) djl[io][l] = 0.0;
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
}

// Tables of one k-point, in its match scratch (ld = ngkmax):
//   rad[((is * PCAP + io) * (LM + 1) + l) * ld + g] = |G+k|^io j_l^(io)(|G+k| R_is),
//     zero for io >= omax(is);
//   yt[lm * ld + g] = conj(Y_lm(G+k)) * 4 pi i^l / sqrt(Omega).
// One thread per (G+k, role): role < nspecies the radial factors of that
// species, role == nspecies the spherical harmonics.
__host__ __device__ inline size_t match_rad_doubles(int nspecies, int lmaxapw, int ld) {
  // rounded so that the harmonics that follow are 256-byte aligned
  return ((size_t)nspecies * PCAP * (lmaxapw + 1) * ld + 31) / 32 * 32;
}

template <int LM, typename PD, typename PR, typename PY>
__device__ __forceinline__ void prep_body(int g, int role, int ngp, int nspecies, int ld,
                                          PD __restrict__ gpc, PD __restrict__ tpgpc,
                                          const double *__restrict__ rmt,
                                          const int *__restrict__ omaxs, double fpso,
                                          PR __restrict__ rad, PY __restrict__ yt) {
  if (g >= ngp) return;
  if (role < nspecies) {
    double djl[PCAP][LM + 1];
    radial_row<LM>(gpc[g], rmt[role], omaxs[role], djl);
#pragma unroll
    for (int io = 0; io < PCAP; io++)
#pragma unroll
      for (int l = 0; l <= LM; l++)
        rad[((size_t)(role * PCAP + io) * (LM + 1) + l) * ld + g] = djl[io][l];
  } else {
    const double th = tpgpc[2 * g], phi = tpgpc[2 * g + 1];
    ylm_all<LM>(sin(th), cos(th), phi, fpso, yt + g, (size_t)ld);
  }
}

template <int LM>
__global__ __launch_bounds__(64) void match_prep_kernel(int ngp, int nspecies, int ld,
                                                        const double *__restrict__ gpc,
                                                        const double *__restrict__ tpgpc,
                                                        const double *__restrict__ rmt,
                                                        const int *__restrict__ omaxs,
                                                        double fpso, double *__restrict__ tab) {
  prep_body<LM>(blockIdx.x * 64 + threadIdx.x, blockIdx.y, ngp, nspecies, ld, gpc, tpgpc, rmt,
                omaxs, fpso, tab,
                (double2 *)(tab + match_rad_doubles(nspecies, LM, ld)));
}

template <int LM>
__global__ __launch_bounds__(64) void match_prep_grouped_kernel(const KDesc *__restrict__ kd) {
  const KDesc &d = kd[blockIdx.z];
  const SDesc &S = *d.sd;
  if (S.lmaxapw != LM || (int)blockIdx.y > S.nspecies) return;
  double *tab = d.mtab;
  prep_body<LM>(blockIdx.x * 64 + threadIdx.x, blockIdx.y, d.ngp, S.nspecies, S.ngkmax,
                gptr(d.gpc), gptr(d.tpgpc), S.rmt, S.omax, S.fpso, gptr(tab),
                gptr((double2 *)(tab + match_rad_doubles(S.nspecies, LM, S.ngkmax))));
}

template <int LM, typename PS, typename PA, typename PR, typename PY>
__device__ __forceinline__ void match_tile(
    int g0, int a0, int ngp, int natmtot, int ld, int pmax, PS __restrict__ sfacgp,
    PA __restrict__ apwalm, const int *__restrict__ atom_species, PR __restrict__ rad,
    PY __restrict__ yt, const int *__restrict__ apword, const double *__restrict__ lu,
    const int *__restrict__ piv) {
  const int na = min(natmtot, a0 + MA) - a0;
  if (g0 >= ngp || na <= 0) return;  // uniform per workgroup
  const int g = g0 + (int)threadIdx.x;
  if (g >= ngp) return;  // no barriers below
  const EXH_C double *LUt = cptr(lu);
  const EXH_C int *PVt = cptr(piv);
  const EXH_C int *APt = cptr(apword);
  const EXH_C int *ASt = cptr(atom_species);
  int is[MA];
  double2 S[MA];
#pragma unroll
  for (int u = 0; u < MA; u++) {
    is[u] = u < na ? ASt[a0 + u] : 0;
    S[u] = u < na ? exh::ld(sfacgp + g + (size_t)ld * (a0 + u)) : z2(0.0, 0.0);
  }
#pragma unroll 1
  for (int l = 0; l <= LM; l++) {
    // boundary solutions of every atom of the tile for this l
    double v[MA][PCAP];
    int p[MA];
#pragma unroll
    for (int u = 0; u < MA; u++) {
      p[u] = 0;
      if (u >= na) continue;
      const int pu = APt[l + (LM + 1) * is[u]];
      p[u] = pu;
      const EXH_C double *LU = LUt + (size_t)PCAP * PCAP * (l + (LM + 1) * (a0 + u));
      const EXH_C int *PV = PVt + PCAP * (l + (LM + 1) * (a0 + u));
      double x[PCAP];
#pragma unroll
      for (int io = 0; io < PCAP; io++)
        x[io] = io < pu ? exh::ld(rad + ((size_t)(is[u] * PCAP + io) * (LM + 1) + l) * ld + g)
                        : 0.0;
      // zgetrs: row interchanges, unit-lower then upper solve
#pragma unroll
      for (int i = 0; i < PCAP; i++)
        if (i < pu) {
          const int ip = PV[i] - 1;
          if (ip != i) {
            double tt = x[i];
#pragma unroll
            for (int j = 0; j < PCAP; j++)
              if (j == ip) {
                x[i] = x[j];
                x[j] = tt;
              }
          }
        }
#pragma unroll
      for (int k = 0; k < PCAP; k++)
#pragma unroll
        for (int i = k + 1; i < PCAP; i++)
          if (i < pu) x[i] = fma(-x[k], LU[i + PCAP * k], x[i]);
#pragma unroll
      for (int k = PCAP - 1; k >= 0; k--)
        if (k < pu) {
          x[k] = x[k] / LU[k + PCAP * k];
#pragma unroll
          for (int i = 0; i < k; i++) x[i] = fma(-x[k], LU[i + PCAP * k], x[i]);
        }
#pragma unroll
      for (int io = 0; io < PCAP; io++) v[u][io] = x[io];
    }
#pragma unroll 1
    for (int lm = l * l; lm < (l + 1) * (l + 1); lm++) {
      const double2 y = exh::ld(yt + (size_t)lm * ld + g);
#pragma unroll
      for (int u = 0; u < MA; u++) {
        if (u >= na) break;
        // explicit fma: the same rounding in every instantiation of the tile
        // (the compiler would otherwise fuse either product)
        const double wr = fma(S[u].x, y.x, -(S[u].y * y.y));
        const double wi = fma(S[u].x, y.y, S[u].y * y.x);
        const PA out = apwalm + g + (size_t)ld * pmax * ((LM + 1) * (LM + 1)) * (a0 + u);
#pragma unroll
        for (int io = 0; io < PCAP; io++)
          if (io < p[u])
            st_nt(out + (size_t)ld * (io + (size_t)pmax * lm), z2(v[u][io] * wr, v[u][io] * wi));
      }
    }
  }
}

template <int LM>
__global__ __launch_bounds__(MTH) void match_kernel(
    int ngp, int natmtot, int nspecies, int ld, int pmax, const double2 *__restrict__ sfacgp,
    double2 *__restrict__ apwalm, const int *__restrict__ atom_species,
    const double *__restrict__ tab, const int *__restrict__ apword,
    const double *__restrict__ lu, const int *__restrict__ piv) {
  match_tile<LM>(blockIdx.x * MG, blockIdx.y * MA, ngp, natmtot, ld, pmax, sfacgp, apwalm,
                 atom_species, tab,
                 (const double2 *)(tab + match_rad_doubles(nspecies, LM, ld)), apword, lu, piv);
}

// Grouped over k-points: blockIdx.z selects the descriptor, blockIdx.y the
// atom chunk.  One launch per lmaxapw present in the group; k-points of
// structures with another lmaxapw exit at once.
template <int LM>
__global__ __launch_bounds__(MTH) void match_grouped_kernel(const KDesc *__restrict__ kd) {
  const KDesc &d = kd[blockIdx.z];
  const EXH_C SDesc &S = *cptr(d.sd);
  if (S.lmaxapw != LM) return;
  const double *tab = d.mtab;
  match_tile<LM>(blockIdx.x * MG, blockIdx.y * MA, d.ngp, S.natmtot, S.ngkmax, S.pmax,
                 gptr(d.sfac), gptr(d.apwalm), S.atom_species, gptr(tab),
                 gptr((const double2 *)(tab + match_rad_doubles(S.nspecies, LM, S.ngkmax))),
                 S.apword, S.lu, S.piv);
}

inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

// Largest lmaxapw with a compiled match kernel.
constexpr int kMatchLmaxCap = 12;

size_t match_table_bytes(const exh_context *c) {
  const int ld = std::max(c->ngkmax, 1);
  return sizeof(double) * match_rad_doubles(std::max(c->nspecies, 1), c->lmaxapw, ld) +
         sizeof(double2) * (size_t)(c->lmaxapw + 1) * (c->lmaxapw + 1) * ld;
}

hipError_t launch_match(hipStream_t s, const exh_context *c, int ngp,
                        const double *gpc, const double *tpgpc,
                        const double2 *sfacgp, double2 *apwalm, double *tab) {
  if (ngp <= 0 || c->natmtot <= 0) return hipSuccess;
  const double fpso = 12.566370614359172954 / sqrt(c->omega);
  dim3 gridp(cdiv(ngp, 64), c->nspecies + 1);
  dim3 grid(cdiv(ngp, MG), cdiv(c->natmtot, MA));
  const int *as = (const int *)c->atom_species_d.p;
  const int *apw = (const int *)c->apword_d.p;
  const int *om = (const int *)c->omax_d.p;
  const double *lu = (const double *)c->mt_match.lu.p;
  const int *pv = (const int *)c->mt_match.piv.p;
  const double *rmt = (const double *)c->rmt_d.p;
#define EXH_M(L)                                                                          \
  case L:                                                                                 \
    hipLaunchKernelGGL(match_prep_kernel<L>, gridp, dim3(64), 0, s, ngp, c->nspecies,     \
                       c->ngkmax, gpc, tpgpc, rmt, om, fpso, tab);                        \
    hipLaunchKernelGGL(match_kernel<L>, grid, dim3(MTH), 0, s, ngp, c->natmtot,           \
                       c->nspecies, c->ngkmax, c->apwordmax, sfacgp, apwalm, as, tab, apw, \
                       lu, pv);                                                           \
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
  dim3 gridp(cdiv(g.ngpmax, 64), g.nspmax + 1, g.nk);
  dim3 grid(cdiv(g.ngpmax, MG), cdiv(g.natmax, MA), g.nk);
  if (g.lmax_mask >> (kMatchLmaxCap + 1)) return hipErrorInvalidValue;
#define EXH_M(L)                                                                  \
  if (g.lmax_mask & (1u << L)) {                                                  \
    hipLaunchKernelGGL(match_prep_grouped_kernel<L>, gridp, dim3(64), 0, s, kd);  \
    hipLaunchKernelGGL(match_grouped_kernel<L>, grid, dim3(MTH), 0, s, kd);       \
    hipError_t e = hipGetLastError();                                             \
    if (e != hipSuccess) return e;                                                \
  }
  EXH_M(0) EXH_M(1) EXH_M(2) EXH_M(3) EXH_M(4) EXH_M(5) EXH_M(6)
  EXH_M(7) EXH_M(8) EXH_M(9) EXH_M(10) EXH_M(11) EXH_M(12)
#undef EXH_M
  return hipSuccess;
}

}  // namespace exh
