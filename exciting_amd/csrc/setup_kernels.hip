// FP64 kernels of the H/O setup (hamiltonsetup.f90 + overlapsetup.f90 +
// hmlistln.f90 + olpistln.f90 + hamiltonoverlapocopy_UL.f90).
//
// Data layout in HBM (all column-major complex128 as double2):
//   apwalm(ngkmax, apwordmax, lmmax, natmtot)  the caller's matching coeffs.
//     Read as A^T: for stacked K index k = (atom, f) the column
//     apwalm + colA[k] holds A_f(G) for G = 0..ngp-1 (contiguous).  When
//     colA[k] = k * ngkmax (uniform apword, lmaxmat = lmaxapw) apwalm is used
//     in place as a K-major matrix; otherwise it is packed into At first.
//   Bt[k * ldb + G] = (aa_atom . A_atom)(f, G)        per-call scratch.
//   H, O (ldh = rank): written once, full square, upper computed, lower
//     mirrored (lower := conj(upper)).
//
// MFMA: v_mfma_f64_16x16x4f64.  Operand lanes: A[i=lane%16][k=lane/16],
// B[k=lane/16][j=lane%16]; result D[i=lane/16+4*r][j=lane%16], r=0..3
// (established by the parity tests against the oracle).
#include <cstring>

#include "exh_device.h"

namespace exh {
namespace {

typedef double dx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ dx4 mfma(double a, double b, dx4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int TM = 64;    // output tile edge (rows and columns)
// K slab staged in LDS per step (bt), a multiple of 4: 8 measured 2.5 %
// faster than 16 and 9 % faster than 32 at the 64-atom shape (same sums)
constexpr int TK = 8;
// K slab of the single-k H/O tile: 8 deep, as the grouped kernel's default
// plan; 183 ms -> 161 ms per 64-atom k-point against 16, same sums in the
// same order
constexpr int TKH = 8;
constexpr int NTH = 256;  // 4 waves; wave w owns a 32x32 quadrant

// ---------------------------------------------------------------------------
// ivg(:, igpig(i)) for the G+k set; w = 0 flags an igpig outside 1..ngrtot.
// ---------------------------------------------------------------------------
template <typename PI, typename PO>
__device__ __forceinline__ void ivgk_body(int i, int ngp, PI __restrict__ igpig,
                                          const int *__restrict__ ivg, int ngrtot,
                                          PO __restrict__ out) {
  if (i >= ngp) return;
  int g = igpig[i] - 1;
  int4 v = make_int4(0, 0, 0, 0);
  if (g >= 0 && g < ngrtot) {
    v.x = ivg[3 * g];
    v.y = ivg[3 * g + 1];
    v.z = ivg[3 * g + 2];
    v.w = 1;
  }
  st(out + i, v);
}

__global__ void ivgk_kernel(int ngp, const int *__restrict__ igpig,
                            const int *__restrict__ ivg, int ngrtot,
                            int4 *__restrict__ out) {
  ivgk_body(blockIdx.x * blockDim.x + threadIdx.x, ngp, igpig, ivg, ngrtot, out);
}

// Grouped over k-points: blockIdx.y selects the descriptor.
__global__ void ivgk_grouped_kernel(const KDesc *__restrict__ kd) {
  const KDesc &d = kd[blockIdx.y];
  const SDesc &S = *d.sd;
  ivgk_body(blockIdx.x * blockDim.x + threadIdx.x, d.ngp, gptr(d.igpig), S.ivg, S.ngrtot,
            gptr(d.ivgk));
}

// ---------------------------------------------------------------------------
// Interstitial lookup table over the ivgig box (one gather per element).
// ---------------------------------------------------------------------------
__global__ void ist_table_kernel(int nbox, const int *__restrict__ ivgig,
                                 int ngvec, const double2 *__restrict__ cf,
                                 const double2 *__restrict__ vf,
                                 const double2 *__restrict__ mf,
                                 const double2 *__restrict__ m2, int relnone,
                                 int h1on, double a2, IstEntry *__restrict__ T) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbox) return;
  int ig = ivgig[b];
  IstEntry e;
  e.veff = z2(0, 0);
  e.kin = z2(0, 0);
  e.ocf = z2(0, 0);
  e.okin = z2(0, 0);
  if (ig > 0 && ig <= ngvec) {
    double2 c = cf[ig - 1];
    e.veff = vf[ig - 1];
    e.ocf = c;
    e.kin = relnone ? c : mf[ig - 1];
    if (h1on) {
      double2 m = m2[ig - 1];
      e.okin = z2(a2 * m.x, a2 * m.y);
    }
  }
  T[b] = e;
}

// ---------------------------------------------------------------------------
// At[k * ldb + G] = apwalm[colA[k] + G]: the K-major operand, packed only
// when the stacked (atom, l, m, io) columns are not already equally spaced
// in apwalm (lmaxmat < lmaxapw or an apword(l) below apwordmax).
// ---------------------------------------------------------------------------
__global__ void pack_kernel(int ngp, const double2 *__restrict__ apwalm,
                            const long long *__restrict__ colA,
                            double2 *__restrict__ at, int ldb) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (g < ngp) at[(size_t)k * ldb + g] = apwalm[colA[k] + g];
}

__global__ void pack_grouped_kernel(const KDesc *__restrict__ kd) {
  const KDesc &d = kd[blockIdx.z];
  const SDesc &S = *d.sd;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (!d.at || k >= S.Ktot) return;  // equally spaced columns: nothing to pack
  if (g < d.ngp)
    st(gptr(d.at) + (size_t)k * S.ldb + g, ld(gptr((const double2 *)d.apwalm) + S.colA[k] + g));
}

// ---------------------------------------------------------------------------
// Bt = (aa . A)^T per atom:  Bt[Koff+f][G] = sum_f' A^T[G][f'] aa(f, f')
// (hamiltonsetup.f90:75-88, zm = aa * apwi).  X holds A^T K-major with
// leading dimension ldx.  Grid (G tiles, f tiles, atoms).
// ---------------------------------------------------------------------------
template <typename PX, typename PB, typename PAA = const double2 *>
__device__ __forceinline__ void bt_body(
    int ias, int I0, int F0, int ngp, PX __restrict__ X, long long ldx,
    const int *__restrict__ Koff, const int *__restrict__ Ms,
    PAA __restrict__ aa, int maxaa, PB __restrict__ bt, int ldb) {
  const int M = Ms[ias];
  if (F0 >= M || I0 >= ngp) return;
  const int koff = Koff[ias];
  const PAA aat = aa + (size_t)ias * maxaa * maxaa;

  __shared__ double sXr[TK][TM + 16], sXi[TK][TM + 16], sYr[TK][TM + 16], sYi[TK][TM + 16];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = w >> 1, wc = w & 1, lr = lane & 15, lk = lane >> 4;
  const int ii = tid & 63, kq = tid >> 6;
  const bool vi = I0 + ii < ngp, vf = F0 + ii < M;
  const PX xp = X + (size_t)koff * ldx + (vi ? I0 + ii : ngp - 1);
  const PAA yp = aat + (vf ? F0 + ii : M - 1);

  dx4 cre[2][2], cim[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      cre[a][b] = dx4{0, 0, 0, 0};
      cim[a][b] = dx4{0, 0, 0, 0};
    }

  // the validity factors are applied at the LDS store, after the slab's
  // MFMAs, so the prefetch loads stay in flight meanwhile
  constexpr int QB = TK / 4;  // slab rows per thread (4 row groups of 64 threads)
  double2 px[QB], py[QB];
  double fxv[QB], fyv[QB];
  auto load = [&](int kb) {
#pragma unroll
    for (int q = 0; q < QB; q++) {
      const int k = kb + kq + 4 * q;
      const bool kv = k < M;
      const int kc = kv ? k : M - 1;
      px[q] = ld(xp + (size_t)kc * ldx);
      py[q] = ld(yp + (size_t)maxaa * kc);
      fxv[q] = (kv && vi) ? 1.0 : 0.0;
      fyv[q] = (kv && vf) ? 1.0 : 0.0;
    }
  };
  load(0);
  for (int kb = 0; kb < M; kb += TK) {
#pragma unroll
    for (int q = 0; q < QB; q++) {
      const int kk = kq + 4 * q;
      sXr[kk][ii] = fxv[q] * px[q].x;
      sXi[kk][ii] = fxv[q] * px[q].y;
      sYr[kk][ii] = fyv[q] * py[q].x;
      sYi[kk][ii] = fyv[q] * py[q].y;
    }
    __syncthreads();
    if (kb + TK < M) load(kb + TK);
#pragma unroll
    for (int k4 = 0; k4 < TK / 4; k4++) {
      const int kk = 4 * k4 + lk;
      double xr[2], xi[2], nxi[2], yr[2], yi[2];
#pragma unroll
      for (int a = 0; a < 2; a++) {
        xr[a] = sXr[kk][wr * 32 + a * 16 + lr];
        xi[a] = sXi[kk][wr * 32 + a * 16 + lr];
        nxi[a] = -xi[a];
      }
#pragma unroll
      for (int b = 0; b < 2; b++) {
        yr[b] = sYr[kk][wc * 32 + b * 16 + lr];
        yi[b] = sYi[kk][wc * 32 + b * 16 + lr];
      }
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          cre[a][b] = mfma(xr[a], yr[b], cre[a][b]);
          cre[a][b] = mfma(nxi[a], yi[b], cre[a][b]);
          cim[a][b] = mfma(xr[a], yi[b], cim[a][b]);
          cim[a][b] = mfma(xi[a], yr[b], cim[a][b]);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int i = I0 + wr * 32 + a * 16 + lk + 4 * r;
        int f = F0 + wc * 32 + b * 16 + lr;
        if (i < ngp && f < M)
          st(bt + (size_t)(koff + f) * ldb + i, z2(cre[a][b][r], cim[a][b][r]));
      }
}

__global__ __launch_bounds__(NTH) void bt_kernel(
    int ngp, const double2 *__restrict__ X, long long ldx,
    const int *__restrict__ Koff, const int *__restrict__ Ms,
    const double2 *__restrict__ aa, int maxaa, double2 *__restrict__ bt, int ldb) {
  bt_body(blockIdx.z, blockIdx.x * TM, blockIdx.y * TM, ngp, X, ldx, Koff, Ms, aa, maxaa, bt,
          ldb);
}

// Grouped: blockIdx.z = ias + natmax * (k-point), natmax the largest atom
// count of the group's structures.
__global__ __launch_bounds__(NTH) void bt_grouped_kernel(const KDesc *__restrict__ kd,
                                                         int natmax) {
  const int k = blockIdx.z / natmax, ias = blockIdx.z - k * natmax;
  const KDesc &d = kd[k];
  const SDesc &S = *d.sd;
  if (ias >= S.natmtot) return;
  bt_body(ias, blockIdx.x * TM, blockIdx.y * TM, d.ngp, gptr(d.x), S.ldx, S.Koff, S.Ms,
          gptr(S.aa), S.maxaa, gptr(d.bt), S.ldb);
}

// ---------------------------------------------------------------------------
// A2t = ((1 + h1) . A)^T: the overlap operand when h1aa is present
// (overlapsetup.f90:82-94).  h1aa(io1, io2, l, ias) couples the radial
// orders of one (l, m).  One thread per (G, atom, f2).
// ---------------------------------------------------------------------------
template <typename PA, typename PT>
__device__ __forceinline__ void a2t_body(
    int g, int ias, int f2, int ngp, PA __restrict__ apwalm,
    const long long *__restrict__ colA, const int *__restrict__ Koff,
    const int *__restrict__ Ms, const int *__restrict__ kinfo, const double *__restrict__ h1aa,
    int pmax, int lmaxapw, PT __restrict__ a2t, int ldb) {
  if (g >= ngp || f2 >= Ms[ias]) return;
  int koff = Koff[ias];
  // kinfo[k] = io | p << 4 | l << 8 | (first f of this (l,m)) << 16
  const int info = kinfo[koff + f2];
  const int io2 = info & 15, p = (info >> 4) & 15, l = (info >> 8) & 255, fb = info >> 16;
  double2 v = z2(0, 0);
  for (int io1 = 0; io1 < p; io1++) {
    double hv = h1aa[io1 + pmax * (io2 + (size_t)pmax * (l + (size_t)(lmaxapw + 1) * ias))];
    double2 a = ld(apwalm + colA[koff + fb + io1] + g);
    v.x += hv * a.x;
    v.y += hv * a.y;
  }
  double2 a = ld(apwalm + colA[koff + f2] + g);
  v.x += a.x;
  v.y += a.y;
  st(a2t + (size_t)(koff + f2) * ldb + g, v);
}

__global__ void a2t_kernel(int ngp, int natmtot, const double2 *__restrict__ apwalm,
                           const long long *__restrict__ colA,
                           const int *__restrict__ Koff, const int *__restrict__ Ms,
                           const int *__restrict__ kinfo, const double *__restrict__ h1aa,
                           int pmax, int lmaxapw, double2 *__restrict__ a2t, int ldb) {
  (void)natmtot;
  a2t_body(blockIdx.x * blockDim.x + threadIdx.x, blockIdx.z, blockIdx.y, ngp, apwalm, colA,
           Koff, Ms, kinfo, h1aa, pmax, lmaxapw, a2t, ldb);
}

// Grouped: blockIdx.z = ias + natmax * (k-point).
__global__ void a2t_grouped_kernel(const KDesc *__restrict__ kd, int natmax) {
  const int k = blockIdx.z / natmax, ias = blockIdx.z - k * natmax;
  const KDesc &d = kd[k];
  const SDesc &S = *d.sd;
  if (!S.have_h1aa || ias >= S.natmtot) return;
  a2t_body(blockIdx.x * blockDim.x + threadIdx.x, ias, blockIdx.y, d.ngp,
           gptr((const double2 *)d.apwalm), S.colA, S.Koff, S.Ms, S.kinfo, S.h1aa, S.pmax,
           S.lmaxapw, gptr(d.a2t), S.ldb);
}

// ---------------------------------------------------------------------------
// Fused upper-triangle H and O tiles.
//   H(i,j) = sum_k conj(X[k][i]) YB[k][j]  + IST_H(i,j)   i <= j < ngp
//   O(i,j) = sum_k conj(X[k][i]) YA[k][j]  + IST_O(i,j)
// X = A^T, YB = (aa A)^T, YA = A^T (or ((1+h1) A)^T), all K-major with K
// running over every atom's (l, m, io) (hamiltonsetup.f90:90-103,
// overlapsetup.f90:95-108).  The epilogue adds the interstitial terms
// (hmlistln.f90, olpistln.f90) and writes lower := conj(upper)
// (hamiltonoverlapocopy_UL.f90).  A workgroup of 32 x 32 waves computes one
// TMI x TMJ upper tile; tiles are dealt to the 8 XCDs in contiguous ranges
// so co-resident tiles share column slabs in L2.  K slabs of TK are
// double-buffered in LDS (one barrier per slab) and prefetched two slabs
// ahead into registers with clamped, branch-free loads.  LDS rows are padded
// by 16 doubles (80 or 144) so the four k rows read by one MFMA operand
// fetch alternate LDS bank halves.
// ---------------------------------------------------------------------------
template <int TMI, int TMJ, int TK>
struct HoShape {
  static constexpr int WJ = TMJ / 32;               // waves across a tile row
  static constexpr int NW = (TMI / 32) * WJ;        // waves per workgroup
  static constexpr int NTH = 64 * NW;
  static constexpr int PX = TMI + 16, PY = TMJ + 16;  // padded LDS rows
  static constexpr int RX = NTH / TMI, RY = NTH / TMJ;  // k rows per load pass
  static constexpr int QX = TK / RX, QY = TK / RY;      // loads per thread
  static_assert(NTH % TMI == 0 && NTH % TMJ == 0, "tile rows must divide the workgroup");
  static_assert(TK % RX == 0 && TK % RY == 0 && TK % 4 == 0, "K slab");
};

// One TMI x TMJ upper tile (I0, J0) of H and O for one k-point.  WH = false
// computes O alone (overlapsetup): YB and H are not touched.
template <int TMI, int TMJ, int TK, typename PX, typename PH, typename B, bool WH = true>
__device__ __forceinline__ void ho_tile(
    int ngp, int Ktot, PX __restrict__ X, long long ldx, PX __restrict__ YB, long long ldb,
    PX __restrict__ YA, long long lda, const B &box, PH __restrict__ H,
    PH __restrict__ O, long long ldh, int accumulate, int I0, int J0) {
  using S = HoShape<TMI, TMJ, TK>;
  // [buffer][re, im][k][row] of X; [buffer][Br, Bi, Ar, Ai][k][row] of YB, YA
  __shared__ double sX[2][2][TK][S::PX];
  __shared__ double sY[2][4][TK][S::PY];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = w / S::WJ, wc = w % S::WJ, lr = lane & 15, lk = lane >> 4;
  const int xr_ = tid % TMI, xk = tid / TMI;
  const int yc = tid % TMJ, yk = tid / TMJ;
  const int ri = min(I0 + xr_, ngp - 1), rj = min(J0 + yc, ngp - 1);
  const PX xp = X + ri;
  const PX bp = YB + rj;
  const PX ap = YA + rj;

  dx4 hre[2][2], him[2][2], ore[2][2], oim[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      hre[a][b] = dx4{0, 0, 0, 0};
      him[a][b] = dx4{0, 0, 0, 0};
      ore[a][b] = dx4{0, 0, 0, 0};
      oim[a][b] = dx4{0, 0, 0, 0};
    }

  // Rows at or past ngp only feed output rows/columns the epilogue never
  // writes, so addresses are clamped but values are not masked; the K tail
  // is zeroed by a 0/1 factor so that every load is unconditional.  The
  // factor is applied when the slab is stored to LDS, one slab after the
  // load was issued, so the loads stay in flight across a whole slab of
  // MFMAs instead of being waited for right away.
  double2 px[S::QX], pb[S::QY], pa[S::QY];
  double fx[S::QX], fy[S::QY];
  auto load = [&](int kb) {
#pragma unroll
    for (int q = 0; q < S::QX; q++) {
      const int k = kb + xk + S::RX * q;
      fx[q] = k < Ktot ? 1.0 : 0.0;
      const size_t kc = k < Ktot ? k : Ktot - 1;
      px[q] = ld(xp + kc * ldx);
    }
#pragma unroll
    for (int q = 0; q < S::QY; q++) {
      const int k = kb + yk + S::RY * q;
      fy[q] = k < Ktot ? 1.0 : 0.0;
      const size_t kc = k < Ktot ? k : Ktot - 1;
      if (WH) pb[q] = ld(bp + kc * ldb);
      pa[q] = ld(ap + kc * lda);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < S::QX; q++) {
      const int kk = xk + S::RX * q;
      sX[buf][0][kk][xr_] = fx[q] * px[q].x;
      sX[buf][1][kk][xr_] = fx[q] * px[q].y;
    }
#pragma unroll
    for (int q = 0; q < S::QY; q++) {
      const int kk = yk + S::RY * q;
      if (WH) {
        sY[buf][0][kk][yc] = fy[q] * pb[q].x;
        sY[buf][1][kk][yc] = fy[q] * pb[q].y;
      }
      sY[buf][2][kk][yc] = fy[q] * pa[q].x;
      sY[buf][3][kk][yc] = fy[q] * pa[q].y;
    }
  };
  auto compute = [&](int buf) {
#pragma unroll
    for (int k4 = 0; k4 < TK / 4; k4++) {
      const int kk = 4 * k4 + lk;
      double xr[2], xi[2], nxi[2], br[2], bi[2], ar[2], ai[2];
#pragma unroll
      for (int a = 0; a < 2; a++) {
        xr[a] = sX[buf][0][kk][wr * 32 + a * 16 + lr];
        xi[a] = sX[buf][1][kk][wr * 32 + a * 16 + lr];
        nxi[a] = -xi[a];
      }
#pragma unroll
      for (int b = 0; b < 2; b++) {
        if (WH) {
          br[b] = sY[buf][0][kk][wc * 32 + b * 16 + lr];
          bi[b] = sY[buf][1][kk][wc * 32 + b * 16 + lr];
        }
        ar[b] = sY[buf][2][kk][wc * 32 + b * 16 + lr];
        ai[b] = sY[buf][3][kk][wc * 32 + b * 16 + lr];
      }
      // conj(x) * y = (xr yr + xi yi) + i (xr yi - xi yr)
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          if (WH) {
            hre[a][b] = mfma(xr[a], br[b], hre[a][b]);
            him[a][b] = mfma(xr[a], bi[b], him[a][b]);
          }
          ore[a][b] = mfma(xr[a], ar[b], ore[a][b]);
          oim[a][b] = mfma(xr[a], ai[b], oim[a][b]);
        }
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          if (WH) {
            hre[a][b] = mfma(xi[a], bi[b], hre[a][b]);
            him[a][b] = mfma(nxi[a], br[b], him[a][b]);
          }
          ore[a][b] = mfma(xi[a], ai[b], ore[a][b]);
          oim[a][b] = mfma(nxi[a], ar[b], oim[a][b]);
        }
    }
  };

  // Loads past the last slab are clamped and zero-weighted, so the loop body
  // is one straight-line block (no per-iteration conditions).  The closing
  // barrier of the last slab also frees the LDS for the next tile of a
  // persistent workgroup.
  const int nslab = (Ktot + TK - 1) / TK;
  load(0);
  store(0);
  __syncthreads();
  load(TK);
  for (int sl = 0; sl < nslab; sl++) {
    compute(sl & 1);
    // keep the scaling and LDS store of the prefetched slab after this
    // slab's MFMAs, so its loads are waited for only here
    __builtin_amdgcn_sched_barrier(0);
    store((sl + 1) & 1);
    load((sl + 2) * TK);
    __syncthreads();
  }

  // ---- epilogue: accumulate, interstitial gather, write upper + mirror ----
  // (accumulate adds the old upper value before the interstitial term, the
  // reference order: zgemm beta = 1, then hmlistln/olpistln)
#pragma unroll
  for (int b = 0; b < 2; b++) {
    const int j = J0 + wc * 32 + b * 16 + lr;
    if (j >= ngp) continue;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = I0 + wr * 32 + a * 16 + lk + 4 * r;
        if (i >= ngp || i > j) continue;
        double hr = hre[a][b][r], hi = him[a][b][r];
        double orr = ore[a][b][r], oi = oim[a][b][r];
        if (accumulate) {
          const size_t up = (size_t)i + (size_t)j * ldh;
          if (WH) {
            const double2 h0 = ld(H + up);
            hr += h0.x;
            hi += h0.y;
          }
          const double2 o0 = ld(O + up);
          orr += o0.x;
          oi += o0.y;
        }
        ist_add(box, i, j, hr, hi, orr, oi);
        if (WH) store_upper(H, O, ldh, i, j, hr, hi, orr, oi, 0);
        else store_upper_o(O, ldh, i, j, orr, oi);
      }
  }
}

// Single k-point: one workgroup per upper 64 x 64 tile pair I <= J, dealt to
// the 8 XCDs in contiguous ranges, J-major within each range.  O alone runs
// on o_kernel's larger tiles instead.
template <bool WH>
__global__ __launch_bounds__(NTH) void ho_kernel(
    int ngp, int Ktot, const double2 *__restrict__ X, long long ldx,
    const double2 *__restrict__ YB, long long ldb,
    const double2 *__restrict__ YA, long long lda,
    const int4 *__restrict__ ivgk, const double *__restrict__ vgpc,
    const IstEntry *__restrict__ T, int lo0, int lo1, int lo2, int n0, int n1,
    int n2, double2 *__restrict__ H, double2 *__restrict__ O, long long ldh,
    int accumulate) {
  const int bid = blockIdx.x, nb = gridDim.x;
  const int xcd = bid & 7, i8 = bid >> 3, q8 = nb >> 3, r8 = nb & 7;
  const int t = xcd * q8 + (xcd < r8 ? xcd : r8) + i8;
  int J = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((J + 1) * (J + 2) / 2 <= t) J++;
  while (J * (J + 1) / 2 > t) J--;
  const int I = t - J * (J + 1) / 2;
  IstBox box{ivgk, vgpc, T, lo0, lo1, lo2, n0, n1, n2};
  ho_tile<TM, TM, TKH, const double2 *, double2 *, IstBox, WH>(
      ngp, Ktot, X, ldx, YB, ldb, YA, lda, box, H, O, ldh, accumulate, I * TM, J * TM);
}

// overlapsetup alone (O without the Hamiltonian's products): per operand
// load it does half the MFMA work of the fused H + O tile, so it takes
// TO x TO tiles (upper tile pairs I <= J, XCD-contiguous like ho_kernel) to
// raise the flops per loaded byte.
template <int TO, int TKO>
__global__ __launch_bounds__(64 * (TO / 32) * (TO / 32)) void o_kernel(
    int ngp, int Ktot, const double2 *__restrict__ X, long long ldx,
    const double2 *__restrict__ YA, long long lda, const int4 *__restrict__ ivgk,
    const double *__restrict__ vgpc, const IstEntry *__restrict__ T, int lo0, int lo1, int lo2,
    int n0, int n1, int n2, double2 *__restrict__ O, long long ldh, int accumulate) {
  const int bid = blockIdx.x, nb = gridDim.x;
  const int xcd = bid & 7, i8 = bid >> 3, q8 = nb >> 3, r8 = nb & 7;
  const int t = xcd * q8 + (xcd < r8 ? xcd : r8) + i8;
  int J = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((J + 1) * (J + 2) / 2 <= t) J++;
  while (J * (J + 1) / 2 > t) J--;
  const int I = t - J * (J + 1) / 2;
  IstBox box{ivgk, vgpc, T, lo0, lo1, lo2, n0, n1, n2};
  ho_tile<TO, TO, TKO, const double2 *, double2 *, IstBox, false>(
      ngp, Ktot, X, ldx, X, ldx, YA, lda, box, nullptr, O, ldh, accumulate, I * TO, J * TO);
}

// Grouped over k-points: a list of (k, I0, J0) tiles (k < 0 = padding) is
// walked by gridDim.x workgroups, workgroup b taking entries b, b + gridDim.x,
// ...  With gridDim.x a multiple of 8, list position t runs on XCD t % 8, so
// the host orders the list to give each XCD contiguous tile ranges.  Every
// tile of one shape accumulates K in the same order whatever the list order
// or grid, so all plans of a shape give bitwise identical results.
template <int TMI, int TMJ, int TK>
__global__ __launch_bounds__(64 * (TMI / 32) * (TMJ / 32)) void ho_grouped_kernel(
    const KDesc *__restrict__ kd, const int4 *__restrict__ tiles, int ntiles, int accumulate) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int4 e = tiles[t];
    const int k = __builtin_amdgcn_readfirstlane(e.x);
    if (k < 0) continue;
    const int I0 = __builtin_amdgcn_readfirstlane(e.y);
    const int J0 = __builtin_amdgcn_readfirstlane(e.z);
    const KDesc &d = kd[k];
    const SDesc &S = *d.sd;
    const int Ktot = __builtin_amdgcn_readfirstlane(S.Ktot);
    const long long ldx = S.ldx, ldb = S.ldb;
    const IstBoxT<const EXH_G int4 *, const EXH_G double *> box{
        gptr((const int4 *)d.ivgk), gptr(d.vgpc), S.ist, S.lo0, S.lo1, S.lo2, S.n0, S.n1, S.n2};
    const EXH_G double2 *X = gptr(d.x);
    const EXH_G double2 *YB = gptr((const double2 *)d.bt);
    const bool use_a2t = S.have_h1aa != 0;
    const EXH_G double2 *YA = use_a2t ? gptr((const double2 *)d.a2t) : X;
    const long long lda = use_a2t ? ldb : ldx;
    ho_tile<TMI, TMJ, TK>(d.ngp, Ktot, X, ldx, YB, ldb, YA, lda, box, gptr(d.h), gptr(d.o),
                          d.ldh, accumulate, I0, J0);
  }
}

// ---------------------------------------------------------------------------
// APW-lo border (hamiltonsetup.f90:124-151, overlapsetup.f90:116-133).
// One thread per (G, atom):  H(lo rows, G) = loa . A   (+ mirror),
//                            O(G, lo cols) = sum_io conj(A_io,lm s_io) (+ mirror)
// with s = oalo + h1loa.
// ---------------------------------------------------------------------------
template <int MAXLO, typename PA, typename PH>
__device__ __forceinline__ void lo_border_body(int g, int ias, int ngp, PA __restrict__ apwalm,
                                               const LoTables &t, PH __restrict__ H,
                                               PH __restrict__ O, long long ldh,
                                               int accumulate) {
  // every table index below is uniform across the wave: scalar loads
  const EXH_C long long *__restrict__ colA = cptr(t.colA);
  const EXH_C int *__restrict__ nlorb = cptr(t.nlorb);
  const EXH_C int *__restrict__ lorbl = cptr(t.lorbl);
  const EXH_C int *__restrict__ apword = cptr(t.apword);
  const EXH_C double *__restrict__ salo = cptr(t.salo);
  const int maxnlo = t.maxnlo, nlomax = t.nlomax, lmaxapw = t.lmaxapw, pmax = t.pmax,
            lmmax = t.lmmax, ngkmax = t.ngkmax;
  const EXH_C int *__restrict__ Koff = cptr(t.Koff);
  const EXH_C int *__restrict__ Ms = cptr(t.Ms);
  const EXH_C int *__restrict__ lobase = cptr(t.lobase);
  const EXH_C double2 *__restrict__ loa = cptr(t.loa);
  const int maxaa = t.maxaa;
  const int is = cptr(t.atom_species)[ias];
  const int los = cptr(t.losize)[is];
  if (los == 0 || g >= ngp) return;
  const int r0 = ngp + lobase[ias];
  const EXH_C long long *col = colA + Koff[ias];
  const int M = Ms[ias];
  const EXH_C double2 *L = loa + (size_t)ias * maxnlo * maxaa;

  if (H) {  // null for overlapsetup alone
  double2 acc[MAXLO];
#pragma unroll
  for (int r = 0; r < MAXLO; r++) acc[r] = z2(0, 0);
  for (int f = 0; f < M; f++) {
    const double2 a = ld(apwalm + col[f] + g);
    const EXH_C double2 *Lf = L + (size_t)maxnlo * f;
#pragma unroll
    for (int r = 0; r < MAXLO; r++)
      if (r < los) {
        const double2 lv = ld(Lf + r);
        acc[r].x += a.x * lv.x - a.y * lv.y;
        acc[r].y += a.x * lv.y + a.y * lv.x;
      }
  }
#pragma unroll
  for (int r = 0; r < MAXLO; r++)
    if (r < los) {
      const size_t e = (size_t)(r0 + r) + (size_t)g * ldh;
      double2 v = acc[r];
      if (accumulate) {
        double2 h0 = ld(H + e);
        v.x += h0.x;
        v.y += h0.y;
      }
      st(H + e, v);
      st(H + (size_t)g + (size_t)(r0 + r) * ldh, z2(v.x, -v.y));
    }
  }

  int c = r0;
  for (int ilo = 0; ilo < nlorb[is]; ilo++) {
    const int l = lorbl[ilo + nlomax * is];
    const int p = apword[l + (lmaxapw + 1) * is];
    for (int mi = 0; mi <= 2 * l; mi++, c++) {
      const int lm = l * l + mi;
      const size_t e = (size_t)g + (size_t)c * ldh;
      double2 v = accumulate ? ld(O + e) : z2(0, 0);
      for (int io = 0; io < p; io++) {
        const double2 a =
            ld(apwalm + g + (size_t)ngkmax * (io + (size_t)pmax * (lm + (size_t)lmmax * ias)));
        const double s = salo[io + pmax * (ilo + (size_t)nlomax * ias)];
        v.x += a.x * s;
        v.y += -(a.y * s);
      }
      st(O + e, v);
      st(O + (size_t)c + (size_t)g * ldh, z2(v.x, -v.y));
    }
  }
}

template <int MAXLO>
__global__ __launch_bounds__(64) void lo_border_kernel(int ngp, const double2 *__restrict__ apwalm,
                                                       LoTables t, double2 *__restrict__ H,
                                                       double2 *__restrict__ O, long long ldh,
                                                       int accumulate) {
  lo_border_body<MAXLO>(blockIdx.x * 64 + threadIdx.x, blockIdx.y, ngp, apwalm, t, H, O, ldh,
                        accumulate);
}

// Grouped: blockIdx.z selects the k-point, blockIdx.y the atom.
template <int MAXLO>
__global__ __launch_bounds__(64) void lo_border_grouped_kernel(const KDesc *__restrict__ kd,
                                                               int accumulate) {
  const KDesc &d = kd[blockIdx.z];
  const EXH_C SDesc &S = *cptr(d.sd);
  if ((int)blockIdx.y >= S.natmtot || S.nlotot == 0) return;
  // field by field: scalar loads from the descriptor
  const EXH_C LoTables &L = S.lo;
  LoTables t;
  t.colA = L.colA;
  t.Koff = L.Koff;
  t.Ms = L.Ms;
  t.atom_species = L.atom_species;
  t.losize = L.losize;
  t.lobase = L.lobase;
  t.loa = L.loa;
  t.maxnlo = L.maxnlo;
  t.maxaa = L.maxaa;
  t.nlorb = L.nlorb;
  t.lorbl = L.lorbl;
  t.nlomax = L.nlomax;
  t.apword = L.apword;
  t.lmaxapw = L.lmaxapw;
  t.pmax = L.pmax;
  t.lmmax = L.lmmax;
  t.ngkmax = L.ngkmax;
  t.salo = L.salo;
  lo_border_body<MAXLO>(blockIdx.x * 64 + threadIdx.x, blockIdx.y, d.ngp,
                        gptr((const double2 *)d.apwalm), t, gptr(d.h), gptr(d.o), d.ldh,
                        accumulate);
}

// ---------------------------------------------------------------------------
// lo-lo block (hamiltonsetup.f90:153-161, overlapsetup.f90:134-151) over the
// whole nlotot x nlotot corner: H += lolo (same atom), O += ololo + h1lolo
// (same atom, same l, same m).  Upper threads write both triangles.
// ---------------------------------------------------------------------------
template <typename PH>
__device__ __forceinline__ void lolo_body(int r, int c, int ngp, int nlotot,
                                          const LoRow *__restrict__ lorow,
                                          const int *__restrict__ lobase,
                                          const double2 *__restrict__ lolo, int maxnlo,
                                          const double *__restrict__ ololo_eff, int nlomax,
                                          PH __restrict__ H, PH __restrict__ O, long long ldh,
                                          int accumulate) {
  if (r >= nlotot || c >= nlotot || r > c) return;
  const LoRow P = lorow[r], Q = lorow[c];
  const size_t up = (size_t)(ngp + r) + (size_t)(ngp + c) * ldh;
  double2 hv = z2(0, 0), ov = z2(0, 0);
  const bool wh = H != nullptr;  // null for overlapsetup alone
  if (accumulate) {
    if (wh) hv = ld(H + up);
    ov = ld(O + up);
  }
  if (P.ias == Q.ias) {
    const int b = lobase[P.ias];
    const double2 lv = lolo[(r - b) + (size_t)maxnlo * ((c - b) + (size_t)maxnlo * P.ias)];
    hv.x += lv.x;
    hv.y += lv.y;
    if (P.l == Q.l && P.mi == Q.mi)
      ov.x += ololo_eff[P.ilo + nlomax * (Q.ilo + (size_t)nlomax * P.ias)];
  }
  if (wh) st(H + up, hv);
  st(O + up, ov);
  if (r != c) {
    const size_t lo = (size_t)(ngp + c) + (size_t)(ngp + r) * ldh;
    if (wh) st(H + lo, z2(hv.x, -hv.y));
    st(O + lo, z2(ov.x, -ov.y));
  }
}

__global__ void lolo_kernel(int ngp, int nlotot, const LoRow *__restrict__ lorow,
                            const int *__restrict__ lobase,
                            const double2 *__restrict__ lolo, int maxnlo,
                            const double *__restrict__ ololo_eff, int nlomax,
                            double2 *__restrict__ H, double2 *__restrict__ O,
                            long long ldh, int accumulate) {
  lolo_body(blockIdx.x * 16 + (threadIdx.x & 15), blockIdx.y * 16 + (threadIdx.x >> 4), ngp,
            nlotot, lorow, lobase, lolo, maxnlo, ololo_eff, nlomax, H, O, ldh, accumulate);
}

__global__ void lolo_grouped_kernel(const KDesc *__restrict__ kd, int accumulate) {
  const KDesc &d = kd[blockIdx.z];
  const SDesc &S = *d.sd;
  lolo_body(blockIdx.x * 16 + (threadIdx.x & 15), blockIdx.y * 16 + (threadIdx.x >> 4), d.ngp,
            S.nlotot, S.lorow, S.lobase, S.lolo, S.maxnlo, S.ololo_eff, S.nlomax, gptr(d.h),
            gptr(d.o), d.ldh, accumulate);
}

inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

hipError_t launch_ivgk(hipStream_t s, int ngp, const int *igpig, const int *ivg,
                       int ngrtot, int4 *ivgk) {
  if (ngp <= 0) return hipSuccess;
  hipLaunchKernelGGL(ivgk_kernel, dim3(cdiv(ngp, 256)), dim3(256), 0, s, ngp,
                     igpig, ivg, ngrtot, ivgk);
  return hipGetLastError();
}

hipError_t launch_ist_table(hipStream_t s, const exh_context *c,
                            const double2 *cf, const double2 *vf,
                            const double2 *mf, const double2 *m2, int relnone,
                            int h1on) {
  const int n0 = c->intgv[3] - c->intgv[0] + 1, n1 = c->intgv[4] - c->intgv[1] + 1,
            n2 = c->intgv[5] - c->intgv[2] + 1;
  const int nbox = n0 * n1 * n2;
  const double alpha = 1.0 / 137.03599911, a2 = 0.5 * alpha * alpha;
  hipLaunchKernelGGL(ist_table_kernel, dim3(cdiv(nbox, 256)), dim3(256), 0, s, nbox,
                     (const int *)c->ivgig.p, c->ngvec, cf, vf, mf, m2, relnone,
                     h1on, a2, (IstEntry *)c->ist.p);
  return hipGetLastError();
}

// K-major view of A for the MFMA kernels: apwalm itself when its stacked
// columns are equally spaced (ld = ngkmax), else the packed copy `at`
// (Ktot x ldb scratch).
static const double2 *k_major_A(hipStream_t s, const exh_context *c, int ngp,
                                const double2 *apwalm, double2 *at, long long *ld,
                                hipError_t *err) {
  *err = hipSuccess;
  if (c->contig) {
    *ld = c->ngkmax;
    return apwalm;
  }
  const int ldb = (c->ngkmax + 63) / 64 * 64;
  *ld = ldb;
  hipLaunchKernelGGL(pack_kernel, dim3(cdiv(ngp, 256), c->Ktot), dim3(256), 0, s, ngp,
                     apwalm, (const long long *)c->colA.p, at, ldb);
  *err = hipGetLastError();
  return at;
}

hipError_t launch_pack(hipStream_t s, const exh_context *c, int ngp,
                       const double2 *apwalm, double2 *at) {
  if (ngp <= 0 || c->contig || c->Ktot == 0) return hipSuccess;
  long long ld;
  hipError_t e;
  (void)k_major_A(s, c, ngp, apwalm, at, &ld, &e);
  return e;
}

static const double2 *k_major_view(const exh_context *c, const double2 *apwalm,
                                   const double2 *at, long long *ld) {
  if (c->contig) {
    *ld = c->ngkmax;
    return apwalm;
  }
  *ld = (c->ngkmax + 63) / 64 * 64;
  return at;
}

const double2 *k_major_operand(const exh_context *c, const double2 *apwalm, const double2 *at,
                               long long *ld) {
  return k_major_view(c, apwalm, at, ld);
}

hipError_t launch_bt(hipStream_t s, const exh_context *c, int ngp,
                     const double2 *apwalm, const double2 *at, double2 *bt, int ldb) {
  if (ngp <= 0 || c->natmtot <= 0) return hipSuccess;
  long long ldx;
  const double2 *X = k_major_view(c, apwalm, at, &ldx);
  dim3 grid(cdiv(ngp, TM), cdiv(c->maxaa, TM), c->natmtot);
  hipLaunchKernelGGL(bt_kernel, grid, dim3(NTH), 0, s, ngp, X, ldx,
                     (const int *)c->katom.p, (const int *)c->katom.p + c->natmtot,
                     (const double2 *)c->aa.p, c->maxaa, bt, ldb);
  return hipGetLastError();
}

hipError_t launch_a2t(hipStream_t s, const exh_context *c, int ngp,
                      const double2 *apwalm, double2 *a2t, int ldb) {
  if (ngp <= 0 || c->natmtot <= 0) return hipSuccess;
  dim3 grid(cdiv(ngp, 256), c->maxaa, c->natmtot);
  hipLaunchKernelGGL(a2t_kernel, grid, dim3(256), 0, s, ngp, c->natmtot, apwalm,
                     (const long long *)c->colA.p, (const int *)c->katom.p,
                     (const int *)c->katom.p + c->natmtot,
                     (const int *)c->katom.p + 2 * c->natmtot, (const double *)c->h1aa.p,
                     c->apwordmax, c->lmaxapw, a2t, ldb);
  return hipGetLastError();
}

hipError_t launch_ho(hipStream_t s, const exh_context *c, int ngp,
                     const double2 *apwalm, const double2 *at, const double2 *bt,
                     const double2 *a2t, int ldb, const int4 *ivgk,
                     const double *vgpc, double2 *h, double2 *o, int ldh,
                     int flags) {
  if (ngp <= 0) return hipSuccess;
  const long long nt = (ngp + TM - 1) / TM;
  const long long npairs = nt * (nt + 1) / 2;
  const int lo0 = c->intgv[0], lo1 = c->intgv[1], lo2 = c->intgv[2];
  const int n0 = c->intgv[3] - lo0 + 1, n1 = c->intgv[4] - lo1 + 1, n2 = c->intgv[5] - lo2 + 1;
  long long ldx;
  const double2 *X = k_major_view(c, apwalm, at, &ldx);
  const double2 *YA = a2t ? a2t : X;
  const long long lda = a2t ? (long long)ldb : ldx;
  if (h)
    hipLaunchKernelGGL(ho_kernel<true>, dim3((unsigned)npairs), dim3(NTH), 0, s, ngp, c->Ktot,
                       X, ldx, bt, (long long)ldb, YA, lda, ivgk, vgpc,
                       (const IstEntry *)c->ist.p, lo0, lo1, lo2, n0, n1, n2, h, o,
                       (long long)ldh, flags & EXH_ACCUMULATE);
  else {  // overlapsetup alone: bt and H are not read
    constexpr int TO = 128, TKO = 8;
    const long long ntO = (ngp + TO - 1) / TO;
    hipLaunchKernelGGL((o_kernel<TO, TKO>), dim3((unsigned)(ntO * (ntO + 1) / 2)),
                       dim3(HoShape<TO, TO, TKO>::NTH), 0, s, ngp, c->Ktot, X, ldx, YA, lda,
                       ivgk, vgpc, (const IstEntry *)c->ist.p, lo0, lo1, lo2, n0, n1, n2, o,
                       (long long)ldh, flags & EXH_ACCUMULATE);
  }
  return hipGetLastError();
}

static LoTables lo_tables(const exh_context *c) {
  const int *ki = (const int *)c->katom.p;
  const int *lo = (const int *)c->lobase_d.p;  // [lobase(natmtot), losize(nspecies), nlorb, lorbl]
  LoTables t;
  t.colA = (const long long *)c->colA.p;
  t.Koff = ki;
  t.Ms = ki + c->natmtot;
  t.atom_species = (const int *)c->atom_species_d.p;
  t.lobase = lo;
  t.losize = lo + c->natmtot;
  t.nlorb = t.losize + c->nspecies;
  t.lorbl = t.nlorb + c->nspecies;
  t.loa = (const double2 *)c->loa.p;
  t.maxnlo = c->maxnlo;
  t.maxaa = c->maxaa;
  t.nlomax = c->nlomax;
  t.apword = (const int *)c->apword_d.p;
  t.lmaxapw = c->lmaxapw;
  t.pmax = c->apwordmax;
  t.lmmax = c->lmmax;
  t.ngkmax = c->ngkmax;
  t.salo = (const double *)c->salo.p;
  return t;
}

hipError_t launch_lo_border(hipStream_t s, const exh_context *c, int ngp,
                            const double2 *apwalm, double2 *h, double2 *o,
                            int ldh, int flags) {
  if (ngp <= 0 || c->nlotot == 0) return hipSuccess;
  dim3 grid(cdiv(ngp, 64), c->natmtot);
  const LoTables t = lo_tables(c);
#define EXH_LOB(N)                                                                     \
  hipLaunchKernelGGL(lo_border_kernel<N>, grid, dim3(64), 0, s, ngp, apwalm, t, h, o,  \
                     (long long)ldh, flags & EXH_ACCUMULATE)
  if (c->maxnlo <= 16) EXH_LOB(16);
  else if (c->maxnlo <= 32) EXH_LOB(32);
  else EXH_LOB(64);
#undef EXH_LOB
  return hipGetLastError();
}

hipError_t launch_lolo(hipStream_t s, const exh_context *c, int ngp, double2 *h,
                       double2 *o, int ldh, int flags) {
  if (c->nlotot == 0) return hipSuccess;
  dim3 grid(cdiv(c->nlotot, 16), cdiv(c->nlotot, 16));
  hipLaunchKernelGGL(lolo_kernel, grid, dim3(256), 0, s, ngp, c->nlotot,
                     (const LoRow *)c->lorow.p, (const int *)c->lobase_d.p,
                     (const double2 *)c->lolo.p, c->maxnlo,
                     (const double *)c->ololo_eff.p, c->nlomax, h, o, (long long)ldh,
                     flags & EXH_ACCUMULATE);
  return hipGetLastError();
}

// hamiltonoverlapocopy_UL.f90:18-22 for one matrix: lower := conj(upper).
__global__ void mirror_kernel(int n, long long ld, double2 *__restrict__ M) {
  const int i = blockIdx.x * 16 + (threadIdx.x & 15), j = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (i >= n || j >= n || i <= j) return;
  const double2 u = M[(size_t)j + (size_t)i * ld];
  M[(size_t)i + (size_t)j * ld] = z2(u.x, -u.y);
}

hipError_t launch_mirror(hipStream_t s, int n, double2 *m, int ld) {
  if (n <= 1) return hipSuccess;
  hipLaunchKernelGGL(mirror_kernel, dim3(cdiv(n, 16), cdiv(n, 16)), dim3(256), 0, s, n,
                     (long long)ld, m);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Grouped launchers (batched k-points)
// ---------------------------------------------------------------------------
const HoShapeInfo kHoShapeInfo[kHoShapes] = {
    {64, 64, 16}, {64, 64, 8}, {128, 64, 16}, {128, 64, 8}, {64, 128, 8}};

// The device addresses and scalars of a context's tables, as the grouped
// kernels read them (one per structure of a grouped launch).
SDesc make_sdesc(const exh_context *c) {
  SDesc S;
  std::memset(&S, 0, sizeof S);
  const int *ki = (const int *)c->katom.p;  // [Koff(natmtot), Ms(natmtot), kinfo(Ktot)]
  S.natmtot = c->natmtot;
  S.nspecies = c->nspecies;
  S.Ktot = c->Ktot;
  S.maxaa = c->maxaa;
  S.lmaxapw = c->lmaxapw;
  S.pmax = c->apwordmax;
  S.lmmax = c->lmmax;
  S.ngkmax = c->ngkmax;
  S.have_h1aa = c->have_h1aa ? 1 : 0;
  S.ldb = (c->ngkmax + 63) / 64 * 64;
  S.ldx = c->contig ? c->ngkmax : S.ldb;
  S.colA = (const long long *)c->colA.p;
  S.Koff = ki;
  S.Ms = ki ? ki + c->natmtot : nullptr;
  S.kinfo = ki ? ki + 2 * c->natmtot : nullptr;
  S.aa = (const double2 *)c->aa.p;
  S.h1aa = (const double *)c->h1aa.p;
  S.atom_species = (const int *)c->atom_species_d.p;
  S.apword = (const int *)c->apword_d.p;
  S.omax = (const int *)c->omax_d.p;
  S.piv = (const int *)c->mt_match.piv.p;
  S.rmt = (const double *)c->rmt_d.p;
  S.lu = (const double *)c->mt_match.lu.p;
  S.fpso = c->omega > 0.0 ? 12.566370614359172954 / sqrt(c->omega) : 0.0;
  S.ivg = (const int *)c->ivg.p;
  S.ngrtot = c->ngrtot;
  S.ist = (const IstEntry *)c->ist.p;
  S.lo0 = c->intgv[0];
  S.lo1 = c->intgv[1];
  S.lo2 = c->intgv[2];
  S.n0 = c->intgv[3] - c->intgv[0] + 1;
  S.n1 = c->intgv[4] - c->intgv[1] + 1;
  S.n2 = c->intgv[5] - c->intgv[2] + 1;
  S.lo = lo_tables(c);
  S.nlotot = c->nlotot;
  S.maxnlo = c->maxnlo;
  S.nlomax = c->nlomax;
  S.lorow = (const LoRow *)c->lorow.p;
  S.lobase = (const int *)c->lobase_d.p;
  S.lolo = (const double2 *)c->lolo.p;
  S.ololo_eff = (const double *)c->ololo_eff.p;
  return S;
}

hipError_t launch_ivgk_grouped(hipStream_t s, const GroupDims &g, const KDesc *kd) {
  if (g.nk <= 0 || g.ngpmax <= 0) return hipSuccess;
  hipLaunchKernelGGL(ivgk_grouped_kernel, dim3(cdiv(g.ngpmax, 256), g.nk), dim3(256), 0, s, kd);
  return hipGetLastError();
}

hipError_t launch_pack_grouped(hipStream_t s, const GroupDims &g, const KDesc *kd) {
  if (g.nk <= 0 || g.ngpmax <= 0 || !g.any_pack || g.Kmax == 0) return hipSuccess;
  hipLaunchKernelGGL(pack_grouped_kernel, dim3(cdiv(g.ngpmax, 256), g.Kmax, g.nk), dim3(256), 0,
                     s, kd);
  return hipGetLastError();
}

hipError_t launch_bt_grouped(hipStream_t s, const GroupDims &g, const KDesc *kd) {
  if (g.nk <= 0 || g.ngpmax <= 0 || g.natmax <= 0) return hipSuccess;
  dim3 grid(cdiv(g.ngpmax, TM), cdiv(g.maxaa, TM), g.natmax * g.nk);
  hipLaunchKernelGGL(bt_grouped_kernel, grid, dim3(NTH), 0, s, kd, g.natmax);
  return hipGetLastError();
}

hipError_t launch_a2t_grouped(hipStream_t s, const GroupDims &g, const KDesc *kd) {
  if (g.nk <= 0 || g.ngpmax <= 0 || g.natmax <= 0 || !g.any_a2t) return hipSuccess;
  dim3 grid(cdiv(g.ngpmax, 256), g.maxaa, g.natmax * g.nk);
  hipLaunchKernelGGL(a2t_grouped_kernel, grid, dim3(256), 0, s, kd, g.natmax);
  return hipGetLastError();
}

template <int TMI, int TMJ, int TK>
static hipError_t ho_grouped_launch(hipStream_t s, const KDesc *kd, const int4 *tiles,
                                    int ntiles, int nwg, int accumulate) {
  hipLaunchKernelGGL((ho_grouped_kernel<TMI, TMJ, TK>), dim3(nwg),
                     dim3(HoShape<TMI, TMJ, TK>::NTH), 0, s, kd, tiles, ntiles, accumulate);
  return hipGetLastError();
}

hipError_t launch_ho_grouped(hipStream_t s, int shape, const KDesc *kd, const int4 *tiles,
                             int ntiles, int nwg, int accumulate) {
  if (ntiles <= 0 || nwg <= 0) return hipSuccess;
  switch (shape) {
    case 0: return ho_grouped_launch<64, 64, 16>(s, kd, tiles, ntiles, nwg, accumulate);
    case 1: return ho_grouped_launch<64, 64, 8>(s, kd, tiles, ntiles, nwg, accumulate);
    case 2: return ho_grouped_launch<128, 64, 16>(s, kd, tiles, ntiles, nwg, accumulate);
    case 3: return ho_grouped_launch<128, 64, 8>(s, kd, tiles, ntiles, nwg, accumulate);
    case 4: return ho_grouped_launch<64, 128, 8>(s, kd, tiles, ntiles, nwg, accumulate);
    default: return hipErrorInvalidValue;
  }
}

// Workgroups of one shape that fit on a CU at once (LDS, registers, waves).
int ho_grouped_occupancy(int shape) {
  int n = 0;
  hipError_t e = hipErrorInvalidValue;
  switch (shape) {
#define EXH_OCC(I, J, K)                                                                    \
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, ho_grouped_kernel<I, J, K>,           \
                                                   HoShape<I, J, K>::NTH, 0);               \
  break;
    case 0: EXH_OCC(64, 64, 16)
    case 1: EXH_OCC(64, 64, 8)
    case 2: EXH_OCC(128, 64, 16)
    case 3: EXH_OCC(128, 64, 8)
    case 4: EXH_OCC(64, 128, 8)
#undef EXH_OCC
    default: break;
  }
  return e == hipSuccess && n > 0 ? n : 1;
}

hipError_t launch_lo_border_grouped(hipStream_t s, const GroupDims &g, const KDesc *kd,
                                    int accumulate) {
  if (g.nk <= 0 || g.ngpmax <= 0 || g.nlotot == 0 || g.natmax <= 0) return hipSuccess;
  dim3 grid(cdiv(g.ngpmax, 64), g.natmax, g.nk);
  if (g.maxnlo <= 16)
    hipLaunchKernelGGL(lo_border_grouped_kernel<16>, grid, dim3(64), 0, s, kd, accumulate);
  else if (g.maxnlo <= 32)
    hipLaunchKernelGGL(lo_border_grouped_kernel<32>, grid, dim3(64), 0, s, kd, accumulate);
  else
    hipLaunchKernelGGL(lo_border_grouped_kernel<64>, grid, dim3(64), 0, s, kd, accumulate);
  return hipGetLastError();
}

hipError_t launch_lolo_grouped(hipStream_t s, const GroupDims &g, const KDesc *kd,
                               int accumulate) {
  if (g.nk <= 0 || g.nlotot == 0) return hipSuccess;
  dim3 grid(cdiv(g.nlotot, 16), cdiv(g.nlotot, 16), g.nk);
  hipLaunchKernelGGL(lolo_grouped_kernel, grid, dim3(256), 0, s, kd, accumulate);
  return hipGetLastError();
}

}  // namespace exh
