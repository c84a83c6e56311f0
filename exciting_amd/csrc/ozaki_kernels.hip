// Split-precision muffin-tin contraction: Ozaki scheme II on int8 MFMA.
//
// The FP64 products H_MT = X^H B and O_MT = X^H X (X = A^T, B = (aa A)^T,
// K = every atom's (l, m, io) stacked) are emulated with integer arithmetic:
//
//  1. Each G column of X and of B gets a power-of-two scale 2^s so that its
//     2-norm is at most 2^(c-1); the scaled values are rounded to integers,
//     which keeps the norm below 2^c.  This rounding is the only
//     approximation: the error is <= 1/2 of 2^-s per component, i.e. about
//     2^-c of the column's norm.
//  2. The integers are reduced modulo N primes p_m <= 251 into int8 residues
//     in [-(p-1)/2, (p-1)/2], stored in 16-byte k blocks per G row: the
//     operand format of v_mfma_i32_32x32x32_i8.
//  3. For each modulus and each of six real products (Karatsuba for the
//     conjugated complex product, exact in modular arithmetic):
//        H: P1 = xr.br, P2 = xi.bi, P3 = (xr+xi).(br-bi)
//        O: Q1 = xr.xr, Q2 = xi.xi, Q3 = (xr+xi).(xr-xi)
//     an upper-triangular int8 GEMM accumulates exactly in int32 and stores
//     the result modulo p_m as one byte per element (operands streamed
//     global -> LDS by direct DMA loads).
//  4. Re = P1+P2 and Im = P1-P2-P3 are formed modulo each p_m and turned
//     back into exact integers by Garner's mixed-radix algorithm with
//     symmetric digits (no cancellation for small results), evaluated in
//     FP64 and rescaled by 2^-(s_i+s_j).  The interstitial terms and the
//     conj mirror are added as in the FP64 path.
//
// The reconstructed integers are Re and Im of x_i^H b_j (the Karatsuba
// products only meet modulo p), bounded by ||x_i|| ||b_j|| <= 2^(2c)
// (Cauchy-Schwarz); c is the largest value with 2^(2c) < prod(p)/2, so every
// element is recovered exactly from its residues.  Against scaling by the
// largest entry and the worst case 6 K 2^(2b) this gains
// (log2(6K) + 2 log2(max/norm ratio of the columns)) / 2 bits per operand.
#include <algorithm>
#include <climits>
#include <cmath>
#include <mutex>

#include "exh_device.h"

namespace exh {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int OZ_T = 128;   // output tile edge
constexpr int OZ_KS = 64;   // k per LDS stage (4 blocks of 16)

__constant__ int c_p[kOzMaxMod];
__constant__ float c_rp[kOzMaxMod];          // 1/p
__constant__ int c_w[kOzMaxMod];             // (p_0..p_{k-1})^{-1} mod p_k
__constant__ int c_c[kOzMaxMod][kOzMaxMod];  // (p_0..p_{j-1}) * w_k mod p_k, j < k

// ---------------------------------------------------------------------------
// Column scales from the 2-norms: part[chunk][g] = sum over the chunk's 256
// k of |x|^2, then per column the chunks in order (a fixed summation order,
// so the scale is the same in every run) and emax[g] = E + 1 with
// ||x_g|| < 2^E: the scale 2^(c - emax) takes the norm below 2^(c-1).
// ---------------------------------------------------------------------------
constexpr int OZ_NCH = 256;  // k per partial sum

__global__ void colnorm_part_kernel(int ngp, int Ktot, const double2 *__restrict__ X,
                                    long long ldx, double *__restrict__ part, int ldp) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngp) return;
  const int k0 = blockIdx.y * OZ_NCH, k1 = min(Ktot, k0 + OZ_NCH);
  double s = 0.0;
  for (int k = k0; k < k1; k++) {
    const double2 v = X[(size_t)k * ldx + g];
    s = fma(v.x, v.x, s);
    s = fma(v.y, v.y, s);
  }
  part[(size_t)blockIdx.y * ldp + g] = s;
}

__global__ void colnorm_final_kernel(int ngp, int nchunk, const double *__restrict__ part,
                                     int ldp, int *__restrict__ emax) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngp) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; c++) s += part[(size_t)c * ldp + g];
  int e = INT_MIN;
  if (s > 0.0) {
    (void)frexp(sqrt(s), &e);  // norm = f 2^e, f in [0.5, 1): norm < 2^e
    e += 1;
  }
  emax[g] = e;
}

__device__ __forceinline__ int smod(double x, int m) {
  // symmetric residue of the integer-valued double x modulo the odd prime p_m
  const double p = (double)c_p[m];
  return (int)(x - p * rint(x * (1.0 / p)));
}

// ---------------------------------------------------------------------------
// Residues: one thread per (16-k block, G row).  X gives four arrays
// (xr, xi, xr+xi, xr-xi), B gives three (br, bi, br-bi), per modulus.
// Rows >= ngp and k >= Ktot are written as zeros (no masking downstream).
// Layout: Q[(m * narr + t) * qstride + (kblk * Gpad + g) * 16 + byte].
// ---------------------------------------------------------------------------
template <int NARR>
__global__ void residue_kernel(int ngp, int Ktot, int Gpad, const double2 *__restrict__ X,
                               long long ldx, const int *__restrict__ emax, int bits, int nmod,
                               size_t qstride, int8_t *__restrict__ Q) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int kb = blockIdx.y;
  if (g >= Gpad) return;
  double re[16], im[16];
  int s = 0;
  if (g < ngp) {
    const int e = emax[g];
    s = (e == INT_MIN) ? 0 : bits - e;
  }
#pragma unroll
  for (int t = 0; t < 16; t++) {
    const int k = kb * 16 + t;
    double2 v = z2(0.0, 0.0);
    if (g < ngp && k < Ktot) v = X[(size_t)k * ldx + g];
    re[t] = rint(ldexp(v.x, s));
    im[t] = rint(ldexp(v.y, s));
  }
  const size_t off = ((size_t)kb * Gpad + g) * 16;
  for (int m = 0; m < nmod; m++) {
#pragma unroll
    for (int a = 0; a < NARR; a++) {
      int w[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        unsigned packed = 0;
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const int k = 4 * q + t;
          double x;
          if (a == 0) x = re[k];
          else if (a == 1) x = im[k];
          else if (a == 2) x = (NARR == 4) ? re[k] + im[k] : re[k] - im[k];
          else x = re[k] - im[k];
          packed |= (unsigned)(smod(x, m) & 0xff) << (8 * t);
        }
        w[q] = (int)packed;
      }
      *(v4i *)(Q + (size_t)(m * NARR + a) * qstride + off) = v4i{w[0], w[1], w[2], w[3]};
    }
  }
}

// ---------------------------------------------------------------------------
// Upper-triangular int8 GEMMs.  One workgroup computes the product pair
// (H_q, O_q) of one modulus (q = 0, 1, 2) for a row panel of the row operand
// (xr, xi or xr+xi) against one 128-column block J of the H and O column
// operands:
//   C_H[i][j] = sum_k A[k][i] BH[k][j],  C_O[i][j] = sum_k A[k][i] BO[k][j]
// (int32, exact), each stored mod p as bytes in the 128x128 tiles on or
// above the diagonal.  The H and O products share the row operand, as the
// FP64 kernel shares conj(A) between H and O.  Each wave owns 96 rows x 64
// columns of both products (3 x 2 blocks of v_mfma_i32_32x32x32_i8 each, 192
// accumulator registers) and reads 7 operand fragments from LDS per 12
// MFMAs; NW waves form NW/2 row groups x 2 column groups, so the panel is
// 48*NW rows.  NW = 8: 384-row panels, one workgroup per CU.  (NW = 4, two
// 192-row workgroups per CU whose barriers fall at different times, gives
// the same products and measured 1 % slower at the 64-atom shape.)
// Operands stream global -> LDS directly (global_load_lds, 16 bytes per
// lane, one 64-row chunk of one k block per instruction), two stages of 64 k
// double-buffered.
// Operand lane layout: row/column = lane%32, 16 k bytes from k block
// 2*ks + lane/32; result D[8*(r/4) + 4*(lane/32) + r%4][lane%32].
// ---------------------------------------------------------------------------
constexpr int OZ_WR = 3;           // 32-row blocks per wave (96 rows)
constexpr int OZ_NWG = 8;          // waves per workgroup
constexpr int OZ_R = 48 * OZ_NWG;  // rows of the row panel per workgroup

#define EXH_LDS(p) ((__attribute__((address_space(3))) void *)(p))
#define EXH_GLB(p) ((const __attribute__((address_space(1))) void *)(p))

// Residues mod p of one 32 x 64 strip (block row a of a wave) per product,
// four consecutive rows packed per 32-bit store, into the 128x128 tile it
// lies in when that tile is on or above the diagonal.
__device__ __forceinline__ void store_residues(const v16i (&acc)[OZ_WR][2], int p, float rp,
                                               int rg, int wc, int lh, int l32, int IT, int J,
                                               uint8_t *Rp) {
#pragma unroll
  for (int a = 0; a < OZ_WR; a++) {
    const int row = IT * OZ_R + rg * 32 * OZ_WR + a * 32;
    const int I = row >> 7;
    if (I > J) continue;  // below the diagonal (uniform per wave)
    uint8_t *out = Rp + ((long long)J * (J + 1) / 2 + I) * (OZ_T * OZ_T);
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int h = 0; h < 4; h++) {
        unsigned packed = 0;
#pragma unroll
        for (int t4 = 0; t4 < 4; t4++) {
          const int x = acc[a][b][4 * h + t4];
          int q = (int)floorf((float)x * rp);
          int r = x - q * p;
          r += (r < 0) ? p : 0;
          r -= (r >= p) ? p : 0;
          packed |= (unsigned)r << (8 * t4);
        }
        const int il = (row & 127) + 8 * h + 4 * lh;
        const int jl = wc * 64 + b * 32 + l32;
        *(unsigned *)(out + il + OZ_T * jl) = packed;
      }
  }
}

// Row panels column block J needs: those starting at or above its last row.
__host__ __device__ inline int oz_npan(int j) { return (j * OZ_T + OZ_T - 1) / OZ_R + 1; }

// Panels per GEMM.
__host__ __device__ inline long long oz_panels(int nt) {
  long long n = 0;
  for (int j = 0; j < nt; j++) n += oz_npan(j);
  return n;
}

// Panel order: super-columns of OZ_SJ column blocks, each walked in chunks
// of OZ_SI row panels (4 x 384 rows), so that the workgroups an XCD runs at
// once share a few row panels and column panels in its L2 (the column-major
// order had 32 co-resident workgroups share about 1 column panel and 19 row
// panels); 3.6 % faster at the 64-atom shape, bitwise the same products.
constexpr int OZ_SJ = 8, OZ_SI = 4 * 384 / OZ_R;

// Panel o of one GEMM -> (column block J, row panel IT), IT < oz_npan(J).
__device__ __forceinline__ void oz_panel_at(int o, int nt, int &J, int &IT) {
  int rem = o, j0 = 0;
  for (;; j0 += OZ_SJ) {
    const int j1 = min(nt, j0 + OZ_SJ);
    int cnt = 0;
    for (int j = j0; j < j1; j++) cnt += oz_npan(j);
    if (rem < cnt) break;
    rem -= cnt;
  }
  const int j1 = min(nt, j0 + OZ_SJ);
  int i0 = 0;
  for (;; i0 += OZ_SI) {
    int cnt = 0;
    for (int j = j0; j < j1; j++) cnt += max(0, min(OZ_SI, oz_npan(j) - i0));
    if (rem < cnt) break;
    rem -= cnt;
  }
  J = j0;
  IT = i0;
  for (int j = j0; j < j1; j++) {
    const int cj = max(0, min(OZ_SI, oz_npan(j) - i0));
    if (rem < cj) {
      J = j;
      IT = i0 + rem;
      return;
    }
    rem -= cj;
  }
}

// two waves per SIMD whatever the workgroup size: at most 256 VGPRs
__global__ __launch_bounds__(64 * OZ_NWG) __attribute__((amdgpu_waves_per_eu(2, 2))) void
ozgemm_kernel(
    int nt, int nslab, int Gpad, int nmod, const int8_t *__restrict__ QX,
    const int8_t *__restrict__ QB, size_t qstride, uint8_t *__restrict__ R) {
  // XCD-contiguous remap over the whole grid, then (pair*nmod+m, panel)
  const int bid = blockIdx.x, nb = gridDim.x;
  const int xcd = bid & 7, i8 = bid >> 3, q8 = nb >> 3, r8 = nb & 7;
  const int t = xcd * q8 + (xcd < r8 ? xcd : r8) + i8;
  const int per = (int)oz_panels(nt);
  const int qm = t / per, o = t - qm * per;
  const int q = qm / nmod, m = qm - q * nmod;
  int J, IT;
  oz_panel_at(o, nt, J, IT);
  const long long npairs = (long long)nt * (nt + 1) / 2;
  const int I0 = IT * OZ_R, J0 = J * OZ_T;

  const int8_t *Ab = QX + (size_t)(m * 4 + q) * qstride;
  const int8_t *Hb = QB + (size_t)(m * 3 + q) * qstride;
  const int8_t *Ob = QX + (size_t)(m * 4 + (q == 2 ? 3 : q)) * qstride;

  // one LDS array per stage and operand: the compiler then sees the DMA into
  // one stage and the reads of the other as disjoint (no wait inserted)
  __shared__ v4i sA0[4][OZ_R], sH0[4][OZ_T], sO0[4][OZ_T];
  __shared__ v4i sA1[4][OZ_R], sH1[4][OZ_T], sO1[4][OZ_T];
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int rg = w >> 1, wc = w & 1, l32 = lane & 31, lh = lane >> 5;
  v16i ah[OZ_WR][2], ao[OZ_WR][2];
#pragma unroll
  for (int a = 0; a < OZ_WR; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      ah[a][b] = v16i{0};
      ao[a][b] = v16i{0};
    }

  // Chunks of 64 rows x 16 bytes per stage: 4 k blocks x OZ_R/64 row chunks
  // of A (3 per wave), 4 k blocks x 2 of H and of O (8 / OZ_NWG each per
  // wave); each lane moves its 16 bytes to chunk base + 16 lane
  constexpr int RC = OZ_R / 64, NBC = 8 / OZ_NWG;
  static_assert(4 * RC == OZ_WR * OZ_NWG && NBC * OZ_NWG == 8, "chunking");
  const int voff = lane * 16;
  auto issue = [&](int s, v4i (*dA)[OZ_R], v4i (*dH)[OZ_T], v4i (*dO)[OZ_T]) {
#pragma unroll
    for (int u = 0; u < OZ_WR; u++) {
      const int ci = w * OZ_WR + u;
      const int kk = ci / RC, rc = ci - RC * kk;
      const int8_t *src = Ab + (((size_t)s * 4 + kk) * Gpad + I0 + rc * 64) * 16;
      __builtin_amdgcn_global_load_lds(EXH_GLB(src + voff), EXH_LDS(&dA[kk][rc * 64]), 16, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < NBC; u++) {
      const int ci = w * NBC + u;
      const int kk = ci >> 1, rc = ci & 1;
      const size_t boff = (((size_t)s * 4 + kk) * Gpad + J0 + rc * 64) * 16;
      __builtin_amdgcn_global_load_lds(EXH_GLB(Hb + boff + voff), EXH_LDS(&dH[kk][rc * 64]), 16, 0, 0);
      __builtin_amdgcn_global_load_lds(EXH_GLB(Ob + boff + voff), EXH_LDS(&dO[kk][rc * 64]), 16, 0, 0);
    }
  };
  auto compute = [&](v4i (*sA)[OZ_R], v4i (*sH)[OZ_T], v4i (*sO)[OZ_T]) {
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
      const int kk = 2 * ks + lh;
      v4i av[OZ_WR], hv[2], ov[2];
#pragma unroll
      for (int a = 0; a < OZ_WR; a++) av[a] = sA[kk][rg * 32 * OZ_WR + a * 32 + l32];
#pragma unroll
      for (int b = 0; b < 2; b++) {
        hv[b] = sH[kk][wc * 64 + b * 32 + l32];
        ov[b] = sO[kk][wc * 64 + b * 32 + l32];
      }
#pragma unroll
      for (int a = 0; a < OZ_WR; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          ah[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av[a], hv[b], ah[a][b], 0, 0, 0);
          ao[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av[a], ov[b], ao[a][b], 0, 0, 0);
        }
    }
  };
  // vmcnt(0): every DMA of this wave has landed in LDS; the barrier then
  // makes all waves' chunks visible and frees the stage read last
  issue(0, sA0, sH0, sO0);
  __builtin_amdgcn_s_waitcnt(0x0F70);
  __syncthreads();
  for (int s = 0; s < nslab; s += 2) {
    if (s + 1 < nslab) issue(s + 1, sA1, sH1, sO1);
    compute(sA0, sH0, sO0);
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __syncthreads();
    if (s + 1 < nslab) {
      if (s + 2 < nslab) issue(s + 2, sA0, sH0, sO0);
      compute(sA1, sH1, sO1);
      __builtin_amdgcn_s_waitcnt(0x0F70);
      __syncthreads();
    }
  }

  const int p = c_p[m];
  const float rp = c_rp[m];
  store_residues(ah, p, rp, rg, wc, lh, l32, IT, J,
                 R + (size_t)(q * nmod + m) * npairs * (OZ_T * OZ_T));
  store_residues(ao, p, rp, rg, wc, lh, l32, IT, J,
                 R + (size_t)((3 + q) * nmod + m) * npairs * (OZ_T * OZ_T));
}

// Garner with symmetric digits: the integer x in (-M/2, M/2) with residues r_k.
template <int NM>
__device__ __forceinline__ double garner(const int *r) {
  int v[NM];
  double x = 0.0;
#pragma unroll
  for (int k = 0; k < NM; k++) {
    const int p = c_p[k];
    int acc = r[k] * c_w[k];
#pragma unroll
    for (int j = 0; j < k; j++) acc -= v[j] * c_c[j][k];
    int q = (int)floorf((float)acc * c_rp[k]);
    int d = acc - q * p;
    d += (d < 0) ? p : 0;
    d -= (d >= p) ? p : 0;
    if (d > (p >> 1)) d -= p;
    v[k] = d;
  }
#pragma unroll
  for (int k = NM - 1; k >= 0; k--) x = fma(x, (double)c_p[k], (double)v[k]);
  return x;
}


__device__ __forceinline__ int modp(int x, int k) {
  const int p = c_p[k];
  int r = x % p;
  return r < 0 ? r + p : r;
}

// ---------------------------------------------------------------------------
// Residues -> FP64 H and O upper elements (+ interstitial, + mirror).
// One workgroup per 128x128 tile, walked in 32x32 sub-blocks: each thread
// reconstructs four elements of a sub-block and writes them to the upper
// triangle (coalesced along i); the sub-block is staged in LDS and written
// back transposed and conjugated to the lower triangle (coalesced along j).
// ---------------------------------------------------------------------------
template <int NM>
__global__ __launch_bounds__(256) void ozcombine_kernel(
    int ngp, int npairs, const uint8_t *__restrict__ R, const int *__restrict__ ex,
    const int *__restrict__ eb, int bits, IstBox box, double2 *__restrict__ H,
    double2 *__restrict__ O, long long ldh, int accumulate) {
  const int tile = blockIdx.x;
  int J = (int)((sqrt(8.0 * (double)tile + 1.0) - 1.0) * 0.5);
  while ((J + 1) * (J + 2) / 2 <= tile) J++;
  while (J * (J + 1) / 2 > tile) J--;
  const int I = tile - J * (J + 1) / 2;
  const int I0 = I * OZ_T, J0 = J * OZ_T;
  const size_t TT = (size_t)OZ_T * OZ_T;
  __shared__ double2 sH[32][33], sO[32][33];
  const int tid = threadIdx.x, lo = tid & 31, hi8 = tid >> 5;
  for (int sb = 0; sb < 16; sb++) {
    const int bi = sb & 3, bj = sb >> 2;
    if (I == J && bi > bj) continue;  // wholly below the diagonal (uniform)
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
      const int il = bi * 32 + lo, jl = bj * 32 + hi8 + 8 * r;
      const int i = I0 + il, j = J0 + jl;
      if (i >= ngp || j >= ngp || i > j) continue;
      const int e = il + OZ_T * jl;
      int hre[NM], him[NM], ore[NM], oim[NM];
#pragma unroll
      for (int k = 0; k < NM; k++) {
        int rr[6];
#pragma unroll
        for (int pr = 0; pr < 6; pr++)
          rr[pr] = R[((size_t)(pr * NM + k) * npairs + tile) * TT + e];
        hre[k] = modp(rr[0] + rr[1], k);
        him[k] = modp(rr[0] - rr[1] - rr[2], k);
        ore[k] = modp(rr[3] + rr[4], k);
        oim[k] = modp(rr[3] - rr[4] - rr[5], k);
      }
      const int si = ex[i] == INT_MIN ? 0 : bits - ex[i];
      const int sjb = eb[j] == INT_MIN ? 0 : bits - eb[j];
      const int sjx = ex[j] == INT_MIN ? 0 : bits - ex[j];
      double hr = ldexp(garner<NM>(hre), -(si + sjb));
      double hv = ldexp(garner<NM>(him), -(si + sjb));
      double orr = ldexp(garner<NM>(ore), -(si + sjx));
      double oi = ldexp(garner<NM>(oim), -(si + sjx));
      const size_t up = (size_t)i + (size_t)j * ldh;
      if (accumulate) {
        const double2 h0 = H[up], o0 = O[up];
        hr += h0.x;
        hv += h0.y;
        orr += o0.x;
        oi += o0.y;
      }
      ist_add(box, i, j, hr, hv, orr, oi);
      H[up] = z2(hr, hv);
      O[up] = z2(orr, oi);
      sH[jl - bj * 32][lo] = z2(hr, -hv);
      sO[jl - bj * 32][lo] = z2(orr, -oi);
    }
    __syncthreads();
    // lower triangle: element (j, i) = conj(upper (i, j)), i < j
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
      const int jl = bj * 32 + lo, il = bi * 32 + hi8 + 8 * r;
      const int i = I0 + il, j = J0 + jl;
      if (i >= ngp || j >= ngp || i >= j) continue;
      const size_t lw = (size_t)j + (size_t)i * ldh;
      H[lw] = sH[lo][hi8 + 8 * r];
      O[lw] = sO[lo][hi8 + 8 * r];
    }
    __syncthreads();
  }
}


inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

const int kOzPrimes[kOzMaxMod] = {251, 241, 239, 233, 229, 227, 223, 211, 199, 197, 193, 191};

int oz_bits(int nmod, int Ktot) {
  (void)Ktot;
  double lm = 0.0;
  for (int k = 0; k < nmod; k++) lm += std::log2((double)kOzPrimes[k]);
  // |Re|, |Im| of x^H b <= ||x|| ||b|| <= 2^(2c) must stay below prod(p)/2
  // (a hair below the exact bound guards the floating-point log2); the
  // scaled entries stay well inside FP64's exact integers
  const double c = std::floor((lm - 1.0) * 0.5 - 1e-9);
  return (int)std::min(50.0, c);
}

// The tables cover all kOzMaxMod moduli whatever nmod a call uses, so they are
// written once per device and process; the mutex keeps contexts on other host
// threads from writing them while this one does.
hipError_t oz_upload_constants(int nmod) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static bool done[kMaxDevices] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> guard(mu);
  if (dev < kMaxDevices && done[dev]) return hipSuccess;
  int p[kOzMaxMod], w[kOzMaxMod], cc[kOzMaxMod][kOzMaxMod];
  float rp[kOzMaxMod];
  auto inv = [](long long a, int m) {  // a^-1 mod m, m prime
    long long r = 1, b = ((a % m) + m) % m, e = m - 2;
    while (e) {
      if (e & 1) r = r * b % m;
      b = b * b % m;
      e >>= 1;
    }
    return (int)r;
  };
  for (int k = 0; k < kOzMaxMod; k++) {
    p[k] = kOzPrimes[k];
    rp[k] = 1.0f / (float)p[k];
    long long P = 1;  // p_0 ... p_{k-1} mod p_k
    for (int j = 0; j < k; j++) P = P * kOzPrimes[j] % kOzPrimes[k];
    w[k] = k == 0 ? 1 : inv(P, kOzPrimes[k]);
    long long Pj = 1;  // p_0 ... p_{j-1} mod p_k
    for (int j = 0; j < kOzMaxMod; j++) {
      cc[j][k] = (j < k) ? (int)(Pj * w[k] % kOzPrimes[k]) : 0;
      if (j < k) Pj = Pj * kOzPrimes[j] % kOzPrimes[k];
    }
  }
  (void)nmod;
  e = hipMemcpyToSymbol(HIP_SYMBOL(c_p), p, sizeof p);
  if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_rp), rp, sizeof rp);
  if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_w), w, sizeof w);
  if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_c), cc, sizeof cc);
  if (e == hipSuccess && dev < kMaxDevices) done[dev] = true;
  return e;
}

OzLayout oz_layout(int ngkmax, int Ktot, int nmod) {
  OzLayout L;
  // rows padded to 384: a multiple of the row panel (384 or 192 rows) and of
  // the 128-column blocks, so every panel and column block loads inside it
  constexpr int GA = 384;
  static_assert(GA % OZ_R == 0 && GA % OZ_T == 0, "row padding");
  L.Gpad = (ngkmax + GA - 1) / GA * GA;
  L.Kpad = (Ktot + OZ_KS - 1) / OZ_KS * OZ_KS;
  L.qstride = (size_t)L.Kpad * L.Gpad;
  const long long nt = (ngkmax + OZ_T - 1) / OZ_T;  // tile columns of the largest basis
  L.max_pairs = nt * (nt + 1) / 2;
  L.qx_bytes = (size_t)nmod * 4 * L.qstride;
  L.qb_bytes = (size_t)nmod * 3 * L.qstride;
  L.r_bytes = (size_t)6 * nmod * L.max_pairs * OZ_T * OZ_T;
  // per operand: the column exponents, then the partial norms
  L.e_part_off = ((size_t)std::max(ngkmax, 1) * sizeof(int) + 255) / 256 * 256;
  L.e_bytes = L.e_part_off +
              sizeof(double) * (size_t)((Ktot + OZ_NCH - 1) / OZ_NCH) * std::max(ngkmax, 1);
  return L;
}

hipError_t launch_ozaki(hipStream_t s, const exh_context *c, int ngp, const double2 *X,
                        long long ldx, const double2 *B, long long ldb, int nmod, int bits,
                        const OzLayout &L, int8_t *QX, int8_t *QB, uint8_t *R, int *ex, int *eb,
                        const IstBox &box, double2 *h, double2 *o, int ldh, int accumulate,
                        int stage_mask) {
  if (ngp <= 0) return hipSuccess;
  const int K = c->Ktot;
  const long long nt = (ngp + OZ_T - 1) / OZ_T;
  const long long npairs = nt * (nt + 1) / 2;
  const int nslab = L.Kpad / OZ_KS;
  if (stage_mask & 1) {
    const int nch = (K + OZ_NCH - 1) / OZ_NCH;
    const int ldp = std::max(c->ngkmax, 1);
    double *px = (double *)((char *)ex + L.e_part_off);
    double *pb = (double *)((char *)eb + L.e_part_off);
    dim3 ge(cdiv(ngp, 256), nch);
    hipLaunchKernelGGL(colnorm_part_kernel, ge, dim3(256), 0, s, ngp, K, X, ldx, px, ldp);
    hipLaunchKernelGGL(colnorm_part_kernel, ge, dim3(256), 0, s, ngp, K, B, ldb, pb, ldp);
    hipLaunchKernelGGL(colnorm_final_kernel, dim3(cdiv(ngp, 256)), dim3(256), 0, s, ngp, nch,
                       px, ldp, ex);
    hipLaunchKernelGGL(colnorm_final_kernel, dim3(cdiv(ngp, 256)), dim3(256), 0, s, ngp, nch,
                       pb, ldp, eb);
    dim3 gr(cdiv(L.Gpad, 256), L.Kpad / 16);
    hipLaunchKernelGGL(residue_kernel<4>, gr, dim3(256), 0, s, ngp, K, L.Gpad, X, ldx, ex, bits,
                       nmod, L.qstride, QX);
    hipLaunchKernelGGL(residue_kernel<3>, gr, dim3(256), 0, s, ngp, K, L.Gpad, B, ldb, eb, bits,
                       nmod, L.qstride, QB);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (stage_mask & 2) {
    const long long per = oz_panels((int)nt);
    hipLaunchKernelGGL(ozgemm_kernel, dim3((unsigned)(per * 3 * nmod)), dim3(64 * OZ_NWG), 0, s,
                       (int)nt,
                       nslab, L.Gpad, nmod, QX, QB, L.qstride, R);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (stage_mask & 4) {
    switch (nmod) {
#define EXH_OZC(N)                                                                          \
  case N:                                                                                   \
    hipLaunchKernelGGL(ozcombine_kernel<N>, dim3((unsigned)npairs), dim3(256), 0, s, ngp,     \
                       (int)npairs, R, ex, eb, bits, box, h, o, (long long)ldh, accumulate); \
    break;
      EXH_OZC(6) EXH_OZC(7) EXH_OZC(8) EXH_OZC(9) EXH_OZC(10) EXH_OZC(11) EXH_OZC(12)
#undef EXH_OZC
      default:
        return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}

}  // namespace exh
