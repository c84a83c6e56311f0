// Dense products of the generalized eigensolve (the step after the setup that
// solvewithlapack hands to ZHEGVX, modfvsystem.f90:585-720): the Chebyshev
// filter of the subspace iteration applies the reduced Hermitian matrix C
// (n x n, n ~ 10^4) to a block of a few dozen vectors, about 130 times per
// solve.  At that shape the vendor GEMM reads C at 0.7-1.5 TB/s whatever
// the block width; this kernel splits K over many workgroups instead, so
// that the whole chip streams C once per product on FP64 MFMA.
#include "exh_device.h"

namespace exh {
namespace {

typedef double dx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ dx4 mfma16(double a, double b, dx4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int CT = 64;    // output tile edge (rows i and columns j)
constexpr int CK = 8;     // K slab staged in LDS per step (16: 2 %, 32: 6 % slower)
constexpr int CNT = 256;  // 4 waves, each a 32 x 32 quadrant of 2 x 2 MFMA blocks

// part[s][i][j] = sum over k in chunk s of conj(A[k][i]) X[k][j] (asign = -1)
// or A[k][i] X[k][j] (asign = 1), for a 64 x 64 tile (i, j) of the output;
// chunk s is [s * kchunk, min(K, (s+1) * kchunk)).
// A(k, i) at A[k * lda + i], X(k, j) at X[k * ldx + j], part[s] row-major with
// leading dimension m.  Loads are clamped to valid addresses and zero-weighted
// outside the tile, so every thread runs the same slab loop.
__global__ __launch_bounds__(CNT) void cx_part_kernel(int nout, int m, int K, int kchunk,
                                                      const double2 *__restrict__ A,
                                                      long long lda, double asign,
                                                      const double2 *__restrict__ X,
                                                      long long ldx,
                                                      double2 *__restrict__ part) {
  const int I0 = blockIdx.x * CT, J0 = blockIdx.y * CT, s = blockIdx.z;
  const int k0 = s * kchunk, k1 = min(K, k0 + kchunk);
  __shared__ double sAr[CK][CT + 16], sAi[CK][CT + 16], sXr[CK][CT + 16], sXi[CK][CT + 16];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = w >> 1, wc = w & 1, lr = lane & 15, lk = lane >> 4;
  const int ii = tid & 63, kq = tid >> 6;  // loads: tile column ii, slab rows kq + 4q
  const bool vi = I0 + ii < nout, vj = J0 + ii < m;
  const double2 *ap = A + (vi ? I0 + ii : nout - 1);
  const double2 *xp = X + (vj ? J0 + ii : m - 1);

  dx4 cre[2][2], cim[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      cre[a][b] = dx4{0, 0, 0, 0};
      cim[a][b] = dx4{0, 0, 0, 0};
    }

  constexpr int QB = CK / 4;
  double2 pa[QB], px[QB];
  double fa[QB], fx[QB];
  auto load = [&](int kb) {
#pragma unroll
    for (int q = 0; q < QB; q++) {
      const int k = kb + kq + 4 * q;
      const bool kv = k < k1;
      const int kc = kv ? k : k1 - 1;
      pa[q] = ap[(size_t)kc * lda];
      px[q] = xp[(size_t)kc * ldx];
      fa[q] = (kv && vi) ? 1.0 : 0.0;
      fx[q] = (kv && vj) ? 1.0 : 0.0;
    }
  };
  load(k0);
  for (int kb = k0; kb < k1; kb += CK) {
    // conj(A): the imaginary part is stored negated (asign = -1)
#pragma unroll
    for (int q = 0; q < QB; q++) {
      const int kk = kq + 4 * q;
      sAr[kk][ii] = fa[q] * pa[q].x;
      sAi[kk][ii] = asign * (fa[q] * pa[q].y);
      sXr[kk][ii] = fx[q] * px[q].x;
      sXi[kk][ii] = fx[q] * px[q].y;
    }
    __syncthreads();
    if (kb + CK < k1) load(kb + CK);
#pragma unroll
    for (int k4 = 0; k4 < CK / 4; k4++) {
      const int kk = 4 * k4 + lk;
      double ar[2], ai[2], nai[2], xr[2], xi[2];
#pragma unroll
      for (int a = 0; a < 2; a++) {
        ar[a] = sAr[kk][wr * 32 + a * 16 + lr];
        ai[a] = sAi[kk][wr * 32 + a * 16 + lr];
        nai[a] = -ai[a];
      }
#pragma unroll
      for (int b = 0; b < 2; b++) {
        xr[b] = sXr[kk][wc * 32 + b * 16 + lr];
        xi[b] = sXi[kk][wc * 32 + b * 16 + lr];
      }
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          cre[a][b] = mfma16(ar[a], xr[b], cre[a][b]);
          cre[a][b] = mfma16(nai[a], xi[b], cre[a][b]);
          cim[a][b] = mfma16(ar[a], xi[b], cim[a][b]);
          cim[a][b] = mfma16(ai[a], xr[b], cim[a][b]);
        }
    }
    __syncthreads();
  }
  double2 *ps = part + (size_t)s * nout * m;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = I0 + wr * 32 + a * 16 + lk + 4 * r;
        const int j = J0 + wc * 32 + b * 16 + lr;
        if (i < nout && j < m) ps[(size_t)i * m + j] = z2(cre[a][b][r], cim[a][b][r]);
      }
}

// One column (m = 1: the matrix-vector products of the Lanczos bound): the
// same partial sums without MFMA, which would spend 63 of 64 columns on
// padding.  One thread per output row i, its reads of A coalesced along i;
// each chunk's x is staged in LDS 256 entries at a time.
constexpr int VT = 256;
__global__ __launch_bounds__(VT) void cx_part_vec_kernel(int nout, int K, int kchunk,
                                                         const double2 *__restrict__ A,
                                                         long long lda, double asign,
                                                         const double2 *__restrict__ X,
                                                         long long ldx,
                                                         double2 *__restrict__ part) {
  const int i = blockIdx.x * VT + threadIdx.x, s = blockIdx.y;
  const int k0 = s * kchunk, k1 = min(K, k0 + kchunk);
  __shared__ double2 sx[VT];
  const double2 *ap = A + min(i, nout - 1);
  double re = 0.0, im = 0.0;
  for (int kb = k0; kb < k1; kb += VT) {
    const int kl = min(VT, k1 - kb);
    __syncthreads();  // the previous block of x is no longer read
    if ((int)threadIdx.x < kl) sx[threadIdx.x] = X[(size_t)(kb + threadIdx.x) * ldx];
    __syncthreads();
#pragma unroll 8
    for (int t = 0; t < kl; t++) {
      const double2 a = ap[(size_t)(kb + t) * lda];
      const double2 x = sx[t];
      const double ai = asign * a.y;  // conj(A): the imaginary part negated
      re = fma(a.x, x.x, re);
      re = fma(-ai, x.y, re);
      im = fma(a.x, x.y, im);
      im = fma(ai, x.x, im);
    }
  }
  if (i < nout) part[(size_t)s * nout + i] = z2(re, im);
}

// Y[i][j] = alpha * sum_s part[s][i][j] (s in order) + beta * Y0[i][j] +
// gamma * Z[i][j]; Y may be Y0 (each element is read before it is written).
__global__ void cx_reduce_kernel(int nout, int m, int S, const double2 *__restrict__ part,
                                 double alpha, double beta, const double2 *y0, long long ld0,
                                 double gamma, const double2 *__restrict__ z, long long ldz,
                                 double2 *y, long long ldy) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)nout * m;
  if (idx >= total) return;
  const int i = (int)(idx / m), j = (int)(idx % m);
  double sr = 0.0, si = 0.0;
  for (int s = 0; s < S; s++) {
    const double2 v = part[(size_t)s * total + idx];
    sr += v.x;
    si += v.y;
  }
  double yr = alpha * sr, yi = alpha * si;
  if (y0) {
    const double2 v = y0[(size_t)i * ld0 + j];
    yr = fma(beta, v.x, yr);
    yi = fma(beta, v.y, yi);
  }
  if (z) {
    const double2 v = z[(size_t)i * ldz + j];
    yr = fma(gamma, v.x, yr);
    yi = fma(gamma, v.y, yi);
  }
  y[(size_t)i * ldy + j] = z2(yr, yi);
}

}  // namespace
}  // namespace exh

extern "C" int exh_zgemm_cn_device(void *stream, int conj_a, int nout, int m, int k,
                                   const exh_z *a, long long lda, const exh_z *x, long long ldx,
                                   double alpha, double beta, const exh_z *y0, long long ldy0,
                                   double gamma, const exh_z *z, long long ldz, exh_z *y,
                                   long long ldy, int split, exh_z *work) {
  using namespace exh;
  if (nout < 0 || m < 0 || k < 1 || split < 1 || split > 4096) return EXH_EINVAL;
  if (nout == 0 || m == 0) return EXH_OK;
  if (!a || !x || !y || !work || lda < nout || ldx < m || ldy < m) return EXH_EINVAL;
  if ((y0 && ldy0 < m) || (z && ldz < m)) return EXH_EINVAL;
  // chunks of whole slabs, none empty
  int kchunk = (k + split - 1) / split;
  kchunk = (kchunk + CK - 1) / CK * CK;
  const int S = (k + kchunk - 1) / kchunk;
  hipStream_t s = (hipStream_t)stream;
  if (m == 1) {
    const dim3 grid((unsigned)((nout + VT - 1) / VT), (unsigned)S);
    hipLaunchKernelGGL(cx_part_vec_kernel, grid, dim3(VT), 0, s, nout, k, kchunk,
                       (const double2 *)a, lda, conj_a ? -1.0 : 1.0, (const double2 *)x, ldx,
                       (double2 *)work);
  } else {
    const dim3 grid((unsigned)((nout + CT - 1) / CT), (unsigned)((m + CT - 1) / CT),
                    (unsigned)S);
    hipLaunchKernelGGL(cx_part_kernel, grid, dim3(CNT), 0, s, nout, m, k, kchunk,
                       (const double2 *)a, lda, conj_a ? -1.0 : 1.0, (const double2 *)x, ldx,
                       (double2 *)work);
  }
  const size_t total = (size_t)nout * m;
  hipLaunchKernelGGL(cx_reduce_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     nout, m, S, (const double2 *)work, alpha, beta, (const double2 *)y0, ldy0,
                     gamma, (const double2 *)z, ldz, (double2 *)y, ldy);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? EXH_OK : (int)e;
}
