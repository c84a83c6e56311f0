// Device helpers shared by the FP64 and split-precision H/O kernels.
#pragma once

#include "exh_internal.h"

namespace exh {

__device__ __forceinline__ double2 z2(double r, double i) {
  double2 v;
  v.x = r;
  v.y = i;
  return v;
}

// Device buffers live in the global address space.  A pointer read from
// memory (a per-k descriptor) is generic to the compiler, and loads through
// it become flat loads, which also count against lgkmcnt and so delay every
// LDS wait.  The grouped kernels therefore pass descriptor pointers on as
// EXH_G-qualified types; the kernel bodies are templated on pointer types.
#define EXH_G __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ EXH_G T *gptr(T *p) {
  return (EXH_G T *)p;
}

// Read-only tables whose addresses come from a descriptor: in the constant
// address space a load at a wave-uniform address becomes a scalar load (as
// for kernel arguments), instead of a vector load of the same value in
// every lane.
#define EXH_C __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ const EXH_C T *cptr(const T *p) {
  return (const EXH_C T *)p;
}

// Loads and stores of double2 / int4 through either kind of pointer (HIP's
// vector classes cannot be assigned in an explicit address space, so the
// EXH_G forms go through clang's native vector types).
typedef double dv2 __attribute__((ext_vector_type(2)));
typedef int iv4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ double2 ld(const double2 *p) { return *p; }
__device__ __forceinline__ double2 ld(const EXH_G double2 *p) {
  const dv2 v = *(const EXH_G dv2 *)p;
  return z2(v.x, v.y);
}
__device__ __forceinline__ double2 ld(const EXH_C double2 *p) {
  const dv2 v = *(const EXH_C dv2 *)p;
  return z2(v.x, v.y);
}
__device__ __forceinline__ double ld(const double *p) { return *p; }
__device__ __forceinline__ double ld(const EXH_G double *p) { return *p; }
__device__ __forceinline__ int4 ld(const int4 *p) { return *p; }
__device__ __forceinline__ int4 ld(const EXH_G int4 *p) {
  const iv4 v = *(const EXH_G iv4 *)p;
  return make_int4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st(double2 *p, double2 v) { *p = v; }
__device__ __forceinline__ void st(EXH_G double2 *p, double2 v) {
  *(EXH_G dv2 *)p = dv2{v.x, v.y};
}
// Streaming stores of data that will not be read from cache again (arrays
// far larger than the caches): non-temporal, one 16-byte store.
__device__ __forceinline__ void st_nt(double2 *p, double2 v) {
  double *q = (double *)p;
  __builtin_nontemporal_store(v.x, q);
  __builtin_nontemporal_store(v.y, q + 1);
}
__device__ __forceinline__ void st_nt(EXH_G double2 *p, double2 v) {
  EXH_G double *q = (EXH_G double *)p;
  __builtin_nontemporal_store(v.x, q);
  __builtin_nontemporal_store(v.y, q + 1);
}
__device__ __forceinline__ void st(int4 *p, int4 v) { *p = v; }
__device__ __forceinline__ void st(EXH_G int4 *p, int4 v) {
  *(EXH_G iv4 *)p = iv4{v.x, v.y, v.z, v.w};
}

// Geometry of the ivgig box for the interstitial gather.
template <typename P4 = const int4 *, typename PD = const double *>
struct IstBoxT {
  P4 ivgk;             // ivg(:, igpig(i)) per G+k vector, w = valid flag
  PD vgpc;             // (3, ngkmax) cartesian G+k
  const IstEntry *T;   // one entry per box point
  int lo0, lo1, lo2, n0, n1, n2;
};
struct IstBox : IstBoxT<> {};

// Interstitial contributions for the upper element (i, j), i <= j
// (hmlistln.f90:61-84, olpistln.f90:53-82): H += veff + t1 kin,
// O += cfun + t1 okin with t1 = (G_i+k).(G_j+k)/2; skipped when the G-G'
// index falls outside the box or beyond ngvec (zero entry).
template <typename B>
__device__ __forceinline__ void ist_add(const B &b, int i, int j, double &hr, double &hi,
                                        double &orr, double &oi) {
  const int4 gi = ld(b.ivgk + i), gj = ld(b.ivgk + j);
  if (!(gi.w & gj.w)) return;
  const int d0 = gi.x - gj.x - b.lo0, d1 = gi.y - gj.y - b.lo1, d2 = gi.z - gj.z - b.lo2;
  if ((unsigned)d0 >= (unsigned)b.n0 || (unsigned)d1 >= (unsigned)b.n1 ||
      (unsigned)d2 >= (unsigned)b.n2)
    return;
  const IstEntry e = b.T[d0 + (size_t)b.n0 * (d1 + (size_t)b.n1 * d2)];
  const auto vi = b.vgpc + 3 * (size_t)i;
  const auto vj = b.vgpc + 3 * (size_t)j;
  const double t1 = 0.5 * (vi[0] * vj[0] + vi[1] * vj[1] + vi[2] * vj[2]);
  hr += e.veff.x + t1 * e.kin.x;
  hi += e.veff.y + t1 * e.kin.y;
  orr += e.ocf.x + t1 * e.okin.x;
  oi += e.ocf.y + t1 * e.okin.y;
}

// Final write of one upper element (i <= j) of H and O and its mirror
// (hamiltonoverlapocopy_UL.f90: lower := conj(upper)).  With accumulate the
// value already in the upper triangle is added first (zgemm beta = 1).
template <typename PH>
__device__ __forceinline__ void store_upper(PH H, PH O, long long ldh, int i, int j,
                                            double hr, double hi, double orr, double oi,
                                            int accumulate) {
  const size_t up = (size_t)i + (size_t)j * ldh;
  if (accumulate) {
    const double2 h0 = ld(H + up), o0 = ld(O + up);
    hr += h0.x;
    hi += h0.y;
    orr += o0.x;
    oi += o0.y;
  }
  st(H + up, z2(hr, hi));
  st(O + up, z2(orr, oi));
  if (i != j) {
    const size_t lo = (size_t)j + (size_t)i * ldh;
    st(H + lo, z2(hr, -hi));
    st(O + lo, z2(orr, -oi));
  }
}

// O alone (overlapsetup without hamiltonsetup): the upper element and its
// mirror.
template <typename PH>
__device__ __forceinline__ void store_upper_o(PH O, long long ldh, int i, int j, double orr,
                                              double oi) {
  st(O + (size_t)i + (size_t)j * ldh, z2(orr, oi));
  if (i != j) st(O + (size_t)j + (size_t)i * ldh, z2(orr, -oi));
}

}  // namespace exh
