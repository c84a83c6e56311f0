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

// Geometry of the ivgig box for the interstitial gather.
struct IstBox {
  const int4 *ivgk;    // ivg(:, igpig(i)) per G+k vector, w = valid flag
  const double *vgpc;  // (3, ngkmax) cartesian G+k
  const IstEntry *T;   // one entry per box point
  int lo0, lo1, lo2, n0, n1, n2;
};

// Interstitial contributions for the upper element (i, j), i <= j
// (hmlistln.f90:61-84, olpistln.f90:53-82): H += veff + t1 kin,
// O += cfun + t1 okin with t1 = (G_i+k).(G_j+k)/2; skipped when the G-G'
// index falls outside the box or beyond ngvec (zero entry).
__device__ __forceinline__ void ist_add(const IstBox &b, int i, int j, double &hr, double &hi,
                                        double &orr, double &oi) {
  const int4 gi = b.ivgk[i], gj = b.ivgk[j];
  if (!(gi.w & gj.w)) return;
  const int d0 = gi.x - gj.x - b.lo0, d1 = gi.y - gj.y - b.lo1, d2 = gi.z - gj.z - b.lo2;
  if ((unsigned)d0 >= (unsigned)b.n0 || (unsigned)d1 >= (unsigned)b.n1 ||
      (unsigned)d2 >= (unsigned)b.n2)
    return;
  const IstEntry e = b.T[d0 + (size_t)b.n0 * (d1 + (size_t)b.n1 * d2)];
  const double *vi = b.vgpc + 3 * (size_t)i, *vj = b.vgpc + 3 * (size_t)j;
  const double t1 = 0.5 * (vi[0] * vj[0] + vi[1] * vj[1] + vi[2] * vj[2]);
  hr += e.veff.x + t1 * e.kin.x;
  hi += e.veff.y + t1 * e.kin.y;
  orr += e.ocf.x + t1 * e.okin.x;
  oi += e.ocf.y + t1 * e.okin.y;
}

// Final write of one upper element (i <= j) of H and O and its mirror
// (hamiltonoverlapocopy_UL.f90: lower := conj(upper)).  With accumulate the
// value already in the upper triangle is added first (zgemm beta = 1).
__device__ __forceinline__ void store_upper(double2 *H, double2 *O, long long ldh, int i, int j,
                                            double hr, double hi, double orr, double oi,
                                            int accumulate) {
  const size_t up = (size_t)i + (size_t)j * ldh;
  if (accumulate) {
    const double2 h0 = H[up], o0 = O[up];
    hr += h0.x;
    hi += h0.y;
    orr += o0.x;
    oi += o0.y;
  }
  H[up] = z2(hr, hi);
  O[up] = z2(orr, oi);
  if (i != j) {
    const size_t lo = (size_t)j + (size_t)i * ldh;
    H[lo] = z2(hr, -hi);
    O[lo] = z2(orr, -oi);
  }
}

}  // namespace exh
