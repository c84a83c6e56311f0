// Interstitial producers on the device: the smooth step function
// (gencfun.f90:13-159), the G components of the effective potential and
// masses (genveffig.f90:13-44, genmeffig.f90:13-61), and the device form of
// exh_set_interstitial.  The closed form, the grid products and the
// grid <-> G maps are kernels of this file; the 3-D transforms are hipFFT
// double-complex plans, batched over the products of one call.
//
// Grid layout: the FFT box of ngrid(1) x ngrid(2) x ngrid(3) points with the
// first index fastest (zfftifc.f90 / genigfft in gengvec.f90:129-167): the G
// vector with integer coordinates (i1,i2,i3) sits at
//   j = j3*n1*n2 + j2*n1 + j1,   j_d = i_d (i_d >= 0) or n_d + i_d,
// which is hipFFT's row-major layout with dimensions (n3, n2, n1).
#include <hipfft/hipfft.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "exh_device.h"

namespace exh {
namespace {

constexpr double kFourPi = 12.566370614359172954;
constexpr double kAlpha = 1.0 / 137.03599911;  // genmeffig.f90:29

// Per-call constants of the closed form.
struct CfunArgs {
  double b[9];   // bvec, column-major (b1, b2, b3)
  double t1;     // 4 pi / omega (gencfun.f90:120)
  double epslat;
  double damp;   // cfdamp / gmaxvr, 0 = no damping
  int n1, n2, n3;
  int nspecies;
};

__device__ __forceinline__ int fft_index(int i1, int i2, int i3, int n1, int n2, int n3) {
  const int j1 = i1 >= 0 ? i1 : n1 + i1;
  const int j2 = i2 >= 0 ? i2 : n2 + i2;
  const int j3 = i3 >= 0 ? i3 : n3 + i3;
  return (j3 * n2 + j2) * n1 + j1;
}

// gencfun.f90:120-157 for one G vector per thread, then :70 (scatter of the
// undamped value to its grid point) and :76-82 (damping of cfunig).  geo
// holds per species {rmt, first atom, end atom} followed by atposc(3,natmtot)
// in species-major order; its reads are uniform across the wave.
__global__ __launch_bounds__(256) void cfunig_kernel(int ng, const int *__restrict__ ivg,
                                                     CfunArgs a,
                                                     const double *__restrict__ geo,
                                                     double2 *__restrict__ cfunig,
                                                     double2 *__restrict__ zfft) {
  const int ig = blockIdx.x * blockDim.x + threadIdx.x;
  if (ig >= ng) return;
  const int i1 = ivg[3 * ig], i2 = ivg[3 * ig + 1], i3 = ivg[3 * ig + 2];
  const double d1 = i1, d2 = i2, d3 = i3;
  double v[3];
#pragma unroll
  for (int d = 0; d < 3; d++) v[d] = d1 * a.b[d] + d2 * a.b[3 + d] + d3 * a.b[6 + d];
  const double gc = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  const bool finite = gc > a.epslat;
  double re = finite ? 0.0 : 1.0, im = 0.0;
  const double *tau = geo + 3 * a.nspecies;
  for (int is = 0; is < a.nspecies; is++) {
    const double rmt = geo[3 * is];
    const int ia0 = (int)geo[3 * is + 1], ia1 = (int)geo[3 * is + 2];
    double ff;
    if (finite) {
      const double t2 = gc * rmt;
      double s, c;
      sincos(t2, &s, &c);
      ff = a.t1 * (s - t2 * c) / (gc * gc * gc);
    } else {
      ff = a.t1 / 3.0 * rmt * rmt * rmt;
    }
    for (int ia = ia0; ia < ia1; ia++) {
      const double t2 = -(v[0] * tau[3 * ia] + v[1] * tau[3 * ia + 1] + v[2] * tau[3 * ia + 2]);
      double s, c;
      sincos(t2, &s, &c);
      re -= ff * c;
      im -= ff * s;
    }
  }
  zfft[fft_index(i1, i2, i3, a.n1, a.n2, a.n3)] = make_double2(re, im);
  if (a.damp > 0.0) {
    const double t = a.damp * gc, f = exp(-t * t);
    re *= f;
    im *= f;
  }
  cfunig[ig] = make_double2(re, im);
}

// gencfun.f90:74: cfunir = Re(zfft).
__global__ __launch_bounds__(256) void real_part_kernel(int n, const double2 *__restrict__ z,
                                                        double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = z[i].x;
}

// genveffig.f90:33 and genmeffig.f90:38,50: the selected products, stored as
// consecutive grids of n complex values in the order veff, meff, m2eff.
__global__ __launch_bounds__(256) void products_kernel(int n, const double *__restrict__ veffir,
                                                       const double *__restrict__ cfunir,
                                                       int which, double a2,
                                                       double2 *__restrict__ z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = veffir[i], c = cfunir[i];
  double2 *p = z + i;
  if (which & EXH_VEFFIG) {
    *p = make_double2(v * c, 0.0);
    p += n;
  }
  if (which & (EXH_MEFFIG | EXH_M2EFFIG)) {
    const double d = 1.0 - v * a2;
    if (which & EXH_MEFFIG) {
      *p = make_double2(c / d, 0.0);
      p += n;
    }
    if (which & EXH_M2EFFIG) *p = make_double2(c / (d * d), 0.0);
  }
}

// genveffig.f90:36-39 (and genmeffig.f90:41-44, :53-56): out_b(ig) = the
// forward transform at G's grid point, times the 1/N of zfftifc.f90:70-77.
__global__ __launch_bounds__(256) void gather_kernel(int ngvec, const int *__restrict__ ivg,
                                                     int n1, int n2, int n3, int nb,
                                                     const double2 *__restrict__ z, double scale,
                                                     double2 *o0, double2 *o1, double2 *o2) {
  const int ig = blockIdx.x * blockDim.x + threadIdx.x;
  if (ig >= ngvec) return;
  const size_t n = (size_t)n1 * n2 * n3;
  const size_t j = fft_index(ivg[3 * ig], ivg[3 * ig + 1], ivg[3 * ig + 2], n1, n2, n3);
  double2 *outs[3] = {o0, o1, o2};
  for (int b = 0; b < nb; b++) {
    const double2 x = z[b * n + j];
    outs[b][ig] = make_double2(x.x * scale, x.y * scale);
  }
}

int cdiv_i(long long a, int b) { return (int)((a + b - 1) / b); }

void grid_of(const exh_context *c, int *n) {
  for (int d = 0; d < 3; d++) n[d] = c->intgv[3 + d] - c->intgv[d] + 1;
}

// The producers transform the whole G box: every box point is one G vector
// (checked by exh_set_gvectors), so the scatter fills the whole grid and
// every kernel index below stays inside it.
int check_grid(exh_context *c, const char *who) {
  if (!c->have_gvec) return fail(c, EXH_ESTATE, std::string(who) + ": call exh_set_gvectors first");
  if (!c->gvec_full)
    return fail(c, EXH_EINVAL,
                std::string(who) + ": the G set is not the whole FFT grid of the G box "
                                   "(ngrtot = box size, ivgig the inverse of ivg)");
  return 0;
}

const char *fft_error(hipfftResult r) {
  switch (r) {
    case HIPFFT_ALLOC_FAILED: return "allocation failed";
    case HIPFFT_INVALID_VALUE: return "invalid value";
    case HIPFFT_INVALID_SIZE: return "invalid size";
    case HIPFFT_EXEC_FAILED: return "execution failed";
    case HIPFFT_SETUP_FAILED: return "setup failed";
    default: return "error";
  }
}

// In-place double-complex plan over the context's grid for nb grids.
int fft_plan(exh_context *c, int nb, hipfftHandle *out) {
  int n[3];
  grid_of(c, n);
  if (std::memcmp(n, c->fft_grid, sizeof(n)) != 0) {
    interstitial_release(c);
    std::memcpy(c->fft_grid, n, sizeof(n));
  }
  if (!c->fft_plan[nb]) {
    hipfftHandle h = nullptr;
    int dims[3] = {n[2], n[1], n[0]};
    const int dist = c->ngrtot;
    hipfftResult r = hipfftPlanMany(&h, 3, dims, nullptr, 1, dist, nullptr, 1, dist,
                                    HIPFFT_Z2Z, nb);
    if (r != HIPFFT_SUCCESS)
      return fail(c, EXH_ESTATE, std::string("hipfftPlanMany: ") + fft_error(r));
    c->fft_plan[nb] = (void *)h;
  }
  *out = (hipfftHandle)c->fft_plan[nb];
  return 0;
}

int run_fft(exh_context *c, hipStream_t s, int nb, double2 *z, int direction) {
  hipfftHandle h;
  if (int rc = fft_plan(c, nb, &h)) return rc;
  hipfftResult r = hipfftSetStream(h, s);
  hipError_t e = timed(c, EXH_K_IFFT, s, [&] {
    if (r == HIPFFT_SUCCESS)
      r = hipfftExecZ2Z(h, (hipfftDoubleComplex *)z, (hipfftDoubleComplex *)z, direction);
    return hipSuccess;
  });
  if (r != HIPFFT_SUCCESS) return fail(c, EXH_ESTATE, std::string("hipfftExecZ2Z: ") + fft_error(r));
  return e == hipSuccess ? 0 : hip_fail(c, e, "fft");
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

}  // namespace

void interstitial_release(exh_context *c) {
  for (int b = 0; b < 4; b++) {
    if (c->fft_plan[b]) (void)hipfftDestroy((hipfftHandle)c->fft_plan[b]);
    c->fft_plan[b] = nullptr;
  }
  std::memset(c->fft_grid, 0, sizeof(c->fft_grid));
}

}  // namespace exh

using namespace exh;

extern "C" {

int exh_gencfun_device(exh_context *c, void *stream, const double *bvec, const double *atposc,
                       int maxatoms, double epslat, double cfdamp, double gmaxvr,
                       exh_z *cfunig, double *cfunir) {
  if (!c) return EXH_EINVAL;
  if (int rc = wait_done(c)) return rc;
  if (int rc = check_grid(c, "gencfun")) return rc;
  if (!c->have_basis || !c->have_radial)
    return fail(c, EXH_ESTATE, "gencfun: basis and radial data not set");
  if (!bvec || !cfunig || !cfunir || (c->natmtot > 0 && !atposc))
    return fail(c, EXH_EINVAL, "gencfun: null array");
  if (!(epslat >= 0.0) || (cfdamp > 0.0 && !(gmaxvr > 0.0)))
    return fail(c, EXH_EINVAL, "gencfun: epslat < 0, or cfdamp > 0 with gmaxvr <= 0");
  int maxna = 0;
  for (int is = 0; is < c->nspecies; is++) maxna = std::max(maxna, c->natoms[is]);
  if (maxatoms < maxna) return fail(c, EXH_EINVAL, "gencfun: maxatoms below natoms");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, EXH_ESTATE, "hipSetDevice");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;

  // geometry: per species {rmt, first, end}, then positions (species-major)
  std::vector<double> geo(3 * (size_t)c->nspecies + 3 * (size_t)c->natmtot);
  size_t k = 3 * (size_t)c->nspecies;
  int ias = 0;
  for (int is = 0; is < c->nspecies; is++) {
    geo[3 * is] = c->rmt[is];
    geo[3 * is + 1] = ias;
    for (int ia = 0; ia < c->natoms[is]; ia++, ias++)
      for (int d = 0; d < 3; d++) geo[k++] = atposc[d + 3 * ((size_t)ia + (size_t)maxatoms * is)];
    geo[3 * is + 2] = ias;
  }
  int rc = copy_in(c, c->igeo, geo.data(), geo.size());
  if (!rc) rc = ensure(c, c->ifz, (size_t)c->ngrtot * sizeof(double2));
  if (rc) return rc;

  CfunArgs a;
  std::memcpy(a.b, bvec, sizeof(a.b));
  a.t1 = kFourPi / c->omega;
  a.epslat = epslat;
  a.damp = cfdamp > 0.0 ? cfdamp / gmaxvr : 0.0;
  grid_of(c, &a.n1);
  a.nspecies = c->nspecies;
  const int ng = c->ngrtot;
  double2 *z = (double2 *)c->ifz.p;
  hipError_t e = timed(c, EXH_K_CFUNIG, s, [&] {
    hipLaunchKernelGGL(cfunig_kernel, dim3(cdiv_i(ng, 256)), dim3(256), 0, s, ng,
                       (const int *)c->ivg.p, a, (const double *)c->igeo.p,
                       (double2 *)cfunig, z);
    return hipGetLastError();
  });
  if (e != hipSuccess) return hip_fail(c, e, "gencfun");
  if ((rc = run_fft(c, s, 1, z, HIPFFT_BACKWARD))) return rc;
  e = timed(c, EXH_K_IGATHER, s, [&] {
    hipLaunchKernelGGL(real_part_kernel, dim3(cdiv_i(ng, 256)), dim3(256), 0, s, ng,
                       (const double2 *)z, cfunir);
    return hipGetLastError();
  });
  if (e == hipSuccess) e = mark_done(c, s);
  return e == hipSuccess ? 0 : hip_fail(c, e, "gencfun");
}

int exh_genveffig_device(exh_context *c, void *stream, const double *veffir,
                         const double *cfunir, int which, exh_z *veffig, exh_z *meffig,
                         exh_z *m2effig) {
  if (!c) return EXH_EINVAL;
  if (int rc = wait_done(c)) return rc;
  if (int rc = check_grid(c, "genveffig")) return rc;
  if (which <= 0 || (which & ~(EXH_VEFFIG | EXH_MEFFIG | EXH_M2EFFIG)))
    return fail(c, EXH_EINVAL, "genveffig: which selects none or unknown outputs");
  if (!veffir || !cfunir || ((which & EXH_VEFFIG) && !veffig) ||
      ((which & EXH_MEFFIG) && !meffig) || ((which & EXH_M2EFFIG) && !m2effig))
    return fail(c, EXH_EINVAL, "genveffig: null array");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, EXH_ESTATE, "hipSetDevice");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  double2 *outs[3] = {nullptr, nullptr, nullptr};
  int nb = 0;
  if (which & EXH_VEFFIG) outs[nb++] = (double2 *)veffig;
  if (which & EXH_MEFFIG) outs[nb++] = (double2 *)meffig;
  if (which & EXH_M2EFFIG) outs[nb++] = (double2 *)m2effig;
  const int n = c->ngrtot;
  if (int rc = ensure(c, c->ifz, (size_t)nb * n * sizeof(double2))) return rc;
  double2 *z = (double2 *)c->ifz.p;
  const double a2 = 0.5 * kAlpha * kAlpha;
  hipError_t e = timed(c, EXH_K_IPROD, s, [&] {
    hipLaunchKernelGGL(products_kernel, dim3(cdiv_i(n, 256)), dim3(256), 0, s, n, veffir,
                       cfunir, which, a2, z);
    return hipGetLastError();
  });
  if (e != hipSuccess) return hip_fail(c, e, "genveffig");
  if (int rc = run_fft(c, s, nb, z, HIPFFT_FORWARD)) return rc;
  int g[3];
  grid_of(c, g);
  const double scale = 1.0 / (double)n;
  e = timed(c, EXH_K_IGATHER, s, [&] {
    hipLaunchKernelGGL(gather_kernel, dim3(cdiv_i(c->ngvec, 256)), dim3(256), 0, s, c->ngvec,
                       (const int *)c->ivg.p, g[0], g[1], g[2], nb, (const double2 *)z, scale,
                       outs[0], outs[1], outs[2]);
    return hipGetLastError();
  });
  if (e == hipSuccess) e = mark_done(c, s);
  return e == hipSuccess ? 0 : hip_fail(c, e, "genveffig");
}

int exh_gencfun(exh_context *c, const double *bvec, const double *atposc, int maxatoms,
                double epslat, double cfdamp, double gmaxvr, exh_z *cfunig, double *cfunir) {
  if (!c) return EXH_EINVAL;
  if (!cfunig || !cfunir) return fail(c, EXH_EINVAL, "gencfun: null array");
  if (int rc = check_grid(c, "gencfun")) return rc;
  const size_t n = (size_t)c->ngrtot;
  int rc = ensure(c, c->i_out0, n * sizeof(double2));
  if (!rc) rc = ensure(c, c->i_out1, n * sizeof(double));
  if (!rc)
    rc = exh_gencfun_device(c, nullptr, bvec, atposc, maxatoms, epslat, cfdamp, gmaxvr,
                            (exh_z *)c->i_out0.p, (double *)c->i_out1.p);
  if (!rc) rc = copy_out(c, cfunig, c->i_out0, n * sizeof(double2));
  if (!rc) rc = copy_out(c, cfunir, c->i_out1, n * sizeof(double));
  return rc;
}

static int host_products(exh_context *c, const double *veffir, const double *cfunir, int which,
                         exh_z *o0, exh_z *o1, exh_z *o2) {
  if (!veffir || !cfunir) return fail(c, EXH_EINVAL, "genveffig: null array");
  if (int rc = check_grid(c, "genveffig")) return rc;
  const size_t n = (size_t)c->ngrtot, ng = (size_t)c->ngvec;
  int rc = copy_in(c, c->i_in0, veffir, n);
  if (!rc) rc = copy_in(c, c->i_in1, cfunir, n);
  DevBuf *bufs[3] = {&c->i_out0, &c->i_out1, &c->i_out2};
  exh_z *dst[3] = {o0, o1, o2};
  for (int b = 0; b < 3 && !rc; b++)
    if (dst[b]) rc = ensure(c, *bufs[b], ng * sizeof(double2));
  if (rc) return rc;
  rc = exh_genveffig_device(c, nullptr, (const double *)c->i_in0.p, (const double *)c->i_in1.p,
                            which, o0 ? (exh_z *)c->i_out0.p : nullptr,
                            o1 ? (exh_z *)c->i_out1.p : nullptr,
                            o2 ? (exh_z *)c->i_out2.p : nullptr);
  for (int b = 0; b < 3 && !rc; b++)
    if (dst[b]) rc = copy_out(c, dst[b], *bufs[b], ng * sizeof(double2));
  return rc;
}

int exh_genveffig(exh_context *c, const double *veffir, const double *cfunir, exh_z *veffig) {
  if (!c) return EXH_EINVAL;
  if (!veffig) return fail(c, EXH_EINVAL, "genveffig: null array");
  return host_products(c, veffir, cfunir, EXH_VEFFIG, veffig, nullptr, nullptr);
}

int exh_genmeffig(exh_context *c, const double *veffir, const double *cfunir,
                  int valence_rel_none, int iora_star, exh_z *meffig, exh_z *m2effig) {
  if (!c) return EXH_EINVAL;
  if (valence_rel_none) return 0;  // genmeffig.f90:34,46: nothing to produce
  if (!meffig || (iora_star && !m2effig)) return fail(c, EXH_EINVAL, "genmeffig: null array");
  return host_products(c, veffir, cfunir, EXH_MEFFIG | (iora_star ? EXH_M2EFFIG : 0), nullptr,
                       meffig, iora_star ? m2effig : nullptr);
}

int exh_set_interstitial_device(exh_context *c, void *stream, const exh_z *cfunig,
                                const exh_z *veffig, const exh_z *meffig,
                                const exh_z *m2effig, int valence_rel_none, int h1on) {
  if (!c) return EXH_EINVAL;
  if (int rc = wait_done(c)) return rc;
  if (!c->have_gvec) return fail(c, EXH_ESTATE, "set_interstitial: call exh_set_gvectors first");
  if (!cfunig || !veffig || (!valence_rel_none && !meffig) || (h1on && !m2effig))
    return fail(c, EXH_EINVAL, "set_interstitial: null array");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, EXH_ESTATE, "hipSetDevice");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  hipError_t e = launch_ist_table(s, c, (const double2 *)cfunig, (const double2 *)veffig,
                                  (const double2 *)meffig, (const double2 *)m2effig,
                                  valence_rel_none ? 1 : 0, h1on ? 1 : 0);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(c, e, "set_interstitial");
  c->have_ist = true;
  return 0;
}

}  // extern "C"
