// C ABI of libexh (declared in include/exh.h): context, module state held
// resident in HBM, and the match / H+O setup entry points.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <system_error>
#include <thread>
#include <vector>

#include "exh_device.h"

namespace exh {

int fail(exh_context *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(exh_context *c, hipError_t e, const char *where) {
  if (c) c->err = std::string(where) + ": " + hipGetErrorString(e);
  return (int)e > 0 ? (int)e : 1;
}

int ensure(exh_context *c, DevBuf &b, size_t bytes) {
  if (bytes <= b.bytes && b.p) return 0;
  if (c) c->epoch++;  // device addresses change: cached batch graphs are stale
  if (b.p) {
    hipError_t e = hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    if (e != hipSuccess) return hip_fail(c, e, "hipFree");
  }
  size_t n = std::max<size_t>(bytes, 256);
  hipError_t e = hipMalloc(&b.p, n);
  if (e != hipSuccess) {
    b.p = nullptr;
    return hip_fail(c, e, "hipMalloc");
  }
  b.bytes = n;
  return 0;
}

void release(DevBuf &b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

template <typename T>
static int upload(exh_context *c, DevBuf &b, const T *src, size_t n) {
  int rc = ensure(c, b, n * sizeof(T));
  if (rc) return rc;
  if (n == 0) return 0;
  hipError_t e = hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? 0 : hip_fail(c, e, "upload");
}

// ---------------------------------------------------------------------------
// Host arrays of the host-pointer entry points are pageable (Fortran or NumPy
// arrays).  A pageable hipMemcpy moves about 12 GB/s; these copies instead go
// through two pinned buffers of kPinBytes, the host side of each piece split
// over threads, the DMA of one piece overlapping the host copy of the next.
// A copy is a 2-D region: `rows` rows of `width` bytes, `hpitch` / `dpitch`
// bytes apart on the host / device side (a 1-D copy is one row).
// ---------------------------------------------------------------------------
constexpr size_t kPinBytes = 64u << 20;
constexpr size_t kPinDirect = 4u << 20;  // smaller copies: one plain hipMemcpy

static int pin_threads() {
  static const int n = [] {
    const char *e = getenv("EXH_COPY_THREADS");
    const int v = e ? atoi(e) : 8;
    return std::max(1, std::min(v, 64));
  }();
  return n;
}

static void pin_release(exh_context *c) {
  for (int b = 0; b < 2; b++) {
    if (c->pin_ev[b]) (void)hipEventDestroy(c->pin_ev[b]);
    if (c->pin[b]) (void)hipHostFree(c->pin[b]);
    c->pin[b] = nullptr;
    c->pin_ev[b] = nullptr;
  }
  c->pin_bytes = 0;
}

// both buffers and their events, or none (a failed call frees what it made)
static hipError_t pin_ensure(exh_context *c) {
  if (c->pin_bytes) return hipSuccess;
  for (int b = 0; b < 2; b++) {
    void *p = nullptr;
    hipError_t e = hipHostMalloc(&p, kPinBytes, hipHostMallocDefault);
    if (e == hipSuccess) c->pin[b] = p;
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->pin_ev[b], hipEventDisableTiming);
    if (e != hipSuccess) {
      pin_release(c);
      return e;
    }
  }
  c->pin_bytes = kPinBytes;
  return hipSuccess;
}

// rows [r0, r1) of `width` bytes from src (pitch sp) to dst (pitch dp), over
// up to pin_threads() threads (by rows, or by bytes of a single row)
static void par_copy(char *dst, size_t dp, const char *src, size_t sp, size_t width, size_t nrow) {
  const size_t total = width * nrow;
  const int nt = (int)std::min<size_t>(pin_threads(), std::max<size_t>(1, total >> 20));
  auto part = [&](int t) {
    if (nrow == 1) {
      const size_t a = width * t / nt, b = width * (t + 1) / nt;
      memcpy(dst + a, src + a, b - a);
      return;
    }
    const size_t a = nrow * t / nt, b = nrow * (t + 1) / nt;
    for (size_t r = a; r < b; r++) memcpy(dst + r * dp, src + r * sp, width);
  };
  if (nt == 1) {
    part(0);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt - 1);
  int t = 1;
  try {
    for (; t < nt; t++) th.emplace_back(part, t);
  } catch (const std::system_error &) {
    for (; t < nt; t++) part(t);  // no thread to be had: this one copies the rest
  }
  part(0);
  for (auto &x : th) x.join();
}

// one DMA of `rows` rows (a plain copy for one row: pitches may be 0 there)
static hipError_t dma(void *dst, size_t dp, const void *src, size_t sp, size_t width,
                      size_t rows, hipMemcpyKind kind, hipStream_t s) {
  if (rows == 1) return hipMemcpyAsync(dst, src, width, kind, s);
  return hipMemcpy2DAsync(dst, dp, src, sp, width, rows, kind, s);
}

// The pieces of a region: whole rows per piece when a row fits a buffer,
// else byte ranges of one row.
struct Piece {
  size_t row, off, len, nrow;
};
static std::vector<Piece> pieces(size_t width, size_t rows) {
  std::vector<Piece> v;
  if (width <= kPinBytes) {
    const size_t per = kPinBytes / width;
    for (size_t r = 0; r < rows; r += per) v.push_back({r, 0, width, std::min(per, rows - r)});
  } else {
    for (size_t r = 0; r < rows; r++)
      for (size_t o = 0; o < width; o += kPinBytes)
        v.push_back({r, o, std::min(kPinBytes, width - o), 1});
  }
  return v;
}

// host -> device, enqueued on s (returns once every piece is in pinned
// memory; the DMAs complete in stream order)
static hipError_t h2d(exh_context *c, hipStream_t s, void *dst, size_t dpitch, const void *src,
                      size_t hpitch, size_t width, size_t rows) {
  if (!width || !rows) return hipSuccess;
  if (width * rows <= kPinDirect)
    return dma(dst, dpitch, src, hpitch, width, rows, hipMemcpyHostToDevice, s);
  hipError_t e = pin_ensure(c);
  if (e != hipSuccess) return e;
  const auto ps = pieces(width, rows);
  for (size_t i = 0; i < ps.size() && e == hipSuccess; i++) {
    const Piece &p = ps[i];
    const int b = (int)(i & 1);
    e = hipEventSynchronize(c->pin_ev[b]);  // the last DMA out of this buffer is done
    if (e != hipSuccess) break;
    par_copy((char *)c->pin[b], p.len, (const char *)src + p.row * hpitch + p.off, hpitch, p.len,
             p.nrow);
    e = dma((char *)dst + p.row * dpitch + p.off, dpitch, c->pin[b], p.len, p.len, p.nrow,
            hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipEventRecord(c->pin_ev[b], s);
  }
  return e;
}

// device -> host after the work already enqueued on s; returns when dst holds
// the data
static hipError_t d2h(exh_context *c, hipStream_t s, void *dst, size_t hpitch, const void *src,
                      size_t dpitch, size_t width, size_t rows) {
  if (!width || !rows) return hipSuccess;
  if (width * rows <= kPinDirect) {
    hipError_t e = dma(dst, hpitch, src, dpitch, width, rows, hipMemcpyDeviceToHost, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
  }
  hipError_t e = pin_ensure(c);
  if (e != hipSuccess) return e;
  const auto ps = pieces(width, rows);
  auto issue = [&](size_t i) {
    const Piece &p = ps[i];
    const int b = (int)(i & 1);
    hipError_t r = dma(c->pin[b], p.len, (const char *)src + p.row * dpitch + p.off, dpitch,
                       p.len, p.nrow, hipMemcpyDeviceToHost, s);
    return r == hipSuccess ? hipEventRecord(c->pin_ev[b], s) : r;
  };
  e = issue(0);
  for (size_t i = 0; i < ps.size() && e == hipSuccess; i++) {
    // buffer (i+1)&1 was last read by the host copy of piece i-1, finished
    if (i + 1 < ps.size()) e = issue(i + 1);
    if (e == hipSuccess) e = hipEventSynchronize(c->pin_ev[i & 1]);
    if (e != hipSuccess) break;
    const Piece &p = ps[i];
    par_copy((char *)dst + p.row * hpitch + p.off, hpitch, (const char *)c->pin[i & 1], p.len,
             p.len, p.nrow);
  }
  return e;
}

// polynom.f90 (m >= 0): m-th derivative at x of the degree np-1 polynomial
// through (xa, ya).  Used for the boundary-derivative matrix of match.
static double polynom_deriv(int m, int np, const double *xa, const double *ya, double x) {
  double c[16];
  if (np == 4) {
    const double x1 = xa[1] - xa[0], x2 = xa[2] - xa[0], x3 = xa[3] - xa[0];
    const double y1 = ya[1] - ya[0], y2 = ya[2] - ya[0], y3 = ya[3] - ya[0];
    double t1 = 1.0 / (x1 * x2 * x3 * (x1 - x2) * (x1 - x3) * (x2 - x3));
    const double t2 = x1 * x2 * y3, t3 = x2 * x3 * y1, t4 = x3 * x1 * y2;
    const double t5 = x1 * x1, t6 = x2 * x2, t7 = x3 * x3;
    c[1] = t1 * (t3 * (x3 * t6 - x2 * t7) + t4 * (x1 * t7 - x3 * t5) + t2 * (x2 * t5 - x1 * t6));
    c[2] = t1 * (t3 * (t7 - t6) + t4 * (t5 - t7) + t2 * (t6 - t5));
    c[3] = t1 * (t3 * (x2 - x3) + t4 * (x3 - x1) + t2 * (x1 - x2));
    t1 = x - xa[0];
    switch (m) {
      case 0: return ya[0] + t1 * (c[1] + t1 * (c[2] + c[3] * t1));
      case 1: return c[1] + t1 * (2.0 * c[2] + 3.0 * c[3] * t1);
      case 2: return 6.0 * c[3] * t1 + 2.0 * c[2];
      case 3: return 6.0 * c[3];
      default: return 0.0;
    }
  }
  if (m >= np) return 0.0;
  for (int i = 0; i < np; i++) c[i] = ya[i];
  for (int i = 2; i <= np; i++)
    for (int j = np; j >= i; j--) c[j - 1] = (c[j - 1] - c[j - 2]) / (xa[j - 1] - xa[j - i]);
  if (m == 0) {
    double sum = c[0], t1 = 1.0;
    for (int i = 2; i <= np; i++) {
      t1 = t1 * (x - xa[i - 2]);
      sum = sum + c[i - 1] * t1;
    }
    return sum;
  }
  const double x0 = xa[0];
  for (int j = 1; j <= np - 1; j++)
    for (int i = 1; i <= np - j; i++) {
      const int k = np - i;
      c[k - 1] = c[k - 1] + (x0 - xa[k - j]) * c[k];
    }
  for (int j = 1; j <= m; j++)
    for (int i = m + 1; i <= np; i++) c[i - 1] = c[i - 1] * (double)(i - j);
  double t1 = c[np - 1];
  const double t2 = x - x0;
  for (int i = np - 1; i >= m + 1; i--) t1 = t1 * t2 + c[i - 1];
  return t1;
}

// zgetf2 on a real p x p matrix (leading dimension ld): partial pivoting by
// largest |a|, reciprocal scaling of the column, rank-1 update.  Returns info.
static int lu_real(int p, double *a, int ld, int *ipiv) {
  int info = 0;
  for (int j = 0; j < p; j++) {
    int jp = j;
    double best = std::fabs(a[j + ld * j]);
    for (int i = j + 1; i < p; i++)
      if (std::fabs(a[i + ld * j]) > best) {
        best = std::fabs(a[i + ld * j]);
        jp = i;
      }
    ipiv[j] = jp + 1;
    if (a[jp + ld * j] != 0.0) {
      if (jp != j)
        for (int k = 0; k < p; k++) std::swap(a[j + ld * k], a[jp + ld * k]);
      if (j < p - 1) {
        const double r = 1.0 / a[j + ld * j];
        for (int i = j + 1; i < p; i++) a[i + ld * j] = r * a[i + ld * j];
      }
    } else if (info == 0) {
      info = j + 1;
    }
    for (int k = j + 1; k < p; k++) {
      const double t = -a[j + ld * k];
      for (int i = j + 1; i < p; i++) a[i + ld * k] += a[i + ld * j] * t;
    }
  }
  return info;
}

// Rebuild the LU tables of match once both basis and radial data are known.
static int build_match_tables(exh_context *c) {
  constexpr int PC = 3;
  const int L = c->lmaxapw, pmax = c->apwordmax;
  const int np = std::max(pmax + 1, 4);
  std::vector<double> lu((size_t)PC * PC * (L + 1) * c->natmtot, 0.0);
  std::vector<int> pv((size_t)PC * (L + 1) * c->natmtot, 1);
  for (int ias = 0; ias < c->natmtot; ias++) {
    const int is = c->atom_species[ias];
    const int ir = c->nrmt_h[is] - np;
    if (ir < 0) return fail(c, EXH_EINVAL, "nrmt smaller than the polynomial order of match");
    for (int l = 0; l <= L; l++) {
      const int p = c->apword[l + (L + 1) * is];
      double *a = lu.data() + (size_t)PC * PC * (l + (L + 1) * ias);
      for (int io2 = 0; io2 < p; io2++) {
        const double *ya = c->apwfr_h.data() + ir +
                           (size_t)c->nrmtmax * (2 * (io2 + (size_t)pmax * (l + (size_t)(L + 1) * ias)));
        for (int io1 = 0; io1 < p; io1++)
          a[io1 + PC * io2] = polynom_deriv(io1, np, c->spr_h.data() + ir + (size_t)c->nrmtmax * is,
                                            ya, c->rmt[is]);
      }
      int info = lu_real(p, a, PC, pv.data() + (size_t)PC * (l + (L + 1) * ias));
      if (info != 0) {
        char buf[160];
        snprintf(buf, sizeof buf,
                 "match: singular boundary-derivative matrix (species %d, atom %d, l %d, "
                 "info %d)", is + 1, ias + 1, l, info);
        return fail(c, EXH_EMATCH, buf);
      }
    }
  }
  int rc = upload(c, c->mt_match.lu, lu.data(), lu.size());
  if (!rc) rc = upload(c, c->mt_match.piv, pv.data(), pv.size());
  return rc;
}

hipEvent_t take_event(exh_context *c) {
  if (!c->evpool.empty()) {
    hipEvent_t e = c->evpool.back();
    c->evpool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Bracket one launch with events when profiling is on.
int drain_records(exh_context *c) {
  for (auto &r : c->recs) {
    hipError_t e = hipEventSynchronize(r.b);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, r.a, r.b);
    if (e != hipSuccess) return hip_fail(c, e, "profile");
    c->prof_ms[r.k] += ms;
    c->prof_n[r.k] += 1;
    c->evpool.push_back(r.a);
    c->evpool.push_back(r.b);
  }
  c->recs.clear();
  return 0;
}

}  // namespace exh

namespace exh {

hipError_t mark_done(exh_context *c, hipStream_t s) {
  if (c->capturing) return hipSuccess;
  if (!c->done_ev) {
    hipError_t e = hipEventCreateWithFlags(&c->done_ev, hipEventDisableTiming);
    if (e != hipSuccess) {
      c->done_ev = nullptr;
      return e;
    }
  }
  hipError_t e = hipEventRecord(c->done_ev, s);
  if (e == hipSuccess) c->done_pending = true;
  return e;
}

int wait_done(exh_context *c) {
  if (!c->done_pending) return 0;
  hipError_t e = hipEventSynchronize(c->done_ev);
  c->done_pending = false;
  return e == hipSuccess ? 0 : hip_fail(c, e, "waiting for queued work");
}

// Constants and scratch of the split-precision path for nmod moduli.
int oz_prepare(exh_context *c, int nmod, OzLayout &L, int &bits) {
  L = oz_layout(c->ngkmax, c->Ktot, nmod);
  bits = oz_bits(nmod, c->Ktot);
  if (!c->oz_consts) {
    hipError_t e = oz_upload_constants(nmod);
    if (e != hipSuccess) return hip_fail(c, e, "split: constants");
    c->oz_consts = true;
  }
  int rc = ensure(c, c->ozqx, L.qx_bytes);
  if (!rc) rc = ensure(c, c->ozqb, L.qb_bytes);
  if (!rc) rc = ensure(c, c->ozr, L.r_bytes);
  if (!rc) rc = ensure(c, c->ozex, L.e_bytes);
  if (!rc) rc = ensure(c, c->ozeb, L.e_bytes);
  return rc;
}

}  // namespace exh

// Device scratch of exh_setup_device_noalloc: one copy per GPU, shared by
// every context on that GPU and held until the process ends.  Host code never
// passes `scratch` itself to a kernel (on the host the name is only a
// stand-in for the symbol): exh_create looks up each device's address once,
// with that device current, and the entry point uses that address.
constexpr size_t kNoallocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_addr[kMaxDevices];

// Address of `scratch` on `device`, which must be the current device.
int scratch_address(int device, void **out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return EXH_ERANGE;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return e != hipSuccess ? (int)e : 1;
    g_scratch_addr[device] = p;
  }
  *out = g_scratch_addr[device];
  return 0;
}

inline size_t align256(size_t n) { return (n + 255) / 256 * 256; }
}  // namespace

using namespace exh;

#define EXH_CHECK_CTX(c) \
  if (!(c)) return EXH_EINVAL

extern "C" {

int exh_create(int device, exh_context **out) {
  if (!out) return EXH_EINVAL;
  *out = nullptr;
  exh_context *c = new (std::nothrow) exh_context();
  if (!c) return EXH_EINVAL;
  c->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        ncu > 0)
      c->ncu = ncu;
  }
  // created here so that no entry point creates it later (mark_done)
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->done_ev, hipEventDisableTiming);
  if (e != hipSuccess) {
    int rc = (int)e;
    if (c->done_ev) (void)hipEventDestroy(c->done_ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return rc > 0 ? rc : 1;
  }
  if (int rc = scratch_address(device, &c->noalloc_scratch)) {
    (void)hipEventDestroy(c->done_ev);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  *out = c;
  return 0;
}

void exh_destroy(exh_context *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  batch_release_all(c);
  interstitial_release(c);
  pin_release(c);
  if (c->done_ev) (void)hipEventDestroy(c->done_ev);
  DevBuf *bufs[] = {&c->colA, &c->katom, &c->atom_species_d, &c->apword_d, &c->omax_d,
                    &c->lorow, &c->lobase_d, &c->rmt_d, &c->mt_match.lu, &c->mt_match.piv,
                    &c->aa, &c->loa, &c->lolo, &c->salo, &c->ololo_eff, &c->h1aa,
                    &c->ivg, &c->ivgig, &c->ist, &c->cf_stage, &c->vf_stage, &c->mf_stage,
                    &c->m2_stage, &c->bt, &c->a2t, &c->at, &c->ivgk, &c->mtab, &c->bmtab, &c->ozqx, &c->ozqb, &c->ozr, &c->ozex, &c->ozeb, &c->s_apwalm, &c->s_igpig,
                    &c->s_vgpc, &c->s_h, &c->s_o, &c->s_gpc, &c->s_tpgpc, &c->s_sfac,
                    &c->bbt, &c->bat, &c->ba2t, &c->bivgk, &c->ifz, &c->igeo, &c->i_in0,
                    &c->i_in1, &c->i_out0, &c->i_out1, &c->i_out2, &c->apwfr_d, &c->hml_fl,
                    &c->hml_arow, &c->hml_lrow, &c->hml_count, &c->hml_gstart, &c->hml_glm2,
                    &c->hml_gval, &c->hml_w, &c->hml_spr, &c->hml_ipot, &c->hml_ikin,
                    &c->hml_vin, &c->hml_loin, &c->hml_aa, &c->hml_loa, &c->hml_lolo,
                    &c->so_dt, &c->so_rowb, &c->so_pairs, &c->so_pstart, &c->so_host};
  for (DevBuf *b : bufs) release(*b);
  for (auto &r : c->recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (hipEvent_t e : c->evpool) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *exh_error(const exh_context *c) { return c ? c->err.c_str() : "null context"; }

void *exh_stream(exh_context *c) { return c ? (void *)c->stream : nullptr; }

int exh_set_basis(exh_context *c, int nspecies, const int *natoms, int lmaxapw, int lmaxmat,
                  int apwordmax, const int *apword, int nlomax, const int *nlorb,
                  const int *lorbl, int ngkmax) {
  EXH_CHECK_CTX(c);
  if (int rcw = wait_done(c)) return rcw;
  if (nspecies <= 0 || !natoms || !apword || !nlorb || (nlomax > 0 && !lorbl))
    return fail(c, EXH_EINVAL, "set_basis: null or empty argument");
  if (lmaxapw < 0 || lmaxmat < 0 || lmaxmat > lmaxapw)
    return fail(c, EXH_EINVAL, "set_basis: need 0 <= lmaxmat <= lmaxapw");
  if (lmaxapw > 12) return fail(c, EXH_ERANGE, "set_basis: lmaxapw > 12 not compiled");
  if (apwordmax < 1 || apwordmax > 3)
    return fail(c, EXH_ERANGE, "set_basis: apwordmax must be 1..3");
  if (ngkmax <= 0) return fail(c, EXH_EINVAL, "set_basis: ngkmax <= 0");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, EXH_ESTATE, "hipSetDevice");
  c->have_basis = c->have_radial = c->have_mt = false;
  c->uniform_apword = true;
  c->apw_host = nullptr;
  c->nspecies = nspecies;
  c->lmaxapw = lmaxapw;
  c->lmaxmat = lmaxmat;
  c->apwordmax = apwordmax;
  c->nlomax = nlomax;
  c->ngkmax = ngkmax;
  c->lmmax = (lmaxapw + 1) * (lmaxapw + 1);
  c->natoms.assign(natoms, natoms + nspecies);
  c->apword.assign(apword, apword + (size_t)(lmaxapw + 1) * nspecies);
  c->nlorb.assign(nlorb, nlorb + nspecies);
  c->lorbl.assign(nlomax > 0 ? lorbl : nlorb, nlomax > 0 ? lorbl + (size_t)nlomax * nspecies : nlorb);
  c->natmtot = 0;
  for (int is = 0; is < nspecies; is++) {
    if (natoms[is] < 0) return fail(c, EXH_EINVAL, "set_basis: natoms < 0");
    c->natmtot += natoms[is];
    if (nlorb[is] < 0 || nlorb[is] > nlomax) return fail(c, EXH_EINVAL, "set_basis: nlorb");
    for (int l = 0; l <= lmaxapw; l++) {
      int p = apword[l + (lmaxapw + 1) * is];
      if (p < 1 || p > apwordmax) return fail(c, EXH_EINVAL, "set_basis: apword out of 1..apwordmax");
      if (p != apwordmax) c->uniform_apword = false;
    }
    for (int ilo = 0; ilo < nlorb[is]; ilo++) {
      int l = lorbl[ilo + nlomax * is];
      if (l < 0 || l > lmaxapw) return fail(c, EXH_EINVAL, "set_basis: lorbl out of range");
    }
  }
  // per-species sizes (MaxAPWs, losize), per-atom offsets, stacked-K map
  std::vector<int> Msp(nspecies, 0);
  c->losize.assign(nspecies, 0);
  c->maxaa = 0;
  c->maxnlo = 0;
  for (int is = 0; is < nspecies; is++) {
    for (int l = 0; l <= lmaxmat; l++) Msp[is] += (2 * l + 1) * apword[l + (lmaxapw + 1) * is];
    for (int ilo = 0; ilo < nlorb[is]; ilo++) c->losize[is] += 2 * lorbl[ilo + nlomax * is] + 1;
    c->maxaa = std::max(c->maxaa, Msp[is]);
    c->maxnlo = std::max(c->maxnlo, c->losize[is]);
  }
  if (c->maxnlo > 64) return fail(c, EXH_ERANGE, "set_basis: more than 64 LO functions per atom");
  c->atom_species.clear();
  c->Ms.clear();
  c->Koff.clear();
  c->lobase.clear();
  c->colA_h.clear();
  c->lorow_h.clear();
  std::vector<int> kinfo;
  int k = 0, lob = 0;
  for (int is = 0; is < nspecies; is++)
    for (int ia = 0; ia < natoms[is]; ia++) {
      const int ias = (int)c->atom_species.size();
      c->atom_species.push_back(is);
      c->Ms.push_back(Msp[is]);
      c->Koff.push_back(k);
      c->lobase.push_back(lob);
      int f = 0;
      for (int l = 0; l <= lmaxmat; l++)
        for (int m = -l; m <= l; m++) {
          const int lm = l * (l + 1) + m;
          const int p = apword[l + (lmaxapw + 1) * is];
          for (int io = 0; io < p; io++) {
            c->colA_h.push_back((int64_t)ngkmax * (io + (int64_t)apwordmax * (lm + (int64_t)c->lmmax * ias)));
            kinfo.push_back(io | (p << 4) | (l << 8) | ((f - io) << 16));
            f++;
          }
        }
      k += Msp[is];
      for (int ilo = 0; ilo < nlorb[is]; ilo++) {
        const int l = lorbl[ilo + nlomax * is];
        for (int mi = 0; mi <= 2 * l; mi++) c->lorow_h.push_back(LoRow{ias, ilo, mi, l});
      }
      lob += c->losize[is];
    }
  c->Ktot = k;
  c->nlotot = lob;
  c->contig = true;
  for (int kk = 0; kk < c->Ktot; kk++)
    if (c->colA_h[kk] != (int64_t)kk * ngkmax) c->contig = false;
  // column table: [colA (Ktot) | k * ldb (Ktot)] with ldb fixed by ngkmax
  const int64_t ldb = ((int64_t)ngkmax + 63) / 64 * 64;
  std::vector<int64_t> cols(c->colA_h);
  for (int kk = 0; kk < c->Ktot; kk++) cols.push_back((int64_t)kk * ldb);
  std::vector<int> katom;
  katom.insert(katom.end(), c->Koff.begin(), c->Koff.end());
  katom.insert(katom.end(), c->Ms.begin(), c->Ms.end());
  katom.insert(katom.end(), kinfo.begin(), kinfo.end());
  std::vector<int> omax(nspecies, 0);
  for (int is = 0; is < nspecies; is++)
    for (int l = 0; l <= lmaxapw; l++) omax[is] = std::max(omax[is], apword[l + (lmaxapw + 1) * is]);
  std::vector<int> lo;  // [lobase(natmtot), losize, nlorb, lorbl]
  lo.insert(lo.end(), c->lobase.begin(), c->lobase.end());
  lo.insert(lo.end(), c->losize.begin(), c->losize.end());
  lo.insert(lo.end(), c->nlorb.begin(), c->nlorb.end());
  lo.insert(lo.end(), c->lorbl.begin(), c->lorbl.end());
  int rc = upload(c, c->colA, cols.data(), cols.size());
  if (!rc) rc = upload(c, c->katom, katom.data(), katom.size());
  if (!rc) rc = upload(c, c->atom_species_d, c->atom_species.data(), c->atom_species.size());
  if (!rc) rc = upload(c, c->apword_d, c->apword.data(), c->apword.size());
  if (!rc) rc = upload(c, c->omax_d, omax.data(), omax.size());
  if (!rc) rc = upload(c, c->lobase_d, lo.data(), lo.size());
  if (!rc) rc = upload(c, c->lorow, c->lorow_h.data(), c->lorow_h.size());
  if (rc) return rc;
  c->have_basis = true;
  c->basis_epoch++;
  return 0;
}

int exh_set_radial(exh_context *c, double omega, const double *rmt, const int *nrmt,
                   int nrmtmax, const double *spr, const double *apwfr) {
  EXH_CHECK_CTX(c);
  if (int rcw = wait_done(c)) return rcw;
  if (!c->have_basis) return fail(c, EXH_ESTATE, "set_radial: call exh_set_basis first");
  if (!(omega > 0.0) || !rmt || !nrmt || !spr || !apwfr || nrmtmax <= 0)
    return fail(c, EXH_EINVAL, "set_radial: bad argument");
  for (int is = 0; is < c->nspecies; is++)
    if (nrmt[is] > nrmtmax || !(rmt[is] > 0.0)) return fail(c, EXH_EINVAL, "set_radial: nrmt/rmt");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, EXH_ESTATE, "hipSetDevice");
  // match and hmlint refuse to run until every radial table below is rebuilt
  c->have_radial = false;
  c->omega = omega;
  c->rmt.assign(rmt, rmt + c->nspecies);
  c->nrmt_h.assign(nrmt, nrmt + c->nspecies);
  c->nrmtmax = nrmtmax;
  c->spr_h.assign(spr, spr + (size_t)nrmtmax * c->nspecies);
  c->apwfr_h.assign(apwfr, apwfr + (size_t)nrmtmax * 2 * c->apwordmax * (c->lmaxapw + 1) * c->natmtot);
  c->radial_epoch++;
  int rc = upload(c, c->rmt_d, c->rmt.data(), c->rmt.size());
  if (!rc) rc = upload(c, c->apwfr_d, c->apwfr_h.data(), c->apwfr_h.size());
  if (!rc) rc = build_match_tables(c);
  if (rc) return rc;
  c->have_radial = true;
  return 0;
}

int exh_set_mt(exh_context *c, int maxaa, int maxnlo, const exh_z *aa, const exh_z *loa,
               const exh_z *lolo, const double *oalo, const double *ololo, const double *h1aa,
               const double *h1loa, const double *h1lolo) {
  EXH_CHECK_CTX(c);
  if (int rcw = wait_done(c)) return rcw;
  if (!c->have_basis) return fail(c, EXH_ESTATE, "set_mt: call exh_set_basis first");
  if (maxaa != c->maxaa || maxnlo != c->maxnlo)
    return fail(c, EXH_EINVAL, "set_mt: maxaa/maxnlo differ from the basis");
  if (!aa || (c->nlotot > 0 && (!loa || !lolo || !oalo || !ololo)))
    return fail(c, EXH_EINVAL, "set_mt: null table");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, EXH_ESTATE, "hipSetDevice");
  const size_t na = c->natmtot;
  const size_t mnlo = std::max(maxnlo, 1), nlom = std::max(c->nlomax, 1);
  int rc = upload(c, c->aa, (const double2 *)aa, (size_t)maxaa * maxaa * na);
  std::vector<double2> zero;
  if (!rc) {
    if (loa) rc = upload(c, c->loa, (const double2 *)loa, mnlo * maxaa * na);
    else rc = ensure(c, c->loa, 16);
  }
  if (!rc) {
    if (lolo) rc = upload(c, c->lolo, (const double2 *)lolo, mnlo * mnlo * na);
    else rc = ensure(c, c->lolo, 16);
  }
  // s = oalo + h1loa ; ololo_eff = ololo + h1lolo
  std::vector<double> salo((size_t)c->apwordmax * nlom * na, 0.0), olo(nlom * nlom * na, 0.0);
  if (oalo)
    for (size_t i = 0; i < salo.size() && c->nlomax > 0; i++)
      salo[i] = oalo[i] + (h1loa ? h1loa[i] : 0.0);
  if (ololo)
    for (size_t i = 0; i < olo.size() && c->nlomax > 0; i++)
      olo[i] = ololo[i] + (h1lolo ? h1lolo[i] : 0.0);
  if (!rc) rc = upload(c, c->salo, salo.data(), salo.size());
  if (!rc) rc = upload(c, c->ololo_eff, olo.data(), olo.size());
  c->have_h1aa = h1aa != nullptr;
  if (!rc && h1aa)
    rc = upload(c, c->h1aa, h1aa, (size_t)c->apwordmax * c->apwordmax * (c->lmaxapw + 1) * na);
  if (rc) return rc;
  c->have_mt = true;
  return 0;
}

int exh_set_gvectors(exh_context *c, int ngrtot, int ngvec, const int *ivg, const int *intgv,
                     const int *ivgig) {
  EXH_CHECK_CTX(c);
  if (int rcw = wait_done(c)) return rcw;
  if (ngrtot <= 0 || ngvec <= 0 || ngvec > ngrtot || !ivg || !intgv || !ivgig)
    return fail(c, EXH_EINVAL, "set_gvectors: bad argument");
  for (int d = 0; d < 3; d++)
    if (intgv[3 + d] < intgv[d]) return fail(c, EXH_EINVAL, "set_gvectors: empty box");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, EXH_ESTATE, "hipSetDevice");
  const size_t nbox = (size_t)(intgv[3] - intgv[0] + 1) * (intgv[4] - intgv[1] + 1) * (intgv[5] - intgv[2] + 1);
  if (nbox > (size_t)1 << 31) return fail(c, EXH_ERANGE, "set_gvectors: box too large");
  c->have_gvec = c->have_ist = false;
  c->ngrtot = ngrtot;
  c->ngvec = ngvec;
  std::memcpy(c->intgv, intgv, sizeof(int) * 6);
  // Is the G set the whole FFT box, with ivgig the inverse of ivg
  // (gengvec.f90:91-101)?  The interstitial producers need exactly that.
  {
    const int n0 = intgv[3] - intgv[0] + 1, n1 = intgv[4] - intgv[1] + 1;
    bool full = nbox == (size_t)ngrtot;
    for (int ig = 0; ig < ngrtot && full; ig++) {
      const int *v = ivg + 3 * (size_t)ig;
      bool inside = true;
      for (int d = 0; d < 3; d++) inside = inside && v[d] >= intgv[d] && v[d] <= intgv[3 + d];
      full = inside && ivgig[(size_t)(v[0] - intgv[0]) +
                             (size_t)n0 * ((v[1] - intgv[1]) + (size_t)n1 * (v[2] - intgv[2]))] ==
                           ig + 1;
    }
    c->gvec_full = full;
  }
  int rc = upload(c, c->ivg, ivg, (size_t)3 * ngrtot);
  if (!rc) rc = upload(c, c->ivgig, ivgig, nbox);
  if (!rc) rc = ensure(c, c->ist, nbox * sizeof(IstEntry));
  if (rc) return rc;
  c->have_gvec = true;
  return 0;
}

int exh_set_interstitial(exh_context *c, const exh_z *cfunig, const exh_z *veffig,
                         const exh_z *meffig, const exh_z *m2effig, int valence_rel_none,
                         int h1on) {
  EXH_CHECK_CTX(c);
  if (int rcw = wait_done(c)) return rcw;
  if (!c->have_gvec) return fail(c, EXH_ESTATE, "set_interstitial: call exh_set_gvectors first");
  if (!cfunig || !veffig || (!valence_rel_none && !meffig) || (h1on && !m2effig))
    return fail(c, EXH_EINVAL, "set_interstitial: null array");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, EXH_ESTATE, "hipSetDevice");
  int rc = upload(c, c->cf_stage, (const double2 *)cfunig, (size_t)c->ngrtot);
  if (!rc) rc = upload(c, c->vf_stage, (const double2 *)veffig, (size_t)c->ngvec);
  if (!rc && !valence_rel_none) rc = upload(c, c->mf_stage, (const double2 *)meffig, (size_t)c->ngvec);
  if (!rc && h1on) rc = upload(c, c->m2_stage, (const double2 *)m2effig, (size_t)c->ngvec);
  if (rc) return rc;
  hipError_t e = launch_ist_table(c->stream, c, (const double2 *)c->cf_stage.p,
                                  (const double2 *)c->vf_stage.p, (const double2 *)c->mf_stage.p,
                                  (const double2 *)c->m2_stage.p, valence_rel_none ? 1 : 0,
                                  h1on ? 1 : 0);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "set_interstitial");
  c->have_ist = true;
  return 0;
}

int exh_match_device(exh_context *c, void *stream, int ngp, const double *gpc,
                     const double *tpgpc, const exh_z *sfacgp, exh_z *apwalm) {
  EXH_CHECK_CTX(c);
  if (!c->have_basis || !c->have_radial)
    return fail(c, EXH_ESTATE, "match: basis and radial data not set");
  if (ngp < 0 || ngp > c->ngkmax) return fail(c, EXH_EINVAL, "match: ngp outside 0..ngkmax");
  if (ngp > 0 && (!gpc || !tpgpc || !sfacgp || !apwalm))
    return fail(c, EXH_EINVAL, "match: null array");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  if (ngp > 0) {
    const int rc = ensure(c, c->mtab, match_table_bytes(c));
    if (rc) return rc;
  }
  hipError_t e = timed(c, EXH_K_MATCH, s, [&] {
    return launch_match(s, c, ngp, gpc, tpgpc, (const double2 *)sfacgp, (double2 *)apwalm,
                        (double *)c->mtab.p);
  });
  if (e == hipSuccess) e = mark_done(c, s);
  return e == hipSuccess ? 0 : hip_fail(c, e, "match kernel");
}

// Scratch of one single-k setup: the K-major operands (aa.A)^T, ((1+h1)A)^T
// and the packed A^T (Ktot x ldb each, only those the context needs), and the
// G+k -> box index map (ngkmax entries).
struct SetupScratch {
  double2 *bt = nullptr, *a2t = nullptr, *at = nullptr;
  int4 *ivgk = nullptr;
};

// The single-k setup.  o_only: overlapsetup.f90 alone -- O accumulated and
// mirrored, H (may be null) only mirrored (hamiltonoverlapocopy_UL is called
// from overlapsetup.f90:164-166 for both matrices).  check_setup validates the
// arguments; setup_launch enqueues the kernels on scratch the caller provides.
static int check_setup(exh_context *c, int ngp, int rank, const exh_z *apwalm, const int *igpig,
                       const double *vgpc, const exh_z *h, const exh_z *o, int flags,
                       bool o_only) {
  EXH_CHECK_CTX(c);
  if (!c->have_basis || !c->have_mt || !c->have_gvec || !c->have_ist)
    return fail(c, EXH_ESTATE, "setup: basis, muffin-tin tables and interstitial data must be set");
  if (ngp < 0 || ngp > c->ngkmax) return fail(c, EXH_EINVAL, "setup: ngp outside 0..ngkmax");
  if (rank != ngp + c->nlotot) return fail(c, EXH_EINVAL, "setup: rank != ngp + nlotot");
  if ((!h && !o_only) || !o || (ngp > 0 && (!apwalm || !igpig || !vgpc)))
    return fail(c, EXH_EINVAL, "setup: null array");
  const int nmod = (flags >> 8) & 0xff;
  if (flags & ~(EXH_ACCUMULATE | 0xff00)) return fail(c, EXH_ENOTSUP, "setup: unknown flag");
  if (nmod && (nmod < 6 || nmod > kOzMaxMod))
    return fail(c, EXH_EINVAL, "setup: EXH_SPLIT(n) needs 6 <= n <= 12");
  if (nmod && c->have_h1aa)
    return fail(c, EXH_ENOTSUP, "setup: split precision with h1aa (iora*) is not implemented");
  if (nmod && o_only) return fail(c, EXH_ENOTSUP, "overlapsetup: split precision is not implemented");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, EXH_ESTATE, "hipSetDevice");
  return 0;
}

// Enqueue the FP64 kernels (and, with EXH_SPLIT, the split-precision stages,
// whose scratch the caller has prepared with oz_prepare).  `record` brackets
// each launch with the profiling events.
static int setup_launch(exh_context *c, hipStream_t s, int ngp, int rank, const exh_z *apwalm,
                        const int *igpig, const double *vgpc, exh_z *h, exh_z *o, int flags,
                        bool o_only, const SetupScratch &w, bool record) {
  const int nmod = (flags >> 8) & 0xff;
  const int ldb = (c->ngkmax + 63) / 64 * 64;
  exh_z *hw = o_only ? nullptr : h;  // H the kernels write
  auto run = [&](int k, auto &&launch) {
    return record ? timed(c, k, s, launch) : launch();
  };
  const double2 *A = (const double2 *)apwalm;
  hipError_t e = run(EXH_K_IVGK, [&] {
    return launch_ivgk(s, ngp, igpig, (const int *)c->ivg.p, c->ngrtot, w.ivgk);
  });
  if (e == hipSuccess && !c->contig)
    e = run(EXH_K_PACK, [&] { return launch_pack(s, c, ngp, A, w.at); });
  if (e == hipSuccess && !o_only)
    e = run(EXH_K_BT, [&] { return launch_bt(s, c, ngp, A, w.at, w.bt, ldb); });
  if (e == hipSuccess && c->have_h1aa)
    e = run(EXH_K_A2T, [&] { return launch_a2t(s, c, ngp, A, w.a2t, ldb); });
  if (e == hipSuccess && nmod == 0)
    e = run(EXH_K_HO, [&] {
      return launch_ho(s, c, ngp, A, w.at, w.bt, c->have_h1aa ? w.a2t : nullptr, ldb, w.ivgk,
                       vgpc, (double2 *)hw, (double2 *)o, rank, flags);
    });
  if (e == hipSuccess && nmod > 0) {
    OzLayout L = oz_layout(c->ngkmax, c->Ktot, nmod);
    const int bits = oz_bits(nmod, c->Ktot);
    long long ldx;
    const double2 *X = k_major_operand(c, A, w.at, &ldx);
    const int lo0 = c->intgv[0], lo1 = c->intgv[1], lo2 = c->intgv[2];
    IstBox box{w.ivgk, vgpc, (const IstEntry *)c->ist.p, lo0, lo1, lo2,
               c->intgv[3] - lo0 + 1, c->intgv[4] - lo1 + 1, c->intgv[5] - lo2 + 1};
    const int kid[3] = {EXH_K_OZRES, EXH_K_OZGEMM, EXH_K_OZCOMB};
    for (int st = 0; st < 3 && e == hipSuccess; st++)
      e = run(kid[st], [&] {
        return launch_ozaki(s, c, ngp, X, ldx, w.bt, ldb, nmod, bits, L, (int8_t *)c->ozqx.p,
                            (int8_t *)c->ozqb.p, (uint8_t *)c->ozr.p, (int *)c->ozex.p,
                            (int *)c->ozeb.p, box, (double2 *)h, (double2 *)o, rank,
                            flags & EXH_ACCUMULATE, 1 << st);
      });
  }
  if (e == hipSuccess)
    e = run(EXH_K_LOBORDER, [&] {
      return launch_lo_border(s, c, ngp, A, (double2 *)hw, (double2 *)o, rank, flags);
    });
  if (e == hipSuccess)
    e = run(EXH_K_LOLO, [&] { return launch_lolo(s, c, ngp, (double2 *)hw, (double2 *)o, rank, flags); });
  if (e == hipSuccess && o_only && h) e = launch_mirror(s, rank, (double2 *)h, rank);
  if (e == hipSuccess) e = mark_done(c, s);
  return e == hipSuccess ? 0 : hip_fail(c, e, "setup kernels");
}

static int setup_impl(exh_context *c, void *stream, int ngp, int rank, const exh_z *apwalm,
                      const int *igpig, const double *vgpc, exh_z *h, exh_z *o, int flags,
                      bool o_only) {
  int rc = check_setup(c, ngp, rank, apwalm, igpig, vgpc, h, o, flags, o_only);
  if (rc) return rc;
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  const int nmod = (flags >> 8) & 0xff;
  const size_t op = (size_t)std::max(c->Ktot, 1) * ((c->ngkmax + 63) / 64 * 64) * sizeof(double2);
  rc = o_only ? 0 : ensure(c, c->bt, op);
  if (!rc && c->have_h1aa) rc = ensure(c, c->a2t, op);
  if (!rc && !c->contig) rc = ensure(c, c->at, op);
  if (!rc) rc = ensure(c, c->ivgk, (size_t)std::max(c->ngkmax, 1) * sizeof(int4));
  if (!rc && nmod > 0) {
    OzLayout L;
    int bits = 0;
    rc = oz_prepare(c, nmod, L, bits);
  }
  if (rc) return rc;
  SetupScratch w;
  w.bt = (double2 *)c->bt.p;
  w.a2t = (double2 *)c->a2t.p;
  w.at = (double2 *)c->at.p;
  w.ivgk = (int4 *)c->ivgk.p;
  return setup_launch(c, s, ngp, rank, apwalm, igpig, vgpc, h, o, flags, o_only, w, true);
}

int exh_setup_device(exh_context *c, void *stream, int ngp, int rank, const exh_z *apwalm,
                     const int *igpig, const double *vgpc, exh_z *h, exh_z *o, int flags) {
  return setup_impl(c, stream, ngp, rank, apwalm, igpig, vgpc, h, o, flags, false);
}

int exh_overlapsetup_device(exh_context *c, void *stream, int ngp, int rank,
                            const exh_z *apwalm, const int *igpig, const double *vgpc,
                            exh_z *h, exh_z *o, int flags) {
  return setup_impl(c, stream, ngp, rank, apwalm, igpig, vgpc, h, o, flags, true);
}

int exh_setup_device_noalloc(exh_context *c, void *stream, int ngp, int rank,
                             const exh_z *apwalm, const int *igpig, const double *vgpc, exh_z *h,
                             exh_z *o, int flags) {
  int rc = check_setup(c, ngp, rank, apwalm, igpig, vgpc, h, o, flags, false);
  if (rc) return rc;
  if (flags & 0xff00)
    return fail(c, EXH_ENOTSUP,
                "setup_noalloc: split precision needs more scratch than the static buffer holds; "
                "use exh_setup_device");
  if (!c->noalloc_scratch) return fail(c, EXH_ESTATE, "setup_noalloc: no scratch on this device");
  // [ivgk | Bt | A2t (iora*) | packed A^T (non-uniform layouts)], 256-B aligned
  const size_t op = align256((size_t)std::max(c->Ktot, 1) * ((c->ngkmax + 63) / 64 * 64) *
                             sizeof(double2));
  const size_t nivgk = align256((size_t)std::max(c->ngkmax, 1) * sizeof(int4));
  const size_t need = nivgk + op * (1 + (c->have_h1aa ? 1 : 0) + (c->contig ? 0 : 1));
  if (need > kNoallocScratchBytes) {
    char buf[200];
    snprintf(buf, sizeof buf,
             "setup_noalloc: this basis needs %zu bytes of scratch (natmtot*maxaa = %d, ngkmax "
             "= %d); the static buffer holds %zu",
             need, c->Ktot, c->ngkmax, kNoallocScratchBytes);
    return fail(c, EXH_ERANGE, buf);
  }
  unsigned char *base = (unsigned char *)c->noalloc_scratch;
  SetupScratch w;
  w.ivgk = (int4 *)base;
  size_t off = nivgk;
  w.bt = (double2 *)(base + off);
  off += op;
  if (c->have_h1aa) {
    w.a2t = (double2 *)(base + off);
    off += op;
  }
  if (!c->contig) w.at = (double2 *)(base + off);
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  return setup_launch(c, s, ngp, rank, apwalm, igpig, vgpc, h, o, flags, false, w, false);
}

// Host setups: bring rows 1..ngp of the host apwalm into s_apwalm, unless
// reuse is on and s_apwalm already holds this array (same address and ngp
// as the last host match or setup, same basis).
static hipError_t stage_apwalm(exh_context *c, hipStream_t s, int ngp, const exh_z *apwalm) {
  if (ngp <= 0) return hipSuccess;
  if (c->apw_reuse && apwalm == c->apw_host && ngp == c->apw_ngp &&
      c->apw_basis_epoch == c->basis_epoch && c->apw_dev == c->s_apwalm.p) {
    c->apw_reused++;
    return hipSuccess;
  }
  const size_t ld = c->ngkmax, ncol = (size_t)c->apwordmax * c->lmmax * c->natmtot;
  hipError_t e = h2d(c, s, c->s_apwalm.p, ld * sizeof(double2), apwalm, ld * sizeof(double2),
                     ngp * sizeof(double2), ncol);
  if (e == hipSuccess) {
    c->apw_host = apwalm;
    c->apw_dev = c->s_apwalm.p;
    c->apw_ngp = ngp;
    c->apw_basis_epoch = c->basis_epoch;
    c->apw_uploads++;
  } else {
    c->apw_host = nullptr;
  }
  return e;
}

int exh_overlapsetup(exh_context *c, int ngp, int rank, const exh_z *apwalm, const int *igpig,
                     const double *vgpc, exh_z *h, exh_z *o) {
  EXH_CHECK_CTX(c);
  if (!c->have_basis) return fail(c, EXH_ESTATE, "setup: basis not set");
  if (ngp < 0 || ngp > c->ngkmax) return fail(c, EXH_EINVAL, "setup: ngp outside 0..ngkmax");
  if (rank != ngp + c->nlotot) return fail(c, EXH_EINVAL, "setup: rank != ngp + nlotot");
  if (!o || (ngp > 0 && (!apwalm || !igpig || !vgpc)))
    return fail(c, EXH_EINVAL, "setup: null array");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, EXH_ESTATE, "hipSetDevice");
  const size_t ld = c->ngkmax, ncol = (size_t)c->apwordmax * c->lmmax * c->natmtot;
  const size_t nn = (size_t)rank * rank;
  int rc = ensure(c, c->s_apwalm, ld * ncol * sizeof(double2));
  if (!rc) rc = ensure(c, c->s_igpig, ld * sizeof(int));
  if (!rc) rc = ensure(c, c->s_vgpc, 3 * ld * sizeof(double));
  if (!rc && h) rc = ensure(c, c->s_h, std::max<size_t>(nn, 1) * sizeof(double2));
  if (!rc) rc = ensure(c, c->s_o, std::max<size_t>(nn, 1) * sizeof(double2));
  if (rc) return rc;
  hipStream_t s = c->stream;
  hipError_t e = stage_apwalm(c, s, ngp, apwalm);
  if (ngp > 0) {
    if (e == hipSuccess)
      e = hipMemcpyAsync(c->s_igpig.p, igpig, ngp * sizeof(int), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(c->s_vgpc.p, vgpc, 3 * (size_t)ngp * sizeof(double), hipMemcpyHostToDevice, s);
  }
  if (e == hipSuccess) e = h2d(c, s, c->s_o.p, 0, o, 0, nn * sizeof(double2), 1);
  if (e == hipSuccess && h) e = h2d(c, s, c->s_h.p, 0, h, 0, nn * sizeof(double2), 1);
  if (e != hipSuccess) return hip_fail(c, e, "overlapsetup: upload");
  rc = setup_impl(c, s, ngp, rank, (const exh_z *)c->s_apwalm.p, (const int *)c->s_igpig.p,
                  (const double *)c->s_vgpc.p, h ? (exh_z *)c->s_h.p : nullptr,
                  (exh_z *)c->s_o.p, EXH_ACCUMULATE, true);
  if (rc) return rc;
  e = d2h(c, s, o, 0, c->s_o.p, 0, nn * sizeof(double2), 1);
  if (e == hipSuccess && h) e = d2h(c, s, h, 0, c->s_h.p, 0, nn * sizeof(double2), 1);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? 0 : hip_fail(c, e, "overlapsetup: download");
}

int exh_match(exh_context *c, int ngp, const double *gpc, const double *tpgpc,
              const exh_z *sfacgp, exh_z *apwalm) {
  EXH_CHECK_CTX(c);
  if (!c->have_basis || !c->have_radial)
    return fail(c, EXH_ESTATE, "match: basis and radial data not set");
  if (ngp < 0 || ngp > c->ngkmax) return fail(c, EXH_EINVAL, "match: ngp outside 0..ngkmax");
  if (ngp == 0) return 0;
  if (!gpc || !tpgpc || !sfacgp || !apwalm) return fail(c, EXH_EINVAL, "match: null array");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, EXH_ESTATE, "hipSetDevice");
  const size_t ld = c->ngkmax, ncol = (size_t)c->apwordmax * c->lmmax * c->natmtot;
  int rc = ensure(c, c->s_gpc, ld * sizeof(double));
  if (!rc) rc = ensure(c, c->s_tpgpc, 2 * ld * sizeof(double));
  if (!rc) rc = ensure(c, c->s_sfac, ld * c->natmtot * sizeof(double2));
  if (!rc) rc = ensure(c, c->s_apwalm, ld * ncol * sizeof(double2));
  if (!rc) rc = ensure(c, c->mtab, match_table_bytes(c));
  if (rc) return rc;
  hipStream_t s = c->stream;
  c->apw_host = nullptr;  // s_apwalm is rewritten below
  hipError_t e = hipMemcpyAsync(c->s_gpc.p, gpc, ngp * sizeof(double), hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(c->s_tpgpc.p, tpgpc, 2 * (size_t)ngp * sizeof(double), hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = h2d(c, s, c->s_sfac.p, ld * sizeof(double2), sfacgp, ld * sizeof(double2),
            ngp * sizeof(double2), c->natmtot);
  // entries with io >= apword(l) are not written by match: keep the caller's
  if (e == hipSuccess && !c->uniform_apword)
    e = h2d(c, s, c->s_apwalm.p, ld * sizeof(double2), apwalm, ld * sizeof(double2),
            ngp * sizeof(double2), ncol);
  if (e != hipSuccess) return hip_fail(c, e, "match: upload");
  e = timed(c, EXH_K_MATCH, s, [&] {
    return launch_match(s, c, ngp, (const double *)c->s_gpc.p, (const double *)c->s_tpgpc.p,
                        (const double2 *)c->s_sfac.p, (double2 *)c->s_apwalm.p,
                        (double *)c->mtab.p);
  });
  if (e == hipSuccess)
    e = d2h(c, s, apwalm, ld * sizeof(double2), c->s_apwalm.p, ld * sizeof(double2),
            ngp * sizeof(double2), ncol);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    c->apw_host = nullptr;
    return hip_fail(c, e, "match");
  }
  // s_apwalm now equals the caller's apwalm (rows 1..ngp)
  c->apw_host = apwalm;
  c->apw_dev = c->s_apwalm.p;
  c->apw_ngp = ngp;
  c->apw_basis_epoch = c->basis_epoch;
  return 0;
}

int exh_hamiltonandoverlapsetup_ex(exh_context *c, int ngp, int rank, const exh_z *apwalm,
                                   const int *igpig, const double *vgpc, exh_z *h, exh_z *o,
                                   int flags) {
  EXH_CHECK_CTX(c);
  if (!c->have_basis) return fail(c, EXH_ESTATE, "setup: basis not set");
  if (ngp < 0 || ngp > c->ngkmax) return fail(c, EXH_EINVAL, "setup: ngp outside 0..ngkmax");
  if (rank != ngp + c->nlotot) return fail(c, EXH_EINVAL, "setup: rank != ngp + nlotot");
  if (!h || !o || (ngp > 0 && (!apwalm || !igpig || !vgpc)))
    return fail(c, EXH_EINVAL, "setup: null array");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, EXH_ESTATE, "hipSetDevice");
  const size_t ld = c->ngkmax, ncol = (size_t)c->apwordmax * c->lmmax * c->natmtot;
  const size_t nn = (size_t)rank * rank;
  int rc = ensure(c, c->s_apwalm, ld * ncol * sizeof(double2));
  if (!rc) rc = ensure(c, c->s_igpig, ld * sizeof(int));
  if (!rc) rc = ensure(c, c->s_vgpc, 3 * ld * sizeof(double));
  if (!rc) rc = ensure(c, c->s_h, std::max<size_t>(nn, 1) * sizeof(double2));
  if (!rc) rc = ensure(c, c->s_o, std::max<size_t>(nn, 1) * sizeof(double2));
  if (rc) return rc;
  hipStream_t s = c->stream;
  hipError_t e = stage_apwalm(c, s, ngp, apwalm);
  if (ngp > 0) {
    if (e == hipSuccess)
      e = hipMemcpyAsync(c->s_igpig.p, igpig, ngp * sizeof(int), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(c->s_vgpc.p, vgpc, 3 * (size_t)ngp * sizeof(double), hipMemcpyHostToDevice, s);
  }
  if (e == hipSuccess && (flags & EXH_ACCUMULATE)) {
    e = h2d(c, s, c->s_h.p, 0, h, 0, nn * sizeof(double2), 1);
    if (e == hipSuccess) e = h2d(c, s, c->s_o.p, 0, o, 0, nn * sizeof(double2), 1);
  }
  if (e != hipSuccess) return hip_fail(c, e, "setup: upload");
  rc = exh_setup_device(c, s, ngp, rank, (const exh_z *)c->s_apwalm.p, (const int *)c->s_igpig.p,
                        (const double *)c->s_vgpc.p, (exh_z *)c->s_h.p, (exh_z *)c->s_o.p, flags);
  if (rc) return rc;
  e = d2h(c, s, h, 0, c->s_h.p, 0, nn * sizeof(double2), 1);
  if (e == hipSuccess) e = d2h(c, s, o, 0, c->s_o.p, 0, nn * sizeof(double2), 1);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? 0 : hip_fail(c, e, "setup: download");
}

int exh_set_apwalm_reuse(exh_context *c, int on) {
  EXH_CHECK_CTX(c);
  c->apw_reuse = on != 0;
  return 0;
}

int exh_apwalm_transfers(exh_context *c, long long *uploads, long long *reused) {
  EXH_CHECK_CTX(c);
  if (uploads) *uploads = c->apw_uploads;
  if (reused) *reused = c->apw_reused;
  return 0;
}

int exh_profile_enable(exh_context *c, int on) {
  EXH_CHECK_CTX(c);
  c->prof = on != 0;
  return 0;
}

int exh_profile_read(exh_context *c, int kernel, double *total_ms, int *launches) {
  EXH_CHECK_CTX(c);
  if (kernel < 0 || kernel >= EXH_K_COUNT) return fail(c, EXH_EINVAL, "profile_read: kernel id");
  int rc = drain_records(c);
  if (rc) return rc;
  if (total_ms) *total_ms = c->prof_ms[kernel];
  if (launches) *launches = c->prof_n[kernel];
  return 0;
}

int exh_profile_reset(exh_context *c) {
  EXH_CHECK_CTX(c);
  int rc = drain_records(c);
  for (int k = 0; k < EXH_K_COUNT; k++) {
    c->prof_ms[k] = 0.0;
    c->prof_n[k] = 0;
  }
  return rc;
}

int exh_hamiltonandoverlapsetup(exh_context *c, int ngp, int rank, const exh_z *apwalm,
                                const int *igpig, const double *vgpc, exh_z *h, exh_z *o) {
  return exh_hamiltonandoverlapsetup_ex(c, ngp, rank, apwalm, igpig, vgpc, h, o, EXH_ACCUMULATE);
}

}  // extern "C"
