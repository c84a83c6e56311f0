// Internal declarations shared by the HIP translation units of libexh.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/exh.h"

namespace exh {

// Device buffer owned by the context; grows, never shrinks.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

// Interstitial lookup entry for one point of the ivgig box:
//   H(i,j) += veff + t1 * kin,  O(i,j) += ocf + t1 * okin,
//   t1 = 0.5 * (G_i+k).(G_j+k).
// 'none' relativity: kin = ocf = cfunig, okin = 0 (hmlistln.f90:73-84,
// olpistln.f90:69-82).  Non-'none': kin = meffig (hmlistln.f90:61-71).
// h1on: okin = alpha^2/2 * m2effig (olpistln.f90:53-66).  Points whose G
// index is 0 or beyond ngvec hold zeros (the reference skips them).
struct IstEntry {
  double2 veff, kin, ocf, okin;
};

// Per LO row of the lo-lo block: owning atom, local orbital, m offset, l.
struct LoRow {
  int ias, ilo, mi, l;
};

// Split-precision (Ozaki scheme II) path.
constexpr int kOzMaxMod = 12;
extern const int kOzPrimes[kOzMaxMod];
struct OzLayout {
  int Gpad = 0, Kpad = 0;       // G rounded to the 128 tile, K to the 64 stage
  size_t qstride = 0;           // bytes of one residue array (Kpad * Gpad)
  long long max_pairs = 0;      // upper tile pairs at ngkmax
  size_t qx_bytes = 0, qb_bytes = 0, r_bytes = 0;
  size_t e_bytes = 0, e_part_off = 0;  // per operand: column exponents + partial norms
};

// APW-lo border tables (hamiltonsetup.f90:124-151, overlapsetup.f90:116-133)
// of one structure, device pointers.
struct LoTables {
  const long long *colA;
  const int *Koff, *Ms, *atom_species, *losize, *lobase;
  const double2 *loa;
  int maxnlo, maxaa;
  const int *nlorb, *lorbl;
  int nlomax;
  const int *apword;
  int lmaxapw, pmax, lmmax, ngkmax;
  const double *salo;
};

// Per-structure descriptor of a grouped launch, in device memory: the
// device addresses and scalars of one context's tables that the grouped
// kernels read, so that one launch can cover k-points of several structures.
struct SDesc {
  // basis and the stacked K index (atom, l, m, io)
  int natmtot, nspecies, Ktot, maxaa, lmaxapw, pmax, lmmax, ngkmax, have_h1aa;
  long long ldx;  // leading dimension of the K-major A operand
  int ldb;        // leading dimension of the batch scratch operands
  const long long *colA;
  const int *Koff, *Ms, *kinfo;
  const double2 *aa;
  const double *h1aa;
  // match
  const int *atom_species, *apword, *omax, *piv;
  const double *rmt, *lu;
  double fpso;  // 4 pi / sqrt(omega)
  // G vectors and the interstitial box
  const int *ivg;
  int ngrtot;
  const IstEntry *ist;
  int lo0, lo1, lo2, n0, n1, n2;
  // local orbitals
  LoTables lo;
  int nlotot, maxnlo, nlomax;
  const LoRow *lorow;
  const int *lobase;
  const double2 *lolo;
  const double *ololo_eff;
};

// Per-k-point descriptor of a grouped (batched) launch, in device memory.
// Match inputs, apwalm, the G+k maps and H/O are the caller's; x, at, bt,
// a2t and ivgk point into the owning context's batch scratch; sd is the
// owning structure's descriptor.
struct KDesc {
  int ngp, ldh;
  const double *gpc, *tpgpc;
  const double2 *sfac;
  double2 *apwalm;
  const int *igpig;
  const double *vgpc;
  double2 *h, *o;
  const double2 *x;  // K-major A: apwalm itself (equally spaced columns) or at
  double2 *at, *bt, *a2t;
  int4 *ivgk;
  double *mtab;  // match: radial and harmonics tables of this k-point (batch scratch)
  const SDesc *sd;
};

// Extent of one group of a grouped launch: the largest sizes over its
// k-points and structures, which set the launch grids (blocks beyond a
// k-point's own sizes exit).
struct GroupDims {
  int nk = 0, ngpmax = 0, natmax = 0, nspmax = 0, maxaa = 0, Kmax = 0, nlotot = 0, maxnlo = 0;
  unsigned lmax_mask = 0;  // bit l: some structure has lmaxapw = l
  bool any_pack = false, any_a2t = false;
};

// Tile shapes of the grouped H/O kernel: output TMI x TMJ per workgroup
// (4 or 8 waves of 32 x 32), K slabs of TK staged in LDS.
struct HoShapeInfo {
  int tmi, tmj, tk;
};
constexpr int kHoShapes = 5;
extern const HoShapeInfo kHoShapeInfo[kHoShapes];

struct MatchTables {
  // LU factors of the boundary-derivative matrix D per (l, ias): pmax*pmax
  // doubles (column-major, zgetf2 layout) and pmax 1-based pivots.
  DevBuf lu, piv;
};

}  // namespace exh

struct exh_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  // state flags
  bool have_basis = false, have_radial = false, have_mt = false;
  bool have_gvec = false, have_ist = false;

  // basis (host)
  int nspecies = 0, natmtot = 0, lmaxapw = 0, lmaxmat = 0, apwordmax = 0;
  int nlomax = 0, ngkmax = 0, maxaa = 0, maxnlo = 0, nlotot = 0, lmmax = 0;
  bool uniform_apword = true;
  std::vector<int> natoms, apword, nlorb, lorbl;
  std::vector<int> atom_species, losize, lobase;  // per species / per atom
  std::vector<int> Ms, Koff;                     // per atom: K width, K offset
  int Ktot = 0;
  std::vector<int64_t> colA_h;                   // stacked K -> apwalm column
  std::vector<exh::LoRow> lorow_h;

  // device copies of basis tables
  exh::DevBuf colA, katom, atom_species_d, apword_d, omax_d, lorow, lobase_d;

  // match
  double omega = 0.0;
  std::vector<double> rmt;
  std::vector<int> nrmt_h;
  int nrmtmax = 0;
  std::vector<double> spr_h, apwfr_h;
  exh::DevBuf rmt_d;
  exh::MatchTables mt_match;

  // muffin-tin tables
  exh::DevBuf aa, loa, lolo, salo, ololo_eff, h1aa;
  bool have_h1aa = false;

  // G vectors / interstitial
  int ngrtot = 0, ngvec = 0;
  int intgv[6] = {0, 0, 0, 0, 0, 0};
  exh::DevBuf ivg, ivgig, ist;  // ist: IstEntry per box point
  exh::DevBuf cf_stage, vf_stage, mf_stage, m2_stage;
  // interstitial producers: FFT work arrays (up to 3 grids), geometry,
  // host-API staging, and hipFFT plans per batch count 1..3 (opaque here)
  exh::DevBuf ifz, igeo, i_in0, i_in1, i_out0, i_out1, i_out2;
  void *fft_plan[4] = {nullptr, nullptr, nullptr, nullptr};
  int fft_grid[3] = {0, 0, 0};
  bool gvec_full = false;  // G set = the whole FFT box, ivgig inverse of ivg
  // muffin-tin table producer (hmlint): apwfr on the device, species and
  // Gaunt tables, spline weights, radial integrals, host-API staging
  exh::DevBuf apwfr_d, hml_fl, hml_arow, hml_lrow, hml_count, hml_gstart, hml_glm2, hml_gval;
  exh::DevBuf hml_w, hml_spr, hml_ipot, hml_ikin, hml_vin, hml_loin, hml_aa, hml_loa, hml_lolo;
  int hml_lmaxvr = -1;
  exh::DevBuf so_dt, so_rowb, so_pairs, so_pstart, so_host;   // mt_so
  long long so_radial_epoch = -1, so_basis_epoch = -1;
  long long radial_epoch = 0, basis_epoch = 0, hml_radial_epoch = -1, hml_basis_epoch = -1;

  // instrumentation: (kernel id, start, stop) per recorded launch
  bool prof = false;
  struct Rec {
    int k;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> evpool;
  double prof_ms[EXH_K_COUNT] = {0};
  int prof_n[EXH_K_COUNT] = {0};

  // batched launches: plan (-1 = default), scratch cap, per-k scratch, and
  // cached descriptor/tile uploads and graphs keyed by the call's arguments
  int ncu = 256;
  int batch_plan = -1;
  size_t batch_ws_limit = size_t(8) << 30;
  uint64_t epoch = 0;  // bumped by every state change and scratch reallocation
  bool capturing = false;
  // completion of the last hot-path work queued on any stream; the setters
  // wait for it before rewriting tables that work may still read
  hipEvent_t done_ev = nullptr;
  bool done_pending = false;
  exh::DevBuf bbt, bat, ba2t, bivgk, bmtab;
  struct BatchEntry;
  std::vector<BatchEntry *> bcache;
  uint64_t bclock = 0;

  // per-call scratch (mtab: the match kernel's radial and harmonics tables)
  exh::DevBuf bt, a2t, at, ivgk, mtab;
  // this device's copy of the static scratch of exh_setup_device_noalloc
  void *noalloc_scratch = nullptr;
  exh::DevBuf ozqx, ozqb, ozr, ozex, ozeb;
  bool oz_consts = false;
  bool contig = true;  // colA[k] == k * ngkmax
  // host-API staging
  exh::DevBuf s_apwalm, s_igpig, s_vgpc, s_h, s_o, s_gpc, s_tpgpc, s_sfac;
  // s_apwalm holds the host array at apw_host (rows 1..apw_ngp) as of the
  // last host match or setup call; with apw_reuse the host setups then skip
  // uploading it again (exh_set_apwalm_reuse)
  bool apw_reuse = false;
  const void *apw_host = nullptr, *apw_dev = nullptr;
  int apw_ngp = -1;
  long long apw_basis_epoch = -1;
  long long apw_uploads = 0, apw_reused = 0;
  // pinned staging of the host-array entry points: two buffers of
  // pin_bytes, each with the event of the last DMA that used it
  void *pin[2] = {nullptr, nullptr};
  hipEvent_t pin_ev[2] = {nullptr, nullptr};
  size_t pin_bytes = 0;
};

namespace exh {

int fail(exh_context *c, int code, const std::string &msg);
int hip_fail(exh_context *c, hipError_t e, const char *where);
int ensure(exh_context *c, DevBuf &b, size_t bytes);
void release(DevBuf &b);
hipEvent_t take_event(exh_context *c);
int drain_records(exh_context *c);

// Enqueue a launch, bracketed by profiling events when profiling is on (not
// while the stream is being captured into a graph).
template <typename F>
hipError_t timed(exh_context *c, int k, hipStream_t s, F &&launch) {
  if (!c->prof || c->capturing) return launch();
  hipEvent_t a = take_event(c), b = take_event(c);
  if (!a || !b) return hipErrorOutOfMemory;
  hipError_t e = hipEventRecord(a, s);
  if (e == hipSuccess) e = launch();
  if (e == hipSuccess) e = hipEventRecord(b, s);
  c->recs.push_back({k, a, b});
  return e;
}
void batch_release_all(exh_context *c);
void interstitial_release(exh_context *c);
hipError_t mark_done(exh_context *c, hipStream_t s);
int wait_done(exh_context *c);

// Kernel launchers (setup_kernels.hip / match_kernel.hip).
hipError_t launch_ivgk(hipStream_t s, int ngp, const int *igpig, const int *ivg,
                       int ngrtot, int4 *ivgk);
hipError_t launch_ist_table(hipStream_t s, const exh_context *c,
                            const double2 *cf, const double2 *vf,
                            const double2 *mf, const double2 *m2, int relnone,
                            int h1on);
// `at` is the Ktot x ldb K-major copy of apwalm that launch_pack writes and
// the others read when the context's stacked columns are not equally spaced
// (unused otherwise).
hipError_t launch_pack(hipStream_t s, const exh_context *c, int ngp,
                       const double2 *apwalm, double2 *at);
hipError_t launch_bt(hipStream_t s, const exh_context *c, int ngp,
                     const double2 *apwalm, const double2 *at, double2 *bt, int ldb);
hipError_t launch_a2t(hipStream_t s, const exh_context *c, int ngp,
                      const double2 *apwalm, double2 *a2t, int ldb);
hipError_t launch_ho(hipStream_t s, const exh_context *c, int ngp,
                     const double2 *apwalm, const double2 *at, const double2 *bt,
                     const double2 *a2t, int ldb, const int4 *ivgk,
                     const double *vgpc, double2 *h, double2 *o, int ldh,
                     int flags);
hipError_t launch_lo_border(hipStream_t s, const exh_context *c, int ngp,
                            const double2 *apwalm, double2 *h, double2 *o,
                            int ldh, int flags);
hipError_t launch_mirror(hipStream_t s, int n, double2 *m, int ld);
hipError_t launch_lolo(hipStream_t s, const exh_context *c, int ngp,
                       double2 *h, double2 *o, int ldh, int flags);
const double2 *k_major_operand(const exh_context *c, const double2 *apwalm, const double2 *at,
                               long long *ld);
int oz_bits(int nmod, int Ktot);
hipError_t oz_upload_constants(int nmod);
OzLayout oz_layout(int ngkmax, int Ktot, int nmod);
int oz_prepare(exh_context *c, int nmod, OzLayout &L, int &bits);
struct IstBox;
hipError_t launch_ozaki(hipStream_t s, const exh_context *c, int ngp, const double2 *X,
                        long long ldx, const double2 *B, long long ldb, int nmod, int bits,
                        const OzLayout &L, int8_t *QX, int8_t *QB, uint8_t *R, int *ex,
                        int *eb, const IstBox &box, double2 *h, double2 *o, int ldh,
                        int accumulate, int stage_mask);
// tab: match_table_bytes(c) of scratch for the k-point's radial factors and
// spherical harmonics
size_t match_table_bytes(const exh_context *c);
hipError_t launch_match(hipStream_t s, const exh_context *c, int ngp,
                        const double *gpc, const double *tpgpc,
                        const double2 *sfacgp, double2 *apwalm, double *tab);

// Grouped launchers: one launch covers the g.nk k-points described by kd
// (device), of one or several structures.
SDesc make_sdesc(const exh_context *c);
hipError_t launch_match_grouped(hipStream_t s, const GroupDims &g, const KDesc *kd);
hipError_t launch_ivgk_grouped(hipStream_t s, const GroupDims &g, const KDesc *kd);
hipError_t launch_pack_grouped(hipStream_t s, const GroupDims &g, const KDesc *kd);
hipError_t launch_bt_grouped(hipStream_t s, const GroupDims &g, const KDesc *kd);
hipError_t launch_a2t_grouped(hipStream_t s, const GroupDims &g, const KDesc *kd);
hipError_t launch_ho_grouped(hipStream_t s, int shape, const KDesc *kd, const int4 *tiles,
                             int ntiles, int nwg, int accumulate);
int ho_grouped_occupancy(int shape);
hipError_t launch_lo_border_grouped(hipStream_t s, const GroupDims &g, const KDesc *kd,
                                    int accumulate);
hipError_t launch_lolo_grouped(hipStream_t s, const GroupDims &g, const KDesc *kd,
                               int accumulate);

}  // namespace exh
