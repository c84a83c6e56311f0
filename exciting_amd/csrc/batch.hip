// Batched k-points: every stage of match + H/O setup as one grouped launch
// over a list of k-points of one structure, the H/O contraction as one
// persistent launch over all their upper tiles, a plan table with an
// autotuner, and a cache of uploaded descriptors and captured HIP graphs.
//
// The reference runs match and hamiltonandoverlapsetup once per k-point in
// the SCF k loop (scf_cycle.f90:256-296, seceqn.f90:57, seceqnfv.f90:76-77);
// per k-point the batch computes exactly the same numbers as those calls.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "exh_device.h"

using exh::DevBuf;
using exh::KDesc;

struct exh_context::BatchEntry {
  std::vector<unsigned char> key;
  std::vector<KDesc> kd;    // host copy (kept alive: source of the upload)
  std::vector<int4> tiles;  // every group's tile list, concatenated
  struct Group {
    int k0, nk, ngpmax, t0, nt, nwg;
  };
  std::vector<Group> groups;
  DevBuf kd_d, tiles_d;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t use = 0;
};

namespace exh {
namespace {

constexpr int kPlans = kHoShapes * 4;  // shape x {tile grid, persistent} x {k-major, XCD}
constexpr int kDefaultPlan = 1;        // 64x64x16, one workgroup per tile, XCD order
constexpr size_t kMaxEntries = 8;

struct PlanInfo {
  int shape, persistent, xcd;
};
PlanInfo plan_info(int p) { return {p / 4, (p / 2) & 1, p & 1}; }

inline long long cdivl(long long a, long long b) { return (a + b - 1) / b; }

void free_entry(exh_context::BatchEntry *E) {
  if (E->exec) (void)hipGraphExecDestroy(E->exec);
  if (E->graph) (void)hipGraphDestroy(E->graph);
  release(E->kd_d);
  release(E->tiles_d);
  delete E;
}

// Upper tiles (I0 <= J0 + TMJ - 1) of the k-points [k0, k0 + nk), k-major
// then J-major, k numbered from 0 within the group.  With xcd the list is
// split into 8 contiguous ranges and interleaved so that list position t
// (which runs on XCD t % 8) walks range t % 8 in order; k < 0 pads.
void build_tiles(const exh_kpoint *kp, int k0, int nk, int tmi, int tmj, bool xcd,
                 std::vector<int4> &out) {
  std::vector<int4> t;
  for (int kl = 0; kl < nk; kl++) {
    const int ngp = kp[k0 + kl].ngp;
    if (ngp <= 0) continue;
    const int nI = (int)cdivl(ngp, tmi), nJ = (int)cdivl(ngp, tmj);
    for (int J = 0; J < nJ; J++)
      for (int I = 0; I < nI && I * tmi <= J * tmj + tmj - 1; I++)
        t.push_back(make_int4(kl, I * tmi, J * tmj, 0));
  }
  if (!xcd) {
    out.insert(out.end(), t.begin(), t.end());
    return;
  }
  const size_t n = t.size(), q = (n + 7) / 8;
  const size_t base = out.size();
  out.resize(base + 8 * q, make_int4(-1, 0, 0, 0));
  for (int x = 0; x < 8; x++)
    for (size_t i = 0; i < q; i++) {
      const size_t src = x * q + i;
      if (src < n) out[base + 8 * i + x] = t[src];
    }
}

// Everything a recorded batch bakes in besides the k-point list: device
// addresses of the tables and scratch, and every scalar passed to a kernel
// or deciding which kernels run.  Tables rewritten in place by exh_set_* keep
// the signature, so a captured graph stays valid across SCF iterations.
uint64_t state_signature(const exh_context *c) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](const void *p, size_t n) {
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ull;
  };
  const DevBuf *bufs[] = {&c->colA, &c->katom, &c->atom_species_d, &c->apword_d, &c->omax_d,
                          &c->lorow, &c->lobase_d, &c->rmt_d, &c->mt_match.lu,
                          &c->mt_match.piv, &c->aa, &c->loa, &c->lolo, &c->salo,
                          &c->ololo_eff, &c->h1aa, &c->ivg, &c->ist, &c->bbt, &c->bat,
                          &c->ba2t, &c->bivgk, &c->ozqx, &c->ozqb, &c->ozr, &c->ozex,
                          &c->ozeb};
  for (const DevBuf *b : bufs) mix(&b->p, sizeof b->p);
  const int sc[] = {c->nspecies, c->natmtot, c->lmaxapw, c->lmaxmat, c->apwordmax, c->nlomax,
                    c->ngkmax, c->maxaa, c->maxnlo, c->nlotot, c->lmmax, c->Ktot,
                    c->ngrtot, c->ngvec, c->intgv[0], c->intgv[1], c->intgv[2], c->intgv[3],
                    c->intgv[4], c->intgv[5], c->contig ? 1 : 0, c->have_h1aa ? 1 : 0,
                    c->uniform_apword ? 1 : 0, c->ncu};
  mix(sc, sizeof sc);
  mix(&c->omega, sizeof c->omega);
  return h;
}

int validate(exh_context *c, int nk, const exh_kpoint *kp, int flags) {
  if (!c->have_basis || !c->have_mt || !c->have_gvec || !c->have_ist)
    return fail(c, EXH_ESTATE, "batch: basis, muffin-tin tables and interstitial data must be set");
  if ((flags & EXH_MATCH) && !c->have_radial)
    return fail(c, EXH_ESTATE, "batch: EXH_MATCH needs the radial data");
  if (nk < 0 || (nk > 0 && !kp)) return fail(c, EXH_EINVAL, "batch: bad k-point list");
  if (flags & ~(EXH_ACCUMULATE | EXH_MATCH | EXH_GRAPH | 0xff00))
    return fail(c, EXH_ENOTSUP, "batch: unknown flag");
  const int nmod = (flags >> 8) & 0xff;
  if (nmod && (nmod < 6 || nmod > kOzMaxMod))
    return fail(c, EXH_EINVAL, "batch: EXH_SPLIT(n) needs 6 <= n <= 12");
  if (nmod && c->have_h1aa)
    return fail(c, EXH_ENOTSUP, "batch: split precision with h1aa (iora*) is not implemented");
  for (int i = 0; i < nk; i++) {
    const exh_kpoint &k = kp[i];
    if (k.ngp < 0 || k.ngp > c->ngkmax) return fail(c, EXH_EINVAL, "batch: ngp outside 0..ngkmax");
    if (k.rank != k.ngp + c->nlotot) return fail(c, EXH_EINVAL, "batch: rank != ngp + nlotot");
    if (!k.h || !k.o) return fail(c, EXH_EINVAL, "batch: null h or o");
    if (k.ngp > 0 && (!k.apwalm || !k.igpig || !k.vgpc))
      return fail(c, EXH_EINVAL, "batch: null apwalm, igpig or vgpc");
    if ((flags & EXH_MATCH) && k.ngp > 0 && (!k.gpc || !k.tpgpc || !k.sfacgp))
      return fail(c, EXH_EINVAL, "batch: EXH_MATCH with null gpc, tpgpc or sfacgp");
  }
  return 0;
}

// Per-k-point scratch bytes and the group size that fits the workspace limit.
size_t perk_bytes(const exh_context *c, int ldb) {
  const size_t op = (size_t)std::max(c->Ktot, 1) * ldb * sizeof(double2);
  return op * (1 + (c->contig ? 0 : 1) + (c->have_h1aa ? 1 : 0)) +
         (size_t)std::max(c->ngkmax, 1) * sizeof(int4);
}

int ensure_scratch(exh_context *c, int nkg, int ldb) {
  const size_t op = (size_t)std::max(c->Ktot, 1) * ldb * sizeof(double2);
  int rc = ensure(c, c->bbt, op * nkg);
  if (!rc && !c->contig) rc = ensure(c, c->bat, op * nkg);
  if (!rc && c->have_h1aa) rc = ensure(c, c->ba2t, op * nkg);
  if (!rc) rc = ensure(c, c->bivgk, (size_t)std::max(c->ngkmax, 1) * sizeof(int4) * nkg);
  return rc;
}

bool scratch_fits(const exh_context *c, int nkg, int ldb) {
  const size_t op = (size_t)std::max(c->Ktot, 1) * ldb * sizeof(double2);
  return c->bbt.p && c->bbt.bytes >= op * nkg && (c->contig || c->bat.bytes >= op * nkg) &&
         (!c->have_h1aa || c->ba2t.bytes >= op * nkg) &&
         c->bivgk.bytes >= (size_t)std::max(c->ngkmax, 1) * sizeof(int4) * nkg;
}

exh_context::BatchEntry *find_entry(exh_context *c, const std::vector<unsigned char> &key) {
  for (auto *E : c->bcache)
    if (E->key == key) return E;
  return nullptr;
}

exh_context::BatchEntry *build_entry(exh_context *c, hipStream_t s, int nk, const exh_kpoint *kp,
                                     int nkg, int ldb, int plan, std::vector<unsigned char> key,
                                     int *rc) {
  *rc = 0;
  if (c->bcache.size() >= kMaxEntries) {
    // the oldest entry may still be read by queued work on any stream
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      *rc = hip_fail(c, e, "batch: cache eviction");
      return nullptr;
    }
    auto it = std::min_element(c->bcache.begin(), c->bcache.end(),
                               [](auto *a, auto *b) { return a->use < b->use; });
    free_entry(*it);
    c->bcache.erase(it);
  }
  auto *E = new exh_context::BatchEntry();
  E->key = std::move(key);
  const PlanInfo P = plan_info(plan);
  const HoShapeInfo &S = kHoShapeInfo[P.shape];
  const int occ = P.persistent ? ho_grouped_occupancy(P.shape) : 0;
  const size_t op = (size_t)std::max(c->Ktot, 1) * ldb;
  E->kd.resize(nk);
  for (int i = 0; i < nk; i++) {
    const exh_kpoint &k = kp[i];
    const int slot = i % nkg;
    KDesc &d = E->kd[i];
    d.ngp = k.ngp;
    d.ldh = k.rank;
    d.gpc = k.gpc;
    d.tpgpc = k.tpgpc;
    d.sfac = (const double2 *)k.sfacgp;
    d.apwalm = (double2 *)k.apwalm;
    d.igpig = k.igpig;
    d.vgpc = k.vgpc;
    d.h = (double2 *)k.h;
    d.o = (double2 *)k.o;
    d.bt = (double2 *)c->bbt.p + slot * op;
    d.at = c->contig ? nullptr : (double2 *)c->bat.p + slot * op;
    d.a2t = c->have_h1aa ? (double2 *)c->ba2t.p + slot * op : nullptr;
    d.ivgk = (int4 *)c->bivgk.p + (size_t)slot * std::max(c->ngkmax, 1);
    d.x = c->contig ? (const double2 *)d.apwalm : d.at;
  }
  for (int k0 = 0; k0 < nk; k0 += nkg) {
    exh_context::BatchEntry::Group g;
    g.k0 = k0;
    g.nk = std::min(nkg, nk - k0);
    g.ngpmax = 0;
    for (int i = k0; i < k0 + g.nk; i++) g.ngpmax = std::max(g.ngpmax, kp[i].ngp);
    g.t0 = (int)E->tiles.size();
    build_tiles(kp, k0, g.nk, S.tmi, S.tmj, P.xcd != 0, E->tiles);
    g.nt = (int)E->tiles.size() - g.t0;
    long long nwg = g.nt;
    if (P.persistent) {
      nwg = std::min<long long>(nwg, (long long)occ * c->ncu);
      if (P.xcd) nwg = std::max<long long>(8, nwg / 8 * 8);
    }
    g.nwg = (int)std::max<long long>(nwg, 0);
    E->groups.push_back(g);
  }
  *rc = ensure(nullptr, E->kd_d, std::max<size_t>(E->kd.size(), 1) * sizeof(KDesc));
  if (!*rc) *rc = ensure(nullptr, E->tiles_d, std::max<size_t>(E->tiles.size(), 1) * sizeof(int4));
  hipError_t e = hipSuccess;
  if (!*rc && nk > 0)
    e = hipMemcpyAsync(E->kd_d.p, E->kd.data(), E->kd.size() * sizeof(KDesc),
                       hipMemcpyHostToDevice, s);
  if (!*rc && e == hipSuccess && !E->tiles.empty())
    e = hipMemcpyAsync(E->tiles_d.p, E->tiles.data(), E->tiles.size() * sizeof(int4),
                       hipMemcpyHostToDevice, s);
  if (*rc || e != hipSuccess) {
    if (!*rc) *rc = hip_fail(c, e, "batch: descriptor upload");
    else fail(c, *rc, "batch: descriptor buffers");
    free_entry(E);
    return nullptr;
  }
  c->bcache.push_back(E);
  return E;
}

hipError_t enqueue(exh_context *c, hipStream_t s, const exh_context::BatchEntry &E, int flags,
                   int plan, int ldb, const OzLayout &L, int bits) {
  const int nmod = (flags >> 8) & 0xff, acc = flags & EXH_ACCUMULATE;
  const long long ldx = c->contig ? c->ngkmax : ldb;
  const PlanInfo P = plan_info(plan);
  const KDesc *kd0 = (const KDesc *)E.kd_d.p;
  const int4 *tiles = (const int4 *)E.tiles_d.p;
  const int lo0 = c->intgv[0], lo1 = c->intgv[1], lo2 = c->intgv[2];
  hipError_t e = hipSuccess;
  for (const auto &g : E.groups) {
    const KDesc *kd = kd0 + g.k0;
    if (e == hipSuccess && (flags & EXH_MATCH))
      e = timed(c, EXH_K_MATCH, s, [&] { return launch_match_grouped(s, c, kd, g.nk, g.ngpmax); });
    if (e == hipSuccess)
      e = timed(c, EXH_K_IVGK, s, [&] { return launch_ivgk_grouped(s, c, kd, g.nk, g.ngpmax); });
    if (e == hipSuccess && !c->contig)
      e = timed(c, EXH_K_PACK, s,
                [&] { return launch_pack_grouped(s, c, kd, g.nk, g.ngpmax, ldb); });
    if (e == hipSuccess)
      e = timed(c, EXH_K_BT, s,
                [&] { return launch_bt_grouped(s, c, kd, g.nk, g.ngpmax, ldx, ldb); });
    if (e == hipSuccess && c->have_h1aa)
      e = timed(c, EXH_K_A2T, s,
                [&] { return launch_a2t_grouped(s, c, kd, g.nk, g.ngpmax, ldb); });
    if (e == hipSuccess && nmod == 0)
      e = timed(c, EXH_K_HOGRP, s, [&] {
        return launch_ho_grouped(s, c, P.shape, kd, tiles + g.t0, g.nt, g.nwg, ldx, ldb, acc);
      });
    if (nmod > 0)
      for (int i = g.k0; i < g.k0 + g.nk && e == hipSuccess; i++) {
        const KDesc &d = E.kd[i];
        IstBox box{d.ivgk, d.vgpc, (const IstEntry *)c->ist.p, lo0, lo1, lo2,
                   c->intgv[3] - lo0 + 1, c->intgv[4] - lo1 + 1, c->intgv[5] - lo2 + 1};
        const int kid[3] = {EXH_K_OZRES, EXH_K_OZGEMM, EXH_K_OZCOMB};
        for (int st = 0; st < 3 && e == hipSuccess; st++)
          e = timed(c, kid[st], s, [&] {
            return launch_ozaki(s, c, d.ngp, d.x, ldx, d.bt, ldb, nmod, bits, L,
                                (int8_t *)c->ozqx.p, (int8_t *)c->ozqb.p, (uint8_t *)c->ozr.p,
                                (int *)c->ozex.p, (int *)c->ozeb.p, box, d.h, d.o, d.ldh, acc,
                                1 << st);
          });
      }
    if (e == hipSuccess)
      e = timed(c, EXH_K_LOBORDER, s,
                [&] { return launch_lo_border_grouped(s, c, kd, g.nk, g.ngpmax, acc); });
    if (e == hipSuccess)
      e = timed(c, EXH_K_LOLO, s, [&] { return launch_lolo_grouped(s, c, kd, g.nk, acc); });
  }
  return e;
}

}  // namespace

void batch_release_all(exh_context *c) {
  for (auto *E : c->bcache) free_entry(E);
  c->bcache.clear();
}

}  // namespace exh

using namespace exh;

extern "C" {

int exh_setup_batch_device(exh_context *c, void *stream, int nk, const exh_kpoint *kp,
                           int flags) {
  if (!c) return EXH_EINVAL;
  int rc = validate(c, nk, kp, flags);
  if (rc) return rc;
  if (nk == 0) return 0;
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, EXH_ESTATE, "hipSetDevice");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cap) != hipSuccess) cap = hipStreamCaptureStatusNone;
  const bool outer_capture = cap == hipStreamCaptureStatusActive;
  const int plan = c->batch_plan >= 0 ? c->batch_plan : kDefaultPlan;
  const int nmod = (flags >> 8) & 0xff;
  const int ldb = (c->ngkmax + 63) / 64 * 64;
  const size_t pk = perk_bytes(c, ldb);
  const int nkg = (int)std::max<size_t>(1, std::min<size_t>(nk, c->batch_ws_limit / pk));
  OzLayout L;
  int bits = 0;
  if (outer_capture) {
    // nothing may be allocated or uploaded inside a caller's capture
    if (!scratch_fits(c, nkg, ldb) || (nmod && (!c->oz_consts || !c->ozr.p)))
      return fail(c, EXH_ESTATE, "batch: make one call outside graph capture first");
    if (nmod) {
      L = oz_layout(c->ngkmax, c->Ktot, nmod);
      bits = oz_bits(nmod, c->Ktot);
      if (c->ozqx.bytes < L.qx_bytes || c->ozqb.bytes < L.qb_bytes || c->ozr.bytes < L.r_bytes)
        return fail(c, EXH_ESTATE, "batch: make one call outside graph capture first");
    }
  } else {
    rc = ensure_scratch(c, nkg, ldb);
    if (!rc && nmod) rc = oz_prepare(c, nmod, L, bits);
    if (rc) return rc;
  }
  // key: arguments, plan, group size, the context's reallocation epoch and
  // the signature of what the kernels take from the context
  std::vector<unsigned char> key(sizeof(int) * 4 + 2 * sizeof(uint64_t) +
                                 nk * sizeof(exh_kpoint));
  {
    const int hdr[4] = {nk, flags & ~EXH_GRAPH, plan, nkg};
    const uint64_t sig[2] = {c->epoch, state_signature(c)};
    std::memcpy(key.data(), hdr, sizeof hdr);
    std::memcpy(key.data() + sizeof hdr, sig, sizeof sig);
    std::memcpy(key.data() + sizeof hdr + sizeof sig, kp, nk * sizeof(exh_kpoint));
  }
  exh_context::BatchEntry *E = find_entry(c, key);
  if (!E) {
    if (outer_capture)
      return fail(c, EXH_ESTATE, "batch: make one call outside graph capture first");
    E = build_entry(c, s, nk, kp, nkg, ldb, plan, std::move(key), &rc);
    if (!E) return rc;
  }
  E->use = ++c->bclock;
  hipError_t e = hipSuccess;
  if ((flags & EXH_GRAPH) && !outer_capture) {
    if (!E->exec) {
      e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) return hip_fail(c, e, "batch: begin capture");
      c->capturing = true;
      hipError_t el = enqueue(c, s, *E, flags, plan, ldb, L, bits);
      c->capturing = false;
      hipGraph_t g = nullptr;
      e = hipStreamEndCapture(s, &g);
      if (el != hipSuccess || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(c, el != hipSuccess ? el : e, "batch: capture");
      }
      e = hipGraphInstantiate(&E->exec, g, nullptr, nullptr, 0);
      if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        E->exec = nullptr;
        return hip_fail(c, e, "batch: graph instantiate");
      }
      E->graph = g;
    }
    e = timed(c, EXH_K_GRAPH, s, [&] { return hipGraphLaunch(E->exec, s); });
    if (e == hipSuccess) e = mark_done(c, s);
    return e == hipSuccess ? 0 : hip_fail(c, e, "batch: graph launch");
  }
  const bool was = c->capturing;
  c->capturing = was || outer_capture;
  e = enqueue(c, s, *E, flags, plan, ldb, L, bits);
  if (e == hipSuccess) e = mark_done(c, s);
  c->capturing = was;
  return e == hipSuccess ? 0 : hip_fail(c, e, "batch kernels");
}

int exh_batch_plan_count(void) { return kPlans; }

int exh_batch_plan_info(int plan, int *tile_rows, int *tile_cols, int *k_slab, int *persistent,
                        int *xcd_order) {
  if (plan < 0 || plan >= kPlans) return EXH_EINVAL;
  const PlanInfo P = plan_info(plan);
  const HoShapeInfo &S = kHoShapeInfo[P.shape];
  if (tile_rows) *tile_rows = S.tmi;
  if (tile_cols) *tile_cols = S.tmj;
  if (k_slab) *k_slab = S.tk;
  if (persistent) *persistent = P.persistent;
  if (xcd_order) *xcd_order = P.xcd;
  return 0;
}

int exh_set_batch_plan(exh_context *c, int plan) {
  if (!c) return EXH_EINVAL;
  if (plan < -1 || plan >= kPlans) return fail(c, EXH_EINVAL, "batch plan out of range");
  c->batch_plan = plan;
  return 0;
}

int exh_get_batch_plan(exh_context *c) {
  if (!c) return EXH_EINVAL;
  return c->batch_plan >= 0 ? c->batch_plan : kDefaultPlan;
}

int exh_set_batch_workspace(exh_context *c, size_t bytes) {
  if (!c) return EXH_EINVAL;
  c->batch_ws_limit = bytes;
  return 0;
}

int exh_tune_batch(exh_context *c, void *stream, int nk, const exh_kpoint *kp, int flags,
                   int *plan) {
  if (!c) return EXH_EINVAL;
  flags &= ~(EXH_ACCUMULATE | EXH_GRAPH);
  int rc = validate(c, nk, kp, flags);
  if (rc) return rc;
  if (nk == 0) return fail(c, EXH_EINVAL, "tune: no k-points");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, EXH_ESTATE, "hipSetDevice");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  // tune on leading k-points worth about 4 TFLOP of contraction (1 to 8)
  int nt = 0;
  double fl = 0.0;
  while (nt < std::min(nk, 8) && (nt == 0 || fl < 4e12)) {
    fl += 8.0 * kp[nt].ngp * (double)kp[nt].ngp * c->Ktot;
    nt++;
  }
  const int saved = c->batch_plan;
  const bool prof = c->prof;
  c->prof = false;
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  int best = -1;
  float best_ms = 0.f;
  for (int p = 0; p < kPlans && e == hipSuccess && !rc; p++) {
    c->batch_plan = p;
    rc = exh_setup_batch_device(c, s, nt, kp, flags);  // warm-up (code load, descriptors)
    if (rc) break;
    e = hipEventRecord(a, s);
    if (e == hipSuccess) rc = exh_setup_batch_device(c, s, nt, kp, flags);
    if (rc) break;
    e = hipEventRecord(b, s);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    if (e == hipSuccess && (best < 0 || ms < best_ms)) {
      best = p;
      best_ms = ms;
    }
  }
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  c->prof = prof;
  if (rc || e != hipSuccess || best < 0) {
    c->batch_plan = saved;
    return rc ? rc : hip_fail(c, e, "tune");
  }
  c->batch_plan = best;
  if (plan) *plan = best;
  return 0;
}

}  // extern "C"
