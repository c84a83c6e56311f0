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
  std::vector<exh_context *> ctxs;  // distinct contexts, in order of first use
  std::vector<int> kci;             // context index of every k-point
  std::vector<exh::KDesc> kd;       // host copies (kept alive: sources of the uploads)
  std::vector<exh::SDesc> sd;
  std::vector<int4> tiles;          // every group's tile list, concatenated
  struct Group {
    int k0, t0, nt, nwg;
    exh::GroupDims dims;
  };
  std::vector<Group> groups;
  DevBuf kd_d, sd_d, tiles_d;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t use = 0;
};

namespace exh {
namespace {

constexpr int kPlans = kHoShapes * 4;  // shape x {tile grid, persistent} x {k-major, XCD}
// 64x64x8, one workgroup per tile, XCD order: the autotuner's choice at 64
// atoms (16 k-points and one) and within 1 % of it at 8 and 16 atoms
constexpr int kDefaultPlan = 5;
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
  release(E->sd_d);
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
                          &c->ba2t, &c->bivgk, &c->bmtab, &c->ozqx, &c->ozqb, &c->ozr, &c->ozex,
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

int validate(exh_context *lead, int nk, const exh_kpoint *kp, exh_context *const *kc,
             int flags) {
  if (!lead->have_basis || !lead->have_mt || !lead->have_gvec || !lead->have_ist)
    return fail(lead, EXH_ESTATE, "batch: basis, muffin-tin tables and interstitial data must be set");
  if (nk < 0 || (nk > 0 && (!kp || !kc))) return fail(lead, EXH_EINVAL, "batch: bad k-point list");
  if (flags & ~(EXH_ACCUMULATE | EXH_MATCH | EXH_GRAPH | 0xff00))
    return fail(lead, EXH_ENOTSUP, "batch: unknown flag");
  const int nmod = (flags >> 8) & 0xff;
  if (nmod && (nmod < 6 || nmod > kOzMaxMod))
    return fail(lead, EXH_EINVAL, "batch: EXH_SPLIT(n) needs 6 <= n <= 12");
  for (int i = 0; i < nk; i++) {
    const exh_context *c = kc[i];
    if (!c) return fail(lead, EXH_EINVAL, "batch: null context of a k-point");
    if (c->device != lead->device)
      return fail(lead, EXH_EINVAL, "batch: contexts on different devices");
    if (!c->have_basis || !c->have_mt || !c->have_gvec || !c->have_ist)
      return fail(lead, EXH_ESTATE,
                  "batch: basis, muffin-tin tables and interstitial data must be set");
    if ((flags & EXH_MATCH) && !c->have_radial)
      return fail(lead, EXH_ESTATE, "batch: EXH_MATCH needs the radial data");
    if (nmod && c->have_h1aa)
      return fail(lead, EXH_ENOTSUP,
                  "batch: split precision with h1aa (iora*) is not implemented");
    if (c->lmaxapw > 12) return fail(lead, EXH_ENOTSUP, "batch: lmaxapw above 12");
    const exh_kpoint &k = kp[i];
    if (k.ngp < 0 || k.ngp > c->ngkmax)
      return fail(lead, EXH_EINVAL, "batch: ngp outside 0..ngkmax");
    if (k.rank != k.ngp + c->nlotot) return fail(lead, EXH_EINVAL, "batch: rank != ngp + nlotot");
    if (!k.h || !k.o) return fail(lead, EXH_EINVAL, "batch: null h or o");
    if (k.ngp > 0 && (!k.apwalm || !k.igpig || !k.vgpc))
      return fail(lead, EXH_EINVAL, "batch: null apwalm, igpig or vgpc");
    if ((flags & EXH_MATCH) && k.ngp > 0 && (!k.gpc || !k.tpgpc || !k.sfacgp))
      return fail(lead, EXH_EINVAL, "batch: EXH_MATCH with null gpc, tpgpc or sfacgp");
  }
  return 0;
}

int ldb_of(const exh_context *c) { return (c->ngkmax + 63) / 64 * 64; }

// Per-k-point scratch bytes of a context.
// Per-k-point scratch bytes of a context: the K-major operands, the G+k box
// indices and the match kernel's radial table.
size_t perk_bytes(const exh_context *c) {
  const size_t op = (size_t)std::max(c->Ktot, 1) * ldb_of(c) * sizeof(double2);
  return op * (1 + (c->contig ? 0 : 1) + (c->have_h1aa ? 1 : 0)) +
         (size_t)std::max(c->ngkmax, 1) * sizeof(int4) + match_table_bytes(c);
}

int ensure_scratch(exh_context *c, int nkg) {
  const size_t op = (size_t)std::max(c->Ktot, 1) * ldb_of(c) * sizeof(double2);
  int rc = ensure(c, c->bbt, op * nkg);
  if (!rc && !c->contig) rc = ensure(c, c->bat, op * nkg);
  if (!rc && c->have_h1aa) rc = ensure(c, c->ba2t, op * nkg);
  if (!rc) rc = ensure(c, c->bivgk, (size_t)std::max(c->ngkmax, 1) * sizeof(int4) * nkg);
  if (!rc) rc = ensure(c, c->bmtab, match_table_bytes(c) * nkg);
  return rc;
}

bool scratch_fits(const exh_context *c, int nkg) {
  const size_t op = (size_t)std::max(c->Ktot, 1) * ldb_of(c) * sizeof(double2);
  return c->bbt.p && c->bbt.bytes >= op * nkg && (c->contig || c->bat.bytes >= op * nkg) &&
         (!c->have_h1aa || c->ba2t.bytes >= op * nkg) &&
         c->bivgk.bytes >= (size_t)std::max(c->ngkmax, 1) * sizeof(int4) * nkg &&
         c->bmtab.bytes >= match_table_bytes(c) * nkg;
}

exh_context::BatchEntry *find_entry(exh_context *c, const std::vector<unsigned char> &key) {
  for (auto *E : c->bcache)
    if (E->key == key) return E;
  return nullptr;
}

// The distinct contexts of a call, the context index of every k-point and
// each context's k-point count.
struct CallContexts {
  std::vector<exh_context *> ctxs;
  std::vector<int> kci, count, nkg;
};

CallContexts call_contexts(int nk, exh_context *const *kc) {
  CallContexts cc;
  cc.kci.resize(nk);
  for (int i = 0; i < nk; i++) {
    int j = 0;
    while (j < (int)cc.ctxs.size() && cc.ctxs[j] != kc[i]) j++;
    if (j == (int)cc.ctxs.size()) {
      cc.ctxs.push_back(kc[i]);
      cc.count.push_back(0);
    }
    cc.kci[i] = j;
    cc.count[j]++;
  }
  // k-points of one context that share a group (the context's scratch slots)
  for (size_t j = 0; j < cc.ctxs.size(); j++)
    cc.nkg.push_back((int)std::max<size_t>(
        1, std::min<size_t>(cc.count[j], cc.ctxs[j]->batch_ws_limit / perk_bytes(cc.ctxs[j]))));
  return cc;
}

exh_context::BatchEntry *build_entry(exh_context *lead, hipStream_t s, int nk,
                                     const exh_kpoint *kp, const CallContexts &cc, int plan,
                                     int flags, std::vector<unsigned char> key, int *rc) {
  *rc = 0;
  if (lead->bcache.size() >= kMaxEntries) {
    // the oldest entry may still be read by queued work on any stream
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      *rc = hip_fail(lead, e, "batch: cache eviction");
      return nullptr;
    }
    auto it = std::min_element(lead->bcache.begin(), lead->bcache.end(),
                               [](auto *a, auto *b) { return a->use < b->use; });
    free_entry(*it);
    lead->bcache.erase(it);
  }
  auto *E = new exh_context::BatchEntry();
  E->key = std::move(key);
  E->ctxs = cc.ctxs;
  E->kci = cc.kci;
  const PlanInfo P = plan_info(plan);
  const HoShapeInfo &S = kHoShapeInfo[P.shape];
  const int occ = P.persistent ? ho_grouped_occupancy(P.shape) : 0;
  const size_t nctx = cc.ctxs.size();
  E->sd.resize(nctx);
  for (size_t j = 0; j < nctx; j++) E->sd[j] = make_sdesc(cc.ctxs[j]);
  *rc = ensure(nullptr, E->kd_d, std::max<size_t>(nk, 1) * sizeof(KDesc));
  if (!*rc) *rc = ensure(nullptr, E->sd_d, std::max<size_t>(nctx, 1) * sizeof(SDesc));
  if (*rc) {
    fail(lead, *rc, "batch: descriptor buffers");
    free_entry(E);
    return nullptr;
  }
  const SDesc *sd_dev = (const SDesc *)E->sd_d.p;
  // k-points are grouped in order; a group closes when one of its contexts
  // has used all its scratch slots
  std::vector<int> cnt(nctx, 0), ingroup(nctx, 0);
  E->kd.resize(nk);
  exh_context::BatchEntry::Group g{};
  g.k0 = 0;
  auto close_group = [&](int kend) {
    g.dims.nk = kend - g.k0;
    g.t0 = (int)E->tiles.size();
    build_tiles(kp, g.k0, g.dims.nk, S.tmi, S.tmj, P.xcd != 0, E->tiles);
    g.nt = (int)E->tiles.size() - g.t0;
    long long nwg = g.nt;
    if (P.persistent) {
      nwg = std::min<long long>(nwg, (long long)occ * lead->ncu);
      if (P.xcd) nwg = std::max<long long>(8, nwg / 8 * 8);
    }
    g.nwg = (int)std::max<long long>(nwg, 0);
    E->groups.push_back(g);
    g = exh_context::BatchEntry::Group{};
    g.k0 = kend;
    std::fill(ingroup.begin(), ingroup.end(), 0);
  };
  for (int i = 0; i < nk; i++) {
    const int j = cc.kci[i];
    exh_context *c = cc.ctxs[j];
    if (ingroup[j] == cc.nkg[j]) close_group(i);
    const int slot = cnt[j]++ % cc.nkg[j];
    ingroup[j]++;
    const exh_kpoint &k = kp[i];
    const int ldb = ldb_of(c);
    const size_t op = (size_t)std::max(c->Ktot, 1) * ldb;
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
    d.mtab = (double *)((char *)c->bmtab.p + (size_t)slot * match_table_bytes(c));
    d.x = c->contig ? (const double2 *)d.apwalm : d.at;
    d.sd = sd_dev + j;
    GroupDims &D = g.dims;
    D.ngpmax = std::max(D.ngpmax, k.ngp);
    D.natmax = std::max(D.natmax, c->natmtot);
    D.nspmax = std::max(D.nspmax, c->nspecies);
    D.maxaa = std::max(D.maxaa, c->maxaa);
    D.Kmax = std::max(D.Kmax, c->Ktot);
    D.nlotot = std::max(D.nlotot, c->nlotot);
    D.maxnlo = std::max(D.maxnlo, c->maxnlo);
    if (flags & EXH_MATCH) D.lmax_mask |= 1u << c->lmaxapw;
    D.any_pack = D.any_pack || !c->contig;
    D.any_a2t = D.any_a2t || c->have_h1aa;
  }
  if (nk > 0) close_group(nk);
  *rc = ensure(nullptr, E->tiles_d, std::max<size_t>(E->tiles.size(), 1) * sizeof(int4));
  hipError_t e = hipSuccess;
  if (!*rc && nk > 0)
    e = hipMemcpyAsync(E->kd_d.p, E->kd.data(), E->kd.size() * sizeof(KDesc),
                       hipMemcpyHostToDevice, s);
  if (!*rc && e == hipSuccess && nctx > 0)
    e = hipMemcpyAsync(E->sd_d.p, E->sd.data(), nctx * sizeof(SDesc), hipMemcpyHostToDevice, s);
  if (!*rc && e == hipSuccess && !E->tiles.empty())
    e = hipMemcpyAsync(E->tiles_d.p, E->tiles.data(), E->tiles.size() * sizeof(int4),
                       hipMemcpyHostToDevice, s);
  if (*rc || e != hipSuccess) {
    if (!*rc) *rc = hip_fail(lead, e, "batch: descriptor upload");
    else fail(lead, *rc, "batch: descriptor buffers");
    free_entry(E);
    return nullptr;
  }
  lead->bcache.push_back(E);
  return E;
}

// Split-precision layout and bit count of each context of a call.
struct OzParams {
  std::vector<OzLayout> L;
  std::vector<int> bits;
};

hipError_t enqueue(exh_context *lead, hipStream_t s, const exh_context::BatchEntry &E, int flags,
                   int plan, const OzParams &oz) {
  const int nmod = (flags >> 8) & 0xff, acc = flags & EXH_ACCUMULATE;
  const PlanInfo P = plan_info(plan);
  const KDesc *kd0 = (const KDesc *)E.kd_d.p;
  const int4 *tiles = (const int4 *)E.tiles_d.p;
  hipError_t e = hipSuccess;
  for (const auto &g : E.groups) {
    const KDesc *kd = kd0 + g.k0;
    const GroupDims &D = g.dims;
    if (e == hipSuccess && (flags & EXH_MATCH))
      e = timed(lead, EXH_K_MATCH, s, [&] { return launch_match_grouped(s, D, kd); });
    if (e == hipSuccess)
      e = timed(lead, EXH_K_IVGK, s, [&] { return launch_ivgk_grouped(s, D, kd); });
    if (e == hipSuccess && D.any_pack)
      e = timed(lead, EXH_K_PACK, s, [&] { return launch_pack_grouped(s, D, kd); });
    if (e == hipSuccess)
      e = timed(lead, EXH_K_BT, s, [&] { return launch_bt_grouped(s, D, kd); });
    if (e == hipSuccess && D.any_a2t)
      e = timed(lead, EXH_K_A2T, s, [&] { return launch_a2t_grouped(s, D, kd); });
    if (e == hipSuccess && nmod == 0)
      e = timed(lead, EXH_K_HOGRP, s, [&] {
        return launch_ho_grouped(s, P.shape, kd, tiles + g.t0, g.nt, g.nwg, acc);
      });
    if (nmod > 0)
      for (int i = g.k0; i < g.k0 + D.nk && e == hipSuccess; i++) {
        const KDesc &d = E.kd[i];
        const int j = E.kci[i];
        const exh_context *c = E.ctxs[j];
        const SDesc &S = E.sd[j];
        IstBox box{d.ivgk, d.vgpc, (const IstEntry *)c->ist.p, S.lo0, S.lo1, S.lo2,
                   S.n0, S.n1, S.n2};
        const int kid[3] = {EXH_K_OZRES, EXH_K_OZGEMM, EXH_K_OZCOMB};
        for (int st = 0; st < 3 && e == hipSuccess; st++)
          e = timed(lead, kid[st], s, [&] {
            return launch_ozaki(s, c, d.ngp, d.x, S.ldx, d.bt, S.ldb, nmod, oz.bits[j],
                                oz.L[j], (int8_t *)c->ozqx.p, (int8_t *)c->ozqb.p,
                                (uint8_t *)c->ozr.p, (int *)c->ozex.p, (int *)c->ozeb.p, box,
                                d.h, d.o, d.ldh, acc, 1 << st);
          });
      }
    if (e == hipSuccess)
      e = timed(lead, EXH_K_LOBORDER, s,
                [&] { return launch_lo_border_grouped(s, D, kd, acc); });
    if (e == hipSuccess)
      e = timed(lead, EXH_K_LOLO, s, [&] { return launch_lolo_grouped(s, D, kd, acc); });
  }
  return e;
}

// One call of the batch: k-point i belongs to kc[i]; the lead context holds
// the plan, the cache of descriptors and graphs, and the profiling records.
int run_batch(exh_context *lead, void *stream, int nk, const exh_kpoint *kp,
              exh_context *const *kc, int flags) {
  int rc = validate(lead, nk, kp, kc, flags);
  if (rc) return rc;
  if (nk == 0) return 0;
  if (hipSetDevice(lead->device) != hipSuccess) return fail(lead, EXH_ESTATE, "hipSetDevice");
  hipStream_t s = stream ? (hipStream_t)stream : lead->stream;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cap) != hipSuccess) cap = hipStreamCaptureStatusNone;
  const bool outer_capture = cap == hipStreamCaptureStatusActive;
  const int plan = lead->batch_plan >= 0 ? lead->batch_plan : kDefaultPlan;
  const int nmod = (flags >> 8) & 0xff;
  const CallContexts cc = call_contexts(nk, kc);
  const size_t nctx = cc.ctxs.size();
  OzParams oz;
  oz.L.resize(nctx);
  oz.bits.assign(nctx, 0);
  for (size_t j = 0; j < nctx; j++) {
    exh_context *c = cc.ctxs[j];
    if (outer_capture) {
      // nothing may be allocated or uploaded inside a caller's capture
      if (!scratch_fits(c, cc.nkg[j]) || (nmod && (!c->oz_consts || !c->ozr.p)))
        return fail(lead, EXH_ESTATE, "batch: make one call outside graph capture first");
      if (nmod) {
        oz.L[j] = oz_layout(c->ngkmax, c->Ktot, nmod);
        oz.bits[j] = oz_bits(nmod, c->Ktot);
        if (c->ozqx.bytes < oz.L[j].qx_bytes || c->ozqb.bytes < oz.L[j].qb_bytes ||
            c->ozr.bytes < oz.L[j].r_bytes || c->ozex.bytes < oz.L[j].e_bytes ||
            c->ozeb.bytes < oz.L[j].e_bytes)
          return fail(lead, EXH_ESTATE, "batch: make one call outside graph capture first");
      }
    } else {
      rc = ensure_scratch(c, cc.nkg[j]);
      if (!rc && nmod) rc = oz_prepare(c, nmod, oz.L[j], oz.bits[j]);
      if (rc) {
        if (c != lead) fail(lead, rc, c->err);
        return rc;
      }
    }
  }
  // key: arguments, plan, and per context its address, group size,
  // reallocation epoch and the signature of what the kernels take from it
  std::vector<unsigned char> key;
  auto put = [&](const void *p, size_t n) {
    const unsigned char *b = (const unsigned char *)p;
    key.insert(key.end(), b, b + n);
  };
  const int hdr[4] = {nk, flags & ~EXH_GRAPH, plan, (int)nctx};
  put(hdr, sizeof hdr);
  for (size_t j = 0; j < nctx; j++) {
    const exh_context *c = cc.ctxs[j];
    const uint64_t sig[3] = {(uint64_t)(uintptr_t)c, c->epoch, state_signature(c)};
    put(sig, sizeof sig);
    put(&cc.nkg[j], sizeof(int));
  }
  put(kp, nk * sizeof(exh_kpoint));
  put(cc.kci.data(), nk * sizeof(int));
  exh_context::BatchEntry *E = find_entry(lead, key);
  if (!E) {
    if (outer_capture)
      return fail(lead, EXH_ESTATE, "batch: make one call outside graph capture first");
    E = build_entry(lead, s, nk, kp, cc, plan, flags, std::move(key), &rc);
    if (!E) return rc;
  }
  E->use = ++lead->bclock;
  hipError_t e = hipSuccess;
  auto mark_all = [&]() {
    hipError_t em = hipSuccess;
    for (exh_context *c : cc.ctxs)
      if (em == hipSuccess) em = mark_done(c, s);
    return em;
  };
  if ((flags & EXH_GRAPH) && !outer_capture) {
    if (!E->exec) {
      e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) return hip_fail(lead, e, "batch: begin capture");
      lead->capturing = true;
      hipError_t el = enqueue(lead, s, *E, flags, plan, oz);
      lead->capturing = false;
      hipGraph_t g = nullptr;
      e = hipStreamEndCapture(s, &g);
      if (el != hipSuccess || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(lead, el != hipSuccess ? el : e, "batch: capture");
      }
      e = hipGraphInstantiate(&E->exec, g, nullptr, nullptr, 0);
      if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        E->exec = nullptr;
        return hip_fail(lead, e, "batch: graph instantiate");
      }
      E->graph = g;
    }
    e = timed(lead, EXH_K_GRAPH, s, [&] { return hipGraphLaunch(E->exec, s); });
    if (e == hipSuccess) e = mark_all();
    return e == hipSuccess ? 0 : hip_fail(lead, e, "batch: graph launch");
  }
  const bool was = lead->capturing;
  lead->capturing = was || outer_capture;
  e = enqueue(lead, s, *E, flags, plan, oz);
  if (e == hipSuccess) e = mark_all();
  lead->capturing = was;
  return e == hipSuccess ? 0 : hip_fail(lead, e, "batch kernels");
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
  if (nk < 0) return fail(c, EXH_EINVAL, "batch: bad k-point list");
  std::vector<exh_context *> kc((size_t)nk, c);
  return run_batch(c, stream, nk, kp, kc.data(), flags);
}

int exh_setup_multi_device(exh_context *const *ctxs, void *stream, int nk,
                           const exh_kpoint *kp, int flags) {
  if (!ctxs || !ctxs[0]) return EXH_EINVAL;
  return run_batch(ctxs[0], stream, nk, kp, ctxs, flags);
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
  std::vector<exh_context *> kc((size_t)std::max(nk, 0), c);
  int rc = validate(c, nk, kp, kc.data(), flags);
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
