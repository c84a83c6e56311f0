/*
 * exh.h -- C ABI of the MI355X-native LAPW Hamiltonian/overlap setup.
 *
 * One library context per GPU holds the module state the reference reads from
 * Fortran globals (basis sizes, muffin-tin tables, G-vector maps, interstitial
 * functions) resident in HBM.  The two hot-path entry points replace
 *
 *   Subroutine match (ngp, gpc, tpgpc, sfacgp, apwalm)          src/match.f90:13
 *   Subroutine hamiltonandoverlapsetup (system, ngp, apwalm, igpig, vgpc)
 *                             src/src_eigensystem/hamiltonandoverlapsetup.f90:8
 *
 * with the semantics of the live pair hamiltonsetup + overlapsetup called at
 * src/seceqnfv.f90:76-77 (src/src_eigensystem/hamiltonsetup.f90:8,
 * overlapsetup.f90:8), including hamiltonoverlapocopy_UL.
 *
 * Conventions
 *   - Arrays use the reference's column-major layouts and 1-based index
 *     VALUES (igpig, ivgig, ivg), exactly as the Fortran caller holds them.
 *   - Atoms are numbered species-major (idxas(ia,is) = sum natoms(<is) + ia).
 *   - Return codes: 0 ok; < 0 invalid argument or state (EXH_E*); > 0 a HIP
 *     runtime error code.  exh_error() returns the last message.
 *   - Functions ending in _device take device pointers and enqueue work on
 *     `stream` (a hipStream_t, NULL = the context's stream) without host
 *     synchronisation.  The others take host pointers and return when the
 *     result is back in host memory.
 *   - A context is used by one host thread at a time.  Work queued by the
 *     _device entry points may run on any stream; an exh_set_* call first
 *     waits for the context's last queued work, so tables are never rewritten
 *     under a running kernel.
 */
#ifndef EXH_H
#define EXH_H

#include <complex.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
typedef struct {
  double re, im;
} exh_z;
#else
typedef double _Complex exh_z;
#endif

typedef struct exh_context exh_context;

enum {
  EXH_OK = 0,
  EXH_EINVAL = -1,   /* bad argument                         */
  EXH_ESTATE = -2,   /* required set_* call missing           */
  EXH_ERANGE = -3,   /* size beyond what the kernels support  */
  EXH_ENOTSUP = -4,  /* option not implemented                */
  EXH_EMATCH = -5    /* singular boundary-derivative matrix   */
};

/* Flags for exh_setup_device / exh_hamiltonandoverlapsetup_ex. */
enum {
  /* Add to the H and O already in memory, as the reference does (zgemm with
   * beta=1 and += gathers; the upper triangle and the lo rows below the APW
   * block are read).  Without it H and O are overwritten, which equals the
   * reference when called right after newsystem() zeroed them
   * (seceqnfv.f90:73). */
  EXH_ACCUMULATE = 1
};
/* Split precision: EXH_SPLIT(n), 6 <= n <= 12, OR-ed into flags, computes the
 * muffin-tin APW-APW products with the Ozaki scheme II on int8 matrix cores:
 * A and aa.A are rounded per G column to b-bit integers (b is the largest
 * value for which n primes <= 251 recover every product exactly; b = 26 for
 * n = 9 at K = 15488) and multiplied exactly through n modular int8 GEMMs.
 * The rounding is the only approximation (relative 2^-b per column).  Match,
 * the lo blocks and the interstitial stay FP64.  Not combined with h1aa. */
#define EXH_SPLIT(n) ((n) << 8)

/* ---- context ------------------------------------------------------------ */
/* No reference counterpart: the reference keeps this state in modules. */
int exh_create(int device, exh_context **ctx);
void exh_destroy(exh_context *ctx);
const char *exh_error(const exh_context *ctx);
/* The context's own stream (hipStream_t), created at default priority. */
void *exh_stream(exh_context *ctx);

/* ---- module state ------------------------------------------------------- */
/* Basis layout: mod_atoms (nspecies, natoms), mod_APW_LO (apwordmax, apword,
 * nlorb, lorbl), input%groundstate%lmaxapw / lmaxmat, mod_Gkvector ngkmax.
 * apword is apword(0:lmaxapw, nspecies); lorbl is lorbl(nlomax, nspecies).
 * Derived here as the reference derives them: idxlo (genidxlo.f90:27-45),
 * losize/maxnlo (mod_eigensystem.F90:206-230), maxaa (MaxAPWs :151-184). */
int exh_set_basis(exh_context *ctx, int nspecies, const int *natoms,
                  int lmaxapw, int lmaxmat, int apwordmax, const int *apword,
                  int nlomax, const int *nlorb, const int *lorbl, int ngkmax);

/* Radial boundary data read by match (match.f90:141-148): omega
 * (mod_lattice), rmt(nspecies), nrmt(nspecies), spr(nrmtmax, nspecies) and
 * apwfr(nrmtmax, 2, apwordmax, 0:lmaxapw, natmtot) (mod_APW_LO; only the last
 * max(apwordmax+1,4) points of component 1 are read).  Changes every SCF
 * iteration (genapwfr, scf_cycle.f90:210). */
int exh_set_radial(exh_context *ctx, double omega, const double *rmt,
                   const int *nrmt, int nrmtmax, const double *spr,
                   const double *apwfr);

/* Muffin-tin tables for the current SCF iteration (hmlint/olprad,
 * scf_cycle.f90:210-222): mt_hscf%main%aa(maxaa,maxaa,natmtot),
 * %loa(maxnlo,maxaa,natmtot), %lolo(maxnlo,maxnlo,natmtot) (complex);
 * oalo(apwordmax,nlomax,natmtot), ololo(nlomax,nlomax,natmtot) (real);
 * h1aa(apwordmax,apwordmax,0:lmaxapw,natmtot), h1loa(apwordmax,nlomax,natmtot),
 * h1lolo(nlomax,nlomax,natmtot) (real, NULL = zero, the non-iora* case).
 * maxaa and maxnlo must equal the values exh_set_basis derived. */
int exh_set_mt(exh_context *ctx, int maxaa, int maxnlo, const exh_z *aa,
               const exh_z *loa, const exh_z *lolo, const double *oalo,
               const double *ololo, const double *h1aa, const double *h1loa,
               const double *h1lolo);

/* G-vector maps (gengvec.f90): ivg(3,ngrtot), the ivgig box with bounds
 * intgv(3,2) (column-major: lo1,lo2,lo3,hi1,hi2,hi3), ngvec.  Once per
 * structure. */
int exh_set_gvectors(exh_context *ctx, int ngrtot, int ngvec, const int *ivg,
                     const int *intgv, const int *ivgig);

/* Interstitial functions for the current SCF iteration: cfunig(ngrtot),
 * veffig(ngvec) (genveffig), meffig(ngvec) read when valence_rel_none == 0
 * (hmlistln.f90:61-71), m2effig(ngvec) read when h1on != 0
 * (olpistln.f90:53-66). */
int exh_set_interstitial(exh_context *ctx, const exh_z *cfunig,
                         const exh_z *veffig, const exh_z *meffig,
                         const exh_z *m2effig, int valence_rel_none, int h1on);
/* Same from device memory, ordered after earlier work on `stream` (NULL =
 * the context's stream); returns once the tables are built.  With the
 * producers below, the interstitial inputs never leave HBM. */
int exh_set_interstitial_device(exh_context *ctx, void *stream, const exh_z *cfunig,
                                const exh_z *veffig, const exh_z *meffig,
                                const exh_z *m2effig, int valence_rel_none, int h1on);

/* Same from device memory (the tables exh_hmlint_device makes, oalo/ololo,
 * h1 terms), ordered after earlier work on `stream`; returns once installed.
 * Tables of the same sizes are rewritten in place. */
int exh_set_mt_device(exh_context *ctx, void *stream, int maxaa, int maxnlo,
                      const exh_z *aa, const exh_z *loa, const exh_z *lolo,
                      const double *oalo, const double *ololo, const double *h1aa,
                      const double *h1loa, const double *h1lolo);

/* ---- muffin-tin table producer ------------------------------------------ */
/* hmlint.f90:53-55 for the spinless Hamiltonian: mt_kin.f90:14-237 then
 * mt_pot.f90:14-323 into aa(maxaa,maxaa,natmtot), loa(maxnlo,maxaa,natmtot),
 * lolo(maxnlo,maxnlo,natmtot) (the layouts of mt_hscf%main; maxnlo >= 1 as
 * allocated).  veffmt(lmmaxvr,nrmtmax,natmtot) with lmmaxvr = (lmaxvr+1)^2
 * (mod_potential_and_density), lofr(nrmtmax,2,nlomax,natmtot) (mod_APW_LO);
 * apwfr, spr and nrmt come from exh_set_radial, the basis from
 * exh_set_basis.  valence_rel_none selects the kinetic mass factor of
 * mt_kin.f90:55-59.  flags: EXH_ACCUMULATE adds to the tables as the
 * reference does (it zeroes them before hmlint), 0 overwrites.  Radial
 * integrals use the spline rule of fderiv.f90; the Gaunt coefficients are
 * those of init1.f90:563-609 (lm2 <= lmmaxvr, lm1 and lm3 <= lmmaxmat).  At
 * most 63 radial functions per atom ((lmaxmat+1)*apwordmax + nlomax). */
int exh_hmlint(exh_context *ctx, int lmaxvr, const double *veffmt, const double *lofr,
               int valence_rel_none, exh_z *aa, exh_z *loa, exh_z *lolo, int flags);
/* Same with every array in device memory, enqueued on `stream`. */
int exh_hmlint_device(exh_context *ctx, void *stream, int lmaxvr, const double *veffmt,
                      const double *lofr, int valence_rel_none, exh_z *aa, exh_z *loa,
                      exh_z *lolo, int flags);
/* mt_pot.f90:14-323 alone, for any real muffin-tin function pot(lmmaxvr,
 * nrmtmax, natmtot): the spin-polarised part of hmlint (hmlint.f90:58-173)
 * calls it with the B_xc components (plus the external fields) into the
 * alpha / ab tables, and sets beta = -alpha and ab = -i ab itself.  Same
 * tables, flags and limits as exh_hmlint. */
int exh_mt_pot(exh_context *ctx, int lmaxvr, const double *pot, const double *lofr,
               exh_z *aa, exh_z *loa, exh_z *lolo, int flags);
int exh_mt_pot_device(exh_context *ctx, void *stream, int lmaxvr, const double *pot,
                      const double *lofr, exh_z *aa, exh_z *loa, exh_z *lolo, int flags);

/* mt_so.f90:11-397, as hmlint.f90:172 calls it (mt_so(veffmt, mt_basis,
 * mt_basis, mt_h, level_zora)) when spin-orbit coupling is on: for every
 * pair of radial functions of equal l >= 1, R = integral of u1 u2 r^2
 * (1/r) d/dr rm(r), with rm = a V00 y00 (rel_level 0), 1/(1 - a V00 y00)
 * (1, zora) or a/(1 - a V00 y00)^2 (2), a = alpha^2/2 and V00 = pot(1,:,ias),
 * the derivative being fderiv.f90's (m = 1).  Per m, alpha += m R/2 and
 * beta -= m R/2 on (m, m); ab += sqrt(l(l+1) - m(m+1)) R/2 on (m, m+1), in
 * aa, loa, lolo and (APW row, LO column) alo.  alpha%alo and beta%alo are
 * then set to conjg(transpose(loa)) over the whole table (:361-366).
 * Tables: aa(maxaa,maxaa,natmtot), loa(maxnlo,maxaa,natmtot),
 * lolo(maxnlo,maxnlo,natmtot), alo(maxaa,maxnlo,natmtot); the LO tables may
 * be null without local orbitals.  EXH_ACCUMULATE adds to them as the
 * reference does (after hmlint's mt_pot calls); 0 zeroes them first.
 * nrmtmax <= 4096. */
int exh_mt_so(exh_context *ctx, int lmaxvr, const double *pot, const double *lofr,
              int rel_level, exh_z *alpha_aa, exh_z *alpha_loa, exh_z *alpha_lolo,
              exh_z *alpha_alo, exh_z *beta_aa, exh_z *beta_loa, exh_z *beta_lolo,
              exh_z *beta_alo, exh_z *ab_aa, exh_z *ab_loa, exh_z *ab_lolo, exh_z *ab_alo,
              int flags);
int exh_mt_so_device(exh_context *ctx, void *stream, int lmaxvr, const double *pot,
                     const double *lofr, int rel_level, exh_z *alpha_aa, exh_z *alpha_loa,
                     exh_z *alpha_lolo, exh_z *alpha_alo, exh_z *beta_aa, exh_z *beta_loa,
                     exh_z *beta_lolo, exh_z *beta_alo, exh_z *ab_aa, exh_z *ab_loa,
                     exh_z *ab_lolo, exh_z *ab_alo, int flags);

/* The kinetic radial integrals alone, as mt_kin.f90:117-170 and
 * kinetic_energy.f90:50-183 (t_aa, t_alo, t_lolo) take them: for the radial
 * functions f of each atom -- APW (l, io) as f = l*apwordmax + io for
 * l <= lmaxmat, then LO ilo as f = (lmaxmat+1)*apwordmax + ilo -- and f1, f2
 * of equal l, the spline integral of 1/2 m(r) (u1' u2' r^2 + l(l+1) u1 u2);
 * zero for pairs of different l.  m(r) = 1/(1 - alpha^2/2 veffmt(1,r) y00)
 * on APW-APW pairs with EXH_KIN_MASS_APW and on pairs with an LO with
 * EXH_KIN_MASS_LO, else 1.  mt_kin uses both for non-'none' relativity and
 * neither for 'none'; kinetic_energy.f90 uses both for ZORA and, for 'none',
 * only EXH_KIN_MASS_LO (its t_alo/t_lolo always carry the factor,
 * :139-150, :159-167).  tkin(nf, nf, natmtot) in device memory,
 * nf = (lmaxmat+1)*apwordmax + nlomax.  veffmt (lmaxvr) is read only with a
 * mass flag. */
enum { EXH_KIN_MASS_APW = 1, EXH_KIN_MASS_LO = 2 };
int exh_mt_kinetic_device(exh_context *ctx, void *stream, int lmaxvr, const double *veffmt,
                          const double *lofr, int mass_flags, double *tkin);

/* ---- interstitial producers --------------------------------------------- */
/* They need exh_set_gvectors with the whole FFT box as the G set (ngrtot =
 * ngrid(1)*ngrid(2)*ngrid(3), ngrid(d) = intgv(d,2) - intgv(d,1) + 1, as
 * gridsize.f90 makes it).  Real-space arrays are on that grid with the first
 * index fastest; FFTs are hipFFT double-complex transforms with the
 * reference's conventions (zfftifc.f90: sign -1 forward and scaled by 1/N,
 * +1 backward unscaled).
 *
 * gencfun.f90:13-85 (gencfun + gencfunig :96-159): the smooth step function
 * cfunig(ngrtot) in the context's G order and cfunir(ngrtot) = Re of its
 * backward transform; cfdamp > 0 damps cfunig by exp(-(cfdamp/gmaxvr*|G|)^2)
 * after the transform, as the reference does.  bvec(3,3) (columns b1,b2,b3,
 * mod_lattice), atposc(3,maxatoms,nspecies) Cartesian (mod_atoms), epslat
 * (input%structure%epslat) are host arrays; omega, rmt, nspecies and natoms
 * come from exh_set_radial and exh_set_basis. */
int exh_gencfun(exh_context *ctx, const double *bvec, const double *atposc, int maxatoms,
                double epslat, double cfdamp, double gmaxvr, exh_z *cfunig, double *cfunir);
/* Same with cfunig and cfunir in device memory, enqueued on `stream`. */
int exh_gencfun_device(exh_context *ctx, void *stream, const double *bvec,
                       const double *atposc, int maxatoms, double epslat, double cfdamp,
                       double gmaxvr, exh_z *cfunig, double *cfunir);
/* genveffig.f90:13-44: veffig(ngvec) from veffir(ngrtot) * cfunir(ngrtot). */
int exh_genveffig(exh_context *ctx, const double *veffir, const double *cfunir,
                  exh_z *veffig);
/* genmeffig.f90:13-61: nothing when valence_rel_none; else meffig(ngvec) from
 * cfunir / (1 - veffir*a2) and, when iora_star, m2effig(ngvec) from
 * cfunir / (1 - veffir*a2)^2, a2 = alpha^2/2. */
int exh_genmeffig(exh_context *ctx, const double *veffir, const double *cfunir,
                  int valence_rel_none, int iora_star, exh_z *meffig, exh_z *m2effig);
/* Device form of both: the products selected by `which` (EXH_VEFFIG,
 * EXH_MEFFIG, EXH_M2EFFIG) are transformed as one batched FFT and gathered;
 * outputs not selected may be NULL.  Enqueued on `stream`. */
enum { EXH_VEFFIG = 1, EXH_MEFFIG = 2, EXH_M2EFFIG = 4 };
int exh_genveffig_device(exh_context *ctx, void *stream, const double *veffir,
                         const double *cfunir, int which, exh_z *veffig, exh_z *meffig,
                         exh_z *m2effig);

/* ---- hot path ----------------------------------------------------------- */
/* Host-array entry points (exh_match, exh_hamiltonandoverlapsetup[_ex],
 * exh_overlapsetup) copy arrays larger than 4 MiB through two pinned 64 MiB
 * host buffers that the context allocates on first use and exh_destroy
 * frees; the host side of each piece runs on EXH_COPY_THREADS threads
 * (environment variable, default 8).  The caller's arrays stay pageable and
 * may be reused as soon as the call returns. */
/* match.f90:13-194.  gpc(ngkmax), tpgpc(2,ngkmax), sfacgp(ngkmax,natmtot) in;
 * apwalm(ngkmax,apwordmax,(lmaxapw+1)^2,natmtot) out (entries with io beyond
 * apword(l) are left untouched, as in the reference). */
int exh_match(exh_context *ctx, int ngp, const double *gpc,
              const double *tpgpc, const exh_z *sfacgp, exh_z *apwalm);
int exh_match_device(exh_context *ctx, void *stream, int ngp,
                     const double *gpc, const double *tpgpc,
                     const exh_z *sfacgp, exh_z *apwalm);

/* hamiltonandoverlapsetup.f90:8 signature with hamiltonsetup.f90 +
 * overlapsetup.f90 semantics.  `system` unpacks to rank = nmat (the order of
 * H and O, ngp + nlotot) and the two rank x rank column-major matrices
 * system%hamilton%za and system%overlap%za.  apwalm(ngkmax,...) as produced
 * by match, igpig(ngkmax) (1-based G-vector indices), vgpc(3,ngkmax).  H and
 * O are accumulated into (reference semantics). */
int exh_hamiltonandoverlapsetup(exh_context *ctx, int ngp, int rank,
                                const exh_z *apwalm, const int *igpig,
                                const double *vgpc, exh_z *h, exh_z *o);
/* Same with EXH_* flags (0 = overwrite). */
int exh_hamiltonandoverlapsetup_ex(exh_context *ctx, int ngp, int rank,
                                   const exh_z *apwalm, const int *igpig,
                                   const double *vgpc, exh_z *h, exh_z *o,
                                   int flags);
/* Device-resident form: every pointer is device memory, ldh = rank. */
int exh_setup_device(exh_context *ctx, void *stream, int ngp, int rank,
                     const exh_z *apwalm, const int *igpig, const double *vgpc,
                     exh_z *h, exh_z *o, int flags);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same setup as exh_setup_device (FP64; flags
 * EXH_ACCUMULATE or 0), with bitwise the same result, but its scratch
 * (G+k index map, aa.A operand, and for iora* or non-uniform layouts one
 * more operand) lives in a static 64 MiB device array instead of the
 * context's buffers: it calls no hipMalloc, hipFree or allocating library
 * routine.  Bases needing more than 64 MiB (16 * natmtot*maxaa * ngkmax
 * rounded up, per operand) are refused with EXH_ERANGE, split precision
 * with EXH_ENOTSUP.  Every context on a GPU shares that GPU's array, so
 * calls that overlap on one GPU (two streams, two host threads) race; the
 * array keeps the last call's contents, and every kernel writes its scratch
 * before reading it.  No profiling events are recorded. */
int exh_setup_device_noalloc(exh_context *ctx, void *stream, int ngp, int rank,
                             const exh_z *apwalm, const int *igpig, const double *vgpc,
                             exh_z *h, exh_z *o, int flags);

/* No counterpart in the reference (its arrays never leave host memory).
 * With reuse on, the host-pointer setups (exh_hamiltonandoverlapsetup[_ex],
 * exh_overlapsetup) skip uploading apwalm when it is the array the last
 * host-pointer exh_match filled, or the last such setup read, on this
 * context: same address, same ngp, no exh_set_basis since.  That saves
 * 16 * ngp * apwordmax * (lmaxapw+1)^2 * natmtot bytes of PCIe per call
 * (2.3 GB for 64 atoms at ngk 9340).  The caller guarantees that apwalm was
 * not modified in between, as exciting's k loop does: match fills it and
 * the setup reads it next (seceqn.f90:57-59, calc_vnlmat.f90:54-60).  Off
 * by default. */
int exh_set_apwalm_reuse(exh_context *ctx, int on);
/* Host-pointer apwalm uploads made and skipped by reuse since the context
 * was created. */
int exh_apwalm_transfers(exh_context *ctx, long long *uploads, long long *reused);

/* overlapsetup.f90:8-166 alone: O(rank,rank) accumulated (muffin-tin, lo and
 * interstitial parts, upper triangle) and mirrored; then, as the reference
 * calls hamiltonoverlapocopy_UL for both matrices (:164-166), the lower
 * triangle of h(rank,rank) is set to the conjugate of its upper one (h may be
 * NULL to leave it alone).  Used without hamiltonsetup by calc_vnlmat
 * (src_hybrids/calc_vnlmat.f90:49-60). */
int exh_overlapsetup(exh_context *ctx, int ngp, int rank, const exh_z *apwalm,
                     const int *igpig, const double *vgpc, exh_z *h, exh_z *o);
/* Same in device memory; flags EXH_ACCUMULATE or 0 (overwrite O). */
int exh_overlapsetup_device(exh_context *ctx, void *stream, int ngp, int rank,
                            const exh_z *apwalm, const int *igpig, const double *vgpc,
                            exh_z *h, exh_z *o, int flags);

/* ---- the eigensolve after the setup ------------------------------------- */
/* No counterpart in the reference as a call (it is inside ZHEGVX,
 * modfvsystem.f90:585-720): the product of the Chebyshev-filtered subspace
 * iteration that solves the reduced problem for the lowest states.
 * y = alpha * op(a) x + beta * y0 + gamma * z, all device memory, complex,
 * enqueued on `stream`, with op(a) = a^H (conj_a != 0) or a^T (conj_a = 0)
 * for a(k, i) at a[k*lda + i] (k < k, i < nout).  A Hermitian matrix C gives
 * C x either way without a copy: row-major storage with conj_a = 1,
 * column-major storage with conj_a = 0.  x(k, j) at x[k*ldx + j] (j < m);
 * y0, z and y row-major nout x m (y0 and z may be NULL, y may be y0).  K is
 * split into at most `split` chunks, each summed by its own workgroups into
 * work (split * nout * m complex; FP64 MFMA tiles, or for m = 1 one thread
 * per output row), then the chunks are added in order: the result does not
 * depend on the launch.  Returns
 * EXH_EINVAL for bad sizes and the HIP error code (> 0) if a launch
 * fails. */
int exh_zgemm_cn_device(void *stream, int conj_a, int nout, int m, int k, const exh_z *a,
                        long long lda, const exh_z *x, long long ldx, double alpha, double beta,
                        const exh_z *y0, long long ldy0, double gamma, const exh_z *z,
                        long long ldz, exh_z *y, long long ldy, int split, exh_z *work);

/* ---- batched k-points ---------------------------------------------------- */
/* No reference counterpart as one call: the reference runs match and
 * hamiltonandoverlapsetup once per k-point inside the k loop of the SCF cycle
 * (scf_cycle.f90:256-296 -> seceqn.f90:57 match, seceqnfv.f90:76-77 setup).
 * A batch takes the whole list of k-points of one structure (uneven ngp) and
 * runs every stage as one grouped launch over all of them; the H/O
 * contraction is one persistent launch over the upper tiles of every
 * k-point.  Per k-point the result is bitwise identical to
 * exh_match_device + exh_setup_device with the same flags.  Device pointers
 * throughout; rank = ngp + nlotot is the order and leading dimension of h and
 * o.  gpc, tpgpc and sfacgp are read only with EXH_MATCH. */
typedef struct exh_kpoint {
  int ngp;
  int rank;
  const double *gpc;     /* gpc(ngkmax)                     (EXH_MATCH) */
  const double *tpgpc;   /* tpgpc(2,ngkmax)                 (EXH_MATCH) */
  const exh_z *sfacgp;   /* sfacgp(ngkmax,natmtot)          (EXH_MATCH) */
  exh_z *apwalm;         /* apwalm(ngkmax,apwordmax,lmmaxapw,natmtot)   */
  const int *igpig;      /* igpig(ngkmax), 1-based                       */
  const double *vgpc;    /* vgpc(3,ngkmax)                               */
  exh_z *h;              /* h(rank,rank)                                 */
  exh_z *o;              /* o(rank,rank)                                 */
} exh_kpoint;
enum {
  /* Batch: compute apwalm with match first (grouped), as seceqn does before
   * the setup. */
  EXH_MATCH = 2,
  /* Batch: capture the launches in a HIP graph on the first call and replay
   * it on later calls with identical arguments (same pointers, sizes, flags
   * and plan).  exh_set_* calls that rewrite tables in place (same sizes, as
   * in every SCF iteration) keep the graph, and replays read the new values;
   * a change of any size, address or flag the kernels take from the context
   * makes the next call capture a new graph. */
  EXH_GRAPH = 4
};
/* Flags: EXH_ACCUMULATE, EXH_SPLIT(n), EXH_MATCH, EXH_GRAPH.  The per-k
 * scratch of one grouped launch is bounded by the batch workspace limit;
 * larger batches run as consecutive groups. */
int exh_setup_batch_device(exh_context *ctx, void *stream, int nk, const exh_kpoint *kpts,
                           int flags);
/* The same over k-points of several structures at once: kpts[i] belongs to
 * ctxs[i] (each context holds one structure's tables; all on one device).
 * Every stage is one grouped launch over all k-points (match once per
 * lmaxapw present), the FP64 H/O contraction one persistent launch over the
 * upper tiles of all of them; per k-point the result is bitwise that of
 * exh_setup_batch_device on its own context.  ctxs[0] supplies the plan,
 * the default stream, the descriptor/graph cache and the profiling records;
 * each context's batch workspace limit bounds its k-points per group.  No
 * reference counterpart: the reference sets up one k-point of one structure
 * per call (seceqnfv.f90:76-77); this is the serving loop's launch. */
int exh_setup_multi_device(exh_context *const *ctxs, void *stream, int nk,
                           const exh_kpoint *kpts, int flags);
/* Plans of the grouped H/O kernel: tile shape (rows x columns x K slab),
 * persistent grid (one workgroup per resident slot on every CU, walking the
 * tile list) or one workgroup per tile, and tile order (k-major, or split
 * into 8 contiguous ranges, one per XCD).  All plans give bitwise identical
 * results. */
int exh_batch_plan_count(void);
int exh_batch_plan_info(int plan, int *tile_rows, int *tile_cols, int *k_slab,
                        int *persistent, int *xcd_order);
/* Plan used by exh_setup_batch_device (-1 = the default, plan 5: 64x64 tiles
 * with K slabs of 8, one workgroup per tile, XCD order). */
int exh_set_batch_plan(exh_context *ctx, int plan);
int exh_get_batch_plan(exh_context *ctx);
/* Autotuner: times every plan on (the first few of) these k-points, with the
 * same flags minus EXH_ACCUMULATE and EXH_GRAPH, keeps the fastest as the
 * context's batch plan and returns it in *plan.  h and o are used as scratch:
 * on return they hold the non-accumulated result. */
int exh_tune_batch(exh_context *ctx, void *stream, int nk, const exh_kpoint *kpts, int flags,
                   int *plan);
/* Upper bound on the batch's per-k-point scratch in bytes (default 8 GiB,
 * at least one k-point's worth is always used). */
int exh_set_batch_workspace(exh_context *ctx, size_t bytes);
/* ---- instrumentation ----------------------------------------------------- */
/* No reference counterpart (the reference accumulates wall-clock timers such
 * as timemat, hamiltonsetup.f90:190).  When enabled, every hot-path kernel
 * launch is bracketed by HIP events on the stream it runs on; read returns
 * the summed device time (ms) and launch count of one kernel since the last
 * reset, synchronising on the recorded events. */
enum {
  EXH_K_MATCH = 0,    /* match kernel                                  */
  EXH_K_IVGK = 1,     /* G+k -> G index prep                           */
  EXH_K_BT = 2,       /* aa . A per atom (FP64 MFMA)                    */
  EXH_K_A2T = 3,      /* (1 + h1aa) . A (iora* only)                    */
  EXH_K_HO = 4,       /* fused upper H/O contraction + interstitial     */
  EXH_K_LOBORDER = 5, /* APW-lo rows/columns                           */
  EXH_K_LOLO = 6,     /* lo-lo block                                   */
  EXH_K_PACK = 7,     /* K-major copy of apwalm (non-uniform layouts)  */
  EXH_K_OZRES = 8,    /* split: column exponents + int8 residues        */
  EXH_K_OZGEMM = 9,   /* split: modular int8 GEMMs (MFMA)               */
  EXH_K_OZCOMB = 10,  /* split: Garner/CRT + interstitial + mirror      */
  EXH_K_HOGRP = 11,   /* batch: grouped persistent H/O contraction      */
  EXH_K_GRAPH = 12,   /* batch: one replay of a captured batch graph    */
  EXH_K_CFUNIG = 13,  /* gencfunig closed form + scatter to the grid    */
  EXH_K_IFFT = 14,    /* interstitial producers: hipFFT transforms      */
  EXH_K_IPROD = 15,   /* genveffig/genmeffig products on the grid       */
  EXH_K_IGATHER = 16, /* grid -> G gathers (cfunir, veffig, meffig...)  */
  EXH_K_HMLRAD = 17,  /* hmlint: radial integrals (potential, kinetic)  */
  EXH_K_HMLCON = 18,  /* hmlint: Gaunt contraction into aa/loa/lolo     */
  EXH_K_SO = 19,      /* mt_so: spin-orbit terms of alpha/beta/ab       */
  EXH_K_COUNT = 20
};
int exh_profile_enable(exh_context *ctx, int on);
int exh_profile_read(exh_context *ctx, int kernel, double *total_ms, int *launches);
int exh_profile_reset(exh_context *ctx);

#ifdef __cplusplus
}
#endif
#endif
