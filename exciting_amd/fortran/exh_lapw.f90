! ISO_C_BINDING interface to libexh (include/exh.h).
!
! A Fortran caller (exciting's seceqn/seceqnfv) uses this module to reach the
! MI355X H/O setup.  Every entry point a Fortran caller needs has an
! interface here with the same argument order; arrays are passed by address
! (c_loc, or implicitly as assumed-size dummies) in the reference's own
! column-major layouts.  (The profiling calls and the single-k device entry
! points are left to C and Python callers.)
!
!   exh_match                    replaces  match (ngp, gpc, tpgpc, sfacgp, apwalm)
!                                          src/match.f90:13
!   exh_hamiltonandoverlapsetup  replaces  hamiltonandoverlapsetup (system, ngp,
!                                          apwalm, igpig, vgpc)
!                                          src/src_eigensystem/hamiltonandoverlapsetup.f90:8
!                                with the semantics of hamiltonsetup followed by
!                                overlapsetup (seceqnfv.f90:76-77).
!   exh_gencfun                  replaces  gencfun (src/gencfun.f90:13)
!   exh_genveffig                replaces  genveffig (src/genveffig.f90:13)
!   exh_genmeffig                replaces  genmeffig (src/genmeffig.f90:13)
!   exh_overlapsetup             replaces  overlapsetup (system, ngp, apwalm, igpig,
!                                          vgpc) called alone, as calc_vnlmat does
!                                          (src/src_hybrids/calc_vnlmat.f90:52-53)
!   exh_hmlint                   replaces  hmlint's spinless part
!                                          (src/src_eigensystem/hmlint.f90:53-55)
!   exh_mt_pot                   replaces  mt_pot (src/src_eigensystem/mt_pot.f90:14),
!                                          as the spin-polarised blocks of hmlint
!                                          call it (hmlint.f90:58-173)
!   exh_mt_so                    replaces  mt_so (src/src_eigensystem/mt_so.f90:11),
!                                          called from hmlint.f90:172
module exh_lapw
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: EXH_ACCUMULATE = 1
  ! EXH_MATCH of include/exh.h (Fortran names are case-blind: exh_match is the
  ! function): batch flag that runs match first
  integer(c_int), parameter, public :: EXH_BATCH_MATCH = 2
  integer(c_int), parameter, public :: EXH_GRAPH = 4   ! batch: capture / replay a HIP graph
  ! outputs of exh_genveffig_device
  integer(c_int), parameter, public :: EXH_VEFFIG = 1, EXH_MEFFIG = 2, EXH_M2EFFIG = 4
  ! mass factors of exh_mt_kinetic_device
  integer(c_int), parameter, public :: EXH_KIN_MASS_APW = 1, EXH_KIN_MASS_LO = 2

  ! One k-point of a batch (exh_kpoint in include/exh.h); device pointers.
  type, bind(C), public :: exh_kpoint
    integer(c_int) :: ngp, rank
    type(c_ptr) :: gpc, tpgpc, sfacgp, apwalm, igpig, vgpc, h, o
  end type

  public :: exh_create, exh_destroy, exh_set_basis, exh_set_radial, exh_set_mt, &
            exh_set_gvectors, exh_set_interstitial, exh_match, &
            exh_hamiltonandoverlapsetup, exh_hamiltonandoverlapsetup_ex, &
            exh_error_message, exh_split, exh_setup_batch_device, exh_setup_multi_device, &
            exh_tune_batch, &
            exh_set_batch_plan, exh_get_batch_plan, exh_set_batch_workspace, &
            exh_set_interstitial_device, exh_gencfun, exh_gencfun_device, exh_genveffig, &
            exh_genmeffig, exh_genveffig_device, exh_hmlint, exh_hmlint_device, &
            exh_set_mt_device, exh_overlapsetup, exh_overlapsetup_device, exh_mt_kinetic_device, &
            exh_mt_pot, exh_mt_pot_device, exh_mt_so, exh_mt_so_device, exh_stream, &
            exh_match_device, exh_setup_device, exh_setup_device_noalloc, &
            exh_set_apwalm_reuse, exh_apwalm_transfers, exh_zgemm_cn_device, &
            exh_batch_plan_count, exh_batch_plan_info, &
            exh_profile_enable, exh_profile_read, exh_profile_reset

  interface
    integer(c_int) function exh_create(device, ctx) bind(C, name="exh_create")
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: ctx
    end function

    subroutine exh_destroy(ctx) bind(C, name="exh_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine

    type(c_ptr) function exh_error_c(ctx) bind(C, name="exh_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function exh_set_basis(ctx, nspecies, natoms, lmaxapw, lmaxmat, &
        apwordmax, apword, nlomax, nlorb, lorbl, ngkmax) bind(C, name="exh_set_basis")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nspecies, lmaxapw, lmaxmat, apwordmax, nlomax, ngkmax
      integer(c_int), intent(in) :: natoms(*), apword(*), nlorb(*), lorbl(*)
    end function

    integer(c_int) function exh_set_radial(ctx, omega, rmt, nrmt, nrmtmax, spr, apwfr) &
        bind(C, name="exh_set_radial")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: omega
      real(c_double), intent(in) :: rmt(*), spr(*), apwfr(*)
      integer(c_int), intent(in) :: nrmt(*)
      integer(c_int), value :: nrmtmax
    end function

    integer(c_int) function exh_set_mt(ctx, maxaa, maxnlo, aa, loa, lolo, oalo, ololo, &
        h1aa, h1loa, h1lolo) bind(C, name="exh_set_mt")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: maxaa, maxnlo
      type(c_ptr), value :: aa, loa, lolo, oalo, ololo, h1aa, h1loa, h1lolo
    end function

    integer(c_int) function exh_set_gvectors(ctx, ngrtot, ngvec, ivg, intgv, ivgig) &
        bind(C, name="exh_set_gvectors")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ngrtot, ngvec
      integer(c_int), intent(in) :: ivg(*), intgv(*), ivgig(*)
    end function

    integer(c_int) function exh_set_interstitial(ctx, cfunig, veffig, meffig, m2effig, &
        valence_rel_none, h1on) bind(C, name="exh_set_interstitial")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, cfunig, veffig, meffig, m2effig
      integer(c_int), value :: valence_rel_none, h1on
    end function

    ! Device pointers, ordered on `stream` (c_null_ptr: the context's own).
    integer(c_int) function exh_set_interstitial_device(ctx, stream, cfunig, veffig, meffig, &
        m2effig, valence_rel_none, h1on) bind(C, name="exh_set_interstitial_device")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, stream, cfunig, veffig, meffig, m2effig
      integer(c_int), value :: valence_rel_none, h1on
    end function

    ! gencfun.f90:13-85: bvec(3,3), atposc(3,maxatoms,nspecies) as in mod_lattice
    ! and mod_atoms; cfunig(ngrtot), cfunir(ngrtot) out.
    integer(c_int) function exh_gencfun(ctx, bvec, atposc, maxatoms, epslat, cfdamp, gmaxvr, &
        cfunig, cfunir) bind(C, name="exh_gencfun")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: bvec(*), atposc(*)
      integer(c_int), value :: maxatoms
      real(c_double), value :: epslat, cfdamp, gmaxvr
      complex(c_double_complex), intent(out) :: cfunig(*)
      real(c_double), intent(out) :: cfunir(*)
    end function

    ! Same with cfunig/cfunir in device memory (bvec and atposc stay host arrays).
    integer(c_int) function exh_gencfun_device(ctx, stream, bvec, atposc, maxatoms, epslat, &
        cfdamp, gmaxvr, cfunig, cfunir) bind(C, name="exh_gencfun_device")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, stream
      real(c_double), intent(in) :: bvec(*), atposc(*)
      integer(c_int), value :: maxatoms
      real(c_double), value :: epslat, cfdamp, gmaxvr
      type(c_ptr), value :: cfunig, cfunir
    end function

    ! genveffig.f90:13-44: veffir(ngrtot), cfunir(ngrtot) in, veffig(ngvec) out.
    integer(c_int) function exh_genveffig(ctx, veffir, cfunir, veffig) &
        bind(C, name="exh_genveffig")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: veffir(*), cfunir(*)
      complex(c_double_complex), intent(out) :: veffig(*)
    end function

    ! genmeffig.f90:13-61: meffig(ngvec) unless valence_rel_none, m2effig(ngvec)
    ! when iora_star.
    integer(c_int) function exh_genmeffig(ctx, veffir, cfunir, valence_rel_none, iora_star, &
        meffig, m2effig) bind(C, name="exh_genmeffig")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: veffir(*), cfunir(*)
      integer(c_int), value :: valence_rel_none, iora_star
      complex(c_double_complex), intent(out) :: meffig(*), m2effig(*)
    end function

    ! overlapsetup.f90 alone: o accumulated; h only mirrored (may be c_null_ptr
    ! through the device form).
    integer(c_int) function exh_overlapsetup(ctx, ngp, rank, apwalm, igpig, vgpc, h, o) &
        bind(C, name="exh_overlapsetup")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: ngp, rank
      complex(c_double_complex), intent(in) :: apwalm(*)
      integer(c_int), intent(in) :: igpig(*)
      real(c_double), intent(in) :: vgpc(*)
      complex(c_double_complex), intent(inout) :: h(*), o(*)
    end function

    integer(c_int) function exh_overlapsetup_device(ctx, stream, ngp, rank, apwalm, igpig, &
        vgpc, h, o, flags) bind(C, name="exh_overlapsetup_device")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, stream, apwalm, igpig, vgpc, h, o
      integer(c_int), value :: ngp, rank, flags
    end function

    ! hmlint.f90:53-55 (mt_kin + mt_pot): veffmt(lmmaxvr,nrmtmax,natmtot),
    ! lofr(nrmtmax,2,nlomax,natmtot) in; mt_hscf%main%aa/loa/lolo out (flags:
    ! EXH_ACCUMULATE adds as the reference does, 0 overwrites).
    integer(c_int) function exh_hmlint(ctx, lmaxvr, veffmt, lofr, valence_rel_none, aa, loa, &
        lolo, flags) bind(C, name="exh_hmlint")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: lmaxvr, valence_rel_none, flags
      real(c_double), intent(in) :: veffmt(*), lofr(*)
      complex(c_double_complex), intent(inout) :: aa(*), loa(*), lolo(*)
    end function

    integer(c_int) function exh_hmlint_device(ctx, stream, lmaxvr, veffmt, lofr, &
        valence_rel_none, aa, loa, lolo, flags) bind(C, name="exh_hmlint_device")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, stream, veffmt, lofr, aa, loa, lolo
      integer(c_int), value :: lmaxvr, valence_rel_none, flags
    end function

    integer(c_int) function exh_mt_pot(ctx, lmaxvr, pot, lofr, aa, loa, lolo, flags) &
        bind(C, name="exh_mt_pot")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: lmaxvr, flags
      real(c_double), intent(in) :: pot(*), lofr(*)
      complex(c_double_complex), intent(inout) :: aa(*), loa(*), lolo(*)
    end function

    integer(c_int) function exh_mt_pot_device(ctx, stream, lmaxvr, pot, lofr, aa, loa, lolo, &
        flags) bind(C, name="exh_mt_pot_device")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, stream, pot, lofr, aa, loa, lolo
      integer(c_int), value :: lmaxvr, flags
    end function

    ! The context's HIP stream (hipStream_t), for ordering device-pointer calls.
    type(c_ptr) function exh_stream(ctx) bind(C, name="exh_stream")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    ! match and the setup on device pointers (c_loc of device arrays), enqueued
    ! on stream (c_null_ptr = the context's stream).
    integer(c_int) function exh_match_device(ctx, stream, ngp, gpc, tpgpc, sfacgp, apwalm) &
        bind(C, name="exh_match_device")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, stream, gpc, tpgpc, sfacgp, apwalm
      integer(c_int), value :: ngp
    end function

    integer(c_int) function exh_setup_device(ctx, stream, ngp, rank, apwalm, igpig, vgpc, h, o, &
        flags) bind(C, name="exh_setup_device")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, stream, apwalm, igpig, vgpc, h, o
      integer(c_int), value :: ngp, rank, flags
    end function

    ! Same as exh_setup_device without any device allocation: its scratch is
    ! a static 64 MiB array per GPU, shared by every context there (calls
    ! that overlap on one GPU race).  FP64 only.
    integer(c_int) function exh_setup_device_noalloc(ctx, stream, ngp, rank, apwalm, igpig, &
        vgpc, h, o, flags) bind(C, name="exh_setup_device_noalloc")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, stream, apwalm, igpig, vgpc, h, o
      integer(c_int), value :: ngp, rank, flags
    end function

    ! Host-array setups reuse the apwalm the last host match left in HBM
    ! (same array, same ngp); the caller guarantees it was not modified.
    integer(c_int) function exh_set_apwalm_reuse(ctx, on) bind(C, name="exh_set_apwalm_reuse")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function

    integer(c_int) function exh_apwalm_transfers(ctx, uploads, reused) &
        bind(C, name="exh_apwalm_transfers")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx
      integer(c_long_long), intent(out) :: uploads, reused
    end function

    ! Batch plans: their number and each one's shape.
    integer(c_int) function exh_batch_plan_count() bind(C, name="exh_batch_plan_count")
      import :: c_int
    end function

    integer(c_int) function exh_batch_plan_info(plan, tile_rows, tile_cols, k_slab, persistent, &
        xcd_order) bind(C, name="exh_batch_plan_info")
      import :: c_int
      integer(c_int), value :: plan
      integer(c_int), intent(out) :: tile_rows, tile_cols, k_slab, persistent, xcd_order
    end function

    ! Per-kernel timing with HIP events (kernel ids EXH_K_* of include/exh.h).
    integer(c_int) function exh_profile_enable(ctx, on) bind(C, name="exh_profile_enable")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function

    integer(c_int) function exh_profile_read(ctx, kernel, total_ms, launches) &
        bind(C, name="exh_profile_read")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: kernel
      real(c_double), intent(out) :: total_ms
      integer(c_int), intent(out) :: launches
    end function

    integer(c_int) function exh_profile_reset(ctx) bind(C, name="exh_profile_reset")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    ! alpha, beta and ab tables (aa, loa, lolo, alo each); rel_level 1 = level_zora
    integer(c_int) function exh_mt_so(ctx, lmaxvr, pot, lofr, rel_level, alpha_aa, alpha_loa, &
        alpha_lolo, alpha_alo, beta_aa, beta_loa, beta_lolo, beta_alo, ab_aa, ab_loa, ab_lolo, &
        ab_alo, flags) bind(C, name="exh_mt_so")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: lmaxvr, rel_level, flags
      real(c_double), intent(in) :: pot(*), lofr(*)
      complex(c_double_complex), intent(inout) :: alpha_aa(*), alpha_loa(*), alpha_lolo(*), &
          alpha_alo(*), beta_aa(*), beta_loa(*), beta_lolo(*), beta_alo(*), ab_aa(*), &
          ab_loa(*), ab_lolo(*), ab_alo(*)
    end function

    integer(c_int) function exh_mt_so_device(ctx, stream, lmaxvr, pot, lofr, rel_level, &
        alpha_aa, alpha_loa, alpha_lolo, alpha_alo, beta_aa, beta_loa, beta_lolo, beta_alo, &
        ab_aa, ab_loa, ab_lolo, ab_alo, flags) bind(C, name="exh_mt_so_device")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, stream, pot, lofr, alpha_aa, alpha_loa, alpha_lolo, alpha_alo, &
          beta_aa, beta_loa, beta_lolo, beta_alo, ab_aa, ab_loa, ab_lolo, ab_alo
      integer(c_int), value :: lmaxvr, rel_level, flags
    end function

    ! The kinetic radial integrals alone (mt_kin, kinetic_energy): tkin(nf,nf,natmtot).
    integer(c_int) function exh_mt_kinetic_device(ctx, stream, lmaxvr, veffmt, lofr, &
        mass_flags, tkin) bind(C, name="exh_mt_kinetic_device")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, stream, veffmt, lofr, tkin
      integer(c_int), value :: lmaxvr, mass_flags
    end function

    ! exh_set_mt from device memory.
    integer(c_int) function exh_set_mt_device(ctx, stream, maxaa, maxnlo, aa, loa, lolo, oalo, &
        ololo, h1aa, h1loa, h1lolo) bind(C, name="exh_set_mt_device")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, stream
      integer(c_int), value :: maxaa, maxnlo
      type(c_ptr), value :: aa, loa, lolo, oalo, ololo, h1aa, h1loa, h1lolo
    end function

    ! Device form of both, outputs selected by `which` (EXH_VEFFIG, ...).
    integer(c_int) function exh_genveffig_device(ctx, stream, veffir, cfunir, which, veffig, &
        meffig, m2effig) bind(C, name="exh_genveffig_device")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, stream, veffir, cfunir
      integer(c_int), value :: which
      type(c_ptr), value :: veffig, meffig, m2effig
    end function

    integer(c_int) function exh_match(ctx, ngp, gpc, tpgpc, sfacgp, apwalm) &
        bind(C, name="exh_match")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: ngp
      real(c_double), intent(in) :: gpc(*), tpgpc(*)
      complex(c_double_complex), intent(in) :: sfacgp(*)
      complex(c_double_complex), intent(inout) :: apwalm(*)
    end function

    integer(c_int) function exh_hamiltonandoverlapsetup(ctx, ngp, rank, apwalm, igpig, vgpc, &
        h, o) bind(C, name="exh_hamiltonandoverlapsetup")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: ngp, rank
      complex(c_double_complex), intent(in) :: apwalm(*)
      integer(c_int), intent(in) :: igpig(*)
      real(c_double), intent(in) :: vgpc(*)
      complex(c_double_complex), intent(inout) :: h(*), o(*)
    end function

    integer(c_int) function exh_hamiltonandoverlapsetup_ex(ctx, ngp, rank, apwalm, igpig, &
        vgpc, h, o, flags) bind(C, name="exh_hamiltonandoverlapsetup_ex")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: ngp, rank, flags
      complex(c_double_complex), intent(in) :: apwalm(*)
      integer(c_int), intent(in) :: igpig(*)
      real(c_double), intent(in) :: vgpc(*)
      complex(c_double_complex), intent(inout) :: h(*), o(*)
    end function

    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function

    ! Batched k-points of one structure (scf_cycle.f90:256-296 k loop).
    integer(c_int) function exh_setup_batch_device(ctx, stream, nk, kpts, flags) &
        bind(C, name="exh_setup_batch_device")
      import :: c_int, c_ptr, exh_kpoint
      type(c_ptr), value :: ctx, stream
      integer(c_int), value :: nk, flags
      type(exh_kpoint), intent(in) :: kpts(*)
    end function

    ! y = alpha op(a) x + beta y0 + gamma z on the device, op(a) = a^H
    ! (conj_a /= 0) or a^T (the eigensolver's filter product; y0 and z may
    ! be c_null_ptr)
    integer(c_int) function exh_zgemm_cn_device(stream, conj_a, nout, m, k, a, lda, x, ldx, &
        alpha, beta, y0, ldy0, gamma, z, ldz, y, ldy, split, work) &
        bind(C, name="exh_zgemm_cn_device")
      import :: c_int, c_ptr, c_long_long, c_double
      type(c_ptr), value :: stream, a, x, y0, z, y, work
      integer(c_int), value :: conj_a, nout, m, k, split
      integer(c_long_long), value :: lda, ldx, ldy0, ldz, ldy
      real(c_double), value :: alpha, beta, gamma
    end function

    ! k-points of several structures in one grouped launch: kpts(i) belongs
    ! to the context ctxs(i).
    integer(c_int) function exh_setup_multi_device(ctxs, stream, nk, kpts, flags) &
        bind(C, name="exh_setup_multi_device")
      import :: c_int, c_ptr, exh_kpoint
      type(c_ptr), intent(in) :: ctxs(*)
      type(c_ptr), value :: stream
      integer(c_int), value :: nk, flags
      type(exh_kpoint), intent(in) :: kpts(*)
    end function

    integer(c_int) function exh_tune_batch(ctx, stream, nk, kpts, flags, plan) &
        bind(C, name="exh_tune_batch")
      import :: c_int, c_ptr, exh_kpoint
      type(c_ptr), value :: ctx, stream
      integer(c_int), value :: nk, flags
      type(exh_kpoint), intent(in) :: kpts(*)
      integer(c_int), intent(out) :: plan
    end function

    integer(c_int) function exh_set_batch_plan(ctx, plan) bind(C, name="exh_set_batch_plan")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: plan
    end function

    integer(c_int) function exh_get_batch_plan(ctx) bind(C, name="exh_get_batch_plan")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function exh_set_batch_workspace(ctx, nbytes) &
        bind(C, name="exh_set_batch_workspace")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx
      integer(c_size_t), value :: nbytes
    end function
  end interface

contains

  !> The library's last error message as a Fortran string.
  function exh_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = exh_error_c(ctx)
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate (character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function

  ! EXH_SPLIT(n) of include/exh.h: split-precision flag with n int8 moduli.
  pure integer(c_int) function exh_split(n)
    integer, intent(in) :: n
    exh_split = ishft(int(n, c_int), 8)
  end function

end module exh_lapw
