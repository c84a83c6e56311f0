! Fortran caller of libexh through the exh_lapw module, in the order the
! reference's k-point loop uses the path: module state once, then for the
! k-point match (seceqn.f90:57, apwalm zeroed at :46-47) and the H/O setup
! (seceqnfv.f90:73-77, H and O zeroed by newsystem).
!
! Then the interstitial producers (gencfun, genveffig, genmeffig for iora*)
! from the cell's bvec, atposc and a real-space veffir, and the muffin-tin
! table producer (hmlint) from veffmt and lofr.
!
! Usage: exh_driver <input.bin> <output.bin>
! The input is a stream of int32/float64/complex128 records written by
! tests/test_gpu_fortran.py; the output holds apwalm, H and O, then cfunig,
! cfunir, veffig, meffig and m2effig, then the hmlint tables aa, loa, lolo.
program exh_driver
  use, intrinsic :: iso_c_binding
  use exh_lapw
  implicit none
  character(len=1024) :: fin, fout
  integer :: u, rc
  integer(c_int) :: nspecies, natmtot, lmaxapw, lmaxmat, apwordmax, nlomax, ngkmax
  integer(c_int) :: nrmtmax, maxaa, maxnlo, ngrtot, ngvec, ngp, nbox, nlotot, rank
  integer(c_int), allocatable :: natoms(:), apword(:,:), nlorb(:), lorbl(:,:), nrmt(:)
  integer(c_int), allocatable :: ivg(:,:), ivgig(:), igpig(:)
  integer(c_int) :: intgv(3, 2)
  real(c_double) :: omega
  real(c_double), allocatable :: rmt(:), spr(:,:), apwfr(:,:,:,:,:), gpc(:), tpgpc(:,:)
  real(c_double), allocatable, target :: oalo(:,:,:), ololo(:,:,:)
  real(c_double), allocatable :: vgpc(:,:)
  complex(c_double_complex), allocatable, target :: aa(:,:,:), loa(:,:,:), lolo(:,:,:)
  complex(c_double_complex), allocatable, target :: cfunig(:), veffig(:)
  complex(c_double_complex), allocatable :: sfacgp(:,:), apwalm(:,:,:,:), h(:,:), o(:,:)
  integer(c_int) :: maxatoms
  real(c_double) :: bvec(3, 3)
  real(c_double), allocatable :: atposc(:,:,:), veffir(:), cfunir(:)
  complex(c_double_complex), allocatable :: cfunig_p(:), veffig_p(:), meffig(:), m2effig(:)
  integer(c_int) :: lmaxvr
  real(c_double), allocatable :: veffmt(:,:,:), lofr(:,:,:,:)
  complex(c_double_complex), allocatable :: aa2(:,:,:), loa2(:,:,:), lolo2(:,:,:)
  type(c_ptr) :: ctx

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read (u) nspecies, natmtot, lmaxapw, lmaxmat, apwordmax, nlomax, ngkmax, nrmtmax, &
           maxaa, maxnlo, ngrtot, ngvec, ngp, nbox, nlotot
  allocate (natoms(nspecies), apword(0:lmaxapw, nspecies), nlorb(nspecies))
  allocate (lorbl(nlomax, nspecies), nrmt(nspecies), rmt(nspecies), spr(nrmtmax, nspecies))
  allocate (apwfr(nrmtmax, 2, apwordmax, 0:lmaxapw, natmtot))
  allocate (aa(maxaa, maxaa, natmtot), loa(maxnlo, maxaa, natmtot), lolo(maxnlo, maxnlo, natmtot))
  allocate (oalo(apwordmax, nlomax, natmtot), ololo(nlomax, nlomax, natmtot))
  allocate (ivg(3, ngrtot), ivgig(nbox), cfunig(ngrtot), veffig(ngvec))
  allocate (gpc(ngkmax), tpgpc(2, ngkmax), sfacgp(ngkmax, natmtot), igpig(ngkmax), vgpc(3, ngkmax))
  read (u) natoms, apword, nlorb, lorbl, omega, rmt, nrmt, spr, apwfr
  read (u) aa, loa, lolo, oalo, ololo
  read (u) ivg, intgv, ivgig, cfunig, veffig
  read (u) gpc, tpgpc, sfacgp, igpig, vgpc
  read (u) maxatoms
  allocate (atposc(3, maxatoms, nspecies), veffir(ngrtot))
  read (u) bvec, atposc, veffir
  read (u) lmaxvr
  allocate (veffmt((lmaxvr + 1)**2, nrmtmax, natmtot), lofr(nrmtmax, 2, max(nlomax, 1), natmtot))
  read (u) veffmt, lofr
  close (u)

  rc = exh_create(0_c_int, ctx)
  call check(rc, 'exh_create')
  rc = exh_set_basis(ctx, nspecies, natoms, lmaxapw, lmaxmat, apwordmax, apword, nlomax, &
                     nlorb, lorbl, ngkmax)
  call check(rc, 'exh_set_basis')
  rc = exh_set_radial(ctx, omega, rmt, nrmt, nrmtmax, spr, apwfr)
  call check(rc, 'exh_set_radial')
  rc = exh_set_mt(ctx, maxaa, maxnlo, c_loc(aa), c_loc(loa), c_loc(lolo), c_loc(oalo), &
                  c_loc(ololo), c_null_ptr, c_null_ptr, c_null_ptr)
  call check(rc, 'exh_set_mt')
  rc = exh_set_gvectors(ctx, ngrtot, ngvec, ivg, intgv, ivgig)
  call check(rc, 'exh_set_gvectors')
  rc = exh_set_interstitial(ctx, c_loc(cfunig), c_loc(veffig), c_null_ptr, c_null_ptr, 1_c_int, &
                            0_c_int)
  call check(rc, 'exh_set_interstitial')

  ! seceqn.f90:46-47 and :57
  allocate (apwalm(ngkmax, apwordmax, (lmaxapw + 1)**2, natmtot))
  apwalm = (0.d0, 0.d0)
  rc = exh_match(ctx, ngp, gpc, tpgpc, sfacgp, apwalm)
  call check(rc, 'exh_match')
  ! seceqnfv.f90:73-77 (newsystem zero-fills, then the setup accumulates)
  rank = ngp + nlotot
  allocate (h(rank, rank), o(rank, rank))
  h = (0.d0, 0.d0)
  o = (0.d0, 0.d0)
  rc = exh_hamiltonandoverlapsetup(ctx, ngp, rank, apwalm, igpig, vgpc, h, o)
  call check(rc, 'exh_hamiltonandoverlapsetup')

  ! gencfun.f90:13 (epslat 1e-6, no damping), genveffig.f90:13, genmeffig.f90:13
  allocate (cfunig_p(ngrtot), cfunir(ngrtot), veffig_p(ngvec), meffig(ngvec), m2effig(ngvec))
  rc = exh_gencfun(ctx, bvec, atposc, maxatoms, 1.d-6, 0.d0, 12.d0, cfunig_p, cfunir)
  call check(rc, 'exh_gencfun')
  rc = exh_genveffig(ctx, veffir, cfunir, veffig_p)
  call check(rc, 'exh_genveffig')
  rc = exh_genmeffig(ctx, veffir, cfunir, 0_c_int, 1_c_int, meffig, m2effig)
  call check(rc, 'exh_genmeffig')

  ! hmlint.f90:53-55 into zeroed tables (MTInitAll zeroes mt_hscf before it)
  allocate (aa2(maxaa, maxaa, natmtot), loa2(maxnlo, maxaa, natmtot), lolo2(maxnlo, maxnlo, natmtot))
  aa2 = (0.d0, 0.d0)
  loa2 = (0.d0, 0.d0)
  lolo2 = (0.d0, 0.d0)
  rc = exh_hmlint(ctx, lmaxvr, veffmt, lofr, 1_c_int, aa2, loa2, lolo2, EXH_ACCUMULATE)
  call check(rc, 'exh_hmlint')
  call exh_destroy(ctx)

  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (u) apwalm, h, o, cfunig_p, cfunir, veffig_p, meffig, m2effig, aa2, loa2, lolo2
  close (u)
  print '(a,i0)', 'exh_driver: done, rank = ', rank

contains

  subroutine check(code, what)
    integer, intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= 0) then
      print '(a,a,i0,a,a)', what, ' failed with code ', code, ': ', exh_error_message(ctx)
      error stop 1
    end if
  end subroutine

end program exh_driver
