! exciting's calling sequence through the drop-in shims, by the reference's
! names: module state through exh_state, then for one k-point
!   match (seceqn.f90:57, apwalm zeroed at :46-47),
!   newsystem + hamiltonandoverlapsetup (seceqnfv.f90:73-77, one call),
!   newsystem + overlapsetup alone (calc_vnlmat.f90:59-60),
!   hamiltonandoverlapsetup once more on the first system (it accumulates).
!
! Usage: exh_shim_driver <input.bin> <output.bin>
! The input is the record stream tests/test_gpu_fortran.py writes for
! exh_driver; the output holds apwalm, H and O, the O of overlapsetup alone,
! H and O after the second call, and the host apwalm uploads made and
! skipped (two int64).
program exh_shim_driver
  use, intrinsic :: iso_c_binding
  use exh_lapw
  use exh_state
  use modfvsystem, only: evsystem, newsystem, deletesystem
  implicit none
  character(len=1024) :: fin, fout
  integer :: u
  integer(c_int) :: nspecies, natmtot_, lmaxapw, lmaxmat, apwordmax_, nlomax, ngkmax_
  integer(c_int) :: nrmtmax, maxaa, maxnlo, ngrtot, ngvec, ngp, nbox, nlotot_
  integer(c_int), allocatable :: natoms(:), apword(:,:), nlorb(:), lorbl(:,:), nrmt(:)
  integer(c_int), allocatable :: ivg(:,:), ivgig(:), igpig(:)
  integer(c_int) :: intgv(3, 2)
  integer(c_long_long) :: uploads, reused
  real(c_double) :: omega
  real(c_double), allocatable :: rmt(:), spr(:,:), apwfr(:,:,:,:,:), gpc(:), tpgpc(:,:)
  real(c_double), allocatable, target :: oalo(:,:,:), ololo(:,:,:)
  real(c_double), allocatable :: vgpc(:,:)
  complex(c_double_complex), allocatable, target :: aa(:,:,:), loa(:,:,:), lolo(:,:,:)
  complex(c_double_complex), allocatable, target :: cfunig(:), veffig(:)
  complex(c_double_complex), allocatable :: sfacgp(:,:), apwalm(:,:,:,:)
  type(evsystem) :: system, olp

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read (u) nspecies, natmtot_, lmaxapw, lmaxmat, apwordmax_, nlomax, ngkmax_, nrmtmax, &
           maxaa, maxnlo, ngrtot, ngvec, ngp, nbox, nlotot_
  allocate (natoms(nspecies), apword(0:lmaxapw, nspecies), nlorb(nspecies))
  allocate (lorbl(nlomax, nspecies), nrmt(nspecies), rmt(nspecies), spr(nrmtmax, nspecies))
  allocate (apwfr(nrmtmax, 2, apwordmax_, 0:lmaxapw, natmtot_))
  allocate (aa(maxaa, maxaa, natmtot_), loa(maxnlo, maxaa, natmtot_))
  allocate (lolo(maxnlo, maxnlo, natmtot_))
  allocate (oalo(apwordmax_, nlomax, natmtot_), ololo(nlomax, nlomax, natmtot_))
  allocate (ivg(3, ngrtot), ivgig(nbox), cfunig(ngrtot), veffig(ngvec))
  allocate (gpc(ngkmax_), tpgpc(2, ngkmax_), sfacgp(ngkmax_, natmtot_), igpig(ngkmax_))
  allocate (vgpc(3, ngkmax_))
  read (u) natoms, apword, nlorb, lorbl, omega, rmt, nrmt, spr, apwfr
  read (u) aa, loa, lolo, oalo, ololo
  read (u) ivg, intgv, ivgig, cfunig, veffig
  read (u) gpc, tpgpc, sfacgp, igpig, vgpc
  close (u)

  ! module state, as exciting would set it after init1 and the producers
  call exh_state_init(0)
  call exh_state_set_basis(nspecies, natoms, lmaxapw, lmaxmat, apwordmax_, apword, nlomax, &
                           nlorb, lorbl, ngkmax_)
  call exh_state_check(exh_set_radial(exh_ctx, omega, rmt, nrmt, nrmtmax, spr, apwfr), &
                       'exh_set_radial')
  call exh_state_check(exh_set_mt(exh_ctx, maxaa, maxnlo, c_loc(aa), c_loc(loa), c_loc(lolo), &
                                  c_loc(oalo), c_loc(ololo), c_null_ptr, c_null_ptr, &
                                  c_null_ptr), 'exh_set_mt')
  call exh_state_check(exh_set_gvectors(exh_ctx, ngrtot, ngvec, ivg, intgv, ivgig), &
                       'exh_set_gvectors')
  call exh_state_check(exh_set_interstitial(exh_ctx, c_loc(cfunig), c_loc(veffig), c_null_ptr, &
                                            c_null_ptr, 1_c_int, 0_c_int), &
                       'exh_set_interstitial')
  if (nlotot /= nlotot_ .or. natmtot /= natmtot_) error stop 'basis bounds differ'

  ! seceqn.f90:46-47, 57 and seceqnfv.f90:73-77
  allocate (apwalm(ngkmax, apwordmax, lmmaxapw, natmtot))
  apwalm = (0.d0, 0.d0)
  call match(ngp, gpc, tpgpc, sfacgp, apwalm)
  call newsystem(system, .false., ngp + nlotot)
  call hamiltonandoverlapsetup(system, ngp, apwalm, igpig, vgpc)
  ! calc_vnlmat.f90:59-60
  call newsystem(olp, .false., ngp + nlotot)
  call overlapsetup(olp, ngp, apwalm, igpig, vgpc)

  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (u) apwalm, system%hamilton%za, system%overlap%za, olp%overlap%za
  call hamiltonandoverlapsetup(system, ngp, apwalm, igpig, vgpc)
  call exh_state_check(exh_apwalm_transfers(exh_ctx, uploads, reused), 'exh_apwalm_transfers')
  write (u) system%hamilton%za, system%overlap%za, uploads, reused
  close (u)
  call deletesystem(system)
  call deletesystem(olp)
  call exh_state_final()
  print '("exh_shim_driver: done, apwalm uploads ", i0, ", reused ", i0)', uploads, reused
end program exh_shim_driver
