! The integration's own module state for the drop-in shims: one libexh
! context per process and the array bounds the reference's dummy arguments
! take from its modules (ngkmax from mod_Gkvector, apwordmax from
! mod_APW_LO, lmmaxapw from mod_muffin_tin, natmtot from mod_atoms).
!
! exciting calls exh_state_init once (e.g. after init0), then
! exh_state_set_basis after init1 (init1.f90:343-481) and the exh_set_*
! setters, passing exh_ctx, after the producers of each SCF iteration:
! genapwfr (scf_cycle.f90:210), olprad + hmlint (:212, 221-222), genveffig
! (:80, 421).  The shims read exh_ctx and the bounds from here.
module exh_state
  use, intrinsic :: iso_c_binding
  use exh_lapw
  implicit none
  private
  type(c_ptr), public :: exh_ctx = c_null_ptr
  integer, public :: ngkmax = 0, apwordmax = 0, lmmaxapw = 0, natmtot = 0, nlotot = 0
  public :: exh_state_init, exh_state_set_basis, exh_state_check, exh_state_final

contains

  ! Context on GPU `device`; host-array setups reuse the apwalm that match
  ! left in HBM, since exciting's k loop calls the setup right after match
  ! on the same, unmodified array (seceqn.f90:57-59, calc_vnlmat.f90:54-60).
  subroutine exh_state_init(device)
    integer, intent(in) :: device
    integer(c_int) :: rc
    if (c_associated(exh_ctx)) return
    rc = exh_create(int(device, c_int), exh_ctx)
    if (rc /= 0) then
      print '("Error(exh_state_init): exh_create failed with code ", i0)', rc
      error stop 1
    end if
    call exh_state_check(exh_set_apwalm_reuse(exh_ctx, 1_c_int), 'exh_state_init')
  end subroutine

  ! exh_set_basis, recording the bounds the shims declare their dummies with.
  subroutine exh_state_set_basis(nspecies, natoms, lmaxapw, lmaxmat, apwordmax_, apword, &
                                 nlomax, nlorb, lorbl, ngkmax_)
    integer, intent(in) :: nspecies, lmaxapw, lmaxmat, apwordmax_, nlomax, ngkmax_
    integer, intent(in) :: natoms(nspecies), apword(0:lmaxapw, nspecies), nlorb(nspecies)
    integer, intent(in) :: lorbl(max(nlomax, 1), nspecies)
    integer :: is, ilo
    integer(c_int) :: rc
    rc = exh_set_basis(exh_ctx, int(nspecies, c_int), int(natoms, c_int), int(lmaxapw, c_int), &
                       int(lmaxmat, c_int), int(apwordmax_, c_int), int(apword, c_int), &
                       int(nlomax, c_int), int(nlorb, c_int), int(lorbl, c_int), &
                       int(ngkmax_, c_int))
    call exh_state_check(rc, 'exh_set_basis')
    ngkmax = ngkmax_
    apwordmax = apwordmax_
    lmmaxapw = (lmaxapw + 1)**2
    natmtot = sum(natoms)
    nlotot = 0
    do is = 1, nspecies
      do ilo = 1, nlorb(is)
        nlotot = nlotot + natoms(is) * (2 * lorbl(ilo, is) + 1)
      end do
    end do
  end subroutine

  ! Stops with libexh's message on a nonzero return code (the reference's
  ! routines stop through terminate, modmpi.F90:202).
  subroutine exh_state_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= 0) then
      print '("Error(", a, "): libexh code ", i0, ": ", a)', where, rc, &
        exh_error_message(exh_ctx)
      error stop 1
    end if
  end subroutine

  subroutine exh_state_final()
    if (c_associated(exh_ctx)) call exh_destroy(exh_ctx)
    exh_ctx = c_null_ptr
  end subroutine

end module exh_state
