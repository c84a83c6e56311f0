! Drop-in replacements, under the reference's names and with its dummy
! arguments, for the three routines of exciting's first-variational setup:
!
!   match (ngp, gpc, tpgpc, sfacgp, apwalm)                  src/match.f90:13
!   hamiltonandoverlapsetup (system, ngp, apwalm, igpig, vgpc)
!         src/src_eigensystem/hamiltonandoverlapsetup.f90:8, with the
!         semantics of hamiltonsetup.f90:8 + overlapsetup.f90:8 (incl. the
!         UL copy) as seceqnfv.f90:76-77 call them back to back
!   overlapsetup (system, ngp, apwalm, igpig, vgpc)
!         src/src_eigensystem/overlapsetup.f90:8, called alone by
!         src_hybrids/calc_vnlmat.f90:60
!
! External subroutines with implicit interfaces, as in the reference: an
! exciting build links this object in place of the reference's and changes
! seceqnfv.f90:76-77 into one call of hamiltonandoverlapsetup.  They run on
! the GPU through libexh; the context and the array bounds come from module
! exh_state, the matrix types from exciting's modfvsystem.  H and O are
! accumulated into (zgemm beta = 1 and += in the reference); the caller's
! newsystem has zeroed them.  Packed storage is refused: the reference's
! live setup writes system%hamilton%za itself (hamiltonsetup.f90:90-103,
! 143-161), which packed storage leaves unallocated (modfvsystem.f90:62-66),
! so that path has no packed form to reproduce.

subroutine match(ngp, gpc, tpgpc, sfacgp, apwalm)
  use exh_lapw, only: exh_match
  use exh_state, only: exh_ctx, exh_state_check, ngkmax, apwordmax, lmmaxapw, natmtot
  implicit none
  integer, intent(in) :: ngp
  real(8), intent(in) :: gpc(ngkmax), tpgpc(2, ngkmax)
  complex(8), intent(in) :: sfacgp(ngkmax, natmtot)
  complex(8), intent(out) :: apwalm(ngkmax, apwordmax, lmmaxapw, natmtot)
  call exh_state_check(exh_match(exh_ctx, ngp, gpc, tpgpc, sfacgp, apwalm), 'match')
end subroutine

subroutine hamiltonandoverlapsetup(system, ngp, apwalm, igpig, vgpc)
  use modfvsystem, only: evsystem
  use exh_lapw, only: exh_hamiltonandoverlapsetup
  use exh_state, only: exh_ctx, exh_state_check, ngkmax, apwordmax, lmmaxapw, natmtot
  implicit none
  type(evsystem) :: system
  integer, intent(in) :: ngp
  complex(8), intent(in) :: apwalm(ngkmax, apwordmax, lmmaxapw, natmtot)
  integer, intent(in) :: igpig(ngkmax)
  real(8), intent(in) :: vgpc(3, ngkmax)
  if (system%hamilton%packed .or. system%overlap%packed) then
    print '("Error(hamiltonandoverlapsetup): packed matrix storage is not supported")'
    error stop 1
  end if
  call exh_state_check(exh_hamiltonandoverlapsetup(exh_ctx, ngp, system%hamilton%rank, apwalm, &
                                                   igpig, vgpc, system%hamilton%za, &
                                                   system%overlap%za), &
                       'hamiltonandoverlapsetup')
end subroutine

subroutine overlapsetup(system, ngp, apwalm, igpig, vgpc)
  use modfvsystem, only: evsystem
  use exh_lapw, only: exh_overlapsetup
  use exh_state, only: exh_ctx, exh_state_check, ngkmax, apwordmax, lmmaxapw, natmtot
  implicit none
  type(evsystem) :: system
  integer, intent(in) :: ngp
  complex(8), intent(in) :: apwalm(ngkmax, apwordmax, lmmaxapw, natmtot)
  integer, intent(in) :: igpig(ngkmax)
  real(8), intent(in) :: vgpc(3, ngkmax)
  if (system%overlap%packed) then
    print '("Error(overlapsetup): packed matrix storage is not supported")'
    error stop 1
  end if
  call exh_state_check(exh_overlapsetup(exh_ctx, ngp, system%overlap%rank, apwalm, igpig, &
                                        vgpc, system%hamilton%za, system%overlap%za), &
                       'overlapsetup')
end subroutine
