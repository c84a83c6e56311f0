! The caller's matrix types for the standalone build of the drop-in shims.
!
! In exciting the shims (exh_shims.f90) compile against exciting's own
! module modfvsystem (src/src_eigensystem/modfvsystem.f90:10-29), whose
! HermitianMatrix and evsystem types they receive from seceqnfv and
! calc_vnlmat.  This file declares a module of the same name with the same
! components, so that the shims build and are tested here without the rest
! of exciting; it is not linked into exciting.  newsystem/deletesystem give
! zero-filled dense (or packed) storage of the requested order, as the
! caller's do before the setup (seceqnfv.f90:73).
module modfvsystem
  implicit none
  private
  public :: HermitianMatrix, evsystem, newsystem, deletesystem

  type HermitianMatrix
    integer :: rank = 0
    logical :: packed = .false., ludecomposed = .false., sp = .false.
    integer, pointer :: ipiv(:) => null()
    complex(8), pointer :: za(:,:) => null(), zap(:) => null()
    complex(8), pointer :: eigenvalues(:) => null()
    complex(4), pointer :: ca(:,:) => null(), cap(:) => null()
    complex(4), pointer :: eigenvaluessp(:) => null()
  end type

  type evsystem
    type(HermitianMatrix) :: hamilton, overlap
  end type

contains

  subroutine alloc_matrix(m, packed, rank)
    type(HermitianMatrix), intent(inout) :: m
    logical, intent(in) :: packed
    integer, intent(in) :: rank
    m%rank = rank
    m%packed = packed
    m%ludecomposed = .false.
    m%sp = .false.
    if (packed) then
      allocate (m%zap(rank * (rank + 1) / 2))
      m%zap = (0.d0, 0.d0)
    else
      allocate (m%za(rank, rank))
      m%za = (0.d0, 0.d0)
    end if
  end subroutine

  subroutine newsystem(self, packed, rank)
    type(evsystem), intent(out) :: self
    logical, intent(in) :: packed
    integer, intent(in) :: rank
    call alloc_matrix(self%hamilton, packed, rank)
    call alloc_matrix(self%overlap, packed, rank)
  end subroutine

  subroutine free_matrix(m)
    type(HermitianMatrix), intent(inout) :: m
    if (associated(m%za)) deallocate (m%za)
    if (associated(m%zap)) deallocate (m%zap)
  end subroutine

  subroutine deletesystem(self)
    type(evsystem), intent(inout) :: self
    call free_matrix(self%hamilton)
    call free_matrix(self%overlap)
  end subroutine

end module modfvsystem
