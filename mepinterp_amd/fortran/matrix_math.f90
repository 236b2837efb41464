! matrix_math.f90 -- the small helpers of the reference's matrix_math module
! that the k-space host needs (matrix_math.f90:74-139, 314-336).
module matrix_math
  use constants, only: dp
  implicit none
  private
  public :: get_linspace, get_levi_civita, crossP
contains
  pure subroutine get_linspace(min, max, n_mesh, linspace)
    real(dp), intent(in) :: min, max
    integer, intent(in) :: n_mesh
    real(dp), allocatable, intent(out) :: linspace(:)
    real(dp) :: delta
    integer :: idx
    if (n_mesh > 1) then
      allocate(linspace(n_mesh))
      delta = (max - min) / real(n_mesh - 1, dp)
      do idx = 1, n_mesh
        linspace(idx) = min + real(idx - 1, dp) * delta
      end do
    else
      allocate(linspace(1))
      linspace = min
    end if
  end subroutine get_linspace

  pure subroutine get_levi_civita(lc_tens)
    integer, intent(out) :: lc_tens(3, 3, 3)
    lc_tens = 0
    lc_tens(1, 2, 3) = 1
    lc_tens(2, 3, 1) = 1
    lc_tens(3, 1, 2) = 1
    lc_tens(1, 3, 2) = -1
    lc_tens(3, 2, 1) = -1
    lc_tens(2, 1, 3) = -1
  end subroutine get_levi_civita

  pure function crossP(a, b) result(c)
    real(dp), intent(in) :: a(3), b(3)
    real(dp) :: c(3)
    c(1) = a(2) * b(3) - a(3) * b(2)
    c(2) = a(3) * b(1) - a(1) * b(3)
    c(3) = a(1) * b(2) - a(2) * b(1)
  end function crossP
end module matrix_math
