! statistics.f90 -- Fermi-Dirac occupations and electron count
! (statistics.f90:41-124 of the reference).  Host-side helpers for callers
! of the per-k interface; the mesh path evaluates them on the GPU.
module statistics
  use constants, only: dp, kBoltz_Eh_K
  implicit none
  private
  public :: fd_stat, fd_get_N_el
  real(dp), parameter :: min_temp = 1e-2_dp

  interface fd_stat
    module procedure d0_fd_stat, d1_fd_stat
  end interface fd_stat
  interface fd_get_N_el
    module procedure d1_fd_get_N_el
  end interface fd_get_N_el
contains
  real(dp) pure function d0_fd_stat(e_band, e_fermi, T_kelvin)
    real(dp), intent(in) :: e_band, e_fermi, T_kelvin
    d0_fd_stat = 0.0_dp
    if (T_kelvin > min_temp) then
      d0_fd_stat = 1.0_dp / (1.0_dp + exp((e_band - e_fermi) / (kBoltz_Eh_K * T_kelvin)))
    else if (e_band < e_fermi) then
      d0_fd_stat = 1.0_dp
    end if
  end function d0_fd_stat

  pure function d1_fd_stat(e_band, ef_lst, T_kelvin) result(fd_lst)
    real(dp), intent(in) :: e_band, ef_lst(:), T_kelvin
    real(dp), allocatable :: fd_lst(:)
    integer :: i
    allocate(fd_lst(size(ef_lst)))
    do i = 1, size(ef_lst)
      fd_lst(i) = d0_fd_stat(e_band, ef_lst(i), T_kelvin)
    end do
  end function d1_fd_stat

  pure function d1_fd_get_N_el(en_k, ef_lst, T_kelvin) result(N_el_lst)
    real(dp), intent(in) :: en_k(:), ef_lst(:), T_kelvin
    real(dp), allocatable :: N_el_lst(:)
    integer :: n
    allocate(N_el_lst(size(ef_lst)))
    N_el_lst = 0.0_dp
    do n = 1, size(en_k)
      N_el_lst = N_el_lst + d1_fd_stat(en_k(n), ef_lst, T_kelvin)
    end do
  end function d1_fd_get_N_el
end module statistics
