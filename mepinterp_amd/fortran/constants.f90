! constants.f90 -- kinds, mathematical and physical constants with the
! reference's values (constants.f90:6-23 of MEPinterp).
module constants
  implicit none
  integer, parameter :: dp = kind(0.d0)
  real(dp), parameter :: fp_acc = 1e-14_dp
  real(dp), parameter :: pi_dp = 4 * atan(1.0_dp)
  complex(dp), parameter :: i_dp = cmplx(0.0_dp, 1.0_dp, dp)
  real(dp), parameter :: aUtoAngstrm = 0.52917721092_dp
  real(dp), parameter :: aUtoEv = 27.211385_dp
  real(dp), parameter :: kBoltz_Eh_K = 8.6173303_dp * 1e-5_dp / aUtoEv
end module constants
