! mep_niu.f90 -- mep_niu_CS / LC / IC with the reference's signatures and
! argument order (mep_niu.f90:20-159), evaluated on the GPU; valence_bands
! and kubo_tol come from input_paras as in the reference.
module mep_niu
  use constants, only: dp
  use mep_capi, only: MEP_JOB_MEP
  use mep_gpu, only: gpu_bandsum_k
  use input_paras, only: valence_bands
  implicit none
  private
  public :: mep_niu_CS, mep_niu_IC, mep_niu_LC
contains
  function mep_niu_CS(A_ka, Om_kab) result(cs_tens)
    complex(dp), allocatable, intent(in) :: A_ka(:, :, :), Om_kab(:, :, :, :)
    real(dp), allocatable :: cs_tens(:, :, :)
    real(dp), allocatable :: e(:), fd(:, :), ahc(:, :, :), kic(:, :, :), klc(:, :, :), &
                             kcs(:, :, :), nic(:, :, :), nlc(:, :, :), ncs(:, :, :)
    complex(dp), allocatable :: V(:, :, :), velo(:, :, :, :), ohc(:, :, :, :)
    integer :: n, i
    allocate(cs_tens(3, 3, valence_bands))
    cs_tens = 0.0_dp
    if (.not. (allocated(A_ka) .and. allocated(Om_kab))) return
    n = size(A_ka, 2)
    allocate(e(n), V(3, n, n), fd(1, n))
    e = [(real(i, dp), i = 1, n)]
    V = (0.0_dp, 0.0_dp)
    fd = 0.0_dp
    call gpu_bandsum_k(e, V, fd, int(MEP_JOB_MEP), [0.0_dp], 0.0_dp, valence_bands, ahc, velo, &
                       ohc, kic, klc, kcs, nic, nlc, ncs, A_ka, Om_kab)
    cs_tens = ncs
  end function mep_niu_CS

  function mep_niu_LC(V_ka, en_k) result(F3)
    complex(dp), intent(in) :: V_ka(:, :, :)
    real(dp), intent(in) :: en_k(:)
    real(dp), allocatable :: F3(:, :, :)
    real(dp), allocatable :: fd(:, :), ahc(:, :, :), kic(:, :, :), klc(:, :, :), kcs(:, :, :), &
                             nic(:, :, :), nlc(:, :, :), ncs(:, :, :)
    complex(dp), allocatable :: velo(:, :, :, :), ohc(:, :, :, :)
    allocate(fd(1, size(en_k)))
    fd = 0.0_dp
    call gpu_bandsum_k(en_k, V_ka, fd, int(MEP_JOB_MEP), [0.0_dp], 0.0_dp, valence_bands, ahc, &
                       velo, ohc, kic, klc, kcs, nic, nlc, ncs)
    F3 = nlc
  end function mep_niu_LC

  function mep_niu_IC(V_ka, en_k) result(F2)
    complex(dp), intent(in) :: V_ka(:, :, :)
    real(dp), intent(in) :: en_k(:)
    real(dp), allocatable :: F2(:, :, :)
    real(dp), allocatable :: fd(:, :), ahc(:, :, :), kic(:, :, :), klc(:, :, :), kcs(:, :, :), &
                             nic(:, :, :), nlc(:, :, :), ncs(:, :, :)
    complex(dp), allocatable :: velo(:, :, :, :), ohc(:, :, :, :)
    allocate(fd(1, size(en_k)))
    fd = 0.0_dp
    call gpu_bandsum_k(en_k, V_ka, fd, int(MEP_JOB_MEP), [0.0_dp], 0.0_dp, valence_bands, ahc, &
                       velo, ohc, kic, klc, kcs, nic, nlc, ncs)
    F2 = nic
  end function mep_niu_IC
end module mep_niu
