! kubo_mep.f90 -- kubo_mep_CS / LC / IC with the reference's signatures
! (kubo_mep.f90:16-149), evaluated on the GPU.
module kubo_mep
  use constants, only: dp
  use mep_capi, only: MEP_JOB_KUBO
  use mep_gpu, only: gpu_bandsum_k
  implicit none
  private
  public :: kubo_mep_CS, kubo_mep_LC, kubo_mep_IC
contains
  function kubo_mep_CS(A_ka, Om_kab, fd_distrib) result(cs_tens)
    complex(dp), allocatable, intent(in) :: A_ka(:, :, :), Om_kab(:, :, :, :)
    real(dp), intent(in) :: fd_distrib(:, :)
    real(dp), allocatable :: cs_tens(:, :, :)
    real(dp), allocatable :: e(:), ahc(:, :, :), kic(:, :, :), klc(:, :, :), kcs(:, :, :), &
                             nic(:, :, :), nlc(:, :, :), ncs(:, :, :)
    complex(dp), allocatable :: V(:, :, :), velo(:, :, :, :), ohc(:, :, :, :)
    integer :: n, i
    allocate(cs_tens(3, 3, size(fd_distrib, 1)))
    cs_tens = 0.0_dp
    if (.not. (allocated(A_ka) .and. allocated(Om_kab))) return
    n = size(A_ka, 2)
    allocate(e(n), V(3, n, n))
    e = [(real(i, dp), i = 1, n)]   ! distinct energies; V = 0 leaves IC = LC = 0
    V = (0.0_dp, 0.0_dp)
    call gpu_bandsum_k(e, V, fd_distrib, int(MEP_JOB_KUBO), [0.0_dp], 0.0_dp, 1, ahc, velo, ohc, &
                       kic, klc, kcs, nic, nlc, ncs, A_ka, Om_kab)
    cs_tens = kcs
  end function kubo_mep_CS

  function kubo_mep_LC(en_k, V_ka, fd_distrib) result(F3)
    real(dp), intent(in) :: en_k(:), fd_distrib(:, :)
    complex(dp), intent(in) :: V_ka(:, :, :)
    real(dp), allocatable :: F3(:, :, :)
    real(dp), allocatable :: ahc(:, :, :), kic(:, :, :), klc(:, :, :), kcs(:, :, :), &
                             nic(:, :, :), nlc(:, :, :), ncs(:, :, :)
    complex(dp), allocatable :: velo(:, :, :, :), ohc(:, :, :, :)
    call gpu_bandsum_k(en_k, V_ka, fd_distrib, int(MEP_JOB_KUBO), [0.0_dp], 0.0_dp, 1, ahc, velo, &
                       ohc, kic, klc, kcs, nic, nlc, ncs)
    F3 = klc
  end function kubo_mep_LC

  function kubo_mep_IC(en_k, V_ka, fd_distrib) result(F2)
    real(dp), intent(in) :: en_k(:), fd_distrib(:, :)
    complex(dp), intent(in) :: V_ka(:, :, :)
    real(dp), allocatable :: F2(:, :, :)
    real(dp), allocatable :: ahc(:, :, :), kic(:, :, :), klc(:, :, :), kcs(:, :, :), &
                             nic(:, :, :), nlc(:, :, :), ncs(:, :, :)
    complex(dp), allocatable :: velo(:, :, :, :), ohc(:, :, :, :)
    call gpu_bandsum_k(en_k, V_ka, fd_distrib, int(MEP_JOB_KUBO), [0.0_dp], 0.0_dp, 1, ahc, velo, &
                       ohc, kic, klc, kcs, nic, nlc, ncs)
    F2 = kic
  end function kubo_mep_IC
end module kubo_mep
