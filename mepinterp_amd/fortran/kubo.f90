! kubo.f90 -- kubo_ahc_tens, velo_ahc_tens, kubo_ohc_tens with the
! reference's signatures (kubo.f90:37-198), evaluated on the GPU.  hw_lst is
! the reference's get_linspace list, regenerated on the device from its ends.
module kubo
  use constants, only: dp
  use mep_capi, only: MEP_JOB_AHC
  use mep_gpu, only: gpu_bandsum_k
  implicit none
  private
  public :: kubo_ahc_tens, velo_ahc_tens, kubo_ohc_tens
contains
  function kubo_ahc_tens(en_k, V_ka, fd_distrib) result(o_ahc)
    real(dp), intent(in) :: en_k(:), fd_distrib(:, :)
    complex(dp), intent(in) :: V_ka(:, :, :)
    real(dp), allocatable :: o_ahc(:, :, :)
    real(dp), allocatable :: ahc(:, :, :), kic(:, :, :), klc(:, :, :), kcs(:, :, :), &
                             nic(:, :, :), nlc(:, :, :), ncs(:, :, :)
    complex(dp), allocatable :: velo(:, :, :, :), ohc(:, :, :, :)
    call gpu_bandsum_k(en_k, V_ka, fd_distrib, int(MEP_JOB_AHC), [0.0_dp], 0.0_dp, 1, ahc, velo, &
                       ohc, kic, klc, kcs, nic, nlc, ncs)
    o_ahc = ahc
  end function kubo_ahc_tens

  function velo_ahc_tens(en_k, V_ka, hw_lst, fd_distrib, i_eta_smr) result(o_ahc)
    real(dp), intent(in) :: en_k(:), hw_lst(:), fd_distrib(:, :)
    complex(dp), intent(in) :: V_ka(:, :, :), i_eta_smr
    complex(dp), allocatable :: o_ahc(:, :, :, :)
    real(dp), allocatable :: ahc(:, :, :), kic(:, :, :), klc(:, :, :), kcs(:, :, :), &
                             nic(:, :, :), nlc(:, :, :), ncs(:, :, :)
    complex(dp), allocatable :: velo(:, :, :, :), ohc(:, :, :, :)
    call gpu_bandsum_k(en_k, V_ka, fd_distrib, int(MEP_JOB_AHC), hw_lst, aimag(i_eta_smr), 1, ahc, &
                       velo, ohc, kic, klc, kcs, nic, nlc, ncs)
    o_ahc = velo
  end function velo_ahc_tens

  function kubo_ohc_tens(en_k, V_ka, hw_lst, fd_distrib, i_eta_smr) result(z_ohc)
    real(dp), intent(in) :: en_k(:), hw_lst(:), fd_distrib(:, :)
    complex(dp), intent(in) :: V_ka(:, :, :), i_eta_smr
    complex(dp), allocatable :: z_ohc(:, :, :, :)
    real(dp), allocatable :: ahc(:, :, :), kic(:, :, :), klc(:, :, :), kcs(:, :, :), &
                             nic(:, :, :), nlc(:, :, :), ncs(:, :, :)
    complex(dp), allocatable :: velo(:, :, :, :), ohc(:, :, :, :)
    call gpu_bandsum_k(en_k, V_ka, fd_distrib, int(MEP_JOB_AHC), hw_lst, aimag(i_eta_smr), 1, ahc, &
                       velo, ohc, kic, klc, kcs, nic, nlc, ncs)
    z_ohc = ohc
  end function kubo_ohc_tens
end module kubo
