! main.f90 -- the mepInterp executable (main.f90:19-50 of the reference):
! read ./input.cfg, run the k-space integration on the GPU, write out/*.npy
! (or, with jobs%plot_bands, the band structure along ./kpts).
program mepInterp
  use input_paras, only: init_parameters, plot_bands
  use core, only: core_worker
  use band_calc, only: band_worker
  use mep_gpu, only: gpu_release
  implicit none
  if (.not. init_parameters()) stop "could not read input.cfg"
  if (plot_bands) then
    call band_worker()
  else
    call core_worker()
  end if
  call gpu_release()
end program mepInterp
