! main.f90 -- the mepInterp executable (main.f90:19-50 of the reference):
! read ./input.cfg, run the k-space integration on the GPU, write out/*.npy.
program mepInterp
  use input_paras, only: init_parameters, plot_bands
  use core, only: core_worker
  use mep_gpu, only: gpu_release
  implicit none
  if (.not. init_parameters()) stop "could not read input.cfg"
  if (plot_bands) then
    write(*, '(a)') "[mepInterp]: band-structure mode (jobs%plot_bands) is not part of this build"
    stop 1
  end if
  call core_worker()
  call gpu_release()
end program mepInterp
