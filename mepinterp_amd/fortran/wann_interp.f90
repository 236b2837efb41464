! wann_interp.f90 -- get_wann_interp with the reference's signature
! (wann_interp.f90:56-108): the Fourier sums, eigensolve and gauge
! transformation run on the GPU (mep_interp_batch with one k-point).
! Job selection by allocation status is kept: allocated(wf_centers) selects
! the tight-binding convention, allocated(A_ka, r_real, Om_kab) the position
! operator, allocated(V_ka) the velocities.
module wann_interp
  use iso_c_binding
  use constants, only: dp
  use mep_capi
  use mep_gpu, only: gpu_model_ctx
  implicit none
  private
  public :: get_wann_interp
contains
  subroutine get_wann_interp(do_gauge_trafo, H_real, r_real, R_frac, wf_centers, kpt_idx, kpt, &
                             e_k, V_ka, A_ka, Om_kab)
    logical, intent(in) :: do_gauge_trafo
    complex(dp), intent(in) :: H_real(:, :, :)
    complex(dp), allocatable, intent(inout) :: r_real(:, :, :, :)
    integer, intent(in) :: kpt_idx
    real(dp), intent(in) :: R_frac(:, :), kpt(3)
    real(dp), allocatable, intent(in) :: wf_centers(:, :)
    real(dp), intent(out) :: e_k(:)
    complex(dp), allocatable, intent(inout) :: V_ka(:, :, :)
    complex(dp), allocatable, intent(inout) :: A_ka(:, :, :), Om_kab(:, :, :, :)
    type(c_ptr) :: ctx, vp, ap, op
    real(dp), target :: k(3)
    real(dp), allocatable, target :: e(:)
    complex(dp), allocatable, target :: V(:, :, :), A(:, :, :), Om(:, :, :, :)
    logical :: use_pos
    integer :: n
    integer(c_int) :: rc
    if (kpt_idx < 0) stop "[get_wann_interp]: negative k-point index"
    n = size(H_real, 1)
    use_pos = allocated(A_ka) .and. allocated(r_real) .and. allocated(Om_kab) .and. allocated(V_ka)
    call gpu_model_ctx(do_gauge_trafo, H_real, r_real, use_pos, R_frac, wf_centers, ctx)
    k = kpt
    allocate(e(n))
    vp = c_null_ptr
    ap = c_null_ptr
    op = c_null_ptr
    if (allocated(V_ka)) then
      allocate(V(3, n, n))
      vp = c_loc(V)
      if (use_pos) then
        allocate(A(3, n, n), Om(3, 3, n, n))
        ap = c_loc(A)
        op = c_loc(Om)
      end if
    end if
    rc = mep_interp_batch(ctx, 1_c_int64_t, c_loc(k), c_loc(e), vp, ap, op)
    call mep_check(rc, "get_wann_interp")
    e_k = e
    if (allocated(V_ka)) V_ka = V
    if (use_pos) then
      A_ka = A
      Om_kab = Om
    end if
  end subroutine get_wann_interp
end module wann_interp
