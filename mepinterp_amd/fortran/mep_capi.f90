! mep_capi.f90 -- iso_c_binding interface to include/mepinterp.h.
module mep_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: MEP_OK = 0
  integer(c_int32_t), parameter :: MEP_F_GAUGE_TRAFO = 1, MEP_F_CART_VELO = 2, &
                                   MEP_F_TEXTBOOK = 4, MEP_F_APPLY_DEG = 8, &
                                   MEP_F_SEPARATE_VELO = 16
  integer(c_int32_t), parameter :: MEP_JOB_AHC = 1, MEP_JOB_KUBO = 2, MEP_JOB_MEP = 4, &
                                   MEP_JOB_OPT = 8, MEP_JOB_GYRO = 16
  integer(c_int32_t), parameter :: MEP_PREC_FP64 = 0, MEP_PREC_MIXED = 1

  type, bind(C) :: mep_run_params
    integer(c_int32_t) :: jobs
    integer(c_int32_t) :: n_eF
    real(c_double) :: eF_min, eF_max
    integer(c_int32_t) :: n_hw
    real(c_double) :: hw_min, hw_max
    real(c_double) :: T_kelvin
    real(c_double) :: eta
    integer(c_int32_t) :: valence_bands
    integer(c_int32_t) :: precision
    type(c_ptr) :: hw_lst = c_null_ptr
  end type mep_run_params

  type, bind(C) :: mep_sums
    type(c_ptr) :: Ne = c_null_ptr, ahc = c_null_ptr, ahc_velo = c_null_ptr, ohc = c_null_ptr
    type(c_ptr) :: kubo_ic = c_null_ptr, kubo_lc = c_null_ptr, kubo_cs = c_null_ptr
    type(c_ptr) :: niu_ic = c_null_ptr, niu_lc = c_null_ptr, niu_cs = c_null_ptr
    integer(c_int64_t) :: n_k = 0
    type(c_ptr) :: opt_s = c_null_ptr, opt_a = c_null_ptr, photo = c_null_ptr
    type(c_ptr) :: gyro_C = c_null_ptr, gyro_D = c_null_ptr, gyro_Dw = c_null_ptr
  end type mep_sums

  interface
    integer(c_int) function mep_ctx_create(n_wf, n_R, R_frac, wf_centers, H_real, r_real, deg, &
                                           a_latt, kubo_tol, flags, ctx) bind(C, name="mep_ctx_create")
      import :: c_int, c_int64_t, c_int32_t, c_ptr, c_double
      integer(c_int64_t), value :: n_wf, n_R
      type(c_ptr), value :: R_frac, wf_centers, H_real, r_real, deg, a_latt
      real(c_double), value :: kubo_tol
      integer(c_int32_t), value :: flags
      type(c_ptr) :: ctx
    end function

    integer(c_int) function mep_ctx_destroy(ctx) bind(C, name="mep_ctx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function mep_last_error() bind(C, name="mep_last_error")
      import :: c_ptr
    end function

    integer(c_int) function mep_interp_batch(ctx, nk, kpts, e_k, V_ka, A_ka, Om_kab) &
        bind(C, name="mep_interp_batch")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nk
      type(c_ptr), value :: kpts, e_k, V_ka, A_ka, Om_kab
    end function

    integer(c_int) function mep_bandsum_batch(ctx, nk, e_k, V_ka, A_ka, Om_kab, fd_distrib, p, out) &
        bind(C, name="mep_bandsum_batch")
      import :: c_int, c_int64_t, c_ptr, mep_run_params, mep_sums
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nk
      type(c_ptr), value :: e_k, V_ka, A_ka, Om_kab, fd_distrib
      type(mep_run_params), intent(in) :: p
      type(mep_sums), intent(inout) :: out
    end function

    integer(c_int) function mep_run_mesh(ctx, mp_grid, p, out) bind(C, name="mep_run_mesh")
      import :: c_int, c_int32_t, c_ptr, mep_run_params, mep_sums
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(in) :: mp_grid(3)
      type(mep_run_params), intent(in) :: p
      type(mep_sums), intent(inout) :: out
    end function

    integer(c_int) function mep_run_kpts(ctx, nk, kpts, weights, p, out) bind(C, name="mep_run_kpts")
      import :: c_int, c_int64_t, c_ptr, mep_run_params, mep_sums
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nk
      type(c_ptr), value :: kpts, weights
      type(mep_run_params), intent(in) :: p
      type(mep_sums), intent(inout) :: out
    end function

    integer(c_int) function mep_set_kblock(ctx, kb) bind(C, name="mep_set_kblock")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: kb
    end function

    function c_strlen(s) bind(C, name="strlen") result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function
  end interface

contains

  ! error stop with the library's message when rc /= MEP_OK
  subroutine mep_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=:), allocatable :: text
    integer :: i, n
    if (rc == MEP_OK) return
    n = int(c_strlen(mep_last_error()))
    call c_f_pointer(mep_last_error(), msg, [max(n, 1)])
    allocate(character(len=n) :: text)
    do i = 1, n
      text(i:i) = msg(i)
    end do
    write(*, '(a,a,a,i0,a,a)') "[", where, "]: GPU call failed (rc=", rc, "): ", text
    error stop "mepinterp_amd: GPU call failed"
  end subroutine mep_check

end module mep_capi
