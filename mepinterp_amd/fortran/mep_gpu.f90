! mep_gpu.f90 -- GPU contexts behind the per-k interface: the model context
! used by get_wann_interp (re-created when the model changes) and the
! band-sum context used by the kubo / kubo_mep / mep_niu functions.
module mep_gpu
  use iso_c_binding
  use constants, only: dp
  use mep_capi
  use input_paras, only: a_latt, kubo_tol, use_cart_velo
  implicit none
  private
  public :: gpu_model_ctx, gpu_bandsum_k, make_run_params, gpu_release

  type(c_ptr), save :: model_ctx = c_null_ptr
  integer, save :: model_n = -1, model_nR = -1, model_flags = -1
  logical, save :: model_tb = .false., model_pos = .false.
  ! the inputs of the resident model, kept for an exact comparison
  complex(dp), allocatable, save :: model_H(:, :, :), model_rpos(:, :, :, :)
  real(dp), allocatable, save :: model_Rvec(:, :), model_wc(:, :)
  real(dp), save :: model_al(3, 3) = 0.0_dp, model_tol = -1.0_dp
  type(c_ptr), save :: band_ctx = c_null_ptr
  integer, save :: band_n = -1
  logical, save :: band_pos = .false.

contains

  subroutine gpu_release()
    integer(c_int) :: rc
    if (c_associated(model_ctx)) rc = mep_ctx_destroy(model_ctx)
    if (c_associated(band_ctx)) rc = mep_ctx_destroy(band_ctx)
    model_ctx = c_null_ptr
    band_ctx = c_null_ptr
    model_n = -1
    band_n = -1
    if (allocated(model_H)) deallocate(model_H)
    if (allocated(model_rpos)) deallocate(model_rpos)
    if (allocated(model_Rvec)) deallocate(model_Rvec)
    if (allocated(model_wc)) deallocate(model_wc)
  end subroutine gpu_release

  ! context for the model (H, optional r, R, optional centres); reused while
  ! the sizes, the flags and every input value are unchanged (compared
  ! element by element against the copies of the resident model)
  subroutine gpu_model_ctx(gauge_trafo, H_real, r_real, use_pos, R_frac, wf_centers, ctx)
    logical, intent(in) :: gauge_trafo, use_pos
    complex(dp), intent(in) :: H_real(:, :, :)
    complex(dp), allocatable, intent(in) :: r_real(:, :, :, :)
    real(dp), intent(in) :: R_frac(:, :)
    real(dp), allocatable, intent(in) :: wf_centers(:, :)
    type(c_ptr), intent(out) :: ctx
    complex(dp), allocatable, target :: Hc(:, :, :), rcopy(:, :, :, :)
    real(dp), allocatable, target :: Rvec_c(:, :), wc(:, :)
    real(dp), target :: al(3, 3)
    type(c_ptr) :: rp, wp
    integer :: n, nR, flags
    integer(c_int) :: rc
    logical :: same
    n = size(H_real, 1)
    nR = size(H_real, 3)
    flags = 0
    if (gauge_trafo) flags = ior(flags, int(MEP_F_GAUGE_TRAFO))
    if (use_cart_velo) flags = ior(flags, int(MEP_F_CART_VELO))
    same = c_associated(model_ctx) .and. n == model_n .and. nR == model_nR .and. &
           flags == model_flags .and. (use_pos .eqv. model_pos) .and. &
           (allocated(wf_centers) .eqv. model_tb)
    if (same) same = all(a_latt == model_al) .and. kubo_tol == model_tol
    if (same) same = all(R_frac == model_Rvec)
    if (same) same = all(H_real == model_H)
    if (same .and. use_pos) same = all(r_real == model_rpos)
    if (same .and. allocated(wf_centers)) same = all(wf_centers(:, 1:n) == model_wc)
    if (same) then
      ctx = model_ctx
      return
    end if
    if (c_associated(model_ctx)) rc = mep_ctx_destroy(model_ctx)
    model_ctx = c_null_ptr
    allocate(Hc(n, n, nR), Rvec_c(3, nR))
    Hc = H_real
    Rvec_c = R_frac
    al = a_latt
    rp = c_null_ptr
    wp = c_null_ptr
    if (use_pos) then
      allocate(rcopy(3, n, n, nR))
      rcopy = r_real
      rp = c_loc(rcopy)
    end if
    if (allocated(wf_centers)) then
      allocate(wc(3, n))
      wc = wf_centers(:, 1:n)
      wp = c_loc(wc)
    end if
    rc = mep_ctx_create(int(n, c_int64_t), int(nR, c_int64_t), c_loc(Rvec_c), wp, c_loc(Hc), rp, &
                        c_null_ptr, c_loc(al), real(kubo_tol, c_double), int(flags, c_int32_t), &
                        model_ctx)
    call mep_check(rc, "mep_ctx_create")
    model_n = n
    model_nR = nR
    model_flags = flags
    model_pos = use_pos
    model_tb = allocated(wf_centers)
    model_H = H_real
    model_Rvec = R_frac
    model_al = a_latt
    model_tol = kubo_tol
    if (use_pos) model_rpos = r_real
    if (allocated(wf_centers)) model_wc = wf_centers(:, 1:n)
    ctx = model_ctx
  end subroutine gpu_model_ctx

  ! context for band sums only (a one-cell zero model of the right size)
  subroutine get_band_ctx(n, pos, ctx)
    integer, intent(in) :: n
    logical, intent(in) :: pos
    type(c_ptr), intent(out) :: ctx
    complex(dp), allocatable, target :: Hc(:, :, :), rz(:, :, :, :)
    real(dp), target :: R0(3, 1), al(3, 3)
    type(c_ptr) :: rp
    integer(c_int) :: rc
    if (c_associated(band_ctx) .and. n == band_n .and. (pos .eqv. band_pos)) then
      ctx = band_ctx
      return
    end if
    if (c_associated(band_ctx)) rc = mep_ctx_destroy(band_ctx)
    band_ctx = c_null_ptr
    allocate(Hc(n, n, 1))
    Hc = cmplx(0.0_dp, 0.0_dp, dp)
    R0 = 0.0_dp
    al = a_latt
    if (all(al == 0.0_dp)) al = reshape([1.0_dp, 0.0_dp, 0.0_dp, 0.0_dp, 1.0_dp, 0.0_dp, &
                                          0.0_dp, 0.0_dp, 1.0_dp], [3, 3])
    rp = c_null_ptr
    if (pos) then
      allocate(rz(3, n, n, 1))
      rz = cmplx(0.0_dp, 0.0_dp, dp)
      rp = c_loc(rz)
    end if
    rc = mep_ctx_create(int(n, c_int64_t), 1_c_int64_t, c_loc(R0), c_null_ptr, c_loc(Hc), rp, &
                        c_null_ptr, c_loc(al), real(kubo_tol, c_double), MEP_F_GAUGE_TRAFO, band_ctx)
    call mep_check(rc, "mep_ctx_create (band sums)")
    band_n = n
    band_pos = pos
    ctx = band_ctx
  end subroutine get_band_ctx

  function make_run_params(jobs, n_ef, ef_min, ef_max, n_hw, hw_min, hw_max, T, eta, nval) result(p)
    integer, intent(in) :: jobs, n_ef, n_hw, nval
    real(dp), intent(in) :: ef_min, ef_max, hw_min, hw_max, T, eta
    type(mep_run_params) :: p
    p%jobs = int(jobs, c_int32_t)
    p%n_eF = int(max(1, n_ef), c_int32_t)
    p%eF_min = ef_min
    p%eF_max = ef_max
    p%n_hw = int(max(1, n_hw), c_int32_t)
    p%hw_min = hw_min
    p%hw_max = hw_max
    p%T_kelvin = T
    p%eta = eta
    p%valence_bands = int(max(1, nval), c_int32_t)
    p%precision = MEP_PREC_FP64
    p%hw_lst = c_null_ptr
  end function make_run_params

  ! band sums of ONE k-point with caller-given occupations fd(n_eF, n)
  subroutine gpu_bandsum_k(en_k, V_ka, fd_distrib, jobs, hw_lst, eta, nval, ahc, velo, ohc, &
                           kic, klc, kcs, nic, nlc, ncs, A_ka, Om_kab)
    real(dp), intent(in) :: en_k(:), fd_distrib(:, :), hw_lst(:), eta
    complex(dp), intent(in) :: V_ka(:, :, :)
    integer, intent(in) :: jobs, nval
    real(dp), allocatable, target, intent(out) :: ahc(:, :, :), kic(:, :, :), klc(:, :, :), &
                                                  kcs(:, :, :), nic(:, :, :), nlc(:, :, :), ncs(:, :, :)
    complex(dp), allocatable, target, intent(out) :: velo(:, :, :, :), ohc(:, :, :, :)
    complex(dp), optional, intent(in) :: A_ka(:, :, :), Om_kab(:, :, :, :)
    real(dp), allocatable, target :: e(:), fd(:, :), Ne(:), hw(:)
    complex(dp), allocatable, target :: V(:, :, :), A(:, :, :), Om(:, :, :, :)
    type(c_ptr) :: ctx, ap, op
    type(mep_run_params) :: p
    type(mep_sums) :: s
    integer :: n, nef, nhw
    logical :: pos
    integer(c_int) :: rc
    n = size(en_k)
    nef = size(fd_distrib, 1)
    nhw = max(1, size(hw_lst))
    pos = present(A_ka) .and. present(Om_kab)
    call get_band_ctx(n, pos, ctx)
    allocate(e(n), fd(nef, n), V(3, n, n), Ne(nef))
    e = en_k
    fd = fd_distrib
    V = V_ka
    allocate(ahc(3, 3, nef), velo(3, 3, nhw, nef), ohc(3, 3, nhw, nef), kic(3, 3, nef), &
             klc(3, 3, nef), kcs(3, 3, nef), nic(3, 3, max(1, nval)), nlc(3, 3, max(1, nval)), &
             ncs(3, 3, max(1, nval)))
    ap = c_null_ptr
    op = c_null_ptr
    if (pos) then
      allocate(A(3, n, n), Om(3, 3, n, n))
      A = A_ka
      Om = Om_kab
      ap = c_loc(A)
      op = c_loc(Om)
    end if
    if (size(hw_lst) >= 1) then
      ! the caller's frequencies as given (any spacing), kubo.f90:86-104, 145-164
      p = make_run_params(jobs, nef, 0.0_dp, 0.0_dp, nhw, hw_lst(1), hw_lst(size(hw_lst)), &
                          0.0_dp, eta, nval)
      allocate(hw(nhw))
      hw = hw_lst
      p%hw_lst = c_loc(hw)
    else
      p = make_run_params(jobs, nef, 0.0_dp, 0.0_dp, 1, 0.0_dp, 0.0_dp, 0.0_dp, eta, nval)
    end if
    s%Ne = c_loc(Ne)
    s%ahc = c_loc(ahc)
    s%ahc_velo = c_loc(velo)
    s%ohc = c_loc(ohc)
    s%kubo_ic = c_loc(kic)
    s%kubo_lc = c_loc(klc)
    s%kubo_cs = c_loc(kcs)
    s%niu_ic = c_loc(nic)
    s%niu_lc = c_loc(nlc)
    s%niu_cs = c_loc(ncs)
    rc = mep_bandsum_batch(ctx, 1_c_int64_t, c_loc(e), c_loc(V), ap, op, c_loc(fd), p, s)
    call mep_check(rc, "mep_bandsum_batch")
  end subroutine gpu_bandsum_k
end module mep_gpu
