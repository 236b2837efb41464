! core.f90 -- core_worker (core.f90:79-298 of the reference): the k-loop over
! the Monkhorst-Pack mesh runs on the GPU in one call (mep_run_mesh); the
! Fermi-level selection, normalisation and NPY output stay on the host, as
! fermi_selector / integrate_over_k_space / mep_sum_bands
! (core.f90:317-603) do.
module core
  use iso_c_binding
  use constants, only: dp, aUtoEv
  use mep_capi
  use mep_gpu, only: gpu_model_ctx, make_run_params
  use matrix_math, only: get_linspace
  use k_space, only: get_mp_grid, normalize_k_int, k_norm, get_kpt_idx, get_rel_kpt
  use input_paras, only: seed_name, valence_bands, do_gauge_trafo, do_mep, do_ahc, do_kubo, &
                         do_opt, do_gyro, do_write_mep_bands, wf_centers, N_hw, hw_min, hw_max, &
                         N_eF, eF_min, eF_max, T_kelvin, i_eta_smr, out_dir, do_write_velo, &
                         my_mkdir
  use band_calc, only: write_velo
  use file_io, only: read_tb_basis, write_hw_list, write_mep_bands, write_mep_tensors, &
                     write_kubo_mep_tensors, write_ahc_tensor, write_opt_tensors, &
                     write_gyro_tensors
  implicit none
  private
  public :: core_worker, fermi_selector
contains

  subroutine core_worker()
    real(dp), allocatable :: R_vect(:, :), ef_lst(:), hw_lst(:)
    complex(dp), allocatable :: H_tb(:, :, :), r_tb(:, :, :, :)
    real(dp), allocatable, target :: Ne(:), ahc(:, :, :), kic(:, :, :), klc(:, :, :), &
                                     kcs(:, :, :), nic(:, :, :), nlc(:, :, :), ncs(:, :, :)
    complex(dp), allocatable, target :: velo(:, :, :, :), ohc(:, :, :, :)
    complex(dp), allocatable, target :: opt_s(:, :, :, :), opt_a(:, :, :, :), photo(:, :, :, :, :), &
                                        gyro_C(:, :, :), gyro_D(:, :, :), gyro_Dw(:, :, :, :)
    type(c_ptr) :: ctx
    type(mep_run_params) :: p
    type(mep_sums) :: s
    integer(c_int32_t), target :: mp(3)
    integer :: jobs, nval, eF_idx, n_ki_glob
    integer(c_int) :: rc
    integer(8) :: t0, t1, rate
    logical :: use_pos

    call get_linspace(eF_min, eF_max, N_eF, ef_lst)
    call get_linspace(hw_min, hw_max, N_hw, hw_lst)
    call read_tb_basis(seed_name, R_vect, H_tb, r_tb)
    use_pos = allocated(r_tb)
    call gpu_model_ctx(do_gauge_trafo, H_tb, r_tb, use_pos, R_vect, wf_centers, ctx)

    jobs = 0
    if (do_ahc) jobs = ior(jobs, int(MEP_JOB_AHC))
    if (do_kubo) jobs = ior(jobs, int(MEP_JOB_KUBO))
    if (do_mep) jobs = ior(jobs, int(MEP_JOB_MEP))
    if (do_opt) jobs = ior(jobs, int(MEP_JOB_OPT))
    if (do_gyro) jobs = ior(jobs, int(MEP_JOB_GYRO))
    nval = max(1, valence_bands)
    p = make_run_params(jobs, N_eF, eF_min, eF_max, N_hw, hw_min, hw_max, T_kelvin, &
                        aimag(i_eta_smr), nval)
    allocate(Ne(N_eF), ahc(3, 3, N_eF), velo(3, 3, N_hw, N_eF), ohc(3, 3, N_hw, N_eF), &
             kic(3, 3, N_eF), klc(3, 3, N_eF), kcs(3, 3, N_eF), nic(3, 3, nval), &
             nlc(3, 3, nval), ncs(3, 3, nval))
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
    allocate(opt_s(3, 3, N_hw, N_eF), opt_a(3, 3, N_hw, N_eF), photo(3, 3, 3, N_hw, N_eF), &
             gyro_C(3, 3, N_eF), gyro_D(3, 3, N_eF), gyro_Dw(3, 3, N_hw, N_eF))
    s%opt_s = c_loc(opt_s)
    s%opt_a = c_loc(opt_a)
    s%photo = c_loc(photo)
    s%gyro_C = c_loc(gyro_C)
    s%gyro_D = c_loc(gyro_D)
    s%gyro_Dw = c_loc(gyro_Dw)
    mp = int(get_mp_grid(), c_int32_t)
    call system_clock(t0, rate)
    rc = mep_run_mesh(ctx, mp, p, s)
    call mep_check(rc, "core_worker k-loop")
    call system_clock(t1)
    n_ki_glob = int(s%n_k)
    write(*, '(a,i0,a,f10.4,a,es12.4,a)') "[core_worker]: k-loop of ", n_ki_glob, " k-points on the GPU in ", &
      real(t1 - t0, dp) / real(rate, dp), " s (", real(n_ki_glob, dp) * real(rate, dp) / &
      max(real(t1 - t0, dp), 1.0_dp), " k-pts/s)"

    ! jobs%do_write_velo: out/velo/velo_Vka.<ki> for every k-point of the
    ! mesh, as the reference's k-loop writes them (core.f90:258)
    if (do_write_velo .and. jobs /= 0) call write_mesh_velo(ctx, size(H_tb, 1), use_pos)

    eF_idx = fermi_selector(n_ki_glob, Ne, ef_lst, valence_bands)
    call integrate_over_k_space(eF_idx, nic, nlc, ncs, kic, klc, kcs, ahc, velo, ohc)
    ! optical tensors at eF_idx, photogalvanic and gyrotropic tensors at every
    ! E_F, all normalised (core.f90:512-531)
    if (do_opt) then
      call write_opt_tensors(opt_s(:, :, :, eF_idx) / k_norm(), opt_a(:, :, :, eF_idx) / k_norm(), &
                             photo / k_norm())
    end if
    if (do_gyro) call write_gyro_tensors(gyro_C / k_norm(), gyro_D / k_norm(), gyro_Dw / k_norm())
  end subroutine core_worker

  ! the velocities of the mesh's k-points, in the reference's visiting order
  ! (kiz slowest, kix fastest; core.f90:165-178), interpolated in batches of
  ! up to 4096 k-points (about 50 MB of V per batch) and written one file per
  ! k-point (file_io.f90:146-183)
  subroutine write_mesh_velo(ctx, n, use_pos)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: n
    logical, intent(in) :: use_pos
    real(dp), allocatable, target :: kp(:, :), en(:, :)
    complex(dp), allocatable, target :: V(:, :, :, :), A(:, :, :, :), Om(:, :, :, :, :)
    integer, allocatable :: idx(:)
    type(c_ptr) :: ap, op
    integer :: g(3), kix, kiy, kiz, q, nb, chunk
    chunk = max(64, min(4096, 50000000 / (48 * n * n)))
    call my_mkdir(out_dir // "/velo/")
    g = get_mp_grid()
    allocate(kp(3, chunk), en(n, chunk), V(3, n, n, chunk), idx(chunk))
    ap = c_null_ptr
    op = c_null_ptr
    if (use_pos) then
      allocate(A(3, n, n, chunk), Om(3, 3, n, n, chunk))
      ap = c_loc(A)
      op = c_loc(Om)
    end if
    nb = 0
    do kiz = 1, g(3)
      do kiy = 1, g(2)
        do kix = 1, g(1)
          nb = nb + 1
          idx(nb) = get_kpt_idx(kix, kiy, kiz)
          kp(:, nb) = get_rel_kpt(idx(nb), kix, kiy, kiz)
          if (nb == chunk .or. (kix == g(1) .and. kiy == g(2) .and. kiz == g(3))) then
            call mep_check(mep_interp_batch(ctx, int(nb, c_int64_t), c_loc(kp), c_loc(en), &
                                            c_loc(V), ap, op), "write_mesh_velo")
            do q = 1, nb
              call write_velo(idx(q), V(:, :, :, q))
            end do
            nb = 0
          end if
        end do
      end do
    end do
  end subroutine write_mesh_velo

  integer function fermi_selector(n_ki_glob, Ne_loc_sum, eF_lst, Ne_sys) result(opt_idx)
    ! core.f90:317-415: E_F minimising |N_el(E_F) - N_el,sys|, first on ties;
    ! writes out/occ.dat
    integer, intent(in) :: n_ki_glob, Ne_sys
    real(dp), intent(in) :: Ne_loc_sum(:), eF_lst(:)
    real(dp), allocatable :: Ne_glob(:)
    real(dp) :: opt_error, new_error
    integer :: i, u
    Ne_glob = Ne_loc_sum / real(n_ki_glob, dp)
    open(newunit=u, file=out_dir // "occ.dat", form="formatted", action="write", status="replace")
    opt_idx = 1
    opt_error = abs(Ne_glob(1) - real(Ne_sys, dp))
    do i = 1, size(eF_lst)
      new_error = abs(Ne_glob(i) - real(Ne_sys, dp))
      write(u, '(f16.8,a,f16.8)') eF_lst(i) * aUtoEv, "	", Ne_glob(i)
      if (new_error < opt_error) then
        opt_idx = i
        opt_error = new_error
      end if
    end do
    close(u)
    write(*, '(a,f12.6,a,es16.8)') "[fermi_selector]: eF_optimal=", eF_lst(opt_idx) * aUtoEv, &
      " eV with abs(N_el(eF_optimal) - N_el)=", opt_error
  end function fermi_selector

  subroutine integrate_over_k_space(eF_idx, nic, nlc, ncs, kic, klc, kcs, ahc, velo, ohc)
    ! core.f90:419-603 (single process: the reductions are the identity)
    integer, intent(in) :: eF_idx
    real(dp), intent(in) :: nic(:, :, :), nlc(:, :, :), ncs(:, :, :), kic(:, :, :), klc(:, :, :), &
                            kcs(:, :, :), ahc(:, :, :)
    complex(dp), intent(in) :: velo(:, :, :, :), ohc(:, :, :, :)
    real(dp), allocatable :: mep_ic(:, :), mep_lc(:, :), mep_cs(:, :), mep_bands(:, :, :), &
                             k_ic(:, :), k_lc(:, :), k_cs(:, :), ahc_g(:, :)
    complex(dp), allocatable :: velo_g(:, :, :), ohc_g(:, :, :)
    integer :: n0
    if (do_mep) then
      allocate(mep_ic(3, 3), mep_lc(3, 3), mep_cs(3, 3))
      mep_ic = 0.0_dp
      mep_lc = 0.0_dp
      mep_cs = 0.0_dp
      if (do_write_mep_bands) allocate(mep_bands(3, 3, valence_bands))
      do n0 = 1, valence_bands
        mep_ic = mep_ic + nic(:, :, n0)
        mep_lc = mep_lc + nlc(:, :, n0)
        mep_cs = mep_cs + ncs(:, :, n0)
        if (do_write_mep_bands) mep_bands(:, :, n0) = nic(:, :, n0) + nlc(:, :, n0) + ncs(:, :, n0)
      end do
      call normalize_k_int(mep_ic)
      call normalize_k_int(mep_lc)
      call normalize_k_int(mep_cs)
      call normalize_k_int(mep_bands)
    end if
    if (do_kubo) then
      k_ic = kic(:, :, eF_idx)
      k_lc = klc(:, :, eF_idx)
      k_cs = kcs(:, :, eF_idx)
      call normalize_k_int(k_ic)
      call normalize_k_int(k_lc)
      call normalize_k_int(k_cs)
    end if
    if (do_ahc) then
      ahc_g = ahc(:, :, eF_idx)
      velo_g = velo(:, :, :, eF_idx)
      ohc_g = ohc(:, :, :, eF_idx)
      call normalize_k_int(ahc_g)
      call normalize_k_int(velo_g)
      call normalize_k_int(ohc_g)
    end if
    call write_hw_list(N_hw, hw_min, hw_max)
    call write_mep_bands(mep_bands)
    call write_mep_tensors(mep_ic, mep_lc, mep_cs)
    call write_kubo_mep_tensors(k_ic, k_lc, k_cs)
    call write_ahc_tensor(ahc_g, velo_g, ohc_g)
    write(*, '(a)') "[core_worker]: wrote the response tensors to " // out_dir
  end subroutine integrate_over_k_space
end module core
