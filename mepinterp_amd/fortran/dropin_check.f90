! dropin_check.f90 -- calls the per-k interface exactly as the reference's
! core_worker does (core.f90:176-230) for the k-points listed in kpts.txt
! and writes every result to out/dropin_*.npy.  Used by the GPU tests.
program dropin_check
  use constants, only: dp
  use input_paras, only: init_parameters, seed_name, do_gauge_trafo, wf_centers, valence_bands, &
                         N_eF, eF_min, eF_max, N_hw, hw_min, hw_max, T_kelvin, i_eta_smr, out_dir
  use matrix_math, only: get_linspace
  use file_io, only: read_tb_basis, save_npy
  use wann_interp, only: get_wann_interp
  use statistics, only: fd_stat
  use kubo, only: kubo_ahc_tens, velo_ahc_tens, kubo_ohc_tens
  use kubo_mep, only: kubo_mep_IC, kubo_mep_LC, kubo_mep_CS
  use mep_niu, only: mep_niu_IC, mep_niu_LC, mep_niu_CS
  use k_space, only: kspace_allocator
  use mep_gpu, only: gpu_release
  implicit none
  real(dp), allocatable :: R_vect(:, :), en_k(:), ef_lst(:), hw_lst(:), fd(:, :), kpts(:, :)
  complex(dp), allocatable :: H_tb(:, :, :), r_tb(:, :, :, :), V_ka(:, :, :), A_ka(:, :, :), &
                              Om_kab(:, :, :, :)
  real(dp), allocatable :: e_all(:, :), ahc(:, :, :, :), kic(:, :, :, :), klc(:, :, :, :), &
                           kcs(:, :, :, :), nic(:, :, :, :), nlc(:, :, :, :), ncs(:, :, :, :)
  complex(dp), allocatable :: V_all(:, :, :, :), velo(:, :, :, :), ohc(:, :, :, :)
  complex(dp), allocatable :: H_conj(:, :, :)
  real(dp), allocatable :: hw_nu(:), e_conj(:, :)
  integer :: u, nk, ik, n, i
  if (.not. init_parameters()) stop "no input.cfg"
  call get_linspace(eF_min, eF_max, N_eF, ef_lst)
  call get_linspace(hw_min, hw_max, N_hw, hw_lst)
  call read_tb_basis(seed_name, R_vect, H_tb, r_tb)
  call kspace_allocator(H_tb, r_tb, en_k, V_ka, A_ka, Om_kab)
  n = size(H_tb, 1)
  open(newunit=u, file="kpts.txt", status="old", action="read")
  read(u, *) nk
  allocate(kpts(3, nk))
  do ik = 1, nk
    read(u, *) kpts(:, ik)
  end do
  close(u)
  allocate(e_all(n, nk), V_all(3 * n, n, nk, 1), ahc(3, 3, N_eF, nk), velo(3, 3, N_hw * N_eF, nk), &
           ohc(3, 3, N_hw * N_eF, nk), kic(3, 3, N_eF, nk), klc(3, 3, N_eF, nk), &
           kcs(3, 3, N_eF, nk), nic(3, 3, valence_bands, nk), nlc(3, 3, valence_bands, nk), &
           ncs(3, 3, valence_bands, nk), fd(N_eF, n))
  do ik = 1, nk
    call get_wann_interp(do_gauge_trafo, H_tb, r_tb, R_vect, wf_centers, ik, kpts(:, ik), en_k, &
                         V_ka, A_ka, Om_kab)
    e_all(:, ik) = en_k
    V_all(:, :, ik, 1) = reshape(V_ka, [3 * n, n])
    do i = 1, n
      fd(:, i) = fd_stat(en_k(i), ef_lst, T_kelvin)
    end do
    ahc(:, :, :, ik) = kubo_ahc_tens(en_k, V_ka, fd)
    velo(:, :, :, ik) = reshape(velo_ahc_tens(en_k, V_ka, hw_lst, fd, i_eta_smr), [3, 3, N_hw * N_eF])
    ohc(:, :, :, ik) = reshape(kubo_ohc_tens(en_k, V_ka, hw_lst, fd, i_eta_smr), [3, 3, N_hw * N_eF])
    kic(:, :, :, ik) = kubo_mep_IC(en_k, V_ka, fd)
    klc(:, :, :, ik) = kubo_mep_LC(en_k, V_ka, fd)
    kcs(:, :, :, ik) = kubo_mep_CS(A_ka, Om_kab, fd)
    nic(:, :, :, ik) = mep_niu_IC(V_ka, en_k)
    nlc(:, :, :, ik) = mep_niu_LC(V_ka, en_k)
    ncs(:, :, :, ik) = mep_niu_CS(A_ka, Om_kab)
  end do
  ! a second model with the same |H(R)| element by element (complex
  ! conjugate): the resident model must not be reused for it
  H_conj = conjg(H_tb)
  allocate(e_conj(n, nk))
  do ik = 1, nk
    call get_wann_interp(do_gauge_trafo, H_conj, r_tb, R_vect, wf_centers, ik, kpts(:, ik), en_k, &
                         V_ka, A_ka, Om_kab)
    e_conj(:, ik) = en_k
  end do
  call save_npy(out_dir // "dropin_e_conj.npy", e_conj)
  ! a non-uniform frequency list, kept as given (kubo.f90:86-104, 145-164)
  allocate(hw_nu(N_hw + 2))
  do i = 1, N_hw + 2
    hw_nu(i) = (hw_max + 0.1_dp) * (real(i - 1, dp) / real(N_hw + 1, dp))**2
  end do
  call get_wann_interp(do_gauge_trafo, H_tb, r_tb, R_vect, wf_centers, 1, kpts(:, 1), en_k, &
                       V_ka, A_ka, Om_kab)
  do i = 1, n
    fd(:, i) = fd_stat(en_k(i), ef_lst, T_kelvin)
  end do
  call save_npy(out_dir // "dropin_hw_nu.npy", hw_nu)
  call save_npy(out_dir // "dropin_velo_nu.npy", &
                reshape(velo_ahc_tens(en_k, V_ka, hw_nu, fd, i_eta_smr), [3, 3, (N_hw + 2) * N_eF]))
  call save_npy(out_dir // "dropin_ohc_nu.npy", &
                reshape(kubo_ohc_tens(en_k, V_ka, hw_nu, fd, i_eta_smr), [3, 3, (N_hw + 2) * N_eF]))
  call save_npy(out_dir // "dropin_e.npy", e_all)
  call save_npy(out_dir // "dropin_V.npy", V_all)
  call save_npy(out_dir // "dropin_ahc.npy", reshape(ahc, [3, 3, N_eF * nk]))
  call save_npy(out_dir // "dropin_velo.npy", velo)
  call save_npy(out_dir // "dropin_ohc.npy", ohc)
  call save_npy(out_dir // "dropin_kic.npy", reshape(kic, [3, 3, N_eF * nk]))
  call save_npy(out_dir // "dropin_klc.npy", reshape(klc, [3, 3, N_eF * nk]))
  call save_npy(out_dir // "dropin_kcs.npy", reshape(kcs, [3, 3, N_eF * nk]))
  call save_npy(out_dir // "dropin_nic.npy", reshape(nic, [3, 3, valence_bands * nk]))
  call save_npy(out_dir // "dropin_nlc.npy", reshape(nlc, [3, 3, valence_bands * nk]))
  call save_npy(out_dir // "dropin_ncs.npy", reshape(ncs, [3, 3, valence_bands * nk]))
  call gpu_release()
  write(*, '(a,i0,a)') "[dropin_check]: ", nk, " k-points done"
end program dropin_check
