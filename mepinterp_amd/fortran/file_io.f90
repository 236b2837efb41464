! file_io.f90 -- the boundary formats of the k-space path:
!   <seed>_hr.dat reader (file_io.f90:663-746 of the reference),
!   <seed>_r.dat  reader (:750-795; opened as <seed>_r.dat, R fractional,
!                  positions in Angstrom),
!   <seed>_tb.dat reader (:555-658; R kept fractional),
!   NPY v2.0 Fortran-order writer with the header layout of the reference's
!   output files, and the response-tensor writers (:186-359).
module file_io
  use iso_fortran_env, only: int8, int32
  use constants, only: dp, fp_acc, aUtoEv, aUtoAngstrm
  use input_paras, only: w90_dir, out_dir, use_R_float, a_latt
  implicit none
  private
  public :: read_hr_file, read_r_file, read_tb_file, read_tb_basis, save_npy, write_hw_list, write_mep_bands, &
            write_mep_tensors, write_kubo_mep_tensors, write_ahc_tensor, write_opt_tensors, &
            write_gyro_tensors

  interface save_npy
    module procedure save_d1, save_d2, save_d3, save_z3, save_z4, save_z5
  end interface save_npy
contains

  subroutine read_hr_file(seed_name, R_vect, H_mat)
    character(len=*), intent(in) :: seed_name
    real(dp), allocatable, intent(inout) :: R_vect(:, :)
    complex(dp), allocatable, intent(inout) :: H_mat(:, :, :)
    integer :: u, f_nwfs, f_nSC, to_read, idx, sc, wf, m, n, ios, int3(3), k
    integer, allocatable :: deg(:)
    real(dp) :: real2(2), real3(3)
    open(newunit=u, file=seed_name // "_hr.dat", form="formatted", action="read", status="old", iostat=ios)
    if (ios /= 0) stop "[read_hr_file]: ERROR could not open the _hr.dat file"
    read(u, *)
    read(u, *) f_nwfs
    read(u, *) f_nSC
    allocate(deg(f_nSC), R_vect(3, f_nSC), H_mat(f_nwfs, f_nwfs, f_nSC))
    H_mat = cmplx(0.0_dp, 0.0_dp, dp)
    to_read = f_nSC
    idx = 1
    do while (to_read > 0)
      k = min(15, to_read)
      read(u, *) deg(idx:idx + k - 1)
      idx = idx + k
      to_read = to_read - k
    end do
    do sc = 1, f_nSC
      do wf = 1, f_nwfs**2
        if (use_R_float) then
          read(u, *) real3(1:3), m, n, real2(1:2)
        else
          read(u, *) int3(1:3), m, n, real2(1:2)
          real3 = real(int3, dp)
        end if
        if (wf == 1) then
          R_vect(:, sc) = real3
        else if (maxval(abs(R_vect(:, sc) - real3)) > fp_acc) then
          write(*, *) "[read_hr_file]: WARNING unexpected new R_vect=", real3
        end if
        if (abs(real2(1)) < 1e-7_dp) real2(1) = 0.0_dp
        if (abs(real2(2)) < 1e-7_dp) real2(2) = 0.0_dp
        H_mat(m, n, sc) = cmplx(real2(1), real2(2), dp)
      end do
    end do
    close(u)
    H_mat = H_mat / aUtoEv
    write(*, '(a,i6,a,i6,a)') "[read_hr_file]: SUCCESS (nWfs=", f_nwfs, "; nrpts=", f_nSC, ")"
  end subroutine read_hr_file

  subroutine read_r_file(seed_name, R_vect, r_mat)
    ! rows "R1 R2 R3 m n Re(x) Im(x) Re(y) Im(y) Re(z) Im(z)", R in the order
    ! of the _hr.dat file, positions in Angstrom
    character(len=*), intent(in) :: seed_name
    real(dp), intent(in) :: R_vect(:, :)
    complex(dp), allocatable, intent(inout) :: r_mat(:, :, :, :)
    integer :: u, f_nwfs, sc, it, m, n, ios, int3(3)
    real(dp) :: real6(6), real3(3)
    open(newunit=u, file=seed_name // "_r.dat", form="formatted", action="read", status="old", iostat=ios)
    if (ios /= 0) stop "[read_r_file]: ERROR could not open the _r.dat file"
    read(u, *)
    read(u, *) f_nwfs
    allocate(r_mat(3, f_nwfs, f_nwfs, size(R_vect, 2)))
    do sc = 1, size(R_vect, 2)
      do it = 1, f_nwfs**2
        if (use_R_float) then
          read(u, *) real3(1:3), m, n, real6(1:6)
        else
          read(u, *) int3(1:3), m, n, real6(1:6)
          real3 = real(int3, dp)
        end if
        if (maxval(abs(R_vect(:, sc) - real3)) > fp_acc) &
          stop "[read_r_file]: different R_vect order in _hr.dat and _r.dat file"
        r_mat(1, m, n, sc) = cmplx(real6(1), real6(2), dp)
        r_mat(2, m, n, sc) = cmplx(real6(3), real6(4), dp)
        r_mat(3, m, n, sc) = cmplx(real6(5), real6(6), dp)
      end do
    end do
    close(u)
    r_mat = r_mat / aUtoAngstrm
  end subroutine read_r_file

  subroutine read_tb_file(seed_name, R_vect, H_mat, r_mat)
    ! Wannier90 _tb.dat: a comment line, the unit cell in Angstrom (one
    ! lattice vector per line, overriding a_latt as file_io.f90:578-579 does),
    ! n_wf, n_R, the degeneracy list (skipped), then per R a blank line, the R
    ! line and n_wf**2 lines "m n Re Im" (eV), then per R a blank line, the R
    ! line and n_wf**2 lines "m n Re(x) Im(x) Re(y) Im(y) Re(z) Im(z)" (Angstrom).
    ! R is kept fractional: the reference stores matmul(unit_cell, R)
    ! (file_io.f90:605) and then uses it as the fractional R of the Fourier
    ! phase (wann_interp.f90:213-219).
    character(len=*), intent(in) :: seed_name
    real(dp), allocatable, intent(inout) :: R_vect(:, :)
    complex(dp), allocatable, intent(inout) :: H_mat(:, :, :), r_mat(:, :, :, :)
    integer :: u, f_nwfs, f_nSC, sc, it, m, n, ios, int3(3), idx(2), line
    real(dp) :: cell(3, 3), real2(2), real6(6), real3(3)
    open(newunit=u, file=seed_name // "_tb.dat", form="formatted", action="read", status="old", iostat=ios)
    if (ios /= 0) stop "[read_tb_file]: ERROR could not open the _tb.dat file"
    read(u, *)
    do m = 1, 3
      read(u, *) cell(m, 1:3)
    end do
    cell = cell / aUtoAngstrm
    if (maxval(abs(cell - a_latt)) > 1e-6_dp * maxval(abs(cell))) &
      write(*, *) "[read_tb_file]: WARNING the _tb.dat unit cell differs from the input.cfg lattice"
    a_latt = cell
    read(u, *) f_nwfs
    read(u, *) f_nSC
    allocate(R_vect(3, f_nSC), H_mat(f_nwfs, f_nwfs, f_nSC), r_mat(3, f_nwfs, f_nwfs, f_nSC))
    H_mat = cmplx(0.0_dp, 0.0_dp, dp)
    r_mat = cmplx(0.0_dp, 0.0_dp, dp)
    do line = 1, (f_nSC + 14) / 15
      read(u, *)
    end do
    do sc = 1, f_nSC
      read(u, *)
      call read_R(real3)
      R_vect(:, sc) = real3
      do it = 1, f_nwfs**2
        read(u, *) idx(1:2), real2(1:2)
        if (abs(real2(1)) < 1e-7_dp) real2(1) = 0.0_dp
        if (abs(real2(2)) < 1e-7_dp) real2(2) = 0.0_dp
        H_mat(idx(1), idx(2), sc) = cmplx(real2(1), real2(2), dp)
      end do
    end do
    do sc = 1, f_nSC
      read(u, *)
      call read_R(real3)
      if (norm2(real3 - R_vect(:, sc)) > 1e-8_dp) &
        write(*, *) "[read_tb_file]: WARNING rHopp has different sc order then tHopp"
      do it = 1, f_nwfs**2
        read(u, *) idx(1:2), real6(1:6)
        m = idx(1)
        n = idx(2)
        r_mat(1, m, n, sc) = cmplx(real6(1), real6(2), dp)
        r_mat(2, m, n, sc) = cmplx(real6(3), real6(4), dp)
        r_mat(3, m, n, sc) = cmplx(real6(5), real6(6), dp)
      end do
    end do
    close(u)
    H_mat = H_mat / aUtoEv
    r_mat = r_mat / aUtoAngstrm
    write(*, '(a,i6,a,i6,a)') "[read_tb_file]: SUCCESS (nWfs=", f_nwfs, "; nrpts=", f_nSC, ")"
  contains
    subroutine read_R(v)
      real(dp), intent(out) :: v(3)
      if (use_R_float) then
        read(u, *) v
      else
        read(u, *) int3
        v = real(int3, dp)
      end if
    end subroutine read_R
  end subroutine read_tb_file

  subroutine read_tb_basis(seed_name, R_vect, H_mat, r_mat)
    ! <seed>_tb.dat when present, else <seed>_hr.dat and an optional
    ! <seed>_r.dat (file_io.f90:800-860)
    character(len=*), intent(in) :: seed_name
    real(dp), allocatable, intent(inout) :: R_vect(:, :)
    complex(dp), allocatable, intent(inout) :: H_mat(:, :, :), r_mat(:, :, :, :)
    logical :: hr_exist, r_exist, tb_exist
    inquire(file=w90_dir // seed_name // "_tb.dat", exist=tb_exist)
    if (tb_exist) then
      write(*, *) '[read_tb_basis]: found data file ' // w90_dir // seed_name // '_tb.dat'
      call read_tb_file(w90_dir // seed_name, R_vect, H_mat, r_mat)
      if (size(H_mat, 1) /= size(H_mat, 2)) stop "[read_tb_basis]: H_mat is not square"
      return
    end if
    inquire(file=w90_dir // seed_name // "_hr.dat", exist=hr_exist)
    if (.not. hr_exist) then
      write(*, *) '[read_tb_basis]: "' // w90_dir // seed_name // '_hr.dat" does not exist'
      stop "[read_tb_basis]: ERROR could not read the real space Hamiltonian"
    end if
    call read_hr_file(w90_dir // seed_name, R_vect, H_mat)
    inquire(file=w90_dir // seed_name // "_r.dat", exist=r_exist)
    if (r_exist) call read_r_file(w90_dir // seed_name, R_vect, r_mat)
    if (size(H_mat, 1) /= size(H_mat, 2)) stop "[read_tb_basis]: H_mat is not square"
  end subroutine read_tb_basis

  ! ---------------------------------------------------------------- NPY
  subroutine npy_open(fname, descr, shp, u)
    character(len=*), intent(in) :: fname, descr
    integer, intent(in) :: shp(:)
    integer, intent(out) :: u
    character(len=:), allocatable :: dict, num
    character(len=24) :: buf
    integer :: i, hlen
    dict = "{'descr': '" // descr // "', 'fortran_order': True, 'shape': ("
    do i = 1, size(shp)
      write(buf, '(i0)') shp(i)
      num = trim(buf)
      dict = dict // num // ","
    end do
    dict = dict // "), }"
    hlen = len(dict) + 1
    do while (mod(10 + hlen, 16) /= 0)
      hlen = hlen + 1
    end do
    dict = dict // repeat(" ", hlen - 1 - len(dict)) // achar(10)
    open(newunit=u, file=fname, form="unformatted", access="stream", status="replace", action="write")
    write(u) char(147) // "NUMPY" // char(2) // char(0)
    write(u) int(hlen, int32)
    write(u) dict
  end subroutine npy_open

  subroutine save_d1(fname, a)
    character(len=*), intent(in) :: fname
    real(dp), intent(in) :: a(:)
    integer :: u
    call npy_open(fname, "<f8", shape(a), u)
    write(u) a
    close(u)
  end subroutine save_d1
  subroutine save_d2(fname, a)
    character(len=*), intent(in) :: fname
    real(dp), intent(in) :: a(:, :)
    integer :: u
    call npy_open(fname, "<f8", shape(a), u)
    write(u) a
    close(u)
  end subroutine save_d2
  subroutine save_d3(fname, a)
    character(len=*), intent(in) :: fname
    real(dp), intent(in) :: a(:, :, :)
    integer :: u
    call npy_open(fname, "<f8", shape(a), u)
    write(u) a
    close(u)
  end subroutine save_d3
  subroutine save_z3(fname, a)
    character(len=*), intent(in) :: fname
    complex(dp), intent(in) :: a(:, :, :)
    integer :: u
    call npy_open(fname, "<c16", shape(a), u)
    write(u) a
    close(u)
  end subroutine save_z3
  subroutine save_z4(fname, a)
    character(len=*), intent(in) :: fname
    complex(dp), intent(in) :: a(:, :, :, :)
    integer :: u
    call npy_open(fname, "<c16", shape(a), u)
    write(u) a
    close(u)
  end subroutine save_z4

  subroutine save_z5(fname, a)
    character(len=*), intent(in) :: fname
    complex(dp), intent(in) :: a(:, :, :, :, :)
    integer :: u
    call npy_open(fname, "<c16", shape(a), u)
    write(u) a
    close(u)
  end subroutine save_z5

  ! write_opt_tensors / write_gyro_tensors (file_io.f90:362-398)
  subroutine write_opt_tensors(s_symm, a_symm, photo_2nd)
    complex(dp), intent(in) :: s_symm(:, :, :), a_symm(:, :, :), photo_2nd(:, :, :, :, :)
    call save_npy(out_dir // "opt_Ssymm.npy", s_symm)
    call save_npy(out_dir // "opt_Asymm.npy", a_symm)
    call save_npy(out_dir // "photoC_2nd.npy", photo_2nd)
  end subroutine write_opt_tensors

  subroutine write_gyro_tensors(C_tens, D_tens, Dw_tens)
    complex(dp), intent(in) :: C_tens(:, :, :), D_tens(:, :, :), Dw_tens(:, :, :, :)
    call save_npy(out_dir // "gyro_C.npy", C_tens)
    call save_npy(out_dir // "gyro_D.npy", D_tens)
    call save_npy(out_dir // "gyro_Dw.npy", Dw_tens)
  end subroutine write_gyro_tensors

  ! ---------------------------------------------------------------- writers
  subroutine write_hw_list(n_hw, hw_min, hw_max)
    integer, intent(in) :: n_hw
    real(dp), intent(in) :: hw_min, hw_max
    real(dp), allocatable :: hw_lst(:)
    integer :: i
    real(dp) :: d
    if (n_hw > 1) then
      allocate(hw_lst(n_hw))
      d = (hw_max - hw_min) / real(n_hw - 1, dp)
      do i = 1, n_hw
        hw_lst(i) = (hw_min + real(i - 1, dp) * d) * aUtoEv
      end do
    else
      allocate(hw_lst(1))
      hw_lst = hw_min
    end if
    call save_npy(out_dir // "hw_lst.npy", hw_lst)
  end subroutine write_hw_list

  subroutine write_mep_bands(mep_bands)
    real(dp), allocatable, intent(in) :: mep_bands(:, :, :)
    if (allocated(mep_bands)) call save_npy(out_dir // "mep_bands.npy", mep_bands)
  end subroutine write_mep_bands

  subroutine write_mep_tensors(mep_ic, mep_lc, mep_cs)
    real(dp), allocatable, intent(in) :: mep_ic(:, :), mep_lc(:, :), mep_cs(:, :)
    if (allocated(mep_ic)) call save_npy(out_dir // "mep_ic.npy", mep_ic)
    if (allocated(mep_lc)) call save_npy(out_dir // "mep_lc.npy", mep_lc)
    if (allocated(mep_cs)) call save_npy(out_dir // "mep_cs.npy", mep_cs)
    if (allocated(mep_ic) .and. allocated(mep_lc) .and. allocated(mep_cs)) &
      call save_npy(out_dir // "mep_tens.npy", mep_ic + mep_lc + mep_cs)
  end subroutine write_mep_tensors

  subroutine write_kubo_mep_tensors(ic, lc, cs)
    real(dp), allocatable, intent(in) :: ic(:, :), lc(:, :), cs(:, :)
    if (allocated(ic)) call save_npy(out_dir // "kubo_mep_ic.npy", ic)
    if (allocated(lc)) call save_npy(out_dir // "kubo_mep_lc.npy", lc)
    if (allocated(cs)) call save_npy(out_dir // "kubo_mep_cs.npy", cs)
    if (allocated(ic) .and. allocated(lc) .and. allocated(cs)) &
      call save_npy(out_dir // "kubo_mep_tens.npy", ic + lc + cs)
  end subroutine write_kubo_mep_tensors

  subroutine write_ahc_tensor(ahc_tens, velo_ahc_tens, ohc_tens)
    real(dp), allocatable, intent(in) :: ahc_tens(:, :)
    complex(dp), allocatable, intent(in) :: velo_ahc_tens(:, :, :), ohc_tens(:, :, :)
    if (allocated(ahc_tens)) call save_npy(out_dir // "ahc_tens.npy", ahc_tens)
    if (allocated(velo_ahc_tens)) call save_npy(out_dir // "ahcVELO.npy", velo_ahc_tens)
    if (allocated(ohc_tens)) call save_npy(out_dir // "ohcVELO.npy", ohc_tens)
  end subroutine write_ahc_tensor
end module file_io
