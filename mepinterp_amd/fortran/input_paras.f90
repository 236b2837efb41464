! input_paras.f90 -- reads ./input.cfg (config_fortran grammar: [section]
! headers, "key = value" lines, whitespace-separated arrays, Fortran or
! Python logicals, '#' comments) into the module variables the reference's
! init_parameters sets (input_paras.f90:79-293), with the same unit handling.
module input_paras
  use constants, only: dp, aUtoEv
  use matrix_math, only: crossP
  use k_space, only: set_recip_latt, set_mp_grid
  implicit none
  private
  public :: init_parameters, my_mkdir, cfg_load, cfg_has, &
            w90_dir, raw_dir, out_dir, &
            plot_bands, debug_mode, use_cart_velo, do_gauge_trafo, do_write_velo, use_R_float, &
            do_write_mep_bands, do_mep, do_ahc, do_kubo, do_opt, do_gyro, &
            wf_centers, seed_name, valence_bands, a_latt, kubo_tol, unit_vol, &
            N_wf, N_hw, hw_min, hw_max, N_eF, eF_min, eF_max, T_kelvin, i_eta_smr, mp_grid_in

  integer, save :: valence_bands = 0, N_wf = 0, N_eF = 1, N_hw = 1, mp_grid_in(3) = 1
  character(len=:), allocatable, save :: seed_name
  character(len=9), save :: w90_dir = "w90files/"
  character(len=4), save :: raw_dir = "raw/"
  character(len=4), save :: out_dir = "out/"
  logical, save :: plot_bands = .false., use_cart_velo = .false., do_gauge_trafo = .true., &
                   use_R_float = .false., do_write_velo = .false., do_write_mep_bands = .false., &
                   debug_mode = .false., do_mep = .false., do_ahc = .false., do_kubo = .false., &
                   do_opt = .false., do_gyro = .false.
  real(dp), save :: a_latt(3, 3) = 0.0_dp, a0 = 1.0_dp, unit_vol = 0.0_dp, kubo_tol = 1e-5_dp, &
                    hw_min = 0.0_dp, hw_max = 0.0_dp, eF_min = 0.0_dp, eF_max = 0.0_dp, &
                    T_kelvin = 0.0_dp
  complex(dp), save :: i_eta_smr = (0.0_dp, 0.0_dp)
  real(dp), allocatable, save :: wf_centers(:, :)

  ! parsed key/value store ("section%key" -> value text)
  integer, parameter :: max_keys = 256, key_len = 96, val_len = 4096
  character(len=key_len), save :: cfg_keys(max_keys)
  character(len=val_len), save :: cfg_vals(max_keys)
  integer, save :: n_cfg = 0

contains

  subroutine cfg_load(fname, ok)
    character(len=*), intent(in) :: fname
    logical, intent(out) :: ok
    integer :: u, ios, p, q
    character(len=val_len) :: line
    character(len=key_len) :: section, key
    n_cfg = 0
    section = ""
    open(newunit=u, file=fname, status="old", action="read", iostat=ios)
    ok = (ios == 0)
    if (.not. ok) return
    do
      read(u, "(a)", iostat=ios) line
      if (ios /= 0) exit
      p = index(line, "#")
      if (p > 0) line(p:) = " "
      p = index(line, ";")
      if (p > 0) line(p:) = " "
      line = adjustl(line)
      if (len_trim(line) == 0) cycle
      if (line(1:1) == "[") then
        q = index(line, "]")
        if (q > 2) section = adjustl(line(2:q - 1))
        cycle
      end if
      p = index(line, "=")
      if (p <= 1) cycle
      key = adjustl(line(1:p - 1))
      if (n_cfg >= max_keys) exit
      n_cfg = n_cfg + 1
      if (len_trim(section) > 0 .and. index(key, "%") == 0) then
        cfg_keys(n_cfg) = trim(section) // "%" // trim(key)
      else
        cfg_keys(n_cfg) = key
      end if
      cfg_vals(n_cfg) = adjustl(line(p + 1:))
    end do
    close(u)
  end subroutine cfg_load

  integer function cfg_find(key)
    character(len=*), intent(in) :: key
    integer :: i
    cfg_find = 0
    do i = n_cfg, 1, -1
      if (trim(cfg_keys(i)) == trim(key)) then
        cfg_find = i
        return
      end if
    end do
  end function cfg_find

  logical function cfg_has(key)
    character(len=*), intent(in) :: key
    cfg_has = cfg_find(key) > 0
  end function cfg_has

  subroutine get_l(key, v)
    character(len=*), intent(in) :: key
    logical, intent(inout) :: v
    integer :: i, ios
    character(len=val_len) :: s
    i = cfg_find(key)
    if (i == 0) return
    s = adjustl(cfg_vals(i))
    if (s(1:1) == "'" .or. s(1:1) == '"') s = adjustl(s(2:))
    read(s, *, iostat=ios) v
    if (ios /= 0) then
      write(*, '(3a)') "[init_parameters]: ERROR cannot read logical ", trim(key), trim(s)
      stop "input.cfg: bad logical"
    end if
  end subroutine get_l

  subroutine get_i(key, v)
    character(len=*), intent(in) :: key
    integer, intent(inout) :: v
    integer :: i, ios
    real(dp) :: r
    i = cfg_find(key)
    if (i == 0) return
    read(cfg_vals(i), *, iostat=ios) v
    if (ios /= 0) then
      read(cfg_vals(i), *, iostat=ios) r
      if (ios /= 0) stop "input.cfg: bad integer"
      v = nint(r)
    end if
  end subroutine get_i

  subroutine get_r(key, v)
    character(len=*), intent(in) :: key
    real(dp), intent(inout) :: v
    integer :: i, ios
    i = cfg_find(key)
    if (i == 0) return
    read(cfg_vals(i), *, iostat=ios) v
    if (ios /= 0) stop "input.cfg: bad real"
  end subroutine get_r

  subroutine get_rv(key, v)
    character(len=*), intent(in) :: key
    real(dp), intent(inout) :: v(:)
    integer :: i, ios
    i = cfg_find(key)
    if (i == 0) return
    read(cfg_vals(i), *, iostat=ios) v
    if (ios /= 0) stop "input.cfg: bad real array"
  end subroutine get_rv

  subroutine get_iv(key, v)
    character(len=*), intent(in) :: key
    integer, intent(inout) :: v(:)
    integer :: i, ios
    i = cfg_find(key)
    if (i == 0) return
    read(cfg_vals(i), *, iostat=ios) v
    if (ios /= 0) stop "input.cfg: bad integer array"
  end subroutine get_iv

  subroutine get_s(key, v)
    character(len=*), intent(in) :: key
    character(len=:), allocatable, intent(inout) :: v
    integer :: i
    character(len=val_len) :: s
    i = cfg_find(key)
    if (i == 0) return
    s = adjustl(cfg_vals(i))
    if (s(1:1) == "'" .or. s(1:1) == '"') s = s(2:)
    i = scan(s, "'"" ")
    if (i > 0) s = s(1:i - 1)
    v = trim(s)
  end subroutine get_s

  logical function init_parameters()
    ! input_paras.f90:79-293: read, convert units, set up k-space
    real(dp) :: a1(3), a2(3), a3(3), eta
    logical :: ok
    init_parameters = .false.
    call cfg_load("./input.cfg", ok)
    if (.not. ok) then
      write(*, '(a)') "[init_parameters]: could not find input file"
      stop "please provide a input.cfg file"
    end if
    call get_l("jobs%plot_bands", plot_bands)
    call get_l("jobs%debug_mode", debug_mode)
    call get_l("jobs%R_vect_float", use_R_float)
    call get_l("jobs%do_write_velo", do_write_velo)
    call get_l("jobs%do_mep", do_mep)
    call get_l("jobs%do_kubo", do_kubo)
    call get_l("jobs%do_ahc", do_ahc)
    call get_l("jobs%do_opt", do_opt)
    call get_l("jobs%do_gyro", do_gyro)
    a1 = 0.0_dp; a2 = 0.0_dp; a3 = 0.0_dp
    call get_rv("unitCell%a1", a1)
    call get_rv("unitCell%a2", a2)
    call get_rv("unitCell%a3", a3)
    call get_r("unitCell%a0", a0)
    seed_name = "wf1"
    call get_s("wannBase%seed_name", seed_name)
    call get_i("wannBase%N_wf", N_wf)
    if (N_wf > 0) then
      allocate(wf_centers(3, N_wf))
      wf_centers = 0.0_dp
      call get_rv("wannBase%wf_centers_x", wf_centers(1, :))
      call get_rv("wannBase%wf_centers_y", wf_centers(2, :))
      call get_rv("wannBase%wf_centers_z", wf_centers(3, :))
    end if
    call get_l("wannInterp%use_cart_velo", use_cart_velo)
    call get_l("wannInterp%doGaugeTrafo", do_gauge_trafo)
    call get_iv("wannInterp%mp_grid", mp_grid_in)
    call get_i("Fermi%N_eF", N_eF)
    call get_r("Fermi%eF_min", eF_min)
    call get_r("Fermi%eF_max", eF_max)
    call get_r("Fermi%Tkelvin", T_kelvin)
    eta = 0.0_dp
    call get_r("Fermi%eta_smearing", eta)
    call get_r("Fermi%kuboTol", kubo_tol)
    call get_i("MEP%valence_bands", valence_bands)
    call get_l("MEP%do_write_mep_bands", do_write_mep_bands)
    call get_i("Laser%N_hw", N_hw)
    call get_r("Laser%hw_min", hw_min)
    call get_r("Laser%hw_max", hw_max)
    ! lattice rows a_i (input_paras.f90:213-216), units (:171-181)
    a_latt(1, :) = a1
    a_latt(2, :) = a2
    a_latt(3, :) = a3
    a_latt = a0 * a_latt
    hw_min = hw_min / aUtoEv
    hw_max = hw_max / aUtoEv
    N_hw = max(1, N_hw)
    N_eF = max(1, N_eF)
    eF_min = eF_min / aUtoEv
    eF_max = eF_max / aUtoEv
    eta = eta / aUtoEv
    i_eta_smr = cmplx(0.0_dp, eta, dp)
    unit_vol = dot_product(crossP(a_latt(1, :), a_latt(2, :)), a_latt(3, :))
    call set_recip_latt(a_latt)
    call set_mp_grid(mp_grid_in)
    call my_mkdir(out_dir)
    init_parameters = .true.
  end function init_parameters

  ! mkdir -p through the C library (no child process)
  subroutine my_mkdir(dir)
    use iso_c_binding, only: c_int, c_char, c_null_char
    character(len=*), intent(in) :: dir
    interface
      integer(c_int) function c_mkdir(path, mode) bind(C, name="mkdir")
        import :: c_int, c_char
        character(kind=c_char), intent(in) :: path(*)
        integer(c_int), value :: mode
      end function c_mkdir
    end interface
    character(len=:), allocatable :: d
    integer :: i
    integer(c_int) :: rc
    d = "./" // trim(dir)
    do i = 2, len(d)
      if (d(i:i) == "/" .or. i == len(d)) rc = c_mkdir(d(1:i) // c_null_char, int(o'755', c_int))
    end do
  end subroutine my_mkdir
end module input_paras
