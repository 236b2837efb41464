! band_calc.f90 -- band-structure mode (band_calc.f90:30-121 of the
! reference): the k-points of ./kpts (kptsgen.pl format,
! file_io.f90:451-478) are interpolated on the GPU in one batched call
! (mep_interp_batch); the outputs keep the reference's files:
!   raw/enK.<k, 7 digits>   eigenvalues, unformatted stream (write_en_binary)
!   out/eBands.dat          k-index, fractional k, energy per band (write_en_global)
!   out//velo/velo_Vka.<k>  formatted V_ka when jobs%do_write_velo (write_velo)
module band_calc
  use iso_c_binding
  use constants, only: dp
  use mep_capi
  use mep_gpu, only: gpu_model_ctx
  use input_paras, only: seed_name, do_gauge_trafo, wf_centers, do_write_velo, raw_dir, out_dir, &
                         my_mkdir
  use file_io, only: read_tb_basis
  implicit none
  private
  public :: band_worker, read_kptsgen_pl_file, write_en_binary, write_en_global, write_velo

  character(len=*), parameter :: kfmt = '(a,i7.7)'

contains

  subroutine band_worker()
    real(dp), allocatable :: rel_kpts(:, :), R_vect(:, :)
    complex(dp), allocatable :: H_tb(:, :, :), r_tb(:, :, :, :)
    real(dp), allocatable, target :: kp(:, :), en(:, :)
    complex(dp), allocatable, target :: V(:, :, :, :), A(:, :, :, :), Om(:, :, :, :, :)
    type(c_ptr) :: ctx, vp, ap, op
    logical :: use_pos
    integer :: n, nk, ki
    write(*, '(a)') "***^^^^ - BANDSTRUCTURE MODE - ^^^^***"
    call my_mkdir(raw_dir)
    call my_mkdir(out_dir)
    if (do_write_velo) call my_mkdir(out_dir // "/velo/")
    if (.not. read_kptsgen_pl_file(rel_kpts)) &
      stop 'for bandstructure calculations a kpts file has to be provided'
    nk = size(rel_kpts, 2)
    call read_tb_basis(seed_name, R_vect, H_tb, r_tb)
    n = size(H_tb, 1)
    use_pos = allocated(r_tb)
    call gpu_model_ctx(do_gauge_trafo, H_tb, r_tb, use_pos, R_vect, wf_centers, ctx)
    allocate(kp(3, nk), en(n, nk))
    kp = rel_kpts
    vp = c_null_ptr
    ap = c_null_ptr
    op = c_null_ptr
    if (do_write_velo .or. use_pos) then
      allocate(V(3, n, n, nk))
      vp = c_loc(V)
      if (use_pos) then
        allocate(A(3, n, n, nk), Om(3, 3, n, n, nk))
        ap = c_loc(A)
        op = c_loc(Om)
      end if
    end if
    call mep_check(mep_interp_batch(ctx, int(nk, c_int64_t), c_loc(kp), c_loc(en), vp, ap, op), &
                   "band_worker")
    do ki = 1, nk
      call write_en_binary(ki, en(:, ki))
      if (do_write_velo) call write_velo(ki, V(:, :, :, ki))
    end do
    call write_en_global(rel_kpts, n)
    write(*, '(a,i10,a)') "[band_worker]: interpolated ", nk, " kpts"
  end subroutine band_worker

  ! ./kpts: "num_kpts <x>" then per line "kx ky kz weight" (kptsgen.pl)
  logical function read_kptsgen_pl_file(kpt_latt)
    real(dp), allocatable, intent(inout) :: kpt_latt(:, :)
    integer :: u, num_kpts, kpt
    real(dp) :: raw(3), tmp
    inquire(file='kpts', exist=read_kptsgen_pl_file)
    if (.not. read_kptsgen_pl_file) return
    open(newunit=u, file='kpts', form='formatted', action='read', status='old')
    read(u, *) num_kpts, raw(1)
    allocate(kpt_latt(3, num_kpts))
    do kpt = 1, num_kpts
      read(u, *) raw(1:3), tmp
      kpt_latt(:, kpt) = raw
    end do
    close(u)
    write(*, '(a,i8)') "[read_kptsgen_pl_file]: success! num_kpts", num_kpts
  end function read_kptsgen_pl_file

  subroutine write_en_binary(kpt_idx, e_bands)
    integer, intent(in) :: kpt_idx
    real(dp), intent(in) :: e_bands(:)
    character(len=24) :: filename
    integer :: u
    write(filename, kfmt) raw_dir // 'enK.', kpt_idx
    open(newunit=u, file=filename, form='unformatted', action='write', access='stream', &
         status='replace')
    write(u) e_bands(:)
    close(u)
  end subroutine write_en_binary

  subroutine write_en_global(kpt_latt, num_bands)
    real(dp), intent(in) :: kpt_latt(:, :)
    integer, intent(in) :: num_bands
    real(dp), allocatable :: ek(:)
    character(len=24) :: filename
    integer :: u, ur, qi, band, x
    allocate(ek(num_bands))
    open(newunit=u, file=out_dir // 'eBands.dat', form='formatted', action='write', &
         access='stream', status='replace')
    ! (1x,a): the list-directed layout of the reference without record wrapping
    write(u, '(1x,a)') '# energies interpolated by MEPinterp program, written by mepInterp (MI355X)'
    write(u, '(1x,a)') '# first 3 columns give the relative k-mesh, 4th column are the enegies'
    write(u, '(1x,a)') '# Kpt_idx  K_x (frac)       K_y (frac)        K_z (frac)       Energy (Hartree)  '
    do qi = 1, size(kpt_latt, 2)
      write(filename, kfmt) raw_dir // 'enK.', qi
      open(newunit=ur, file=filename, form='unformatted', action='read', access='stream', &
           status='old')
      read(ur) ek(:)
      close(ur)
      do band = 1, num_bands
        write(u, '(i7,a)', advance="no") qi, char(9)
        write(u, '(200(f16.8))', advance="no") (kpt_latt(x, qi), x=1, size(kpt_latt, 1))
        write(u, '(a,f18.8)') char(9), ek(band)
      end do
    end do
    close(u)
  end subroutine write_en_global

  subroutine write_velo(kpt_idx, V_ka)
    integer, intent(in) :: kpt_idx
    complex(dp), intent(in) :: V_ka(:, :, :)
    character(len=40) :: fpath
    integer :: u, x, n, m
    write(fpath, kfmt) out_dir // "/velo/" // "velo_Vka" // ".", kpt_idx
    open(newunit=u, file=trim(fpath), form='formatted', action='write', access='stream', &
         status='replace')
    write(u, '(1x,a)') "#interpolated velocities"
    write(u, '(1x,a)') '# x |	 n		|	m 	|	real(V^x_nm)  |		imag(Vx_nm)'
    write(u, '(1x,a)') '#-----------------------------------------------------------------------'
    do x = 1, 3
      do m = 1, size(V_ka, 3)
        do n = 1, size(V_ka, 2)
          write(u, '(a,i1,a,i3,a,i3,a,f16.8,a,f16.8)') char(9), x, char(9)//char(9), n, &
            char(9)//char(9), m, char(9)//char(9), real(V_ka(x, n, m), dp), char(9), aimag(V_ka(x, n, m))
        end do
      end do
      write(u, '(1x,a)') "#----------------------------------------------------------------------"
    end do
    close(u)
  end subroutine write_velo

end module band_calc
