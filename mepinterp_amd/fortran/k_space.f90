! k_space.f90 -- Monkhorst-Pack mesh, reciprocal lattice and the BZ
! normalisation (k_space.f90:81-414 of the reference).
module k_space
  use constants, only: dp, pi_dp
  use matrix_math, only: crossP
  implicit none
  private
  public :: set_mp_grid, get_mp_grid, set_recip_latt, get_recip_latt, get_kpt_idx, get_rel_kpt, &
            get_cart_kpt, get_bz_vol, get_unit_vol, normalize_k_int, kspace_allocator, k_norm

  interface normalize_k_int
    module procedure d2_normalize, d3_normalize, z3_normalize, z4_normalize
  end interface normalize_k_int

  integer, save :: mp_grid(3) = 0
  real(dp), save :: recip_latt(3, 3) = 0.0_dp, bz_vol = 0.0_dp, unit_vol = 0.0_dp
  logical, save :: recip_latt_set = .false.

contains

  subroutine set_mp_grid(input_grid)
    integer, intent(in) :: input_grid(3)
    mp_grid = input_grid
  end subroutine set_mp_grid

  function get_mp_grid() result(g)
    integer :: g(3)
    g = mp_grid
  end function get_mp_grid

  subroutine set_recip_latt(a_latt)
    ! a_i = a_latt(:, i) and recip_latt(:, i) = b_i as in the reference
    real(dp), intent(in) :: a_latt(3, 3)
    real(dp) :: a1(3), a2(3), a3(3)
    a1 = a_latt(:, 1)
    a2 = a_latt(:, 2)
    a3 = a_latt(:, 3)
    unit_vol = dot_product(crossP(a1, a2), a3)
    recip_latt(:, 1) = 2.0_dp * pi_dp * crossP(a2, a3) / unit_vol
    recip_latt(:, 2) = 2.0_dp * pi_dp * crossP(a3, a1) / unit_vol
    recip_latt(:, 3) = 2.0_dp * pi_dp * crossP(a1, a2) / unit_vol
    bz_vol = dot_product(crossP(recip_latt(:, 1), recip_latt(:, 2)), recip_latt(:, 3))
    recip_latt_set = .true.
  end subroutine set_recip_latt

  function get_recip_latt() result(r)
    real(dp) :: r(3, 3)
    r = recip_latt
  end function get_recip_latt

  function get_bz_vol() result(v)
    real(dp) :: v
    v = bz_vol
  end function get_bz_vol

  function get_unit_vol() result(v)
    real(dp) :: v
    v = unit_vol
  end function get_unit_vol

  integer pure function get_kpt_idx(qix, qiy, qiz)
    integer, intent(in) :: qix, qiy, qiz
    get_kpt_idx = 1 + (qix - 1) + mp_grid(2) * ((qiy - 1) + mp_grid(3) * (qiz - 1))
  end function get_kpt_idx

  function get_rel_kpt(qi_idx, qix, qiy, qiz) result(kpt)
    integer, intent(in) :: qi_idx, qix, qiy, qiz
    real(dp) :: kpt(3)
    if (qi_idx < 0 .or. qi_idx > mp_grid(1) * mp_grid(2) * mp_grid(3)) &
      stop "[get_rel_kpt]: ERROR got illegal kpt idx"
    kpt(1) = (2.0_dp * real(qix, dp) - real(mp_grid(1), dp) - 1.0_dp) / (2.0_dp * real(mp_grid(1), dp))
    kpt(2) = (2.0_dp * real(qiy, dp) - real(mp_grid(2), dp) - 1.0_dp) / (2.0_dp * real(mp_grid(2), dp))
    kpt(3) = (2.0_dp * real(qiz, dp) - real(mp_grid(3), dp) - 1.0_dp) / (2.0_dp * real(mp_grid(3), dp))
  end function get_rel_kpt

  function get_cart_kpt(kpt) result(kpt_cart)
    real(dp), intent(in) :: kpt(3)
    real(dp) :: kpt_cart(3)
    if (.not. recip_latt_set) stop "[kspace/get_cart_kpt]: ERROR lattice not set"
    kpt_cart = matmul(recip_latt, kpt)
  end function get_cart_kpt

  real(dp) function k_norm()
    integer :: n_k
    n_k = mp_grid(1) * mp_grid(2) * mp_grid(3)
    k_norm = abs(unit_vol * real(n_k, dp))
  end function k_norm

  subroutine d2_normalize(t)
    real(dp), allocatable, intent(inout) :: t(:, :)
    if (allocated(t)) t = t / k_norm()
  end subroutine d2_normalize
  subroutine d3_normalize(t)
    real(dp), allocatable, intent(inout) :: t(:, :, :)
    if (allocated(t)) t = t / k_norm()
  end subroutine d3_normalize
  subroutine z3_normalize(t)
    complex(dp), allocatable, intent(inout) :: t(:, :, :)
    if (allocated(t)) t = t / k_norm()
  end subroutine z3_normalize
  subroutine z4_normalize(t)
    complex(dp), allocatable, intent(inout) :: t(:, :, :, :)
    if (allocated(t)) t = t / k_norm()
  end subroutine z4_normalize

  subroutine kspace_allocator(H_tb, r_tb, en_k, V_ka, A_ka, Om_kab)
    complex(dp), allocatable, intent(inout) :: H_tb(:, :, :), r_tb(:, :, :, :), V_ka(:, :, :), &
                                               A_ka(:, :, :), Om_kab(:, :, :, :)
    real(dp), allocatable, intent(inout) :: en_k(:)
    allocate(en_k(size(H_tb, 2)))
    allocate(V_ka(3, size(H_tb, 1), size(H_tb, 2)))
    if (allocated(r_tb)) then
      allocate(A_ka(3, size(r_tb, 2), size(r_tb, 3)))
      allocate(Om_kab(3, 3, size(r_tb, 2), size(r_tb, 3)))
    end if
  end subroutine kspace_allocator
end module k_space
