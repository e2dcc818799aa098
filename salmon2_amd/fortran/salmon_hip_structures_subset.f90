!
! Interface-compatible subset of SALMON's `structures` module, for building and
! testing the drop-in wrappers (salmon_hip_dropin.f90) outside SALMON.
!
! Only the derived types and components that the propagation hot path reads
! are declared, with SALMON's names and array shapes (src/common/structures.f90:
! s_scalar :36-38, s_dft_system :45-66, s_rgrid :86-94, s_parallel_info :127-146,
! s_orbital :191-196, s_stencil :198-203, s_pp_grid :254-319, s_pp_nlcc :321-324,
! s_xc_functional :327-344, s_reciprocal_grid :346-353, s_poisson :356-381,
! s_rt :472-492), so the same wrapper source compiles unchanged against SALMON's
! full module.
! Inside SALMON this file is not used.
!
module structures
  implicit none

  type s_scalar
    real(8), allocatable :: f(:,:,:)
  end type s_scalar

  type s_dft_system
    logical :: if_real_orbital = .false.
    integer :: ngrid, nspin, no, nk, nion
    real(8) :: hvol, hgs(3), primitive_a(3,3)
    real(8), allocatable :: vec_k(:,:)    ! (3, nk)
    real(8), allocatable :: wtk(:)        ! (nk)
    real(8), allocatable :: rocc(:,:,:)   ! (no, nk, nspin)
    real(8) :: vec_Ac(3)
  end type s_dft_system

  type s_rgrid
    integer :: ndir, Nd
    integer, dimension(3) :: is, ie, num, is_overlap, ie_overlap, is_array, ie_array
    integer, allocatable :: idx(:), idy(:), idz(:)
  end type s_rgrid

  type s_parallel_info
    logical :: if_divide_rspace = .false.
    integer :: im_s, im_e, numm
    integer :: ik_s, ik_e, numk
    integer :: io_s, io_e, numo
  end type s_parallel_info

  type s_sendrecv_grid
    integer :: unused = 0
  end type s_sendrecv_grid

  type s_orbital
    real(8), allocatable :: rwf(:,:,:,:,:,:,:)
    complex(8), allocatable :: zwf(:,:,:,:,:,:,:)   ! (ix,iy,iz,ispin,io,ik,im)
    logical :: update_zwf_overlap
  end type s_orbital

  type s_stencil
    logical :: if_orthogonal
    real(8) :: coef_lap0, coef_lap(4,3), coef_nab(4,3)
  end type s_stencil

  type s_pp_grid
    integer :: nps
    integer, allocatable :: mps(:)
    integer, allocatable :: jxyz(:,:,:)
    real(8), allocatable :: rxyz(:,:,:)
    real(8), allocatable :: uv(:,:)
    integer :: nlma
    integer, allocatable :: ia_tbl(:)
    real(8), allocatable :: rinv_uvu(:)
    complex(8), allocatable :: zekr_uv(:,:,:)
  end type s_pp_grid

  type s_pp_nlcc
    real(8), allocatable :: rho_nlcc(:,:,:)
  end type s_pp_nlcc

  type s_xc_functional
    integer :: xctype(3)
    integer :: ispin
  end type s_xc_functional

  type s_reciprocal_grid
    real(8), allocatable :: coef(:,:,:)
  end type s_reciprocal_grid

  type s_poisson
    integer :: iterVh
  end type s_poisson

  type s_rt
    complex(8), allocatable :: zc(:)
  end type s_rt

end module structures
