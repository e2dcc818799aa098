!
! salmon_hip_mod.f90 -- iso_c_binding interfaces to the C ABI of
! include/salmon_hip.h (the MI355X real-time propagation hot path).
!
! Modelled on the bind(c) interfaces SALMON already carries for its CUDA
! kernels (src/common/nonlocal_potential.f90:194-225 zpseudo_cuda,
! src/common/density_matrix.f90:254-286 stencil_current_core_gpu): scalar
! bounds by value, arrays by address.  Device buffers are opaque c_ptr
! handles.  Every function returns 0 on success; salmon_hip_check turns a
! failure into a stop with the library's message.
!
module salmon_hip_mod
  use iso_c_binding
  implicit none
  private :: c_strlen

  interface
    function salmon_hip_last_error() bind(c, name='salmon_hip_last_error') result(p)
      import
      type(c_ptr) :: p
    end function

    function salmon_hip_create(nx, ny, nz, ly, nspin, no, nk, coef_lap, coef_nab, coef_lap0, &
                               vec_k, rocc, wtk, hgs, hvol, h) bind(c, name='salmon_hip_create') result(rc)
      import
      integer(c_int), value :: nx, ny, nz, ly, nspin, no, nk
      real(c_double), intent(in) :: coef_lap(12), coef_nab(12)
      real(c_double), value :: coef_lap0
      real(c_double), intent(in) :: vec_k(*), rocc(*), wtk(*), hgs(3)
      real(c_double), value :: hvol
      type(c_ptr), intent(out) :: h
      integer(c_int) :: rc
    end function

    function salmon_hip_destroy(h) bind(c, name='salmon_hip_destroy') result(rc)
      import
      type(c_ptr), value :: h
      integer(c_int) :: rc
    end function

    function salmon_hip_set_nonlocal(h, natom, nps, mps, jxyz, rxyz, nlma, ia_tbl, rinv_uvu, uv) &
        bind(c, name='salmon_hip_set_nonlocal') result(rc)
      import
      type(c_ptr), value :: h
      integer(c_int), value :: natom, nps, nlma
      integer(c_int), intent(in) :: mps(*), jxyz(*), ia_tbl(*)
      real(c_double), intent(in) :: rxyz(*), rinv_uvu(*), uv(*)
      integer(c_int) :: rc
    end function

    function salmon_hip_set_vlocal(h, vlocal) bind(c, name='salmon_hip_set_vlocal') result(rc)
      import
      type(c_ptr), value :: h
      real(c_double), intent(in) :: vlocal(*)
      integer(c_int) :: rc
    end function

    function salmon_hip_update_kvector(h, vec_ac) bind(c, name='salmon_hip_update_kvector') result(rc)
      import
      type(c_ptr), value :: h
      real(c_double), intent(in) :: vec_ac(3)
      integer(c_int) :: rc
    end function

    function salmon_hip_set_vec_ac(h, vec_ac) bind(c, name='salmon_hip_set_vec_ac') result(rc)
      import
      type(c_ptr), value :: h
      real(c_double), intent(in) :: vec_ac(3)
      integer(c_int) :: rc
    end function

    function salmon_hip_hpsi(h, tpsi, htpsi) bind(c, name='salmon_hip_hpsi') result(rc)
      import
      type(c_ptr), value :: h, tpsi, htpsi
      integer(c_int) :: rc
    end function

    function salmon_hip_taylor(h, zc, n_hamil, tspsi_in, tspsi_out, sshtpsi) bind(c, name='salmon_hip_taylor') result(rc)
      import
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: zc(*)
      integer(c_int), value :: n_hamil
      type(c_ptr), value :: tspsi_in, tspsi_out, sshtpsi
      integer(c_int) :: rc
    end function

    function salmon_hip_density(h, psi, rho_dev) bind(c, name='salmon_hip_density') result(rc)
      import
      type(c_ptr), value :: h, psi, rho_dev
      integer(c_int) :: rc
    end function

    function salmon_hip_current(h, vec_ac, psi, curr_dev) bind(c, name='salmon_hip_current') result(rc)
      import
      type(c_ptr), value :: h
      real(c_double), intent(in) :: vec_ac(3)
      type(c_ptr), value :: psi, curr_dev
      integer(c_int) :: rc
    end function

    function salmon_hip_dipole(h, rho_dev, out_dev) bind(c, name='salmon_hip_dipole') result(rc)
      import
      type(c_ptr), value :: h, rho_dev, out_dev
      integer(c_int) :: rc
    end function

    function salmon_hip_set_poisson(h, al) bind(c, name='salmon_hip_set_poisson') result(rc)
      import
      type(c_ptr), value :: h
      real(c_double), intent(in) :: al(3)
      integer(c_int) :: rc
    end function

    function salmon_hip_hartree(h, rho_dev, vh_dev) bind(c, name='salmon_hip_hartree') result(rc)
      import
      type(c_ptr), value :: h, rho_dev, vh_dev
      integer(c_int) :: rc
    end function

    function salmon_hip_xc_pz(h, rho_dev, vxc_dev, exc_dev) bind(c, name='salmon_hip_xc_pz') result(rc)
      import
      type(c_ptr), value :: h, rho_dev, vxc_dev, exc_dev
      integer(c_int) :: rc
    end function

    function salmon_hip_update_vlocal(h, vpsl_dev, vh_dev, vxc_dev, vmix_dev, vout_dev) &
        bind(c, name='salmon_hip_update_vlocal') result(rc)
      import
      type(c_ptr), value :: h, vpsl_dev, vh_dev, vxc_dev, vmix_dev, vout_dev
      integer(c_int) :: rc
    end function

    function salmon_hip_update_potential(h, rho_dev, vpsl_dev, vmix_dev, vout_dev) &
        bind(c, name='salmon_hip_update_potential') result(rc)
      import
      type(c_ptr), value :: h, rho_dev, vpsl_dev, vmix_dev, vout_dev
      integer(c_int) :: rc
    end function

    function salmon_hip_vlocal_extrapolate(h, vold1_dev, vold2_dev, vout_dev) &
        bind(c, name='salmon_hip_vlocal_extrapolate') result(rc)
      import
      type(c_ptr), value :: h, vold1_dev, vold2_dev, vout_dev
      integer(c_int) :: rc
    end function

    function salmon_hip_set_vlocal_device(h, vlocal_dev) bind(c, name='salmon_hip_set_vlocal_device') result(rc)
      import
      type(c_ptr), value :: h, vlocal_dev
      integer(c_int) :: rc
    end function

    function salmon_hip_taylor_reuse(h, zc, n_hamil, tspsi_in, tspsi_out, sshtpsi, dv_dev) &
        bind(c, name='salmon_hip_taylor_reuse') result(rc)
      import
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: zc(*)
      integer(c_int), value :: n_hamil
      type(c_ptr), value :: tspsi_in, tspsi_out, sshtpsi, dv_dev
      integer(c_int) :: rc
    end function

    ! mode: SALMON_HIP_PREC_FP64 = 0, SALMON_HIP_PREC_MIXED = 1
    function salmon_hip_set_precision(h, mode) bind(c, name='salmon_hip_set_precision') result(rc)
      import
      type(c_ptr), value :: h
      integer(c_int), value :: mode
      integer(c_int) :: rc
    end function

    function salmon_hip_synchronize(h) bind(c, name='salmon_hip_synchronize') result(rc)
      import
      type(c_ptr), value :: h
      integer(c_int) :: rc
    end function

    function salmon_hip_malloc(dptr, bytes) bind(c, name='salmon_hip_malloc') result(rc)
      import
      type(c_ptr), intent(out) :: dptr
      integer(c_size_t), value :: bytes
      integer(c_int) :: rc
    end function

    function salmon_hip_free(dptr) bind(c, name='salmon_hip_free') result(rc)
      import
      type(c_ptr), value :: dptr
      integer(c_int) :: rc
    end function

    ! host <-> device copies of whole arrays (assumed-type: any host array by address)
    function salmon_hip_memcpy_htod(h, dst, src, bytes) bind(c, name='salmon_hip_memcpy_htod') result(rc)
      import
      type(c_ptr), value :: h, dst
      type(*), dimension(*), intent(in) :: src
      integer(c_size_t), value :: bytes
      integer(c_int) :: rc
    end function

    function salmon_hip_memcpy_dtoh(h, dst, src, bytes) bind(c, name='salmon_hip_memcpy_dtoh') result(rc)
      import
      type(c_ptr), value :: h, src
      type(*), dimension(*), intent(inout) :: dst
      integer(c_size_t), value :: bytes
      integer(c_int) :: rc
    end function

    function c_strlen(s) bind(c, name='strlen') result(n)
      import
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function
  end interface

contains

  ! stop with the library's message when rc /= 0 (SALMON reports fatal states with stop)
  subroutine salmon_hip_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=:), allocatable :: text
    integer :: n, i
    if (rc == 0) return
    n = int(c_strlen(salmon_hip_last_error()))
    call c_f_pointer(salmon_hip_last_error(), msg, [n])
    allocate(character(len=n) :: text)
    do i = 1, n
      text(i:i) = msg(i)
    end do
    write(*, '(a,a,a,i0,a,a)') 'salmon_hip: ', what, ' failed (code ', rc, '): ', text
    stop 'salmon_hip error'
  end subroutine salmon_hip_check

end module salmon_hip_mod
