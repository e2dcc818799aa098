!
! salmon_hip_dropin.f90 -- SALMON-signature wrappers over the MI355X hot path.
!
! Each public routine keeps the argument list of the SALMON routine it
! replaces, so a SALMON build swaps them in by name:
!   taylor_hip                      <- taylor                    (src/rt/taylor.f90:21-38)
!   hpsi_hip                        <- hpsi                      (src/common/hamiltonian.f90:27-53)
!   update_kvector_nonlocalpt_hip   <- update_kvector_nonlocalpt (src/common/hamiltonian.f90:847-860)
!   calc_density_hip                <- calc_density              (src/common/density_matrix.f90:26-42)
!   calc_current_hip                <- calc_current              (src/common/density_matrix.f90:240-300)
! The orbital arguments are SALMON's host arrays; these wrappers copy them to
! and from device buffers owned by this module at every call (PCIe-inclusive,
! the simple drop-in).  A driver that keeps orbitals resident uses the
! salmon_hip_mod interfaces directly (see the RT step in salmon2_amd/rt.py).
!
! Supported: periodic single-process runs (no r-space division, no halo:
! is_array = is), one im, orthogonal lattice, collinear spin, no DFT+U -- the
! branch of hpsi the propagation takes for the benchmark systems.
!
module salmon_hip_dropin
  use iso_c_binding
  use salmon_hip_mod
  use structures
  implicit none
  private
  public :: salmon_hip_dropin_setup, salmon_hip_dropin_release
  public :: taylor_hip, hpsi_hip, update_kvector_nonlocalpt_hip, calc_density_hip, calc_current_hip

  type(c_ptr), save :: h = c_null_ptr
  type(c_ptr), save :: d_a = c_null_ptr, d_b = c_null_ptr, d_w = c_null_ptr
  type(c_ptr), save :: d_rho = c_null_ptr, d_cur = c_null_ptr
  integer(c_size_t), save :: orb_bytes = 0
  integer, save :: nx_s, ny_s, nz_s, nspin_s

contains

  ! Create the device handle for SALMON's grid, stencil, k-points and projectors.
  subroutine salmon_hip_dropin_setup(mg, system, info, stencil, ppg)
    type(s_rgrid), intent(in) :: mg
    type(s_dft_system), intent(in) :: system
    type(s_parallel_info), intent(in) :: info
    type(s_stencil), intent(in) :: stencil
    type(s_pp_grid), intent(in) :: ppg
    integer(c_int) :: ly, natom
    if (c_associated(h)) return
    if (info%if_divide_rspace) stop 'salmon_hip: r-space division is not supported'
    if (any(mg%is_array /= mg%is)) stop 'salmon_hip: halo arrays (is_array /= is) are not supported'
    if (.not. stencil%if_orthogonal) stop 'salmon_hip: non-orthogonal lattice is not supported'
    if (info%im_s /= info%im_e) stop 'salmon_hip: one im only'
    if (info%io_s /= 1 .or. info%io_e /= system%no .or. info%ik_s /= 1 .or. info%ik_e /= system%nk) &
      stop 'salmon_hip: orbital / k-point division is not supported'
    nx_s = mg%num(1)
    ny_s = mg%num(2)
    nz_s = mg%num(3)
    nspin_s = system%nspin
    ly = int(mg%ie_array(2) - mg%is_array(2) + 1, c_int)
    call salmon_hip_check(salmon_hip_create(int(nx_s, c_int), int(ny_s, c_int), int(nz_s, c_int), ly, &
        int(system%nspin, c_int), int(system%no, c_int), int(system%nk, c_int), stencil%coef_lap, &
        stencil%coef_nab, stencil%coef_lap0, system%vec_k, system%rocc, system%wtk, system%hgs, &
        system%hvol, h), 'create')
    natom = 0
    if (allocated(ppg%mps)) natom = int(size(ppg%mps), c_int)
    if (natom > 0 .and. ppg%nlma > 0) then
      call salmon_hip_check(salmon_hip_set_nonlocal(h, natom, int(ppg%nps, c_int), ppg%mps, ppg%jxyz, &
          ppg%rxyz, int(ppg%nlma, c_int), ppg%ia_tbl, ppg%rinv_uvu, ppg%uv), 'set_nonlocal')
    end if
    orb_bytes = int(nx_s, c_size_t) * int(ly, c_size_t) * int(nz_s, c_size_t) * int(system%nspin, c_size_t) &
              * int(system%no, c_size_t) * int(system%nk, c_size_t) * 16_c_size_t
    call salmon_hip_check(salmon_hip_malloc(d_a, orb_bytes), 'malloc')
    call salmon_hip_check(salmon_hip_malloc(d_b, orb_bytes), 'malloc')
    call salmon_hip_check(salmon_hip_malloc(d_w, orb_bytes), 'malloc')
    call salmon_hip_check(salmon_hip_malloc(d_rho, int(nx_s, c_size_t) * ny_s * nz_s * nspin_s * 8_c_size_t), &
        'malloc')
    call salmon_hip_check(salmon_hip_malloc(d_cur, int(3 * nspin_s, c_size_t) * 8_c_size_t), 'malloc')
  end subroutine salmon_hip_dropin_setup

  subroutine salmon_hip_dropin_release()
    if (.not. c_associated(h)) return
    call salmon_hip_check(salmon_hip_free(d_a), 'free')
    call salmon_hip_check(salmon_hip_free(d_b), 'free')
    call salmon_hip_check(salmon_hip_free(d_w), 'free')
    call salmon_hip_check(salmon_hip_free(d_rho), 'free')
    call salmon_hip_check(salmon_hip_free(d_cur), 'free')
    call salmon_hip_check(salmon_hip_destroy(h), 'destroy')
    h = c_null_ptr
  end subroutine salmon_hip_dropin_release

  ! V_local(ispin)%f(is:ie) -> one (nx,ny,nz,nspin) block on the device
  subroutine upload_vlocal(V_local)
    type(s_scalar), intent(in) :: V_local(:)
    real(8), allocatable :: v(:,:,:,:)
    integer :: is
    allocate(v(nx_s, ny_s, nz_s, nspin_s))
    do is = 1, nspin_s
      v(:,:,:,is) = V_local(is)%f
    end do
    call salmon_hip_check(salmon_hip_set_vlocal(h, v), 'set_vlocal')
    deallocate(v)
  end subroutine upload_vlocal

  subroutine update_kvector_nonlocalpt_hip(ik_s, ik_e, system, ppg)
    integer, intent(in) :: ik_s, ik_e
    type(s_dft_system), intent(in) :: system
    type(s_pp_grid) :: ppg
    if (.not. c_associated(h)) stop 'salmon_hip: salmon_hip_dropin_setup not called'
    if (ik_s /= 1 .or. ik_e /= system%nk) stop 'salmon_hip: k-point division is not supported'
    call salmon_hip_check(salmon_hip_update_kvector(h, system%vec_Ac), 'update_kvector')
  end subroutine update_kvector_nonlocalpt_hip

  subroutine taylor_hip(mg, system, info, stencil, srg, tspsi_in, tspsi_out, sshtpsi, ppg, V_local, rt)
    type(s_rgrid), intent(in) :: mg
    type(s_dft_system), intent(in) :: system
    type(s_parallel_info), intent(in) :: info
    type(s_stencil), intent(in) :: stencil
    type(s_sendrecv_grid), intent(inout) :: srg
    type(s_orbital), intent(inout) :: tspsi_in
    type(s_orbital), intent(inout) :: tspsi_out
    type(s_orbital), intent(inout) :: sshtpsi
    type(s_pp_grid), intent(in) :: ppg
    type(s_scalar), intent(in) :: V_local(system%nspin)
    type(s_rt), intent(in) :: rt
    call salmon_hip_dropin_setup(mg, system, info, stencil, ppg)
    call upload_vlocal(V_local)
    call salmon_hip_check(salmon_hip_set_vec_ac(h, system%vec_Ac), 'set_vec_ac')
    call salmon_hip_check(salmon_hip_memcpy_htod(h, d_a, tspsi_in%zwf, orb_bytes), 'upload tspsi_in')
    call salmon_hip_check(salmon_hip_taylor(h, rt%zc, int(size(rt%zc), c_int), d_a, d_b, d_w), 'taylor')
    call salmon_hip_check(salmon_hip_memcpy_dtoh(h, tspsi_out%zwf, d_b, orb_bytes), 'download tspsi_out')
    call salmon_hip_check(salmon_hip_memcpy_dtoh(h, sshtpsi%zwf, d_w, orb_bytes), 'download sshtpsi')
  end subroutine taylor_hip

  subroutine hpsi_hip(tpsi, htpsi, info, mg, V_local, system, stencil, srg, ppg, ttpsi)
    type(s_dft_system), intent(in) :: system
    type(s_parallel_info), intent(in) :: info
    type(s_rgrid), intent(in) :: mg
    type(s_scalar), intent(in) :: V_local(system%nspin)
    type(s_stencil), intent(in) :: stencil
    type(s_sendrecv_grid), intent(inout) :: srg
    type(s_pp_grid), intent(in) :: ppg
    type(s_orbital) :: tpsi, htpsi
    type(s_orbital), optional :: ttpsi
    integer :: is, io, ik, im
    call salmon_hip_dropin_setup(mg, system, info, stencil, ppg)
    call upload_vlocal(V_local)
    call salmon_hip_check(salmon_hip_set_vec_ac(h, system%vec_Ac), 'set_vec_ac')
    call salmon_hip_check(salmon_hip_memcpy_htod(h, d_a, tpsi%zwf, orb_bytes), 'upload tpsi')
    call salmon_hip_check(salmon_hip_hpsi(h, d_a, d_b), 'hpsi')
    call salmon_hip_check(salmon_hip_memcpy_dtoh(h, htpsi%zwf, d_b, orb_bytes), 'download htpsi')
    if (present(ttpsi)) then
      ! ttpsi = htpsi - V_local tpsi (hamiltonian.f90:396-413)
      do im = lbound(tpsi%zwf, 7), ubound(tpsi%zwf, 7)
      do ik = lbound(tpsi%zwf, 6), ubound(tpsi%zwf, 6)
      do io = lbound(tpsi%zwf, 5), ubound(tpsi%zwf, 5)
      do is = 1, system%nspin
        ttpsi%zwf(mg%is(1):mg%ie(1), mg%is(2):mg%ie(2), mg%is(3):mg%ie(3), is, io, ik, im) = &
            htpsi%zwf(mg%is(1):mg%ie(1), mg%is(2):mg%ie(2), mg%is(3):mg%ie(3), is, io, ik, im) &
            - V_local(is)%f * tpsi%zwf(mg%is(1):mg%ie(1), mg%is(2):mg%ie(2), mg%is(3):mg%ie(3), is, io, ik, im)
      end do
      end do
      end do
      end do
    end if
  end subroutine hpsi_hip

  subroutine calc_density_hip(system, rho, psi, info, mg)
    type(s_dft_system), intent(in) :: system
    type(s_parallel_info), intent(in) :: info
    type(s_rgrid), intent(in) :: mg
    type(s_orbital), intent(in) :: psi
    type(s_scalar) :: rho(system%nspin, info%im_s:info%im_e)
    real(8), allocatable :: r(:,:,:,:)
    integer :: is
    if (.not. c_associated(h)) stop 'salmon_hip: salmon_hip_dropin_setup not called'
    call salmon_hip_check(salmon_hip_memcpy_htod(h, d_a, psi%zwf, orb_bytes), 'upload psi')
    call salmon_hip_check(salmon_hip_density(h, d_a, d_rho), 'density')
    allocate(r(nx_s, ny_s, nz_s, nspin_s))
    call salmon_hip_check(salmon_hip_memcpy_dtoh(h, r, d_rho, int(size(r), c_size_t) * 8_c_size_t), 'download rho')
    do is = 1, system%nspin
      rho(is, info%im_s)%f(mg%is(1):mg%ie(1), mg%is(2):mg%ie(2), mg%is(3):mg%ie(3)) = r(:,:,:,is)
    end do
    deallocate(r)
  end subroutine calc_density_hip

  subroutine calc_current_hip(system, mg, stencil, info, srg, psi, ppg, curr)
    type(s_dft_system), intent(in) :: system
    type(s_rgrid), intent(in) :: mg
    type(s_stencil), intent(in) :: stencil
    type(s_parallel_info), intent(in) :: info
    type(s_sendrecv_grid) :: srg
    type(s_orbital) :: psi
    type(s_pp_grid), intent(in) :: ppg
    real(8) :: curr(3, system%nspin, info%im_s:info%im_e)
    real(8) :: c(3, system%nspin)
    if (.not. c_associated(h)) stop 'salmon_hip: salmon_hip_dropin_setup not called'
    call salmon_hip_check(salmon_hip_memcpy_htod(h, d_a, psi%zwf, orb_bytes), 'upload psi')
    call salmon_hip_check(salmon_hip_current(h, system%vec_Ac, d_a, d_cur), 'current')
    call salmon_hip_check(salmon_hip_memcpy_dtoh(h, c, d_cur, int(size(c), c_size_t) * 8_c_size_t), 'download curr')
    curr(:, :, info%im_s) = c
  end subroutine calc_current_hip

end module salmon_hip_dropin
