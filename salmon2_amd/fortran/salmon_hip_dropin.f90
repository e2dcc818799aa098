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
!   hartree_hip                     <- hartree                   (src/poisson/hartree.f90:24; periodic:
!                                                                 poisson_ft, poisson_periodic.f90:23-242)
!   exchange_correlation_hip        <- exchange_correlation      (src/xc/salmon_xc.f90:58-343), xc = 'PZ'
!   update_vlocal_hip               <- update_vlocal             (src/common/hamiltonian.f90:813-843)
!
! The three potential wrappers take and return SALMON's host s_scalar arrays (a few
! hundred KiB per call); update_vlocal_hip also leaves the new V_local on the device
! as the handle's potential, so the next taylor_hip / hpsi_hip uploads nothing.
!
! Residency (the reference's OpenACC path: `!$acc enter data copyin(spsi_in,
! spsi_out,tpsi)` in src/rt/initialization_rt.f90:234-240, then `present`
! clauses): salmon_hip_enter_data(orb) gives an s_orbital a device mirror,
! keyed by the address of its zwf array, and uploads it.  A wrapper whose
! orbital arguments have mirrors works on them in place -- no host<->device
! copy of the orbitals -- and the host arrays are then stale until
! salmon_hip_update_host(orb) (`!$acc update host`); salmon_hip_update_device
! re-uploads after a host write; salmon_hip_exit_data drops the mirror.
! Orbital arguments without a mirror are copied in and out at the call (the
! simple drop-in).  rho and curr are host outputs and are downloaded at every
! call; V_local is uploaded by taylor_hip/hpsi_hip only when it differs from
! the last upload.
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
  public :: salmon_hip_enter_data, salmon_hip_exit_data, salmon_hip_update_host, salmon_hip_update_device
  public :: salmon_hip_dropin_sync
  public :: taylor_hip, hpsi_hip, update_kvector_nonlocalpt_hip, calc_density_hip, calc_current_hip
  public :: hartree_hip, exchange_correlation_hip, update_vlocal_hip

  type(c_ptr), save :: h = c_null_ptr
  type(c_ptr), save :: d_a = c_null_ptr, d_b = c_null_ptr, d_w = c_null_ptr
  type(c_ptr), save :: d_rho = c_null_ptr, d_cur = c_null_ptr
  integer(c_size_t), save :: orb_bytes = 0
  integer, save :: nx_s, ny_s, nz_s, nspin_s
  ! device mirrors of s_orbital arrays: key = address of the host zwf
  integer, parameter :: kmaxres = 16
  integer(c_intptr_t), save :: res_key(kmaxres) = 0
  type(c_ptr), save :: res_dev(kmaxres)
  ! the last V_local uploaded (uploads are skipped while it is unchanged)
  real(8), allocatable, save :: v_last(:,:,:,:)
  logical, save :: have_v = .false.
  ! potential wrappers: device buffers (first use) and the cell of the last set_poisson
  type(c_ptr), save :: d_vh = c_null_ptr, d_vxc = c_null_ptr, d_exc = c_null_ptr, d_vpsl = c_null_ptr
  type(c_ptr), save :: d_vout = c_null_ptr
  real(8), save :: al_last(3) = 0d0
  ! salmon_xctype_pz / salmon_xctype_none (src/xc/salmon_xc.f90:42-43)
  integer, parameter :: xctype_none = 0, xctype_pz = 1

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
    call salmon_hip_check(salmon_hip_malloc(d_rho, int(nx_s, c_size_t) * ny_s * nz_s * nspin_s * 8_c_size_t), &
        'malloc')
    call salmon_hip_check(salmon_hip_malloc(d_cur, int(3 * nspin_s, c_size_t) * 8_c_size_t), 'malloc')
    allocate(v_last(nx_s, ny_s, nz_s, nspin_s))
    have_v = .false.
  end subroutine salmon_hip_dropin_setup

  subroutine salmon_hip_dropin_release()
    integer :: i
    if (.not. c_associated(h)) return
    call salmon_hip_check(salmon_hip_synchronize(h), 'synchronize')
    do i = 1, kmaxres
      if (res_key(i) /= 0) call salmon_hip_check(salmon_hip_free(res_dev(i)), 'free')
      res_key(i) = 0
    end do
    if (c_associated(d_a)) call salmon_hip_check(salmon_hip_free(d_a), 'free')
    if (c_associated(d_b)) call salmon_hip_check(salmon_hip_free(d_b), 'free')
    if (c_associated(d_w)) call salmon_hip_check(salmon_hip_free(d_w), 'free')
    d_a = c_null_ptr; d_b = c_null_ptr; d_w = c_null_ptr
    call salmon_hip_check(salmon_hip_free(d_rho), 'free')
    call salmon_hip_check(salmon_hip_free(d_cur), 'free')
    if (c_associated(d_vh)) then
      call salmon_hip_check(salmon_hip_free(d_vh), 'free')
      call salmon_hip_check(salmon_hip_free(d_vxc), 'free')
      call salmon_hip_check(salmon_hip_free(d_exc), 'free')
      call salmon_hip_check(salmon_hip_free(d_vpsl), 'free')
      call salmon_hip_check(salmon_hip_free(d_vout), 'free')
    end if
    d_vh = c_null_ptr; d_vxc = c_null_ptr; d_exc = c_null_ptr; d_vpsl = c_null_ptr; d_vout = c_null_ptr
    al_last = 0d0
    call salmon_hip_check(salmon_hip_destroy(h), 'destroy')
    if (allocated(v_last)) deallocate(v_last)
    h = c_null_ptr
  end subroutine salmon_hip_dropin_release

  ! Wait for every enqueued device step (host timers, host reads of resident results).
  subroutine salmon_hip_dropin_sync()
    if (.not. c_associated(h)) stop 'salmon_hip: salmon_hip_dropin_setup not called'
    call salmon_hip_check(salmon_hip_synchronize(h), 'synchronize')
  end subroutine salmon_hip_dropin_sync

  integer(c_intptr_t) function key_of(orb)
    type(s_orbital), intent(in), target :: orb
    key_of = transfer(c_loc(orb%zwf), 0_c_intptr_t)
  end function key_of

  ! Device mirror of orb, or c_null_ptr.
  type(c_ptr) function mirror(orb)
    type(s_orbital), intent(in), target :: orb
    integer(c_intptr_t) :: k
    integer :: i
    mirror = c_null_ptr
    if (.not. allocated(orb%zwf)) return
    k = key_of(orb)
    do i = 1, kmaxres
      if (res_key(i) == k) then
        mirror = res_dev(i)
        return
      end if
    end do
  end function mirror

  subroutine check_shape(orb)
    type(s_orbital), intent(in) :: orb
    if (.not. allocated(orb%zwf)) stop 'salmon_hip: orbital array not allocated'
    if (int(size(orb%zwf), c_size_t) * 16_c_size_t /= orb_bytes) stop 'salmon_hip: orbital array shape'
  end subroutine check_shape

  ! `!$acc enter data copyin(orb)`: a device mirror of orb%zwf, uploaded.
  subroutine salmon_hip_enter_data(orb)
    type(s_orbital), intent(in), target :: orb
    integer :: i
    if (.not. c_associated(h)) stop 'salmon_hip: salmon_hip_dropin_setup before enter_data'
    call check_shape(orb)
    if (c_associated(mirror(orb))) return
    do i = 1, kmaxres
      if (res_key(i) == 0) then
        call salmon_hip_check(salmon_hip_malloc(res_dev(i), orb_bytes), 'malloc (enter_data)')
        call salmon_hip_check(salmon_hip_memcpy_htod(h, res_dev(i), orb%zwf, orb_bytes), 'upload (enter_data)')
        res_key(i) = key_of(orb)
        return
      end if
    end do
    stop 'salmon_hip: too many resident orbital arrays'
  end subroutine salmon_hip_enter_data

  ! `!$acc exit data`: drop the mirror (copyout = .true.: download it first).
  subroutine salmon_hip_exit_data(orb, copyout)
    type(s_orbital), intent(inout), target :: orb
    logical, intent(in), optional :: copyout
    integer(c_intptr_t) :: k
    integer :: i
    if (.not. c_associated(h)) return
    k = key_of(orb)
    do i = 1, kmaxres
      if (res_key(i) == k) then
        if (present(copyout)) then
          if (copyout) call salmon_hip_check(salmon_hip_memcpy_dtoh(h, orb%zwf, res_dev(i), orb_bytes), 'download')
        end if
        call salmon_hip_check(salmon_hip_synchronize(h), 'synchronize')
        call salmon_hip_check(salmon_hip_free(res_dev(i)), 'free (exit_data)')
        res_key(i) = 0
        return
      end if
    end do
  end subroutine salmon_hip_exit_data

  ! `!$acc update host(orb)`
  subroutine salmon_hip_update_host(orb)
    type(s_orbital), intent(inout), target :: orb
    type(c_ptr) :: d
    d = mirror(orb)
    if (.not. c_associated(d)) stop 'salmon_hip: update_host of an array without a device mirror'
    call salmon_hip_check(salmon_hip_memcpy_dtoh(h, orb%zwf, d, orb_bytes), 'update_host')
  end subroutine salmon_hip_update_host

  ! `!$acc update device(orb)`
  subroutine salmon_hip_update_device(orb)
    type(s_orbital), intent(in), target :: orb
    type(c_ptr) :: d
    d = mirror(orb)
    if (.not. c_associated(d)) stop 'salmon_hip: update_device of an array without a device mirror'
    call salmon_hip_check(salmon_hip_memcpy_htod(h, d, orb%zwf, orb_bytes), 'update_device')
  end subroutine salmon_hip_update_device

  ! Device address of orb: its mirror, else `scratch` (allocated on first use) with orb
  ! uploaded when `upload`.
  type(c_ptr) function on_device(orb, scratch, upload) result(d)
    type(s_orbital), intent(in), target :: orb
    type(c_ptr), intent(inout) :: scratch
    logical, intent(in) :: upload
    d = mirror(orb)
    if (c_associated(d)) return
    call check_shape(orb)
    if (.not. c_associated(scratch)) call salmon_hip_check(salmon_hip_malloc(scratch, orb_bytes), 'malloc')
    d = scratch
    if (upload) call salmon_hip_check(salmon_hip_memcpy_htod(h, d, orb%zwf, orb_bytes), 'upload')
  end function on_device

  ! Download orb from d unless d is orb's own mirror.
  subroutine to_host(orb, d)
    type(s_orbital), intent(inout), target :: orb
    type(c_ptr), intent(in) :: d
    if (c_associated(mirror(orb), d)) return
    call salmon_hip_check(salmon_hip_memcpy_dtoh(h, orb%zwf, d, orb_bytes), 'download')
  end subroutine to_host

  ! V_local(ispin)%f(is:ie) -> one (nx,ny,nz,nspin) block on the device, when it changed
  subroutine upload_vlocal(V_local)
    type(s_scalar), intent(in) :: V_local(:)
    integer :: is
    logical :: same
    same = have_v
    do is = 1, nspin_s
      if (same) same = all(v_last(:,:,:,is) == V_local(is)%f)
    end do
    if (same) return
    do is = 1, nspin_s
      v_last(:,:,:,is) = V_local(is)%f
    end do
    call salmon_hip_check(salmon_hip_set_vlocal(h, v_last), 'set_vlocal')
    have_v = .true.
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
    type(s_orbital), intent(inout), target :: tspsi_in
    type(s_orbital), intent(inout), target :: tspsi_out
    type(s_orbital), intent(inout), target :: sshtpsi
    type(s_pp_grid), intent(in) :: ppg
    type(s_scalar), intent(in) :: V_local(system%nspin)
    type(s_rt), intent(in) :: rt
    type(c_ptr) :: din, dout, dw
    call salmon_hip_dropin_setup(mg, system, info, stencil, ppg)
    call upload_vlocal(V_local)
    call salmon_hip_check(salmon_hip_set_vec_ac(h, system%vec_Ac), 'set_vec_ac')
    din = on_device(tspsi_in, d_a, .true.)
    dout = on_device(tspsi_out, d_b, .false.)
    dw = on_device(sshtpsi, d_w, .false.)
    call salmon_hip_check(salmon_hip_taylor(h, rt%zc, int(size(rt%zc), c_int), din, dout, dw), 'taylor')
    call to_host(tspsi_out, dout)
    call to_host(sshtpsi, dw)
  end subroutine taylor_hip

  subroutine hpsi_hip(tpsi, htpsi, info, mg, V_local, system, stencil, srg, ppg, ttpsi)
    type(s_dft_system), intent(in) :: system
    type(s_parallel_info), intent(in) :: info
    type(s_rgrid), intent(in) :: mg
    type(s_scalar), intent(in) :: V_local(system%nspin)
    type(s_stencil), intent(in) :: stencil
    type(s_sendrecv_grid), intent(inout) :: srg
    type(s_pp_grid), intent(in) :: ppg
    type(s_orbital), target :: tpsi, htpsi
    type(s_orbital), optional :: ttpsi
    type(c_ptr) :: din, dout
    integer :: is, io, ik, im
    call salmon_hip_dropin_setup(mg, system, info, stencil, ppg)
    call upload_vlocal(V_local)
    call salmon_hip_check(salmon_hip_set_vec_ac(h, system%vec_Ac), 'set_vec_ac')
    din = on_device(tpsi, d_a, .true.)
    dout = on_device(htpsi, d_b, .false.)
    call salmon_hip_check(salmon_hip_hpsi(h, din, dout), 'hpsi')
    call to_host(htpsi, dout)
    if (present(ttpsi)) then
      ! ttpsi = htpsi - V_local tpsi (hamiltonian.f90:396-413), on the host arrays
      if (c_associated(mirror(tpsi)) .or. c_associated(mirror(htpsi))) &
        stop 'salmon_hip: hpsi with ttpsi needs host orbitals (no device mirror)'
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
    type(s_orbital), intent(in), target :: psi
    type(s_scalar) :: rho(system%nspin, info%im_s:info%im_e)
    real(8), allocatable :: r(:,:,:,:)
    type(c_ptr) :: d
    integer :: is
    if (.not. c_associated(h)) stop 'salmon_hip: salmon_hip_dropin_setup not called'
    d = on_device(psi, d_a, .true.)
    call salmon_hip_check(salmon_hip_density(h, d, d_rho), 'density')
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
    type(s_orbital), target :: psi
    type(s_pp_grid), intent(in) :: ppg
    real(8) :: curr(3, system%nspin, info%im_s:info%im_e)
    real(8) :: c(3, system%nspin)
    type(c_ptr) :: d
    if (.not. c_associated(h)) stop 'salmon_hip: salmon_hip_dropin_setup not called'
    d = on_device(psi, d_a, .true.)
    call salmon_hip_check(salmon_hip_current(h, system%vec_Ac, d, d_cur), 'current')
    call salmon_hip_check(salmon_hip_memcpy_dtoh(h, c, d_cur, int(size(c), c_size_t) * 8_c_size_t), 'download curr')
    curr(:, :, info%im_s) = c
  end subroutine calc_current_hip

  subroutine pot_buffers()
    integer(c_size_t) :: g
    if (.not. c_associated(h)) stop 'salmon_hip: salmon_hip_dropin_setup not called'
    if (c_associated(d_vh)) return
    g = int(nx_s, c_size_t) * int(ny_s, c_size_t) * int(nz_s, c_size_t) * 8_c_size_t
    call salmon_hip_check(salmon_hip_malloc(d_vh, g), 'malloc')
    call salmon_hip_check(salmon_hip_malloc(d_vxc, g * nspin_s), 'malloc')
    call salmon_hip_check(salmon_hip_malloc(d_exc, g), 'malloc')
    call salmon_hip_check(salmon_hip_malloc(d_vpsl, g), 'malloc')
    call salmon_hip_check(salmon_hip_malloc(d_vout, g * nspin_s), 'malloc')
  end subroutine pot_buffers

  ! f(is:ie) of one s_scalar -> the device array at d (one channel, nx*ny*nz doubles)
  subroutine scalar_up(mg, x, d)
    type(s_rgrid), intent(in) :: mg
    type(s_scalar), intent(in) :: x
    type(c_ptr), intent(in) :: d
    real(8), allocatable :: t(:,:,:)
    allocate(t(nx_s, ny_s, nz_s))
    t = x%f(mg%is(1):mg%ie(1), mg%is(2):mg%ie(2), mg%is(3):mg%ie(3))
    call salmon_hip_check(salmon_hip_memcpy_htod(h, d, t, int(size(t), c_size_t) * 8_c_size_t), 'upload')
  end subroutine scalar_up

  subroutine scalar_down(mg, d, x)
    type(s_rgrid), intent(in) :: mg
    type(c_ptr), intent(in) :: d
    type(s_scalar), intent(inout) :: x
    real(8), allocatable :: t(:,:,:)
    allocate(t(nx_s, ny_s, nz_s))
    call salmon_hip_check(salmon_hip_memcpy_dtoh(h, t, d, int(size(t), c_size_t) * 8_c_size_t), 'download')
    x%f(mg%is(1):mg%ie(1), mg%is(2):mg%ie(2), mg%is(3):mg%ie(3)) = t
  end subroutine scalar_down

  ! address of channel `is` (1-based) of a device array of nspin channels
  type(c_ptr) function channel(d, is)
    type(c_ptr), intent(in) :: d
    integer, intent(in) :: is
    channel = transfer(transfer(d, 0_c_intptr_t) + int(is - 1, c_intptr_t) * int(nx_s, c_intptr_t) &
                       * int(ny_s, c_intptr_t) * int(nz_s, c_intptr_t) * 8_c_intptr_t, c_null_ptr)
  end function channel

  ! Vh of the total density rho (periodic orthogonal cell: the FFT Poisson solver of
  ! poisson_ft, with the cell lengths from system%primitive_a)
  subroutine hartree_hip(lg, mg, info, system, fg, poisson, srg_scalar, stencil, rho, Vh)
    type(s_rgrid), intent(in) :: lg, mg
    type(s_parallel_info), intent(in) :: info
    type(s_dft_system), intent(in) :: system
    type(s_reciprocal_grid), intent(in) :: fg
    type(s_poisson), intent(inout) :: poisson
    type(s_sendrecv_grid), intent(inout) :: srg_scalar
    type(s_stencil), intent(in) :: stencil
    type(s_scalar), intent(in) :: rho
    type(s_scalar), intent(inout) :: Vh
    real(8) :: al(3)
    integer :: i, j
    call pot_buffers()
    if (any(lg%num /= mg%num)) stop 'salmon_hip: r-space division is not supported'
    do j = 1, 3
      do i = 1, 3
        if (i /= j .and. system%primitive_a(i, j) /= 0d0) stop 'salmon_hip: non-orthogonal cell in hartree_hip'
      end do
      al(j) = system%primitive_a(j, j)
    end do
    if (any(al /= al_last)) then
      call salmon_hip_check(salmon_hip_set_poisson(h, al), 'set_poisson')
      al_last = al
    end if
    call scalar_up(mg, rho, d_rho)
    call salmon_hip_check(salmon_hip_hartree(h, d_rho, d_vh), 'hartree')
    call scalar_down(mg, d_vh, Vh)
  end subroutine hartree_hip

  ! Vxc(ispin) and E_xc of PZ81 (nspin = 1) or its spin-polarised form (nspin = 2) from
  ! the channel densities rho_s; E_xc = hvol sum_r e_xc trho in grid order
  ! (salmon_xc.f90:234-253; trho = rho, or rho_1 + rho_2 + 1e-20 as builtin_pz_sp.f90:35)
  subroutine exchange_correlation_hip(system, xc_func, mg, srg_scalar, srg, rho_s, ppn, info, spsi, stencil, &
                                      Vxc, E_xc)
    type(s_dft_system), intent(in) :: system
    type(s_xc_functional), intent(in) :: xc_func
    type(s_rgrid), intent(in) :: mg
    type(s_sendrecv_grid), intent(inout) :: srg_scalar, srg
    type(s_scalar), intent(in) :: rho_s(system%nspin)
    type(s_pp_nlcc), intent(in) :: ppn
    type(s_parallel_info), intent(in) :: info
    type(s_orbital), intent(in) :: spsi
    type(s_stencil), intent(in) :: stencil
    type(s_scalar), intent(inout) :: Vxc(system%nspin)
    real(8), intent(out) :: E_xc
    real(8), allocatable :: e(:,:,:), r(:,:,:,:)
    real(8) :: tot, trho
    integer :: is, ix, iy, iz
    call pot_buffers()
    if (xc_func%xctype(1) /= xctype_pz .or. xc_func%xctype(2) /= xctype_none .or. &
        xc_func%xctype(3) /= xctype_none) stop 'salmon_hip: exchange_correlation_hip supports xc = PZ only'
    if (xc_func%ispin /= system%nspin - 1) stop 'salmon_hip: xc_func%ispin does not match nspin'
    if (allocated(ppn%rho_nlcc)) then
      if (any(ppn%rho_nlcc /= 0d0)) stop 'salmon_hip: nonlinear core correction is not supported'
    end if
    do is = 1, nspin_s
      call scalar_up(mg, rho_s(is), channel(d_rho, is))
    end do
    call salmon_hip_check(salmon_hip_xc_pz(h, d_rho, d_vxc, d_exc), 'xc_pz')
    do is = 1, nspin_s
      call scalar_down(mg, channel(d_vxc, is), Vxc(is))
    end do
    allocate(e(nx_s, ny_s, nz_s), r(nx_s, ny_s, nz_s, nspin_s))
    call salmon_hip_check(salmon_hip_memcpy_dtoh(h, e, d_exc, int(size(e), c_size_t) * 8_c_size_t), 'download')
    do is = 1, nspin_s
      r(:,:,:,is) = rho_s(is)%f(mg%is(1):mg%ie(1), mg%is(2):mg%ie(2), mg%is(3):mg%ie(3))
    end do
    tot = 0d0
    do iz = 1, nz_s
    do iy = 1, ny_s
    do ix = 1, nx_s
      if (nspin_s == 2) then
        trho = r(ix, iy, iz, 1) + r(ix, iy, iz, 2) + 1d-20
      else
        trho = 2 * (r(ix, iy, iz, 1) * 0.5d0)
      end if
      tot = tot + e(ix, iy, iz) * trho
    end do
    end do
    end do
    E_xc = tot * system%hvol
  end subroutine exchange_correlation_hip

  ! V_local(is) = Vpsl + Vh + Vxc(is) on the device; the result becomes the handle's
  ! potential (no upload at the next taylor_hip / hpsi_hip) and is written to Vlocal
  subroutine update_vlocal_hip(mg, nspin, Vh, Vpsl, Vxc, Vlocal)
    type(s_rgrid), intent(in) :: mg
    integer, intent(in) :: nspin
    type(s_scalar), intent(in) :: Vh, Vpsl, Vxc(nspin)
    type(s_scalar) :: Vlocal(nspin)
    integer :: is
    call pot_buffers()
    if (nspin /= nspin_s) stop 'salmon_hip: nspin does not match the handle'
    call scalar_up(mg, Vh, d_vh)
    call scalar_up(mg, Vpsl, d_vpsl)
    do is = 1, nspin
      call scalar_up(mg, Vxc(is), channel(d_vxc, is))
    end do
    call salmon_hip_check(salmon_hip_update_vlocal(h, d_vpsl, d_vh, d_vxc, c_null_ptr, d_vout), 'update_vlocal')
    do is = 1, nspin
      call scalar_down(mg, channel(d_vout, is), Vlocal(is))
      v_last(:,:,:,is) = Vlocal(is)%f(mg%is(1):mg%ie(1), mg%is(2):mg%ie(2), mg%is(3):mg%ie(3))
    end do
    have_v = .true.
  end subroutine update_vlocal_hip

end module salmon_hip_dropin
