!
! Driver for the drop-in wrappers (test infrastructure).  Reads a seeded system
! (stream file written by tests/fortran/dropin_io.py), fills SALMON's types and
! calls the wrappers as SALMON's RT code would.
!
!   test_dropin check <input> <output>
!       update_kvector_nonlocalpt_hip, hpsi_hip, taylor_hip, calc_density_hip and
!       calc_current_hip on host arrays (copied in and out at every call), then the
!       same Taylor step again with the orbitals resident (salmon_hip_enter_data);
!       writes H psi, psi(t+dt) of both, rho and curr.
!   test_dropin rt <input> <output>
!       the fixed-potential RT loop of main_tddft (src/rt/main_tddft.f90:89-130) /
!       time_evolution_step ('tr', middlepoint; src/rt/time_evolution_step.f90:116-127,
!       146-170, 243-253) with spsi_in, spsi_out and tpsi resident (`!$acc enter data`,
!       src/rt/initialization_rt.f90:234-240): per step vec_Ac = (Ac(itt) + Ac(itt-1))/2,
!       update_kvector_nonlocalpt_hip, taylor_hip (buffers swapped by step parity),
!       calc_density_hip, vec_Ac = Ac(itt), calc_current_hip.  The steps after the
!       warm-up are timed between two device synchronisations; writes the elapsed
!       seconds, J(t) of every step and the final orbitals.
!   test_dropin scf <input> <output>
!       the self-consistent potential update of time_evolution_step (yn_fix_func = 'n',
!       src/rt/time_evolution_step.f90:160-207) through the wrappers: calc_density_hip,
!       rho = rho_s(1) [+ rho_s(2)], hartree_hip, exchange_correlation_hip (PZ),
!       update_vlocal_hip with Vpsl = the input V_local(1), then taylor_hip with the new
!       V_local; writes Vh, Vxc, E_xc, V_local and psi(t+dt).
!
program test_dropin
  use iso_c_binding
  use structures
  use salmon_hip_dropin
  implicit none
  type(s_rgrid) :: mg
  type(s_dft_system) :: system
  type(s_parallel_info) :: info
  type(s_stencil) :: stencil
  type(s_sendrecv_grid) :: srg, srg_scalar
  type(s_rgrid) :: lg
  type(s_reciprocal_grid) :: fg
  type(s_poisson) :: poisson
  type(s_xc_functional) :: xc_func
  type(s_pp_nlcc) :: ppn
  type(s_scalar) :: Vh, Vpsl, rhot
  type(s_scalar), allocatable :: Vxc(:)
  real(8) :: al(3), E_xc
  type(s_pp_grid) :: ppg
  type(s_rt) :: rt
  type(s_orbital), target :: psi, hpsi_out, tin, tout, work
  type(s_scalar), allocatable :: V_local(:), rho(:,:)
  real(8), allocatable :: v(:,:,:,:), curr(:,:,:), ac(:,:), jt(:,:)
  integer :: hdr(11), nx, ny, nz, ly, nspin, no, nk, natom, nps, nlma, n_hamil, is, nt, nwarm, itt
  integer(8) :: c0, c1, crate
  real(8) :: hgs(3), hvol, lap0, vec_ac(3), elapsed
  character(len=512) :: mode, fin, fout
  integer :: u

  call get_command_argument(1, mode)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  nx = hdr(1); ny = hdr(2); nz = hdr(3); ly = hdr(4); nspin = hdr(5); no = hdr(6); nk = hdr(7)
  natom = hdr(8); nps = hdr(9); nlma = hdr(10); n_hamil = hdr(11)
  read(u) hgs, hvol, lap0
  stencil%if_orthogonal = .true.
  stencil%coef_lap0 = lap0
  read(u) stencil%coef_lap, stencil%coef_nab, vec_ac

  mg%is = 1
  mg%ie = [nx, ny, nz]
  mg%num = mg%ie - mg%is + 1
  mg%is_array = mg%is
  mg%ie_array = [nx, ly, nz]
  info%if_divide_rspace = .false.
  info%im_s = 1; info%im_e = 1; info%numm = 1
  info%ik_s = 1; info%ik_e = nk; info%numk = nk
  info%io_s = 1; info%io_e = no; info%numo = no

  system%ngrid = nx * ny * nz
  system%nspin = nspin; system%no = no; system%nk = nk
  system%hvol = hvol; system%hgs = hgs
  allocate(system%vec_k(3, nk), system%wtk(nk), system%rocc(no, nk, nspin))
  read(u) system%vec_k, system%wtk, system%rocc
  allocate(v(nx, ny, nz, nspin), V_local(nspin))
  read(u) v
  do is = 1, nspin
    allocate(V_local(is)%f(nx, ny, nz))
    V_local(is)%f = v(:,:,:,is)
  end do
  ppg%nps = nps
  ppg%nlma = nlma
  allocate(ppg%mps(natom), ppg%jxyz(3, nps, natom), ppg%rxyz(3, nps, natom), ppg%ia_tbl(nlma), &
           ppg%rinv_uvu(nlma), ppg%uv(nps, nlma))
  read(u) ppg%mps, ppg%jxyz, ppg%ia_tbl, ppg%rxyz, ppg%rinv_uvu, ppg%uv
  allocate(rt%zc(n_hamil))
  read(u) rt%zc
  allocate(psi%zwf(nx, ly, nz, nspin, no, nk, 1))
  read(u) psi%zwf
  read(u) nt, nwarm
  allocate(ac(3, 0:nt + 1))
  read(u) ac
  read(u) al
  close(u)
  system%primitive_a = 0d0
  do is = 1, 3
    system%primitive_a(is, is) = al(is)
  end do
  allocate(tin%zwf, source=psi%zwf)
  allocate(tout%zwf, mold=psi%zwf)
  allocate(work%zwf, mold=psi%zwf)
  tout%zwf = 0; work%zwf = 0
  allocate(rho(nspin, 1))
  do is = 1, nspin
    allocate(rho(is, 1)%f(nx, ny, nz))
  end do
  allocate(curr(3, nspin, 1))
  system%vec_Ac = vec_ac
  call salmon_hip_dropin_setup(mg, system, info, stencil, ppg)

  if (trim(mode) == 'check') then
    allocate(hpsi_out%zwf, mold=psi%zwf)
    hpsi_out%zwf = 0
    ! time_evolution_step order: vec_Ac set, phase table, taylor, density, current (host arrays)
    call update_kvector_nonlocalpt_hip(info%ik_s, info%ik_e, system, ppg)
    call hpsi_hip(psi, hpsi_out, info, mg, V_local, system, stencil, srg, ppg)
    call taylor_hip(mg, system, info, stencil, srg, tin, tout, work, ppg, V_local, rt)
    call calc_density_hip(system, rho, tout, info, mg)
    call calc_current_hip(system, mg, stencil, info, srg, tout, ppg, curr)
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) hpsi_out%zwf
    write(u) tout%zwf
    do is = 1, nspin
      write(u) rho(is, 1)%f
    end do
    write(u) curr
    ! the same Taylor step with resident orbitals: the host output is stale until update_host
    tout%zwf = 0
    call salmon_hip_enter_data(tin)
    call salmon_hip_enter_data(tout)
    call salmon_hip_enter_data(work)
    call taylor_hip(mg, system, info, stencil, srg, tin, tout, work, ppg, V_local, rt)
    if (any(tout%zwf /= 0)) stop 'resident taylor_hip wrote the host array'
    call salmon_hip_update_host(tout)
    write(u) tout%zwf
    close(u)
    call salmon_hip_exit_data(tin)
    call salmon_hip_exit_data(tout)
    call salmon_hip_exit_data(work)
  else if (trim(mode) == 'rt') then
    allocate(jt(3, nt))
    call salmon_hip_enter_data(tin)
    call salmon_hip_enter_data(tout)
    call salmon_hip_enter_data(work)
    call salmon_hip_dropin_sync()
    call system_clock(c0, crate)
    do itt = 1, nt
      if (itt == nwarm + 1) then
        call salmon_hip_dropin_sync()
        call system_clock(c0, crate)
      end if
      system%vec_Ac = 0.5d0 * (ac(:, itt) + ac(:, itt - 1))
      call update_kvector_nonlocalpt_hip(info%ik_s, info%ik_e, system, ppg)
      if (mod(itt, 2) == 1) then
        call taylor_hip(mg, system, info, stencil, srg, tin, tout, work, ppg, V_local, rt)
        call calc_density_hip(system, rho, tout, info, mg)
        system%vec_Ac = ac(:, itt)
        call calc_current_hip(system, mg, stencil, info, srg, tout, ppg, curr)
      else
        call taylor_hip(mg, system, info, stencil, srg, tout, tin, work, ppg, V_local, rt)
        call calc_density_hip(system, rho, tin, info, mg)
        system%vec_Ac = ac(:, itt)
        call calc_current_hip(system, mg, stencil, info, srg, tin, ppg, curr)
      end if
      jt(:, itt) = sum(curr(:, :, 1), dim=2)
    end do
    call salmon_hip_dropin_sync()
    call system_clock(c1)
    elapsed = real(c1 - c0, 8) / real(crate, 8)
    if (mod(nt, 2) == 1) then
      call salmon_hip_update_host(tout)
    else
      call salmon_hip_update_host(tin)
    end if
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) elapsed
    write(u) jt
    if (mod(nt, 2) == 1) then
      write(u) tout%zwf
    else
      write(u) tin%zwf
    end if
    close(u)
  else if (trim(mode) == 'scf') then
    lg = mg
    xc_func%xctype = [1, 0, 0]  ! salmon_xctype_pz, none, none
    xc_func%ispin = nspin - 1
    allocate(Vh%f(nx, ny, nz), Vpsl%f(nx, ny, nz), rhot%f(nx, ny, nz), Vxc(nspin))
    do is = 1, nspin
      allocate(Vxc(is)%f(nx, ny, nz))
    end do
    Vpsl%f = V_local(1)%f
    call update_kvector_nonlocalpt_hip(info%ik_s, info%ik_e, system, ppg)
    call calc_density_hip(system, rho, psi, info, mg)
    rhot%f = rho(1, 1)%f
    if (nspin == 2) rhot%f = rho(1, 1)%f + rho(2, 1)%f
    call hartree_hip(lg, mg, info, system, fg, poisson, srg_scalar, stencil, rhot, Vh)
    call exchange_correlation_hip(system, xc_func, mg, srg_scalar, srg, rho(:, 1), ppn, info, psi, stencil, &
                                  Vxc, E_xc)
    call update_vlocal_hip(mg, nspin, Vh, Vpsl, Vxc, V_local)
    call taylor_hip(mg, system, info, stencil, srg, tin, tout, work, ppg, V_local, rt)
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) Vh%f
    do is = 1, nspin
      write(u) Vxc(is)%f
    end do
    write(u) E_xc
    do is = 1, nspin
      write(u) V_local(is)%f
    end do
    write(u) tout%zwf
    close(u)
  else
    stop 'usage: test_dropin check|rt|scf <input> <output>'
  end if
  call salmon_hip_dropin_release()
end program test_dropin
