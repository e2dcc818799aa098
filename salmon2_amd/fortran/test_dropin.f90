!
! Driver for the drop-in wrappers: reads a seeded system (stream file written
! by tests/test_gpu_fortran.py), fills SALMON's types, calls
! update_kvector_nonlocalpt_hip, hpsi_hip, taylor_hip, calc_density_hip and
! calc_current_hip exactly as time_evolution_step would, and writes the results.
!   usage: test_dropin <input> <output>
!
program test_dropin
  use structures
  use salmon_hip_dropin
  implicit none
  type(s_rgrid) :: mg
  type(s_dft_system) :: system
  type(s_parallel_info) :: info
  type(s_stencil) :: stencil
  type(s_sendrecv_grid) :: srg
  type(s_pp_grid) :: ppg
  type(s_rt) :: rt
  type(s_orbital) :: psi, hpsi_out, tin, tout, work
  type(s_scalar), allocatable :: V_local(:), rho(:,:)
  real(8), allocatable :: v(:,:,:,:), curr(:,:,:)
  integer :: hdr(11), nx, ny, nz, ly, nspin, no, nk, natom, nps, nlma, n_hamil, is
  real(8) :: hgs(3), hvol, lap0, vec_ac(3)
  character(len=512) :: fin, fout
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
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
  close(u)
  allocate(hpsi_out%zwf, mold=psi%zwf)
  allocate(tin%zwf, source=psi%zwf)
  allocate(tout%zwf, mold=psi%zwf)
  allocate(work%zwf, mold=psi%zwf)
  hpsi_out%zwf = 0; tout%zwf = 0; work%zwf = 0

  ! time_evolution_step order: vec_Ac set, phase table, taylor, density, current
  system%vec_Ac = vec_ac
  call salmon_hip_dropin_setup(mg, system, info, stencil, ppg)
  call update_kvector_nonlocalpt_hip(info%ik_s, info%ik_e, system, ppg)
  call hpsi_hip(psi, hpsi_out, info, mg, V_local, system, stencil, srg, ppg)
  call taylor_hip(mg, system, info, stencil, srg, tin, tout, work, ppg, V_local, rt)
  allocate(rho(nspin, 1))
  do is = 1, nspin
    allocate(rho(is, 1)%f(nx, ny, nz))
  end do
  call calc_density_hip(system, rho, tout, info, mg)
  allocate(curr(3, nspin, 1))
  call calc_current_hip(system, mg, stencil, info, srg, tout, ppg, curr)
  call salmon_hip_dropin_release()

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) hpsi_out%zwf
  write(u) tout%zwf
  do is = 1, nspin
    write(u) rho(is, 1)%f
  end do
  write(u) curr
  close(u)
end program test_dropin
