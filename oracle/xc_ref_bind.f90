! TEST INFRASTRUCTURE ONLY: C entry points to SALMON's PZ81 routines, which are
! compiled unmodified from the reference sources (src/xc/builtin_pz.f90,
! src/xc/builtin_pz_sp.f90) into oracle/_ref (see Makefile, target ref).  The
! tests compare the oracle's restatement with them; nothing else calls them.
subroutine ref_exc_cor_pz(nl, rho_s, exc, eexc, vexc) bind(c, name='ref_exc_cor_pz')
  use iso_c_binding, only: c_int, c_double
  use builtin_pz, only: exc_cor_pz
  implicit none
  integer(c_int), value :: nl
  real(c_double), intent(in) :: rho_s(nl)
  real(c_double), intent(out) :: exc(nl), eexc(nl), vexc(nl)
  call exc_cor_pz(nl, rho_s, exc, eexc, vexc)
end subroutine ref_exc_cor_pz

subroutine ref_exc_cor_pz_sp(nl, rho_s, exc, eexc, vexc_s) bind(c, name='ref_exc_cor_pz_sp')
  use iso_c_binding, only: c_int, c_double
  use builtin_pz_sp, only: exc_cor_pz_sp
  implicit none
  integer(c_int), value :: nl
  real(c_double), intent(in) :: rho_s(nl, 2)
  real(c_double), intent(out) :: exc(nl), eexc(nl), vexc_s(nl, 2)
  call exc_cor_pz_sp(nl, rho_s, exc, eexc, vexc_s)
end subroutine ref_exc_cor_pz_sp
