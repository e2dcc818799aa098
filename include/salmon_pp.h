/*
 * Projector-table builder for KY-format norm-conserving pseudopotentials ("*_rps.dat")
 * on an orthogonal periodic cell: the host-side set-up whose outputs feed
 * salmon_hip_set_nonlocal (uv, sphere points) and salmon_hip_set_vlocal /
 * salmon_hip_update_potential (the local pseudopotential Vpsl).
 *
 * Replaces, for this format and cell shape:
 *   read_ps_ky                 src/atom/pp/input_pp.f90:288-341
 *   input_pp (after reading)   src/atom/pp/input_pp.f90:92-205   (radial grid, rps, anorm, inorm)
 *   making_ps_without_masking  src/atom/pp/input_pp.f90:1111-1248 (vloctbl, udvtbl)
 *   init_ps: calc_nps, calc_jxyz, set_lma, calc_uv  src/atom/pp/prep_pp.f90:195-525
 *   calc_vpsl_periodic         src/atom/pp/prep_pp.f90:798-940 (separable DFT branch)
 * Not covered: other file formats, masking (yn_psmask = 'y'), spin-orbit, DFT+U,
 * non-orthogonal cells and the force tables (dudvtbl).
 *
 * Arrays use SALMON's Fortran layouts (first index fastest); element and atom indices
 * passed in and returned are 1-based like SALMON's kion / ia_tbl / jxyz.  Every call
 * returns 0 or a SALMON_HIP_E* code with the message in salmon_hip_last_error().
 */
#ifndef SALMON_PP_H
#define SALMON_PP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct salmon_pp *salmon_pp_t;

/* Read nelem KY files (files[i]) with SALMON's Lmax_ps / Lloc_ps per element (lmax_ps < 0:
 * the file's value) and build the radial tables. */
int salmon_pp_create(int nelem, const char *const *files, const int *lmax_ps, const int *lloc_ps,
                     salmon_pp_t *out);
int salmon_pp_destroy(salmon_pp_t pp);

/* Element summary (ielem 1-based): valence charge, maximum l, local l, cut-off radius (bohr),
 * radial points inside it, projectors per atom (channels with inorm != 0, 2l+1 each). */
int salmon_pp_element(salmon_pp_t pp, int ielem, int *zps, int *mlps, int *lref, double *rps, int *nrps,
                      int *nlma_atom);

/* Sizes for salmon_pp_nonlocal: *nps = largest sphere (calc_nps), *nlma = projectors (set_lma).
 * kion(natom) element of each atom (1-based), rion(3, natom) positions (bohr), al(3) cell
 * edges (bohr), n(3) grid points per edge. */
int salmon_pp_sizes(salmon_pp_t pp, int natom, const int *kion, const double *rion, const double *al,
                    const int *n, int *nps, int *nlma);

/* calc_jxyz / set_lma / calc_uv: mps(natom), jxyz(3, nps, natom) (1-based grid indices),
 * rxyz(3, nps, natom), ia_tbl(nlma), rinv_uvu(nlma) = inorm * hvol, uv(nps, nlma); the
 * arguments of salmon_hip_set_nonlocal. */
int salmon_pp_nonlocal(salmon_pp_t pp, int natom, const int *kion, const double *rion, const double *al,
                       const int *n, int nps, int nlma, int *mps, int *jxyz, double *rxyz, int *ia_tbl,
                       double *rinv_uvu, double *uv);

/* calc_vpsl_periodic: the local pseudopotential vpsl(nx, ny, nz) of the atoms. */
int salmon_pp_vpsl(salmon_pp_t pp, int natom, const int *kion, const double *rion, const double *al,
                   const int *n, double *vpsl);

#ifdef __cplusplus
}
#endif

#endif
