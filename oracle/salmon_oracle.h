/*
 * TEST INFRASTRUCTURE ONLY — the CPU oracle for the real-time Kohn-Sham
 * propagation hot path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker / reported CPU
 * baseline.  The product (salmon2_amd) never calls into it.
 *
 * Plain C restatement of SALMON's CPU loops (file:line citations are relative
 * to the SALMON v2.2.0 source tree).  Every array is in SALMON's Fortran
 * column-major layout:
 *   zwf(ix, iy, iz, ispin, io, ik)  complex(8) as interleaved (re, im) doubles,
 *   x unit-stride, y leading extent `ly` (ny + 1 with USE_OPT_ARRAY_PADDING,
 *   src/common/initialization.f90:606-609), z extent nz.
 *   V_local(ix, iy, iz, ispin)      real(8), unpadded (src/common/structures.f90:539-553)
 *   jxyz/rxyz(3, nps, natom), uv(nps, nlma), zekr_uV(nps, nlma, nk)
 * Index tables (jxyz, ia_tbl) are 1-based exactly as SALMON stores them.
 *
 * Parity status of each function (what pins it) is listed in the header of
 * salmon_oracle.c.
 */
#ifndef SALMON_ORACLE_H
#define SALMON_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_system {
  /* grid (src/common/structures.f90:86-94) */
  int nx, ny, nz, ly;
  int nspin, no, nk;
  /* stencil (src/common/structures.f90:198-203) */
  double coef_lap[12];   /* coef_lap(4,3) */
  double coef_nab[12];   /* coef_nab(4,3) */
  double coef_lap0;
  double hgs[3];
  double hvol;
  int stencil_mode;      /* -1: reference rule (N >= 32^3 -> omp order), 0: seq, 1: omp */
  /* k-points and occupations */
  const double *vec_k;   /* (3, nk) */
  const double *rocc;    /* (no, nk, nspin) */
  const double *wtk;     /* (nk) */
  const double *vlocal;  /* (nx, ny, nz, nspin) */
  /* projectors (src/common/structures.f90:254-319) */
  int natom, nps, nlma;
  const int *mps;        /* (natom) */
  const int *jxyz;       /* (3, nps, natom), 1-based */
  const double *rxyz;    /* (3, nps, natom) */
  const int *ia_tbl;     /* (nlma), 1-based */
  const double *rinv_uvu;/* (nlma) */
  const double *uv;      /* (nps, nlma) */
  double *zekr;          /* (2, nps, nlma, nk) written by orc_update_kvector */
} orc_system;

void orc_stencil_coefs(const double hgs[3], double coef_lap[12], double coef_nab[12],
                       double *coef_lap0);

void orc_zstencil_seq(int nx, int ny, int nz, int ly, const double *tpsi, double *htpsi,
                      const double *V, double lap0, const double lapt[12], const double nabt[12]);
void orc_zstencil_omp(int nx, int ny, int nz, int ly, const double *tpsi, double *htpsi,
                      const double *V, double lap0, const double lapt[12], const double nabt[12]);

void orc_update_kvector(orc_system *s, const double vec_Ac[3]);
void orc_zpseudo(const orc_system *s, const double *tpsi, double *htpsi);
void orc_hpsi(const orc_system *s, const double vec_Ac[3], const double *tpsi, double *htpsi);
void orc_taylor(const orc_system *s, const double vec_Ac[3], const double *zc, int n_hamil,
                double *tspsi_in, double *tspsi_out, double *sshtpsi);
void orc_taylor_coefs(double dt, int n_hamil, int aetrs, double *zc);

void orc_density(const orc_system *s, const double *psi, double *rho);
void orc_current(const orc_system *s, const double vec_Ac[3], const double *psi, double *curr);
void orc_dipole(const orc_system *s, const double *rho, double out[4]);

void orc_ac_asin2cos(double A0, double omega, double tw, double phi_cep, const double epdir[3],
                     double dt, int nt, double *ac /* (3, 0:nt+1) */);

int orc_rt_run(orc_system *s, double dt, int n_hamil, int nt, const double *ac_tot,
               double *psi, double *curr_t, double *dip_t, double *rne_t);

void orc_hartree(int nx, int ny, int nz, const double al[3], const double *rho, double *vh);
void orc_xc_pz(long n, const double *rho, double *vxc, double *exc);
void orc_xc_pz_sp(long n, const double *rho_s, double *vxc_s, double *exc);

/* RT loop with the Hartree + PZ-LDA potential updated every step (yn_fix_func = 'n'),
   optionally with the predictor-corrector; vpsl is the fixed local pseudopotential part,
   s->vlocal must point to a writable (nx,ny,nz) buffer that receives V_local. */
int orc_rt_run_full(orc_system *s, int scf, int pc, int lo, const double *vpsl, double *vloc, const double al[3],
                    double dt, int n_hamil, int nt, double *ac, double *psi, double *curr_t, double *dip_t,
                    double *rne_t);
/* The same with the propagator: aetrs = 0 'middlepoint', 1 'aetrs'. */
int orc_rt_run_prop(orc_system *s, int scf, int pc, int lo, int aetrs, const double *vpsl, double *vloc,
                    const double al[3], double dt, int n_hamil, int nt, double *ac, double *psi, double *curr_t,
                    double *dip_t, double *rne_t);
int orc_rt_run_scf(orc_system *s, const double *vpsl, double *vloc, const double al[3], int pc, double dt,
                   int n_hamil, int nt, const double *ac_tot, double *psi, double *curr_t, double *dip_t,
                   double *rne_t);

#ifdef __cplusplus
}
#endif
#endif
