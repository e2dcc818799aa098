/*
 * TEST INFRASTRUCTURE ONLY — CPU oracle of SALMON's real-time propagation hot
 * path.  Used by tests/ (as the parity checker), __graft_entry__.smoke() and
 * bench.py (cpu_baseline leg).  Never linked into or called by salmon2_amd.
 *
 * Each routine restates a SALMON v2.2.0 CPU loop in plain C, keeping the
 * reference's floating-point operation order (left-to-right Fortran
 * evaluation, no FMA contraction: build with -ffp-contract=off).
 * Parallelism (OpenMP) is only over independent orbitals or grid points, so
 * every result is bitwise independent of the thread count; where the
 * reference reduces in an OpenMP-schedule-dependent order the oracle uses the
 * serial order of the same loop nest.
 *
 * What pins each function (tests/test_oracle_*.py):
 *   orc_stencil_coefs   initialization.f90:240-255,1324-1371; closed-form
 *                       consistency (cN(0)+2 sum cN(n) = 0, bN moments).
 *   orc_zstencil_seq    bitwise vs zstencil_typical_seq compiled from the
 *                       reference source (oracle/_ref), plane-wave eigenvalue.
 *   orc_zstencil_omp    bitwise vs zstencil_typical_omp (oracle/_ref).
 *   orc_update_kvector  closed form of conj(exp(i(k+A).r)) uv.
 *   orc_zpseudo         dense brute-force V_NL = sum |p> rinv <p| (numpy).
 *   orc_hpsi            sum of the two above; Hermiticity on random input.
 *   orc_taylor          dense-H truncated polynomial sum_n (-i H dt)^n/n! psi.
 *   orc_density         sum rule sum(rho) hvol = sum rocc wtk; dense |psi|^2.
 *   orc_current         plane-wave closed form; real psi at Gamma gives 0;
 *                       dense-matrix form of the nonlocal current.
 *   orc_dipole          closed form on analytic densities.
 *   orc_ac_asin2cos     closed form of em_field.f90:177-187.
 *   orc_rt_run          composition of the above (time_evolution_step.f90).
 *   orc_hartree         closed form 4 pi/|G|^2 cos(G.r) on cosine densities; the
 *                       periodic Poisson equation through the FD-free spectral
 *                       Laplacian (G^2 V(G) = 4 pi rho(G)); G = 0 removed.
 *   orc_xc_pz           v = d(rho e)/d rho by finite differences of rho*exc; the
 *                       published PZ81 values at rs = 1, 2, 5; exc_cor_pz compiled
 *                       from the reference source (oracle/_ref).
 *   orc_xc_pz_sp        exc_cor_pz_sp compiled from the reference source
 *                       (oracle/_ref); v_s = d(rho e)/d rho_s by finite differences;
 *                       zeta = 0 gives the unpolarised PZ81, zeta = 1 the published
 *                       polarised parameters; exchanging the channels exchanges v.
 *   orc_rt_run_scf      composition (time_evolution_step.f90:160-207, 354-423); with
 *                       nspin = 2 and both channels holding the nspin = 1 orbitals it
 *                       follows the nspin = 1 run (to the PZ offsets' 1e-10 scale).
 *   orc_rt_run_full     the above plus the 'lo' field feedback: the discrete plasma
 *                       oscillation cos(w dt) = 1 - 2 pi n dt^2 of one constant orbital.
 */
#include "salmon_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define C2(i) (2 * (size_t)(i))

/* cN(0:4,4), bN(1:4,4): src/common/initialization.f90:1337-1340, 1367-1371 */
static const double CN4[5] = {-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0};
static const double BN4[5] = {0.0, 4.0 / 5.0, -1.0 / 5.0, 4.0 / 105.0, -1.0 / 280.0};

/* src/common/initialization.f90:240-255 (orthogonal lattice) */
void orc_stencil_coefs(const double hgs[3], double coef_lap[12], double coef_nab[12],
                       double *coef_lap0) {
  *coef_lap0 = -0.5 * CN4[0] *
               (1.0 / (hgs[0] * hgs[0]) + 1.0 / (hgs[1] * hgs[1]) + 1.0 / (hgs[2] * hgs[2]));
  for (int jj = 0; jj < 3; ++jj)
    for (int ii = 0; ii < 4; ++ii) {
      coef_lap[ii + 4 * jj] = CN4[ii + 1] / (hgs[jj] * hgs[jj]);
      coef_nab[ii + 4 * jj] = BN4[ii + 1] / hgs[jj];
    }
}

/* periodic wrap: idx(j) = mod(j+n-1, n)+1 (initialization.f90:574-587), 0-based here */
static inline int wrap(int j, int n) { return ((j % n) + n) % n; }

typedef struct { double re, im; } cpx;

static inline cpx ld(const double *p, size_t i) { cpx c = {p[C2(i)], p[C2(i) + 1]}; return c; }

/* v = lapt(a)*(t[p]+t[m]) accumulation helpers, Fortran left-to-right order */
static inline cpx radd(cpx a, cpx b) { cpx c = {a.re + b.re, a.im + b.im}; return c; }
static inline cpx rsub(cpx a, cpx b) { cpx c = {a.re - b.re, a.im - b.im}; return c; }
static inline cpx rscale(double s, cpx a) { cpx c = {s * a.re, s * a.im}; return c; }

/*
 * One axis of zstencil_typical_{seq,omp} (zstencil_core/typical/seq.f90:64-118):
 *   v = l1*(p1+m1) + l2*(p2+m2) + l3*(p3+m3) + l4*(p4+m4)
 *   w = n1*(p1-m1) + ...
 * p_n = psi(r + n e_d), m_n = psi(r - n e_d).
 */
static inline void axis_vw(const double *l, const double *nb, const cpx p[4], const cpx m[4],
                           cpx *v, cpx *w) {
  cpx vv = rscale(l[0], radd(p[0], m[0]));
  vv = radd(vv, rscale(l[1], radd(p[1], m[1])));
  vv = radd(vv, rscale(l[2], radd(p[2], m[2])));
  vv = radd(vv, rscale(l[3], radd(p[3], m[3])));
  cpx ww = rscale(nb[0], rsub(p[0], m[0]));
  ww = radd(ww, rscale(nb[1], rsub(p[1], m[1])));
  ww = radd(ww, rscale(nb[2], rsub(p[2], m[2])));
  ww = radd(ww, rscale(nb[3], rsub(p[3], m[3])));
  *v = vv;
  *w = ww;
}

/* idx/idy/idz(is-4:ie+4) of initialization.f90:574-587, 0-based: tab[j + 4] = mod(j, n) for
   j = -4 .. n+3 (the reference reads its neighbours through these tables) */
static int *wrap_table(int n) {
  int *t = (int *)malloc(sizeof(int) * (size_t)(n + 8));
  if (t)
    for (int j = -4; j < n + 4; ++j) t[j + 4] = wrap(j, n);
  return t;
}

/* The neighbours of point (ix, iy, iz) along one axis through the tables: p[n-1] = psi(r + n e),
   m[n-1] = psi(r - n e); rowy[k] = offset of row (iy + k - 4, iz), rowz[k] = of (iy, iz + k - 4). */
static inline void gather_x(const double *t, const int *tx, int ix, size_t row, cpx p[4], cpx m[4]) {
  for (int n = 1; n <= 4; ++n) {
    p[n - 1] = ld(t, row + (size_t)tx[ix + n + 4]);
    m[n - 1] = ld(t, row + (size_t)tx[ix - n + 4]);
  }
}
static inline void gather_r(const double *t, int ix, const size_t rows[9], cpx p[4], cpx m[4]) {
  for (int n = 1; n <= 4; ++n) {
    p[n - 1] = ld(t, rows[4 + n] + (size_t)ix);
    m[n - 1] = ld(t, rows[4 - n] + (size_t)ix);
  }
}

static void row_offsets(int iy, int iz, int nx, int ly, const int *ty, const int *tz, size_t rowy[9],
                        size_t rowz[9]) {
  for (int k = 0; k < 9; ++k) {
    rowy[k] = (size_t)nx * ((size_t)ty[iy + k] + (size_t)ly * iz);
    rowz[k] = (size_t)nx * ((size_t)iy + (size_t)ly * tz[iz + k]);
  }
}

/* zstencil_core/typical/seq.f90:17-127 (one orbital, periodic, no halo) */
void orc_zstencil_seq(int nx, int ny, int nz, int ly, const double *tpsi, double *htpsi,
                      const double *V, double lap0, const double lapt[12], const double nabt[12]) {
  int *tx = wrap_table(nx), *ty = wrap_table(ny), *tz = wrap_table(nz);
  for (int iz = 0; iz < nz; ++iz)
    for (int iy = 0; iy < ny; ++iy) {
      size_t rowy[9], rowz[9];
      row_offsets(iy, iz, nx, ly, ty, tz, rowy, rowz);
      const size_t row = rowy[4];
      for (int ix = 0; ix < nx; ++ix) {
        cpx p[4], m[4], v, w, v2, w2;
        gather_x(tpsi, tx, ix, row, p, m);
        axis_vw(lapt + 0, nabt + 0, p, m, &v, &w);
        gather_r(tpsi, ix, rowy, p, m);
        axis_vw(lapt + 4, nabt + 4, p, m, &v2, &w2);
        v = radd(v2, v);
        w = radd(w2, w);
        gather_r(tpsi, ix, rowz, p, m);
        axis_vw(lapt + 8, nabt + 8, p, m, &v2, &w2);
        v = radd(v2, v);
        w = radd(w2, w);
        size_t i = row + (size_t)ix;
        cpx t = ld(tpsi, i);
        double vl = V[(size_t)ix + (size_t)nx * ((size_t)iy + (size_t)ny * iz)];
        /* htpsi = V*t + lap0*t - 0.5d0*v - zI*w ;  zI*w = (-w.im, w.re) */
        double hre = ((vl * t.re + lap0 * t.re) - 0.5 * v.re) - (-w.im);
        double him = ((vl * t.im + lap0 * t.im) - 0.5 * v.im) - w.re;
        htpsi[C2(i)] = hre;
        htpsi[C2(i) + 1] = him;
      }
    }
  free(tx);
  free(ty);
  free(tz);
}

/* zstencil_core/typical/omp.f90:22-164: x+y first, z added in a second pass */
void orc_zstencil_omp(int nx, int ny, int nz, int ly, const double *tpsi, double *htpsi,
                      const double *V, double lap0, const double lapt[12], const double nabt[12]) {
  int *tx = wrap_table(nx), *ty = wrap_table(ny), *tz = wrap_table(nz);
  for (int iz = 0; iz < nz; ++iz)
    for (int iy = 0; iy < ny; ++iy) {
      size_t rowy[9], rowz[9];
      row_offsets(iy, iz, nx, ly, ty, tz, rowy, rowz);
      const size_t row = rowy[4];
      for (int ix = 0; ix < nx; ++ix) {
        cpx p[4], m[4], v, w, v2, w2;
        gather_x(tpsi, tx, ix, row, p, m);
        axis_vw(lapt + 0, nabt + 0, p, m, &v, &w);
        gather_r(tpsi, ix, rowy, p, m);
        axis_vw(lapt + 4, nabt + 4, p, m, &v2, &w2);
        v = radd(v2, v);
        w = radd(w2, w);
        size_t i = row + (size_t)ix;
        cpx t = ld(tpsi, i);
        double vl = V[(size_t)ix + (size_t)nx * ((size_t)iy + (size_t)ny * iz)];
        htpsi[C2(i)] = ((vl * t.re + lap0 * t.re) - 0.5 * v.re) - (-w.im);
        htpsi[C2(i) + 1] = ((vl * t.im + lap0 * t.im) - 0.5 * v.im) - w.re;
      }
      for (int ix = 0; ix < nx; ++ix) {
        cpx p[4], m[4], v, w;
        gather_r(tpsi, ix, rowz, p, m);
        axis_vw(lapt + 8, nabt + 8, p, m, &v, &w);
        size_t i = row + (size_t)ix;
        /* htpsi = htpsi - 0.5d0*v - zI*w */
        htpsi[C2(i)] = (htpsi[C2(i)] - 0.5 * v.re) - (-w.im);
        htpsi[C2(i) + 1] = (htpsi[C2(i) + 1] - 0.5 * v.im) - w.re;
      }
    }
  free(tx);
  free(ty);
  free(tz);
}

static size_t orb_size(const orc_system *s) { return (size_t)s->nx * s->ly * s->nz; }
static size_t grid_size(const orc_system *s) { return (size_t)s->nx * s->ny * s->nz; }

/* src/common/hamiltonian.f90:847-903 (update_kvector_nonlocalpt, collinear branch) */
void orc_update_kvector(orc_system *s, const double vec_Ac[3]) {
  const int nk = s->nk, nlma = s->nlma, nps = s->nps;
#pragma omp parallel for collapse(2) schedule(static)
  for (int ik = 0; ik < nk; ++ik)
    for (int ilma = 0; ilma < nlma; ++ilma) {
      double kAc[3];
      for (int d = 0; d < 3; ++d) kAc[d] = s->vec_k[3 * ik + d] + vec_Ac[d];
      int ia = s->ia_tbl[ilma] - 1;
      for (int j = 0; j < s->mps[ia]; ++j) {
        const double *r = s->rxyz + 3 * ((size_t)j + (size_t)nps * ia);
        double th = kAc[0] * r[0] + kAc[1] * r[1] + kAc[2] * r[2];
        double c = cos(th), sn = sin(th);
        double u = s->uv[(size_t)j + (size_t)nps * ilma];
        /* conjg(ekr) * uv */
        size_t o = (size_t)j + (size_t)nps * ((size_t)ilma + (size_t)nlma * ik);
        s->zekr[C2(o)] = c * u;
        s->zekr[C2(o) + 1] = -sn * u;
      }
    }
}

/* projection of one orbital: uVpsi = rinv_uvu * sum_j conjg(zekr) psi(jxyz)
   (src/common/nonlocal_potential.f90:360-369) */
static cpx project(const orc_system *s, int ilma, int ik, const double *psi) {
  int ia = s->ia_tbl[ilma] - 1;
  const double *z = s->zekr + C2((size_t)s->nps * ((size_t)ilma + (size_t)s->nlma * ik));
  cpx acc = {0.0, 0.0};
  for (int j = 0; j < s->mps[ia]; ++j) {
    const int *jx = s->jxyz + 3 * ((size_t)j + (size_t)s->nps * ia);
    size_t i = (size_t)(jx[0] - 1) + (size_t)s->nx * ((size_t)(jx[1] - 1) + (size_t)s->ly * (jx[2] - 1));
    double zr = z[C2(j)], zi = -z[C2(j) + 1]; /* conjg(zekr) */
    double pr = psi[C2(i)], pi = psi[C2(i) + 1];
    acc.re = acc.re + (zr * pr - zi * pi);
    acc.im = acc.im + (zr * pi + zi * pr);
  }
  double r = s->rinv_uvu[ilma];
  acc.re = acc.re * r;
  acc.im = acc.im * r;
  return acc;
}

/* src/common/nonlocal_potential.f90:351-382 (CPU branch, one orbital per call) */
static void zpseudo_orbital(const orc_system *s, int ik, const double *tpsi, double *htpsi) {
  for (int ilma = 0; ilma < s->nlma; ++ilma) {
    int ia = s->ia_tbl[ilma] - 1;
    cpx u = project(s, ilma, ik, tpsi);
    const double *z = s->zekr + C2((size_t)s->nps * ((size_t)ilma + (size_t)s->nlma * ik));
    for (int j = 0; j < s->mps[ia]; ++j) {
      const int *jx = s->jxyz + 3 * ((size_t)j + (size_t)s->nps * ia);
      size_t i = (size_t)(jx[0] - 1) + (size_t)s->nx * ((size_t)(jx[1] - 1) + (size_t)s->ly * (jx[2] - 1));
      double zr = z[C2(j)], zi = z[C2(j) + 1];
      /* wrk = uVpsi * zekr_uV ; htpsi = htpsi + wrk */
      double wr = u.re * zr - u.im * zi;
      double wi = u.re * zi + u.im * zr;
      htpsi[C2(i)] = htpsi[C2(i)] + wr;
      htpsi[C2(i) + 1] = htpsi[C2(i) + 1] + wi;
    }
  }
}

void orc_zpseudo(const orc_system *s, const double *tpsi, double *htpsi) {
  const size_t osz = orb_size(s);
  const int nb = s->nk * s->no * s->nspin;
#pragma omp parallel for schedule(dynamic)
  for (int b = 0; b < nb; ++b) {
    int ik = b / (s->no * s->nspin);
    zpseudo_orbital(s, ik, tpsi + C2(osz * b), htpsi + C2(osz * b));
  }
}

static int use_omp_order(const orc_system *s) {
  if (s->stencil_mode >= 0) return s->stencil_mode;
  /* code_optimization.f90:63 */
  return (long)s->nx * s->ny * s->nz >= 32L * 32L * 32L;
}

/* k_lap0 / k_nabt: src/common/hamiltonian.f90:146-155 (identical at :209-218) */
static void k_coefs(const orc_system *s, int ik, const double vec_Ac[3], double *k_lap0,
                    double k_nabt[12]) {
  double kAc[3];
  for (int d = 0; d < 3; ++d) kAc[d] = s->vec_k[3 * ik + d] + vec_Ac[d];
  /* sum(kAc(1:3)**2), Fortran sum intrinsic in element order */
  double ss = kAc[0] * kAc[0];
  ss = ss + kAc[1] * kAc[1];
  ss = ss + kAc[2] * kAc[2];
  *k_lap0 = s->coef_lap0 + 0.5 * ss;
  for (int d = 0; d < 3; ++d)
    for (int n = 0; n < 4; ++n) k_nabt[n + 4 * d] = kAc[d] * s->coef_nab[n + 4 * d];
}

/* src/common/hamiltonian.f90:27-229,424-435: orthogonal periodic hpsi */
void orc_hpsi(const orc_system *s, const double vec_Ac[3], const double *tpsi, double *htpsi) {
  const size_t osz = orb_size(s);
  const int nb = s->nk * s->no * s->nspin;
  const int omp_order = use_omp_order(s);
#pragma omp parallel for schedule(dynamic)
  for (int b = 0; b < nb; ++b) {
    int ispin = b % s->nspin;
    int ik = b / (s->no * s->nspin);
    double k_lap0, k_nabt[12];
    k_coefs(s, ik, vec_Ac, &k_lap0, k_nabt);
    const double *V = s->vlocal + grid_size(s) * ispin;
    if (omp_order)
      orc_zstencil_omp(s->nx, s->ny, s->nz, s->ly, tpsi + C2(osz * b), htpsi + C2(osz * b), V,
                       k_lap0, s->coef_lap, k_nabt);
    else
      orc_zstencil_seq(s->nx, s->ny, s->nz, s->ly, tpsi + C2(osz * b), htpsi + C2(osz * b), V,
                       k_lap0, s->coef_lap, k_nabt);
    if (s->natom > 0 && s->nlma > 0) zpseudo_orbital(s, ik, tpsi + C2(osz * b), htpsi + C2(osz * b));
  }
}

/* rt%zc: src/rt/initialization_rt.f90:492-508 */
void orc_taylor_coefs(double dt, int n_hamil, int aetrs, double *zc) {
  double h = aetrs ? dt / 2.0 : dt;
  for (int n = 1; n <= n_hamil; ++n) {
    /* (-zi*dt)**n: Fortran integer power of a complex by repeated multiplication */
    double re = 0.0, im = -h;
    double pr = re, pi = im;
    for (int k = 2; k <= n; ++k) {
      double t = pr * re - pi * im;
      pi = pr * im + pi * re;
      pr = t;
    }
    for (int m = 1; m <= n; ++m) {
      pr = pr / m;
      pi = pi / m;
    }
    zc[2 * (n - 1)] = pr;
    zc[2 * (n - 1) + 1] = pi;
  }
}

/* out = out + zc*h over the interior of every orbital block (taylor.f90:53-114) */
static void zaxpy_interior(const orc_system *s, const double *zc, const double *h, double *out,
                           const double *base) {
  const size_t osz = orb_size(s);
  const int nb = s->nk * s->no * s->nspin;
#pragma omp parallel for collapse(2) schedule(static)
  for (int b = 0; b < nb; ++b)
    for (int iz = 0; iz < s->nz; ++iz)
      for (int iy = 0; iy < s->ny; ++iy)
        for (int ix = 0; ix < s->nx; ++ix) {
          size_t i = osz * b + (size_t)ix + (size_t)s->nx * ((size_t)iy + (size_t)s->ly * iz);
          double hr = h[C2(i)], hi = h[C2(i) + 1];
          double br = base[C2(i)], bi = base[C2(i) + 1];
          out[C2(i)] = br + (zc[0] * hr - zc[1] * hi);
          out[C2(i) + 1] = bi + (zc[0] * hi + zc[1] * hr);
        }
}

/* src/rt/taylor.f90:43-119 */
void orc_taylor(const orc_system *s, const double vec_Ac[3], const double *zc, int n_hamil,
                double *tspsi_in, double *tspsi_out, double *sshtpsi) {
  for (int nn = 1; nn <= n_hamil; ++nn) {
    const double *z = zc + 2 * (nn - 1);
    if (nn % 2 == 1) {
      orc_hpsi(s, vec_Ac, tspsi_in, sshtpsi);
      if (nn == 1)
        zaxpy_interior(s, z, sshtpsi, tspsi_out, tspsi_in);
      else
        zaxpy_interior(s, z, sshtpsi, tspsi_out, tspsi_out);
    } else {
      orc_hpsi(s, vec_Ac, sshtpsi, tspsi_in);
      zaxpy_interior(s, z, tspsi_in, tspsi_out, tspsi_out);
    }
  }
}

/* abs(z)**2 as Fortran evaluates it: |z| = hypot, then squared */
static inline double abs2(double re, double im) {
  double a = hypot(re, im);
  return a * a;
}

/* src/common/density_matrix.f90:150-230 (complex branch), serial (ik, io) order per point */
void orc_density(const orc_system *s, const double *psi, double *rho) {
  const size_t osz = orb_size(s), g = grid_size(s);
  for (int ispin = 0; ispin < s->nspin; ++ispin) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = 0; iz < s->nz; ++iz)
      for (int iy = 0; iy < s->ny; ++iy)
        for (int ix = 0; ix < s->nx; ++ix) {
          double acc = 0.0;
          size_t io_ = (size_t)ix + (size_t)s->nx * ((size_t)iy + (size_t)s->ly * iz);
          for (int ik = 0; ik < s->nk; ++ik)
            for (int io = 0; io < s->no; ++io) {
              size_t b = (size_t)ispin + (size_t)s->nspin * ((size_t)io + (size_t)s->no * ik);
              size_t i = osz * b + io_;
              double w2 = abs2(psi[C2(i)], psi[C2(i) + 1]);
              double rocc = s->rocc[(size_t)io + (size_t)s->no * ((size_t)ik + (size_t)s->nk * ispin)];
              acc = acc + (w2 * rocc) * s->wtk[ik];
            }
          rho[g * ispin + (size_t)ix + (size_t)s->nx * ((size_t)iy + (size_t)s->ny * iz)] = acc;
        }
  }
}

/* stencil_current: src/common/density_matrix.f90:465-530 (one orbital) */
static void stencil_current(const orc_system *s, const double *kAc, const double *psi,
                            double j1[3], double j2[3]) {
  const int nx = s->nx, ny = s->ny, nz = s->nz, ly = s->ly;
  double rtmp = 0.0;
  cpx t[3] = {{0, 0}, {0, 0}, {0, 0}};
  for (int iz = 0; iz < nz; ++iz)
    for (int iy = 0; iy < ny; ++iy) {
      for (int ix = 0; ix < nx; ++ix) {
        size_t i = (size_t)ix + (size_t)nx * ((size_t)iy + (size_t)ly * iz);
        rtmp = rtmp + abs2(psi[C2(i)], psi[C2(i) + 1]);
      }
      for (int d = 0; d < 3; ++d)
        for (int ix = 0; ix < nx; ++ix) {
          size_t i = (size_t)ix + (size_t)nx * ((size_t)iy + (size_t)ly * iz);
          double cr = psi[C2(i)], ci = -psi[C2(i) + 1]; /* cpsi = conjg(psi) */
          cpx acc = t[d];
          for (int n = 1; n <= 4; ++n) {
            int x = ix, y = iy, z = iz;
            if (d == 0) x = wrap(ix + n, nx);
            if (d == 1) y = wrap(iy + n, ny);
            if (d == 2) z = wrap(iz + n, nz);
            size_t k = (size_t)x + (size_t)nx * ((size_t)y + (size_t)ly * z);
            double nb = s->coef_nab[(n - 1) + 4 * d];
            /* (nabt*cpsi) * psi(neighbour) */
            double ar = nb * cr, ai = nb * ci;
            double pr = psi[C2(k)], pi = psi[C2(k) + 1];
            acc.re = acc.re + (ar * pr - ai * pi);
            acc.im = acc.im + (ar * pi + ai * pr);
          }
          t[d] = acc;
        }
    }
  for (int d = 0; d < 3; ++d) {
    j1[d] = kAc[d] * rtmp;
    j2[d] = t[d].im * 2.0; /* aimag(xtmp*2d0) */
  }
}

/* calc_current_nonlocal: src/common/density_matrix.f90:532-591 */
static void current_nonlocal(const orc_system *s, int ik, const double *psi, double jw[3]) {
  jw[0] = jw[1] = jw[2] = 0.0;
  for (int ilma = 0; ilma < s->nlma; ++ilma) {
    int ia = s->ia_tbl[ilma] - 1;
    const double *z = s->zekr + C2((size_t)s->nps * ((size_t)ilma + (size_t)s->nlma * ik));
    cpx u = {0, 0}, ur[3] = {{0, 0}, {0, 0}, {0, 0}};
    for (int j = 0; j < s->mps[ia]; ++j) {
      const double *r = s->rxyz + 3 * ((size_t)j + (size_t)s->nps * ia);
      const int *jx = s->jxyz + 3 * ((size_t)j + (size_t)s->nps * ia);
      size_t i = (size_t)(jx[0] - 1) + (size_t)s->nx * ((size_t)(jx[1] - 1) + (size_t)s->ly * (jx[2] - 1));
      double zr = z[C2(j)], zi = -z[C2(j) + 1];
      double pr = psi[C2(i)], pi = psi[C2(i) + 1];
      u.re = u.re + (zr * pr - zi * pi);
      u.im = u.im + (zr * pi + zi * pr);
      for (int d = 0; d < 3; ++d) {
        double ar = zr * r[d], ai = zi * r[d]; /* conjg(zekr)*x */
        ur[d].re = ur[d].re + (ar * pr - ai * pi);
        ur[d].im = ur[d].im + (ar * pi + ai * pr);
      }
    }
    double rv = s->rinv_uvu[ilma];
    u.re = u.re * rv;
    u.im = u.im * rv;
    for (int d = 0; d < 3; ++d) {
      /* aimag(conjg(uVpsi_r)*uVpsi) */
      double ar = ur[d].re, ai = -ur[d].im;
      jw[d] = jw[d] + (ar * u.im + ai * u.re);
    }
  }
  for (int d = 0; d < 3; ++d) jw[d] = jw[d] * 2.0;
}

/* calc_current: src/common/density_matrix.f90:240-463 (orthogonal: B = I) */
void orc_current(const orc_system *s, const double vec_Ac[3], const double *psi, double *curr) {
  const size_t osz = orb_size(s);
  const int nb = s->nk * s->no;
  double *parts = (double *)malloc(sizeof(double) * 9 * (size_t)nb);
  long ngrid = (long)s->nx * s->ny * s->nz;
  for (int ispin = 0; ispin < s->nspin; ++ispin) {
#pragma omp parallel for schedule(dynamic)
    for (int b = 0; b < nb; ++b) {
      int ik = b / s->no, io = b % s->no;
      double kAc[3];
      for (int d = 0; d < 3; ++d) kAc[d] = s->vec_k[3 * ik + d] + vec_Ac[d];
      size_t blk = (size_t)ispin + (size_t)s->nspin * ((size_t)io + (size_t)s->no * ik);
      const double *p = psi + C2(osz * blk);
      stencil_current(s, kAc, p, parts + 9 * b, parts + 9 * b + 3);
      if (s->natom > 0 && s->nlma > 0)
        current_nonlocal(s, ik, p, parts + 9 * b + 6);
      else
        parts[9 * b + 6] = parts[9 * b + 7] = parts[9 * b + 8] = 0.0;
    }
    double w4[3] = {0, 0, 0};
    for (int b = 0; b < nb; ++b) {
      int ik = b / s->no, io = b % s->no;
      double rocc = s->rocc[(size_t)io + (size_t)s->no * ((size_t)ik + (size_t)s->nk * ispin)];
      for (int d = 0; d < 3; ++d) {
        double j = (parts[9 * b + d] + parts[9 * b + 3 + d]) + parts[9 * b + 6 + d];
        w4[d] = w4[d] + (j * rocc) * s->wtk[ik];
      }
    }
    for (int d = 0; d < 3; ++d) curr[3 * ispin + d] = w4[d] / (double)ngrid;
  }
  free(parts);
}

/* calc_dip + subdip: src/rt/dip.f90:23-112, 114-225.
   out[0:3] = Dp_e = -rbox(1:3)*Hgs*Hvol, out[3] = rNe = rbox(4)*Hvol */
void orc_dipole(const orc_system *s, const double *rho, double out[4]) {
  const int n[3] = {s->nx, s->ny, s->nz};
  double rbox[4] = {0, 0, 0, 0};
  for (int d = 0; d < 4; ++d) {
    double acc = 0.0;
    for (int is = 0; is < s->nspin; ++is)
      for (int iz = 0; iz < s->nz; ++iz)
        for (int iy = 0; iy < s->ny; ++iy)
          for (int ix = 0; ix < s->nx; ++ix) {
            double r = rho[(size_t)ix + (size_t)s->nx * ((size_t)iy + (size_t)s->ny * ((size_t)iz + (size_t)s->nz * is))];
            if (d == 3) {
              acc = acc + r;
            } else {
              int i1 = (d == 0 ? ix : d == 1 ? iy : iz) + 1; /* Fortran index */
              double pos = (n[d] % 2 == 1) ? (double)i1 : ((double)i1 - 0.5);
              acc = acc + pos * r;
            }
          }
    rbox[d] = acc;
  }
  for (int d = 0; d < 3; ++d) out[d] = -rbox[d] * s->hgs[d] * s->hvol;
  out[3] = rbox[3] * s->hvol;
}

/* 'Asin2cos' pulse: src/rt/em_field.f90:177-187, A(t) tabulated at t = i*dt, i = 0..nt+1
   (calc_Ac_ext_t(0, dt, 0, nt+1, ...) in initialization_rt); f0/omega = A0. */
void orc_ac_asin2cos(double A0, double omega, double tw, double phi_cep, const double epdir[3],
                     double dt, int nt, double *ac) {
  const double pi = 3.141592653589793238462643383279502884;
  for (int i = 0; i <= nt + 1; ++i) {
    double t = 0.0 + i * dt;
    for (int d = 0; d < 3; ++d) ac[3 * i + d] = 0.0;
    if (t < tw) {
      double sn = sin(pi * t / tw);
      double env = sn * sn;
      double c = cos(omega * t + phi_cep * 2.0 * pi);
      for (int d = 0; d < 3; ++d) ac[3 * i + d] = -epdir[d] * A0 * env * c;
    }
  }
}

/*
 * Plain RT loop with a fixed local potential (yn_fix_func = 'y'), periodic,
 * middlepoint propagator, trans_longi = 'tr' (Ac_tot = Ac_ext):
 * src/rt/main_tddft.f90:89-130 and src/rt/time_evolution_step.f90:116-127,
 * 146-149, 160-170, 243-253, 264.  psi holds the initial orbitals on entry and
 * the final orbitals on exit.  Per step itt = 1..nt it records the current
 * (calc_current with vec_Ac = Ac_tot(itt) and the phase table of the
 * midpoint A, as the reference does), the electronic dipole and rNe.
 */
int orc_rt_run(orc_system *s, double dt, int n_hamil, int nt, const double *ac_tot, double *psi,
               double *curr_t, double *dip_t, double *rne_t) {
  const size_t tot = C2(orb_size(s) * (size_t)s->nspin * s->no * s->nk);
  double *a = psi;
  double *b = (double *)calloc(tot, sizeof(double));
  double *w = (double *)calloc(tot, sizeof(double));
  double *rho = (double *)malloc(sizeof(double) * grid_size(s) * s->nspin);
  double *zc = (double *)malloc(sizeof(double) * 2 * n_hamil);
  if (!b || !w || !rho || !zc) return -1;
  orc_taylor_coefs(dt, n_hamil, 0, zc);
  for (int itt = 1; itt <= nt; ++itt) {
    double A[3];
    for (int d = 0; d < 3; ++d) A[d] = 0.5 * (ac_tot[3 * itt + d] + ac_tot[3 * (itt - 1) + d]);
    if (s->natom > 0) orc_update_kvector(s, A);
    orc_taylor(s, A, zc, n_hamil, a, b, w);
    orc_density(s, b, rho);
    double Ac[3] = {ac_tot[3 * itt], ac_tot[3 * itt + 1], ac_tot[3 * itt + 2]};
    double cur[6];
    orc_current(s, Ac, b, cur);
    for (int d = 0; d < 3; ++d) {
      double c = cur[d];
      if (s->nspin == 2) c = c + cur[3 + d];
      curr_t[3 * (itt - 1) + d] = c;
    }
    double dp[4];
    orc_dipole(s, rho, dp);
    for (int d = 0; d < 3; ++d) dip_t[3 * (itt - 1) + d] = dp[d];
    rne_t[itt - 1] = dp[3];
    double *t = a;
    a = b;
    b = t;
  }
  if (a != psi) {
    memcpy(psi, a, sizeof(double) * tot);
    free(a);
  } else {
    free(b);
  }
  free(w);
  free(rho);
  free(zc);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Twiddle table e(k, j) = exp(i 2 pi (j k) / n), stored [k*n + j]
   (src/common/initialization.f90:761-784: egx(kx,ix), egxc = conjg) */
static void twiddles(int n, double *er, double *ei) {
  const double pi = 3.141592653589793238462643383279502884;
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k) {
      const double ang = 2.0 * pi * (double)((long)j * k) / (double)n;
      er[(size_t)k * n + j] = cos(ang);
      ei[(size_t)k * n + j] = sin(ang);
    }
}

/* out(k along axis d) = sum_j conj?(e(k, j)) in(j along axis d) for every line of axis d */
static void dft_axis(const int n[3], int d, int conj_tw, const double *er, const double *ei, const double *inr,
                     const double *ini, double *outr, double *outi) {
  const int m = n[d];
  const size_t stride = d == 0 ? 1 : (d == 1 ? (size_t)n[0] : (size_t)n[0] * n[1]);
  const size_t nlines = (size_t)n[0] * n[1] * n[2] / (size_t)m;
  const double sg = conj_tw ? -1.0 : 1.0;
#pragma omp parallel for schedule(static)
  for (long l = 0; l < (long)nlines; ++l) {
    /* line start: the point with index 0 along d */
    size_t base;
    if (d == 0) base = (size_t)l * m;
    else if (d == 1) base = (size_t)(l % n[0]) + (size_t)(l / n[0]) * (size_t)n[0] * n[1];
    else base = (size_t)l;
    for (int k = 0; k < m; ++k) {
      double sr = 0.0, si = 0.0;
      for (int j = 0; j < m; ++j) {
        const double c = er[(size_t)k * m + j], t = sg * ei[(size_t)k * m + j];
        const double xr = inr[base + stride * j], xi = ini[base + stride * j];
        sr = sr + (c * xr - t * xi);
        si = si + (c * xi + t * xr);
      }
      outr[base + stride * k] = sr;
      outi[base + stride * k] = si;
    }
  }
}

/* Hartree potential of a periodic orthogonal cell: poisson_ft
   (src/poisson/poisson_periodic.f90:23-242) as separable discrete Fourier sums:
   forward z, y, x with egzc/egyc/egxc and 1/N at the x stage, V(G) = coef(G) rho(G)
   with coef = 4 pi/|G|^2, 0 at G = 0, G index i -> i - n for i >= (n+1)/2
   (src/common/initialization.f90:705-724), inverse z, y, x with egz/egy/egx, Vh = Re. */
void orc_hartree(int nx, int ny, int nz, const double al[3], const double *rho, double *vh) {
  const double pi = 3.141592653589793238462643383279502884;
  const int n[3] = {nx, ny, nz};
  const size_t N = (size_t)nx * ny * nz;
  double *ar = (double *)malloc(sizeof(double) * N), *ai = (double *)malloc(sizeof(double) * N);
  double *br = (double *)malloc(sizeof(double) * N), *bi = (double *)malloc(sizeof(double) * N);
  double *er[3], *ei[3];
  for (int d = 0; d < 3; ++d) {
    er[d] = (double *)malloc(sizeof(double) * n[d] * n[d]);
    ei[d] = (double *)malloc(sizeof(double) * n[d] * n[d]);
    twiddles(n[d], er[d], ei[d]);
  }
  for (size_t i = 0; i < N; ++i) { ar[i] = rho[i]; ai[i] = 0.0; }
  dft_axis(n, 2, 1, er[2], ei[2], ar, ai, br, bi);
  dft_axis(n, 1, 1, er[1], ei[1], br, bi, ar, ai);
  dft_axis(n, 0, 1, er[0], ei[0], ar, ai, br, bi);
  for (int kz = 0; kz < nz; ++kz)
    for (int ky = 0; ky < ny; ++ky)
      for (int kx = 0; kx < nx; ++kx) {
        const int ik[3] = {kx, ky, kz};
        double g2 = 0.0;
        for (int d = 0; d < 3; ++d) {
          const int ii = ik[d] >= (n[d] + 1) / 2 ? ik[d] - n[d] : ik[d];
          const double g = (double)ii * (2.0 * pi / al[d]);
          g2 = g2 + g * g;
        }
        const size_t i = (size_t)kx + (size_t)nx * ((size_t)ky + (size_t)ny * kz);
        const double coef = (kx == 0 && ky == 0 && kz == 0) ? 0.0 : 4.0 * pi / g2;
        br[i] = coef * (br[i] / (double)N);
        bi[i] = coef * (bi[i] / (double)N);
      }
  dft_axis(n, 2, 0, er[2], ei[2], br, bi, ar, ai);
  dft_axis(n, 1, 0, er[1], ei[1], ar, ai, br, bi);
  dft_axis(n, 0, 0, er[0], ei[0], br, bi, ar, ai);
  for (size_t i = 0; i < N; ++i) vh[i] = ar[i];
  for (int d = 0; d < 3; ++d) { free(er[d]); free(ei[d]); }
  free(ar); free(ai); free(br); free(bi);
}

/* PZ81 LDA, unpolarised: exc_cor_pz / PZxc (src/xc/builtin_pz.f90:23-83) with
   trho = rho (rho_s = rho/2 then trho = 2 rho_s, salmon_xc.f90:778-795). */
void orc_xc_pz(long n, const double *rho, double *vxc, double *exc_out) {
  const double Pi = 3.141592653589793;
  const double gammaU = -0.1423, beta1U = 1.0529, beta2U = 0.3334, AU = 0.0311, BU = -0.048;
  const double CU = 0.002, DU = -0.0116;
  const double cst = 0.75 * pow(3.0 / (2.0 * Pi), 2.0 / 3.0);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    const double rho_s = rho[i] * 0.5;
    const double trho = 2 * rho_s;
    const double ttrho = trho + 1e-10;
    const double rs = pow(3.0 / (4 * Pi * ttrho), 1.0 / 3.0);
    double exc = -cst / rs;
    double dexc = exc / (3 * ttrho);
    if (rs > 1.0) {
      const double rssq = sqrt(rs);
      exc = exc + gammaU / (1.0 + beta1U * rssq + beta2U * rs);
      dexc = dexc + gammaU * (0.5 * beta1U * rssq + beta2U * rs) / (3 * ttrho) /
                        ((1 + beta1U * rssq + beta2U * rs) * (1 + beta1U * rssq + beta2U * rs));
    } else {
      const double rsln = log(rs);
      exc = exc + AU * rsln + BU + CU * rs * rsln + DU * rs;
      dexc = dexc - rs / (3 * ttrho) * (AU / rs + CU * (rsln + 1) + DU);
    }
    vxc[i] = exc + trho * dexc;
    if (exc_out) exc_out[i] = exc;
  }
}

/* PZ81 LDA, spin-polarised (collinear): exc_cor_pz_sp / PZxc_sp
   (src/xc/builtin_pz_sp.f90:23-98), called with the channel densities as they are
   (salmon_xc.f90:780, 797).  rho_s, vxc_s: (n, 2), channel-major; exc_out may be NULL.
   Terms are added in the order the Fortran expressions give them (left to right). */
void orc_xc_pz_sp(long n, const double *rho_s, double *vxc_s, double *exc_out) {
  const double Pi = 3.141592653589793;
  const double gammaU = -0.1423, gammaP = -0.0843;
  const double beta1U = 1.0529, beta2U = 0.3334, beta1P = 1.3981, beta2P = 0.2611;
  const double AU = 0.0311, BU = -0.048, CU = 0.002, DU = -0.0116;
  const double AP = 0.01555, BP = -0.0269, CP = 0.0007, DP = -0.0048;
  const double cx = 3.0 / 4.0 * pow(3.0 / (2 * Pi), 2.0 / 3.0);
  const double c13 = -pow(2.0, 1.0 / 3.0) + 1.0;  /* (2^(1/3) - 1) with the sign of the source */
  const double fden = pow(2.0, 4.0 / 3.0) - 2.0;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    const double trho = rho_s[i] + rho_s[n + i] + 1e-20;
    const double tu = rho_s[i] + 5e-21, td = rho_s[n + i] + 5e-21;
    const double rs = pow(3.0 / (4.0 * Pi * trho), 1.0 / 3.0);
    const double zeta = (tu - td) / trho;
    const double sf = (pow(1.0 + zeta, 4.0 / 3.0) + pow(1.0 - zeta, 4.0 / 3.0) - 2.0) / fden;
    const double dsf = 4.0 / 3.0 * (pow(1.0 + zeta, 1.0 / 3.0) - pow(1.0 - zeta, 1.0 / 3.0)) / fden;
    double exc = -cx / rs + c13 * cx / rs * sf;
    double v[2];
    for (int js = 0; js < 2; ++js) {
      const double sg = js == 0 ? 1.0 : -1.0;
      v[js] = exc - 1.0 / 3.0 * cx / rs + 1.0 / 3.0 * c13 * cx / rs * sf + c13 * cx / rs * (sg - zeta) * dsf;
    }
    double ecu, ecp, vcu, vcp;
    if (rs > 1.0) {
      const double du = 1.0 + beta1U * sqrt(rs) + beta2U * rs, dp = 1.0 + beta1P * sqrt(rs) + beta2P * rs;
      ecu = gammaU / du;
      ecp = gammaP / dp;
      vcu = gammaU * (1.0 + 7.0 / 6.0 * beta1U * sqrt(rs) + 4.0 / 3.0 * beta2U * rs) / (du * du);
      vcp = gammaP * (1.0 + 7.0 / 6.0 * beta1P * sqrt(rs) + 4.0 / 3.0 * beta2P * rs) / (dp * dp);
    } else {
      const double L = log(rs);
      ecu = AU * L + BU + CU * rs * L + DU * rs;
      ecp = AP * L + BP + CP * rs * L + DP * rs;
      vcu = AU * L + (BU - AU / 3.0) + 2.0 / 3.0 * CU * rs * L + (2.0 * DU - CU) / 3.0 * rs;
      vcp = AP * L + (BP - AP / 3.0) + 2.0 / 3.0 * CP * rs * L + (2.0 * DP - CP) / 3.0 * rs;
    }
    exc = exc + ecu + sf * (ecp - ecu);
    for (int js = 0; js < 2; ++js) {
      const double sg = js == 0 ? 1.0 : -1.0;
      vxc_s[(size_t)js * n + i] = v[js] + vcu + sf * (vcp - vcu) + (ecp - ecu) * (sg - zeta) * dsf;
    }
    if (exc_out) exc_out[i] = exc;
  }
}

/* V_local(is) = Vpsl + Vh + Vxc(is) (update_vlocal, src/common/hamiltonian.f90:813-843) from
   the channel densities: Vh of the total density rho_s(1) + rho_s(2)
   (time_evolution_step.f90:162-169), Vxc from PZ (nspin = 1) or spin-polarised PZ. */
static void vlocal_from_rho(const orc_system *s, const double *vpsl, const double al[3], const double *rho,
                            double *vloc) {
  const size_t g = grid_size(s);
  const int ns = s->nspin;
  double *vh = (double *)malloc(sizeof(double) * g), *vxc = (double *)malloc(sizeof(double) * g * ns);
  if (ns == 2) {
    double *rt = (double *)malloc(sizeof(double) * g);
    for (size_t i = 0; i < g; ++i) rt[i] = rho[i] + rho[g + i];
    orc_hartree(s->nx, s->ny, s->nz, al, rt, vh);
    orc_xc_pz_sp((long)g, rho, vxc, NULL);
    free(rt);
  } else {
    orc_hartree(s->nx, s->ny, s->nz, al, rho, vh);
    orc_xc_pz((long)g, rho, vxc, NULL);
  }
  for (int is = 0; is < ns; ++is)
    for (size_t i = 0; i < g; ++i) vloc[(size_t)is * g + i] = (vpsl[i] + vh[i]) + vxc[(size_t)is * g + i];
  free(vh);
  free(vxc);
}

/*
 * General RT loop: fixed V_local (scf = 0, s->vlocal as given) or V_local rebuilt from the
 * density every step (scf = 1, yn_fix_func = 'n': src/rt/time_evolution_step.f90:160-166,
 * 182-207, initial potential from the initial density, src/rt/initialization_rt.f90), with
 * the optional predictor_corrector (:141-148, 354-423), and the field update of
 * calc_emfields (src/rt/em_field.f90:24-70): trans_longi 'tr' (lo = 0, Ac_tot = Ac_ext) or
 * 'lo' (lo = 1, Ac_ind(itt+1) = 2 Ac_ind(itt) - Ac_ind(itt-1) - 4 pi J(itt) dt^2).
 * ac: (3, 0:nt+1); on entry Ac_ext, on exit Ac_tot.  vpsl / vloc / al only with scf = 1.
 */
int orc_rt_run_full(orc_system *s, int scf, int pc, int lo, const double *vpsl, double *vloc, const double al[3],
                    double dt, int n_hamil, int nt, double *ac, double *psi, double *curr_t, double *dip_t,
                    double *rne_t) {
  return orc_rt_run_prop(s, scf, pc, lo, 0, vpsl, vloc, al, dt, n_hamil, nt, ac, psi, curr_t, dip_t, rne_t);
}

/*
 * The same loop with the propagator chosen: aetrs = 0 'middlepoint' as above; aetrs = 1
 * 'aetrs' (time_evolution_step.f90:119-123, 140-141, 425-475): zc of dt/2
 * (initialization_rt.f90:492-508); per step vec_Ac = Ac_tot(itt), a Taylor half step with
 * V_local(t), then V_local extrapolated, V_new = 3 V - 3 V_old1 + V_old2 with the history
 * shifted (V_old2 = V_old1, V_old1 = V; both start at the initial V_local,
 * initialization_rt.f90:290-291), vec_Ac = Ac_ext(itt+1) + 2 Ac_ind(itt) - Ac_ind(itt-1), and
 * the second half step; the predictor-corrector is not used with aetrs.  With a fixed
 * potential the extrapolation is the identity (3V - 3V + V = V exactly) and is skipped.
 */
int orc_rt_run_prop(orc_system *s, int scf, int pc, int lo, int aetrs, const double *vpsl, double *vloc,
                    const double al[3], double dt, int n_hamil, int nt, double *ac, double *psi, double *curr_t,
                    double *dip_t, double *rne_t) {
  const double pi = 3.141592653589793238462643383279502884;
  pc = pc && scf && !aetrs;
  const size_t tot = C2(orb_size(s) * (size_t)s->nspin * s->no * s->nk);
  const size_t g = grid_size(s);
  double *a = psi;
  double *b = (double *)calloc(tot, sizeof(double));
  double *w = (double *)calloc(tot, sizeof(double));
  double *psi_tmp = pc ? (double *)malloc(sizeof(double) * tot) : NULL;
  double *rho = (double *)malloc(sizeof(double) * g * s->nspin);
  const size_t gv = g * (size_t)s->nspin;  /* V_local(ix,iy,iz,ispin) */
  double *v_tmp = (double *)malloc(sizeof(double) * gv);
  double *zc = (double *)malloc(sizeof(double) * 2 * n_hamil);
  double *ac_ind = (double *)calloc(3 * (size_t)(nt + 2), sizeof(double));
  double *vold = (aetrs && scf) ? (double *)malloc(sizeof(double) * 2 * gv) : NULL;
  if (!b || !w || !rho || !v_tmp || !zc || !ac_ind || (pc && !psi_tmp) || (aetrs && scf && !vold)) return -1;
  orc_taylor_coefs(dt, n_hamil, aetrs, zc);
  if (scf) {
    s->vlocal = vloc;
    orc_density(s, a, rho);
    vlocal_from_rho(s, vpsl, al, rho, vloc);
    if (vold) {
      memcpy(vold, vloc, sizeof(double) * gv);
      memcpy(vold + gv, vloc, sizeof(double) * gv);
    }
  }
  for (int itt = 1; itt <= nt; ++itt) {
    double A[3];
    for (int d = 0; d < 3; ++d)
      A[d] = aetrs ? ac[3 * itt + d] : 0.5 * (ac[3 * itt + d] + ac[3 * (itt - 1) + d]);
    if (s->natom > 0) orc_update_kvector(s, A);
    if (aetrs) {
      orc_taylor(s, A, zc, n_hamil, a, b, w);
      if (vold) {
        double *v1 = vold, *v2 = vold + gv;
        for (size_t i = 0; i < gv; ++i) {
          const double vn = 3.0 * vloc[i] - 3.0 * v1[i] + v2[i];
          v2[i] = v1[i];
          v1[i] = vloc[i];
          vloc[i] = vn;
        }
      }
      /* Ac_ext(itt+1) + 2 Ac_ind(itt) - Ac_ind(itt-1); ac[itt+1] holds Ac_ext(itt+1) until the
         field update at the end of this step */
      for (int d = 0; d < 3; ++d)
        A[d] = ac[3 * (itt + 1) + d] + 2.0 * ac_ind[3 * itt + d] - ac_ind[3 * (itt - 1) + d];
      if (s->natom > 0) orc_update_kvector(s, A);
      double *t = a;
      a = b;
      b = t;
    }
    if (pc) {
      memcpy(v_tmp, vloc, sizeof(double) * gv);
      memcpy(psi_tmp, a, sizeof(double) * tot);
      orc_taylor(s, A, zc, n_hamil, a, b, w);
      orc_density(s, b, rho);
      vlocal_from_rho(s, vpsl, al, rho, vloc);
      for (size_t i = 0; i < gv; ++i) vloc[i] = 0.5 * (v_tmp[i] + vloc[i]);
      memcpy(a, psi_tmp, sizeof(double) * tot);
    }
    orc_taylor(s, A, zc, n_hamil, a, b, w);
    orc_density(s, b, rho);
    if (scf) vlocal_from_rho(s, vpsl, al, rho, vloc);
    double Ac[3] = {ac[3 * itt], ac[3 * itt + 1], ac[3 * itt + 2]};
    double cur[6];
    orc_current(s, Ac, b, cur);
    for (int d = 0; d < 3; ++d) {
      double c = cur[d];
      if (s->nspin == 2) c = c + cur[3 + d];
      curr_t[3 * (itt - 1) + d] = c;
    }
    if (lo) {
      for (int d = 0; d < 3; ++d) {
        double *ai = ac_ind + 3 * (size_t)(itt + 1) + d;
        *ai = 2.0 * ac_ind[3 * itt + d] - ac_ind[3 * (itt - 1) + d] - 4.0 * pi * curr_t[3 * (itt - 1) + d] * (dt * dt);
        ac[3 * (itt + 1) + d] = ac[3 * (itt + 1) + d] + *ai;  /* Ac_tot = Ac_ext + Ac_ind */
      }
    }
    double dp[4];
    orc_dipole(s, rho, dp);
    for (int d = 0; d < 3; ++d) dip_t[3 * (itt - 1) + d] = dp[d];
    rne_t[itt - 1] = dp[3];
    double *t = a;
    a = b;
    b = t;
  }
  if (a != psi) {
    memcpy(psi, a, sizeof(double) * tot);
    free(a);
  } else {
    free(b);
  }
  free(w); free(psi_tmp); free(rho); free(v_tmp); free(zc); free(ac_ind); free(vold);
  return 0;
}

int orc_rt_run_scf(orc_system *s, const double *vpsl, double *vloc, const double al[3], int pc, double dt,
                   int n_hamil, int nt, const double *ac_tot, double *psi, double *curr_t, double *dip_t,
                   double *rne_t) {
  double *ac = (double *)malloc(sizeof(double) * 3 * (size_t)(nt + 2));
  if (!ac) return -1;
  memcpy(ac, ac_tot, sizeof(double) * 3 * (size_t)(nt + 2));
  int rc = orc_rt_run_full(s, 1, pc, 0, vpsl, vloc, al, dt, n_hamil, nt, ac, psi, curr_t, dip_t, rne_t);
  free(ac);
  return rc;
}
