// Projector-table builder for KY-format pseudopotentials on an orthogonal periodic cell
// (include/salmon_pp.h).  Host-side set-up: it runs once per geometry and hands its tables
// to the device path (salmon_hip_set_nonlocal, salmon_hip_set_vlocal).
//
// Follows the reference routine by routine (file:line in include/salmon_pp.h):
//   read_ps_ky + input_pp post-processing + making_ps_without_masking -> element tables
//   spline (natural cubic, salmon_math.f90:494-584): the tridiagonal system is solved by
//     the Thomas algorithm where the reference inverts it with LAPACK: the same equations
//   calc_nps / calc_jxyz: sphere points over periodic images, images outer, z, y, x inner
//   set_lma / calc_uv: uv = spline(r) * Ylm, rinv_uvu = inorm * hvol
//   calc_vpsl_periodic: V_ion(G) from the radial local potential, inverse DFT z, y, x
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "salmon_hip.h"
#include "salmon_pp.h"

void salmon_hip_internal_error(const char *msg);  // salmon_hip.hip: sets salmon_hip_last_error()

namespace {

constexpr double kAuLengthAA = 0.52917721067;  // src/io/inputoutput.f90:27
constexpr double kAuEnergyEV = 27.21138505;    // src/io/inputoutput.f90:26
constexpr double kPi = 3.14159265358979323846;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  salmon_hip_internal_error(buf);
  return code;
}

struct Element {
  int mr = 0, mlps = 0, lref = 0, zps = 0, nrps = 0;
  double rps = 0.0;
  std::vector<double> rad;                  // rad(i) = (i-1) step, 1-based (index 0 unused)
  std::vector<double> vloctbl;              // (1..nrps)
  std::vector<std::vector<double>> udvtbl;  // [l](1..nrps)
  std::vector<double> anorm;
  std::vector<int> inorm;
  // natural-spline coefficients of udvtbl per channel, interval s = 0 .. nrps-2 between
  // rad(s+1) and rad(s+2): a xx^3 + b xx^2 + c xx + d
  std::vector<std::vector<double>> sa, sb, sc, sd;
  int nlma_atom() const {
    int n = 0;
    for (int l = 0; l <= mlps; ++l)
      if (inorm[l] != 0) n += 2 * l + 1;
    return n;
  }
};

// list-directed read: whitespace-separated values, Fortran 'D' exponents accepted
bool next_value(std::ifstream &f, double *v) {
  std::string tok;
  if (!(f >> tok)) return false;
  for (char &ch : tok)
    if (ch == 'D' || ch == 'd') ch = 'E';
  char *end = nullptr;
  *v = strtod(tok.c_str(), &end);
  return end && *end == '\0';
}

// salmon_math.f90:494-584 (natural cubic spline through (xn, yn), Np points)
void spline(const std::vector<double> &xn, const std::vector<double> &yn, std::vector<double> &an,
            std::vector<double> &bn, std::vector<double> &cn, std::vector<double> &dn) {
  const int Np = (int)xn.size();
  const int Npm2 = Np - 2;
  std::vector<double> dxn(Np - 1), dyn(Np - 1);
  for (int i = 0; i <= Np - 2; ++i) {
    dxn[i] = xn[i + 1] - xn[i];
    dyn[i] = yn[i + 1] - yn[i];
  }
  // tridiagonal system A u = v for u(1..Npm2) (the reference's Amat), Thomas algorithm
  std::vector<double> lo(Npm2 + 1, 0.0), di(Npm2 + 1, 0.0), up(Npm2 + 1, 0.0), v(Npm2 + 1, 0.0);
  for (int i = 1; i <= Npm2; ++i) {
    v[i] = 6.0 * (dyn[i] / dxn[i] - dyn[i - 1] / dxn[i - 1]);
    di[i] = 2.0 * (dxn[i] + dxn[i - 1]);
    if (i < Npm2) up[i] = dxn[i];
    if (i > 1) lo[i] = dxn[i - 1];
  }
  std::vector<double> u(Npm2 + 2, 0.0);
  for (int i = 2; i <= Npm2; ++i) {
    const double w = lo[i] / di[i - 1];
    di[i] -= w * up[i - 1];
    v[i] -= w * v[i - 1];
  }
  for (int i = Npm2; i >= 1; --i) u[i] = (v[i] - (i < Npm2 ? up[i] * u[i + 1] : 0.0)) / di[i];
  an.assign(Npm2 + 1, 0.0);
  bn.assign(Npm2 + 1, 0.0);
  cn.assign(Npm2 + 1, 0.0);
  dn.assign(Npm2 + 1, 0.0);
  for (int i = 1; i <= Npm2; ++i) bn[i] = 0.5 * u[i];
  for (int i = 0; i <= Npm2 - 1; ++i) an[i] = (u[i + 1] - 2.0 * bn[i]) / (6.0 * dxn[i]);
  an[Npm2] = (0.0 - 2.0 * bn[Npm2]) / (6.0 * dxn[Npm2]);
  for (int i = 0; i <= Npm2; ++i) dn[i] = yn[i];
  cn[0] = dyn[0] / dxn[0] - dxn[0] * (u[1] + 2.0 * 0.0) / 6.0;
  for (int i = 1; i <= Npm2 - 1; ++i) cn[i] = dyn[i] / dxn[i] - dxn[i] * (u[i + 1] + 2.0 * u[i]) / 6.0;
  cn[Npm2] = dyn[Npm2] / dxn[Npm2] - dxn[Npm2] * (0.0 + 2.0 * u[Npm2]) / 6.0;
}

// salmon_math.f90:386-426 (monomial Ylm) for l <= 2
double ylm(double x, double y, double z, int l, int m) {
  const double r2 = x * x + y * y + z * z;
  switch (l * l + l + 1 + m) {
    case 1: return 1.0;
    case 2: return -y;
    case 3: return z;
    case 4: return -x;
    case 5: return std::sqrt(3.0) * x * y;
    case 6: return -std::sqrt(3.0) * y * z;
    case 7: return (3.0 * z * z - r2) / 2.0;
    case 8: return -std::sqrt(3.0) * x * z;
    case 9: return std::sqrt(3.0 / 4.0) * (x * x - y * y);
    default: return 0.0;
  }
}

int read_ky(const char *path, int lmax_ps, int lloc_ps, Element &e) {
  std::ifstream f(path);
  if (!f) return fail(SALMON_HIP_EINVAL, "cannot open pseudopotential file %s", path);
  double v[4];
  for (double &x : v)
    if (!next_value(f, &x)) return fail(SALMON_HIP_EINVAL, "%s: bad header", path);
  const int mr = (int)v[0];
  double step = v[1];
  int mlps = (int)v[2];
  e.zps = (int)(v[3] + 1e-10);
  if (mr < 4 || mr > 100000 || mlps < 0 || mlps > 2 || step <= 0.0)
    return fail(SALMON_HIP_EUNSUPPORTED, "%s: mr=%d mlps=%d step=%g outside the supported range", path, mr, mlps, step);
  double irpc = 0.0;
  std::vector<double> rrc(mlps + 1);
  if (!next_value(f, &irpc)) return fail(SALMON_HIP_EINVAL, "%s: bad rrc line", path);
  for (double &x : rrc)
    if (!next_value(f, &x)) return fail(SALMON_HIP_EINVAL, "%s: bad rrc line", path);
  std::vector<std::vector<double>> vpp(mlps + 1, std::vector<double>(mr + 2, 0.0)), upp = vpp;
  for (int i = 0; i <= mr; ++i) {
    double r, rho;
    if (!next_value(f, &r) || !next_value(f, &rho)) return fail(SALMON_HIP_EINVAL, "%s: short potential table", path);
    for (int l = 0; l <= mlps; ++l)
      if (!next_value(f, &vpp[l][i])) return fail(SALMON_HIP_EINVAL, "%s: short potential table", path);
  }
  for (int i = 0; i <= mr; ++i) {
    double r;
    if (!next_value(f, &r)) return fail(SALMON_HIP_EINVAL, "%s: short wave-function table", path);
    for (int l = 0; l <= mlps; ++l)
      if (!next_value(f, &upp[l][i])) return fail(SALMON_HIP_EINVAL, "%s: short wave-function table", path);
  }
  // atomic units (input_pp.f90:320-324)
  step = step / kAuLengthAA;
  for (double &x : rrc) x = x / kAuLengthAA;
  for (int l = 0; l <= mlps; ++l)
    for (int i = 0; i <= mr; ++i) {
      vpp[l][i] = vpp[l][i] / kAuEnergyEV;
      upp[l][i] = upp[l][i] * std::sqrt(kAuLengthAA);
    }
  const int nrmax = mr + 2;
  e.rad.assign(nrmax + 1, 0.0);
  for (int i = 1; i <= nrmax; ++i) e.rad[i] = (i - 1) * step;
  if (lmax_ps >= 0) {
    if (lmax_ps > mlps) return fail(SALMON_HIP_EINVAL, "%s: Lmax_ps=%d above the file's %d", path, lmax_ps, mlps);
    mlps = lmax_ps;
  }
  if (lloc_ps < 0 || lloc_ps > mlps) return fail(SALMON_HIP_EINVAL, "%s: Lloc_ps=%d outside 0..%d", path, lloc_ps, mlps);
  e.mr = mr;
  e.mlps = mlps;
  e.lref = lloc_ps;
  // rps = max rrc, nrps = first i with rad(i) > rps (input_pp.f90:114-119)
  double rps = 0.0;
  bool any = false;
  for (int l = 0; l <= mlps; ++l) {
    any = any || rrc[l] != 0.0;
    rps = l == 0 ? rrc[0] : std::max(rps, rrc[l]);
  }
  if (!any) return fail(SALMON_HIP_EUNSUPPORTED, "%s: no cut-off radius in the file", path);
  int i = 1;
  while (i <= nrmax && !(e.rad[i] > rps)) ++i;
  if (i >= nrmax) return fail(SALMON_HIP_EINVAL, "%s: NRps >= Nrmax", path);
  e.rps = rps;
  e.nrps = i;
  const int lref = e.lref;
  // anorm / inorm before the radial functions are rescaled (input_pp.f90:163-178)
  e.anorm.assign(mlps + 1, 0.0);
  e.inorm.assign(mlps + 1, 0);
  for (int l = 0; l <= mlps; ++l) {
    double a = 0.0;
    for (int k = 1; k <= mr - 1; ++k) {
      const double r1 = e.rad[k + 1] - e.rad[k];
      a = a + (upp[l][k] * upp[l][k] * (vpp[l][k] - vpp[lref][k]) +
               upp[l][k - 1] * upp[l][k - 1] * (vpp[l][k - 1] - vpp[lref][k - 1])) * r1;
    }
    a = 0.5 * a;
    e.inorm[l] = 1;
    if (a < 0.0) {
      a = -a;
      e.inorm[l] = -1;
    }
    if (std::fabs(a) < 1e-10) e.inorm[l] = 0;
    e.anorm[l] = std::sqrt(a);
  }
  // making_ps_without_masking: upp * sqrt((2l+1)/4pi) / r^(l+1); vloctbl, udvtbl / anorm
  for (int l = 0; l <= mlps; ++l) {
    for (int k = 1; k <= mr; ++k)
      upp[l][k] = upp[l][k] * std::sqrt((2 * l + 1.0) / (4 * kPi)) / std::pow(e.rad[k + 1], l + 1);
    upp[l][0] = upp[l][1];
  }
  e.vloctbl.assign(e.nrps + 1, 0.0);
  e.udvtbl.assign(mlps + 1, std::vector<double>(e.nrps + 1, 0.0));
  for (int l = 0; l <= mlps; ++l) {
    for (int k = 1; k <= e.nrps; ++k) {
      e.vloctbl[k] = vpp[lref][k - 1];
      e.udvtbl[l][k] = (vpp[l][k - 1] - vpp[lref][k - 1]) * upp[l][k - 1];
    }
    if (e.inorm[l] == 0) continue;
    for (int k = 1; k <= e.nrps; ++k) e.udvtbl[l][k] = e.udvtbl[l][k] / e.anorm[l];
  }
  // spline coefficients of calc_uv (prep_pp.f90:434-455)
  std::vector<double> xn(e.rad.begin() + 1, e.rad.begin() + 1 + e.nrps);
  e.sa.resize(mlps + 1);
  e.sb.resize(mlps + 1);
  e.sc.resize(mlps + 1);
  e.sd.resize(mlps + 1);
  for (int l = 0; l <= mlps; ++l) {
    std::vector<double> yn(e.udvtbl[l].begin() + 1, e.udvtbl[l].begin() + 1 + e.nrps);
    spline(xn, yn, e.sa[l], e.sb[l], e.sc[l], e.sd[l]);
  }
  return 0;
}

struct Cell {
  double al[3], hgs[3], hvol;
  int n[3];
};

int check_cell(const salmon_pp *pp, int natom, const int *kion, const double *rion, const double *al, const int *n,
               Cell *c);

}  // namespace

struct salmon_pp {
  std::vector<Element> el;
};

namespace {

int check_cell(const salmon_pp *pp, int natom, const int *kion, const double *rion, const double *al, const int *n,
               Cell *c) {
  if (!pp) return fail(SALMON_HIP_EINVAL, "null pseudopotential handle");
  if (natom < 1 || !kion || !rion || !al || !n) return fail(SALMON_HIP_EINVAL, "invalid atoms or cell");
  for (int d = 0; d < 3; ++d) {
    if (!(al[d] > 0.0) || n[d] < 1) return fail(SALMON_HIP_EINVAL, "cell edge %d: al=%g n=%d", d + 1, al[d], n[d]);
    c->al[d] = al[d];
    c->n[d] = n[d];
    c->hgs[d] = al[d] / n[d];
  }
  for (int ia = 0; ia < natom; ++ia)
    if (kion[ia] < 1 || kion[ia] > (int)pp->el.size())
      return fail(SALMON_HIP_EINVAL, "kion(%d)=%d outside 1..%zu", ia + 1, kion[ia], pp->el.size());
  c->hvol = al[0] * al[1] * al[2] / ((double)n[0] * n[1] * n[2]);
  return 0;
}

// calc_nps / calc_jxyz (prep_pp.f90:195-370) for an orthogonal cell; f(ia, i1, i2, i3, x, y, z) per point
template <typename F>
void sphere_points(const salmon_pp *pp, int natom, const int *kion, const double *rion, const Cell &c, F f) {
  double rps_max = 0.0;
  for (const Element &e : pp->el) rps_max = std::max(rps_max, e.rps);
  rps_max = rps_max + 1.5 * std::max(c.hgs[0], std::max(c.hgs[1], c.hgs[2])) + 1e-2;
  int nc[3];
  for (int d = 0; d < 3; ++d) {
    double rmin = rion[d], rmax = rion[d];
    for (int ia = 1; ia < natom; ++ia) {
      rmin = std::min(rmin, rion[3 * ia + d]);
      rmax = std::max(rmax, rion[3 * ia + d]);
    }
    nc[d] = (rmin + 2.0 * c.al[d] > c.al[d] + rps_max && rmax - 2.0 * c.al[d] < -rps_max) ? 1 : 2;
  }
  for (int ia = 0; ia < natom; ++ia) {
    const double rps = pp->el[kion[ia] - 1].rps;
    const double *R = rion + 3 * ia;
    for (int j1 = -nc[0]; j1 <= nc[0]; ++j1) {
      const double x1 = R[0] + j1 * c.al[0];
      if (x1 <= 1 * c.hgs[0] - rps_max || x1 >= c.n[0] * c.hgs[0] + rps_max) continue;
      for (int j2 = -nc[1]; j2 <= nc[1]; ++j2) {
        const double x2 = R[1] + j2 * c.al[1];
        if (x2 <= 1 * c.hgs[1] - rps_max || x2 >= c.n[1] * c.hgs[1] + rps_max) continue;
        for (int j3 = -nc[2]; j3 <= nc[2]; ++j3) {
          const double x3 = R[2] + j3 * c.al[2];
          if (x3 <= 1 * c.hgs[2] - rps_max || x3 >= c.n[2] * c.hgs[2] + rps_max) continue;
          for (int i3 = 1; i3 <= c.n[2]; ++i3)
            for (int i2 = 1; i2 <= c.n[1]; ++i2)
              for (int i1 = 1; i1 <= c.n[0]; ++i1) {
                const double x = (i1 * c.hgs[0] + (-c.hgs[0])) - x1;
                const double y = (i2 * c.hgs[1] + (-c.hgs[1])) - x2;
                const double z = (i3 * c.hgs[2] + (-c.hgs[2])) - x3;
                const double r = std::sqrt(x * x + y * y + z * z);
                if (r < rps + 1e-12) f(ia, i1, i2, i3, x, y, z);
              }
        }
      }
    }
  }
}

}  // namespace

extern "C" {

int salmon_pp_create(int nelem, const char *const *files, const int *lmax_ps, const int *lloc_ps,
                     salmon_pp_t *out) {
  if (!out || nelem < 1 || !files || !lloc_ps) return fail(SALMON_HIP_EINVAL, "salmon_pp_create: invalid arguments");
  *out = nullptr;
  salmon_pp *pp = new salmon_pp();
  pp->el.resize(nelem);
  for (int k = 0; k < nelem; ++k) {
    if (!files[k]) {
      delete pp;
      return fail(SALMON_HIP_EINVAL, "file of element %d is null", k + 1);
    }
    if (int rc = read_ky(files[k], lmax_ps ? lmax_ps[k] : -1, lloc_ps[k], pp->el[k])) {
      delete pp;
      return rc;
    }
  }
  *out = pp;
  return 0;
}

int salmon_pp_destroy(salmon_pp_t pp) {
  delete pp;
  return 0;
}

int salmon_pp_element(salmon_pp_t pp, int ielem, int *zps, int *mlps, int *lref, double *rps, int *nrps,
                      int *nlma_atom) {
  if (!pp || ielem < 1 || ielem > (int)pp->el.size()) return fail(SALMON_HIP_EINVAL, "element %d", ielem);
  const Element &e = pp->el[ielem - 1];
  if (zps) *zps = e.zps;
  if (mlps) *mlps = e.mlps;
  if (lref) *lref = e.lref;
  if (rps) *rps = e.rps;
  if (nrps) *nrps = e.nrps;
  if (nlma_atom) *nlma_atom = e.nlma_atom();
  return 0;
}

int salmon_pp_sizes(salmon_pp_t pp, int natom, const int *kion, const double *rion, const double *al,
                    const int *n, int *nps, int *nlma) {
  Cell c;
  if (int rc = check_cell(pp, natom, kion, rion, al, n, &c)) return rc;
  std::vector<int> cnt(natom, 0);
  sphere_points(pp, natom, kion, rion, c, [&](int ia, int, int, int, double, double, double) { ++cnt[ia]; });
  int m = 0, l = 0;
  for (int ia = 0; ia < natom; ++ia) {
    m = std::max(m, cnt[ia]);
    l += pp->el[kion[ia] - 1].nlma_atom();
  }
  if (nps) *nps = m;
  if (nlma) *nlma = l;
  return 0;
}

int salmon_pp_nonlocal(salmon_pp_t pp, int natom, const int *kion, const double *rion, const double *al,
                       const int *n, int nps, int nlma, int *mps, int *jxyz, double *rxyz, int *ia_tbl,
                       double *rinv_uvu, double *uv) {
  Cell c;
  if (int rc = check_cell(pp, natom, kion, rion, al, n, &c)) return rc;
  if (!mps || !jxyz || !rxyz || !ia_tbl || !rinv_uvu || !uv) return fail(SALMON_HIP_EINVAL, "null output");
  int want_nps = 0, want_nlma = 0;
  if (int rc = salmon_pp_sizes(pp, natom, kion, rion, al, n, &want_nps, &want_nlma)) return rc;
  if (nps < want_nps || nlma != want_nlma)
    return fail(SALMON_HIP_EINVAL, "nps=%d nlma=%d, need nps >= %d and nlma = %d", nps, nlma, want_nps, want_nlma);
  memset(jxyz, 0, sizeof(int) * 3 * (size_t)nps * natom);
  memset(rxyz, 0, sizeof(double) * 3 * (size_t)nps * natom);
  memset(uv, 0, sizeof(double) * (size_t)nps * nlma);
  for (int ia = 0; ia < natom; ++ia) mps[ia] = 0;
  sphere_points(pp, natom, kion, rion, c, [&](int ia, int i1, int i2, int i3, double x, double y, double z) {
    const size_t o = 3 * ((size_t)mps[ia] + (size_t)nps * ia);
    jxyz[o] = i1; jxyz[o + 1] = i2; jxyz[o + 2] = i3;
    rxyz[o] = x; rxyz[o + 1] = y; rxyz[o + 2] = z;
    mps[ia] += 1;
  });
  // set_lma / calc_uv (prep_pp.f90:374-525), one projector per channel
  int lma = 0;
  for (int ia = 0; ia < natom; ++ia) {
    const Element &e = pp->el[kion[ia] - 1];
    for (int l = 0; l <= e.mlps; ++l) {
      if (e.inorm[l] == 0) continue;
      for (int m = -l; m <= l; ++m) {
        ia_tbl[lma] = ia + 1;
        rinv_uvu[lma] = (double)e.inorm[l] * c.hvol;
        for (int j = 0; j < mps[ia]; ++j) {
          const double *p = rxyz + 3 * ((size_t)j + (size_t)nps * ia);
          const double x = p[0], y = p[1], z = p[2];
          const double r = std::sqrt(x * x + y * y + z * z) + 1e-50;
          int ir = 1;
          while (ir <= e.nrps && !(e.rad[ir] > r)) ++ir;
          const int intr = ir - 1;
          if (intr < 1 || intr >= e.nrps)
            return fail(SALMON_HIP_EINVAL, "sphere point %d of atom %d outside the radial table", j + 1, ia + 1);
          const double xx = r - e.rad[intr];
          const int s = intr - 1;
          const double uvr = e.sa[l][s] * (xx * xx * xx) + e.sb[l][s] * (xx * xx) + e.sc[l][s] * xx + e.sd[l][s];
          uv[(size_t)j + (size_t)nps * lma] = uvr * ylm(x, y, z, l, m);
        }
        ++lma;
      }
    }
  }
  return 0;
}

int salmon_pp_vpsl(salmon_pp_t pp, int natom, const int *kion, const double *rion, const double *al,
                   const int *n, double *vpsl) {
  Cell c;
  if (int rc = check_cell(pp, natom, kion, rion, al, n, &c)) return rc;
  if (!vpsl) return fail(SALMON_HIP_EINVAL, "null output");
  typedef std::complex<double> cd;
  const int nx = c.n[0], ny = c.n[1], nz = c.n[2];
  const size_t ng = (size_t)nx * ny * nz;
  const double det = c.al[0] * c.al[1] * c.al[2];
  double brl[3];
  for (int d = 0; d < 3; ++d) brl[d] = 2.0 * kPi / c.al[d];
  auto gidx = [](int i, int N) { return i - 1 - N * (1 + ((i - 1 - (N + 1) / 2) >= 0 ? 1 : -1)) / 2; };
  const int nel = (int)pp->el.size();
  std::vector<cd> vg(ng);
  std::vector<double> zvg(nel);
  for (int iz = 1; iz <= nz; ++iz)
    for (int iy = 1; iy <= ny; ++iy)
      for (int ix = 1; ix <= nx; ++ix) {
        const double g[3] = {gidx(ix, nx) * brl[0], gidx(iy, ny) * brl[1], gidx(iz, nz) * brl[2]};
        const double g2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
        const bool gzero = (ix - 1) * (ix - 1) + (iy - 1) * (iy - 1) + (iz - 1) * (iz - 1) == 0;
        const double coef = gzero ? 0.0 : 4.0 * kPi / g2;
        const double gs = std::sqrt(g2);
        for (int k = 0; k < nel; ++k) {  // zVG_ion (prep_pp.f90:826-846)
          const Element &e = pp->el[k];
          double s = 0.0;
          for (int i = 2; i <= e.nrps; ++i) {
            const double r1 = 0.5 * (e.rad[i] + e.rad[i - 1]);
            const double dr = e.rad[i] - e.rad[i - 1];
            const double va = 0.5 * (e.vloctbl[i] + e.vloctbl[i - 1]);
            if (gzero) s = s + 4.0 * kPi * (r1 * r1 * va + r1 * e.zps) * dr;
            else s = s + 4.0 * kPi * std::sin(gs * r1) / gs * (r1 * va + e.zps) * dr;
          }
          zvg[k] = s;
        }
        cd acc(0.0, 0.0);
        for (int ia = 0; ia < natom; ++ia) {  // V_ion(G) (prep_pp.f90:855-870)
          const Element &e = pp->el[kion[ia] - 1];
          const double gd = g[0] * rion[3 * ia] + g[1] * rion[3 * ia + 1] + g[2] * rion[3 * ia + 2];
          acc = acc + (zvg[kion[ia] - 1] - coef * e.zps) * (cd(std::cos(gd), -std::sin(gd)) / det);
        }
        vg[(size_t)(ix - 1) + (size_t)nx * ((size_t)(iy - 1) + (size_t)ny * (iz - 1))] = acc;
      }
  // inverse DFT with e(k, i) = exp(i 2 pi (i-1)(k-1)/N): z, then y, then x (prep_pp.f90:903-938)
  auto table = [](int N) {
    std::vector<cd> t((size_t)N * N);
    for (int i = 0; i < N; ++i)
      for (int k = 0; k < N; ++k) t[(size_t)k + (size_t)N * i] = std::exp(cd(0.0, 2.0 * kPi * (double)(i * k) / N));
    return t;
  };
  const std::vector<cd> ex = table(nx), ey = table(ny), ez = table(nz);
  std::vector<cd> t1(ng), t2(ng);
  for (int iz = 0; iz < nz; ++iz)
    for (int ky = 0; ky < ny; ++ky)
      for (int kx = 0; kx < nx; ++kx) {
        cd s(0.0, 0.0);
        for (int kz = 0; kz < nz; ++kz) s += ez[kz + (size_t)nz * iz] * vg[kx + (size_t)nx * (ky + (size_t)ny * kz)];
        t1[kx + (size_t)nx * (ky + (size_t)ny * iz)] = s;
      }
  for (int iz = 0; iz < nz; ++iz)
    for (int iy = 0; iy < ny; ++iy)
      for (int kx = 0; kx < nx; ++kx) {
        cd s(0.0, 0.0);
        for (int ky = 0; ky < ny; ++ky) s += ey[ky + (size_t)ny * iy] * t1[kx + (size_t)nx * (ky + (size_t)ny * iz)];
        t2[kx + (size_t)nx * (iy + (size_t)ny * iz)] = s;
      }
  for (int iz = 0; iz < nz; ++iz)
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        cd s(0.0, 0.0);
        for (int kx = 0; kx < nx; ++kx) s += ex[kx + (size_t)nx * ix] * t2[kx + (size_t)nx * (iy + (size_t)ny * iz)];
        vpsl[ix + (size_t)nx * (iy + (size_t)ny * iz)] = s.real();
      }
  return 0;
}

}  // extern "C"
