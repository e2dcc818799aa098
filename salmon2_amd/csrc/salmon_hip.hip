// MI355X (gfx950) real-time Kohn-Sham propagation hot path: device state,
// kernels and the C ABI declared in include/salmon_hip.h.
//
// Kernels (one launch each, all deterministic: no floating-point atomics):
//   k_update_kvector  phase table zekr_uV and per-k stencil terms
//                     (update_kvector_nonlocalpt, src/common/hamiltonian.f90:847-903)
//   k_proj<NW>        batched projector x orbital contraction on FP64 MFMA
//                     (v_mfma_f64_16x16x4f64): uVpsi of zpseudo
//                     (src/common/nonlocal_potential.f90:360-369) for NW = 1 and the
//                     four moments of calc_current_nonlocal (density_matrix.f90:532-591)
//                     for NW = 4
//   k_stencil<HC,NC>  25-point stencil + V_local + k+A gradient
//                     (zstencil_typical_seq, zstencil_core/typical/seq.f90:17-127) fed by
//                     direct global->LDS DMA rings; in the last Taylor term it also forms
//                     the sum of taylor.f90:53-114
//   k_nl_scatter      the KB scatter of zpseudo (nonlocal_potential.f90:370-378) as a
//                     deterministic per-point gather over a point->(atom, j) inverse map
//   k_density_*       calc_density (density_matrix.f90:150-230)
//   k_current_*       stencil_current + calc_current (density_matrix.f90:240-530)
//   k_dipole_part/_sum calc_dip (src/rt/dip.f90:114-225)
//   k_dft_axis        Hartree potential (poisson_ft, src/poisson/poisson_periodic.f90:23-242)
//   k_xc_pz           PZ81 LDA (src/xc/builtin_pz.f90:23-83) and its spin-polarised form
//                     (src/xc/builtin_pz_sp.f90:23-98)
//   k_update_vlocal   update_vlocal (src/common/hamiltonian.f90:813-843) with the
//                     predictor-corrector average (src/rt/time_evolution_step.f90:404)
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "salmon_hip.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(x)                                                                            \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess) return fail(SALMON_HIP_EHIP, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

#define LAUNCHCHK(name)                                                                          \
  do {                                                                                           \
    hipError_t e_ = hipGetLastError();                                                           \
    if (e_ != hipSuccess) return fail(SALMON_HIP_EHIP, "launch %s: %s", name, hipGetErrorString(e_)); \
  } while (0)

constexpr int kBlock = 256;
constexpr int kProjJC = 32;      // sphere points staged per LDS chunk in k_proj
constexpr int kProjOrb = 64;     // orbitals per k_proj block (4 waves x 16)
constexpr int kMaxLmaPerAtom = 16;
constexpr int kDensGroup = 32;   // orbitals per density partial

__device__ __forceinline__ int wrapi(int v, int n) { return (v + 8 * n) % n; }

// Complex element types: double2 (FP64) or float2 (FP32, mixed-precision chain).
// A float2 element is moved by two dword DMAs per 64-point wave chunk (lane i ->
// point 32k + i/2, half i%2), a double2 element by one dwordx4 DMA (lane i -> point i).
template <typename R> struct Cx;
template <> struct Cx<double> { using T = double2; };
template <> struct Cx<float> { using T = float2; };
template <typename E> struct DmaW;
template <> struct DmaW<double2> { static constexpr int W = 1; };
template <> struct DmaW<float2> { static constexpr int W = 2; };

template <typename D, typename S>
__device__ __forceinline__ D cvt2(const S &v) {
  D r;
  r.x = v.x;
  r.y = v.y;
  return r;
}

// ---------------------------------------------------------------------------
// Phase table and per-k stencil coefficients
// ---------------------------------------------------------------------------
struct KvecArgs {
  const double *veck;    // (3, nk)
  const double *rxyz;    // (3, nps, natom)
  const double *uv;      // (nps, nlma)
  const int *mps;        // (natom)
  const int *lma_atom;   // (nlma) 0-based atom of each projector
  const int2 *atom_lma;  // (natom): first ilma, count
  const int *entof;      // (nps, natom): CSR entry of each sphere point
  double2 *zekr;         // (nps, nlma, nk)
  double2 *zent;         // (cmax, nent, nk)
  double *kpar;          // (16, nk): k_lap0, k_nabt(12), kAc(3)
  double coef_nab[12];
  double coef_lap0;
  double ax, ay, az;
  // with actab: A = (actab(:, itt) + actab(:, itt-1)) / 2 at the device step counter *itt
  const double *actab;
  const int *itt;
  int tmax;              // last row of actab (nt + 1); the step counter is clamped to 1..tmax
  int nlma, nps, nk, nent, cmax;
};

// the midpoint field of the current step from the device table, else the by-value field
__device__ __forceinline__ void kvec_field(const KvecArgs &a, double &ax, double &ay, double &az) {
  if (a.actab) {
    const int t = min(max(*a.itt, 1), a.tmax);
    ax = 0.5 * (a.actab[3 * t + 0] + a.actab[3 * (t - 1) + 0]);
    ay = 0.5 * (a.actab[3 * t + 1] + a.actab[3 * (t - 1) + 1]);
    az = 0.5 * (a.actab[3 * t + 2] + a.actab[3 * (t - 1) + 2]);
  } else {
    ax = a.ax;
    ay = a.ay;
    az = a.az;
  }
}

__device__ __forceinline__ void update_kvector_body(const KvecArgs &a, int ilma, int ik) {
  double ax, ay, az;
  kvec_field(a, ax, ay, az);
  const double kx = a.veck[3 * ik + 0] + ax;
  const double ky = a.veck[3 * ik + 1] + ay;
  const double kz = a.veck[3 * ik + 2] + az;
  if (ilma == 0 && threadIdx.x == 0) {
    double *kp = a.kpar + 16 * ik;
    double ss = kx * kx;
    ss = ss + ky * ky;
    ss = ss + kz * kz;
    kp[0] = a.coef_lap0 + 0.5 * ss;
    for (int n = 0; n < 4; ++n) {
      kp[1 + n] = kx * a.coef_nab[n];
      kp[5 + n] = ky * a.coef_nab[4 + n];
      kp[9 + n] = kz * a.coef_nab[8 + n];
    }
    kp[13] = kx;
    kp[14] = ky;
    kp[15] = kz;
  }
  if (ilma >= a.nlma) return;
  const int ia = a.lma_atom[ilma];
  const int m = a.mps[ia];
  for (int j = threadIdx.x; j < m; j += blockDim.x) {
    const double *r = a.rxyz + 3 * ((size_t)j + (size_t)a.nps * ia);
    const double th = kx * r[0] + ky * r[1] + kz * r[2];
    double s, c;
    sincos(th, &s, &c);
    const double u = a.uv[(size_t)j + (size_t)a.nps * ilma];
    const double2 z = make_double2(c * u, -s * u);
    a.zekr[(size_t)j + (size_t)a.nps * ((size_t)ilma + (size_t)a.nlma * ik)] = z;
    const int e = a.entof[(size_t)j + (size_t)a.nps * ia];
    a.zent[(size_t)(ilma - a.atom_lma[ia].x) + (size_t)a.cmax * ((size_t)e + (size_t)a.nent * ik)] = z;
  }
}

__global__ __launch_bounds__(kBlock) void k_update_kvector(KvecArgs a) { update_kvector_body(a, blockIdx.x, blockIdx.y); }

// Per-k stencil terms only (hpsi reads system%vec_Ac at every call,
// src/common/hamiltonian.f90:146-155; the phase table is left untouched).
__global__ void k_set_kpar(KvecArgs a) {
  const int ik = threadIdx.x;
  if (ik >= a.nk) return;
  double ax, ay, az;
  kvec_field(a, ax, ay, az);
  const double kx = a.veck[3 * ik + 0] + ax;
  const double ky = a.veck[3 * ik + 1] + ay;
  const double kz = a.veck[3 * ik + 2] + az;
  double *kp = a.kpar + 16 * ik;
  double ss = kx * kx;
  ss = ss + ky * ky;
  ss = ss + kz * kz;
  kp[0] = a.coef_lap0 + 0.5 * ss;
  for (int n = 0; n < 4; ++n) {
    kp[1 + n] = kx * a.coef_nab[n];
    kp[5 + n] = ky * a.coef_nab[4 + n];
    kp[9 + n] = kz * a.coef_nab[8 + n];
  }
  kp[13] = kx;
  kp[14] = ky;
  kp[15] = kz;
}

// ---------------------------------------------------------------------------
// Projector x orbital contraction on FP64 MFMA (v_mfma_f64_16x16x4f64).
// D[n][o] = sum_K A[n][K] B[K][o] with K = (sphere point j, re/im),
// o = 16 orbitals per wave, n = (lma-weight index lw, re/im part).
//   part re:  Re(conj(Z) r_w psi) = Zr r_w psi_re + Zi r_w psi_im
//   part im:  Im(conj(Z) r_w psi) = Zr r_w psi_im - Zi r_w psi_re
// r_0 = 1 (plain projection), r_1..3 = rxyz components (current moments).
// Operand layout: lane L holds A[L%16][L/16] and B[L/16][L%16]; the result
// lands as D[L/16 + 4r][L%16] in register r.  Rows are therefore assigned as
// n = 8h + 4p + q (q = 0..3, p = part, h = half) <-> lw = 8t + 4h + q, so the
// re and im rows of one lw sit in registers (2h, 2h+1) of the same lane.
// ---------------------------------------------------------------------------
struct ProjArgs {
  const void *psi;        // orbital block base (double2, or float2 in the mixed-precision chain)
  const double2 *zekr;    // (nps, nlma, nk)
  const double *rxyz;     // (3, nps, natom)
  const int *jidx;        // (nps, natom) offsets within an orbital block
  const int *mps;         // (natom)
  const int2 *atom_lma;   // (natom): first ilma, count
  const double *rinv;     // (nlma)
  double2 *out;           // NW=1: c(nb, nlma, nk) scaled by rinv; NW=4, 5: (4, nlma, nb, nk) raw
  size_t osz;
  int nb, nk, natom, nps, nlma, ngrp;
  const int *skip;        // device flag: the launch does nothing when it is set (null: never)
  // NW=5: weight 4 is the plain projection under a second phase table (the next step's), written
  // as c(nb, nlma, nk) scaled by rinv
  const double2 *zekr2;
  double2 *out2;
};

// XCD-aware order of a launch of nblk blocks (block b runs on XCD b % 8): the blocks of one XCD
// take a contiguous range of the work items, so the cache lines that neighbouring items share
// are read through one L2
__device__ __forceinline__ int xcd_order(int b, int nblk) {
  const int nfull = (nblk / 8) * 8;
  return b < nfull ? (b % 8) * (nblk / 8) + b / 8 : b;
}

// LMX bounds the projectors of one atom (the tile and staging sizes)
template <int NW, typename TP, int LMX>
__device__ __forceinline__ void proj_body(const ProjArgs &a, int bid) {
  static_assert(LMX <= kMaxLmaPerAtom, "LMX bounds the projectors of one atom");
  constexpr int MAXT = (NW * LMX + 7) / 8;
  __shared__ double2 sp[kProjJC][kProjOrb];
  __shared__ double2 zsh[kProjJC][LMX];
  __shared__ double2 zsh2[NW == 5 ? kProjJC : 1][LMX];
  __shared__ double rsh[kProjJC][3];
  const int grp = bid % a.ngrp;
  const int t = bid / a.ngrp;
  const int ia = t % a.natom;
  const int ik = t / a.natom;
  const int2 al = a.atom_lma[ia];
  const int m = a.mps[ia];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int nlw = NW * al.y;            // (lma, weight) pairs of this atom
  const int ntile = (nlw + 7) >> 3;     // 8 lw (16 rows) per MFMA tile
  const size_t kbase = (size_t)ik * a.nb;

  typedef double __attribute__((ext_vector_type(4))) d4;
  d4 acc[MAXT];
#pragma unroll
  for (int tt = 0; tt < MAXT; ++tt) acc[tt] = d4{0.0, 0.0, 0.0, 0.0};

  const int kq = lane >> 4;      // K row of this lane (0..3)
  const int comp = kq & 1;       // 0: re, 1: im of psi
  const int oc = lane & 15;      // orbital column (B) / output row (A)

  // psi at kProjJC sphere points for the block's 64 orbitals: chunk j0 + kProjJC is loaded into
  // registers while the MFMAs consume chunk j0 from LDS
  constexpr int kPer = (kProjJC * kProjOrb) / kBlock;
  constexpr int kZPer = (kProjJC * LMX + kBlock - 1) / kBlock;
  TP pre[kPer];
  double2 zpre[kZPer], zpre2[NW == 5 ? kZPer : 1];
  double rpre = 0.0;
  const TP *psi = (const TP *)a.psi;
  const int *jx = a.jidx + (size_t)a.nps * ia;
  const double2 *zk = a.zekr + (size_t)a.nps * ((size_t)al.x + (size_t)a.nlma * ik);
  const double2 *zk2 = NW == 5 ? a.zekr2 + (size_t)a.nps * ((size_t)al.x + (size_t)a.nlma * ik) : nullptr;
  // every thread loads the same sphere point of each chunk (kBlock is a multiple of kProjJC);
  // its grid index is fetched one chunk ahead, so the psi gathers wait on one load, not two
  static_assert(kBlock % kProjJC == 0, "one sphere point per thread and chunk");
  const int jt = threadIdx.x % kProjJC;
  int jnext = jt < m ? jx[jt] : 0;
  auto load_chunk = [&](int j0) {
    const int jv = jnext;
    jnext = j0 + kProjJC + jt < m ? jx[j0 + kProjJC + jt] : 0;
#pragma unroll
    for (int r = 0; r < kZPer; ++r) {
      const int e = threadIdx.x + kBlock * r;  // (l, jj) = (e / kProjJC, e % kProjJC)
      const int jj = e % kProjJC, l = e / kProjJC;
      const int j = j0 + jj;
      zpre[r] = (j < m && l < al.y) ? zk[(size_t)j + (size_t)a.nps * l] : make_double2(0.0, 0.0);
      if constexpr (NW == 5)
        zpre2[r] = (j < m && l < al.y) ? zk2[(size_t)j + (size_t)a.nps * l] : make_double2(0.0, 0.0);
    }
    if (NW > 1 && threadIdx.x < 3 * kProjJC) {
      const int j = j0 + threadIdx.x / 3;
      rpre = j < m ? a.rxyz[3 * ((size_t)j0 + (size_t)a.nps * ia) + threadIdx.x] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const int e = threadIdx.x + kBlock * r;
      const int jj = e % kProjJC;
      const int o = e / kProjJC;
      const int j = j0 + jj;
      const int ob = grp * kProjOrb + o;
      TP v;
      v.x = 0;
      v.y = 0;
      if (j < m && ob < a.nb) v = psi[(kbase + ob) * a.osz + jv];
      pre[r] = v;
    }
  };
  load_chunk(0);
  for (int j0 = 0; j0 < m; j0 += kProjJC) {
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const int e = threadIdx.x + kBlock * r;
      sp[e % kProjJC][e / kProjJC] = cvt2<double2>(pre[r]);
    }
#pragma unroll
    for (int r = 0; r < kZPer; ++r) {
      const int e = threadIdx.x + kBlock * r;
      if (e < kProjJC * LMX) zsh[e % kProjJC][e / kProjJC] = zpre[r];
      if constexpr (NW == 5)
        if (e < kProjJC * LMX) zsh2[e % kProjJC][e / kProjJC] = zpre2[r];
    }
    if (NW > 1 && threadIdx.x < 3 * kProjJC) rsh[threadIdx.x / 3][threadIdx.x % 3] = rpre;
    __syncthreads();
    if (j0 + kProjJC < m) load_chunk(j0 + kProjJC);
#pragma unroll 2
    for (int ks = 0; ks < kProjJC / 2; ++ks) {
      const int jp = 2 * ks + (kq >> 1);
      const int j = j0 + jp;
      const double2 pv = sp[jp][wave * 16 + oc];
      const double bval = comp ? pv.y : pv.x;
      double zr = 0.0, zi = 0.0, rw[3] = {0.0, 0.0, 0.0};
      const bool jv = j < m;
      if (jv && NW > 1) {
        rw[0] = rsh[jp][0];
        rw[1] = rsh[jp][1];
        rw[2] = rsh[jp][2];
      }
#pragma unroll
      for (int tt = 0; tt < MAXT; ++tt) {
        if (tt < ntile) {
          const int part = (oc >> 2) & 1;
          const int lw = tt * 8 + (oc >> 3) * 4 + (oc & 3);
          double aval = 0.0;
          if (jv && lw < nlw) {
            const int w = lw % NW;
            const int l = lw / NW;
            const double2 z = (NW == 5 && w == 4) ? zsh2[jp][l] : zsh[jp][l];
            const double r = (w == 0 || w == 4) ? 1.0 : rw[w - 1];
            zr = z.x * r;
            zi = z.y * r;
            aval = part == 0 ? (comp ? zi : zr) : (comp ? zr : -zi);
          }
          acc[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(aval, bval, acc[tt], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  // lane (kq, oc) holds rows kq + 4r: r = 2h -> re of lw = 8t + 4h + kq, r = 2h+1 -> its im
  const int ob = grp * kProjOrb + wave * 16 + oc;
  if (ob >= a.nb) return;
#pragma unroll
  for (int tt = 0; tt < MAXT; ++tt) {
    if (tt >= ntile) break;
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
      const int lw = tt * 8 + hh * 4 + kq;
      if (lw >= nlw) continue;
      const int w = lw % NW;
      const int l = lw / NW;
      const int ilma = al.x + l;
      const double re = acc[tt][2 * hh], im = acc[tt][2 * hh + 1];
      if (NW == 1) {
        const double s = a.rinv[ilma];
        a.out[((size_t)ik * a.nlma + ilma) * a.nb + ob] = make_double2(re * s, im * s);
      } else if (NW == 5 && w == 4) {
        const double s = a.rinv[ilma];
        a.out2[((size_t)ik * a.nlma + ilma) * a.nb + ob] = make_double2(re * s, im * s);
      } else {
        a.out[((kbase + ob) * a.nlma + ilma) * 4 + w] = make_double2(re, im);
      }
    }
  }
}

template <int NW, typename TP, int LMX>
__global__ __launch_bounds__(kBlock) void k_proj(ProjArgs a) {
  if (a.skip && *a.skip) return;  // a device-side gate switched this launch off
  proj_body<NW, TP, LMX>(a, xcd_order(blockIdx.x, gridDim.x));
}

// ---------------------------------------------------------------------------
// KB term of one H application fused with the projection of its result: for a Taylor term
// h = H h_prev the pass adds s * sum_lma c_prev(lma) zekr(j, lma) at the sphere points that one
// (atom, point) pair covers alone and projects the completed h onto the atom's projectors (the
// c of the next term), reading h once.  Block = (k, atom, 64 orbitals) as in k_proj<1>: each
// thread owns one sphere point of every 32-point chunk and 8 of the orbitals; after the chunk is
// staged in LDS it completes its points (exclusive ones: KB sum in ascending ilma, in FP64, then
// rounded to TP as k_nl_scatter stores it), writes them back and puts the rounded values into the
// LDS tile the MFMAs read.  Points covered by several pairs are completed beforehand by
// k_nl_scatter over their list, and no point is written by two blocks (an exclusive point
// belongs to one atom, and the blocks of one atom split the orbitals), so the pass is
// deterministic and race-free.  c_prev for the block's orbitals is staged in dynamic LDS
// (cmax x 64 elements).
// ---------------------------------------------------------------------------
struct ProjScatArgs {
  ProjArgs p;              // projection of h (p.psi = h, p.out = its c, scaled by rinv)
  const double2 *cin;      // c of h_prev (nb, nlma, nk), scaled by rinv
  const int *jexcl;        // (nps, natom): offset within an orbital block if one pair covers the point, else -1
  double2 s;               // h += s * sum (1 for the terms without the Taylor sum)
  int cmax;                // projectors per atom, at most (the row count of the staged c)
};

template <typename TP, int LM>
__device__ __forceinline__ void proj_scatter_body(const ProjScatArgs &A, int bid) {
  static_assert(LM <= kMaxLmaPerAtom, "LM bounds the projectors of one atom");
  const ProjArgs &a = A.p;
  constexpr int MAXT = (LM + 7) / 8;
  // rows padded by one element: the per-point accesses of the completion step (lanes on points)
  // then fall on different banks
  __shared__ double2 sp[kProjJC][kProjOrb + 1];
  __shared__ double2 zsh[kProjJC][LM + 1];
  extern __shared__ double2 csh[];  // cmax x kProjOrb
  const int grp = bid % a.ngrp;
  const int t = bid / a.ngrp;
  const int ia = t % a.natom;
  const int ik = t / a.natom;
  const int2 al = a.atom_lma[ia];
  const int m = a.mps[ia];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int nlw = al.y;
  const int ntile = (nlw + 7) >> 3;
  const size_t kbase = (size_t)ik * a.nb;

  for (int e = threadIdx.x; e < al.y * kProjOrb; e += kBlock) {
    const int l = e / kProjOrb, o = e % kProjOrb, ob = grp * kProjOrb + o;
    csh[e] = ob < a.nb ? A.cin[((size_t)ik * a.nlma + al.x + l) * a.nb + ob] : make_double2(0.0, 0.0);
  }
  __syncthreads();

  typedef double __attribute__((ext_vector_type(4))) d4;
  d4 acc[MAXT];
#pragma unroll
  for (int tt = 0; tt < MAXT; ++tt) acc[tt] = d4{0.0, 0.0, 0.0, 0.0};
  const int kq = lane >> 4;
  const int comp = kq & 1;
  const int oc = lane & 15;

  constexpr int kPer = (kProjJC * kProjOrb) / kBlock;
  constexpr int kZPer = (kProjJC * LM + kBlock - 1) / kBlock;
  TP pre[kPer];
  double2 zpre[kZPer];
  TP *h = (TP *)a.psi;
  const int *jx = a.jidx + (size_t)a.nps * ia;
  const int *jxe = A.jexcl + (size_t)a.nps * ia;
  const double2 *zk = a.zekr + (size_t)a.nps * ((size_t)al.x + (size_t)a.nlma * ik);
  static_assert(kBlock % kProjJC == 0, "one sphere point per thread and chunk");
  const int jt = threadIdx.x % kProjJC;
  const int o0 = threadIdx.x / kProjJC;  // this thread's orbitals: o0 + (kBlock / kProjJC) r
  int jnext = jt < m ? jx[jt] : 0;
  int xnext = jt < m ? jxe[jt] : -1;
  int jload = 0, xload = -1;  // point offset and exclusive offset of the chunk last loaded
  double2 zown[LM];           // zekr of this thread's point of that chunk (exclusive points only)
  auto load_chunk = [&](int j0) {
    jload = jnext;
    xload = xnext;
#pragma unroll
    for (int l = 0; l < LM; ++l)
      zown[l] = (xload >= 0 && l < al.y) ? zk[(size_t)(j0 + jt) + (size_t)a.nps * l] : make_double2(0.0, 0.0);
    const bool more = j0 + kProjJC + jt < m;
    jnext = more ? jx[j0 + kProjJC + jt] : 0;
    xnext = more ? jxe[j0 + kProjJC + jt] : -1;
#pragma unroll
    for (int r = 0; r < kZPer; ++r) {
      const int e = threadIdx.x + kBlock * r;
      const int jj = e % kProjJC, l = e / kProjJC;
      const int j = j0 + jj;
      zpre[r] = (j < m && l < al.y) ? zk[(size_t)j + (size_t)a.nps * l] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const int ob = grp * kProjOrb + o0 + (kBlock / kProjJC) * r;
      TP v;
      v.x = 0;
      v.y = 0;
      if (j0 + jt < m && ob < a.nb) v = h[(kbase + ob) * a.osz + jload];
      pre[r] = v;
    }
  };
  load_chunk(0);
  for (int j0 = 0; j0 < m; j0 += kProjJC) {
    // complete this thread's point if one pair covers it (xload >= 0 implies j0 + jt < m) while
    // staging the chunk
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const int o = o0 + (kBlock / kProjJC) * r;
      const int ob = grp * kProjOrb + o;
      double2 v = cvt2<double2>(pre[r]);
      if (xload >= 0 && ob < a.nb) {
        double2 sum = make_double2(0.0, 0.0);
#pragma unroll
        for (int l = 0; l < LM; ++l) {
          if (l < al.y) {
            const double2 cv = csh[l * kProjOrb + o];
            sum.x += cv.x * zown[l].x - cv.y * zown[l].y;
            sum.y += cv.x * zown[l].y + cv.y * zown[l].x;
          }
        }
        v.x += A.s.x * sum.x - A.s.y * sum.y;
        v.y += A.s.x * sum.y + A.s.y * sum.x;
        const TP vt = cvt2<TP>(v);
        h[(kbase + ob) * a.osz + jload] = vt;
        v = cvt2<double2>(vt);
      }
      sp[jt][o] = v;
    }
#pragma unroll
    for (int r = 0; r < kZPer; ++r) {
      const int e = threadIdx.x + kBlock * r;
      if (e < kProjJC * LM) zsh[e % kProjJC][e / kProjJC] = zpre[r];
    }
    __syncthreads();
    if (j0 + kProjJC < m) load_chunk(j0 + kProjJC);
#pragma unroll 2
    for (int ks = 0; ks < kProjJC / 2; ++ks) {
      const int jp = 2 * ks + (kq >> 1);
      const int j = j0 + jp;
      const double2 pv = sp[jp][wave * 16 + oc];
      const double bval = comp ? pv.y : pv.x;
      const bool jv = j < m;
#pragma unroll
      for (int tt = 0; tt < MAXT; ++tt) {
        if (tt < ntile) {
          const int part = (oc >> 2) & 1;
          const int lw = tt * 8 + (oc >> 3) * 4 + (oc & 3);
          double aval = 0.0;
          if (jv && lw < nlw) {
            const double2 z = zsh[jp][lw];
            aval = part == 0 ? (comp ? z.y : z.x) : (comp ? z.x : -z.y);
          }
          acc[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(aval, bval, acc[tt], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  const int ob = grp * kProjOrb + wave * 16 + oc;
  if (ob >= a.nb) return;
#pragma unroll
  for (int tt = 0; tt < MAXT; ++tt) {
    if (tt >= ntile) break;
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
      const int lw = tt * 8 + hh * 4 + kq;
      if (lw >= nlw) continue;
      const int ilma = al.x + lw;
      const double re = acc[tt][2 * hh], im = acc[tt][2 * hh + 1];
      const double sc = a.rinv[ilma];
      a.out[((size_t)ik * a.nlma + ilma) * a.nb + ob] = make_double2(re * sc, im * sc);
    }
  }
}

template <typename TP, int LM>
__global__ __launch_bounds__(kBlock) void k_proj_scatter(ProjScatArgs A) {
  if (A.p.skip && *A.p.skip) return;  // a device-side gate switched this launch off
  proj_scatter_body<TP, LM>(A, xcd_order(blockIdx.x, gridDim.x));
}

constexpr int kMaxComb = 5;   // Taylor order up to 6

// ---------------------------------------------------------------------------
// Local hpsi: 25-point stencil + V + k.grad (zstencil_typical_seq,
// zstencil_core/typical/seq.f90:17-127), with the Taylor combination folded
// into the last term.  The KB term is added afterwards by k_nl_scatter.
//
// Block = full x rows (nx) x TY rows in y (one point per thread), streaming z
// over [z0, z0+TZ).  Every per-plane input arrives by direct global->LDS DMA
// (global_load_lds, issued through inline asm so that only explicit
// s_waitcnt vmcnt(N) waits gate it) kDmaDepth planes ahead:
//   tile ring  (kRingT slots): psi rows of planes z .. z+4+D  (x/y neighbours, queue head)
//   halo ring  (D+1 slots):    the 4+4 y-halo rows of planes z .. z+D
//   V ring     (D+1 slots):    V_local of planes z .. z+D
//   comb ring  (2 slots, last term only): psi, h1 .. h_{N-2} of planes z, z+1
// z neighbours come from a 9-deep register queue fed from the tile ring.  The
// only compiler-visible global memory operation in the loop is the store, so
// the in-order vmcnt counter never forces an early wait on the ring.
// NC < 0:  hout = H_loc in
// NC >= 0: hout = (((psi + c1 h1) + ... + c_{N-1} in) + c_N H_loc in), N = NC + 1
// ---------------------------------------------------------------------------
constexpr int kDmaDepth = 2;
// The per-plane wait lets the DMAs issued in the previous iteration stay in flight: the data a
// plane needs must have been issued at least two iterations earlier.
static_assert(kDmaDepth >= 2, "the ring wait assumes a look-ahead of at least two planes");
constexpr int kRingT = 5 + kDmaDepth;
constexpr int kRingH = kDmaDepth + 1;
constexpr int kTile = 256;  // tile slot size (elements), = 4 waves x 64 lanes

typedef __attribute__((address_space(3))) void lds_void;

__device__ __forceinline__ unsigned lds_addr(const void *p) {
  return (unsigned)(uintptr_t)((lds_void *)p);
}
// lane i of the wave writes its element to LDS byte address m0 + i * size
// (M0 is not used by compiler-generated gfx950 code in these kernels: LDS
// instructions do not read it)
__device__ __forceinline__ void dma16(const void *g, unsigned m0) {
  asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(m0), "v"(g) : "memory");
}
__device__ __forceinline__ void dma4(const void *g, unsigned m0) {
  asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dword %1, off" ::"s"(m0), "v"(g) : "memory");
}
template <int N>
__device__ __forceinline__ void wait_vm() {
  asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}


// One element chunk of 64 points (this wave's) into LDS at byte address m0:
// W = 1: off[0] = the lane's point; W = 2: off[k] = point 32k + lane/2.
template <typename E>
__device__ __forceinline__ void dma_chunk(const E *base, const size_t *off, unsigned m0, int lane) {
  if constexpr (DmaW<E>::W == 1) {
    dma16(base + off[0], m0);
  } else {
    dma4((const float *)(base + off[0]) + (lane & 1), m0);
    dma4((const float *)(base + off[1]) + (lane & 1), m0 + 256u);
  }
}

struct StencilArgs {
  const void *in;
  void *hout;
  double2 *sink;          // one element per thread absorbing the stores of lanes without a point
  const void *comb[kMaxComb];  // comb[0] double2 (psi); comb[n >= 1] the chain type
  double2 zc[kMaxComb + 1];
  const double *V;        // (nx, ny, nz, nspin)
  const double *kpar;     // (16, nk)
  double lap[12];         // coef_lap(4,3)
  size_t osz;
  int nx, ny, nz, ly, nspin, nb, nk;
  int TY, TZ, nyt, nzt, nblk;
  const int *skip;        // device flag: the launch does nothing when it is set (null: never)
};

// TI: element type of `in`; R: arithmetic of the stencil and local-potential
// products; TO: element type of `hout`.  Plain FP64: (double2, double, double2).
// Mixed form: (double2, float, float2) for the first term, (float2, float, float2)
// for the middle ones and (float2, float, double2) for the last, whose Taylor sum
// is formed in FP64 from psi (FP64) and h_1 .. h_N (FP32).  (FP64 arithmetic in the
// first term, rounded once at the store, measured 10% slower at 32^3.)
// One tile (orbital block, y-tile, z-chunk) = `bid` of the system described by `a`.
template <typename TI, typename R, typename TO, int HC, int NC, int BS>
__device__ __forceinline__ void stencil_tile(const StencilArgs &a, int bid, double2 *lds) {
  static_assert(HC == 0 || BS == kTile, "y-halo tiles are 4 waves wide");
  using CH = typename Cx<R>::T;
  constexpr bool FINAL = NC >= 0;
  constexpr int D = kDmaDepth;
  constexpr int WI = DmaW<TI>::W;
  constexpr int WC = DmaW<CH>::W;
  constexpr unsigned EB = sizeof(TI);
  // DMA instructions per wave per plane, in issue order: comb, tile (WI), halo (HC*WI), V (2 x dword);
  // then 1 store
  constexpr int kOpsAfterComb = WI + HC * WI + 2 + 1;
  TI *tileR = (TI *)lds;                                  // kRingT x BS
  TI *haloR = tileR + kRingT * BS;                     // kRingH x HC*BS
  double *vR = (double *)(haloR + kRingH * HC * BS);   // kRingH x BS doubles
  char *combR = (char *)(vR + kRingH * BS);            // 2 x NC x BS x 16 bytes

  const int ytile = bid % a.nyt;
  const int zt = (bid / a.nyt) % a.nzt;
  const int b = bid / (a.nyt * a.nzt);
  const int ik = b / a.nb;
  const int ispin = b % a.nspin;

  const int nx = a.nx, ny = a.ny, nz = a.nz, ly = a.ly;
  const int TY = a.TY;
  const int y0 = ytile * TY;
  // With one z chunk the column is a full period and may start anywhere: y-tile t starts 4t planes
  // earlier, so that its upper y-halo rows of plane p are fetched in the same iteration as the
  // next tile's own rows of p (4 planes ahead of their use) and come from the L2.  Planes are
  // counted unwrapped in [z0, z1) and wrapped on every access.
  const int zph = (HC > 0 && a.nzt == 1) ? (nz - (4 * ytile) % nz) % nz : 0;
  const int z0 = zt * a.TZ + zph;
  const int z1 = min(zt * a.TZ + a.TZ, nz) + zph;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int ntile = TY * nx;
  const bool active = tid < ntile;
  const int te = active ? tid : 0;          // element this lane computes (dummy lanes repeat element 0)
  const int tx = te % nx;
  const int ty = te / nx;
  const int y = y0 + ty;
  const bool outY = active && y < ny;
  const int yw = wrapi(y, ny);

  const size_t zs = (size_t)nx * ly;
  const size_t orbOff = (size_t)b * a.osz;
  const TI *in = (const TI *)a.in + orbOff;
  const size_t colOff = (size_t)tx + (size_t)nx * yw;
  const size_t gplane = (size_t)nx * ny;
  // tile points this lane moves (W = 1: its own; W = 2: points 32k + lane/2 of the wave's chunk)
  auto tile_off = [&](int W, size_t *off) {
    if (W == 1) {
      off[0] = colOff;
      return;
    }
    for (int k = 0; k < 2; ++k) {
      const int e = 64 * wave + 32 * k + (lane >> 1);
      const int ec = e < ntile ? e : 0;
      off[k] = (size_t)(ec % nx) + (size_t)nx * wrapi(y0 + ec / nx, ny);
    }
  };
  size_t toffI[2], toffC[2];
  tile_off(WI, toffI);
  tile_off(WC, toffC);
  // V: two dword DMAs per wave, each covering 32 points of the wave's 64 (lane i -> point i/2, half i%2)
  const double *Vb = a.V + (size_t)ispin * gplane * nz;
  size_t voff[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int e = 64 * wave + 32 * k + (lane >> 1);
    const int ec = e < ntile ? e : 0;
    voff[k] = 2 * ((size_t)(ec % nx) + (size_t)nx * wrapi(y0 + ec / nx, ny)) + (lane & 1);
  }

  // halo points of this lane: chunk c = wave + 4k, points 64c + (lane | 32kk + lane/2)
  size_t hoff[HC > 0 ? HC : 1][2];
#pragma unroll
  for (int k = 0; k < HC; ++k)
#pragma unroll
    for (int kk = 0; kk < WI; ++kk) {
      const int e = 64 * (wave + 4 * k) + (WI == 1 ? lane : 32 * kk + (lane >> 1));
      const int ec = e < 8 * nx ? e : 0;
      const int hr = ec / nx, hx = ec % nx;
      const int gy = hr < 4 ? y0 - 4 + hr : y0 + TY + (hr - 4);
      hoff[k][kk] = (size_t)hx + (size_t)nx * wrapi(gy, ny);
    }
  // LDS byte addresses of this wave's chunk in each ring (uniform per wave)
  const unsigned tileA = __builtin_amdgcn_readfirstlane(lds_addr(tileR) + EB * 64u * wave);
  const unsigned haloA = __builtin_amdgcn_readfirstlane(lds_addr(haloR) + EB * 64u * wave);
  const unsigned vA = __builtin_amdgcn_readfirstlane(lds_addr(vR) + 8u * 64u * wave);
  const unsigned combA = __builtin_amdgcn_readfirstlane(lds_addr(combR) + 16u * 64u * wave);
  const unsigned combAC = __builtin_amdgcn_readfirstlane(lds_addr(combR) + (unsigned)sizeof(CH) * 64u * wave);

  auto dma_tile = [&](int plane, int slot) {
    dma_chunk<TI>(in + zs * wrapi(plane, nz), toffI, tileA + EB * BS * slot, lane);
  };
  auto dma_halo = [&](int plane, int slot) {
#pragma unroll
    for (int k = 0; k < HC; ++k)
      dma_chunk<TI>(in + zs * wrapi(plane, nz), hoff[k], haloA + EB * (HC * BS * slot + 4 * 64 * k), lane);
  };
  auto dma_v = [&](int plane, int slot) {
    const float *vb = (const float *)Vb + 2 * gplane * wrapi(plane, nz);
    dma4(vb + voff[0], vA + 8u * BS * slot);
    dma4(vb + voff[1], vA + 8u * BS * slot + 256u);
  };
  auto dma_comb = [&](int plane, int slot) {
    const size_t pl = orbOff + zs * wrapi(plane, nz);
#pragma unroll
    for (int n = 0; n < (NC > 0 ? NC : 1); ++n) {
      if (n >= NC) break;
      // every comb array owns BS x 16 bytes per slot; psi is double2, the chain arrays CH
      if (n == 0) {
        size_t o0[1] = {colOff};
        dma_chunk<double2>((const double2 *)a.comb[0] + pl, o0, combA + 16u * BS * (NC * slot), lane);
      } else {
        dma_chunk<CH>((const CH *)a.comb[n] + pl, toffC, combAC + 16u * BS * (NC * slot + n), lane);
      }
    }
  };

  // y-neighbour LDS offsets: rows ty+n (n = 1..4, -1..-4) in the tile slot or the halo slot
  int yoff[8];
  bool yhal[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int n = i < 4 ? i + 1 : -(i - 3);
    const int r = ty + n;
    if (HC == 0) {  // the tile is the whole xy-plane (TY == ny): y wraps inside it
      yoff[i] = (r < 0 ? r + TY : (r >= TY ? r - TY : r)) * nx + tx;
      yhal[i] = false;
    } else if (r >= 0 && r < TY) {
      yoff[i] = r * nx + tx;
      yhal[i] = false;
    } else {
      yoff[i] = (r < 0 ? r + 4 : 4 + r - TY) * nx + tx;
      yhal[i] = true;
    }
  }
  int xo[8];
#pragma unroll
  for (int n = 1; n <= 4; ++n) {
    xo[n - 1] = ty * nx + wrapi(tx + n, nx);
    xo[3 + n] = ty * nx + wrapi(tx - n, nx);
  }
  const double *kp = a.kpar + 16 * ik;
  const R klap0 = (R)kp[0];
  R nab[12], lap[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    nab[i] = (R)kp[1 + i];
    lap[i] = (R)a.lap[i];
  }

  // prologue: queue planes z0-4 .. z0+3 (plain loads), rings for planes z0 .. z0+3+D
  CH q[9];
#pragma unroll
  for (int i = 0; i < 8; ++i) q[i] = cvt2<CH>(in[colOff + zs * wrapi(z0 - 4 + i, nz)]);
  __builtin_amdgcn_s_waitcnt(0);  // compiler-visible: the queue loads are complete
  if (FINAL) dma_comb(z0, z0 & 1);
  for (int p = 0; p < 4 + D; ++p) dma_tile(z0 + p, (z0 + p) % kRingT);
  for (int p = 0; p < D; ++p) {
    dma_halo(z0 + p, (z0 + p) % kRingH);
    dma_v(z0 + p, (z0 + p) % kRingH);
  }
  wait_vm<0>();

  int sT = z0 % kRingT, sH = z0 % kRingH;  // ring slots of plane z
  for (int z = z0; z < z1; ++z) {
    // planes z (halo, V, comb) and z+4 (tile) were issued >= one full plane ago
    wait_vm<kOpsAfterComb>();
    __syncthreads();
    const int sTn = (sT + 4 + D) % kRingT;  // slot of plane z+4+D (= slot of z-1)
    const int sHn = (sH + D) % kRingH;      // slot of plane z+D   (= slot of z-1)
    // past the chunk's last needed plane the look-ahead re-reads that plane (an L2 hit) into the
    // slot, which is never read: every wave still issues the same memory operations per plane
    if (FINAL) dma_comb(min(z + 1, z1 - 1), (z + 1) & 1);
    dma_tile(min(z + 4 + D, z1 + 3), sTn);
    dma_halo(min(z + D, z1 - 1), sHn);
    dma_v(min(z + D, z1 - 1), sHn);

    const TI *T = tileR + BS * sT;
    const TI *H = haloR + HC * BS * sH;
    q[8] = cvt2<CH>(tileR[BS * ((sT + 4) % kRingT) + te]);
    CH v, w;
    {
      const CH p1 = cvt2<CH>(T[xo[0]]), p2 = cvt2<CH>(T[xo[1]]), p3 = cvt2<CH>(T[xo[2]]), p4 = cvt2<CH>(T[xo[3]]);
      const CH m1 = cvt2<CH>(T[xo[4]]), m2 = cvt2<CH>(T[xo[5]]), m3 = cvt2<CH>(T[xo[6]]), m4 = cvt2<CH>(T[xo[7]]);
      v.x = lap[0] * (p1.x + m1.x) + lap[1] * (p2.x + m2.x) + lap[2] * (p3.x + m3.x) + lap[3] * (p4.x + m4.x);
      v.y = lap[0] * (p1.y + m1.y) + lap[1] * (p2.y + m2.y) + lap[2] * (p3.y + m3.y) + lap[3] * (p4.y + m4.y);
      w.x = nab[0] * (p1.x - m1.x) + nab[1] * (p2.x - m2.x) + nab[2] * (p3.x - m3.x) + nab[3] * (p4.x - m4.x);
      w.y = nab[0] * (p1.y - m1.y) + nab[1] * (p2.y - m2.y) + nab[2] * (p3.y - m3.y) + nab[3] * (p4.y - m4.y);
    }
    {
      const CH p1 = cvt2<CH>((yhal[0] ? H : T)[yoff[0]]), p2 = cvt2<CH>((yhal[1] ? H : T)[yoff[1]]);
      const CH p3 = cvt2<CH>((yhal[2] ? H : T)[yoff[2]]), p4 = cvt2<CH>((yhal[3] ? H : T)[yoff[3]]);
      const CH m1 = cvt2<CH>((yhal[4] ? H : T)[yoff[4]]), m2 = cvt2<CH>((yhal[5] ? H : T)[yoff[5]]);
      const CH m3 = cvt2<CH>((yhal[6] ? H : T)[yoff[6]]), m4 = cvt2<CH>((yhal[7] ? H : T)[yoff[7]]);
      v.x += lap[4] * (p1.x + m1.x) + lap[5] * (p2.x + m2.x) + lap[6] * (p3.x + m3.x) + lap[7] * (p4.x + m4.x);
      v.y += lap[4] * (p1.y + m1.y) + lap[5] * (p2.y + m2.y) + lap[6] * (p3.y + m3.y) + lap[7] * (p4.y + m4.y);
      w.x += nab[4] * (p1.x - m1.x) + nab[5] * (p2.x - m2.x) + nab[6] * (p3.x - m3.x) + nab[7] * (p4.x - m4.x);
      w.y += nab[4] * (p1.y - m1.y) + nab[5] * (p2.y - m2.y) + nab[6] * (p3.y - m3.y) + nab[7] * (p4.y - m4.y);
    }
    v.x += lap[8] * (q[5].x + q[3].x) + lap[9] * (q[6].x + q[2].x) + lap[10] * (q[7].x + q[1].x) + lap[11] * (q[8].x + q[0].x);
    v.y += lap[8] * (q[5].y + q[3].y) + lap[9] * (q[6].y + q[2].y) + lap[10] * (q[7].y + q[1].y) + lap[11] * (q[8].y + q[0].y);
    w.x += nab[8] * (q[5].x - q[3].x) + nab[9] * (q[6].x - q[2].x) + nab[10] * (q[7].x - q[1].x) + nab[11] * (q[8].x - q[0].x);
    w.y += nab[8] * (q[5].y - q[3].y) + nab[9] * (q[6].y - q[2].y) + nab[10] * (q[7].y - q[1].y) + nab[11] * (q[8].y - q[0].y);

    const CH t = q[4];
    const R vcur = (R)vR[BS * sH + te];
    CH hh;
    hh.x = (vcur * t.x + klap0 * t.x) - (R)0.5 * v.x + w.y;
    hh.y = (vcur * t.y + klap0 * t.y) - (R)0.5 * v.y - w.x;
    TO o;
    if constexpr (FINAL) {
      constexpr int N = NC + 1;
      const double2 td = cvt2<double2>(t), h = cvt2<double2>(hh);
      const char *Cb = combR + 16u * BS * (NC * (z & 1));
      double2 od;
      if (N == 1) {
        od = td;
      } else {
        od = ((const double2 *)Cb)[te];
#pragma unroll
        for (int n = 1; n < N - 1; ++n) {
          const double2 cc = cvt2<double2>(((const CH *)(Cb + 16u * BS * n))[te]);
          const double2 zc = a.zc[n - 1];
          od.x = od.x + (zc.x * cc.x - zc.y * cc.y);
          od.y = od.y + (zc.x * cc.y + zc.y * cc.x);
        }
        const double2 zc = a.zc[N > 1 ? N - 2 : 0];
        od.x = od.x + (zc.x * td.x - zc.y * td.y);
        od.y = od.y + (zc.x * td.y + zc.y * td.x);
      }
      const double2 zc = a.zc[N - 1];
      od.x = od.x + (zc.x * h.x - zc.y * h.y);
      od.y = od.y + (zc.x * h.y + zc.y * h.x);
      o = cvt2<TO>(od);
    } else {
      o = cvt2<TO>(hh);
    }
    // the store is the one compiler-visible vm op per plane (counted in kOpsAfterComb):
    // every lane stores, lanes without a point into the sink, so each wave issues it
    TO *dst = outY ? (TO *)a.hout + orbOff + colOff + zs * wrapi(z, nz) : (TO *)a.sink + tid;
    *dst = o;
#pragma unroll
    for (int i = 0; i < 8; ++i) q[i] = q[i + 1];
    sT = (sT + 1 == kRingT) ? 0 : sT + 1;
    sH = (sH + 1 == kRingH) ? 0 : sH + 1;
  }
  wait_vm<0>();  // drain the look-ahead DMAs before the LDS is released
}

// ---------------------------------------------------------------------------
// Whole-plane form of the terms without the Taylor sum, for planes of up to PY * BS
// points: lane `tid` computes the plane elements tid + k * BS (k < PY), so a 32 x 32
// plane takes 512 threads with two points each.  There is no y halo (y wraps inside
// the plane held in LDS), so every orbital plane is read once, plus the 4 + 4 wrap
// planes of the z period.  LDS: tile ring (5 + D) x PY*BS elements, V ring (D + 1) x
// PY*BS doubles, D = look-ahead in planes (D = 2: 136 KB for 32 x 32 in FP64, one block of
// 8 waves per CU, two waves per SIMD, which leaves 256 registers per lane for the two z
// queues).  The wait at the top of a plane lets the DMAs of the last D - 1 planes stay in
// flight.  The DMA issue discipline is the one of stencil_tile; lanes without a point
// store into the sink.
// ---------------------------------------------------------------------------
template <typename TI, typename R, typename TO, int PY, int BS, int D>
__device__ __forceinline__ void stencil_plane(const StencilArgs &a, int bid, double2 *lds) {
  static_assert(PY * BS <= 1024, "the sink holds 1024 elements");
  static_assert(D >= 2, "the ring wait assumes a look-ahead of at least two planes");
  using CH = typename Cx<R>::T;
  constexpr int kRingT = 5 + D, kRingH = D + 1;
  constexpr int WI = DmaW<TI>::W;
  constexpr unsigned EB = sizeof(TI);
  constexpr int SL = PY * BS;  // elements per ring slot
  // memory instructions per wave per plane, in issue order: tile DMAs (PY x WI), V DMAs (PY x 2), PY stores
  constexpr int kOps = PY * WI + 2 * PY + PY;
  TI *tileR = (TI *)lds;                       // kRingT x SL
  double *vR = (double *)(tileR + kRingT * SL);  // kRingH x SL

  const int zt = bid % a.nzt;
  const int b = bid / a.nzt;
  const int ik = b / a.nb;
  const int ispin = b % a.nspin;
  const int nx = a.nx, ny = a.ny, nz = a.nz, ly = a.ly;
  const int z0 = zt * a.TZ;
  const int z1 = min(z0 + a.TZ, nz);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int ntile = nx * ny;
  const size_t zs = (size_t)nx * ly;
  const size_t orbOff = (size_t)b * a.osz;
  const TI *in = (const TI *)a.in + orbOff;
  const size_t gplane = (size_t)nx * ny;
  // in-plane offset x + nx y of element e (elements past the plane repeat element 0)
  auto pofs = [&](int e) {
    const int ec = e < ntile ? e : 0;
    return (size_t)(ec % nx) + (size_t)nx * (ec / nx);
  };

  int te[PY], xo[PY][8], yo[PY][8];
  bool act[PY];
  size_t col[PY], toff[PY][2], voff[PY][2];
#pragma unroll
  for (int k = 0; k < PY; ++k) {
    const int e = tid + k * BS;
    act[k] = e < ntile;
    te[k] = act[k] ? e : 0;
    const int tx = te[k] % nx, ty = te[k] / nx;
    col[k] = (size_t)tx + (size_t)nx * ty;
#pragma unroll
    for (int n = 1; n <= 4; ++n) {
      xo[k][n - 1] = ty * nx + wrapi(tx + n, nx);
      xo[k][3 + n] = ty * nx + wrapi(tx - n, nx);
      yo[k][n - 1] = wrapi(ty + n, ny) * nx + tx;
      yo[k][3 + n] = wrapi(ty - n, ny) * nx + tx;
    }
    // DMA sources: W = 1 the lane's own point; W = 2 (and V) points 32 kk + lane/2 of the wave's chunk
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      const size_t p = pofs(64 * wave + k * BS + 32 * kk + (lane >> 1));
      toff[k][kk] = WI == 1 ? col[k] : p;
      voff[k][kk] = 2 * p + (lane & 1);
    }
  }
  const double *Vb = a.V + (size_t)ispin * gplane * nz;
  const unsigned tileA = __builtin_amdgcn_readfirstlane(lds_addr(tileR) + EB * 64u * wave);
  const unsigned vA = __builtin_amdgcn_readfirstlane(lds_addr(vR) + 8u * 64u * wave);
  auto dma_tile = [&](int plane, int slot) {
    const TI *src = in + zs * wrapi(plane, nz);
#pragma unroll
    for (int k = 0; k < PY; ++k) dma_chunk<TI>(src, toff[k], tileA + EB * (unsigned)(SL * slot + k * BS), lane);
  };
  auto dma_v = [&](int plane, int slot) {
    const float *vb = (const float *)Vb + 2 * gplane * wrapi(plane, nz);
#pragma unroll
    for (int k = 0; k < PY; ++k) {
      const unsigned m0 = vA + 8u * (unsigned)(SL * slot + k * BS);
      dma4(vb + voff[k][0], m0);
      dma4(vb + voff[k][1], m0 + 256u);
    }
  };
  const double *kp = a.kpar + 16 * ik;
  const R klap0 = (R)kp[0];
  R nab[12], lap[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    nab[i] = (R)kp[1 + i];
    lap[i] = (R)a.lap[i];
  }

  // prologue: queues of planes z0-4 .. z0+3 (plain loads), rings for planes z0 .. z0+3+D
  CH q[PY][9];
#pragma unroll
  for (int k = 0; k < PY; ++k)
#pragma unroll
    for (int i = 0; i < 8; ++i) q[k][i] = cvt2<CH>(in[col[k] + zs * wrapi(z0 - 4 + i, nz)]);
  __builtin_amdgcn_s_waitcnt(0);
  for (int p = 0; p < 4 + D; ++p) dma_tile(z0 + p, (z0 + p) % kRingT);
  for (int p = 0; p < D; ++p) dma_v(z0 + p, (z0 + p) % kRingH);
  wait_vm<0>();

  int sT = z0 % kRingT, sH = z0 % kRingH;
  for (int z = z0; z < z1; ++z) {
    wait_vm<(D - 1) * kOps>();
    __syncthreads();
    // past the chunk's last needed plane the look-ahead re-reads that plane into a slot never read
    dma_tile(min(z + 4 + D, z1 + 3), (sT + 4 + D) % kRingT);
    dma_v(min(z + D, z1 - 1), (sH + D) % kRingH);
    const TI *T = tileR + SL * sT;
    const TI *T4 = tileR + SL * ((sT + 4) % kRingT);
    const double *Vs = vR + SL * sH;
#pragma unroll
    for (int k = 0; k < PY; ++k) {
      q[k][8] = cvt2<CH>(T4[te[k]]);
      CH v, w;
      {
        const CH p1 = cvt2<CH>(T[xo[k][0]]), p2 = cvt2<CH>(T[xo[k][1]]), p3 = cvt2<CH>(T[xo[k][2]]),
                 p4 = cvt2<CH>(T[xo[k][3]]);
        const CH m1 = cvt2<CH>(T[xo[k][4]]), m2 = cvt2<CH>(T[xo[k][5]]), m3 = cvt2<CH>(T[xo[k][6]]),
                 m4 = cvt2<CH>(T[xo[k][7]]);
        v.x = lap[0] * (p1.x + m1.x) + lap[1] * (p2.x + m2.x) + lap[2] * (p3.x + m3.x) + lap[3] * (p4.x + m4.x);
        v.y = lap[0] * (p1.y + m1.y) + lap[1] * (p2.y + m2.y) + lap[2] * (p3.y + m3.y) + lap[3] * (p4.y + m4.y);
        w.x = nab[0] * (p1.x - m1.x) + nab[1] * (p2.x - m2.x) + nab[2] * (p3.x - m3.x) + nab[3] * (p4.x - m4.x);
        w.y = nab[0] * (p1.y - m1.y) + nab[1] * (p2.y - m2.y) + nab[2] * (p3.y - m3.y) + nab[3] * (p4.y - m4.y);
      }
      {
        const CH p1 = cvt2<CH>(T[yo[k][0]]), p2 = cvt2<CH>(T[yo[k][1]]), p3 = cvt2<CH>(T[yo[k][2]]),
                 p4 = cvt2<CH>(T[yo[k][3]]);
        const CH m1 = cvt2<CH>(T[yo[k][4]]), m2 = cvt2<CH>(T[yo[k][5]]), m3 = cvt2<CH>(T[yo[k][6]]),
                 m4 = cvt2<CH>(T[yo[k][7]]);
        v.x += lap[4] * (p1.x + m1.x) + lap[5] * (p2.x + m2.x) + lap[6] * (p3.x + m3.x) + lap[7] * (p4.x + m4.x);
        v.y += lap[4] * (p1.y + m1.y) + lap[5] * (p2.y + m2.y) + lap[6] * (p3.y + m3.y) + lap[7] * (p4.y + m4.y);
        w.x += nab[4] * (p1.x - m1.x) + nab[5] * (p2.x - m2.x) + nab[6] * (p3.x - m3.x) + nab[7] * (p4.x - m4.x);
        w.y += nab[4] * (p1.y - m1.y) + nab[5] * (p2.y - m2.y) + nab[6] * (p3.y - m3.y) + nab[7] * (p4.y - m4.y);
      }
      const CH *qq = q[k];
      v.x += lap[8] * (qq[5].x + qq[3].x) + lap[9] * (qq[6].x + qq[2].x) + lap[10] * (qq[7].x + qq[1].x) +
             lap[11] * (qq[8].x + qq[0].x);
      v.y += lap[8] * (qq[5].y + qq[3].y) + lap[9] * (qq[6].y + qq[2].y) + lap[10] * (qq[7].y + qq[1].y) +
             lap[11] * (qq[8].y + qq[0].y);
      w.x += nab[8] * (qq[5].x - qq[3].x) + nab[9] * (qq[6].x - qq[2].x) + nab[10] * (qq[7].x - qq[1].x) +
             nab[11] * (qq[8].x - qq[0].x);
      w.y += nab[8] * (qq[5].y - qq[3].y) + nab[9] * (qq[6].y - qq[2].y) + nab[10] * (qq[7].y - qq[1].y) +
             nab[11] * (qq[8].y - qq[0].y);
      const CH t = qq[4];
      const R vcur = (R)Vs[te[k]];
      CH hh;
      hh.x = (vcur * t.x + klap0 * t.x) - (R)0.5 * v.x + w.y;
      hh.y = (vcur * t.y + klap0 * t.y) - (R)0.5 * v.y - w.x;
      TO *dst = act[k] ? (TO *)a.hout + orbOff + col[k] + zs * z : (TO *)a.sink + tid + k * BS;
      *dst = cvt2<TO>(hh);
#pragma unroll
      for (int i = 0; i < 8; ++i) q[k][i] = q[k][i + 1];
    }
    sT = (sT + 1 == kRingT) ? 0 : sT + 1;
    sH = (sH + 1 == kRingH) ? 0 : sH + 1;
  }
  wait_vm<0>();
}

// ---------------------------------------------------------------------------
// Terms without the Taylor sum on 256-point y-tiles (TY rows, 4 + 4 halo rows, nx <= 32),
// with every LDS address fixed at compile time: one ring slot holds a plane's TY + 8 rows
// (tile and halo together, so y neighbours are row offsets in one slot), the ring has 8
// slots and the z loop is unrolled by 8, so slot offsets become immediate offsets of the
// LDS reads; z neighbours above come from the ring, those below from a 4-deep register
// queue whose rotation is resolved by the unrolling (no moves).  Issue order per plane and
// wave: the tile DMAs of plane z+7 (NCH x W, or with PR ceil(NCH/2) paired dwordx4 ones), the V
// DMAs of plane z+3 (2 dword, or with PR 1 dwordx4), 1 store; the wait at the top of a plane
// leaves the last two planes' operations in flight.  Needs TZ % 8 == 0.
// Used for the middle terms of the mixed chain (instruction-issue bound): the pair sums of
// the x, y and z neighbours go into one fused multiply-add chain per component, a regrouping
// of stencil_tile's sum that saves two instructions per component.
// ---------------------------------------------------------------------------
template <int P> struct Ph { static constexpr int v = P; };

template <typename TI, typename R, typename TO, int NCH, bool PR>
__device__ __forceinline__ void stencil_tile8(const StencilArgs &a, int bid, double2 *lds) {
  using CH = typename Cx<R>::T;
  constexpr int BS = kTile;
  constexpr int WI = DmaW<TI>::W;
  constexpr unsigned EB = sizeof(TI);
  constexpr int SLOT = NCH * BS;              // elements per slot: (TY + 8) * nx <= NCH * 256
  // PR (FP32 elements, rows of even length): the wave's 64 NCH slot elements move by paired dwordx4
  // DMAs (lane l: elements 2l and 2l + 1 of a 128-element piece, one row, 16-byte aligned), one
  // instruction for NCH = 2 and two for NCH = 3 (the second on lanes < 32), and the V of its 64
  // points by one more (lanes < 32: points 64 w + 2l, 2l + 1)
  constexpr bool X4 = PR;
  static_assert(!PR || WI == 2, "paired DMAs move FP32 elements");
  constexpr int NT4 = (NCH + 1) / 2;
  constexpr int kOps = X4 ? NT4 + 1 + 1 : NCH * WI + 2 + 1;
  TI *tileR = (TI *)lds;                      // 8 x SLOT
  double *vR = (double *)(tileR + 8 * SLOT);  // 4 x BS

  const int ytile = bid % a.nyt;
  const int zt = (bid / a.nyt) % a.nzt;
  const int b = bid / (a.nyt * a.nzt);
  const int ik = b / a.nb;
  const int ispin = b % a.nspin;
  const int nx = a.nx, ny = a.ny, nz = a.nz, ly = a.ly;
  const int TY = a.TY;
  const int y0 = ytile * TY;
  const int z0 = zt * a.TZ;
  const int z1 = min(z0 + a.TZ, nz);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int ntile = TY * nx;
  const bool active = tid < ntile;
  const int te = active ? tid : 0;
  const int tx = te % nx, ty = te / nx;
  const int y = y0 + ty;
  const bool outY = active && y < ny;
  const size_t zs = (size_t)nx * ly;
  const size_t orbOff = (size_t)b * a.osz;
  const TI *in = (const TI *)a.in + orbOff;
  const size_t colOff = (size_t)tx + (size_t)nx * wrapi(y, ny);
  const size_t gplane = (size_t)nx * ny;
  const int nslot = (TY + 8) * nx;
  // slot element e -> in-plane source offset (row r = global row y0 - 4 + r)
  auto sofs = [&](int e) {
    const int ec = e < nslot ? e : 0;
    return (size_t)(ec % nx) + (size_t)nx * wrapi(y0 - 4 + ec / nx, ny);
  };
  size_t toff[NCH][2];
#pragma unroll
  for (int k = 0; k < NCH; ++k)
#pragma unroll
    for (int kk = 0; kk < 2; ++kk)
      toff[k][kk] = sofs(64 * (wave + 4 * k) + (WI == 1 ? lane : 32 * kk + (lane >> 1)));
  size_t voff[2];
#pragma unroll
  for (int kk = 0; kk < 2; ++kk) {
    const int e = 64 * wave + 32 * kk + (lane >> 1);
    const int ec = e < ntile ? e : 0;
    voff[kk] = 2 * ((size_t)(ec % nx) + (size_t)nx * wrapi(y0 + ec / nx, ny)) + (lane & 1);
  }
  size_t toff4[2] = {0, 0}, voff4 = 0;
  if (X4) {
#pragma unroll
    for (int i = 0; i < NT4; ++i) toff4[i] = sofs(64 * NCH * wave + 128 * i + 2 * lane);
    const int e = 64 * wave + 2 * (lane & 31);
    const int ec = e < ntile ? e : 0;
    voff4 = (size_t)(ec % nx) + (size_t)nx * wrapi(y0 + ec / nx, ny);
  }
  const double *Vb = a.V + (size_t)ispin * gplane * nz;
  const unsigned tileA = __builtin_amdgcn_readfirstlane(lds_addr(tileR) + EB * (X4 ? 64u * NCH : 64u) * wave);
  const unsigned vA = __builtin_amdgcn_readfirstlane(lds_addr(vR) + 8u * 64u * wave);
  auto dma_tile = [&](int plane, int slot) {
    const TI *src = in + zs * wrapi(plane, nz);
    if constexpr (X4) {
      dma16(src + toff4[0], tileA + EB * (unsigned)(SLOT * slot));
      if constexpr (NT4 > 1) {
        if (lane < 32) dma16(src + toff4[1], tileA + EB * (unsigned)(SLOT * slot + 128));
      }
    } else {
#pragma unroll
      for (int k = 0; k < NCH; ++k) dma_chunk<TI>(src, toff[k], tileA + EB * (unsigned)(SLOT * slot + 4 * 64 * k), lane);
    }
  };
  auto dma_v = [&](int plane, int slot) {
    if constexpr (X4) {
      if (lane < 32) dma16(Vb + gplane * wrapi(plane, nz) + voff4, vA + 8u * BS * slot);
    } else {
      const float *vb = (const float *)Vb + 2 * gplane * wrapi(plane, nz);
      dma4(vb + voff[0], vA + 8u * BS * slot);
      dma4(vb + voff[1], vA + 8u * BS * slot + 256u);
    }
  };
  // element indices in a slot: own point, x and y neighbours
  const int own = (ty + 4) * nx + tx;
  int xo[8], yo[8];
#pragma unroll
  for (int n = 1; n <= 4; ++n) {
    xo[n - 1] = (ty + 4) * nx + wrapi(tx + n, nx);
    xo[3 + n] = (ty + 4) * nx + wrapi(tx - n, nx);
    yo[n - 1] = (ty + 4 + n) * nx + tx;
    yo[3 + n] = (ty + 4 - n) * nx + tx;
  }
  const double *kp = a.kpar + 16 * ik;
  const R klap0 = (R)kp[0];
  R nab[12], lap[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    nab[i] = (R)kp[1 + i];
    lap[i] = (R)a.lap[i];
  }
  CH q[4];  // planes z-4 .. z-1 at the own point; plane z-4+k sits in q[(phase + k) & 3]
#pragma unroll
  for (int i = 0; i < 4; ++i) q[i] = cvt2<CH>(in[colOff + zs * wrapi(z0 - 4 + i, nz)]);
  __builtin_amdgcn_s_waitcnt(0);
  for (int p = 0; p < 7; ++p) dma_tile(z0 + p, p);
  for (int p = 0; p < 3; ++p) dma_v(z0 + p, p);
  wait_vm<0>();

  auto body = [&](int z, auto ph) {
    constexpr int P = decltype(ph)::v;
    wait_vm<2 * kOps>();
    __syncthreads();
    dma_tile(min(z + 7, z1 + 3), (P + 7) & 7);
    dma_v(min(z + 3, z1 - 1), (P + 3) & 3);
    const TI *T = tileR + SLOT * P;
    // one accumulation chain per component (x, then y, then z pairs: 12 fused multiply-adds), then
    // hh = (V t + klap0 t) - v/2 + i w
    R vx, vy, wx, wy;
    {
      const CH p1 = cvt2<CH>(T[xo[0]]), p2 = cvt2<CH>(T[xo[1]]), p3 = cvt2<CH>(T[xo[2]]), p4 = cvt2<CH>(T[xo[3]]);
      const CH m1 = cvt2<CH>(T[xo[4]]), m2 = cvt2<CH>(T[xo[5]]), m3 = cvt2<CH>(T[xo[6]]), m4 = cvt2<CH>(T[xo[7]]);
      vx = lap[0] * (p1.x + m1.x);
      vy = lap[0] * (p1.y + m1.y);
      wx = nab[0] * (p1.x - m1.x);
      wy = nab[0] * (p1.y - m1.y);
      vx += lap[1] * (p2.x + m2.x); vy += lap[1] * (p2.y + m2.y); wx += nab[1] * (p2.x - m2.x); wy += nab[1] * (p2.y - m2.y);
      vx += lap[2] * (p3.x + m3.x); vy += lap[2] * (p3.y + m3.y); wx += nab[2] * (p3.x - m3.x); wy += nab[2] * (p3.y - m3.y);
      vx += lap[3] * (p4.x + m4.x); vy += lap[3] * (p4.y + m4.y); wx += nab[3] * (p4.x - m4.x); wy += nab[3] * (p4.y - m4.y);
    }
    {
      const CH p1 = cvt2<CH>(T[yo[0]]), p2 = cvt2<CH>(T[yo[1]]), p3 = cvt2<CH>(T[yo[2]]), p4 = cvt2<CH>(T[yo[3]]);
      const CH m1 = cvt2<CH>(T[yo[4]]), m2 = cvt2<CH>(T[yo[5]]), m3 = cvt2<CH>(T[yo[6]]), m4 = cvt2<CH>(T[yo[7]]);
      vx += lap[4] * (p1.x + m1.x); vy += lap[4] * (p1.y + m1.y); wx += nab[4] * (p1.x - m1.x); wy += nab[4] * (p1.y - m1.y);
      vx += lap[5] * (p2.x + m2.x); vy += lap[5] * (p2.y + m2.y); wx += nab[5] * (p2.x - m2.x); wy += nab[5] * (p2.y - m2.y);
      vx += lap[6] * (p3.x + m3.x); vy += lap[6] * (p3.y + m3.y); wx += nab[6] * (p3.x - m3.x); wy += nab[6] * (p3.y - m3.y);
      vx += lap[7] * (p4.x + m4.x); vy += lap[7] * (p4.y + m4.y); wx += nab[7] * (p4.x - m4.x); wy += nab[7] * (p4.y - m4.y);
    }
    const CH t = cvt2<CH>(T[own]);
    const CH u1 = cvt2<CH>(tileR[SLOT * ((P + 1) & 7) + own]), u2 = cvt2<CH>(tileR[SLOT * ((P + 2) & 7) + own]);
    const CH u3 = cvt2<CH>(tileR[SLOT * ((P + 3) & 7) + own]), u4 = cvt2<CH>(tileR[SLOT * ((P + 4) & 7) + own]);
    const CH d1 = q[(P + 3) & 3], d2 = q[(P + 2) & 3], d3 = q[(P + 1) & 3], d4 = q[P & 3];
    vx += lap[8] * (u1.x + d1.x); vy += lap[8] * (u1.y + d1.y); wx += nab[8] * (u1.x - d1.x); wy += nab[8] * (u1.y - d1.y);
    vx += lap[9] * (u2.x + d2.x); vy += lap[9] * (u2.y + d2.y); wx += nab[9] * (u2.x - d2.x); wy += nab[9] * (u2.y - d2.y);
    vx += lap[10] * (u3.x + d3.x); vy += lap[10] * (u3.y + d3.y); wx += nab[10] * (u3.x - d3.x); wy += nab[10] * (u3.y - d3.y);
    vx += lap[11] * (u4.x + d4.x); vy += lap[11] * (u4.y + d4.y); wx += nab[11] * (u4.x - d4.x); wy += nab[11] * (u4.y - d4.y);
    const R vcur = (R)vR[BS * (P & 3) + te];
    CH hh;
    hh.x = (vcur * t.x + klap0 * t.x) - (R)0.5 * vx + wy;
    hh.y = (vcur * t.y + klap0 * t.y) - (R)0.5 * vy - wx;
    TO *dst = outY ? (TO *)a.hout + orbOff + colOff + zs * z : (TO *)a.sink + tid;
    *dst = cvt2<TO>(hh);
    q[P & 3] = t;
  };
  for (int z = z0; z < z1; z += 8) {
    body(z, Ph<0>());
    body(z + 1, Ph<1>());
    body(z + 2, Ph<2>());
    body(z + 3, Ph<3>());
    body(z + 4, Ph<4>());
    body(z + 5, Ph<5>());
    body(z + 6, Ph<6>());
    body(z + 7, Ph<7>());
  }
  wait_vm<0>();
}

// ---------------------------------------------------------------------------
// The same terms with the z part of each output plane formed 4 planes late, so that a plane's
// LDS slot (tile and halo rows) is needed only while that plane is the current one: the ring
// has 4 slots (plane p read, p+1 and p+2 in flight, p+3 being written) and the block takes
// 40 KB of LDS in FP64 instead of 72 KB (four blocks per CU instead of two, twice the bytes in
// flight).  At plane p the kernel forms the x and y part and the centre term of plane p from
// the slot (kept in a 4-deep register queue), pushes p's own value into a 9-deep queue of own
// values (planes p-8 .. p) and completes plane p-4 with its z part.  The centre, x/y and z
// partial sums are added in that order (hh = ((V t + klap0 t) - v_xy/2 + i w_xy) - v_z/2 + i w_z),
// a regrouping of stencil_tile's sum.  Needs TZ % 4 == 0; the z loop runs TZ + 4 planes,
// unrolled by 4 so that slot and queue indices are compile-time.  A slot holds NCK chunks of 64
// elements (8: rows of up to 32 points; 10: 13 rows of 48 points, so that the 48-point rows keep
// three blocks per CU; 12: up to 64 points); chunk c is moved by wave c % 4, and a wave with no
// chunk left for its last DMA of a plane moves its previous chunk again (the same data to the
// same place), so that every wave issues the same number of operations per plane.
// ---------------------------------------------------------------------------
template <typename TI, typename R, typename TO, int NCK>
__device__ __forceinline__ void stencil_tilez(const StencilArgs &a, int bid, double2 *lds) {
  using CH = typename Cx<R>::T;
  constexpr int BS = kTile;
  constexpr int WI = DmaW<TI>::W;
  constexpr unsigned EB = sizeof(TI);
  constexpr int SLOT = 64 * NCK;              // elements per slot: (TY + 8) * nx <= 64 NCK
  constexpr int NW = (NCK + 3) / 4;           // tile DMA chunks per wave and plane
  constexpr int kOps = NW * WI + 2 + 1;
  TI *tileR = (TI *)lds;                      // 4 x SLOT
  double *vR = (double *)(tileR + 4 * SLOT);  // 4 x BS

  const int ytile = bid % a.nyt;
  const int zt = (bid / a.nyt) % a.nzt;
  const int b = bid / (a.nyt * a.nzt);
  const int ik = b / a.nb;
  const int ispin = b % a.nspin;
  const int nx = a.nx, ny = a.ny, nz = a.nz, ly = a.ly;
  const int TY = a.TY;
  const int y0 = ytile * TY;
  const int z0 = zt * a.TZ;
  const int z1 = min(z0 + a.TZ, nz);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int ntile = TY * nx;
  const bool active = tid < ntile;
  const int te = active ? tid : 0;
  const int tx = te % nx, ty = te / nx;
  const int y = y0 + ty;
  const bool outY = active && y < ny;
  const size_t zs = (size_t)nx * ly;
  const size_t orbOff = (size_t)b * a.osz;
  const TI *in = (const TI *)a.in + orbOff;
  const size_t colOff = (size_t)tx + (size_t)nx * wrapi(y, ny);
  const size_t gplane = (size_t)nx * ny;
  const int nslot = (TY + 8) * nx;
  auto sofs = [&](int e) {
    const int ec = e < nslot ? e : 0;
    return (size_t)(ec % nx) + (size_t)nx * wrapi(y0 - 4 + ec / nx, ny);
  };
  size_t toff[NW][2];
  unsigned kofs[NW];  // slot offset (elements) of the chunk of this wave's k-th DMA, wave-uniform
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    const int kc = wave + 4 * k < NCK ? k : k - 1;
    kofs[k] = __builtin_amdgcn_readfirstlane(256u * (unsigned)kc);
#pragma unroll
    for (int kk = 0; kk < 2; ++kk)
      toff[k][kk] = sofs(64 * (wave + 4 * kc) + (WI == 1 ? lane : 32 * kk + (lane >> 1)));
  }
  size_t voff[2];
#pragma unroll
  for (int kk = 0; kk < 2; ++kk) {
    const int e = 64 * wave + 32 * kk + (lane >> 1);
    const int ec = e < ntile ? e : 0;
    voff[kk] = 2 * ((size_t)(ec % nx) + (size_t)nx * wrapi(y0 + ec / nx, ny)) + (lane & 1);
  }
  const double *Vb = a.V + (size_t)ispin * gplane * nz;
  const unsigned tileA = __builtin_amdgcn_readfirstlane(lds_addr(tileR) + EB * 64u * wave);
  const unsigned vA = __builtin_amdgcn_readfirstlane(lds_addr(vR) + 8u * 64u * wave);
  auto dma_tile = [&](int plane, int slot) {
    const TI *src = in + zs * wrapi(plane, nz);
#pragma unroll
    for (int k = 0; k < NW; ++k) dma_chunk<TI>(src, toff[k], tileA + EB * (unsigned)(SLOT * slot) + EB * kofs[k], lane);
  };
  auto dma_v = [&](int plane, int slot) {
    const float *vb = (const float *)Vb + 2 * gplane * wrapi(plane, nz);
    dma4(vb + voff[0], vA + 8u * BS * slot);
    dma4(vb + voff[1], vA + 8u * BS * slot + 256u);
  };
  const int own = (ty + 4) * nx + tx;
  int xo[8], yo[8];
#pragma unroll
  for (int n = 1; n <= 4; ++n) {
    xo[n - 1] = (ty + 4) * nx + wrapi(tx + n, nx);
    xo[3 + n] = (ty + 4) * nx + wrapi(tx - n, nx);
    yo[n - 1] = (ty + 4 + n) * nx + tx;
    yo[3 + n] = (ty + 4 - n) * nx + tx;
  }
  const double *kp = a.kpar + 16 * ik;
  const R klap0 = (R)kp[0];
  R nab[12], lap[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    nab[i] = (R)kp[1 + i];
    lap[i] = (R)a.lap[i];
  }
  CH q[9];   // own values of planes p-8 .. p (q[8] = p) after the push at plane p
  CH hq[4];  // centre + x/y part of planes p-3 .. p; plane p's sits in hq[(p - z0) & 3]
  // With the whole column in one chunk the planes past its end are its own first four planes
  // (periodic z), whose own values are kept from the first pass: the tail needs only those
  // (no x/y part is formed there), so it fetches nothing from HBM (its slot DMAs re-read the
  // last plane, an L2 hit, to keep the per-plane memory-op count fixed).
  const bool wrapkeep = a.nzt == 1;
  CH keep[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    q[i] = cvt2<CH>(make_double2(0.0, 0.0));
    hq[i] = q[i];
    keep[i] = q[i];
    q[5 + i] = cvt2<CH>(in[colOff + zs * wrapi(z0 - 4 + i, nz)]);
  }
  q[4] = q[0];
  __builtin_amdgcn_s_waitcnt(0);
  for (int p = 0; p < 3; ++p) {
    dma_tile(z0 + p, p);
    dma_v(z0 + p, p);
  }
  wait_vm<0>();

  auto body = [&](int p, auto ph) {
    constexpr int P = decltype(ph)::v;
    wait_vm<2 * kOps>();
    __syncthreads();
    dma_tile(p + 3 < z1 ? p + 3 : (wrapkeep ? z1 - 1 : min(p + 3, z1 + 3)), (P + 3) & 3);
    dma_v(min(p + 3, z1 - 1), (P + 3) & 3);
    const TI *T = tileR + SLOT * P;
#pragma unroll
    for (int i = 0; i < 8; ++i) q[i] = q[i + 1];
    // in the tail of a whole-column chunk p = z1 + P wraps to plane P (z0 = 0, TZ % 4 == 0)
    q[8] = (wrapkeep && p >= z1) ? keep[P] : cvt2<CH>(T[own]);
    if (p < z0 + 4) keep[P] = q[8];
    const CH hp = hq[P];  // plane p - 4
    if (p < z1) {
      CH v, w;
      {
        const CH p1 = cvt2<CH>(T[xo[0]]), p2 = cvt2<CH>(T[xo[1]]), p3 = cvt2<CH>(T[xo[2]]), p4 = cvt2<CH>(T[xo[3]]);
        const CH m1 = cvt2<CH>(T[xo[4]]), m2 = cvt2<CH>(T[xo[5]]), m3 = cvt2<CH>(T[xo[6]]), m4 = cvt2<CH>(T[xo[7]]);
        v.x = lap[0] * (p1.x + m1.x) + lap[1] * (p2.x + m2.x) + lap[2] * (p3.x + m3.x) + lap[3] * (p4.x + m4.x);
        v.y = lap[0] * (p1.y + m1.y) + lap[1] * (p2.y + m2.y) + lap[2] * (p3.y + m3.y) + lap[3] * (p4.y + m4.y);
        w.x = nab[0] * (p1.x - m1.x) + nab[1] * (p2.x - m2.x) + nab[2] * (p3.x - m3.x) + nab[3] * (p4.x - m4.x);
        w.y = nab[0] * (p1.y - m1.y) + nab[1] * (p2.y - m2.y) + nab[2] * (p3.y - m3.y) + nab[3] * (p4.y - m4.y);
      }
      {
        const CH p1 = cvt2<CH>(T[yo[0]]), p2 = cvt2<CH>(T[yo[1]]), p3 = cvt2<CH>(T[yo[2]]), p4 = cvt2<CH>(T[yo[3]]);
        const CH m1 = cvt2<CH>(T[yo[4]]), m2 = cvt2<CH>(T[yo[5]]), m3 = cvt2<CH>(T[yo[6]]), m4 = cvt2<CH>(T[yo[7]]);
        v.x += lap[4] * (p1.x + m1.x) + lap[5] * (p2.x + m2.x) + lap[6] * (p3.x + m3.x) + lap[7] * (p4.x + m4.x);
        v.y += lap[4] * (p1.y + m1.y) + lap[5] * (p2.y + m2.y) + lap[6] * (p3.y + m3.y) + lap[7] * (p4.y + m4.y);
        w.x += nab[4] * (p1.x - m1.x) + nab[5] * (p2.x - m2.x) + nab[6] * (p3.x - m3.x) + nab[7] * (p4.x - m4.x);
        w.y += nab[4] * (p1.y - m1.y) + nab[5] * (p2.y - m2.y) + nab[6] * (p3.y - m3.y) + nab[7] * (p4.y - m4.y);
      }
      const CH t = q[8];
      const R vcur = (R)vR[BS * P + te];
      CH hc;
      hc.x = (vcur * t.x + klap0 * t.x) - (R)0.5 * v.x + w.y;
      hc.y = (vcur * t.y + klap0 * t.y) - (R)0.5 * v.y - w.x;
      hq[P] = hc;
    }
    CH v, w;
    v.x = lap[8] * (q[5].x + q[3].x) + lap[9] * (q[6].x + q[2].x) + lap[10] * (q[7].x + q[1].x) + lap[11] * (q[8].x + q[0].x);
    v.y = lap[8] * (q[5].y + q[3].y) + lap[9] * (q[6].y + q[2].y) + lap[10] * (q[7].y + q[1].y) + lap[11] * (q[8].y + q[0].y);
    w.x = nab[8] * (q[5].x - q[3].x) + nab[9] * (q[6].x - q[2].x) + nab[10] * (q[7].x - q[1].x) + nab[11] * (q[8].x - q[0].x);
    w.y = nab[8] * (q[5].y - q[3].y) + nab[9] * (q[6].y - q[2].y) + nab[10] * (q[7].y - q[1].y) + nab[11] * (q[8].y - q[0].y);
    CH hh;
    hh.x = hp.x - (R)0.5 * v.x + w.y;
    hh.y = hp.y - (R)0.5 * v.y - w.x;
    // every lane stores every plane (the vm count per plane is fixed): the first four planes
    // of the chunk and lanes without a point store into the sink
    const bool outP = outY && p >= z0 + 4;
    TO *dst = outP ? (TO *)a.hout + orbOff + colOff + zs * (p - 4) : (TO *)a.sink + tid;
    *dst = cvt2<TO>(hh);
  };
  for (int p = z0; p < z1 + 4; p += 4) {
    body(p, Ph<0>());
    body(p + 1, Ph<1>());
    body(p + 2, Ph<2>());
    body(p + 3, Ph<3>());
  }
  wait_vm<0>();
}

template <typename TI, typename R, typename TO, int NCK>
__global__ __launch_bounds__(kTile) void k_stencilz(StencilArgs a) {
  extern __shared__ double2 lds[];
  if (a.skip && *a.skip) return;  // a device-side gate switched this launch off
  int bid = blockIdx.x;
  const int nfull = (a.nblk / 8) * 8;
  if (bid < nfull) bid = (bid % 8) * (a.nblk / 8) + bid / 8;  // XCD-aware tile order
  stencil_tilez<TI, R, TO, NCK>(a, bid, lds);
}

// Persistent form over many systems (the batched step's terms without the Taylor sum): block b
// takes tiles b, b + grid, ... of `tiles` = (system index, tile index).  stencil_tilez returns
// with its DMAs drained (vmcnt 0); the barrier keeps the next tile's DMAs off slots that other
// waves still read.
template <typename TI, typename R, typename TO, int NCK>
__global__ __launch_bounds__(kTile) void k_stencilz_packed(const StencilArgs *__restrict__ sys,
                                                           const int2 *__restrict__ tiles, int ntiles) {
  extern __shared__ double2 lds[];
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int2 tr = tiles[t];
    const int si = __builtin_amdgcn_readfirstlane(tr.x);
    const int bid = __builtin_amdgcn_readfirstlane(tr.y);
    const StencilArgs a = sys[si];
    stencil_tilez<TI, R, TO, NCK>(a, bid, lds);
    __syncthreads();
  }
}

template <typename TI, typename R, typename TO, int NCH, bool PR>
__global__ __launch_bounds__(kTile) void k_stencil8(StencilArgs a) {
  extern __shared__ double2 lds[];
  if (a.skip && *a.skip) return;  // a device-side gate switched this launch off
  int bid = blockIdx.x;
  const int nfull = (a.nblk / 8) * 8;
  if (bid < nfull) bid = (bid % 8) * (a.nblk / 8) + bid / 8;  // XCD-aware tile order
  stencil_tile8<TI, R, TO, NCH, PR>(a, bid, lds);
}

template <typename TI, typename R, typename TO, int HC, int NC, int BS>
__global__ __launch_bounds__(BS) void k_stencil(StencilArgs a) {
  extern __shared__ double2 lds[];
  if (a.skip && *a.skip) return;  // a device-side gate switched this launch off
  int bid = blockIdx.x;
  const int nfull = (a.nblk / 8) * 8;
  if (bid < nfull) bid = (bid % 8) * (a.nblk / 8) + bid / 8;  // XCD-aware tile order
  stencil_tile<TI, R, TO, HC, NC, BS>(a, bid, lds);
}

template <typename TI, typename R, typename TO, int PY, int BS, int D>
__global__ __launch_bounds__(BS) void k_stencil_plane(StencilArgs a) {
  extern __shared__ double2 lds[];
  if (a.skip && *a.skip) return;  // a device-side gate switched this launch off
  int bid = blockIdx.x;
  const int nfull = (a.nblk / 8) * 8;
  if (bid < nfull) bid = (bid % 8) * (a.nblk / 8) + bid / 8;  // XCD-aware tile order
  stencil_plane<TI, R, TO, PY, BS, D>(a, bid, lds);
}

// Persistent form over many systems (the batched step): block b takes tiles b, b + grid, ...
// of `tiles` = (system index, tile index); all systems share the template shape.
template <typename TI, typename R, typename TO, int HC, int NC, int BS>
__global__ __launch_bounds__(BS) void k_stencil_packed(const StencilArgs *__restrict__ sys,
                                                       const int2 *__restrict__ tiles, int ntiles) {
  extern __shared__ double2 lds[];
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int2 tr = tiles[t];
    // the tile entry is the same for the whole block: uniform indices let the descriptor come in
    // through scalar loads, once per tile
    const int si = __builtin_amdgcn_readfirstlane(tr.x);
    const int bid = __builtin_amdgcn_readfirstlane(tr.y);
    const StencilArgs a = sys[si];
    stencil_tile<TI, R, TO, HC, NC, BS>(a, bid, lds);
    __syncthreads();  // every wave is done with the LDS rings before the next tile's DMAs
  }
}

// ---------------------------------------------------------------------------
// KB term of hpsi at the sphere points only (zpseudo's scatter,
// nonlocal_potential.f90:370-378), as a deterministic per-point gather over
// the point -> (atom, j) inverse map, entries in ascending ilma:
//   hout(p) += s * sum_e sum_l c(ilma0_e + l) zekr(j_e, ilma0_e + l)
// with s = 1 (plain / intermediate terms) or s = c_N (last Taylor term).
// One thread per touched point and G orbitals of one k-point.  The kernel is
// bound by the latency of its dependent loads, so the per-point record
// (orbital-block offset, first entry, entry count, first entry's ilma0 | cnt)
// comes in one 16-byte load, and with CM > 0 the CM projectors of an entry are
// loaded together (clamped to the entry's own, unused ones weighted by zero are
// skipped in the sum, so the order of the additions is the sequential one).
// ---------------------------------------------------------------------------
// Orbitals per thread and projectors per entry loaded together (entries with cmax <= CM), per
// element type of the chain: FP64 <8, 4>, FP32 <4, 0> (the fastest of G in {4, 8}, CM in {0, 4}
// measured at 32^3 with 64 s+p atoms).
constexpr int kScatG = 8, kScatCM = 4;
constexpr int kScatG32 = 4, kScatCM32 = 0;

struct ScatterArgs {
  void *hout;             // double2, or float2 in the mixed-precision chain
  const double2 *c;       // (nb, nlma, nk), orbital fastest
  const double2 *zent;    // (cmax, nent, nk)
  const int4 *tp_info;    // (ntp) orbital-block offset, first entry, entry count, pt_lma of the first entry
  const int *pt_lma;      // (nent) ilma0 | cnt << 24
  double2 s;
  size_t osz;
  int ntp, nb, nlma, nent, cmax, ngrp;
  const int *skip;        // device flag: the launch does nothing when it is set (null: never)
};

template <typename TH, int G, int CM>
__device__ __forceinline__ void nl_scatter_body(const ScatterArgs &a, int bx, int by) {
  const int i = bx * kBlock + threadIdx.x;
  const int ik = by / a.ngrp;
  const int b0 = (by % a.ngrp) * G;
  if (i >= a.ntp) return;
  const int4 inf = a.tp_info[i];
  const int nob = min(G, a.nb - b0);
  // the read half of the read-modify-write is issued first and overlaps the KB sum
  TH *p = (TH *)a.hout + ((size_t)ik * a.nb + b0) * a.osz + inf.x;
  double2 old[G];
#pragma unroll
  for (int o = 0; o < G; ++o)
    if (o < nob) old[o] = cvt2<double2>(p[(size_t)o * a.osz]);
  const double2 *zb = a.zent + (size_t)ik * a.nent * a.cmax;
  const double2 *cb = a.c + (size_t)ik * a.nlma * a.nb + b0;
  double2 acc[G];
#pragma unroll
  for (int o = 0; o < G; ++o) acc[o] = make_double2(0.0, 0.0);
  int lc = inf.w;
  for (int q = 0; q < inf.z; ++q) {
    const int e = inf.y + q;
    if (q > 0) lc = a.pt_lma[e];
    const int ilma0 = lc & 0xffffff;
    const int cnt = lc >> 24;
    const double2 *ze = zb + (size_t)e * a.cmax;
    if (CM > 0) {
      double2 zv[CM > 0 ? CM : 1], cv[CM > 0 ? CM : 1][G];
#pragma unroll
      for (int l = 0; l < CM; ++l) {
        const int lx = min(l, cnt - 1);
        zv[l] = ze[lx];
#pragma unroll
        for (int o = 0; o < G; ++o) cv[l][o] = cb[(size_t)(ilma0 + lx) * a.nb + min(o, nob - 1)];
      }
#pragma unroll
      for (int l = 0; l < CM; ++l) {
        if (l < cnt) {
#pragma unroll
          for (int o = 0; o < G; ++o) {
            acc[o].x += cv[l][o].x * zv[l].x - cv[l][o].y * zv[l].y;
            acc[o].y += cv[l][o].x * zv[l].y + cv[l][o].y * zv[l].x;
          }
        }
      }
    } else {
      for (int l = 0; l < cnt; ++l) {
        const double2 zvl = ze[l];
        const double2 *cl = cb + (size_t)(ilma0 + l) * a.nb;
#pragma unroll
        for (int o = 0; o < G; ++o) {
          if (o < nob) {
            const double2 c = cl[o];
            acc[o].x += c.x * zvl.x - c.y * zvl.y;
            acc[o].y += c.x * zvl.y + c.y * zvl.x;
          }
        }
      }
    }
  }
#pragma unroll
  for (int o = 0; o < G; ++o) {
    if (o < nob) {
      double2 v = old[o];
      v.x += a.s.x * acc[o].x - a.s.y * acc[o].y;
      v.y += a.s.x * acc[o].y + a.s.y * acc[o].x;
      p[(size_t)o * a.osz] = cvt2<TH>(v);
    }
  }
}

template <typename TH, int G, int CM>
__global__ __launch_bounds__(kBlock) void k_nl_scatter(ScatterArgs a) {
  if (a.skip && *a.skip) return;  // a device-side gate switched this launch off
  nl_scatter_body<TH, G, CM>(a, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// h += dv * psi over the orbital blocks (dv per point and spin; padding rows untouched):
// turns H_old psi into H_new psi when only V_local changed (salmon_hip_taylor_reuse).
// ---------------------------------------------------------------------------
template <typename TH>
__global__ __launch_bounds__(kBlock) void k_vshift(TH *h, const double2 *psi, const double *dv, size_t osz, int nx,
                                                   int ny, int ly, int nz, int nspin, size_t ntot, const int *skip) {
  if (skip && *skip) return;  // a device-side gate switched this launch off
  const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= ntot) return;
  const size_t b = i / osz;
  const size_t r = i % osz;
  const int x = (int)(r % nx);
  const int y = (int)((r / nx) % ly);
  const int z = (int)(r / ((size_t)nx * ly));
  if (y >= ny || z >= nz) return;
  const int ispin = (int)(b % nspin);
  const double d = dv[(size_t)ispin * nx * ny * nz + x + (size_t)nx * (y + (size_t)ny * z)];
  const double2 p = psi[i];
  double2 v = cvt2<double2>(h[i]);
  v.x += d * p.x;
  v.y += d * p.y;
  h[i] = cvt2<TH>(v);
}

// ---------------------------------------------------------------------------
// Density: rho(p) = sum_{ik, io} (|psi|^2 rocc) wtk, orbitals in (ik, io) order,
// in groups of gsz (kDensGroup by default) summed in order, groups combined in order.
// ---------------------------------------------------------------------------
struct DensArgs {
  const double2 *psi;
  const double *rocc;   // (no, nk, nspin)
  const double *wtk;
  double *part;         // (ngrid, nspin, ngrp)
  double *rho;          // (ngrid, nspin)
  size_t osz;
  int nx, ny, nz, ly, nspin, no, nk, ngrp, gsz;
  const int *skip;      // device flag: the launch does nothing when it is set (null: never)
};

__device__ __forceinline__ void density_part_body(const DensArgs &a, int bx, int grp, int ispin) {
  const size_t ngrid = (size_t)a.nx * a.ny * a.nz;
  const size_t g = (size_t)bx * kBlock + threadIdx.x;
  if (g >= ngrid) return;
  const int x = g % a.nx;
  const int yz = g / a.nx;
  const int y = yz % a.ny;
  const int z = yz / a.ny;
  const size_t off = (size_t)x + (size_t)a.nx * ((size_t)y + (size_t)a.ly * z);
  const int M = a.no * a.nk;
  const int m0 = grp * a.gsz, m1 = min(M, m0 + a.gsz);
  const double2 *pb = a.psi + off;
  const size_t bstride = (size_t)a.nspin * a.osz;  // one orbital to the next
  const double *occ = a.rocc + (size_t)a.no * a.nk * ispin;
  int io = m0 % a.no, ik = m0 / a.no;
  double acc = 0.0;
  int mm = m0;
  // U orbitals' loads in flight before the in-order sum consumes them
  constexpr int U = 8;
  for (; mm + U <= m1; mm += U) {
    double2 p[U];
#pragma unroll
    for (int u = 0; u < U; ++u) p[u] = pb[((size_t)ispin * a.osz) + (size_t)(mm + u) * bstride];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double w2 = p[u].x * p[u].x + p[u].y * p[u].y;
      acc = acc + (w2 * occ[io + (size_t)a.no * ik]) * a.wtk[ik];
      if (++io == a.no) { io = 0; ++ik; }
    }
  }
  for (; mm < m1; ++mm) {
    const double2 p = pb[((size_t)ispin * a.osz) + (size_t)mm * bstride];
    const double w2 = p.x * p.x + p.y * p.y;
    acc = acc + (w2 * occ[io + (size_t)a.no * ik]) * a.wtk[ik];
    if (++io == a.no) { io = 0; ++ik; }
  }
  if (a.ngrp == 1)
    a.rho[g + ngrid * ispin] = acc;
  else
    a.part[g + ngrid * ((size_t)ispin + (size_t)a.nspin * grp)] = acc;
}

__global__ __launch_bounds__(kBlock) void k_density_part(DensArgs a) {
  if (a.skip && *a.skip) return;  // a device-side gate switched this launch off
  density_part_body(a, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Partial sums of salmon_hip_density_noalloc: a file-scope device array (one copy per GPU,
// held until the process ends) in place of the handle's partial buffer.  Host code never
// passes its name to a kernel: the per-device address from hipGetSymbolAddress, resolved
// once per device when a handle is created there, is what the launches use.
constexpr size_t kNoAllocScratch = (size_t)8 << 20;  // doubles (64 MiB)
static __device__ __align__(256) double scratch[kNoAllocScratch];

__device__ __forceinline__ void density_sum_body(const DensArgs &a, int bx) {
  const size_t ngrid = (size_t)a.nx * a.ny * a.nz;
  const size_t i = (size_t)bx * kBlock + threadIdx.x;
  if (i >= ngrid * a.nspin) return;
  const size_t gs = ngrid * a.nspin;  // one group's partials to the next
  double acc = 0.0;
  int grp = 0;
  // 8 partials' loads in flight ahead of the in-order sum
  for (; grp + 8 <= a.ngrp; grp += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = a.part[i + gs * (size_t)(grp + u)];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc = acc + v[u];
  }
  for (; grp < a.ngrp; ++grp) acc = acc + a.part[i + gs * (size_t)grp];
  a.rho[i] = acc;
}

__global__ __launch_bounds__(kBlock) void k_density_sum(DensArgs a) {
  if (a.skip && *a.skip) return;  // a device-side gate switched this launch off
  density_sum_body(a, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Current, local part (stencil_current, density_matrix.f90:465-530): per
// orbital sum |psi|^2 and Im sum_n (nab_n conj(psi)) psi(r + n e_d) over
// forward neighbours.  Block = full x rows x TY rows, streaming z over the
// whole column: z-forward neighbours from a 5-deep register queue, x and
// y-forward neighbours from an LDS plane carrying 4 extra rows below the tile.
// Each block writes its 4 partial sums (fixed-order tree) to part(4, tile, orbital).
// ---------------------------------------------------------------------------
struct CurArgs {
  const double2 *psi;
  double *part;            // (4, nyt, norb)
  double coef_nab[12];
  size_t osz;
  int nx, ny, nz, ly, TY, nyt;
};

// f(integral_constant<int, I>) for I = 0, 1, ... in order while it returns true: a loop body
// whose iteration index is a compile-time constant (register-array slots stay registers).
template <int... I, typename F>
__device__ __forceinline__ void unroll_while(std::integer_sequence<int, I...>, F &&f) {
  (void)(f(std::integral_constant<int, I>{}) && ...);
}

__device__ __forceinline__ double block_sum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ void current_local_body(const CurArgs &a, int bid) {
  extern __shared__ double2 lds[];
  __shared__ double sh[kBlock];
  const int ytile = bid % a.nyt;
  const int b = bid / a.nyt;
  const int nx = a.nx, ny = a.ny, nz = a.nz, TY = a.TY;
  const int y0 = ytile * TY;
  const int tid = threadIdx.x;
  const bool active = tid < TY * nx;
  const int te = active ? tid : 0;
  const int tx = te % nx, ty = te / nx;
  const int y = y0 + ty;
  const bool own = active && y < ny;  // this lane's point belongs to the tile
  const int yw = wrapi(y, ny);
  const size_t zs = (size_t)nx * a.ly;
  const double2 *p = a.psi + (size_t)b * a.osz;
  const size_t col = (size_t)tx + (size_t)nx * yw;
  const int planeL = (TY + 4) * nx;
  // z neighbours: a register ring of R planes, plane z in slot z % R (the z loop is unrolled by R,
  // so every slot index is a constant and no register moves wait on a load); plane z + R - 1 is
  // loaded R - 5 iterations before its first use.  Halo: rows y0+TY .. y0+TY+3, 4*nx elements,
  // element e = tid + k * kBlock (k < nh <= 4, nx <= 256), loaded one plane ahead into a
  // two-deep register ring (R even: slot parity = plane parity).
  constexpr int R = 8;
  double2 q[R];
#pragma unroll
  for (int i = 0; i < R - 1; ++i) q[i] = p[col + zs * wrapi(i, nz)];
  const int nh = (4 * nx + kBlock - 1) / kBlock;
  int hoff[4];     // halo element k's offset within a plane (element 0's where the lane has none: loaded, unused)
  bool hown[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int e = tid + k * kBlock;
    hown[k] = k < nh && e < 4 * nx;
    const int ee = hown[k] ? e : 0;
    hoff[k] = ee % nx + nx * wrapi(y0 + TY + ee / nx, ny);
  }
  // halo rows of the even (hA*) and odd (hB*) planes; named registers, not arrays
  double2 hA0 = {}, hA1 = {}, hA2 = {}, hA3 = {}, hB0 = {}, hB1 = {}, hB2 = {}, hB3 = {};
  auto halo_get = [&](size_t zoff, double2 &v0, double2 &v1, double2 &v2, double2 &v3)
                      __attribute__((always_inline)) {
    v0 = p[(size_t)hoff[0] + zoff];
    if (nh > 1) v1 = p[(size_t)hoff[1] + zoff];
    if (nh > 2) v2 = p[(size_t)hoff[2] + zoff];
    if (nh > 3) v3 = p[(size_t)hoff[3] + zoff];
  };
  auto halo_put = [&](double2 *P, const double2 &v0, const double2 &v1, const double2 &v2, const double2 &v3)
                      __attribute__((always_inline)) {
    double2 *H = P + TY * nx + tid;
    if (hown[0]) H[0] = v0;
    if (hown[1]) H[kBlock] = v1;
    if (hown[2]) H[2 * kBlock] = v2;
    if (hown[3]) H[3 * kBlock] = v3;
  };
  halo_get(0, hA0, hA1, hA2, hA3);
  double r0 = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
  for (int z0 = 0; z0 < nz; z0 += R) {
    // plane z = z0 + u for the constants u = 0 .. R-1, while z < nz
    unroll_while(std::make_integer_sequence<int, R>{}, [&](auto uc) __attribute__((always_inline)) {
      constexpr int u = decltype(uc)::value;
      const int z = z0 + u;
      if (z >= nz) return false;
      double2 *P = lds + (u & 1) * planeL;
      if (active) P[ty * nx + tx] = q[u];
      const size_t zn = zs * wrapi(z + 1, nz);
      if constexpr ((u & 1) == 0) {
        halo_put(P, hA0, hA1, hA2, hA3);
        halo_get(zn, hB0, hB1, hB2, hB3);
      } else {
        halo_put(P, hB0, hB1, hB2, hB3);
        halo_get(zn, hA0, hA1, hA2, hA3);
      }
      q[(u + R - 1) % R] = p[col + zs * wrapi(z + R - 1, nz)];
      __syncthreads();
      if (own) {
        const double2 c = q[u];
        r0 += c.x * c.x + c.y * c.y;
        const double cr = c.x, ci = -c.y;
#pragma unroll
        for (int n = 1; n <= 4; ++n) {
          const double2 px = P[ty * nx + wrapi(tx + n, nx)];
          const double2 py = P[(ty + n) * nx + tx];
          const double2 pz = q[(u + n) % R];
          const double ax = a.coef_nab[n - 1] * cr, bx = a.coef_nab[n - 1] * ci;
          const double ay = a.coef_nab[3 + n] * cr, by = a.coef_nab[3 + n] * ci;
          const double az = a.coef_nab[7 + n] * cr, bz = a.coef_nab[7 + n] * ci;
          sx += ax * px.y + bx * px.x;
          sy += ay * py.y + by * py.x;
          sz += az * pz.y + bz * pz.x;
        }
      }
      return true;
    });
  }
  r0 = block_sum(r0, sh);
  sx = block_sum(sx, sh);
  sy = block_sum(sy, sh);
  sz = block_sum(sz, sh);
  if (tid == 0) {
    double *o = a.part + 4 * ((size_t)ytile + (size_t)a.nyt * b);
    o[0] = r0;
    o[1] = sx;
    o[2] = sy;
    o[3] = sz;
  }
}

__global__ __launch_bounds__(kBlock) void k_current_local(CurArgs a) { current_local_body(a, blockIdx.x); }

struct CurFinArgs {
  const double *part;      // (4, nslab, norb)
  const double2 *nl;       // (4, nlma, nb, nk) raw moments, or null
  const double *rinv;      // (nlma)
  const double *veck;      // (3, nk)
  const double *rocc;      // (no, nk, nspin)
  const double *wtk;       // (nk)
  double *contrib;         // (3, no*nk, nspin) scratch
  double *curr;            // (3, nspin)
  double ax, ay, az;
  const double *actab;     // with actab: vec_Ac = actab(:, *itt) (Ac_tot at the device step counter)
  const int *itt;
  int tmax;                // last row of actab (nt + 1); the step counter is clamped to 0..tmax
  double inv_ngrid;
  int nslab, nspin, no, nk, nlma;
};

// One wave per orbital (ik, io) of spin ispin: its weighted current
// ((j1 + j2) + jNL) rocc wtk (calc_current, density_matrix.f90:405-450), with
// jNL = 2 sum_lma Im(conj(uVpsi_r) uVpsi) (calc_current_nonlocal :532-591)
// reduced across the wave in a fixed tree.
__device__ __forceinline__ void current_orb_body(const CurFinArgs &a, int bx, int ispin) {
  const int M = a.no * a.nk;
  const int mm = bx * (kBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (mm >= M) return;
  const int io = mm % a.no, ik = mm / a.no;
  double jw[3] = {0.0, 0.0, 0.0};
  if (a.nl) {
    const int nb = a.no * a.nspin;
    const size_t ob = (size_t)ik * nb + (size_t)(ispin + a.nspin * io);
    for (int l = lane; l < a.nlma; l += 64) {
      const double2 *u4 = a.nl + (ob * a.nlma + l) * 4;
      const double rv = a.rinv[l];
      const double ur = u4[0].x * rv, ui = u4[0].y * rv;
      for (int d = 0; d < 3; ++d) jw[d] += u4[1 + d].x * ui - u4[1 + d].y * ur;
    }
    for (int d = 0; d < 3; ++d)
      for (int off = 32; off > 0; off >>= 1) jw[d] += __shfl_down(jw[d], off, 64);
  }
  if (lane == 0) {
    const size_t b = (size_t)ispin + (size_t)a.nspin * mm;
    double rt = 0.0, s[3] = {0.0, 0.0, 0.0};
    for (int sl = 0; sl < a.nslab; ++sl) {
      const double *o = a.part + 4 * ((size_t)sl + (size_t)a.nslab * b);
      rt = rt + o[0];
      s[0] = s[0] + o[1];
      s[1] = s[1] + o[2];
      s[2] = s[2] + o[3];
    }
    double A[3] = {a.ax, a.ay, a.az};
    if (a.actab) {
      const int t = min(max(*a.itt, 0), a.tmax);
      A[0] = a.actab[3 * t];
      A[1] = a.actab[3 * t + 1];
      A[2] = a.actab[3 * t + 2];
    }
    const double kA[3] = {a.veck[3 * ik] + A[0], a.veck[3 * ik + 1] + A[1], a.veck[3 * ik + 2] + A[2]};
    const double occ = a.rocc[io + (size_t)a.no * (ik + (size_t)a.nk * ispin)];
    for (int d = 0; d < 3; ++d) {
      const double j1 = kA[d] * rt;
      const double j2 = s[d] * 2.0;
      a.contrib[3 * ((size_t)mm + (size_t)M * ispin) + d] = (((j1 + j2) + jw[d] * 2.0) * occ) * a.wtk[ik];
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_current_orb(CurFinArgs a, int ispin) { current_orb_body(a, blockIdx.x, ispin); }

// Sum of the orbital contributions in (ik, io) order, divided by N_grid
// (contributions staged through LDS in chunks so the serial sum reads LDS).
__device__ __forceinline__ void current_sum_body(const CurFinArgs &a, int ispin) {
  __shared__ double sh[3 * 1024];
  const int M = a.no * a.nk;
  const double *cs = a.contrib + 3 * (size_t)M * ispin;  // this spin's contributions
  double w4 = 0.0;
  for (int m0 = 0; m0 < M; m0 += 1024) {
    const int n = min(1024, M - m0);
    for (int i = threadIdx.x; i < 3 * n; i += kBlock) sh[i] = cs[3 * (size_t)m0 + i];
    __syncthreads();
    if (threadIdx.x < 3)
      for (int mm = 0; mm < n; ++mm) w4 = w4 + sh[3 * mm + threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x < 3) a.curr[3 * ispin + threadIdx.x] = w4 * a.inv_ngrid;
}

__global__ __launch_bounds__(kBlock) void k_current_sum(CurFinArgs a, int ispin) { current_sum_body(a, ispin); }

// ---------------------------------------------------------------------------
// Dipole: out[0:3] = -sum(pos_d rho) Hgs_d Hvol, out[3] = sum(rho) Hvol
// pos = i (odd n) or i - 1/2 (even n), i the 1-based grid index.
// ---------------------------------------------------------------------------
// Dipole moment and electron count (calc_dip / subdip, src/rt/dip.f90:114-225): per block a
// strided share of the points (32-bit indices), tree-reduced in LDS into part(4, block); the
// final kernel sums the block partials in block order.  Positions are the grid indices with
// the reference's half-cell shift on even extents.
constexpr int kDipBlocks = 128;

struct DipArgs {
  const double *rho;
  double *part, *out;    // part(4, kDipBlocks); out: dipole (3) and electron count
  int nx, ny, nz, nspin;
  double hx, hy, hz, hvol;
};

__device__ __forceinline__ void dipole_part_body(const DipArgs &a, int bx) {
  const double *rho = a.rho;
  double *part = a.part;
  const int nx = a.nx, ny = a.ny, nz = a.nz, nspin = a.nspin;
  __shared__ double sh[4][256];
  const int ngrid = nx * ny * nz;
  const double sx = (nx % 2 == 1) ? 1.0 : 0.5, sy = (ny % 2 == 1) ? 1.0 : 0.5, sz = (nz % 2 == 1) ? 1.0 : 0.5;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int is = 0; is < nspin; ++is)
    for (int g = bx * 256 + threadIdx.x; g < ngrid; g += 256 * kDipBlocks) {
      const int x = g % nx;
      const int yz = g / nx;
      const int y = yz % ny;
      const int z = yz / ny;
      const double r = rho[(size_t)g + (size_t)ngrid * is];
      acc[0] += ((double)x + sx) * r;
      acc[1] += ((double)y + sy) * r;
      acc[2] += ((double)z + sz) * r;
      acc[3] += r;
    }
  for (int d = 0; d < 4; ++d) sh[d][threadIdx.x] = acc[d];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int d = 0; d < 4; ++d) sh[d][threadIdx.x] = sh[d][threadIdx.x] + sh[d][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 4) part[4 * bx + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_dipole_part(DipArgs a) { dipole_part_body(a, blockIdx.x); }

// the final sum, on the first kDipBlocks threads of the block
__device__ __forceinline__ void dipole_sum_body(const DipArgs &a) {
  __shared__ double sh[4][kDipBlocks];
  const bool on = threadIdx.x < kDipBlocks;
  for (int d = 0; d < 4; ++d)
    if (on) sh[d][threadIdx.x] = a.part[4 * threadIdx.x + d];
  __syncthreads();
  for (int s = kDipBlocks / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int d = 0; d < 4; ++d) sh[d][threadIdx.x] = sh[d][threadIdx.x] + sh[d][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.out[0] = -sh[0][0] * a.hx * a.hvol;
    a.out[1] = -sh[1][0] * a.hy * a.hvol;
    a.out[2] = -sh[2][0] * a.hz * a.hvol;
    a.out[3] = sh[3][0] * a.hvol;
  }
}

__global__ __launch_bounds__(256) void k_dipole_sum(DipArgs a) { dipole_sum_body(a); }

// ---------------------------------------------------------------------------
// Hartree potential: poisson_ft (src/poisson/poisson_periodic.f90:23-242) as six
// separable DFT passes (forward z, y, x with conj twiddles, 1/N and the 4 pi/|G|^2
// table folded into the x pass; inverse z, y, x, real part out).  Per pass the grid
// is viewed as (outer, n, inner), inner = stride of the transformed axis; a block
// owns `cols` consecutive columns c = (o, i), stages the [n][cols] tile in LDS and
// forms out(k) = sum_j tw(k, j) in(j) with the n x n table tw = exp(i 2 pi j k / n)
// (src/common/initialization.f90:761-784).  Reads and writes are unit-stride along
// c for the y and z passes and along j for the x pass.
// ---------------------------------------------------------------------------
constexpr int kDftMaxN = 256;

struct DftArgs {
  const double2 *in;  // complex input, or null with rin
  const double *rin;  // real input (first forward pass)
  const double *rin2; // with rin: a second real input added to it (the other spin channel), or null
  double2 *out;       // complex output, or null with rout
  double *rout;       // real part output (last inverse pass)
  const double2 *tw;  // n x n, tw[k * n + j]
  const double *coef; // per-point multiplier after the pass (x forward pass), or null
  double nrm;         // divisor applied before coef
  long long ncols;    // outer * inner
  int n, inner, cols;
  double sg;          // -1: conj(tw) (forward), +1: tw (inverse)
  const int *skip;    // device flag: the launch does nothing when it is set (null: never)
};

__global__ __launch_bounds__(256) void k_dft_axis(DftArgs a) {
  extern __shared__ double2 dft_tile[];  // [n][cols]
  if (a.skip && *a.skip) return;  // a device-side gate switched this launch off
  const int n = a.n, inner = a.inner, cols = a.cols;
  const long long c0 = (long long)blockIdx.x * cols;
  const int ncl = (int)min((long long)cols, a.ncols - c0);
  const int tot = n * cols;
  for (int t = threadIdx.x; t < tot; t += blockDim.x) {
    int j, cl;
    if (inner == 1) { cl = t / n; j = t - cl * n; } else { j = t / cols; cl = t - j * cols; }
    double2 v = make_double2(0.0, 0.0);
    if (cl < ncl) {
      const long long c = c0 + cl;
      const long long o = c / inner, i = c - o * inner;
      const size_t g = ((size_t)o * n + j) * inner + i;
      v = a.rin ? make_double2(a.rin2 ? a.rin[g] + a.rin2[g] : a.rin[g], 0.0) : a.in[g];
    }
    dft_tile[j * cols + cl] = v;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < tot; t += blockDim.x) {
    int k, cl;
    if (inner == 1) { cl = t / n; k = t - cl * n; } else { k = t / cols; cl = t - k * cols; }
    if (cl >= ncl) continue;
    const double2 *w = a.tw + (size_t)k * n;
    double sr = 0.0, si = 0.0;
    int j = 0;
    // 8 twiddle loads in flight ahead of the in-order sum
    for (; j + 8 <= n; j += 8) {
      double2 e[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) e[u] = w[j + u];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double2 x = dft_tile[(j + u) * cols + cl];
        const double ei = a.sg * e[u].y;
        sr = sr + (e[u].x * x.x - ei * x.y);
        si = si + (e[u].x * x.y + ei * x.x);
      }
    }
    for (; j < n; ++j) {
      const double2 e = w[j];
      const double2 x = dft_tile[j * cols + cl];
      const double ei = a.sg * e.y;
      sr = sr + (e.x * x.x - ei * x.y);
      si = si + (e.x * x.y + ei * x.x);
    }
    const long long c = c0 + cl;
    const long long o = c / inner, i = c - o * inner;
    const size_t g = ((size_t)o * n + k) * inner + i;
    if (a.coef) {
      const double cf = a.coef[g];
      sr = cf * (sr / a.nrm);
      si = cf * (si / a.nrm);
    }
    if (a.rout) a.rout[g] = sr;
    else a.out[g] = make_double2(sr, si);
  }
}

// PZ81 LDA, unpolarised (exc_cor_pz / PZxc, src/xc/builtin_pz.f90:23-83), trho = rho.
__device__ __forceinline__ void pz_point(double rho, double &vxc, double &exc) {
  const double Pi = 3.141592653589793;
  const double gammaU = -0.1423, beta1U = 1.0529, beta2U = 0.3334, AU = 0.0311, BU = -0.048;
  const double CU = 0.002, DU = -0.0116;
  const double cst = 0.75 * pow(3.0 / (2.0 * Pi), 2.0 / 3.0);
  const double trho = 2.0 * (rho * 0.5);
  const double ttrho = trho + 1e-10;
  const double rs = pow(3.0 / (4 * Pi * ttrho), 1.0 / 3.0);
  double e = -cst / rs;
  double de = e / (3 * ttrho);
  if (rs > 1.0) {
    const double rssq = sqrt(rs);
    const double den = 1.0 + beta1U * rssq + beta2U * rs;
    e = e + gammaU / den;
    de = de + gammaU * (0.5 * beta1U * rssq + beta2U * rs) / (3 * ttrho) / (den * den);
  } else {
    const double rsln = log(rs);
    e = e + AU * rsln + BU + CU * rs * rsln + DU * rs;
    de = de - rs / (3 * ttrho) * (AU / rs + CU * (rsln + 1) + DU);
  }
  vxc = e + trho * de;
  exc = e;
}

// PZ81 LDA, spin-polarised (exc_cor_pz_sp / PZxc_sp, src/xc/builtin_pz_sp.f90:23-98) at one
// point from the two channel densities; terms in the order of the source's expressions.
__device__ __forceinline__ void pz_sp_point(double ru, double rd, double &vu, double &vd, double &exc) {
  const double Pi = 3.141592653589793;
  const double gammaU = -0.1423, gammaP = -0.0843;
  const double beta1U = 1.0529, beta2U = 0.3334, beta1P = 1.3981, beta2P = 0.2611;
  const double AU = 0.0311, BU = -0.048, CU = 0.002, DU = -0.0116;
  const double AP = 0.01555, BP = -0.0269, CP = 0.0007, DP = -0.0048;
  const double cx = 3.0 / 4.0 * pow(3.0 / (2 * Pi), 2.0 / 3.0);
  const double c13 = -pow(2.0, 1.0 / 3.0) + 1.0;
  const double fden = pow(2.0, 4.0 / 3.0) - 2.0;
  const double trho = ru + rd + 1e-20;
  const double tu = ru + 5e-21, td = rd + 5e-21;
  const double rs = pow(3.0 / (4.0 * Pi * trho), 1.0 / 3.0);
  const double zeta = (tu - td) / trho;
  const double sf = (pow(1.0 + zeta, 4.0 / 3.0) + pow(1.0 - zeta, 4.0 / 3.0) - 2.0) / fden;
  const double dsf = 4.0 / 3.0 * (pow(1.0 + zeta, 1.0 / 3.0) - pow(1.0 - zeta, 1.0 / 3.0)) / fden;
  double e = -cx / rs + c13 * cx / rs * sf;
  const double vs = e - 1.0 / 3.0 * cx / rs + 1.0 / 3.0 * c13 * cx / rs * sf;  // common to both channels
  double v0 = vs + c13 * cx / rs * (1.0 - zeta) * dsf;
  double v1 = vs + c13 * cx / rs * (-1.0 - zeta) * dsf;
  double ecu, ecp, vcu, vcp;
  if (rs > 1.0) {
    const double sq = sqrt(rs);
    const double du = 1.0 + beta1U * sq + beta2U * rs, dp = 1.0 + beta1P * sq + beta2P * rs;
    ecu = gammaU / du;
    ecp = gammaP / dp;
    vcu = gammaU * (1.0 + 7.0 / 6.0 * beta1U * sq + 4.0 / 3.0 * beta2U * rs) / (du * du);
    vcp = gammaP * (1.0 + 7.0 / 6.0 * beta1P * sq + 4.0 / 3.0 * beta2P * rs) / (dp * dp);
  } else {
    const double L = log(rs);
    ecu = AU * L + BU + CU * rs * L + DU * rs;
    ecp = AP * L + BP + CP * rs * L + DP * rs;
    vcu = AU * L + (BU - AU / 3.0) + 2.0 / 3.0 * CU * rs * L + (2.0 * DU - CU) / 3.0 * rs;
    vcp = AP * L + (BP - AP / 3.0) + 2.0 / 3.0 * CP * rs * L + (2.0 * DP - CP) / 3.0 * rs;
  }
  e = e + ecu + sf * (ecp - ecu);
  vu = v0 + vcu + sf * (vcp - vcu) + (ecp - ecu) * (1.0 - zeta) * dsf;
  vd = v1 + vcu + sf * (vcp - vcu) + (ecp - ecu) * (-1.0 - zeta) * dsf;
  exc = e;
}

// nspin = 1: rho, vxc (n); nspin = 2: rho_s, vxc_s (n, 2) channel-major; exc (n) or null
__global__ __launch_bounds__(256) void k_xc_pz(const double *rho, double *vxc, double *exc, size_t n, int nspin) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double e;
  if (nspin == 2) {
    double vu, vd;
    pz_sp_point(rho[i], rho[n + i], vu, vd, e);
    vxc[i] = vu;
    vxc[n + i] = vd;
  } else {
    double v;
    pz_point(rho[i], v, e);
    vxc[i] = v;
  }
  if (exc) exc[i] = e;
}

// update_vlocal (src/common/hamiltonian.f90:813-843): V(is) = Vpsl + Vh + Vxc(is), with Vxc
// either given or formed from rho on the fly (PZ, or spin-polarised PZ from the two channel
// densities); optionally the predictor-corrector average V(is) = (Vmix(is) + V(is))/2
// (src/rt/time_evolution_step.f90:404-409).  vpsl, vh: one channel; vxc, rho, vmix, V, vout:
// nspin channels of n points.
struct VlocArgs {
  const double *vpsl, *vh, *vxc, *rho, *vmix;
  double *V, *vout;
  size_t n;
  int nspin;
  const int *skip;    // device flag: the launch does nothing when it is set (null: never)
};

__global__ __launch_bounds__(256) void k_update_vlocal(VlocArgs a) {
  if (a.skip && *a.skip) return;  // a device-side gate switched this launch off
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double vx[2];
  if (a.rho) {
    double e;
    if (a.nspin == 2) pz_sp_point(a.rho[i], a.rho[a.n + i], vx[0], vx[1], e);
    else pz_point(a.rho[i], vx[0], e);
  } else {
    vx[0] = a.vxc[i];
    if (a.nspin == 2) vx[1] = a.vxc[a.n + i];
  }
  const double base = a.vpsl[i] + a.vh[i];
  for (int is = 0; is < a.nspin; ++is) {
    const size_t j = (size_t)is * a.n + i;
    double v = base + vx[is];
    if (a.vmix) v = 0.5 * (a.vmix[j] + v);
    a.V[j] = v;
    if (a.vout) a.vout[j] = v;
  }
}

// The aetrs propagator's potential extrapolation (src/rt/time_evolution_step.f90:432-445):
// V_new = 3 V - 3 V_old1 + V_old2, then V_old2 = V_old1, V_old1 = V, V = V_new, over all
// spin channels of the handle's potential.
__global__ __launch_bounds__(256) void k_vextrap(double *V, double *v1, double *v2, double *vout, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = V[i];
  const double vn = 3.0 * v - 3.0 * v1[i] + v2[i];
  v2[i] = v1[i];
  v1[i] = v;
  V[i] = vn;
  if (vout) vout[i] = vn;
}

// Streaming copy of n 16-byte elements (HBM ceiling of the box): dwordx4 loads and stores,
// non-temporal, four elements per thread in flight.
typedef unsigned int u4v __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_stream_copy(u4v *__restrict__ dst, const u4v *__restrict__ src, size_t n) {
  const size_t stride = (size_t)gridDim.x * 256;
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    u4v v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = __builtin_nontemporal_load(src + i + k * stride);
#pragma unroll
    for (int k = 0; k < 4; ++k) __builtin_nontemporal_store(v[k], dst + i + k * stride);
  }
  for (; i < n; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// dst = src where the device flag is set (the accepted branch of a device-side gate)
__global__ __launch_bounds__(256) void k_copy_if(const int *flag, double *dst, const double *src, size_t n) {
  if (!*flag) return;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

template <typename T>
int dalloc(T **p, size_t n) {
  *p = nullptr;
  if (n == 0) return 0;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(SALMON_HIP_ENOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
  }
  return 0;
}

template <typename T>
void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

}  // namespace

struct salmon_hip_ctx {
  int nx = 0, ny = 0, nz = 0, ly = 0, nspin = 0, no = 0, nk = 0;
  int nb = 0, norb = 0;
  size_t osz = 0, ngrid = 0;
  double coef_lap[12] = {}, coef_nab[12] = {}, coef_lap0 = 0, hgs[3] = {}, hvol = 0;
  hipStream_t own = nullptr, stream = nullptr;
  // projections run on `side`, forked from and joined back into `stream` by events, so
  // that they overlap the stencil pass reading the same input
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool join_pending = false;
  bool overlap = false;  // salmon_hip_set_overlap
  bool have_V = false, have_k = false;
  double *d_veck = nullptr, *d_rocc = nullptr, *d_wtk = nullptr, *d_V = nullptr, *d_kpar = nullptr;
  // nonlocal
  int natom = 0, nps = 0, nlma = 0, nent = 0;
  int cmax = 0;
  int *d_mps = nullptr, *d_jidx = nullptr, *d_lma_atom = nullptr, *d_pt_off = nullptr;
  int *d_pt_lma = nullptr, *d_entof = nullptr;
  int4 *d_tp_info = nullptr;  // per touched point: orbital-block offset, first entry, entry count, its pt_lma
  int ntp = 0;
  // fused KB pass (k_proj_scatter): per sphere point its offset if one (atom, point) pair covers
  // it, else -1; the touched points with several pairs (completed by k_nl_scatter first)
  int *d_jexcl = nullptr;
  int4 *d_tp_shared = nullptr;
  int ntp_shared = 0;
  bool kb_fused = true;
  double2 *d_sink = nullptr;
  int2 *d_atom_lma = nullptr;
  double *d_rxyz = nullptr, *d_uv = nullptr, *d_rinv = nullptr;
  double2 *d_zekr = nullptr, *d_zent = nullptr, *d_c = nullptr, *d_curnl = nullptr;
  // salmon_hip_current_next: the next step's phase table built ahead into d_*_n (staged, for the
  // field staged_A) and the projection of that step's orbitals pre_psi under it left in d_c;
  // pre_state 1 = made, table not yet promoted; 2 = promoted by update_kvector(staged_A), d_c
  // holds c of pre_psi for its first Taylor term; 0 = none
  double2 *d_zekr_n = nullptr, *d_zent_n = nullptr;
  double *d_kpar_n = nullptr;
  bool staged = false;
  double staged_A[3] = {0.0, 0.0, 0.0};
  const void *pre_psi = nullptr;
  int pre_state = 0;
  bool in_batch = false;  // a batch's packed records hold this handle's table pointers
  double2 *d_c2 = nullptr;  // the fused pass writes the next term's c here; the two are then swapped
  // Taylor scratch: H^2 psi .. H^(N-2) psi (allocated on first use); FP32 chain h_1 .. h_{N-1}
  std::vector<double2 *> tscr;
  std::vector<float2 *> tscr32;
  int prec = SALMON_HIP_PREC_FP64;
  // observables scratch
  double *d_curpart = nullptr, *d_contrib = nullptr, *d_denspart = nullptr;
  double *d_scratch = nullptr;  // this device's copy of `scratch` (salmon_hip_density_noalloc)
  double *d_dippart = nullptr;  // dipole block partials (4 x kDipBlocks)
  int dens_ngrp = 1;
  // Hartree (salmon_hip_set_poisson): twiddle tables per axis, 4 pi/|G|^2 table, work
  bool have_cell = false;
  double2 *d_tw[3] = {nullptr, nullptr, nullptr};
  double *d_gcoef = nullptr, *d_vh = nullptr;
  double2 *d_fft[2] = {nullptr, nullptr};
  // stencil launch shape
  int TY = 1, TZ = 1, nyt = 1, nzt = 1;
  // launch shape of the terms without the Taylor sum (salmon_hip_autotune): tile_plane -1 = rule
  // (whole plane when it fits 256 threads), 0 = y tiles, 1 = whole plane; mid_nzt 0 = rule
  int tile_plane = -1, mid_nzt = 0;
  bool tile8 = true;  // stencil_tile8 for the mixed middle terms where the shape allows it
  bool tilez = true;  // stencil_tilez for the FP64 ones and the mixed first term (before stencil_tile8)
  // salmon_hip_set_skip: device flag read by the launches of hpsi/taylor, density and the
  // potential update (they do nothing while it is set); null = none
  const int *d_skip = nullptr;
  // kernel timing (salmon_hip_prof_*): start/stop event pairs per launch
  bool prof = false;
  std::vector<hipEvent_t> ev;
  std::vector<int> ev_kind;
  size_t ev_used = 0;
};

// Events for one timed launch, or nullptr pair when timing is off / the pool is full.
static void prof_events(salmon_hip_ctx *c, int kind, hipEvent_t *s, hipEvent_t *e) {
  *s = nullptr;
  *e = nullptr;
  if (!c->prof || 2 * (c->ev_used + 1) > c->ev.size()) return;
  *s = c->ev[2 * c->ev_used];
  *e = c->ev[2 * c->ev_used + 1];
  c->ev_kind[c->ev_used] = kind;
  c->ev_used += 1;
}

static int check(salmon_hip_t h) {
  if (!h) return fail(SALMON_HIP_EINVAL, "null handle");
  return 0;
}

// Per-device address of `scratch`, resolved once per device with that device current.
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_dev[kMaxDevices] = {};

static int resolve_scratch(double **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(SALMON_HIP_EHIP, "hipGetDevice failed or device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(SALMON_HIP_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_dev[dev] = (double *)ptr;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

// Error text from the other translation units of the library (the predictor).
void salmon_hip_internal_error(const char *msg) { g_err = msg; }

extern "C" {

const char *salmon_hip_last_error(void) { return g_err.c_str(); }

int salmon_hip_create(int nx, int ny, int nz, int ly, int nspin, int no, int nk, const double *coef_lap,
                      const double *coef_nab, double coef_lap0, const double *vec_k, const double *rocc,
                      const double *wtk, const double *hgs, double hvol, salmon_hip_t *out) {
  if (!out) return fail(SALMON_HIP_EINVAL, "out is null");
  *out = nullptr;
  if (nx < 4 || ny < 4 || nz < 4) return fail(SALMON_HIP_EINVAL, "grid %dx%dx%d below Nd=4", nx, ny, nz);
  if (nx > kBlock) return fail(SALMON_HIP_EUNSUPPORTED, "nx=%d above %d", nx, kBlock);
  if (ly < ny) return fail(SALMON_HIP_EINVAL, "ly=%d < ny=%d", ly, ny);
  if ((size_t)nx * ly * nz > (size_t)INT_MAX) return fail(SALMON_HIP_EUNSUPPORTED, "grid of %dx%dx%d above 2^31 points", nx, ly, nz);
  if (nspin != 1 && nspin != 2) return fail(SALMON_HIP_EINVAL, "nspin=%d", nspin);
  if (no < 1 || nk < 1) return fail(SALMON_HIP_EINVAL, "no=%d nk=%d", no, nk);
  if (!coef_lap || !coef_nab || !vec_k || !rocc || !wtk || !hgs) return fail(SALMON_HIP_EINVAL, "null input");
  salmon_hip_ctx *c = new salmon_hip_ctx();
  c->nx = nx; c->ny = ny; c->nz = nz; c->ly = ly;
  c->nspin = nspin; c->no = no; c->nk = nk;
  c->nb = nspin * no;
  c->norb = c->nb * nk;
  c->osz = (size_t)nx * ly * nz;
  c->ngrid = (size_t)nx * ny * nz;
  memcpy(c->coef_lap, coef_lap, sizeof(c->coef_lap));
  memcpy(c->coef_nab, coef_nab, sizeof(c->coef_nab));
  c->coef_lap0 = coef_lap0;
  memcpy(c->hgs, hgs, sizeof(c->hgs));
  c->hvol = hvol;
  int rc = 0;
  auto bail = [&](int code) { salmon_hip_destroy(c); return code; };
  if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess)
    return bail(fail(SALMON_HIP_EHIP, "stream / event creation failed"));
  c->stream = c->own;
  if ((rc = dalloc(&c->d_veck, 3 * (size_t)nk))) return bail(rc);
  if ((rc = dalloc(&c->d_rocc, (size_t)no * nk * nspin))) return bail(rc);
  if ((rc = dalloc(&c->d_wtk, (size_t)nk))) return bail(rc);
  if ((rc = dalloc(&c->d_V, c->ngrid * nspin))) return bail(rc);
  if ((rc = dalloc(&c->d_kpar, 16 * (size_t)nk))) return bail(rc);
  if (hipMemcpy(c->d_veck, vec_k, sizeof(double) * 3 * nk, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_rocc, rocc, sizeof(double) * no * nk * nspin, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_wtk, wtk, sizeof(double) * nk, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(SALMON_HIP_EHIP, "upload of k-point tables failed"));
  // stencil tile shape: full x rows, TY rows, z chunks so that the launch has
  // at least ~2048 blocks where the orbital count allows it
  c->TY = std::min(ny, std::max(1, kBlock / nx));
  c->nyt = (ny + c->TY - 1) / c->TY;
  int want = 2048;
  int nzt = (want + c->norb * c->nyt - 1) / (c->norb * c->nyt);
  nzt = std::max(1, std::min(nzt, std::max(1, nz / 8)));
  c->TZ = (nz + nzt - 1) / nzt;
  c->nzt = (nz + c->TZ - 1) / c->TZ;
  // observables scratch
  c->dens_ngrp = (no * nk + kDensGroup - 1) / kDensGroup;
  if (c->dens_ngrp > 1 && (rc = dalloc(&c->d_denspart, c->ngrid * nspin * c->dens_ngrp))) return bail(rc);
  if ((rc = dalloc(&c->d_curpart, 4 * (size_t)c->nyt * c->norb))) return bail(rc);
  // per spin: the batched step forms both spins' orbital contributions in one launch
  if ((rc = dalloc(&c->d_contrib, 3 * (size_t)no * nk * nspin))) return bail(rc);
  if ((rc = dalloc(&c->d_sink, 1024))) return bail(rc);  // one element per thread of the largest tile
  if ((rc = dalloc(&c->d_dippart, 4 * (size_t)kDipBlocks))) return bail(rc);
  if ((rc = resolve_scratch(&c->d_scratch))) return bail(rc);
  *out = c;
  return 0;
}

int salmon_hip_destroy(salmon_hip_t c) {
  if (!c) return 0;
  if (c->own) (void)hipStreamSynchronize(c->own);
  if (c->stream && c->stream != c->own) (void)hipStreamSynchronize(c->stream);
  dfree(c->d_veck); dfree(c->d_rocc); dfree(c->d_wtk); dfree(c->d_V); dfree(c->d_kpar); dfree(c->d_kpar_n);
  dfree(c->d_mps); dfree(c->d_jidx); dfree(c->d_lma_atom); dfree(c->d_pt_off);
  dfree(c->d_atom_lma); dfree(c->d_pt_lma); dfree(c->d_entof); dfree(c->d_rxyz); dfree(c->d_uv);
  dfree(c->d_rinv); dfree(c->d_zekr); dfree(c->d_zent); dfree(c->d_c); dfree(c->d_curnl);
  dfree(c->d_zekr_n); dfree(c->d_zent_n);
  c->staged = false;
  c->pre_state = 0;
  dfree(c->d_c2); dfree(c->d_jexcl); dfree(c->d_tp_shared);
  dfree(c->d_tp_info); dfree(c->d_sink);
  dfree(c->d_curpart); dfree(c->d_contrib); dfree(c->d_denspart); dfree(c->d_dippart);
  for (int d = 0; d < 3; ++d) dfree(c->d_tw[d]);
  dfree(c->d_gcoef); dfree(c->d_vh); dfree(c->d_fft[0]); dfree(c->d_fft[1]);
  for (double2 *&p : c->tscr) dfree(p);
  for (float2 *&p : c->tscr32) dfree(p);
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  if (c->side) (void)hipStreamSynchronize(c->side);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->own) (void)hipStreamDestroy(c->own);
  delete c;
  return 0;
}

int salmon_hip_set_stream(salmon_hip_t c, void *stream) {
  if (int rc = check(c)) return rc;
  c->stream = (hipStream_t)stream;  // NULL: HIP's null stream (ordered with every blocking stream)
  return 0;
}

void *salmon_hip_get_stream(salmon_hip_t c) { return c ? (void *)c->stream : nullptr; }

int salmon_hip_set_nonlocal(salmon_hip_t c, int natom, int nps, const int *mps, const int *jxyz,
                            const double *rxyz, int nlma, const int *ia_tbl, const double *rinv_uvu,
                            const double *uv) {
  if (int rc = check(c)) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  dfree(c->d_mps); dfree(c->d_jidx); dfree(c->d_lma_atom); dfree(c->d_pt_off);
  dfree(c->d_atom_lma); dfree(c->d_pt_lma); dfree(c->d_entof); dfree(c->d_rxyz); dfree(c->d_uv);
  dfree(c->d_rinv); dfree(c->d_zekr); dfree(c->d_zent); dfree(c->d_c); dfree(c->d_curnl);
  dfree(c->d_zekr_n); dfree(c->d_zent_n);
  c->staged = false;
  c->pre_state = 0;
  dfree(c->d_c2); dfree(c->d_jexcl); dfree(c->d_tp_shared);
  dfree(c->d_tp_info);
  c->natom = c->nps = c->nlma = c->nent = c->cmax = c->ntp = c->ntp_shared = 0;
  c->have_k = false;
  if (natom == 0 || nlma == 0) return 0;
  if (natom < 0 || nps < 1 || nlma < 0 || !mps || !jxyz || !rxyz || !ia_tbl || !rinv_uvu || !uv)
    return fail(SALMON_HIP_EINVAL, "invalid projector tables");
  // validate: sphere sizes and grid indices, projectors contiguous per atom
  std::vector<int2> atom_lma(natom, make_int2(0, 0));
  std::vector<int> lma_atom(nlma);
  for (int l = 0; l < nlma; ++l) {
    const int ia = ia_tbl[l] - 1;
    if (ia < 0 || ia >= natom) return fail(SALMON_HIP_EINVAL, "ia_tbl(%d)=%d out of range", l + 1, ia + 1);
    if (l > 0 && ia < ia_tbl[l - 1] - 1) return fail(SALMON_HIP_EUNSUPPORTED, "ia_tbl not atom-ordered");
    if (atom_lma[ia].y == 0) atom_lma[ia].x = l;
    else if (atom_lma[ia].x + atom_lma[ia].y != l)
      return fail(SALMON_HIP_EUNSUPPORTED, "projectors of atom %d not contiguous", ia + 1);
    atom_lma[ia].y += 1;
    lma_atom[l] = ia;
  }
  for (int ia = 0; ia < natom; ++ia)
    if (atom_lma[ia].y > kMaxLmaPerAtom)
      return fail(SALMON_HIP_EUNSUPPORTED, "atom %d has %d projectors (max %d)", ia + 1, atom_lma[ia].y,
                  kMaxLmaPerAtom);
  if (nlma >= (1 << 24)) return fail(SALMON_HIP_EUNSUPPORTED, "nlma too large");
  std::vector<int> jidx((size_t)natom * nps, 0);
  std::vector<int> cnt(c->ngrid + 1, 0);
  for (int ia = 0; ia < natom; ++ia) {
    if (mps[ia] < 0 || mps[ia] > nps) return fail(SALMON_HIP_EINVAL, "mps(%d)=%d", ia + 1, mps[ia]);
    for (int j = 0; j < mps[ia]; ++j) {
      const int *p = jxyz + 3 * ((size_t)j + (size_t)nps * ia);
      const int x = p[0] - 1, y = p[1] - 1, z = p[2] - 1;
      if (x < 0 || x >= c->nx || y < 0 || y >= c->ny || z < 0 || z >= c->nz)
        return fail(SALMON_HIP_EINVAL, "jxyz(:,%d,%d) outside the grid", j + 1, ia + 1);
      jidx[(size_t)nps * ia + j] = x + c->nx * (y + c->ly * z);
      const size_t g = (size_t)x + (size_t)c->nx * ((size_t)y + (size_t)c->ny * z);
      if (atom_lma[ia].y > 0) cnt[g + 1] += 1;
    }
  }
  for (size_t g = 0; g < c->ngrid; ++g) cnt[g + 1] += cnt[g];
  const int nent = cnt[c->ngrid];
  int cmax = 1;
  for (int ia = 0; ia < natom; ++ia) cmax = std::max(cmax, atom_lma[ia].y);
  // CSR point -> (atom, j) entries; atom-major fill keeps each point's entries
  // in ascending ilma, the order zpseudo adds them in
  std::vector<int> pt_lma(std::max(nent, 1), 0);
  std::vector<int> entof((size_t)natom * nps, 0);
  std::vector<int> fillp(cnt.begin(), cnt.end() - 1);
  for (int ia = 0; ia < natom; ++ia) {
    if (atom_lma[ia].y == 0) continue;
    for (int j = 0; j < mps[ia]; ++j) {
      const int *p = jxyz + 3 * ((size_t)j + (size_t)nps * ia);
      const size_t g = (size_t)(p[0] - 1) + (size_t)c->nx * ((size_t)(p[1] - 1) + (size_t)c->ny * (p[2] - 1));
      const int e = fillp[g]++;
      pt_lma[e] = atom_lma[ia].x | (atom_lma[ia].y << 24);
      entof[(size_t)nps * ia + j] = e;
    }
  }
  std::vector<int4> tp_info;
  for (int z = 0; z < c->nz; ++z)
    for (int y = 0; y < c->ny; ++y)
      for (int x = 0; x < c->nx; ++x) {
        const size_t g = (size_t)x + (size_t)c->nx * ((size_t)y + (size_t)c->ny * z);
        if (cnt[g + 1] > cnt[g])
          tp_info.push_back(make_int4(x + c->nx * (y + c->ly * z), cnt[g], cnt[g + 1] - cnt[g], pt_lma[cnt[g]]));
      }
  // fused-pass tables: a sphere point's offset where its grid point has one entry, else -1;
  // the touched points with several entries
  std::vector<int> jexcl((size_t)natom * nps, -1);
  for (int ia = 0; ia < natom; ++ia) {
    if (atom_lma[ia].y == 0) continue;
    for (int j = 0; j < mps[ia]; ++j) {
      const int *p = jxyz + 3 * ((size_t)j + (size_t)nps * ia);
      const size_t g = (size_t)(p[0] - 1) + (size_t)c->nx * ((size_t)(p[1] - 1) + (size_t)c->ny * (p[2] - 1));
      if (cnt[g + 1] - cnt[g] == 1) jexcl[(size_t)nps * ia + j] = jidx[(size_t)nps * ia + j];
    }
  }
  std::vector<int4> tp_shared;
  for (const int4 &t : tp_info)
    if (t.z > 1) tp_shared.push_back(t);
  int rc = 0;
  if (!tp_info.empty()) {
    if ((rc = dalloc(&c->d_tp_info, tp_info.size()))) return rc;
    HIPCHK(hipMemcpy(c->d_tp_info, tp_info.data(), sizeof(int4) * tp_info.size(), hipMemcpyHostToDevice));
  }
  if (!tp_shared.empty()) {
    if ((rc = dalloc(&c->d_tp_shared, tp_shared.size()))) return rc;
    HIPCHK(hipMemcpy(c->d_tp_shared, tp_shared.data(), sizeof(int4) * tp_shared.size(), hipMemcpyHostToDevice));
  }
  if ((rc = dalloc(&c->d_jexcl, jexcl.size())) || (rc = dalloc(&c->d_c2, (size_t)nlma * c->norb))) return rc;
  HIPCHK(hipMemcpy(c->d_jexcl, jexcl.data(), sizeof(int) * jexcl.size(), hipMemcpyHostToDevice));
  if ((rc = dalloc(&c->d_mps, natom)) || (rc = dalloc(&c->d_jidx, (size_t)natom * nps)) ||
      (rc = dalloc(&c->d_lma_atom, nlma)) || (rc = dalloc(&c->d_pt_off, c->ngrid + 1)) ||
      (rc = dalloc(&c->d_atom_lma, natom)) || (rc = dalloc(&c->d_pt_lma, pt_lma.size())) ||
      (rc = dalloc(&c->d_entof, entof.size())) ||
      (rc = dalloc(&c->d_rxyz, 3 * (size_t)nps * natom)) || (rc = dalloc(&c->d_uv, (size_t)nps * nlma)) ||
      (rc = dalloc(&c->d_rinv, nlma)) || (rc = dalloc(&c->d_zekr, (size_t)nps * nlma * c->nk)) ||
      (rc = dalloc(&c->d_zent, (size_t)cmax * std::max(nent, 1) * c->nk)) ||
      (rc = dalloc(&c->d_c, (size_t)nlma * c->norb)) || (rc = dalloc(&c->d_curnl, 4 * (size_t)nlma * c->norb)))
    return rc;
  HIPCHK(hipMemcpy(c->d_mps, mps, sizeof(int) * natom, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_jidx, jidx.data(), sizeof(int) * jidx.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_lma_atom, lma_atom.data(), sizeof(int) * nlma, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_pt_off, cnt.data(), sizeof(int) * cnt.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_atom_lma, atom_lma.data(), sizeof(int2) * natom, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_pt_lma, pt_lma.data(), sizeof(int) * pt_lma.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_entof, entof.data(), sizeof(int) * entof.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_rxyz, rxyz, sizeof(double) * 3 * nps * natom, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_uv, uv, sizeof(double) * (size_t)nps * nlma, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_rinv, rinv_uvu, sizeof(double) * nlma, hipMemcpyHostToDevice));
  c->natom = natom;
  c->nps = nps;
  c->nlma = nlma;
  c->nent = nent;
  c->cmax = cmax;
  c->ntp = (int)tp_info.size();
  c->ntp_shared = (int)tp_shared.size();
  return 0;
}

int salmon_hip_set_vlocal(salmon_hip_t c, const double *vlocal) {
  if (int rc = check(c)) return rc;
  if (!vlocal) return fail(SALMON_HIP_EINVAL, "vlocal is null");
  HIPCHK(hipMemcpyAsync(c->d_V, vlocal, sizeof(double) * c->ngrid * c->nspin, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->have_V = true;
  return 0;
}

int salmon_hip_set_vlocal_device(salmon_hip_t c, const double *vlocal_dev) {
  if (int rc = check(c)) return rc;
  if (!vlocal_dev) return fail(SALMON_HIP_EINVAL, "vlocal is null");
  HIPCHK(hipMemcpyAsync(c->d_V, vlocal_dev, sizeof(double) * c->ngrid * c->nspin, hipMemcpyDeviceToDevice,
                        c->stream));
  c->have_V = true;
  return 0;
}

int salmon_hip_stream_copy(salmon_hip_t c, void *dst_dev, const void *src_dev, size_t bytes, int reps,
                           double *avg_ms) {
  if (int rc = check(c)) return rc;
  if (!dst_dev || !src_dev || bytes % 16 != 0 || bytes == 0 || reps < 1)
    return fail(SALMON_HIP_EINVAL, "stream_copy: null buffer, %zu bytes (a positive multiple of 16) or reps %d",
                bytes, reps);
  const size_t n = bytes / 16;
  const unsigned grid = (unsigned)std::min<size_t>((n + 4 * 256 - 1) / (4 * 256), 1u << 20);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, c->stream, (u4v *)dst_dev, (const u4v *)src_dev, n);
  int rc = 0;
  if (hipEventRecord(e0, c->stream) != hipSuccess) rc = fail(SALMON_HIP_EHIP, "event");
  for (int r = 0; r < reps && !rc; ++r)
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, c->stream, (u4v *)dst_dev, (const u4v *)src_dev, n);
  float ms = 0.f;
  if (!rc && (hipGetLastError() != hipSuccess || hipEventRecord(e1, c->stream) != hipSuccess ||
              hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
    rc = fail(SALMON_HIP_EHIP, "stream_copy timing");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (avg_ms) *avg_ms = (double)ms / reps;
  return 0;
}

int salmon_hip_set_skip(salmon_hip_t c, const int *flag_dev) {
  if (int rc = check(c)) return rc;
  c->d_skip = flag_dev;
  return 0;
}

int salmon_hip_copy_if(salmon_hip_t c, const int *flag_dev, double *dst_dev, const double *src_dev, size_t n) {
  if (int rc = check(c)) return rc;
  if (!flag_dev || !dst_dev || !src_dev) return fail(SALMON_HIP_EINVAL, "null pointer");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_copy_if, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, flag_dev, dst_dev,
                     src_dev, n);
  LAUNCHCHK("k_copy_if");
  return 0;
}

static KvecArgs kvec_args(salmon_hip_t c, const double *vec_Ac, const double *actab, const int *itt, int tmax) {
  KvecArgs a;
  a.veck = c->d_veck;
  a.rxyz = c->d_rxyz;
  a.uv = c->d_uv;
  a.mps = c->d_mps;
  a.lma_atom = c->d_lma_atom;
  a.atom_lma = c->d_atom_lma;
  a.entof = c->d_entof;
  a.zekr = c->d_zekr;
  a.zent = c->d_zent;
  a.nent = c->nent;
  a.cmax = c->cmax;
  a.kpar = c->d_kpar;
  memcpy(a.coef_nab, c->coef_nab, sizeof(a.coef_nab));
  a.coef_lap0 = c->coef_lap0;
  a.ax = vec_Ac ? vec_Ac[0] : 0.0;
  a.ay = vec_Ac ? vec_Ac[1] : 0.0;
  a.az = vec_Ac ? vec_Ac[2] : 0.0;
  a.actab = actab;
  a.itt = itt;
  a.tmax = tmax;
  a.nlma = c->nlma;
  a.nps = c->nps;
  a.nk = c->nk;
  return a;
}

static int update_kvector_impl(salmon_hip_t c, const double *vec_Ac, const double *actab, const int *itt,
                               int tmax = 0) {
  if (!actab && vec_Ac && c->staged && memcmp(vec_Ac, c->staged_A, sizeof(c->staged_A)) == 0) {
    // the table salmon_hip_current_next built ahead for this field becomes the active one
    std::swap(c->d_zekr, c->d_zekr_n);
    std::swap(c->d_zent, c->d_zent_n);
    std::swap(c->d_kpar, c->d_kpar_n);
    c->staged = false;
    if (c->pre_state == 1) c->pre_state = 2;
    c->have_k = true;
    return 0;
  }
  c->staged = false;
  c->pre_state = 0;
  const KvecArgs a = kvec_args(c, vec_Ac, actab, itt, tmax);
  dim3 grid(std::max(1, c->nlma), c->nk);
  hipLaunchKernelGGL(k_update_kvector, grid, dim3(kBlock), 0, c->stream, a);
  LAUNCHCHK("k_update_kvector");
  c->have_k = true;
  return 0;
}

int salmon_hip_update_kvector(salmon_hip_t c, const double *vec_Ac) {
  if (int rc = check(c)) return rc;
  if (!vec_Ac) return fail(SALMON_HIP_EINVAL, "vec_Ac is null");
  return update_kvector_impl(c, vec_Ac, nullptr, nullptr);
}

int salmon_hip_set_vec_ac(salmon_hip_t c, const double *vec_Ac) {
  if (int rc = check(c)) return rc;
  if (!vec_Ac) return fail(SALMON_HIP_EINVAL, "vec_Ac is null");
  if (c->nk > 1024) return fail(SALMON_HIP_EUNSUPPORTED, "nk=%d above 1024", c->nk);
  if (c->natom > 0 && !c->have_k) return fail(SALMON_HIP_ESTATE, "update_kvector not called yet");
  KvecArgs a;
  memset(&a, 0, sizeof(a));
  a.veck = c->d_veck;
  a.kpar = c->d_kpar;
  memcpy(a.coef_nab, c->coef_nab, sizeof(a.coef_nab));
  a.coef_lap0 = c->coef_lap0;
  a.ax = vec_Ac[0];
  a.ay = vec_Ac[1];
  a.az = vec_Ac[2];
  a.nk = c->nk;
  hipLaunchKernelGGL(k_set_kpar, dim3(1), dim3(1024), 0, c->stream, a);
  LAUNCHCHK("k_set_kpar");
  if (c->natom == 0) c->have_k = true;
  return 0;
}

// projection of psi into out (c(nb, nlma, nk) scaled by rinv, or the current moments)
static ProjArgs proj_args(salmon_hip_t c, const void *psi, double2 *out) {
  ProjArgs a;
  a.psi = psi;
  a.zekr = c->d_zekr;
  a.rxyz = c->d_rxyz;
  a.jidx = c->d_jidx;
  a.mps = c->d_mps;
  a.atom_lma = c->d_atom_lma;
  a.rinv = c->d_rinv;
  a.out = out;
  a.osz = c->osz;
  a.nb = c->nb;
  a.nk = c->nk;
  a.natom = c->natom;
  a.nps = c->nps;
  a.nlma = c->nlma;
  a.ngrp = (c->nb + kProjOrb - 1) / kProjOrb;
  a.skip = c->d_skip;
  a.zekr2 = nullptr;
  a.out2 = nullptr;
  return a;
}

// nw = 1: c of psi into d_c (skipped when salmon_hip_current_next left it there); 4: the current
// moments into d_curnl; 5: the moments and c of psi under the staged table into d_c
static int launch_proj(salmon_hip_t c, const void *psi, int nw, bool f32 = false) {
  if (nw == 1) {
    const bool pre = !f32 && c->pre_state == 2 && psi == c->pre_psi;
    c->pre_state = 0;
    if (pre) return 0;
  }
  ProjArgs a = proj_args(c, psi, nw == 1 ? c->d_c : c->d_curnl);
  if (nw == 5) {
    a.zekr2 = c->d_zekr_n;
    a.out2 = c->d_c;
  }
  dim3 grid((unsigned)(a.ngrp * c->natom * c->nk));
  hipEvent_t e0, e1;
  prof_events(c, nw == 1 ? SALMON_HIP_KERNEL_PROJ : SALMON_HIP_KERNEL_PROJ_CURRENT, &e0, &e1);
  // with overlap: fork onto the side stream after everything enqueued on `stream` so far (the
  // producer of psi, the last reader of the projection buffer); joined before the scatter
  hipStream_t st = c->stream;
  if (c->overlap) {
    HIPCHK(hipEventRecord(c->ev_fork, c->stream));
    HIPCHK(hipStreamWaitEvent(c->side, c->ev_fork, 0));
    st = c->side;
  }
  const int lmx = c->cmax <= 4 ? 4 : c->cmax <= 9 ? 9 : 16;
#define PROJ_LAUNCH(NWv, TPv)                                                                             \
  do {                                                                                                   \
    if (lmx == 4) hipExtLaunchKernelGGL((k_proj<NWv, TPv, 4>), grid, dim3(kBlock), 0, st, e0, e1, 0, a);   \
    else if (lmx == 9) hipExtLaunchKernelGGL((k_proj<NWv, TPv, 9>), grid, dim3(kBlock), 0, st, e0, e1, 0, a); \
    else hipExtLaunchKernelGGL((k_proj<NWv, TPv, 16>), grid, dim3(kBlock), 0, st, e0, e1, 0, a);           \
  } while (0)
  if (nw == 1 && f32) PROJ_LAUNCH(1, float2);
  else if (nw == 1) PROJ_LAUNCH(1, double2);
  else if (nw == 5) PROJ_LAUNCH(5, double2);
  else PROJ_LAUNCH(4, double2);
#undef PROJ_LAUNCH
  LAUNCHCHK("k_proj");
  if (c->overlap) {
    HIPCHK(hipEventRecord(c->ev_join, c->side));
    c->join_pending = true;
  }
  return 0;
}

// join: later work on `stream` waits for the last projection
static int join_proj(salmon_hip_t c) {
  if (!c->join_pending) return 0;
  HIPCHK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  c->join_pending = false;
  return 0;
}

}  // extern "C"

template <typename TI, int HC, int NC, int BS>
static size_t stencil_lds_bytes() {
  return sizeof(TI) * (kRingT * BS + kRingH * HC * BS) + sizeof(double2) * (NC >= 0 ? 2 * NC * BS : 0) +
         sizeof(double) * kRingH * BS;
}

// The dynamic-LDS limit of kernel f (hipFuncAttributeMaxDynamicSharedMemorySize), set once per
// device: the attribute belongs to the kernel on the device that is current when it is set.
static int lds_attr(const void *f, std::atomic<unsigned long long> &done, size_t bytes) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  const unsigned long long bit = 1ull << (dev & 63);
  if (done.load(std::memory_order_relaxed) & bit) return 0;
  HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  done.fetch_or(bit);
  return 0;
}

template <typename TI, typename R, typename TO, int HC, int NC, int BS = kTile>
static int launch_stencil_t(salmon_hip_ctx *c, dim3 grid, hipEvent_t e0, hipEvent_t e1, const StencilArgs &a) {
  const size_t shmem = stencil_lds_bytes<TI, HC, NC, BS>();
  static std::atomic<unsigned long long> attr_dev{0};
  if (int rc = lds_attr((const void *)k_stencil<TI, R, TO, HC, NC, BS>, attr_dev, shmem)) return rc;
  hipExtLaunchKernelGGL((k_stencil<TI, R, TO, HC, NC, BS>), grid, dim3(BS), shmem, c->stream, e0, e1, 0, a);
  return 0;
}

// Stencil stages: plain FP64, and the three stages of the mixed-precision chain.
enum StencilStage { kStagePlain = 0, kStageMixedFirst = 1, kStageMixedMiddle = 2, kStageMixedFinal = 3 };

// Whole-plane tiles (no y halo) for the terms without the Taylor sum, nx * ny <= BS = 256.
template <int BS>
static int launch_stencil_plane(salmon_hip_ctx *c, int stage, dim3 grid, hipEvent_t e0, hipEvent_t e1,
                                const StencilArgs &a) {
  if (stage == kStageMixedFirst) return launch_stencil_t<double2, float, float2, 0, -1, BS>(c, grid, e0, e1, a);
  if (stage == kStageMixedMiddle) return launch_stencil_t<float2, float, float2, 0, -1, BS>(c, grid, e0, e1, a);
  return launch_stencil_t<double2, double, double2, 0, -1, BS>(c, grid, e0, e1, a);
}


// Whole planes of 256 < nx*ny <= 1024 points: 512 threads with PY points each (stencil_plane).
constexpr int kPlaneDepth = 2;  // look-ahead of the whole-plane kernel (planes; 3 measured no faster)

template <typename TI, typename R, typename TO, int PY>
static int launch_plane_t(salmon_hip_ctx *c, dim3 grid, hipEvent_t e0, hipEvent_t e1, const StencilArgs &a) {
  constexpr int BS = 512, D = kPlaneDepth;
  const size_t shmem = sizeof(TI) * (5 + D) * PY * BS + sizeof(double) * (D + 1) * PY * BS;
  static std::atomic<unsigned long long> attr_dev{0};
  if (int rc = lds_attr((const void *)k_stencil_plane<TI, R, TO, PY, BS, D>, attr_dev, shmem)) return rc;
  hipExtLaunchKernelGGL((k_stencil_plane<TI, R, TO, PY, BS, D>), grid, dim3(BS), shmem, c->stream, e0, e1, 0, a);
  return 0;
}

template <int PY>
static int launch_stencil_wp(salmon_hip_ctx *c, int stage, dim3 grid, hipEvent_t e0, hipEvent_t e1,
                             const StencilArgs &a) {
  if (stage == kStageMixedFirst) return launch_plane_t<double2, float, float2, PY>(c, grid, e0, e1, a);
  if (stage == kStageMixedMiddle) return launch_plane_t<float2, float, float2, PY>(c, grid, e0, e1, a);
  return launch_plane_t<double2, double, double2, PY>(c, grid, e0, e1, a);
}

// y tiles with the merged halo and compile-time ring slots (stencil_tile8): the middle terms of
// the mixed chain, rows of at most 64 points, (TY + 8) * nx <= 768 (slots of two or three
// 256-element chunks), z chunks of 8k planes.
static bool tile8_fits(const salmon_hip_ctx *c, int stage, int nc, const StencilArgs &a) {
  (void)nc;
  return stage == kStageMixedMiddle && a.nx <= 64 && (a.TY + 8) * a.nx <= 3 * kTile && a.TY * a.nx <= kTile &&
         a.TZ % 8 == 0 && a.nz % 8 == 0 && c->tile8;
}

// the paired dwordx4 DMAs need rows of even length and 16-byte aligned orbital blocks
static bool tile8_pairs(const StencilArgs &a) { return a.nx % 2 == 0 && a.osz % 2 == 0; }

// y tiles with the delayed z part (stencil_tilez): the FP64 terms without the Taylor sum and the
// mixed first term (FP64 input), z chunks of 4k planes; a plane's tile and halo rows fill a slot
// of 8, 10 or 12 64-element chunks (rows of up to 32 points; 48-point rows take 5 tile rows,
// 13 x 48 = 624 elements in 10 chunks; up to 64 points).
static bool tilez_fits(const salmon_hip_ctx *c, int stage, int nc, const StencilArgs &a) {
  return ((stage == kStagePlain && nc < 0) || stage == kStageMixedFirst) && a.nx <= 64 &&
         (a.TY + 8) * a.nx <= 3 * kTile && a.TY * a.nx <= kTile && a.TZ % 4 == 0 && a.nz % 4 == 0 && c->tilez;
}

template <typename TI, typename R, typename TO, int NCK>
static int launch_tilez_n(salmon_hip_ctx *c, dim3 grid, hipEvent_t e0, hipEvent_t e1, const StencilArgs &a) {
  const size_t shmem = sizeof(TI) * 4 * 64 * NCK + sizeof(double) * 4 * kTile;
  static std::atomic<unsigned long long> attr_dev{0};
  if (int rc = lds_attr((const void *)k_stencilz<TI, R, TO, NCK>, attr_dev, shmem)) return rc;
  hipExtLaunchKernelGGL((k_stencilz<TI, R, TO, NCK>), grid, dim3(kTile), shmem, c->stream, e0, e1, 0, a);
  return 0;
}

// 64-element chunks per delayed-z slot for (TY + 8) rows of nx points: 8, 10 or 12
static int tilez_chunks(int TY, int nx) {
  const int need = (TY + 8) * nx;
  return need <= 8 * 64 ? 8 : (need <= 10 * 64 ? 10 : 12);
}

template <typename TI, typename R, typename TO>
static int launch_tilez_t(salmon_hip_ctx *c, dim3 grid, hipEvent_t e0, hipEvent_t e1, const StencilArgs &a) {
  switch (tilez_chunks(a.TY, a.nx)) {
    case 8: return launch_tilez_n<TI, R, TO, 8>(c, grid, e0, e1, a);
    case 10: return launch_tilez_n<TI, R, TO, 10>(c, grid, e0, e1, a);
    default: return launch_tilez_n<TI, R, TO, 12>(c, grid, e0, e1, a);
  }
}

template <typename TI, typename R, typename TO, int NCH, bool PR>
static int launch_tile8_n(salmon_hip_ctx *c, dim3 grid, hipEvent_t e0, hipEvent_t e1, const StencilArgs &a) {
  const size_t shmem = sizeof(TI) * 8 * NCH * kTile + sizeof(double) * 4 * kTile;
  static std::atomic<unsigned long long> attr_dev{0};
  if (int rc = lds_attr((const void *)k_stencil8<TI, R, TO, NCH, PR>, attr_dev, shmem)) return rc;
  hipExtLaunchKernelGGL((k_stencil8<TI, R, TO, NCH, PR>), grid, dim3(kTile), shmem, c->stream, e0, e1, 0, a);
  return 0;
}

template <typename TI, typename R, typename TO>
static int launch_tile8_t(salmon_hip_ctx *c, dim3 grid, hipEvent_t e0, hipEvent_t e1, const StencilArgs &a) {
  const bool pr = tile8_pairs(a);
  if ((a.TY + 8) * a.nx <= 2 * kTile)
    return pr ? launch_tile8_n<TI, R, TO, 2, true>(c, grid, e0, e1, a) : launch_tile8_n<TI, R, TO, 2, false>(c, grid, e0, e1, a);
  return pr ? launch_tile8_n<TI, R, TO, 3, true>(c, grid, e0, e1, a) : launch_tile8_n<TI, R, TO, 3, false>(c, grid, e0, e1, a);
}

template <int HC>
static int launch_stencil_hc(salmon_hip_ctx *c, int stage, int nc, dim3 grid, hipEvent_t e0, hipEvent_t e1,
                             const StencilArgs &a) {
  if (stage == kStageMixedFirst) return launch_stencil_t<double2, float, float2, HC, -1>(c, grid, e0, e1, a);
  if (stage == kStageMixedMiddle) return launch_stencil_t<float2, float, float2, HC, -1>(c, grid, e0, e1, a);
  if (stage == kStageMixedFinal) {
    switch (nc) {
      case 1: return launch_stencil_t<float2, float, double2, HC, 1>(c, grid, e0, e1, a);
      case 2: return launch_stencil_t<float2, float, double2, HC, 2>(c, grid, e0, e1, a);
      case 3: return launch_stencil_t<float2, float, double2, HC, 3>(c, grid, e0, e1, a);
      case 4: return launch_stencil_t<float2, float, double2, HC, 4>(c, grid, e0, e1, a);
      default: return launch_stencil_t<float2, float, double2, HC, 5>(c, grid, e0, e1, a);
    }
  }
  switch (nc) {
    case -1: return launch_stencil_t<double2, double, double2, HC, -1>(c, grid, e0, e1, a);
    case 0: return launch_stencil_t<double2, double, double2, HC, 0>(c, grid, e0, e1, a);
    case 1: return launch_stencil_t<double2, double, double2, HC, 1>(c, grid, e0, e1, a);
    case 2: return launch_stencil_t<double2, double, double2, HC, 2>(c, grid, e0, e1, a);
    case 3: return launch_stencil_t<double2, double, double2, HC, 3>(c, grid, e0, e1, a);
    case 4: return launch_stencil_t<double2, double, double2, HC, 4>(c, grid, e0, e1, a);
    default: return launch_stencil_t<double2, double, double2, HC, 5>(c, grid, e0, e1, a);
  }
}

extern "C" {

// shared_only: the touched points with several (atom, point) pairs (before k_proj_scatter);
// e0/e1 override the profiling events (nullptr: this launch's own)
// KB scatter hout += s * sum c zekr at the touched points (or only those several pairs cover)
static ScatterArgs scatter_args(salmon_hip_t c, void *hout, double2 s, const double2 *cbuf, bool shared_only, bool f32) {
  ScatterArgs a;
  a.hout = hout;
  a.c = cbuf;
  a.zent = c->d_zent;
  a.tp_info = shared_only ? c->d_tp_shared : c->d_tp_info;
  a.pt_lma = c->d_pt_lma;
  a.s = s;
  a.osz = c->osz;
  a.ntp = shared_only ? c->ntp_shared : c->ntp;
  a.nb = c->nb;
  a.nlma = c->nlma;
  a.nent = c->nent;
  a.cmax = c->cmax;
  a.skip = c->d_skip;
  const int G = f32 ? kScatG32 : kScatG;
  a.ngrp = (c->nb + G - 1) / G;
  return a;
}

static int launch_scatter(salmon_hip_t c, void *hout, double2 s, bool f32 = false, bool shared_only = false,
                          hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr, bool own_events = true) {
  if ((shared_only ? c->ntp_shared : c->ntp) == 0) return 0;
  const ScatterArgs a = scatter_args(c, hout, s, c->d_c, shared_only, f32);
  hipEvent_t e0 = ev0, e1 = ev1;
  if (own_events) prof_events(c, SALMON_HIP_KERNEL_NL_SCATTER, &e0, &e1);
  dim3 grid((unsigned)((a.ntp + kBlock - 1) / kBlock), (unsigned)(c->nk * a.ngrp));
  if (f32) {
    if (c->cmax <= kScatCM32)
      hipExtLaunchKernelGGL((k_nl_scatter<float2, kScatG32, kScatCM32>), grid, dim3(kBlock), 0, c->stream, e0, e1, 0, a);
    else
      hipExtLaunchKernelGGL((k_nl_scatter<float2, kScatG32, 0>), grid, dim3(kBlock), 0, c->stream, e0, e1, 0, a);
  } else {
    if (c->cmax <= kScatCM)
      hipExtLaunchKernelGGL((k_nl_scatter<double2, kScatG, kScatCM>), grid, dim3(kBlock), 0, c->stream, e0, e1, 0, a);
    else
      hipExtLaunchKernelGGL((k_nl_scatter<double2, kScatG, 0>), grid, dim3(kBlock), 0, c->stream, e0, e1, 0, a);
  }
  LAUNCHCHK("k_nl_scatter");
  return 0;
}

// the fused KB pass of hout (cin = c of the previous term, cout = c of hout)
static ProjScatArgs projscat_args(salmon_hip_t c, void *hout, const double2 *cin, double2 *cout) {
  ProjScatArgs A;
  A.p = proj_args(c, hout, cout);
  A.cin = cin;
  A.jexcl = c->d_jexcl;
  A.s = make_double2(1.0, 0.0);
  A.cmax = c->cmax;
  return A;
}

// The KB term of hout = H in (c of `in` current in d_c) fused with the projection of hout: the
// points with several pairs by k_nl_scatter, then k_proj_scatter; afterwards d_c holds the c of
// hout.  Profiled as one SALMON_HIP_KERNEL_KB_FUSED interval.
static int launch_kb_fused(salmon_hip_t c, void *hout, bool f32) {
  hipEvent_t e0, e1;
  prof_events(c, SALMON_HIP_KERNEL_KB_FUSED, &e0, &e1);
  const bool shared = c->ntp_shared > 0;
  if (shared)
    if (int rc = launch_scatter(c, hout, make_double2(1.0, 0.0), f32, true, e0, nullptr, false)) return rc;
  const ProjScatArgs A = projscat_args(c, hout, c->d_c, c->d_c2);
  const ProjArgs &a = A.p;
  const size_t shmem = sizeof(double2) * (size_t)c->cmax * kProjOrb;
  dim3 grid((unsigned)(a.ngrp * c->natom * c->nk));
  hipEvent_t s0 = shared ? nullptr : e0;
  // LM: registers for the projector values of one point (s+p: 4, up to d: 9, up to f: 16)
  if (f32) {
    if (c->cmax <= 4)
      hipExtLaunchKernelGGL((k_proj_scatter<float2, 4>), grid, dim3(kBlock), shmem, c->stream, s0, e1, 0, A);
    else if (c->cmax <= 9)
      hipExtLaunchKernelGGL((k_proj_scatter<float2, 9>), grid, dim3(kBlock), shmem, c->stream, s0, e1, 0, A);
    else
      hipExtLaunchKernelGGL((k_proj_scatter<float2, 16>), grid, dim3(kBlock), shmem, c->stream, s0, e1, 0, A);
  } else {
    if (c->cmax <= 4)
      hipExtLaunchKernelGGL((k_proj_scatter<double2, 4>), grid, dim3(kBlock), shmem, c->stream, s0, e1, 0, A);
    else if (c->cmax <= 9)
      hipExtLaunchKernelGGL((k_proj_scatter<double2, 9>), grid, dim3(kBlock), shmem, c->stream, s0, e1, 0, A);
    else
      hipExtLaunchKernelGGL((k_proj_scatter<double2, 16>), grid, dim3(kBlock), shmem, c->stream, s0, e1, 0, A);
  }
  LAUNCHCHK("k_proj_scatter");
  std::swap(c->d_c, c->d_c2);
  c->pre_state = 0;
  return 0;
}

// The fused KB pass is used inside Taylor chains when the projections run on the handle's stream.
static bool kb_fused_on(const salmon_hip_ctx *c) { return c->kb_fused && !c->overlap && c->natom > 0 && c->nlma > 0; }

// One H application: hout = H in (fin = false), or the last Taylor term
// hout = psi + c1 h1 + ... + c_N H in (fin = true, comb = [psi, h1 .. h_{N-2}]).
// The projections c of `in` must be current (launch_proj).
// Stencil descriptor of one H application; *bs_out = whole-plane block size (0 = y tiles).
// tiled_only forces the y-tile shape (the batched step).
static StencilArgs stencil_args(salmon_hip_t c, const void *in, void *hout, bool fin, void *const *comb, int ncomb,
                                const double *zc, int *bs_out, bool tiled_only) {
  StencilArgs a;
  memset(&a, 0, sizeof(a));
  a.in = in;
  a.hout = hout;
  a.sink = c->d_sink;
  a.skip = tiled_only ? nullptr : c->d_skip;
  const int nc = fin ? ncomb : -1;
  for (int n = 0; n < (fin ? ncomb : 0); ++n) a.comb[n] = comb[n];
  if (fin)
    for (int n = 0; n <= ncomb; ++n) a.zc[n] = make_double2(zc[2 * n], zc[2 * n + 1]);
  a.V = c->d_V;
  a.kpar = c->d_kpar;
  memcpy(a.lap, c->coef_lap, sizeof(a.lap));
  a.osz = c->osz;
  a.nx = c->nx; a.ny = c->ny; a.nz = c->nz; a.ly = c->ly;
  a.nspin = c->nspin; a.nb = c->nb; a.nk = c->nk;
  a.TY = c->TY; a.TZ = c->TZ; a.nyt = c->nyt; a.nzt = c->nzt;
  // terms without the Taylor sum use whole xy-plane tiles when a plane fits one block
  const int plane = c->nx * c->ny;
  // rule (tile_plane = -1): whole planes only up to 256 points (one 256-thread block); the
  // 1024-thread FP64 block spills and holds one block per CU (slower at 32x32, measured)
  const bool use_plane = !tiled_only && nc < 0 && plane <= 1024 &&
                         (c->tile_plane == 1 || (c->tile_plane == -1 && plane <= 256));
  const int bs = !use_plane ? 0 : (plane <= 256 ? 256 : (plane <= 512 ? 512 : 1024));
  if (bs) {
    a.TY = c->ny;
    a.nyt = 1;
    a.nzt = std::max(1, std::min((512 + c->norb - 1) / c->norb, c->nz / 8));
  }
  if (nc < 0 && c->mid_nzt > 0) a.nzt = std::min(c->mid_nzt, c->nz);
  if (bs || (nc < 0 && c->mid_nzt > 0)) {
    a.TZ = (c->nz + a.nzt - 1) / a.nzt;
    a.nzt = (c->nz + a.TZ - 1) / a.TZ;
  }
  a.nblk = c->norb * a.nyt * a.nzt;
  *bs_out = bs;
  return a;
}

// fuse_next: finish with the fused KB pass (hout's KB term and its projection) instead of the scatter.
static int launch_hpsi(salmon_hip_t c, int kind, const void *in, void *hout, bool fin, void *const *comb, int ncomb,
                       const double *zc, int stage = kStagePlain, bool fuse_next = false) {
  int bs = 0;
  const int nc = fin ? ncomb : -1;
  StencilArgs a = stencil_args(c, in, hout, fin, comb, ncomb, zc, &bs, false);
  dim3 grid((unsigned)a.nblk);
  hipEvent_t e0, e1;
  prof_events(c, kind, &e0, &e1);
  const int hc = (8 * c->nx + kTile - 1) / kTile;
  int rc = 0;
  if (bs == 256) rc = launch_stencil_plane<256>(c, stage, grid, e0, e1, a);
  else if (bs == 512) rc = launch_stencil_wp<1>(c, stage, grid, e0, e1, a);
  else if (bs == 1024) rc = launch_stencil_wp<2>(c, stage, grid, e0, e1, a);
  else if (tilez_fits(c, stage, nc, a))
    rc = stage == kStagePlain ? launch_tilez_t<double2, double, double2>(c, grid, e0, e1, a)
                              : launch_tilez_t<double2, float, float2>(c, grid, e0, e1, a);
  else if (tile8_fits(c, stage, nc, a)) rc = launch_tile8_t<float2, float, float2>(c, grid, e0, e1, a);
  else if (hc <= 1) rc = launch_stencil_hc<1>(c, stage, nc, grid, e0, e1, a);
  else if (hc <= 2) rc = launch_stencil_hc<2>(c, stage, nc, grid, e0, e1, a);
  else if (hc <= 4) rc = launch_stencil_hc<4>(c, stage, nc, grid, e0, e1, a);
  else rc = launch_stencil_hc<8>(c, stage, nc, grid, e0, e1, a);
  if (rc) return rc;
  LAUNCHCHK("k_stencil");
  const bool f32out = stage == kStageMixedFirst || stage == kStageMixedMiddle;
  if (int rc = join_proj(c)) return rc;
  if (fuse_next && !fin && kb_fused_on(c)) return launch_kb_fused(c, hout, f32out);
  if (c->natom > 0 && c->nlma > 0)
    return launch_scatter(c, hout, fin ? a.zc[ncomb] : make_double2(1.0, 0.0), f32out);
  return 0;
}

static int ready(salmon_hip_t c) {
  if (!c->have_V) return fail(SALMON_HIP_ESTATE, "V_local not set");
  if (!c->have_k) return fail(SALMON_HIP_ESTATE, "update_kvector (vec_Ac) not set");
  return 0;
}

int salmon_hip_hpsi(salmon_hip_t c, const void *tpsi, void *htpsi) {
  if (int rc = check(c)) return rc;
  if (!tpsi || !htpsi || tpsi == htpsi) return fail(SALMON_HIP_EINVAL, "tpsi/htpsi invalid or aliased");
  if (int rc = ready(c)) return rc;
  if (c->natom > 0)
    if (int rc = launch_proj(c, tpsi, 1)) return rc;
  return launch_hpsi(c, SALMON_HIP_KERNEL_HPSI0, tpsi, htpsi, false, nullptr, 0, nullptr);
}

// Mixed-precision Taylor step: h_n = H^n psi in complex FP32 (stencil and
// local-potential products in FP32; projections, the KB term and the Taylor sum
// in FP64), out = psi + sum_n zc(n) h_n formed in FP64 by the last stencil pass.
// first = 1: the whole chain; first = 2: h_1 (the handle's FP32 H psi) is already current
static int taylor_mixed(salmon_hip_t c, const double *zc, int n_hamil, void *tin, void *tout, int first = 1) {
  if (n_hamil < 2) return fail(SALMON_HIP_EUNSUPPORTED, "mixed precision needs n_hamil >= 2");
  while ((int)c->tscr32.size() < n_hamil - 1) {
    float2 *p = nullptr;
    if (int rc = dalloc(&p, c->osz * c->norb)) return rc;
    c->tscr32.push_back(p);
  }
  void *h[kMaxComb + 1];
  h[0] = tin;
  for (int n = 1; n < n_hamil; ++n) h[n] = c->tscr32[n - 1];
  // with the fused KB pass each term's projection comes out of the previous term's KB pass
  const bool fuse = kb_fused_on(c);
  bool have_c = false;
  for (int n = first; n < n_hamil; ++n) {
    if (c->natom > 0 && !have_c)
      if (int rc = launch_proj(c, h[n - 1], 1, n > 1)) return rc;
    const int kind = n == 1 ? SALMON_HIP_KERNEL_HPSI_TAYLOR1 : SALMON_HIP_KERNEL_HPSI_TAYLOR;
    const int stage = n == 1 ? kStageMixedFirst : kStageMixedMiddle;
    if (int rc = launch_hpsi(c, kind, h[n - 1], h[n], false, nullptr, 0, nullptr, stage, fuse)) return rc;
    have_c = fuse;
  }
  if (c->natom > 0 && !have_c)
    if (int rc = launch_proj(c, h[n_hamil - 1], 1, true)) return rc;
  return launch_hpsi(c, SALMON_HIP_KERNEL_HPSI_TAYLOR_LAST, h[n_hamil - 1], tout, true, h, n_hamil - 1, zc,
                     kStageMixedFinal);
}

int salmon_hip_set_precision(salmon_hip_t c, int mode) {
  if (int rc = check(c)) return rc;
  if (mode != SALMON_HIP_PREC_FP64 && mode != SALMON_HIP_PREC_MIXED)
    return fail(SALMON_HIP_EINVAL, "precision mode %d", mode);
  c->prec = mode;
  return 0;
}

// first = 1: the whole chain; first = 2: h_1 = sshtpsi is already current
static int taylor_fp64(salmon_hip_t c, const double *zc, int n_hamil, void *tspsi_in, void *tspsi_out, void *sshtpsi,
                       int first) {
  // h_1 = H psi -> sshtpsi, h_2 .. h_{N-1} -> internal scratch
  const int nscr = std::max(0, n_hamil - 2);
  while ((int)c->tscr.size() < nscr) {
    double2 *p = nullptr;
    if (int rc = dalloc(&p, c->osz * c->norb)) return rc;
    c->tscr.push_back(p);
  }
  void *h[kMaxComb + 1];
  h[0] = tspsi_in;
  if (n_hamil >= 2) h[1] = sshtpsi;
  for (int n = 2; n < n_hamil; ++n) h[n] = c->tscr[n - 2];
  const bool fuse = kb_fused_on(c);
  bool have_c = false;
  for (int n = first; n < n_hamil; ++n) {
    if (c->natom > 0 && !have_c)
      if (int rc = launch_proj(c, h[n - 1], 1)) return rc;
    const int kind = n == 1 ? SALMON_HIP_KERNEL_HPSI_TAYLOR1 : SALMON_HIP_KERNEL_HPSI_TAYLOR;
    if (int rc = launch_hpsi(c, kind, h[n - 1], h[n], false, nullptr, 0, nullptr, kStagePlain, fuse)) return rc;
    have_c = fuse;
  }
  if (c->natom > 0 && !have_c)
    if (int rc = launch_proj(c, h[n_hamil - 1], 1)) return rc;
  return launch_hpsi(c, SALMON_HIP_KERNEL_HPSI_TAYLOR_LAST, h[n_hamil - 1], tspsi_out, true, h, n_hamil - 1, zc);
}

int salmon_hip_taylor(salmon_hip_t c, const double *zc, int n_hamil, void *tspsi_in, void *tspsi_out,
                      void *sshtpsi) {
  if (int rc = check(c)) return rc;
  if (!zc || n_hamil < 1 || n_hamil > kMaxComb + 1)
    return fail(SALMON_HIP_EINVAL, "zc null or n_hamil=%d outside 1..%d", n_hamil, kMaxComb + 1);
  if (!tspsi_in || !tspsi_out || !sshtpsi || tspsi_in == tspsi_out || tspsi_in == sshtpsi || tspsi_out == sshtpsi)
    return fail(SALMON_HIP_EINVAL, "taylor buffers invalid or aliased");
  if (int rc = ready(c)) return rc;
  if (c->prec == SALMON_HIP_PREC_MIXED) return taylor_mixed(c, zc, n_hamil, tspsi_in, tspsi_out);
  return taylor_fp64(c, zc, n_hamil, tspsi_in, tspsi_out, sshtpsi, 1);
}

int salmon_hip_taylor_reuse(salmon_hip_t c, const double *zc, int n_hamil, void *tspsi_in, void *tspsi_out,
                            void *sshtpsi, const double *dv_dev) {
  if (int rc = check(c)) return rc;
  if (!zc || n_hamil < 2 || n_hamil > kMaxComb + 1)
    return fail(SALMON_HIP_EINVAL, "zc null or n_hamil=%d outside 2..%d", n_hamil, kMaxComb + 1);
  if (!tspsi_in || !tspsi_out || !dv_dev || tspsi_in == tspsi_out)
    return fail(SALMON_HIP_EINVAL, "taylor_reuse buffers invalid or aliased");
  if (int rc = ready(c)) return rc;
  const bool mixed = c->prec == SALMON_HIP_PREC_MIXED;
  void *h1 = mixed ? (c->tscr32.empty() ? nullptr : (void *)c->tscr32[0]) : sshtpsi;
  if (!h1 || (!mixed && (sshtpsi == tspsi_in || sshtpsi == tspsi_out)))
    return fail(SALMON_HIP_ESTATE, "taylor_reuse needs H psi of a previous taylor call (sshtpsi, or the FP32 chain)");
  const size_t ntot = c->osz * c->norb;
  const dim3 grid((unsigned)((ntot + kBlock - 1) / kBlock));
  if (mixed)
    hipLaunchKernelGGL(k_vshift<float2>, grid, dim3(kBlock), 0, c->stream, (float2 *)h1, (const double2 *)tspsi_in,
                       dv_dev, c->osz, c->nx, c->ny, c->ly, c->nz, c->nspin, ntot, c->d_skip);
  else
    hipLaunchKernelGGL(k_vshift<double2>, grid, dim3(kBlock), 0, c->stream, (double2 *)h1, (const double2 *)tspsi_in,
                       dv_dev, c->osz, c->nx, c->ny, c->ly, c->nz, c->nspin, ntot, c->d_skip);
  LAUNCHCHK("k_vshift");
  if (mixed) return taylor_mixed(c, zc, n_hamil, tspsi_in, tspsi_out, 2);
  return taylor_fp64(c, zc, n_hamil, tspsi_in, tspsi_out, sshtpsi, 2);
}



int salmon_hip_autotune(salmon_hip_t c, int reps, int *choice, double *best_ms) {
  if (int rc = check(c)) return rc;
  if (int rc = ready(c)) return rc;
  if (reps < 1) reps = 3;
  // two orbital buffers of scratch (zero orbitals: the timing does not depend on the values)
  double2 *bin = nullptr, *bout = nullptr;
  int rc = dalloc(&bin, c->osz * c->norb);
  if (!rc) rc = dalloc(&bout, c->osz * c->norb);
  if (rc) { dfree(bin); return rc; }
  (void)hipMemsetAsync(bin, 0, sizeof(double2) * c->osz * c->norb, c->stream);
  struct Cand { int plane, nzt; };
  std::vector<Cand> cands;
  if (c->nx * c->ny <= 1024)
    for (int n : {1, 2}) cands.push_back({1, n});
  for (int n : {1, 2, 4})
    if (n <= std::max(1, c->nz / 8)) cands.push_back({0, n});
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const int save_plane = c->tile_plane, save_nzt = c->mid_nzt;
  const bool save_prof = c->prof;
  c->prof = false;
  double best = 1e30;
  int besti = 0;
  for (size_t i = 0; i < cands.size() && !rc; ++i) {
    c->tile_plane = cands[i].plane;
    c->mid_nzt = cands[i].nzt;
    // local stencil only (the KB scatter does not depend on the tile shape)
    const int nat = c->natom;
    c->natom = 0;
    rc = launch_hpsi(c, SALMON_HIP_KERNEL_HPSI0, bin, bout, false, nullptr, 0, nullptr);
    if (!rc && hipEventRecord(e0, c->stream) != hipSuccess) rc = fail(SALMON_HIP_EHIP, "event record");
    for (int r = 0; r < reps && !rc; ++r) rc = launch_hpsi(c, SALMON_HIP_KERNEL_HPSI0, bin, bout, false, nullptr, 0, nullptr);
    c->natom = nat;
    if (rc) break;
    float ms = 0.f;
    if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      rc = fail(SALMON_HIP_EHIP, "autotune timing failed");
      break;
    }
    const double t = ms / reps;
    if (t < best) { best = t; besti = (int)i; }
  }
  c->prof = save_prof;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(c->stream);
  dfree(bin);
  dfree(bout);
  if (rc) {
    c->tile_plane = save_plane;
    c->mid_nzt = save_nzt;
    return rc;
  }
  c->tile_plane = cands[besti].plane;
  c->mid_nzt = cands[besti].nzt;
  if (choice) *choice = 10 * cands[besti].plane + cands[besti].nzt;
  if (best_ms) *best_ms = best;
  return 0;
}

int salmon_hip_set_overlap(salmon_hip_t c, int on) {
  if (int rc = check(c)) return rc;
  if (int rc = join_proj(c)) return rc;
  c->overlap = on != 0;
  return 0;
}

int salmon_hip_set_launch_shape(salmon_hip_t c, int whole_plane, int z_chunks) {
  if (int rc = check(c)) return rc;
  if (whole_plane < -1 || whole_plane > 1 || z_chunks < 0) return fail(SALMON_HIP_EINVAL, "launch shape");
  c->tile_plane = whole_plane;
  c->mid_nzt = z_chunks;
  return 0;
}

static DensArgs dens_args(salmon_hip_t c, const void *psi, double *rho_dev, double *part) {
  DensArgs a;
  a.psi = (const double2 *)psi;
  a.rocc = c->d_rocc;
  a.wtk = c->d_wtk;
  a.part = part;
  a.rho = rho_dev;
  a.osz = c->osz;
  a.nx = c->nx; a.ny = c->ny; a.nz = c->nz; a.ly = c->ly;
  a.nspin = c->nspin; a.no = c->no; a.nk = c->nk;
  a.ngrp = c->dens_ngrp;
  a.gsz = kDensGroup;
  a.skip = c->d_skip;
  return a;
}

static int density_impl(salmon_hip_t c, const void *psi, double *rho_dev, double *part) {
  const DensArgs a = dens_args(c, psi, rho_dev, part);
  dim3 grid((unsigned)((c->ngrid + kBlock - 1) / kBlock), a.ngrp, c->nspin);
  hipLaunchKernelGGL(k_density_part, grid, dim3(kBlock), 0, c->stream, a);
  LAUNCHCHK("k_density_part");
  if (a.ngrp > 1) {
    dim3 g2((unsigned)((c->ngrid * c->nspin + kBlock - 1) / kBlock));
    hipLaunchKernelGGL(k_density_sum, g2, dim3(kBlock), 0, c->stream, a);
    LAUNCHCHK("k_density_sum");
  }
  return 0;
}

int salmon_hip_density(salmon_hip_t c, const void *psi, double *rho_dev) {
  if (int rc = check(c)) return rc;
  if (!psi || !rho_dev) return fail(SALMON_HIP_EINVAL, "null pointer");
  return density_impl(c, psi, rho_dev, c->d_denspart);
}

int salmon_hip_density_noalloc(salmon_hip_t c, const void *psi, double *rho_dev) {
  if (int rc = check(c)) return rc;
  if (!psi || !rho_dev) return fail(SALMON_HIP_EINVAL, "null pointer");
  if (!c->d_scratch) return fail(SALMON_HIP_ESTATE, "device scratch not resolved");
  const size_t need = c->dens_ngrp > 1 ? c->ngrid * (size_t)c->nspin * (size_t)c->dens_ngrp : 0;
  if (need > kNoAllocScratch)
    return fail(SALMON_HIP_EINVAL,
                "density_noalloc: %zu partial sums (%d groups of %d orbitals x %zu points x %d spins) exceed the "
                "device scratch of %zu doubles",
                need, c->dens_ngrp, kDensGroup, c->ngrid, c->nspin, kNoAllocScratch);
  return density_impl(c, psi, rho_dev, c->d_scratch);
}

static CurArgs cur_args(salmon_hip_t c, const void *psi) {
  CurArgs a;
  a.psi = (const double2 *)psi;
  a.part = c->d_curpart;
  memcpy(a.coef_nab, c->coef_nab, sizeof(a.coef_nab));
  a.osz = c->osz;
  a.nx = c->nx; a.ny = c->ny; a.nz = c->nz; a.ly = c->ly;
  a.TY = c->TY;
  a.nyt = c->nyt;
  return a;
}

static size_t cur_shmem(salmon_hip_t c) { return sizeof(double2) * 2 * (size_t)(c->TY + 4) * c->nx; }

static CurFinArgs curfin_args(salmon_hip_t c, const double *vec_Ac, const double *actab, const int *itt,
                              double *curr_dev, int tmax) {
  CurFinArgs f;
  f.part = c->d_curpart;
  f.nl = c->natom > 0 ? c->d_curnl : nullptr;
  f.rinv = c->d_rinv;
  f.veck = c->d_veck;
  f.rocc = c->d_rocc;
  f.wtk = c->d_wtk;
  f.contrib = c->d_contrib;
  f.curr = curr_dev;
  f.ax = vec_Ac ? vec_Ac[0] : 0.0;
  f.ay = vec_Ac ? vec_Ac[1] : 0.0;
  f.az = vec_Ac ? vec_Ac[2] : 0.0;
  f.actab = actab;
  f.itt = itt;
  f.tmax = tmax;
  f.inv_ngrid = 1.0 / (double)c->ngrid;
  f.nslab = c->nyt;
  f.nspin = c->nspin;
  f.no = c->no;
  f.nk = c->nk;
  f.nlma = c->nlma;
  return f;
}

static int current_impl(salmon_hip_t c, const double *vec_Ac, const double *actab, const int *itt, const void *psi,
                        double *curr_dev, int tmax = 0, bool next = false) {
  const CurArgs a = cur_args(c, psi);
  // the projector moments run on the side stream, overlapping the local part
  if (c->natom > 0)
    if (int rc = launch_proj(c, psi, next ? 5 : 4)) return rc;
  hipLaunchKernelGGL(k_current_local, dim3((unsigned)(c->norb * c->nyt)), dim3(kBlock), cur_shmem(c), c->stream, a);
  LAUNCHCHK("k_current_local");
  if (int rc = join_proj(c)) return rc;
  const CurFinArgs f = curfin_args(c, vec_Ac, actab, itt, curr_dev, tmax);
  const int M = c->no * c->nk;
  for (int is = 0; is < c->nspin; ++is) {
    hipLaunchKernelGGL(k_current_orb, dim3((unsigned)((M + 3) / 4)), dim3(kBlock), 0, c->stream, f, is);
    LAUNCHCHK("k_current_orb");
    hipLaunchKernelGGL(k_current_sum, dim3(1), dim3(kBlock), 0, c->stream, f, is);
    LAUNCHCHK("k_current_sum");
  }
  return 0;
}

int salmon_hip_current(salmon_hip_t c, const double *vec_Ac, const void *psi, double *curr_dev) {
  if (int rc = check(c)) return rc;
  if (!vec_Ac || !psi || !curr_dev) return fail(SALMON_HIP_EINVAL, "null pointer");
  if (c->natom > 0 && !c->have_k) return fail(SALMON_HIP_ESTATE, "update_kvector not called");
  return current_impl(c, vec_Ac, nullptr, nullptr, psi, curr_dev);
}

static DipArgs dip_args(salmon_hip_t c, const double *rho_dev, double *out_dev) {
  DipArgs a;
  a.rho = rho_dev;
  a.part = c->d_dippart;
  a.out = out_dev;
  a.nx = c->nx; a.ny = c->ny; a.nz = c->nz; a.nspin = c->nspin;
  a.hx = c->hgs[0]; a.hy = c->hgs[1]; a.hz = c->hgs[2]; a.hvol = c->hvol;
  return a;
}

int salmon_hip_current_next(salmon_hip_t c, const double *vec_Ac, const void *psi, double *curr_dev,
                            const double *vec_Ac_next) {
  if (int rc = check(c)) return rc;
  if (!vec_Ac || !psi || !curr_dev || !vec_Ac_next) return fail(SALMON_HIP_EINVAL, "null pointer");
  if (c->natom > 0 && !c->have_k) return fail(SALMON_HIP_ESTATE, "update_kvector not called");
  if (c->in_batch) return fail(SALMON_HIP_ESTATE, "current_next on a handle of a batch");
  if (c->natom == 0 || c->nlma == 0) return current_impl(c, vec_Ac, nullptr, nullptr, psi, curr_dev);
  if (!c->d_zekr_n) {
    int rc = dalloc(&c->d_zekr_n, (size_t)c->nps * c->nlma * c->nk);
    if (!rc) rc = dalloc(&c->d_zent_n, (size_t)c->cmax * std::max(c->nent, 1) * c->nk);
    if (!rc && !c->d_kpar_n) rc = dalloc(&c->d_kpar_n, 16 * (size_t)c->nk);
    if (rc) return rc;
  }
  // the next step's phase table and stencil terms into the staging buffers
  KvecArgs ka = kvec_args(c, vec_Ac_next, nullptr, nullptr, 0);
  ka.zekr = c->d_zekr_n;
  ka.zent = c->d_zent_n;
  ka.kpar = c->d_kpar_n;
  hipLaunchKernelGGL(k_update_kvector, dim3(std::max(1, c->nlma), c->nk), dim3(kBlock), 0, c->stream, ka);
  LAUNCHCHK("k_update_kvector");
  c->staged = true;
  memcpy(c->staged_A, vec_Ac_next, sizeof(c->staged_A));
  if (int rc = current_impl(c, vec_Ac, nullptr, nullptr, psi, curr_dev, 0, true)) return rc;
  c->pre_state = 1;
  c->pre_psi = psi;
  return 0;
}

int salmon_hip_dipole(salmon_hip_t c, const double *rho_dev, double *out_dev) {
  if (int rc = check(c)) return rc;
  if (!rho_dev || !out_dev) return fail(SALMON_HIP_EINVAL, "null pointer");
  const DipArgs a = dip_args(c, rho_dev, out_dev);
  hipLaunchKernelGGL(k_dipole_part, dim3(kDipBlocks), dim3(256), 0, c->stream, a);
  LAUNCHCHK("k_dipole_part");
  hipLaunchKernelGGL(k_dipole_sum, dim3(1), dim3(256), 0, c->stream, a);
  LAUNCHCHK("k_dipole_sum");
  return 0;
}

int salmon_hip_set_poisson(salmon_hip_t c, const double *al) {
  if (int rc = check(c)) return rc;
  if (!al) return fail(SALMON_HIP_EINVAL, "null pointer");
  for (int d = 0; d < 3; ++d)
    if (!(al[d] > 0.0)) return fail(SALMON_HIP_EINVAL, "cell length al[%d]=%g", d, al[d]);
  const int n[3] = {c->nx, c->ny, c->nz};
  for (int d = 0; d < 3; ++d)
    if (n[d] > kDftMaxN) return fail(SALMON_HIP_EUNSUPPORTED, "grid extent %d above %d", n[d], kDftMaxN);
  const double pi = 3.141592653589793238462643383279502884;
  int rc = 0;
  for (int d = 0; d < 3; ++d) {
    std::vector<double2> tw((size_t)n[d] * n[d]);
    for (int j = 0; j < n[d]; ++j)
      for (int k = 0; k < n[d]; ++k) {
        const double ang = 2.0 * pi * (double)((long)j * k) / (double)n[d];
        tw[(size_t)k * n[d] + j] = make_double2(cos(ang), sin(ang));
      }
    if (!c->d_tw[d] && (rc = dalloc(&c->d_tw[d], tw.size()))) return rc;
    HIPCHK(hipMemcpy(c->d_tw[d], tw.data(), sizeof(double2) * tw.size(), hipMemcpyHostToDevice));
  }
  // fg%coef (src/common/initialization.f90:705-724): index i -> i - n for i >= (n+1)/2
  std::vector<double> coef(c->ngrid);
  for (int kz = 0; kz < c->nz; ++kz)
    for (int ky = 0; ky < c->ny; ++ky)
      for (int kx = 0; kx < c->nx; ++kx) {
        const int ik[3] = {kx, ky, kz};
        double g2 = 0.0;
        for (int d = 0; d < 3; ++d) {
          const int ii = ik[d] >= (n[d] + 1) / 2 ? ik[d] - n[d] : ik[d];
          const double g = (double)ii * (2.0 * pi / al[d]);
          g2 = g2 + g * g;
        }
        coef[(size_t)kx + (size_t)c->nx * ((size_t)ky + (size_t)c->ny * kz)] =
            (kx == 0 && ky == 0 && kz == 0) ? 0.0 : 4.0 * pi / g2;
      }
  if (!c->d_gcoef && (rc = dalloc(&c->d_gcoef, c->ngrid))) return rc;
  HIPCHK(hipMemcpy(c->d_gcoef, coef.data(), sizeof(double) * c->ngrid, hipMemcpyHostToDevice));
  if (!c->d_fft[0] && (rc = dalloc(&c->d_fft[0], c->ngrid))) return rc;
  if (!c->d_fft[1] && (rc = dalloc(&c->d_fft[1], c->ngrid))) return rc;
  if (!c->d_vh && (rc = dalloc(&c->d_vh, c->ngrid))) return rc;
  c->have_cell = true;
  return 0;
}

static int launch_dft(salmon_hip_t c, int axis, double sg, const double *rin, const double2 *in, double2 *out,
                      double *rout, const double *coef, const double *rin2 = nullptr) {
  const int n[3] = {c->nx, c->ny, c->nz};
  const int inner = axis == 0 ? 1 : (axis == 1 ? c->nx : c->nx * c->ny);
  DftArgs a;
  a.in = in; a.rin = rin; a.rin2 = rin2; a.out = out; a.rout = rout;
  a.tw = c->d_tw[axis];
  a.coef = coef;
  a.nrm = (double)c->ngrid;
  a.n = n[axis];
  a.inner = inner;
  a.ncols = (long long)(c->ngrid / (size_t)n[axis]);
  a.sg = sg;
  a.skip = c->d_skip;
  // about one output per thread; columns per block a power of two in [1, 32]
  int cols = 1;
  while (cols < 32 && 2 * cols * a.n <= 256) cols *= 2;
  a.cols = cols;
  const unsigned nblk = (unsigned)((a.ncols + cols - 1) / cols);
  const size_t shm = sizeof(double2) * (size_t)a.n * cols;
  hipLaunchKernelGGL(k_dft_axis, dim3(nblk), dim3(256), shm, c->stream, a);
  LAUNCHCHK("k_dft_axis");
  return 0;
}

// Hartree potential of rho (+ rho2 when not null: the total density of two spin channels,
// rho_s(1) + rho_s(2), src/rt/time_evolution_step.f90:166-169, formed in the first pass)
static int hartree_impl(salmon_hip_t c, const double *rho_dev, const double *rho2_dev, double *vh_dev) {
  double2 *b0 = c->d_fft[0], *b1 = c->d_fft[1];
  int rc;
  if ((rc = launch_dft(c, 2, -1.0, rho_dev, nullptr, b0, nullptr, nullptr, rho2_dev))) return rc;
  if ((rc = launch_dft(c, 1, -1.0, nullptr, b0, b1, nullptr, nullptr))) return rc;
  if ((rc = launch_dft(c, 0, -1.0, nullptr, b1, b0, nullptr, c->d_gcoef))) return rc;
  if ((rc = launch_dft(c, 2, 1.0, nullptr, b0, b1, nullptr, nullptr))) return rc;
  if ((rc = launch_dft(c, 1, 1.0, nullptr, b1, b0, nullptr, nullptr))) return rc;
  return launch_dft(c, 0, 1.0, nullptr, b0, nullptr, vh_dev, nullptr);
}

int salmon_hip_hartree(salmon_hip_t c, const double *rho_dev, double *vh_dev) {
  if (int rc = check(c)) return rc;
  if (!rho_dev || !vh_dev) return fail(SALMON_HIP_EINVAL, "null pointer");
  if (!c->have_cell) return fail(SALMON_HIP_ESTATE, "set_poisson not called");
  return hartree_impl(c, rho_dev, nullptr, vh_dev);
}

int salmon_hip_xc_pz(salmon_hip_t c, const double *rho_dev, double *vxc_dev, double *exc_dev) {
  if (int rc = check(c)) return rc;
  if (!rho_dev || !vxc_dev) return fail(SALMON_HIP_EINVAL, "null pointer");
  hipLaunchKernelGGL(k_xc_pz, dim3((unsigned)((c->ngrid + 255) / 256)), dim3(256), 0, c->stream, rho_dev,
                     vxc_dev, exc_dev, c->ngrid, c->nspin);
  LAUNCHCHK("k_xc_pz");
  return 0;
}

static int launch_vlocal(salmon_hip_t c, const double *vpsl, const double *vh, const double *vxc,
                         const double *rho, const double *vmix, double *vout) {
  VlocArgs a;
  a.vpsl = vpsl; a.vh = vh; a.vxc = vxc; a.rho = rho; a.vmix = vmix;
  a.V = c->d_V;
  a.vout = vout;
  a.n = c->ngrid;
  a.nspin = c->nspin;
  a.skip = c->d_skip;
  hipLaunchKernelGGL(k_update_vlocal, dim3((unsigned)((c->ngrid + 255) / 256)), dim3(256), 0, c->stream, a);
  LAUNCHCHK("k_update_vlocal");
  c->have_V = true;
  return 0;
}

int salmon_hip_update_vlocal(salmon_hip_t c, const double *vpsl_dev, const double *vh_dev, const double *vxc_dev,
                             const double *vmix_dev, double *vout_dev) {
  if (int rc = check(c)) return rc;
  if (!vpsl_dev || !vh_dev || !vxc_dev) return fail(SALMON_HIP_EINVAL, "null pointer");
  return launch_vlocal(c, vpsl_dev, vh_dev, vxc_dev, nullptr, vmix_dev, vout_dev);
}

int salmon_hip_vlocal_extrapolate(salmon_hip_t c, double *vold1_dev, double *vold2_dev, double *vout_dev) {
  if (int rc = check(c)) return rc;
  if (!vold1_dev || !vold2_dev) return fail(SALMON_HIP_EINVAL, "null pointer");
  if (!c->have_V) return fail(SALMON_HIP_ESTATE, "V_local not set");
  const size_t n = c->ngrid * (size_t)c->nspin;
  hipLaunchKernelGGL(k_vextrap, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->d_V, vold1_dev,
                     vold2_dev, vout_dev, n);
  LAUNCHCHK("k_vextrap");
  return 0;
}

int salmon_hip_update_potential(salmon_hip_t c, const double *rho_dev, const double *vpsl_dev,
                                const double *vmix_dev, double *vout_dev) {
  if (int rc = check(c)) return rc;
  if (!rho_dev || !vpsl_dev) return fail(SALMON_HIP_EINVAL, "null pointer");
  if (!c->have_cell) return fail(SALMON_HIP_ESTATE, "set_poisson not called");
  if (int rc = hartree_impl(c, rho_dev, c->nspin == 2 ? rho_dev + c->ngrid : nullptr, c->d_vh)) return rc;
  return launch_vlocal(c, vpsl_dev, c->d_vh, nullptr, rho_dev, vmix_dev, vout_dev);
}

// ---------------------------------------------------------------------------
// Batched step over many independent systems (no counterpart in the reference).
// ---------------------------------------------------------------------------
}  // extern "C"

namespace {

// record the per-step observables at the device step counter, advance it
__global__ void k_record(const double *const *csrc, double *const *cdst, const int *cw, const double *const *dsrc,
                         double *const *ddst, const int *nt, int nsys, int *itt) {
  const int i = threadIdx.x;
  const int t = *itt;
  if (i < nsys && t >= 1 && t <= nt[i]) {
    for (int k = 0; k < cw[i]; ++k) cdst[i][(size_t)(t - 1) * cw[i] + k] = csrc[i][k];
    for (int k = 0; k < 4; ++k) ddst[i][(size_t)(t - 1) * 4 + k] = dsrc[i][k];
  }
  __syncthreads();
  if (i == 0) *itt = t + 1;
}

__global__ void k_set_itt(int *itt, int v) { *itt = v; }

template <int HC, int NC>
int launch_packed_t(hipStream_t st, const StencilArgs *sys, const int2 *tiles, int ntiles, int grid) {
  const size_t shmem = stencil_lds_bytes<double2, HC, NC, kTile>();
  static std::atomic<unsigned long long> attr_dev{0};
  if (int rc = lds_attr((const void *)k_stencil_packed<double2, double, double2, HC, NC, kTile>, attr_dev, shmem))
    return rc;
  hipLaunchKernelGGL((k_stencil_packed<double2, double, double2, HC, NC, kTile>), dim3(grid), dim3(kTile), shmem, st,
                     sys, tiles, ntiles);
  LAUNCHCHK("k_stencil_packed");
  return 0;
}

template <int HC>
int launch_packed_hc(hipStream_t st, int nc, const StencilArgs *sys, const int2 *tiles, int ntiles, int grid) {
  switch (nc) {
    case -1: return launch_packed_t<HC, -1>(st, sys, tiles, ntiles, grid);
    case 0: return launch_packed_t<HC, 0>(st, sys, tiles, ntiles, grid);
    case 1: return launch_packed_t<HC, 1>(st, sys, tiles, ntiles, grid);
    case 2: return launch_packed_t<HC, 2>(st, sys, tiles, ntiles, grid);
    case 3: return launch_packed_t<HC, 3>(st, sys, tiles, ntiles, grid);
    case 4: return launch_packed_t<HC, 4>(st, sys, tiles, ntiles, grid);
    default: return launch_packed_t<HC, 5>(st, sys, tiles, ntiles, grid);
  }
}

int launch_packed(hipStream_t st, int hc, int nc, const StencilArgs *sys, const int2 *tiles, int ntiles, int grid) {
  if (hc <= 1) return launch_packed_hc<1>(st, nc, sys, tiles, ntiles, grid);
  if (hc <= 2) return launch_packed_hc<2>(st, nc, sys, tiles, ntiles, grid);
  if (hc <= 4) return launch_packed_hc<4>(st, nc, sys, tiles, ntiles, grid);
  return launch_packed_hc<8>(st, nc, sys, tiles, ntiles, grid);
}

template <int NCK>
int launch_packed_z_t(hipStream_t st, const StencilArgs *sys, const int2 *tiles, int ntiles, int grid) {
  const size_t shmem = sizeof(double2) * 4 * 64 * NCK + sizeof(double) * 4 * kTile;
  static std::atomic<unsigned long long> attr_dev{0};
  if (int rc = lds_attr((const void *)k_stencilz_packed<double2, double, double2, NCK>, attr_dev, shmem))
    return rc;
  hipLaunchKernelGGL((k_stencilz_packed<double2, double, double2, NCK>), dim3(grid), dim3(kTile), shmem, st, sys,
                     tiles, ntiles);
  LAUNCHCHK("k_stencilz_packed");
  return 0;
}

// the delayed-z y tiles for the terms without the Taylor sum, slots of nck 64-element chunks
int launch_packed_z(hipStream_t st, int nck, const StencilArgs *sys, const int2 *tiles, int ntiles, int grid) {
  switch (nck) {
    case 8: return launch_packed_z_t<8>(st, sys, tiles, ntiles, grid);
    case 10: return launch_packed_z_t<10>(st, sys, tiles, ntiles, grid);
    default: return launch_packed_z_t<12>(st, sys, tiles, ntiles, grid);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Packed launches of the batched step: one launch runs a kernel body for every system of the
// batch.  The systems' argument records sit in a device array and system s owns the blocks
// [off[s], off[s+1]) of the launch; a block finds its system by binary search over off (block-
// uniform, scalar loads), copies the record and runs the body with its local block index.  The
// ranges of the bodies that order their blocks for the XCDs (projections, fused KB pass) are
// padded to multiples of 8 blocks, so a local block index keeps its XCD (block b runs on XCD
// b % 8); padding blocks return.  The per-system kernels and these run the same bodies on the
// same records, so a packed step equals the per-system step bit for bit.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int pack_find(const int *off, int nsys, int b) {
  int lo = 0, hi = nsys;  // off[lo] <= b < off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(kBlock) void kp_update_kvector(const KvecArgs *args, const int *off, int nsys) {
  const int s = pack_find(off, nsys, blockIdx.x);
  const KvecArgs a = args[s];
  const int lb = (int)blockIdx.x - off[s], gx = max(1, a.nlma);
  update_kvector_body(a, lb % gx, lb / gx);
}

template <int NW, int LMX>
__global__ __launch_bounds__(kBlock) void kp_proj(const ProjArgs *args, const int *off, int nsys) {
  const int s = pack_find(off, nsys, blockIdx.x);
  const ProjArgs a = args[s];
  const int item = xcd_order((int)blockIdx.x - off[s], off[s + 1] - off[s]);
  if (item >= a.ngrp * a.natom * a.nk) return;
  proj_body<NW, double2, LMX>(a, item);
}

template <int LM>
__global__ __launch_bounds__(kBlock) void kp_proj_scatter(const ProjScatArgs *args, const int *off, int nsys) {
  const int s = pack_find(off, nsys, blockIdx.x);
  const ProjScatArgs A = args[s];
  const int item = xcd_order((int)blockIdx.x - off[s], off[s + 1] - off[s]);
  if (item >= A.p.ngrp * A.p.natom * A.p.nk) return;
  proj_scatter_body<double2, LM>(A, item);
}

template <int CM>
__global__ __launch_bounds__(kBlock) void kp_nl_scatter(const ScatterArgs *args, const int *off, int nsys) {
  const int s = pack_find(off, nsys, blockIdx.x);
  const ScatterArgs a = args[s];
  const int lb = (int)blockIdx.x - off[s], gx = (a.ntp + kBlock - 1) / kBlock;
  nl_scatter_body<double2, kScatG, CM>(a, lb % gx, lb / gx);
}

__global__ __launch_bounds__(kBlock) void kp_density_part(const DensArgs *args, const int *off, int nsys) {
  const int s = pack_find(off, nsys, blockIdx.x);
  const DensArgs a = args[s];
  const int lb = (int)blockIdx.x - off[s];
  const int gx = (a.nx * a.ny * a.nz + kBlock - 1) / kBlock;
  density_part_body(a, lb % gx, (lb / gx) % a.ngrp, lb / (gx * a.ngrp));
}

__global__ __launch_bounds__(kBlock) void kp_density_sum(const DensArgs *args, const int *off, int nsys) {
  const int s = pack_find(off, nsys, blockIdx.x);
  const DensArgs a = args[s];
  density_sum_body(a, (int)blockIdx.x - off[s]);
}

__global__ __launch_bounds__(kBlock) void kp_current_local(const CurArgs *args, const int *off, int nsys) {
  const int s = pack_find(off, nsys, blockIdx.x);
  const CurArgs a = args[s];
  current_local_body(a, (int)blockIdx.x - off[s]);
}

__global__ __launch_bounds__(kBlock) void kp_current_orb(const CurFinArgs *args, const int *off, int nsys) {
  const int s = pack_find(off, nsys, blockIdx.x);
  const CurFinArgs a = args[s];
  const int lb = (int)blockIdx.x - off[s], gx = (a.no * a.nk + 3) / 4;
  current_orb_body(a, lb % gx, lb / gx);
}

__global__ __launch_bounds__(kBlock) void kp_current_sum(const CurFinArgs *args, const int *off, int nsys) {
  const int s = pack_find(off, nsys, blockIdx.x);
  const CurFinArgs a = args[s];
  current_sum_body(a, (int)blockIdx.x - off[s]);
}

__global__ __launch_bounds__(256) void kp_dipole_part(const DipArgs *args, const int *off, int nsys) {
  const int s = pack_find(off, nsys, blockIdx.x);
  const DipArgs a = args[s];
  dipole_part_body(a, (int)blockIdx.x - off[s]);
}

__global__ __launch_bounds__(256) void kp_dipole_sum(const DipArgs *args, const int *off, int nsys) {
  const int s = pack_find(off, nsys, blockIdx.x);
  const DipArgs a = args[s];
  dipole_sum_body(a);
}

// one packed launch: device records, block offsets, grid, dynamic LDS
struct Packed {
  void *args = nullptr;
  int *off = nullptr;
  int nblk = 0;
  size_t shmem = 0;
};

struct salmon_hip_batch {
  std::vector<salmon_hip_t> h;
  hipStream_t stream = nullptr;
  struct Sys {
    void *psi[2] = {nullptr, nullptr};
    void *work = nullptr;
    double *rho = nullptr, *curr_t = nullptr, *dip_t = nullptr;
    const double *actab = nullptr;
    int nt = 0;
    double *curr_s = nullptr, *dip_s = nullptr;  // this step's observables (scratch)
  };
  std::vector<Sys> sys;
  struct Group {
    int hc = 1;
    std::vector<int> mem;               // systems of this halo class
    int2 *d_tiles[2] = {nullptr, nullptr};  // middle terms, last term (different z-chunking)
    int ntiles[2] = {0, 0}, grid[2] = {0, 0};
    StencilArgs *d_args = nullptr;      // [parity][term][member]
    int zch = 0;                        // middle terms on the delayed-z kernel, slots of zch 64-element chunks (0: no)
  };
  std::vector<Group> groups;
  int n_hamil = 0;
  std::vector<double> zc;
  int z_chunks = 0;                     // 0: each system's own rule
  int grid_mult = 4;                    // persistent blocks = grid_mult x 256 (capped at the tile count)
  int *d_itt = nullptr;
  int itt = 1;                          // host copy of the step counter (the next step's index)
  int nt_min = 0;                       // steps the shortest field table covers
  // k_record tables
  double **d_csrc = nullptr, **d_cdst = nullptr, **d_dsrc = nullptr, **d_ddst = nullptr;
  int *d_cw = nullptr, *d_nt = nullptr;
  int parity = 0;
  bool prepared = false;
  std::vector<void *> allocs;
  // the per-system kernels of a phase (projections, scatters, observables) run on kLanes side
  // streams forked from and joined back to `stream`, so small systems overlap (captured as
  // parallel branches of the graph)
  std::vector<hipStream_t> lanes;
  std::vector<hipEvent_t> lane_ev;
  hipEvent_t fork_ev = nullptr;
  // packed launches of the per-system phases ([parity] = which orbital buffer is the input;
  // [term] for the fused KB passes); used when every system with projectors runs the fused
  // KB pass, else the step forks per-system launches onto the lanes
  bool packed = false;
  int lm = 4, cm = 0;                     // template choices common to the systems
  Packed p_kvec, p_proj[2], p_kbsh[2][kMaxComb], p_kbf[2][kMaxComb], p_last[2];
  Packed p_dpart[2], p_dsum[2], p_proj4[2], p_curl[2], p_curo, p_curs, p_dipp, p_dips;
};

namespace {

constexpr int kLanes = 4;  // side streams of a batch

template <typename T>
int balloc(salmon_hip_batch *b, T **p, size_t n) {
  if (int rc = dalloc(p, std::max<size_t>(n, 1))) return rc;
  b->allocs.push_back((void *)*p);
  return 0;
}

// Taylor buffers of system i at parity p: h[0] = input, h[1] = work, h[2..] = handle scratch, out
void batch_buffers(salmon_hip_batch *b, int i, int p, void **hh, void **out) {
  salmon_hip_t c = b->h[i];
  hh[0] = b->sys[i].psi[p];
  hh[1] = b->sys[i].work;
  for (int n = 2; n < b->n_hamil; ++n) hh[n] = c->tscr[n - 2];
  *out = b->sys[i].psi[1 - p];
}

// (re)build tile lists and descriptor tables for the current z_chunks / grid_mult
int batch_tables(salmon_hip_batch *b) {
  for (auto &g : b->groups) {
    // the terms without the Taylor sum run on the delayed-z tiles when every member's rows fit
    // their slots (nx <= 64; the group takes its largest member's chunk count) and its z extent
    // is a multiple of 4; z chunks are then rounded up to whole multiples of 4 planes
    int zch = 0;
    bool zok = g.hc <= 2;
    for (int i : g.mem) {
      const salmon_hip_t c = b->h[i];
      zok = zok && c->tilez && c->nx <= 64 && c->nz % 4 == 0 && c->TY * c->nx <= kTile &&
            (c->TY + 8) * c->nx <= 3 * kTile;
      zch = std::max(zch, tilez_chunks(c->TY, c->nx));
    }
    g.zch = zok ? zch : 0;
    std::vector<int2> tiles[2];
    std::vector<StencilArgs> args;
    for (int p = 0; p < 2; ++p)
      for (int n = 1; n <= b->n_hamil; ++n)
        for (size_t m = 0; m < g.mem.size(); ++m) {
          const int i = g.mem[m];
          salmon_hip_t c = b->h[i];
          void *hh[kMaxComb + 1], *out;
          batch_buffers(b, i, p, hh, &out);
          const bool fin = n == b->n_hamil;
          int bs = 0;
          const int save = c->mid_nzt;
          c->mid_nzt = b->z_chunks;
          StencilArgs a = stencil_args(c, hh[n - 1], fin ? out : hh[n], fin, hh, n - 1, b->zc.data(), &bs, true);
          c->mid_nzt = save;
          if (g.zch && !fin) {
            a.TZ = (a.TZ + 3) / 4 * 4;
            a.nzt = (a.nz + a.TZ - 1) / a.TZ;
            a.nblk = c->norb * a.nyt * a.nzt;
          }
          args.push_back(a);
          if (p == 0 && (n == 1 || fin))
            for (int t = 0; t < a.nblk; ++t) tiles[fin ? 1 : 0].push_back(make_int2((int)m, t));
        }
    for (int k = 0; k < 2; ++k) {
      dfree(g.d_tiles[k]);
      if (int rc = dalloc(&g.d_tiles[k], tiles[k].size())) return rc;
      HIPCHK(hipMemcpy(g.d_tiles[k], tiles[k].data(), sizeof(int2) * tiles[k].size(), hipMemcpyHostToDevice));
      g.ntiles[k] = (int)tiles[k].size();
      g.grid[k] = std::min(g.ntiles[k], b->grid_mult * 256);
    }
    dfree(g.d_args);
    if (int rc = dalloc(&g.d_args, args.size())) return rc;
    HIPCHK(hipMemcpy(g.d_args, args.data(), sizeof(StencilArgs) * args.size(), hipMemcpyHostToDevice));
  }
  return 0;
}

int batch_stencil(salmon_hip_batch *b, int p, int n) {
  const int nc = n == b->n_hamil ? b->n_hamil - 1 : -1;
  for (auto &g : b->groups) {
    const StencilArgs *args = g.d_args + ((size_t)p * b->n_hamil + (n - 1)) * g.mem.size();
    const int k = nc >= 0 ? 1 : 0;
    const int rc = (nc < 0 && g.zch) ? launch_packed_z(b->stream, g.zch, args, g.d_tiles[0], g.ntiles[0], g.grid[0])
                                     : launch_packed(b->stream, g.hc, nc, args, g.d_tiles[k], g.ntiles[k], g.grid[k]);
    if (rc) return rc;
  }
  return 0;
}

// fork: system i's kernels go to lane i % kLanes after all work enqueued on the batch stream
int batch_fork(salmon_hip_batch *b) {
  HIPCHK(hipEventRecord(b->fork_ev, b->stream));
  for (size_t k = 0; k < b->lanes.size(); ++k) HIPCHK(hipStreamWaitEvent(b->lanes[k], b->fork_ev, 0));
  for (size_t i = 0; i < b->h.size(); ++i) b->h[i]->stream = b->lanes[i % b->lanes.size()];
  return 0;
}

// join: later work on the batch stream waits for every lane
int batch_join(salmon_hip_batch *b) {
  for (size_t i = 0; i < b->h.size(); ++i) b->h[i]->stream = b->stream;
  for (size_t k = 0; k < b->lanes.size(); ++k) {
    HIPCHK(hipEventRecord(b->lane_ev[k], b->lanes[k]));
    HIPCHK(hipStreamWaitEvent(b->stream, b->lane_ev[k], 0));
  }
  return 0;
}

int batch_step(salmon_hip_batch *b, int p) {
  const int N = b->n_hamil;
  const size_t ns = b->h.size();
  int rc = batch_fork(b);
  for (size_t i = 0; i < ns && !rc; ++i) rc = update_kvector_impl(b->h[i], nullptr, b->sys[i].actab, b->d_itt, b->sys[i].nt + 1);
  if (int r2 = batch_join(b)) return r2;
  if (rc) return rc;
  // the KB term of every term but the last runs fused with the next term's projection, so a
  // system's projection phase runs only for the first term (or without the fused pass)
  for (int n = 1; n <= N; ++n) {
    bool any = false;
    for (size_t i = 0; i < ns; ++i) any = any || (b->h[i]->natom > 0 && (n == 1 || !kb_fused_on(b->h[i])));
    if (any) {
      rc = batch_fork(b);
      for (size_t i = 0; i < ns && !rc; ++i) {
        salmon_hip_t c = b->h[i];
        void *hh[kMaxComb + 1], *out;
        batch_buffers(b, (int)i, p, hh, &out);
        if (c->natom > 0 && (n == 1 || !kb_fused_on(c))) rc = launch_proj(c, hh[n - 1], 1);
      }
      if (int r2 = batch_join(b)) return r2;
      if (rc) return rc;
    }
    if (int r2 = batch_stencil(b, p, n)) return r2;
    rc = batch_fork(b);
    for (size_t i = 0; i < ns && !rc; ++i) {
      salmon_hip_t c = b->h[i];
      void *hh[kMaxComb + 1], *out;
      batch_buffers(b, (int)i, p, hh, &out);
      rc = join_proj(c);
      if (!rc && n < N && kb_fused_on(c)) {
        rc = launch_kb_fused(c, hh[n], false);
      } else if (!rc && c->natom > 0 && c->nlma > 0) {
        const double2 sc = n == N ? make_double2(b->zc[2 * (N - 1)], b->zc[2 * (N - 1) + 1]) : make_double2(1.0, 0.0);
        rc = launch_scatter(c, n == N ? out : hh[n], sc);
      }
    }
    if (int r2 = batch_join(b)) return r2;
    if (rc) return rc;
  }
  rc = batch_fork(b);
  for (size_t i = 0; i < ns && !rc; ++i) {
    salmon_hip_t c = b->h[i];
    void *out = b->sys[i].psi[1 - p];
    rc = salmon_hip_density(c, out, b->sys[i].rho);
    if (!rc) rc = current_impl(c, nullptr, b->sys[i].actab, b->d_itt, out, b->sys[i].curr_s, b->sys[i].nt + 1);
    if (!rc) rc = salmon_hip_dipole(c, b->sys[i].rho, b->sys[i].dip_s);
  }
  if (int r2 = batch_join(b)) return r2;
  if (rc) return rc;
  hipLaunchKernelGGL(k_record, dim3(1), dim3(std::max<size_t>(64, (ns + 63) / 64 * 64)), 0, b->stream, b->d_csrc,
                     b->d_cdst, b->d_cw, b->d_dsrc, b->d_ddst, b->d_nt, (int)ns, b->d_itt);
  LAUNCHCHK("k_record");
  return 0;
}

template <typename A>
int pack_make(salmon_hip_batch *b, Packed &P, const std::vector<A> &recs, const std::vector<int> &cnt, bool pad8,
              size_t shmem = 0) {
  const int n = (int)recs.size();
  std::vector<int> off(n + 1, 0);
  for (int i = 0; i < n; ++i) off[i + 1] = off[i] + (pad8 ? (cnt[i] + 7) / 8 * 8 : cnt[i]);
  A *da = nullptr;
  int *doff = nullptr;
  if (int rc = balloc(b, &da, (size_t)n)) return rc;
  if (int rc = balloc(b, &doff, (size_t)n + 1)) return rc;
  HIPCHK(hipMemcpy(da, recs.data(), sizeof(A) * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(doff, off.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
  P.args = da;
  P.off = doff;
  P.nblk = off[n];
  P.shmem = shmem;
  return 0;
}

// Records of the packed launches for both buffer parities.  The two c buffers of each handle
// alternate over the Taylor terms exactly as in launch_kb_fused: the first projection writes X,
// the fused pass of term n reads (n odd ? X : Y) and writes the other, the last scatter reads
// what the pass of term N-1 wrote.
int batch_packed_tables(salmon_hip_batch *b) {
  b->packed = false;
  const int ns = (int)b->h.size(), N = b->n_hamil;
  int cmax = 0;
  for (int i = 0; i < ns; ++i) {
    salmon_hip_t c = b->h[i];
    if (c->natom > 0 && c->nlma > 0 && !kb_fused_on(c)) return 0;  // per-system launches on the lanes
    cmax = std::max(cmax, c->natom > 0 ? c->cmax : 0);
  }
  if (cmax > kMaxLmaPerAtom) return 0;
  b->lm = cmax <= 4 ? 4 : cmax <= 9 ? 9 : 16;
  b->cm = cmax <= kScatCM ? kScatCM : 0;
  auto kb = [&](salmon_hip_t c) { return kb_fused_on(c); };
  auto scat_blocks = [&](salmon_hip_t c, int ntp) {
    return ntp > 0 ? ((ntp + kBlock - 1) / kBlock) * c->nk * ((c->nb + kScatG - 1) / kScatG) : 0;
  };
  auto proj_blocks = [&](salmon_hip_t c) { return ((c->nb + kProjOrb - 1) / kProjOrb) * c->natom * c->nk; };
  std::vector<int> cnt(ns);
  {
    std::vector<KvecArgs> r(ns);
    for (int i = 0; i < ns; ++i) {
      r[i] = kvec_args(b->h[i], nullptr, b->sys[i].actab, b->d_itt, b->sys[i].nt + 1);
      cnt[i] = std::max(1, b->h[i]->nlma) * b->h[i]->nk;
    }
    if (int rc = pack_make(b, b->p_kvec, r, cnt, false)) return rc;
  }
  size_t kb_shm = 0, cur_shm = 0;
  for (int i = 0; i < ns; ++i) {
    kb_shm = std::max(kb_shm, sizeof(double2) * (size_t)b->h[i]->cmax * kProjOrb);
    cur_shm = std::max(cur_shm, cur_shmem(b->h[i]));
  }
  for (int p = 0; p < 2; ++p) {
    std::vector<ProjArgs> pr(ns), p4(ns);
    std::vector<ScatterArgs> last(ns);
    std::vector<DensArgs> de(ns);
    std::vector<CurArgs> cl(ns);
    std::vector<int> c4(ns), cl_n(ns), cd(ns), cs(ns), cls(ns);
    for (int i = 0; i < ns; ++i) {
      salmon_hip_t c = b->h[i];
      void *hh[kMaxComb + 1], *out;
      batch_buffers(b, i, p, hh, &out);
      pr[i] = proj_args(c, hh[0], c->d_c);
      cnt[i] = kb(c) ? proj_blocks(c) : 0;
      const double2 *cl_buf = ((N - 1) % 2 == 1) ? c->d_c2 : c->d_c;
      last[i] = scatter_args(c, out, make_double2(b->zc[2 * (N - 1)], b->zc[2 * (N - 1) + 1]), cl_buf, false, false);
      cls[i] = kb(c) ? scat_blocks(c, c->ntp) : 0;
      de[i] = dens_args(c, out, b->sys[i].rho, c->d_denspart);
      const int gx = (int)((c->ngrid + kBlock - 1) / kBlock);
      cd[i] = gx * c->dens_ngrp * c->nspin;
      cs[i] = c->dens_ngrp > 1 ? (int)((c->ngrid * c->nspin + kBlock - 1) / kBlock) : 0;
      p4[i] = proj_args(c, out, c->d_curnl);
      c4[i] = c->natom > 0 ? proj_blocks(c) : 0;
      cl[i] = cur_args(c, out);
      cl_n[i] = c->norb * c->nyt;
    }
    if (int rc = pack_make(b, b->p_proj[p], pr, cnt, true)) return rc;
    if (int rc = pack_make(b, b->p_last[p], last, cls, false)) return rc;
    if (int rc = pack_make(b, b->p_dpart[p], de, cd, false)) return rc;
    if (int rc = pack_make(b, b->p_dsum[p], de, cs, false)) return rc;
    if (int rc = pack_make(b, b->p_proj4[p], p4, c4, true)) return rc;
    if (int rc = pack_make(b, b->p_curl[p], cl, cl_n, false, cur_shm)) return rc;
    for (int n = 1; n < N; ++n) {
      std::vector<ScatterArgs> sh(ns);
      std::vector<ProjScatArgs> fu(ns);
      std::vector<int> shn(ns), fun(ns);
      for (int i = 0; i < ns; ++i) {
        salmon_hip_t c = b->h[i];
        void *hh[kMaxComb + 1], *out;
        batch_buffers(b, i, p, hh, &out);
        const double2 *cin = (n % 2 == 1) ? c->d_c : c->d_c2;
        double2 *cout = (n % 2 == 1) ? c->d_c2 : c->d_c;
        sh[i] = scatter_args(c, hh[n], make_double2(1.0, 0.0), cin, true, false);
        shn[i] = kb(c) ? scat_blocks(c, c->ntp_shared) : 0;
        fu[i] = projscat_args(c, hh[n], cin, cout);
        fun[i] = kb(c) ? proj_blocks(c) : 0;
      }
      if (int rc = pack_make(b, b->p_kbsh[p][n - 1], sh, shn, false)) return rc;
      if (int rc = pack_make(b, b->p_kbf[p][n - 1], fu, fun, true, kb_shm)) return rc;
    }
  }
  std::vector<CurFinArgs> cf(ns);
  std::vector<DipArgs> dp(ns);
  std::vector<int> co(ns), csum(ns), dpn(ns, kDipBlocks), dsn(ns, 1);
  for (int i = 0; i < ns; ++i) {
    salmon_hip_t c = b->h[i];
    cf[i] = curfin_args(c, nullptr, b->sys[i].actab, b->d_itt, b->sys[i].curr_s, b->sys[i].nt + 1);
    co[i] = ((c->no * c->nk + 3) / 4) * c->nspin;
    csum[i] = c->nspin;
    dp[i] = dip_args(c, b->sys[i].rho, b->sys[i].dip_s);
  }
  if (int rc = pack_make(b, b->p_curo, cf, co, false)) return rc;
  if (int rc = pack_make(b, b->p_curs, cf, csum, false)) return rc;
  if (int rc = pack_make(b, b->p_dipp, dp, dpn, false)) return rc;
  if (int rc = pack_make(b, b->p_dips, dp, dsn, false)) return rc;
  b->packed = true;
  return 0;
}

#define PACKED_LAUNCH(kern, P, A)                                                                      \
  do {                                                                                                 \
    if ((P).nblk > 0) {                                                                                \
      hipLaunchKernelGGL(kern, dim3((unsigned)(P).nblk), dim3(kBlock), (P).shmem, b->stream,          \
                         (const A *)(P).args, (const int *)(P).off, ns);                               \
      LAUNCHCHK(#kern);                                                                                \
    }                                                                                                  \
  } while (0)

// One step of every system with one launch per phase (the packed stencil per halo class and term).
int batch_step_packed(salmon_hip_batch *b, int p) {
  const int N = b->n_hamil;
  const int ns = (int)b->h.size();
  PACKED_LAUNCH(kp_update_kvector, b->p_kvec, KvecArgs);
  if (b->lm == 4) PACKED_LAUNCH((kp_proj<1, 4>), b->p_proj[p], ProjArgs);
  else if (b->lm == 9) PACKED_LAUNCH((kp_proj<1, 9>), b->p_proj[p], ProjArgs);
  else PACKED_LAUNCH((kp_proj<1, 16>), b->p_proj[p], ProjArgs);
  for (int n = 1; n <= N; ++n) {
    if (int rc = batch_stencil(b, p, n)) return rc;
    const Packed &S = n < N ? b->p_kbsh[p][n - 1] : b->p_last[p];
    if (b->cm == kScatCM) PACKED_LAUNCH(kp_nl_scatter<kScatCM>, S, ScatterArgs);
    else PACKED_LAUNCH(kp_nl_scatter<0>, S, ScatterArgs);
    if (n < N) {
      const Packed &F = b->p_kbf[p][n - 1];
      if (b->lm == 4) PACKED_LAUNCH(kp_proj_scatter<4>, F, ProjScatArgs);
      else if (b->lm == 9) PACKED_LAUNCH(kp_proj_scatter<9>, F, ProjScatArgs);
      else PACKED_LAUNCH(kp_proj_scatter<16>, F, ProjScatArgs);
    }
  }
  PACKED_LAUNCH(kp_density_part, b->p_dpart[p], DensArgs);
  PACKED_LAUNCH(kp_density_sum, b->p_dsum[p], DensArgs);
  if (b->lm == 4) PACKED_LAUNCH((kp_proj<4, 4>), b->p_proj4[p], ProjArgs);
  else if (b->lm == 9) PACKED_LAUNCH((kp_proj<4, 9>), b->p_proj4[p], ProjArgs);
  else PACKED_LAUNCH((kp_proj<4, 16>), b->p_proj4[p], ProjArgs);
  PACKED_LAUNCH(kp_current_local, b->p_curl[p], CurArgs);
  PACKED_LAUNCH(kp_current_orb, b->p_curo, CurFinArgs);
  PACKED_LAUNCH(kp_current_sum, b->p_curs, CurFinArgs);
  PACKED_LAUNCH(kp_dipole_part, b->p_dipp, DipArgs);
  PACKED_LAUNCH(kp_dipole_sum, b->p_dips, DipArgs);
  hipLaunchKernelGGL(k_record, dim3(1), dim3(std::max<size_t>(64, (ns + 63) / 64 * 64)), 0, b->stream, b->d_csrc,
                     b->d_cdst, b->d_cw, b->d_dsrc, b->d_ddst, b->d_nt, ns, b->d_itt);
  LAUNCHCHK("k_record");
  return 0;
}

int bcheck(salmon_hip_batch_t b) {
  if (!b) return fail(SALMON_HIP_EINVAL, "null batch");
  return 0;
}

}  // namespace

extern "C" {

int salmon_hip_batch_create(const salmon_hip_t *handles, int n, void *stream, salmon_hip_batch_t *out) {
  if (!out || !handles || n < 1 || n > 1024) return fail(SALMON_HIP_EINVAL, "batch of %d systems", n);
  *out = nullptr;
  salmon_hip_batch *b = new salmon_hip_batch();
  for (int i = 0; i < n; ++i)
    if (!handles[i]) { delete b; return fail(SALMON_HIP_EINVAL, "handle %d is null", i); }
  for (int i = 0; i < n; ++i) {
    handles[i]->in_batch = true;
    handles[i]->staged = false;
    handles[i]->pre_state = 0;
  }
  // the null stream cannot be captured into a graph: NULL selects the first handle's own stream
  b->stream = stream ? (hipStream_t)stream : handles[0]->own;
  for (int i = 0; i < n; ++i) {
    b->h.push_back(handles[i]);
    handles[i]->stream = b->stream;
  }
  b->sys.resize(n);
  for (int k = 0; k < std::min(n, kLanes); ++k) {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      salmon_hip_batch_destroy(b);
      return fail(SALMON_HIP_EHIP, "batch lane stream");
    }
    b->lanes.push_back(st);
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      salmon_hip_batch_destroy(b);
      return fail(SALMON_HIP_EHIP, "batch lane event");
    }
    b->lane_ev.push_back(ev);
  }
  if (hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming) != hipSuccess) {
    salmon_hip_batch_destroy(b);
    return fail(SALMON_HIP_EHIP, "batch fork event");
  }
  int rc = balloc(b, &b->d_itt, 1);
  if (!rc) rc = balloc(b, &b->d_csrc, n);
  if (!rc) rc = balloc(b, &b->d_cdst, n);
  if (!rc) rc = balloc(b, &b->d_dsrc, n);
  if (!rc) rc = balloc(b, &b->d_ddst, n);
  if (!rc) rc = balloc(b, &b->d_cw, n);
  if (!rc) rc = balloc(b, &b->d_nt, n);
  for (int i = 0; i < n && !rc; ++i) {
    rc = balloc(b, &b->sys[i].curr_s, 3 * (size_t)handles[i]->nspin);
    if (!rc) rc = balloc(b, &b->sys[i].dip_s, 4);
  }
  if (rc) { salmon_hip_batch_destroy(b); return rc; }
  *out = b;
  return 0;
}

int salmon_hip_batch_destroy(salmon_hip_batch_t b) {
  if (!b) return 0;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (hipStream_t st : b->lanes) (void)hipStreamSynchronize(st);
  for (auto &g : b->groups) { dfree(g.d_tiles[0]); dfree(g.d_tiles[1]); dfree(g.d_args); }
  for (void *p : b->allocs) (void)hipFree(p);
  for (hipStream_t st : b->lanes) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : b->lane_ev) (void)hipEventDestroy(ev);
  if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
  for (salmon_hip_t c : b->h) {
    c->stream = b->stream;
    c->in_batch = false;
  }
  delete b;
  return 0;
}

int salmon_hip_batch_set_system(salmon_hip_batch_t b, int i, void *psi_a, void *psi_b, void *work, double *rho_dev,
                                const double *ac_dev, int nt, double *curr_t_dev, double *dip_t_dev) {
  if (int rc = bcheck(b)) return rc;
  if (i < 0 || i >= (int)b->h.size()) return fail(SALMON_HIP_EINVAL, "system %d of %zu", i, b->h.size());
  if (!psi_a || !psi_b || !work || !rho_dev || !ac_dev || !curr_t_dev || !dip_t_dev || nt < 1)
    return fail(SALMON_HIP_EINVAL, "system %d: null buffer or nt < 1", i);
  auto &y = b->sys[i];
  y.psi[0] = psi_a; y.psi[1] = psi_b; y.work = work; y.rho = rho_dev; y.actab = ac_dev; y.nt = nt;
  y.curr_t = curr_t_dev; y.dip_t = dip_t_dev;
  b->prepared = false;
  return 0;
}

int salmon_hip_batch_prepare(salmon_hip_batch_t b, const double *zc, int n_hamil) {
  if (int rc = bcheck(b)) return rc;
  if (!zc || n_hamil < 2 || n_hamil > kMaxComb + 1)
    return fail(SALMON_HIP_EINVAL, "zc null or n_hamil=%d outside 2..%d", n_hamil, kMaxComb + 1);
  const int ns = (int)b->h.size();
  for (int i = 0; i < ns; ++i) {
    salmon_hip_t c = b->h[i];
    if (!b->sys[i].psi[0]) return fail(SALMON_HIP_ESTATE, "system %d: set_system not called", i);
    if (!c->have_V) return fail(SALMON_HIP_ESTATE, "system %d: V_local not set", i);
    if (c->prec != SALMON_HIP_PREC_FP64) return fail(SALMON_HIP_EUNSUPPORTED, "batched step is FP64");
    while ((int)c->tscr.size() < n_hamil - 2) {
      double2 *q = nullptr;
      if (int rc = dalloc(&q, c->osz * c->norb)) return rc;
      c->tscr.push_back(q);
    }
    c->have_k = true;  // the phase table is rebuilt from the device field table every step
  }
  b->n_hamil = n_hamil;
  b->zc.assign(zc, zc + 2 * n_hamil);
  for (auto &g : b->groups) { dfree(g.d_tiles[0]); dfree(g.d_tiles[1]); dfree(g.d_args); }
  b->groups.clear();
  for (int i = 0; i < ns; ++i) {
    const int hc = (8 * b->h[i]->nx + kTile - 1) / kTile;
    const int cls = hc <= 1 ? 1 : hc <= 2 ? 2 : hc <= 4 ? 4 : 8;
    salmon_hip_batch::Group *g = nullptr;
    for (auto &q : b->groups)
      if (q.hc == cls) g = &q;
    if (!g) {
      b->groups.push_back(salmon_hip_batch::Group());
      g = &b->groups.back();
      g->hc = cls;
    }
    g->mem.push_back(i);
  }
  if (int rc = batch_tables(b)) return rc;
  if (int rc = batch_packed_tables(b)) return rc;
  std::vector<const double *> cs(ns), ds(ns);
  std::vector<double *> cd(ns), dd(ns);
  std::vector<int> cw(ns), nt(ns);
  for (int i = 0; i < ns; ++i) {
    cs[i] = b->sys[i].curr_s; ds[i] = b->sys[i].dip_s; cd[i] = b->sys[i].curr_t; dd[i] = b->sys[i].dip_t;
    cw[i] = 3 * b->h[i]->nspin; nt[i] = b->sys[i].nt;
  }
  HIPCHK(hipMemcpy(b->d_csrc, cs.data(), sizeof(double *) * ns, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->d_dsrc, ds.data(), sizeof(double *) * ns, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->d_cdst, cd.data(), sizeof(double *) * ns, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->d_ddst, dd.data(), sizeof(double *) * ns, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->d_cw, cw.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->d_nt, nt.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
  b->nt_min = nt[0];
  for (int i = 1; i < ns; ++i) b->nt_min = std::min(b->nt_min, nt[i]);
  b->prepared = true;
  b->parity = 0;
  return salmon_hip_batch_set_time(b, 1);
}

int salmon_hip_batch_set_time(salmon_hip_batch_t b, int itt) {
  if (int rc = bcheck(b)) return rc;
  if (itt < 1 || (b->prepared && itt > b->nt_min + 1))
    return fail(SALMON_HIP_EINVAL, "step index %d outside 1..%d (the field tables cover %d steps)", itt,
                b->nt_min + 1, b->nt_min);
  b->itt = itt;
  hipLaunchKernelGGL(k_set_itt, dim3(1), dim3(1), 0, b->stream, b->d_itt, itt);
  LAUNCHCHK("k_set_itt");
  b->parity = (itt - 1) & 1;
  return 0;
}

int salmon_hip_batch_step(salmon_hip_batch_t b) {
  if (int rc = bcheck(b)) return rc;
  if (!b->prepared) return fail(SALMON_HIP_ESTATE, "batch_prepare not called");
  if (b->itt > b->nt_min)
    return fail(SALMON_HIP_EINVAL, "step %d past the field tables (%d steps)", b->itt, b->nt_min);
  if (int rc = b->packed ? batch_step_packed(b, b->parity) : batch_step(b, b->parity)) return rc;
  b->parity ^= 1;
  b->itt += 1;
  return 0;
}

// Graph capture of the batched step is switched off: one capture (stream capture of the step
// with the per-system launches forked onto side streams and joined by events) ended in a host
// segmentation fault inside the capture whose cause was not found.  Both entry points report
// SALMON_HIP_EUNSUPPORTED; salmon_hip_batch_step runs the same step with eager launches.
int salmon_hip_batch_capture(salmon_hip_batch_t b) {
  if (int rc = bcheck(b)) return rc;
  return fail(SALMON_HIP_EUNSUPPORTED, "graph capture of the batched step is switched off (use batch_step)");
}

int salmon_hip_batch_replay(salmon_hip_batch_t b, int npairs) {
  if (int rc = bcheck(b)) return rc;
  (void)npairs;
  return fail(SALMON_HIP_EUNSUPPORTED, "graph capture of the batched step is switched off (use batch_step)");
}

int salmon_hip_batch_autotune(salmon_hip_batch_t b, int reps, int *choice, double *best_ms) {
  if (int rc = bcheck(b)) return rc;
  if (!b->prepared) return fail(SALMON_HIP_ESTATE, "batch_prepare not called");
  if (reps < 1) reps = 3;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  double best = 1e30;
  int bz = 0, bg = 4;
  int rc = 0;
  for (int z : {0, 1, 2}) {
    for (int gm : {1, 2, 4, 8, 1 << 20}) {
      b->z_chunks = z;
      b->grid_mult = gm;
      if ((rc = batch_tables(b))) break;
      // middle-term packed stencil of every group (parity 0, term 2: reads h1, writes h2)
      if ((rc = batch_stencil(b, 0, 2))) break;
      if (hipEventRecord(e0, b->stream) != hipSuccess) { rc = fail(SALMON_HIP_EHIP, "event"); break; }
      for (int r = 0; r < reps && !rc; ++r) rc = batch_stencil(b, 0, 2);
      float ms = 0.f;
      if (rc || hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        if (!rc) rc = fail(SALMON_HIP_EHIP, "autotune timing");
        break;
      }
      if (ms / reps < best) { best = ms / reps; bz = z; bg = gm; }
    }
    if (rc) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  b->z_chunks = bz;
  b->grid_mult = bg;
  if (int rc2 = batch_tables(b)) return rc2;
  if (choice) *choice = 100 * bz + (bg > 64 ? 99 : bg);
  if (best_ms) *best_ms = best;
  return 0;
}

int salmon_hip_prof_start(salmon_hip_t c, int max_launches) {
  if (int rc = check(c)) return rc;
  if (max_launches < 1) return fail(SALMON_HIP_EINVAL, "max_launches=%d", max_launches);
  HIPCHK(hipStreamSynchronize(c->stream));
  while (c->ev.size() < 2 * (size_t)max_launches) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    c->ev.push_back(e);
  }
  c->ev_kind.assign(c->ev.size() / 2, -1);
  c->ev_used = 0;
  c->prof = true;
  return 0;
}

int salmon_hip_prof_stop(salmon_hip_t c, int kind, double *total_ms, int *count) {
  if (int rc = check(c)) return rc;
  if (!total_ms || !count) return fail(SALMON_HIP_EINVAL, "null output");
  HIPCHK(hipStreamSynchronize(c->stream));
  double tot = 0.0;
  int n = 0;
  for (size_t i = 0; i < c->ev_used; ++i) {
    if (c->ev_kind[i] != kind) continue;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[2 * i], c->ev[2 * i + 1]));
    tot += ms;
    n += 1;
  }
  *total_ms = tot;
  *count = n;
  return 0;
}

int salmon_hip_prof_end(salmon_hip_t c) {
  if (int rc = check(c)) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  c->prof = false;
  c->ev_used = 0;
  return 0;
}

int salmon_hip_synchronize(salmon_hip_t c) {
  if (int rc = check(c)) return rc;
  HIPCHK(hipStreamSynchronize(c->side));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int salmon_hip_malloc(void **dptr, size_t bytes) {
  if (!dptr) return fail(SALMON_HIP_EINVAL, "dptr is null");
  hipError_t e = hipMalloc(dptr, bytes);
  if (e != hipSuccess) return fail(SALMON_HIP_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  return 0;
}

int salmon_hip_free(void *dptr) {
  HIPCHK(hipFree(dptr));
  return 0;
}

int salmon_hip_memcpy_htod(salmon_hip_t c, void *dst, const void *src, size_t bytes) {
  if (int rc = check(c)) return rc;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int salmon_hip_memcpy_dtoh(salmon_hip_t c, void *dst, const void *src, size_t bytes) {
  if (int rc = check(c)) return rc;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

}  // extern "C"
