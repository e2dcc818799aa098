/*
 * salmon_hip.h — C ABI of the MI355X (gfx950) real-time Kohn-Sham propagation
 * hot path.  Plain C types only (ints, doubles, raw pointers); callable from
 * Fortran through iso_c_binding (salmon2_amd/fortran/salmon_hip_mod.f90), from
 * C/C++, or from Python ctypes.
 *
 * Layout contract (identical to SALMON's, so the Fortran driver hands over its
 * own arrays with plain memory copies):
 *   orbitals  zwf(ix, iy, iz, ispin, io, ik): complex(8), x unit-stride, y
 *             leading extent `ly` (ny or ny + 1 with USE_OPT_ARRAY_PADDING),
 *             z extent nz; src/common/structures.f90:191-196,
 *             src/common/initialization.f90:606-609.  The padding row is never
 *             read and never written.
 *   V_local   (ix, iy, iz, ispin) real(8), unpadded (structures.f90:539-553).
 *   jxyz      (3, nps, natom) 1-based grid indices, rxyz (3, nps, natom),
 *   uv        (nps, nlma), ia_tbl (nlma) 1-based, rinv_uvu (nlma)
 *             (structures.f90:254-319).
 * Orbital arguments of hpsi/taylor/density/current are DEVICE pointers to the
 * whole block of nspin*no*nk orbitals; set-up arguments are HOST pointers.
 * The periodic single-process case is the one supported: no halo, wrap-around
 * by modulo (idx(j) = mod(j+n-1, n)+1, initialization.f90:574-587), orthogonal
 * lattice, collinear spin, no DFT+U.
 *
 * Every function returns 0 on success or a SALMON_HIP_E* code; the message of
 * the last failure on the calling thread is salmon_hip_last_error().  Work is
 * enqueued on the handle's stream; results are visible after
 * salmon_hip_synchronize() or to later work on the same stream.
 */
#ifndef SALMON_HIP_H
#define SALMON_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SALMON_HIP_OK 0
#define SALMON_HIP_EINVAL 1       /* invalid argument / shape */
#define SALMON_HIP_EHIP 2         /* HIP runtime error */
#define SALMON_HIP_ENOMEM 3       /* device allocation failed */
#define SALMON_HIP_EUNSUPPORTED 4 /* configuration outside the supported path */
#define SALMON_HIP_ESTATE 5       /* call order violated (e.g. hpsi before set_vlocal) */

typedef struct salmon_hip_ctx *salmon_hip_t;

/* Message of the last failure on this thread ("" if none). */
const char *salmon_hip_last_error(void);

/*
 * Create a handle for one grid and orbital set (replaces the device set-up the
 * reference does with `!$acc enter data` in src/rt/initialization_rt.f90:234-240).
 *   coef_lap, coef_nab: s_stencil%coef_lap(4,3), %coef_nab(4,3); coef_lap0: %coef_lap0
 *   (src/common/initialization.f90:240-255).
 *   vec_k (3, nk), rocc (no, nk, nspin), wtk (nk): s_dft_system%vec_k/rocc/wtk.
 *   hgs (3), hvol: grid spacing and volume element (src/common/lattice.f90:40).
 */
int salmon_hip_create(int nx, int ny, int nz, int ly, int nspin, int no, int nk,
                      const double *coef_lap, const double *coef_nab, double coef_lap0,
                      const double *vec_k, const double *rocc, const double *wtk,
                      const double *hgs, double hvol, salmon_hip_t *out);

/* Release every device resource of the handle. */
int salmon_hip_destroy(salmon_hip_t h);

/* Enqueue subsequent work on `stream` (a hipStream_t; NULL = HIP's null stream, ordered with
   every blocking stream).  A new handle starts on a stream of its own (salmon_hip_get_stream). */
int salmon_hip_set_stream(salmon_hip_t h, void *stream);

/* The handle's current stream (a hipStream_t). */
void *salmon_hip_get_stream(salmon_hip_t h);

/*
 * Kleinman-Bylander projector tables, s_pp_grid%{mps,jxyz,rxyz,ia_tbl,rinv_uvu,uv}
 * (set up by init_ps, src/atom/pp/prep_pp.f90:22-527).  Projectors of one atom
 * must be contiguous in ilma (as set_lma orders them).  natom = 0 disables the
 * nonlocal term.
 */
int salmon_hip_set_nonlocal(salmon_hip_t h, int natom, int nps, const int *mps, const int *jxyz,
                            const double *rxyz, int nlma, const int *ia_tbl,
                            const double *rinv_uvu, const double *uv);

/* V_local(ix,iy,iz,ispin), host array (update_vlocal, src/common/hamiltonian.f90:813-843). */
int salmon_hip_set_vlocal(salmon_hip_t h, const double *vlocal);

/* Same from a device array (e.g. a potential produced on the GPU). */
int salmon_hip_set_vlocal_device(salmon_hip_t h, const double *vlocal_dev);

/*
 * Set system%vec_Ac and rebuild the phase table
 * zekr_uV(j,ilma,ik) = conjg(exp(i (k+A).r_j)) uv(j,ilma) and the per-k stencil
 * terms k_lap0 = lap0 + |k+A|^2/2, k_nabt(n,d) = (k+A)_d coef_nab(n,d):
 * update_kvector_nonlocalpt (src/common/hamiltonian.f90:847-903) and
 * src/common/hamiltonian.f90:146-155.
 */
int salmon_hip_update_kvector(salmon_hip_t h, const double *vec_Ac);

/*
 * Set system%vec_Ac for the stencil terms only (k_lap0, k_nabt), leaving the
 * phase table of the last update_kvector call in place: hpsi reads vec_Ac at
 * every call (src/common/hamiltonian.f90:146-155) while zekr_uV changes only in
 * update_kvector_nonlocalpt.
 */
int salmon_hip_set_vec_ac(salmon_hip_t h, const double *vec_Ac);

/*
 * htpsi = H tpsi for every orbital (hpsi, src/common/hamiltonian.f90:27-229,424-435:
 * 25-point stencil + V_local + k+A gradient, then zpseudo,
 * src/common/nonlocal_potential.f90:175-390).  Device pointers; tpsi != htpsi.
 */
int salmon_hip_hpsi(salmon_hip_t h, const void *tpsi, void *htpsi);

/*
 * Taylor propagation tspsi_out = sum_{n=0..n_hamil} zc(n) H^n tspsi_in with
 * zc(0) = 1 (taylor, src/rt/taylor.f90:21-121), summed in the reference's
 * order (((psi + zc1 H psi) + zc2 H^2 psi) + ...).  zc: host complex(8)[n_hamil]
 * as interleaved (re, im); 1 <= n_hamil <= 6.  Buffer contract, stricter than
 * the reference's (which leaves H^n_hamil psi in tspsi_in; its RT driver never
 * reads it, src/rt/main_tddft.f90:91-99): tspsi_in is left unchanged and
 * sshtpsi holds H tspsi_in on return (n_hamil >= 2).  H^2 .. H^(n_hamil-1)
 * psi live in scratch owned by the handle (allocated on first use).
 */
int salmon_hip_taylor(salmon_hip_t h, const double *zc, int n_hamil, void *tspsi_in,
                      void *tspsi_out, void *sshtpsi);

/*
 * The Taylor step of tspsi_in when only V_local changed since the last salmon_hip_taylor
 * call on the same tspsi_in and k-vector (no counterpart in the reference): the corrector
 * of the predictor-corrector (time_evolution_step.f90:354-423) repeats the predictor's
 * Taylor step with the midpoint potential, and H_mid psi = H psi + (V_mid - V) psi, so the
 * first H application is replaced by h_1 += dv * psi with dv_dev = V_mid - V (device array
 * (nx, ny, nz, nspin)); the later terms run as in salmon_hip_taylor with the handle's
 * current V_local, which must already be V_mid.  h_1 is sshtpsi in FP64 mode (it must
 * still hold H tspsi_in from that call) and the handle's FP32 copy in mixed mode.
 */
int salmon_hip_taylor_reuse(salmon_hip_t h, const double *zc, int n_hamil, void *tspsi_in, void *tspsi_out,
                            void *sshtpsi, const double *dv_dev);

/*
 * Device-side gate (no counterpart in the reference; for a predictor whose acceptance is
 * decided on the device, with no host synchronisation).  While flag_dev (a device int) is
 * set here, the launches of salmon_hip_hpsi / _taylor / _taylor_reuse, salmon_hip_density
 * and salmon_hip_hartree / _update_vlocal / _update_potential read *flag_dev when they run
 * and do nothing if it is non-zero; current, dipole and the phase table are never gated.
 * flag_dev = NULL clears it.  The flag's value is read on the device, in stream order.
 */
int salmon_hip_set_skip(salmon_hip_t h, const int *flag_dev);

/* dst_dev[i] = src_dev[i] for i < n when *flag_dev != 0 (device arrays; on the handle's stream). */
int salmon_hip_copy_if(salmon_hip_t h, const int *flag_dev, double *dst_dev, const double *src_dev, size_t n);

/*
 * Streaming copy of `bytes` (a multiple of 16) from src_dev to dst_dev with dwordx4
 * non-temporal loads and stores, one warm-up launch then `reps` timed launches on the
 * handle's stream; *avg_ms = device time per launch (no counterpart in the reference: the
 * HBM ceiling the memory-bound kernels are measured against).  Synchronises.
 */
int salmon_hip_stream_copy(salmon_hip_t h, void *dst_dev, const void *src_dev, size_t bytes, int reps,
                           double *avg_ms);

/*
 * Arithmetic of salmon_hip_taylor (no counterpart in the reference, which is
 * FP64 throughout).  SALMON_HIP_PREC_FP64 (default): as above.
 * SALMON_HIP_PREC_MIXED: H psi, H^2 psi, ... are formed and kept as complex FP32
 * (25-point stencil, V_local and k+A products in FP32), while the projections
 * <p|H^n psi>, the KB term, the Taylor sum psi + sum zc(n) H^n psi and the
 * orbitals themselves stay complex FP64; needs n_hamil >= 2 and leaves sshtpsi
 * unwritten (the FP32 chain lives in handle scratch).  hpsi, density and
 * current are FP64 in both modes.
 */
#define SALMON_HIP_PREC_FP64 0
#define SALMON_HIP_PREC_MIXED 1
int salmon_hip_set_precision(salmon_hip_t h, int mode);

/*
 * Project autotuner (no counterpart in the reference): times the stencil pass of the
 * terms without the Taylor sum under each candidate launch shape — whole xy-plane
 * tiles (when nx*ny <= 1024) with 1 or 2 z-chunks per orbital, 256-point y-tiles with
 * 1, 2 or 4 z-chunks — `reps` launches each on handle-allocated scratch, and keeps the
 * fastest for later hpsi/taylor calls.  *choice = 10*plane + z_chunks, *best_ms its
 * time per launch (either may be null).  Needs V_local and vec_Ac set; synchronises.
 */
int salmon_hip_autotune(salmon_hip_t h, int reps, int *choice, double *best_ms);

/* Set the launch shape directly: whole_plane -1 (rule) / 0 / 1, z_chunks 0 (rule) or a count. */
int salmon_hip_set_launch_shape(salmon_hip_t h, int whole_plane, int z_chunks);

/* Run the KB projections on a second stream of the handle, concurrently with the stencil
   pass that reads the same orbitals (joined by events before the scatter).  Off by default.
   With it on, Taylor chains keep a separate projection and scatter per term instead of the
   pass that fuses a term's KB scatter with the next term's projection (faster at 64 atoms,
   32^3 points, 1024 orbital blocks: 177 against about 172 steps/s in FP64). */
int salmon_hip_set_overlap(salmon_hip_t h, int on);

/*
 * rho(ix,iy,iz,ispin) = sum_{ik,io} rocc*wtk*|psi|^2 into a DEVICE array
 * (calc_density, src/common/density_matrix.f90:26-236).
 */
int salmon_hip_density(salmon_hip_t h, const void *psi, double *rho_dev);

/*
 * salmon_hip_density with no device allocation anywhere in the call (no counterpart in
 * the reference; for callers that cannot allocate: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The same kernels and summation order,
 * so rho equals salmon_hip_density's bit for bit; the orbital-group partial sums go to a
 * file-scope device array of 8 Mi doubles (64 MiB) per GPU, held until the process ends,
 * instead of handle memory.  Returns SALMON_HIP_EINVAL when the partials
 * (ceil(no*nk/32) groups x nx*ny*nz x nspin doubles, none for one group) exceed it.
 * Every call on one GPU, from any handle, shares that array, whose contents are whatever
 * the last call left (every partial a call reads it writes first): calls that overlap on
 * one GPU (different streams or host threads) race.
 */
int salmon_hip_density_noalloc(salmon_hip_t h, const void *psi, double *rho_dev);

/*
 * curr(3, nspin) into a DEVICE array (calc_current, src/common/density_matrix.f90:240-463,
 * stencil_current :465-530, calc_current_nonlocal :532-591): the kinetic part uses
 * k + vec_Ac, the nonlocal part the phase table of the last update_kvector call,
 * exactly as time_evolution_step.f90:243-246 does.  Result divided by N_grid.
 */
int salmon_hip_current(salmon_hip_t h, const double *vec_Ac, const void *psi, double *curr_dev);

/*
 * salmon_hip_current, and with the same gather of psi the projection of psi under the phase
 * table of the field vec_Ac_next, which is built here ahead of time (no counterpart in the
 * reference: in time_evolution_step the current of step itt (:243-246) and the phase table
 * and first projection of step itt+1 (:126, taylor -> hpsi -> zpseudo) read the same
 * orbitals twice).  The next salmon_hip_update_kvector whose field equals vec_Ac_next bit
 * for bit takes that table instead of building it, and the next salmon_hip_taylor /
 * salmon_hip_hpsi of psi takes that projection instead of forming it; any other
 * update_kvector, projection or fused KB pass on the handle discards them.  psi must not
 * change in between.  Results equal the unfused sequence bit for bit.  Not for handles of
 * a batch (SALMON_HIP_ESTATE).
 */
int salmon_hip_current_next(salmon_hip_t h, const double *vec_Ac, const void *psi, double *curr_dev,
                            const double *vec_Ac_next);

/*
 * Electronic dipole and electron count from rho (DEVICE arrays): out[0:3] =
 * -sum(pos_d * rho) * Hgs_d * Hvol, out[3] = sum(rho) * Hvol (calc_dip and
 * subdip, src/rt/dip.f90:23-112,114-225).
 */
int salmon_hip_dipole(salmon_hip_t h, const double *rho_dev, double *out_dev);

/*
 * Reciprocal-space tables of the periodic Poisson solver for an orthogonal cell
 * of lengths al[3] (bohr): the DFT phase tables egx/egy/egz and
 * coef = 4 pi/|G|^2 with G = 0 removed (init_reciprocal_grid,
 * src/common/initialization.f90:640-784).  Required before hartree and
 * update_potential; may be called again when the cell changes.
 */
int salmon_hip_set_poisson(salmon_hip_t h, const double *al);

/*
 * Hartree potential Vh(nx,ny,nz) of the total density rho(nx,ny,nz), both DEVICE
 * arrays (hartree -> poisson_ft, src/poisson/hartree.f90, poisson_periodic.f90:23-242):
 * separable forward DFT z, y, x with 1/N, multiply by coef, inverse DFT, real part.
 */
int salmon_hip_hartree(salmon_hip_t h, const double *rho_dev, double *vh_dev);

/*
 * PZ81 LDA (exchange_correlation with xc = 'PZ', src/xc/salmon_xc.f90:58-343): for
 * nspin = 1 builtin_pz.f90:23-83 on rho(nx,ny,nz) into vxc(nx,ny,nz); for nspin = 2 the
 * spin-polarised form builtin_pz_sp.f90:23-98 on the channel densities rho_s(nx,ny,nz,2)
 * into vxc_s(nx,ny,nz,2).  If exc_dev is not null, the energy density per electron
 * (nx,ny,nz).  DEVICE arrays.
 */
int salmon_hip_xc_pz(salmon_hip_t h, const double *rho_dev, double *vxc_dev, double *exc_dev);

/*
 * V_local(ispin) = Vpsl + Vh + Vxc(ispin) into the handle's local potential (update_vlocal,
 * src/common/hamiltonian.f90:813-843).  With vmix_dev not null the result is the
 * predictor-corrector average (vmix + V)/2 per channel (src/rt/time_evolution_step.f90:402-409).
 * If vout_dev is not null the result is also written there.  All arrays on the DEVICE:
 * vpsl, vh (nx,ny,nz); vxc, vmix, vout (nx,ny,nz,nspin).
 */
int salmon_hip_update_vlocal(salmon_hip_t h, const double *vpsl_dev, const double *vh_dev,
                             const double *vxc_dev, const double *vmix_dev, double *vout_dev);

/*
 * The 'aetrs' propagator's potential extrapolation between its two half steps
 * (time_evolution_half_step_etrs, src/rt/time_evolution_step.f90:432-445) on the handle's
 * local potential: V_new = 3 V - 3 V_old1 + V_old2, V_old2 = V_old1, V_old1 = V, V = V_new.
 * vold1_dev, vold2_dev (nx,ny,nz,nspin) DEVICE arrays, updated in place (start both at the
 * initial V_local, initialization_rt.f90:290-291); V_new also to vout_dev if not null.
 */
int salmon_hip_vlocal_extrapolate(salmon_hip_t h, double *vold1_dev, double *vold2_dev, double *vout_dev);

/*
 * hartree + exchange_correlation (PZ) + update_vlocal in one call from the density
 * rho_s(nx,ny,nz,nspin) as calc_density returns it, as time_evolution_step.f90:160-207
 * does every step with yn_fix_func = 'n': Vh of the total density (rho_s(1) + rho_s(2)
 * for nspin = 2, summed in the first DFT pass), Vxc of the channel densities formed in
 * the update kernel, Vh kept in handle scratch.  vmix, vout (nx,ny,nz,nspin) as above.
 */
int salmon_hip_update_potential(salmon_hip_t h, const double *rho_dev, const double *vpsl_dev,
                                const double *vmix_dev, double *vout_dev);

/*
 * Batched RT step over many independent systems (no counterpart in the reference,
 * which runs one system per process).  Each system keeps its own handle (grid,
 * orbitals, projectors, V_local); the batch runs the plain fixed-potential step
 * (update_kvector, taylor, calc_density, calc_current, calc_dip of
 * src/rt/time_evolution_step.f90) for all of them together: the stencil passes of
 * all systems of one halo class run as ONE persistent kernel over a combined tile
 * list (the terms without the Taylor sum on the single-system path's delayed-z tiles
 * when every member's rows fit them, z chunks rounded up to multiples of 4 planes),
 * every other phase (phase table, projections, fused KB passes, scatter,
 * density, current, dipole) runs as ONE launch for all systems (per-system argument
 * records and block offsets on the device; when a system with projectors cannot run
 * the fused KB pass, those phases fall back to per-system launches forked onto four
 * side streams), A(t) is read from per-system DEVICE tables (3, 0:nt+1) at a device step
 * counter, and the per-step current curr(3*nspin) and dipole/electron count (4) are
 * written into per-system DEVICE series at row itt-1, so that a captured step can be
 * replayed.  FP64 only.  All handles are moved onto `stream` (NULL: the first
 * handle's own stream).  Calls on a batch and
 * its handles must not overlap.
 *   batch_set_system: buffers of system i (psi_a holds the initial orbitals; the two
 *     orbital buffers alternate by step), Ac_tot table, series of nt rows.
 *   batch_prepare: Taylor coefficients; builds the tile lists, descriptor tables and
 *     the packed launches' records, sets the step counter to 1.
 *   batch_autotune: times the packed stencil for each candidate (z-chunks per orbital:
 *     the systems' rule, 1 or 2; persistent blocks: 256, 512, 1024, 2048 or one per
 *     tile) and keeps the fastest; *choice = 100*z + blocks/256 (99 = one per tile).
 *   batch_step: enqueue one step; SALMON_HIP_EINVAL once the step counter has passed the
 *     shortest system's nt (the field tables end there).
 *   batch_set_time: set the step counter (1 .. nt_min + 1).
 *   batch_capture / batch_replay: graph capture of two consecutive steps and its replay;
 *     switched off (they return SALMON_HIP_EUNSUPPORTED): a stream capture of the step,
 *     with its per-system launches forked onto side streams, once ended in a host
 *     segmentation fault whose cause was not found.
 */
typedef struct salmon_hip_batch *salmon_hip_batch_t;
int salmon_hip_batch_create(const salmon_hip_t *handles, int n, void *stream, salmon_hip_batch_t *out);
int salmon_hip_batch_destroy(salmon_hip_batch_t b);
int salmon_hip_batch_set_system(salmon_hip_batch_t b, int i, void *psi_a, void *psi_b, void *work, double *rho_dev,
                                const double *ac_dev, int nt, double *curr_t_dev, double *dip_t_dev);
int salmon_hip_batch_prepare(salmon_hip_batch_t b, const double *zc, int n_hamil);
int salmon_hip_batch_autotune(salmon_hip_batch_t b, int reps, int *choice, double *best_ms);
int salmon_hip_batch_set_time(salmon_hip_batch_t b, int itt);
int salmon_hip_batch_step(salmon_hip_batch_t b);
int salmon_hip_batch_capture(salmon_hip_batch_t b);
int salmon_hip_batch_replay(salmon_hip_batch_t b, int npairs);

/*
 * Kernel timing (no counterpart in the reference, whose timers are host
 * wall-clock regions, src/misc/timer.f90:297-310): between prof_start and
 * prof_end every launch of the listed kernels is bracketed by start/stop
 * events recorded by the launch itself (hipExtLaunchKernelGGL), up to
 * max_launches launches.  prof_stop drains the stream and returns the summed
 * device time and launch count of one kernel kind.
 */
#define SALMON_HIP_KERNEL_HPSI0 0          /* k_hpsi, plain hpsi */
#define SALMON_HIP_KERNEL_HPSI_TAYLOR1 1   /* k_hpsi, first Taylor term */
#define SALMON_HIP_KERNEL_HPSI_TAYLOR 2    /* k_hpsi, middle Taylor terms */
#define SALMON_HIP_KERNEL_HPSI_TAYLOR_LAST 3
#define SALMON_HIP_KERNEL_PROJ 4           /* k_proj, KB projection */
#define SALMON_HIP_KERNEL_PROJ_CURRENT 5   /* k_proj, nonlocal-current moments */
#define SALMON_HIP_KERNEL_NL_SCATTER 6     /* k_nl_scatter, KB term at sphere points */
#define SALMON_HIP_KERNEL_KB_FUSED 7       /* KB term of a Taylor term fused with its projection */
int salmon_hip_prof_start(salmon_hip_t h, int max_launches);
int salmon_hip_prof_stop(salmon_hip_t h, int kind, double *total_ms, int *count);
int salmon_hip_prof_end(salmon_hip_t h);

/* Block the host until the handle's stream has drained. */
int salmon_hip_synchronize(salmon_hip_t h);

/* Device memory helpers for hosts without their own allocator (e.g. Fortran). */
int salmon_hip_malloc(void **dptr, size_t bytes);
int salmon_hip_free(void *dptr);
int salmon_hip_memcpy_htod(salmon_hip_t h, void *dst_dev, const void *src_host, size_t bytes);
int salmon_hip_memcpy_dtoh(salmon_hip_t h, void *dst_host, const void *src_dev, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* SALMON_HIP_H */
