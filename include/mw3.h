/*
 * mw3.h -- C ABI of libmw3, the MI355X (gfx950) Chebyshev H.psi propagator.
 *
 * libmw3 replaces the hot path of MADWAVE3:
 *   - the H.psi operator of module mod_Hphi_01y2 (SRC/mod_Hphi_01y2.f:1-943),
 *   - the damped real Chebyshev recursion driven by program madwave3
 *     (SRC/main_madwave3.f:231-368),
 *   - the Chebyshev-to-energy projections of mod_flux_01y2 and mod_PsiE_01y2
 *     (SRC/mod_flux_01y2.f:585-816, SRC/mod_PsiE_01y2.f:92-123).
 *
 * Conventions
 *   - Plain C types only.  Every array argument is a host pointer; the library
 *     owns its device copies (one entry point, mw3_cheby_spectrum_noalloc,
 *     takes device buffers instead).  All floating-point data are double
 *     precision.
 *   - Multi-dimensional arrays are passed in the memory order of the reference's
 *     Fortran arrays: a Fortran array A(n1,n2,n3) is the C array A[n3][n2][n1].
 *     A Fortran caller therefore passes its module arrays unchanged.
 *   - Integer index arrays hold the reference's 1-based values.
 *   - Every entry point returning int returns 0 on success and a negative
 *     MW3_E* code on failure; mw3_last_error() gives the message.  The reference
 *     writes a message and stops (e.g. SRC/mod_Hphi_01y2.f:70-73); a Fortran
 *     caller does the same with the message.
 *   - One host thread per context; calls on one context are not re-entrant.
 *   - The compressed state vector ("A0 layout") is the reference's rpaqproc at
 *     nproc = 1 (SRC/mod_pot_01y2.f:435-455): for each local angle slot
 *     iang_proc = 1..nangu (global angle iang = indangreal(iang_proc); the
 *     reference assigns slots in decreasing iang, SRC/mod_baseYfunciones_01y2.f
 *     :371-375), for each channel ican = 1..ncan, the npunreal(iang) kept grid
 *     points of that angle in file order (R outer, r inner;
 *     SRC/mod_pot_01y2.f:188-190).  ntot = ncan * sum(npunreal).
 */
#ifndef MW3_H
#define MW3_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define MW3_API __attribute__((visibility("default")))
#else
#define MW3_API
#endif

#define MW3_ABI_VERSION 1

/* status codes */
#define MW3_OK 0
#define MW3_EINVAL (-1)    /* invalid or inconsistent argument             */
#define MW3_ENOMEM (-2)    /* device or host allocation failed             */
#define MW3_EDEVICE (-3)   /* HIP runtime error (no device, launch failure) */
#define MW3_ESTATE (-4)    /* call out of order (e.g. steps before begin)  */
#define MW3_EUNSUP (-5)    /* input the library does not implement         */

/* arithmetic of a propagator context */
#define MW3_FP64 0  /* every operator term and the recursion in FP64            */
#define MW3_MIXED 1 /* radial DST, angular contraction and V in FP32; Chebyshev
                       state, update, spectral scaling and accumulators in FP64 */

/*
 * mw3_setup: the reference's module state that set_vectors_Hphi, difs, the
 * recursion and ini_flux read (SRC/mod_Hphi_01y2.f:38-163, 623-943;
 * SRC/mod_flux_01y2.f:54-334).  Units are the reference's: lengths in
 * angstrom, masses in amu, energies in zots (SRC/mod_gridYpara_01y2.f:13-31).
 * The pointers are read during mw3_create only.
 */
typedef struct mw3_setup {
  int32_t abi_version; /* MW3_ABI_VERSION */

  /* grid and basis sizes (mod_gridYpara_01y2, SRC/mod_gridYpara_01y2.f:92-101) */
  int32_t npun1, npun1min, npun2, nangu, inc;
  int32_t Jtot, iparity, iommin, iommax, j0;
  int32_t nelecmax;
  double rmis1, ah1; /* r = R1 grid: r(ir1) = rmis1 + (ir1-1)*ah1 */
  double rmis2, ah2; /* R = R2 grid */
  double hbr;        /* hbar in zot/amu/angstrom units */

  /* masses, cuts, spectral range (mod_pot_01y2, SRC/mod_pot_01y2.f:17-29) */
  double xm1red, xm2red;
  double radcutmax, rotcutmax;
  double emindlt, delta2; /* Chebyshev scaling, SRC/mod_pot_01y2.f:502-505 */
  int32_t proj_j2_up;     /* must be 0 (SRC/mod_Hphi_01y2.f:274-298 not built) */
  const int32_t *iomdiat, *iomatom; /* [nelecmax] */
  const double *sigdiat, *sigatom;  /* [nelecmax] */

  /* compressed grid (SRC/mod_pot_01y2.f:40-44, 435-455) */
  int32_t npuntot;            /* max over angles of npunreal */
  const int32_t *indangreal;  /* [nangu] global angle of local slot, 1-based */
  const int32_t *npunreal;    /* [nangu] */
  const int32_t *indr1;       /* indr1(npuntot,nangu): [nangu][npuntot], 1-based */
  const int32_t *indr2;       /* indr2(npuntot,nangu) */
  const double *VVV;          /* VVV(npuntot,nangu,nelecmax,nelecmax), second
                                 index the local slot iang_proc:
                                 [nelecmax][nelecmax][nangu][npuntot] */

  /* channels and angular functions (mod_baseYfunciones_01y2,
     SRC/mod_baseYfunciones_01y2.f:50-239, 454-551) */
  int32_t ncan;
  const int32_t *iombas;  /* [ncan] Omega of channel          */
  const int32_t *nelebas; /* [ncan] electronic state, 1-based  */
  const int32_t *nojbas;  /* [ncan] number of j in the channel */
  const int32_t *j00;     /* [ncan] first j of the channel     */
  const int32_t *jbas;    /* jbas(nangu,ncan): [ncan][nangu]   */
  int32_t ndj;            /* j extent of Djmmp: nangu2+1       */
  const double *Djmmp;    /* Djmmp(nangu,0:nangu2,nelecmax,0:Jtot):
                             [Jtot+1][nelecmax][ndj][nangu]; entries the
                             reference leaves unset must be 0 */

  /* absorption (SRC/mod_absorcion_01y2.f:72-100) */
  const double *absfr1; /* [npun1] */
  const double *absfr2; /* [npun2] */
  int32_t ir2abs;       /* first absorbing R index, 1-based */

  /* energy projection (mod_flux_01y2, SRC/mod_flux_01y2.f:54-334) */
  int32_t netot;
  const double *etots2;  /* [netot] energies, zots                  */
  const double *reacfct; /* [netot] flux normalisation (Smatini)     */
  double r1flux;         /* r of the flux line, angstrom (r1flux_ang) */
  int32_t nvini, nvmax, jini, jmax;
  const int32_t *noBCstates; /* noBCstates(jini:jmax,nelecmax) */
  const double *fd; /* fd(npun1,nvini:nvmax,jini:jmax,nelecmax) */
  int32_t iphoto, iprod, kminelastic;
  int32_t icanref; /* reference channel, 1-based; 0 if none */
  int32_t nvref, jref, iomref;

  /* psi(E) at resonance energies (mod_PsiE_01y2, SRC/mod_PsiE_01y2.f:19-88) */
  int32_t nofe;
  const double *efe; /* [nofe] zots */
} mw3_setup;

typedef struct mw3_ctx mw3_ctx;

/* Library ABI version (MW3_ABI_VERSION of the build). */
MW3_API int32_t mw3_abi_version(void);

/* Message of the last failure of a context; of the last failed mw3_create when
   ctx is NULL. */
MW3_API const char *mw3_last_error(const mw3_ctx *ctx);

/* <- set_vectors_Hphi (SRC/mod_Hphi_01y2.f:38-163: angkinj2 166-311,
   bfgridl2mat 315-484, Tradial 487-620) and the flux-line set-up of ini_flux
   (SRC/mod_flux_01y2.f:99-257).  Builds the operator tables on the current HIP
   device.  precision is MW3_FP64 or MW3_MIXED. */
MW3_API int32_t mw3_create(const mw3_setup *setup, int32_t precision, mw3_ctx **out);

/* Frees every device and host resource of ctx.  NULL is accepted. */
MW3_API void mw3_destroy(mw3_ctx *ctx);

/* Length of the compressed state vector (ntotproc at nproc = 1). */
MW3_API int64_t mw3_ntot(const mw3_ctx *ctx);

/* Number of flux points (nflux, SRC/mod_flux_01y2.f:150-178). */
MW3_API int64_t mw3_nflux(const mw3_ctx *ctx);

/* <- difs (SRC/mod_Hphi_01y2.f:623-943): hpsi = H psi in zots, both [ntot]
   host arrays in A0 layout. */
MW3_API int32_t mw3_hphi(mw3_ctx *ctx, const double *psi, double *hpsi);

/* <- SRC/main_madwave3.f:231-290: psi_0 = psi0, psi_1 = (H psi_0 -
   emindlt psi_0)/delta2 (no absorption), k = 1; rpaq0 = psi0 for the
   autocorrelation.  stats3 (may be NULL) receives <psi0|H|psi0>, |psi0|^2 and
   |H psi0|^2 (SRC/main_madwave3.f:265-275).  Clears every accumulator. */
MW3_API int32_t mw3_cheby_begin(mw3_ctx *ctx, const double *psi0, double *stats3);

/* <- SRC/main_madwave3.f:324-353: nsteps iterations k -> k+1 of the damped
   recursion, each followed by the projections CoefEcheby, totalflux_k,
   Cvjflux_k, check and funE at the new k.  norms[nsteps] (may be NULL) receives
   |psi_k|^2 (check, SRC/main_madwave3.f:430-442); autocorr[nsteps] (may be
   NULL) receives <psi_0|psi_k> (SRC/main_madwave3.f:444-455). */
MW3_API int32_t mw3_cheby_steps(mw3_ctx *ctx, int32_t nsteps, double *norms,
                        double *autocorr);

/* Advances every context of ctxs[0..n) by nsteps Chebyshev steps, as
   mw3_cheby_steps does for one, enqueuing step s of every context before step
   s+1 of any so that the kernels of different contexts share the GPU.  Each
   context keeps its own state and accumulators.  norms (may be NULL) receives
   [n][nsteps] |psi_k|^2.  No counterpart in the reference, which propagates
   one total J per run (SRC/main_madwave3.f): it is for batches of J / Omega
   blocks of uneven size on one GPU.  All contexts must be on one device. */
MW3_API int32_t mw3_cheby_steps_batch(mw3_ctx **ctxs, int32_t n, int32_t nsteps, double *norms);

/* Current Chebyshev index k (1 after mw3_cheby_begin). */
MW3_API int32_t mw3_cheby_k(const mw3_ctx *ctx);

/* <- totalflux (SRC/mod_flux_01y2.f:778-816): S2prodtot[netot] and
   S2prodelectot(netot,nelecmax) as [nelecmax][netot]; either may be NULL. */
MW3_API int32_t mw3_flux_result(mw3_ctx *ctx, double *S2prodtot, double *S2prodelec);

/* The E > V(flux point) test of totalflux_k (SRC/mod_flux_01y2.f:640):
   on = 1 (default) keeps a flux point's terms only at the energies above its
   potential, as the reference's current source does; on = 0 keeps them at
   every energy, as the earlier version of the reference that made the
   HD+H example's published probability table did.  The test is applied when
   the sums are read (mw3_flux_result, mw3_flux_accum_get), so it may be set
   at any time and applies to every step. */
MW3_API int32_t mw3_flux_energy_test(mw3_ctx *ctx, int32_t on);

/* The per-step records the reference writes to Cvj.NNNN and Cvjprod.NNNN
   ("write(15) it,Cvj" and "write(16) it,Cvjprod", SRC/main_madwave3.f:353-365):
   with on = 1, every ring flush (at the end of every mw3_cheby_steps call and
   every 256 steps inside one) appends one record per step it holds to a list
   in host memory: the Chebyshev index k, Cvj(nvini:nvmax, jini:jmax, ncan)
   of Cvjflux_k (SRC/mod_flux_01y2.f:663-733) as [ncan][nj][nv], and, with a
   product projection attached, Cvjprod as mw3_cvjprod_get lays it out.
   on = 0 stops recording and drops the held records; so do
   mw3_cheby_begin and mw3_prod_attach. */
MW3_API int32_t mw3_records(mw3_ctx *ctx, int32_t on);

/* Moves the oldest min(maxrec, held) records out, oldest first: it[i],
   cvj[i][ncan*nj*nv], cvjprod[i][nomb*nj'*nv'] (it, cvj or cvjprod may be
   NULL: not copied; cvjprod is left untouched without a product
   projection); *nrec = the number moved. */
MW3_API int32_t mw3_records_get(mw3_ctx *ctx, int32_t maxrec, int32_t *it, double *cvj,
                                double *cvjprod, int32_t *nrec);

/* zS2(netot,nvini:nvmax,jini:jmax,ncan) of Cvjflux_k
   (SRC/mod_flux_01y2.f:663-775), complex as interleaved (re, im):
   [ncan][jmax-jini+1][nvmax-nvini+1][netot][2]. */
MW3_API int32_t mw3_zs2_get(mw3_ctx *ctx, double *zS2);

/* zCR, zCdR(netot,nflux) of totalflux_k (SRC/mod_flux_01y2.f:585-660), complex
   interleaved: [nflux][netot][2] each.  The E > V(flux point) test is
   applied unless mw3_flux_energy_test turned it off.  Either may be NULL.  _put restores them (restart,
   SRC/mod_flux_01y2.f:298-322). */
MW3_API int32_t mw3_flux_accum_get(mw3_ctx *ctx, double *zCR, double *zCdR);
MW3_API int32_t mw3_flux_accum_put(mw3_ctx *ctx, const double *zCR, const double *zCdR,
                           const double *zS2);

/* zfe(nbastotproc,nofe) of funE (SRC/mod_PsiE_01y2.f:92-123), complex
   interleaved [nofe][ntot][2]. */
MW3_API int32_t mw3_psie_get(mw3_ctx *ctx, double *zfe);

/* <- wvpcont / restart (SRC/main_madwave3.f:223-261, 494-552): psi_{k-1}
   (rflanz0proc) and psi_k (rpaqproc), [ntot] each, and k. */
MW3_API int32_t mw3_state_get(mw3_ctx *ctx, double *psi_km1, double *psi_k,
                      int32_t *k);
MW3_API int32_t mw3_state_put(mw3_ctx *ctx, const double *psi_km1, const double *psi_k,
                      int32_t k, const double *psi0);

/*
 * Product-state projection of iprod = 2 (module mod_coortrans01_02): the
 * tables that ini_transcoor / changeini / bfreac2prod build
 * (SRC/mod_coortrans01_02.f:31-690) and that prodpaq / prodcvj read every
 * step (SRC/mod_coortrans01_02.f:694-825).  Ranges are the reference's
 * 1-based inclusive bounds; the pointers are read during mw3_prod_attach only.
 */
typedef struct mw3_prod_setup {
  int32_t n2prod0, n2prod1;      /* R points of the projection              */
  int32_t nangproj0, nangproj1;  /* global angles of the projection         */
  int32_t iomminprod, iommaxprod; /* product helicities Omega'              */
  int32_t nviniprod, nvmaxprod;  /* product vibrational range               */
  int32_t jiniprod, jmaxprod;    /* product rotational range                */
  const int32_t *noBCprod;       /* noBCprod(jiniprod:jmaxprod)             */
  const double *transRgjac;      /* [ntot] in A0 layout                     */
  const double *prodwf;          /* prodwf(n2prod0:n2prod1, nangproj0:nangproj1,
                                    nelecmax, nviniprod:nvmaxprod,
                                    jiniprod:jmaxprod, iomminprod:iommaxprod) */
  const double *dtrans;          /* dtrans(n2prod0:n2prod1, nangproj0:nangproj1,
                                    iomminprod:iommaxprod, ncan)            */
} mw3_prod_setup;

/* <- ini_transcoor (SRC/mod_coortrans01_02.f:31-209): hands the product
   tables to a propagation context created with iprod = 2 (npun1, npun2 > 1).
   From the next Chebyshev step on, every step also runs prodpaq and prodcvj
   (SRC/mod_coortrans01_02.f:694-825, called at SRC/main_madwave3.f:346-349).
   Clears zS2prod; mw3_cheby_begin clears it too.  States the reference
   leaves out (v' > noBCprod(j'), j' < Omega') get no amplitude, whatever the
   table holds for them. */
MW3_API int32_t mw3_prod_attach(mw3_ctx *ctx, const mw3_prod_setup *prod);

/* zS2prod(netot, nviniprod:nvmaxprod, jiniprod:jmaxprod, iomminprod:iommaxprod)
   of prodcvj (SRC/mod_coortrans01_02.f:791-822), complex interleaved:
   [nomb][nj'][nv'][netot][2].  _put restores it (restart,
   SRC/mod_coortrans01_02.f:186-193). */
MW3_API int32_t mw3_zs2prod_get(mw3_ctx *ctx, double *zS2prod);
MW3_API int32_t mw3_zs2prod_put(mw3_ctx *ctx, const double *zS2prod);

/* Cvjprod(nviniprod:nvmaxprod, jiniprod:jmaxprod, iomminprod:iommaxprod) of
   the last step (prodpaq, SRC/mod_coortrans01_02.f:694-788), the record the
   reference writes to Cvjprod.NNNN each step: [nomb][nj'][nv']. */
MW3_API int32_t mw3_cvjprod_get(mw3_ctx *ctx, double *cvjprod);

/* <- cheby-spectra (SRC/photo/cheby-spectra.f:58-97): the spectrum from the
   autocorrelation auto[0..kmax] (auto[0] = 1, auto[k] = <psi_0|psi_k>):
     spectrum[i] = [auto[0] hbr + sum_{k=1..kmax} 2 hbr cos(k acos Es)
                    e^{-k xkgam} auto[k]] / (sqrt(1 - Es^2) delta2),
     Es = (e[i] - emindlt)/delta2, e in zots;
   a NaN (|Es| >= 1) gives 0, as in the reference.  Stateless; runs on the
   current HIP device.  Errors: mw3_last_error(NULL). */
MW3_API int32_t mw3_cheby_spectrum(const double *autoc, int32_t kmax, double emindlt,
                                   double delta2, double hbr, double xkgam, const double *e,
                                   int32_t ne, double *spectrum);

/* Execution options of a context (no counterpart in the reference; results
   do not depend on them beyond rounding order where stated).
   MW3_OPT_GRAPHS  1 (default): Chebyshev steps run as captured HIP graphs,
                   one per step, whenever kernel timing is off and nofe = 0;
                   0: kernel by kernel.  Same kernels, same order of every
                   sum: results are bitwise identical.
   MW3_OPT_RADIAL  1 (default): the radial kinetic correction by the
                   low-mode kernels (R passes with the sine table in LDS, one
                   wave per 16-row work item; the two r passes in one kernel)
                   where the grid allows them; 0: as batched GEMMs.
   MW3_OPT_ANGULAR 0 (default): the angular term by the parity-folded
                   one-pass kernel (with Coriolis-coupled channels, in
                   MW3_MIXED with FP16 products: its forward and back parts
                   around the coefficient pass); 1: the unfolded one-pass
                   kernel; 2: two GEMMs with the FBR coefficients between
                   them.  A variant
                   not built for the context falls back to the next one.
   MW3_OPT_SIDE_STREAM  1: the angular term on its own stream beside the
                   radial passes (default in MW3_MIXED); 0: one stream
                   (default in MW3_FP64).
   MW3_OPT_ANGULAR_GRID 1 (default) or 2: the folded angular kernel runs
                   its blocks per channel divided by this value.
   MW3_OPT_RADIAL_GRID  1 (default), 2 or 4: the staged persistent R-pass
                   kernels run the resident block count divided by this
                   value (each block loads its waves' sine-table slices
                   once).  The R passes with the sine table in LDS (FP32,
                   and FP64 on devices of 16 or more compute units) run a
                   grid of their own and ignore it; mw3_get_option then
                   reports 1.
   MW3_OPT_FP16_PRODUCTS  MW3_MIXED only.  1 (default): the inverse R pass
                   and the folded angular kernel multiply on the FP16 matrix
                   cores, each FP32 operand split into FP16 hi and lo parts
                   (22 significant bits, data scaled per row or column by a
                   power of two) and hi hi + hi lo + lo hi summed in FP32;
                   0: those products in FP32.  A precision choice: the
                   autotuner keeps the value set.  Reported 0 in MW3_FP64.
   The radial and angular variants sum in different orders: results agree
   to rounding.  mw3_get_option returns the variant in effect (after
   fallbacks).  mw3_set_option finishes the steps in flight first. */
#define MW3_OPT_GRAPHS 0
#define MW3_OPT_RADIAL 1
#define MW3_OPT_ANGULAR 2
#define MW3_OPT_SIDE_STREAM 3
#define MW3_OPT_RADIAL_GRID 4
#define MW3_OPT_ANGULAR_GRID 5
#define MW3_OPT_FP16_PRODUCTS 6
#define MW3_NOPTIONS 7
MW3_API int32_t mw3_set_option(mw3_ctx *ctx, int32_t option, int32_t value);
MW3_API int32_t mw3_get_option(const mw3_ctx *ctx, int32_t option, int32_t *value);

/* The autotuner (no counterpart in the reference): times trial_steps
   (2..254, rounded down to even) Chebyshev steps of the context under every
   combination of the options above that the context can run, from its
   current state, and keeps the fastest.  The state, k and every accumulator
   are as before the call (the trial steps are discarded: no ring flush, no
   psi(E) sum); steps already buffered are accumulated first.  chosen (may be
   NULL) receives the MW3_NOPTIONS option values in effect, ms_per_step (may
   be NULL) the best trial time per step. */
MW3_API int32_t mw3_autotune(mw3_ctx *ctx, int32_t trial_steps, int32_t *chosen,
                             double *ms_per_step);

/* No counterpart in the reference.  For callers that cannot allocate device
   memory -- code under HIP graph capture, serving loops, harnesses that
   forbid hipMalloc: the sum of mw3_cheby_spectrum on device buffers the
   caller owns, d_autoc[kmax+1], d_e[ne] and d_spectrum[ne], enqueued on
   `stream` (a hipStream_t; NULL is the null stream) of the current device.
   It returns once the work is enqueued; the caller synchronizes.  It calls
   no allocating routine: the coefficients go to a device array of the
   library (2^17 doubles, 1 MiB per device, held until the process ends), so
   kmax is at most 131071 (MW3_EINVAL beyond), and calls that overlap on one
   device race on that array.  Errors: mw3_last_error(NULL). */
MW3_API int32_t mw3_cheby_spectrum_noalloc(const double *d_autoc, int32_t kmax, double emindlt,
                                           double delta2, double hbr, double xkgam,
                                           const double *d_e, int32_t ne, double *d_spectrum,
                                           void *stream);

/* Per-kernel timing with HIP events on the stream each kernel runs on.  on = 0
   disables it; on = 1 brackets every launch of a timed kernel with an event
   pair; on = N > 1 does so on every N-th Chebyshev step only (sampling keeps
   the event overhead out of the step time).  mw3_timing_get adds the elapsed
   times and counts the timed launches.  ids: 0 .. MW3_NKERNELS-1.  No
   counterpart in the reference (its timing is wall-clock per loop,
   SRC/main_madwave3.f:370-372). */
#define MW3_NKERNELS 11
MW3_API int32_t mw3_timing_enable(mw3_ctx *ctx, int32_t on);
MW3_API int32_t mw3_timing_get(mw3_ctx *ctx, int32_t id, double *total_ms,
                       int64_t *launches);
MW3_API const char *mw3_kernel_name(int32_t id);

#ifdef __cplusplus
}
#endif

#endif /* MW3_H */
