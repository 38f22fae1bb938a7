/*
 * bastr.h -- C-ABI of the MI355X-native Bastr pseudo-spectral time step.
 *
 * This library takes the place of the reference's FFT boundary module
 * (src/fftwlink.F90) together with the solver core it feeds
 * (src/solution.F90, src/tool.F90).  The reference keeps every field on the
 * host and calls FFTW-MPI 49-91 times per step; here the whole periodic
 * domain stays resident in one GPU's HBM, and one call advances the field by
 * whole iterations of the reference's main loop, diagnostics included.
 *
 * Conventions (all entry points):
 *   - return 0 on success, a negative BASTR_E* code on failure; the message of
 *     the last failure on the calling thread is bastr_last_error().  This is
 *     the fail-fast contract of the reference's `stop '...'`
 *     (fftwlink.F90:79,300,319,419): a Fortran caller turns a nonzero return
 *     into `error stop bastr_last_error()`.
 *   - real fields are Fortran column-major, x (index i) fastest:
 *     u(i,j) at u[i + ia*j] (2-D), u(i,j,k) at u[i + ia*(j + ja*k)] (3-D),
 *     exactly the layout of u1/u2/u3 in src/commvar.F90:27.
 *   - calls are synchronous on return; one HIP stream per handle; a handle
 *     is not thread-safe.  Device memory is owned by the library.
 *   - no torch or HIP types appear in any signature.
 */
#ifndef BASTR_H
#define BASTR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define BASTR_OK            0
#define BASTR_STOP_NAN      1   /* Es+Ed was NaN: the loop stopped like the reference's
                                   `stop '"E" is a NaN'` (solution.F90:69-81) */
#define BASTR_EINVAL      (-1)  /* bad argument / unsupported configuration        */
#define BASTR_EDEVICE     (-2)  /* HIP runtime error                                */
#define BASTR_ENOMEM      (-3)  /* device allocation failed                         */

#define BASTR_PREC_F64      0   /* everything in FP64 (parity mode)                 */
#define BASTR_PREC_F32      1   /* FFTs and pointwise work in FP32; energy/spectrum
                                   reductions, forcing normalisation in FP64         */

/*
 * Run parameters.  Each field is the reference variable of the same meaning
 * (src/commvar.F90:16-66, parsed from datin/input.dat by readinput,
 * src/readwrite.F90:19-112).
 */
typedef struct bastr_params {
    int32_t ia, ja, ka;        /* grid; ka == 0 selects 2-D (readwrite.F90:69-80)       */
    double  reynolds;          /* nu = 1/reynolds (solution.F90:1306-1310; miucal(1)=1) */
    double  deltat;            /* time step                                             */
    int32_t timemethod;        /* 1 = Kutta RK3 (solution.F90:172), 2 = 6-sweep CN (:205) */
    int32_t forcemethod;       /* 0 none, 1 rescale, 2 band, 5 projected band (:303-540) */
    int32_t forcek;            /* forcing wavenumber (band 4.5<|k|<5.5 for forcek=5)    */
    double  target_energy;     /* forcing target energy                                 */
    int32_t lproject;          /* lprojectd: project tendency on dilatational part     */
    int32_t kmax;              /* allkmax = min(int(ia/3), kmax) (readwrite.F90:75-83) */
    int32_t precision;         /* BASTR_PREC_F64 / BASTR_PREC_F32                       */
} bastr_params;

/*
 * Diagnostics of one loop iteration, in the reference's order
 * (solution.F90:54-97): spectra and velocity-gradient statistics are taken
 * from the field at `nstep` BEFORE that iteration's step; the forcing values
 * come after the step and belong to log line nstep+1 (listwrite after
 * `nstep = nstep + 1`).
 */
typedef struct bastr_diag {
    int64_t nstep;             /* nstep of the pre-step diagnostics                      */
    double  time;              /* time  of the pre-step diagnostics                      */
    /* spectra_compute (solution.F90:930-1016): log/stat?d_specall.dat               */
    double  Esspe, Edspe, k2Edspe, kMEdspe;
    /* velgrad_calculate (solution.F90:1110-1286): log/stat?d_velgrad.dat            */
    double  umumtheta2, umumijji, u2theta, dissp, eta_min, Taylength, ReTay, Kollength;
    /* forcing (solution.F90:303-540): log/factor_output.dat (methods 1,2,5)         */
    double  force_factor, force_E, force_energy, force_Fen;
    int32_t nan_flag;          /* 1 when isnan(Esspe+Edspe)                              */
    int32_t pad_;
} bastr_diag;

typedef struct bastr_handle bastr_handle;

/* Last error message of the calling thread ("" if none). */
const char* bastr_last_error(void);

/*
 * Create a solver: replaces mpisizedis_fftw + allocation + prepareplan_fftw_plan
 * + quantity_prepare (fftwlink.F90:34,370,284; solution.F90:1288).
 * Supported grids: 2-D ia, ja powers of two in [16, 4096]; 3-D ia == ja == ka,
 * a power of two in [16, 512].  timemethod 1 or 2; forcemethod 0, 1, 2 or 5
 * (methods 3 and 4 draw unseeded random forcing in the reference and are
 * rejected with BASTR_EINVAL).
 */
int bastr_create(const bastr_params* p, bastr_handle** out);
int bastr_destroy(bastr_handle* h);

/* allkmax actually used (min(int(ia/3), kmax)). */
int bastr_allkmax(const bastr_handle* h);

/*
 * Load the initial (or restart) velocity field: replaces read_initial_field /
 * read_continue_field (readwrite.F90:114-220).  u3 may be NULL in 2-D.
 * The next bastr_step first performs mainloop's entry dealias
 * (solution.F90:41-52 / 105-122), as every reference run does.
 */
int bastr_set_fields(bastr_handle* h, const double* u1, const double* u2, const double* u3);
/*
 * Seeded random-phase initial field, built on the device: the construction of
 * Bastrpp's hitgen2d/hitgen3d (create_initial_field_2d/3d, Bastrpp.F90:601-808)
 * with k0 in the role of forcek, `energy` in the role of target_energy and a
 * reproducible noise source in place of the reference's clock-seeded
 * random_number (Bastrpp.F90:626-637):
 *   noise(m) = (splitmix64(seed * 0x100000001B3 + m) >> 11) * 2^-53, m the
 *   Fortran-order linear index; theta(k) = atan2 of fft(noise)(k);
 *   u_hat_i(k) = A(|k|) k_i/|k| (sin theta + i cos theta), u_hat(0) = 0,
 *   A = sqrt(E/(2 pi k)) (2-D) or sqrt(E/(4 pi k^2)) (3-D),
 *   E(k) = k^4 exp(-2 (k/k0)^2) for k >= k0, k0^6 k^-2 e^-2 below;
 *   u = real(ifft(u_hat)) (unnormalised), rescaled to <sum_i u_i^2> = energy
 *   (energy = 1 is u_rms = 1).
 * The field is then loaded as by bastr_set_fields (entry dealias pending),
 * with nstep = 0, time = 0.  energy and k0 must be > 0.
 */
int bastr_init_random_phase(bastr_handle* h, double k0, uint64_t seed, double energy);

/*
 * Spectrum law of the handle's later bastr_init_random_phase(_member) calls:
 *   BASTR_LAW_HITGEN (the default): the piecewise law of Bastrpp's hitgen
 *     above (Bastrpp.F90:654-658);
 *   BASTR_LAW_PURE: E(k) = k^4 exp(-2 (k/k0)^2) at every k (no k^-2 branch
 *     below k0), the law the benchmark workloads state.
 * Everything else in the construction is unchanged.
 */
#define BASTR_LAW_HITGEN    0
#define BASTR_LAW_PURE      1
int bastr_set_spectrum_law(bastr_handle* h, int32_t law);

/* Physical velocity field (Fortran order) at the current nstep. */
int bastr_get_fields(bastr_handle* h, double* u1, double* u2, double* u3);

/* nstep/time of the next iteration (restart sets them from the dump file). */
int bastr_set_time(bastr_handle* h, int64_t nstep, double time);
int bastr_get_time(const bastr_handle* h, int64_t* nstep, double* time);

/*
 * Run `nsteps` iterations of the reference main loop body (solution.F90:54-97
 * for 2-D, :124-169 for 3-D): diagnostics at nstep, NaN stop, nstep+=1,
 * time+=deltat, RK3 (or CN) step, forcing.
 *   diags   : NULL or nsteps entries.
 *   spectra : NULL or nsteps * 3 * allkmax doubles; iteration s writes
 *             kn[1..allkmax], Es[1..allkmax], Ed[1..allkmax] at
 *             spectra + s*3*allkmax (+0, +allkmax, +2*allkmax).
 * Returns BASTR_STOP_NAN when an iteration's Es+Ed is NaN: that iteration's
 * diagnostics are filled, the field is left as it was before that
 * iteration's step (the state the reference dumps before stopping) and
 * nstep/time are those of that iteration.
 */
int bastr_step(bastr_handle* h, int32_t nsteps, bastr_diag* diags, double* spectra);

/*
 * bastr_step on several handles at once (no counterpart in the reference):
 * every handle's iterations are launched on its own stream before any is
 * waited for, so runs of different sizes (e.g. the ensembles of one
 * workload) overlap on the device.  diags[i] / spectra[i]: NULL or handle
 * i's arrays, laid out as bastr_step lays them out for that handle (the
 * pointer arrays themselves may be NULL).  Returns a negative code on the
 * first failure, else BASTR_STOP_NAN if any handle stopped on a NaN, else 0.
 * A handle may appear only once.
 */
int bastr_step_group(bastr_handle* const* hs, int32_t nh, int32_t nsteps, bastr_diag* const* diags,
                     double* const* spectra);

/*
 * bastr_step without device allocation.  No counterpart in the reference:
 * for callers that may not allocate device memory once set up (tight serving
 * loops, harnesses that forbid hipMalloc).  Results, layout, NaN semantics
 * and nstep/time bookkeeping are those of bastr_step on the same handle,
 * which must be a one-field handle (ensembles take bastr_step).  The step's
 * per-call scratch (NaN status, partial and reduced sums, forcing factor,
 * per-iteration diagnostic and spectrum slots) lives in a fixed 16 MiB device
 * array of the library, one per device, instead of handle buffers; nothing is
 * allocated or freed and no graph is instantiated (launches are direct).
 * Needs about 256 B per iteration of diagnostics, 24*allkmax B per iteration
 * of spectra and 16*(allkmax+1) B per spectra block (at most 512 blocks);
 * a call whose total exceeds 16 MiB (e.g. more than ~3300 iterations at
 * allkmax = 200) returns BASTR_EINVAL: split it.  Calls on one device share
 * the array and are serialised by the library; the array keeps the last
 * call's contents, which no call reads before writing.
 */
int bastr_step_noalloc(bastr_handle* h, int32_t nsteps, bastr_diag* diags, double* spectra);

/*
 * Post-processing of a field, e.g. a dump (Bastrpp's commands; the solver state,
 * nstep and time of the handle are left as they were).  Fields are Fortran-order
 * real arrays as for bastr_set_fields, used as given (no entry dealias).
 *
 * bastr_pp_spectra: spectra2d (Bastrpp.F90:299-400), 2-D only: over the full
 *   spectrum, without the 1/2 of the solver's diagnostics,
 *   shells[0..K-1] = Es(1..K), shells[K..2K-1] = Ed(1..K) (K = allkmax);
 *   totals = Esspe, Edspe, k2Edspe, kM1Edspe.
 * bastr_pp_transfer: transfer2d / transfer3d (Bastrpp.F90:174-297, 402-599):
 *   out[0..K-1] = T(1..K), out[K..2K-1] = T1, out[2K..3K-1] = T2, shell sums of
 *   Re(conj(u_hat_i) fft(-u_j d_j u_i)), Im(k_j fft(u_i u_j) conj(u_hat_i)) and
 *   Re(conj(u_hat_i) fft(u_i d_j u_j)).
 */
int bastr_pp_spectra(bastr_handle* h, const double* u1, const double* u2, const double* u3, double* shells,
                     double* totals);
int bastr_pp_transfer(bastr_handle* h, const double* u1, const double* u2, const double* u3, double* out);

/*
 * Plans of the column (y, z) passes; no counterpart in the reference, whose
 * transforms are planned by FFTW_MEASURE (fftwlink.F90:284-303).  The
 * single-term column passes of length n run 16 points per thread (radix-16
 * passes) on blocks of the default width unless another plan is set for
 * (n, precision): 8 points per thread (radix-8 passes) and/or blocks of twice
 * the width (longer row segments; where the ncols of a pass is not a whole
 * number of wide blocks it keeps the default width).  The table is
 * process-wide and applies to every handle's later steps (captured step graphs
 * are remade).
 * bastr_autotune times the plans of length n on the current device in the
 * shape of a real pass (ncols columns, a multiple of 16, of each of nplanes
 * planes, entries past n/3 along the pass pruned as the step prunes them),
 * launched as one 2-D RK3 iteration launches its column passes (a 7-job and
 * two 4-job inverses, three 2-job forwards), each launch on arrays the one
 * before it did not touch so that it streams from HBM, and keeps the fastest
 * (the default plan unless another is at least 3 % faster);
 * ms[0..3]: milliseconds per such mix for the plans (points, width) =
 * (16, 1), (8, 1), (16, 2), (8, 2), NaN
 * where a plan does not exist (8 points below n = 256; a plan needing more than
 * 1024 threads or 160 KiB of LDS per workgroup).  Setting BASTR_TUNE=1 in the
 * environment at bastr_create makes the handle's first bastr_step call run
 * bastr_autotune_step (8 iterations per plan in 2-D, 2 in 3-D) once per
 * column length and precision in the process.  bastr_tune_set fixes a plan,
 * bastr_tune_get reads the one in use.  The 8-point plans differ in rounding
 * (within the parity tolerances); the block width changes no arithmetic.  A
 * fixed table gives reproducible bits.
 */
int bastr_autotune(int32_t n, int32_t precision, int32_t ncols, int32_t nplanes, int32_t* points, int32_t* width,
                   double* ms);
/*
 * bastr_autotune_step times the same four plans on the handle's own step: for
 * each plan valid for every column length of the handle (ja, and ka in 3-D),
 * `iters` iterations (1..256) from the handle's current state, once untimed
 * and once between HIP events, in two interleaved rounds; the spectral state
 * is restored after every run, so nstep, time and the results of later calls
 * are those of an untuned handle (the fields must be set).  It keeps the
 * fastest plan for those lengths (the default unless another is at least 3 %
 * faster); ms[0..3]: milliseconds per iteration, as for bastr_autotune.  A
 * pass timed alone misses what the step around it does to it (the caches the
 * previous kernel leaves, how evenly a pass of one plane fills the CUs), so
 * this is the tuner to use where a handle exists.  Fails with BASTR_EINVAL,
 * leaving the plans as they were, if a run reaches a NaN iteration.
 */
int bastr_autotune_step(bastr_handle* h, int32_t iters, int32_t* points, int32_t* width, double* ms);
int bastr_tune_set(int32_t n, int32_t precision, int32_t points, int32_t width);
int bastr_tune_get(int32_t n, int32_t precision, int32_t* points, int32_t* width);

/* Number of modes per shell 1..allkmax (Ecount, solution.F90:976). */
int bastr_shell_counts(const bastr_handle* h, int64_t* counts);

/*
 * fftwlink-compatible whole-field transforms on caller memory
 * (fftwlink.F90:324-368): in-place c2c over (ia, ja[, ka]) complex128 values
 * (interleaved re,im; Fortran order).  Forward is e^{-ik.x} followed by the
 * 1/(ia*ja[*ka]) normalisation of fft2d/fft3d; backward is e^{+ik.x},
 * unnormalised, as ifft2d/ifft3d.  Always FP64.
 */
int bastr_fft_forward(bastr_handle* h, double* inout);
int bastr_fft_backward(bastr_handle* h, double* inout);

/*
 * Ensembles of independent 2-D runs on one grid.  No counterpart in the
 * reference, which advances one field set (src/commvar.F90:27-29); each member
 * is exactly the run a one-field handle with the same parameters would make
 * (same kernels, same per-member summation order), and all members share
 * every launch of the step.
 *
 * bastr_create_ensemble: p as for bastr_create (ka must be 0); nmembers in
 *   [1, 65535]; forced: NULL (every member forced as p->forcemethod says) or
 *   nmembers flags, 0 = this member is not forced (free decay), 1 = forced.
 * bastr_set_member_fields / bastr_init_random_phase_member: as the one-field
 *   calls, for one member, whose entry dealias (solution.F90:41-52) is applied
 *   at once; every member must be set before the first bastr_step.
 * bastr_get_member_fields / bastr_get_member_time: one member's field / clock.
 * bastr_step on an ensemble handle: diags holds nmembers * nsteps entries,
 *   member m's at diags + m*nsteps; spectra nmembers * nsteps * 3 * allkmax
 *   doubles, member m's at spectra + m*nsteps*3*allkmax.  A member whose Es+Ed
 *   is NaN stops there as a one-field run would (its later entries keep
 *   zeros with NaN forcing values) while the others go on; the call then
 *   returns BASTR_STOP_NAN.
 * The one-field entry points bastr_set_fields, bastr_get_fields,
 * bastr_init_random_phase, bastr_fft_*, bastr_pp_* reject ensemble handles of
 * more than one member.
 */
int bastr_create_ensemble(const bastr_params* p, int32_t nmembers, const int32_t* forced, bastr_handle** out);
int bastr_nmembers(const bastr_handle* h);
int bastr_set_member_fields(bastr_handle* h, int32_t member, const double* u1, const double* u2, const double* u3);
int bastr_get_member_fields(bastr_handle* h, int32_t member, double* u1, double* u2, double* u3);
int bastr_init_random_phase_member(bastr_handle* h, int32_t member, double k0, uint64_t seed, double energy);
int bastr_get_member_time(const bastr_handle* h, int32_t member, int64_t* nstep, double* time);

#ifdef __cplusplus
}
#endif

#endif /* BASTR_H */
