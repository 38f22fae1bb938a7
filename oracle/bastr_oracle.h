/*
 * bastr_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of the reference Bastr time step (src/solution.F90,
 * src/tool.F90, src/fftwlink.F90), used as the correctness oracle for the
 * HIP library and as the CPU baseline in bench.py.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path (bastr_amd) never calls it.
 *
 * The API mirrors include/bastr.h with a bo_ prefix so both libraries can be
 * loaded in one process and driven by identical calls.
 */
#ifndef BASTR_ORACLE_H
#define BASTR_ORACLE_H

#include "../include/bastr.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bo_state bo_state;

const char* bo_last_error(void);
int bo_create(const bastr_params* p, bo_state** out);
int bo_destroy(bo_state* s);
int bo_allkmax(const bo_state* s);
int bo_set_fields(bo_state* s, const double* u1, const double* u2, const double* u3);
int bo_init_random_phase(bo_state* s, double k0, uint64_t seed, double energy);
int bo_set_spectrum_law(bo_state* s, int32_t law);
/* The learned closure of include/bastr_closure.h added to every compute_ut
 * (2-D, square grids): params NPARAM floats, scales 6 (NULL params: none). */
int bo_set_closure(bo_state* s, const float* params, const float* scales);
/* 1: the closure force is evaluated at RK3 stage A only and added again at
 * stages B and C (bastr_closure_per_step); 0 (default): at every stage. */
int bo_set_closure_per_step(bo_state* s, int32_t per_step);
int bo_get_fields(bo_state* s, double* u1, double* u2, double* u3);
int bo_set_time(bo_state* s, int64_t nstep, double time);
int bo_get_time(const bo_state* s, int64_t* nstep, double* time);
int bo_step(bo_state* s, int32_t nsteps, bastr_diag* diags, double* spectra);
int bo_shell_counts(const bo_state* s, int64_t* counts);
int bo_fft_forward(bo_state* s, double* inout);
int bo_fft_backward(bo_state* s, double* inout);

/* Post-processing of a field (Bastrpp's spectra2d, transfer2d/3d); the
 * solver state is left as it was. */
int bo_pp_spectra(bo_state* s, const double* u1, const double* u2, const double* u3, double* shells,
                  double* totals);
int bo_pp_transfer(bo_state* s, const double* u1, const double* u2, const double* u3, double* out);

/* Stand-alone c2c transform of a Fortran-order (n0, n1[, n2]) complex128 array
 * with fft2d/ifft2d conventions: dir=-1 forward and normalised, dir=+1
 * backward unnormalised.  n2 == 0 means 2-D. */
int bo_fft_c2c(int n0, int n1, int n2, double* inout, int dir);

/* Threads used by the OpenMP loops (0 = library default). */
void bo_set_threads(int nthreads);
int bo_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
