/*
 * bastr_prof.h -- per-kernel timing of a bastr handle (no counterpart in the
 * reference, which has no timers beyond an unused MPI_Wtime,
 * parallel.F90:2658-2664).  For benchmarks: when a kernel class is enabled,
 * every launch of it is bracketed by HIP events on the stream it runs on (a
 * 2-D handle created with BASTR_OVERLAP=1 in the environment runs the spectra,
 * the dissipation moment and the diagnostics reduction on a second stream) and
 * its elapsed time is accumulated at the end of each bastr_step call.
 */
#ifndef BASTR_PROF_H
#define BASTR_PROF_H

#include <stdint.h>
#include "bastr.h"

#ifdef __cplusplus
extern "C" {
#endif

enum bastr_kclass {
    BASTR_K_SPECTRA = 0,     /* shell spectra partial sums                     */
    BASTR_K_COL_INV_VG,      /* y-inverse of stage A + velgrad fields           */
    BASTR_K_VELGRAD,         /* x-inverse + velocity-gradient moments           */
    BASTR_K_FINALIZE,        /* diagnostics reduction                           */
    BASTR_K_COL_INV,         /* y-inverse of stages B, C                        */
    BASTR_K_ROW_STAGE,       /* x-inverse, nonlinear term, x-forward            */
    BASTR_K_COL_FWD_A,       /* y-forward + RK stage A update                   */
    BASTR_K_COL_FWD_B,
    BASTR_K_COL_FWD_C,
    BASTR_K_FORCE,           /* band forcing                                    */
    BASTR_K_VG_DIS,          /* not launched: VELGRAD's pass forms every
                                velocity-gradient moment, the dissipation one
                                included (the number is kept)                    */
    BASTR_K_CLOSURE,         /* learned subgrid force of a stage (bastr_closure.h) */
    BASTR_K_NCLASS
};

/* Enable event timing for the classes in mask (bit c = class c); 0 disables.
 * Resets the accumulated counters. */
int bastr_prof_enable(bastr_handle* h, uint32_t mask);
/* Launch count, summed kernel time (ms) and compulsory HBM bytes per launch
 * (every array element the kernel must read or write, counted once). */
int bastr_prof_read(const bastr_handle* h, int kclass, int64_t* launches, double* total_ms,
                    double* bytes_per_launch);
const char* bastr_prof_name(int kclass);
/* 1 if the handle's step calls replay a captured graph (unprofiled calls of
 * up to 256 iterations), 0 if they launch every kernel directly: the default
 * replays grids of at most 2^21 points over all members; BASTR_GRAPH=0 or 1
 * in the environment at bastr_create forces either. */
int bastr_prof_graph(const bastr_handle* h);
/* Device-to-device copy bandwidth on the handle's stream: `reps` copies of
 * `bytes` between two scratch buffers of that size (allocated and freed by
 * the call), timed with HIP events; *gbps = 2 * bytes * reps / time / 1e9
 * (each byte read once and written once). */
int bastr_prof_copy_gbps(bastr_handle* h, uint64_t bytes, int reps, double* gbps);

/* The closure network's kernels (bastr_closure.h) timed with HIP events on its
 * stream: `reps` launches of the training kernel on `batch` random 32-point
 * tiles of the pairs (forward and backward, no parameter update) and `reps`
 * launches of the inference kernel over pair 0; mean milliseconds per launch.
 * Algorithmic work: 13056 flops per point forward (2 (36*64 + 64*64 + 64*2)),
 * three times that per point trained. */
struct bastr_closure;
struct bastr_pairs;
int bastr_prof_closure(struct bastr_closure* net, const struct bastr_pairs* ps, int reps, int batch,
                       double* train_ms, double* forward_ms);

#ifdef __cplusplus
}
#endif
#endif
