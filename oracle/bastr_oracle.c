/*
 * bastr_oracle.c -- TEST INFRASTRUCTURE ONLY: the correctness oracle.
 *
 * A from-scratch FP64 CPU restatement of the reference Bastr main loop,
 * following the Fortran statement by statement so that its rounding is as
 * close to the reference as a different FFT allows.  Each function cites the
 * reference lines it restates.  It is loaded only by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
 * checker or the CPU baseline, never as part of the product path.
 *
 * Pinning (what checks each part independently of this file's own formulas)
 * lives in tests/test_oracle.py: brute-force DFT and numpy.fft for the
 * transforms, closed forms for wavenumbers / dealias counts / shell counts,
 * the exact RK3 amplification of a shear flow, projection invariants,
 * Parseval for the spectra, single-mode velocity-gradient moments and the
 * band-forcing energy identity.
 *
 * Third-party boundary: the reference's transforms are FFTW3 / fftw3-mpi
 * (not vendored; no version pinned in its Makefile:15,27; call sites
 * fftwlink.F90:65,72,294-298,333,346,356,366).  This file uses its own
 * radix-2 FFT with the same sign and normalisation conventions, pinned as
 * above; the rounding of FFTW's own plans (below 1e-14 relative) is parity
 * unpinned.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off
 * -fcx-fortran-rules; no fast-math, no FMA contraction, to keep the
 * reference's statement-level rounding).
 */
#include "bastr_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

static __thread char g_err[512];
static int g_threads = 0;

static int fail(int code, const char* msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return code;
}

const char* bo_last_error(void) { return g_err; }

void bo_set_threads(int n) { g_threads = n; }
int bo_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}
#define OMP_FOR _Pragma("omp parallel for schedule(static) num_threads(bo_get_threads())")

/* ------------------------------------------------------------------------- */
/* 1-D FFT (own implementation: radix-2 for powers of two, direct DFT else)   */
/* ------------------------------------------------------------------------- */

typedef struct {
    int n;
    int pow2;
    cplx* tw; /* tw[m] = exp(-2 pi i m / n), m < n */
} plan1d;

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static void plan1d_init(plan1d* p, int n) {
    p->n = n;
    p->pow2 = is_pow2(n);
    p->tw = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    for (int m = 0; m < n; ++m) {
        /* angle reduced exactly by integer arithmetic before the libm call */
        double a = 2.0 * M_PI * (double)m / (double)n;
        p->tw[m] = cos(a) - I * sin(a);
    }
}

static void plan1d_free(plan1d* p) { free(p->tw); p->tw = NULL; }

/* sign = -1: sum x e^{-i..}; sign = +1: sum x e^{+i..}; no normalisation */
static void fft_line(cplx* x, const plan1d* p, int sign, cplx* tmp) {
    const int n = p->n;
    if (n == 1) return;
    if (p->pow2) {
        for (int i = 1, j = 0; i < n; ++i) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) { cplx t = x[i]; x[i] = x[j]; x[j] = t; }
        }
        for (int len = 2; len <= n; len <<= 1) {
            const int half = len >> 1, step = n / len;
            for (int i = 0; i < n; i += len) {
                for (int j = 0; j < half; ++j) {
                    cplx w = p->tw[j * step];
                    if (sign > 0) w = conj(w);
                    cplx u = x[i + j], v = x[i + j + half] * w;
                    x[i + j] = u + v;
                    x[i + j + half] = u - v;
                }
            }
        }
    } else {
        for (int k = 0; k < n; ++k) {
            cplx s = 0;
            for (int m = 0; m < n; ++m) {
                cplx w = p->tw[(int)(((long long)m * k) % n)];
                if (sign > 0) w = conj(w);
                s += x[m] * w;
            }
            tmp[k] = s;
        }
        memcpy(x, tmp, sizeof(cplx) * (size_t)n);
    }
}

typedef struct {
    int nd;
    int n[3];
    plan1d p[3];
} plannd;

static void plannd_init(plannd* P, int n0, int n1, int n2) {
    P->nd = n2 > 0 ? 3 : 2;
    P->n[0] = n0; P->n[1] = n1; P->n[2] = n2 > 0 ? n2 : 1;
    for (int a = 0; a < 3; ++a) plan1d_init(&P->p[a], P->n[a]);
}

static void plannd_free(plannd* P) {
    for (int a = 0; a < 3; ++a) plan1d_free(&P->p[a]);
}

/* In-place unnormalised c2c over a Fortran-order (n0,n1,n2) array.  Lines
 * along axes 1 and 2 are gathered CB at a time (CB adjacent axis-0 indices
 * share each cache line of the gather), and every thread allocates its line
 * buffers once; each line's transform is fft_line's, so the result does not
 * depend on the blocking or the thread count. */
#define CB 4
static void fft_nd(const plannd* P, cplx* a, int sign) {
    const int n0 = P->n[0], n1 = P->n[1], n2 = P->n[2];
    const int nmax = n0 > n1 ? (n0 > n2 ? n0 : n2) : (n1 > n2 ? n1 : n2);
    _Pragma("omp parallel num_threads(bo_get_threads())")
    {
        cplx* buf = (cplx*)malloc(sizeof(cplx) * (size_t)nmax * (CB + 1));
        cplx* tmp = buf + (size_t)nmax * CB;
        /* axis 0: contiguous lines */
        _Pragma("omp for schedule(static)")
        for (long line = 0; line < (long)n1 * n2; ++line)
            fft_line(a + (size_t)line * n0, &P->p[0], sign, tmp);
        /* axis 1: blocks of CB lines at consecutive i */
        const long nb0 = (n0 + CB - 1) / CB;
        _Pragma("omp for schedule(static)")
        for (long blk = 0; blk < nb0 * n2; ++blk) {
            const long i0 = (blk % nb0) * CB, k = blk / nb0;
            const int nb = (int)(n0 - i0 < CB ? n0 - i0 : CB);
            for (int j = 0; j < n1; ++j) {
                const cplx* src = a + i0 + (size_t)n0 * (j + (size_t)n1 * k);
                for (int b = 0; b < nb; ++b) buf[(size_t)b * nmax + j] = src[b];
            }
            for (int b = 0; b < nb; ++b) fft_line(buf + (size_t)b * nmax, &P->p[1], sign, tmp);
            for (int j = 0; j < n1; ++j) {
                cplx* dst = a + i0 + (size_t)n0 * (j + (size_t)n1 * k);
                for (int b = 0; b < nb; ++b) dst[b] = buf[(size_t)b * nmax + j];
            }
        }
        if (n2 > 1) {
            /* axis 2: blocks of CB lines at consecutive (i, j) offsets */
            const long nl = (long)n0 * n1, nb2 = (nl + CB - 1) / CB;
            _Pragma("omp for schedule(static)")
            for (long blk = 0; blk < nb2; ++blk) {
                const long l0 = blk * CB;
                const int nb = (int)(nl - l0 < CB ? nl - l0 : CB);
                for (int k = 0; k < n2; ++k) {
                    const cplx* src = a + (size_t)l0 + (size_t)nl * k;
                    for (int b = 0; b < nb; ++b) buf[(size_t)b * nmax + k] = src[b];
                }
                for (int b = 0; b < nb; ++b) fft_line(buf + (size_t)b * nmax, &P->p[2], sign, tmp);
                for (int k = 0; k < n2; ++k) {
                    cplx* dst = a + (size_t)l0 + (size_t)nl * k;
                    for (int b = 0; b < nb; ++b) dst[b] = buf[(size_t)b * nmax + k];
                }
            }
        }
        free(buf);
    }
}
#undef CB

int bo_fft_c2c(int n0, int n1, int n2, double* inout, int dir) {
    if (n0 <= 0 || n1 <= 0 || n2 < 0) return fail(BASTR_EINVAL, "bo_fft_c2c: bad extents");
    plannd P;
    plannd_init(&P, n0, n1, n2);
    cplx* a = (cplx*)inout;
    const size_t n = (size_t)n0 * n1 * (n2 > 0 ? n2 : 1);
    if (dir < 0) {
        fft_nd(&P, a, -1);
        /* fftwlink.F90:334 / :347: array = array/(1.d0*ia*ja[*ka]) */
        const double den = 1.0 * n0 * n1 * (n2 > 0 ? n2 : 1);
        for (size_t m = 0; m < n; ++m) a[m] = a[m] / den;
    } else {
        fft_nd(&P, a, +1);
    }
    plannd_free(&P);
    return BASTR_OK;
}

/* ------------------------------------------------------------------------- */
/* solver state (commvar.F90:27-66, buffers of fftwlink.F90:389-467)          */
/* ------------------------------------------------------------------------- */

struct bo_state {
    bastr_params p;
    int nd, ia, ja, ka, nz;
    size_t n;
    int allkmax;
    double nu, dt;
    plannd plan;
    double* u[3];
    double* uold[3];
    double* kw[3];                   /* k1, k2, k3 */
    cplx* spe[3];                    /* u1spe.. */
    cplx* ux[3][3];                  /* ux[i][j] = u_i,x_j */
    cplx* uxixi[3];
    cplx* thetaxixi;
    cplx* tA[3]; cplx* tB[3]; cplx* tC[3];
    cplx* force[3];
    double* Es; double* Ed; double* kn; int64_t* Ecount;
    double Esspe, Edspe;
    int law;                         /* spectrum law of bo_init_random_phase */
    /* learned closure (2-D): FP32 master parameters, scales; NULL = none */
    float* cl_params;
    float cl_scales[6];
    double* cl_fld;                  /* [4][n] u1, u2, theta, omega of a stage */
    double* cl_out;                  /* [2][n] the force of the step's stage A (per-step cadence) */
    int cl_per_step;                 /* 1: the force of stage A is added at the three stages */
    int cl_stage;                    /* RK3 stage of the current compute_ut (0, 1, 2) */
    int64_t nstep;
    double time;
    int prologue_pending;
    int have_fields;
};

#define NCHUNK 64

static void* xcalloc(size_t n, size_t sz, int* ok) {
    void* p = calloc(n, sz);
    if (!p) *ok = 0;
    return p;
}

/* GenerateWave_2D / _3D (tool.F90:17-96) with j0f = k0f = 0 (single rank). */
static void generate_wave(bo_state* s) {
    const int ia = s->ia, ja = s->ja, nz = s->nz;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ja; ++j)
            for (int i = 0; i < ia; ++i) {
                const size_t m = (size_t)i + (size_t)ia * (j + (size_t)ja * k);
                const int I1 = i + 1, J1 = j + 1, K1 = k + 1;
                s->kw[0][m] = (I1 <= ia / 2 + 1) ? (double)(I1 - 1) : (double)(I1 - ia - 1);
                s->kw[1][m] = (J1 <= ja / 2 + 1) ? (double)(J1 - 1) : (double)(J1 - ja - 1);
                if (s->nd == 3)
                    s->kw[2][m] = (K1 <= s->ka / 2 + 1) ? (double)(K1 - 1) : (double)(K1 - s->ka - 1);
            }
}

static double kabs(const bo_state* s, size_t m) {
    if (s->nd == 2) return sqrt(s->kw[0][m] * s->kw[0][m] + s->kw[1][m] * s->kw[1][m]);
    return sqrt(s->kw[0][m] * s->kw[0][m] + s->kw[1][m] * s->kw[1][m] + s->kw[2][m] * s->kw[2][m]);
}

/* fft2d/fft3d: forward then 1/(ia*ja[*ka]) (fftwlink.F90:324-348) */
static void fftf(bo_state* s, cplx* a) {
    fft_nd(&s->plan, a, -1);
    const double den = s->nd == 2 ? 1.0 * s->ia * s->ja : 1.0 * s->ia * s->ja * s->ka;
    OMP_FOR
    for (long m = 0; m < (long)s->n; ++m) a[m] = a[m] / den;
}
/* ifft2d/ifft3d: backward, unnormalised (fftwlink.F90:350-368) */
static void ffti(bo_state* s, cplx* a) { fft_nd(&s->plan, a, +1); }

/* dealiasing2D/3D (tool.F90:149-217), factor 3 */
static void dealias(bo_state* s, cplx* f) {
    const int cutoff = (int)floor((double)s->ia / 3.0);
    OMP_FOR
    for (long m = 0; m < (long)s->n; ++m)
        if (kabs(s, (size_t)m) >= cutoff) f[m] = 0.0;
}

/* projection2D/3D (tool.F90:219-270) */
static void projection(bo_state* s, cplx** t) {
    OMP_FOR
    for (long mm = 0; mm < (long)s->n; ++mm) {
        const size_t m = (size_t)mm;
        const double kk = kabs(s, m);
        if (kk > 0.5) {
            cplx ud = 0;
            if (s->nd == 2) {
                ud = t[0][m] * s->kw[0][m] / kk + t[1][m] * s->kw[1][m] / kk;
                t[0][m] = ud * s->kw[0][m] / kk;
                t[1][m] = ud * s->kw[1][m] / kk;
            } else {
                ud = t[0][m] * s->kw[0][m] / kk + t[1][m] * s->kw[1][m] / kk + t[2][m] * s->kw[2][m] / kk;
                t[0][m] = ud * s->kw[0][m] / kk;
                t[1][m] = ud * s->kw[1][m] / kk;
                t[2][m] = ud * s->kw[2][m] / kk;
            }
        } else {
            for (int c = 0; c < s->nd; ++c) t[c][m] = 0.0;
        }
    }
}

/* The closure network of include/bastr_closure.h on one point: FP64 with the
 * device's BF16 roundings (weights, scaled features, layer-1 activations). */
static float bf16r(float x) {
    union { float f; uint32_t u; } v = {x};
    if ((v.u & 0x7fffffffu) > 0x7f800000u) return x;
    v.u = ((v.u + 0x7fffu + ((v.u >> 16) & 1u)) >> 16) << 16;
    return v.f;
}
static double gelu_t(double z) { return 0.5 * z * (1.0 + tanh(0.7978845608028654 * (z + 0.044715 * z * z * z))); }

static void closure_point(const bo_state* s, const float* x, double* out) {
    const float* p = s->cl_params;
    double h1[64], h2[64];
    for (int h = 0; h < 64; ++h) {
        double z = p[2304 + h];
        for (int f = 0; f < 36; ++f) z += (double)bf16r(p[h * 36 + f]) * (double)x[f];
        h1[h] = (double)bf16r((float)gelu_t(z));
    }
    for (int h = 0; h < 64; ++h) {
        double z = p[2368 + 4096 + h];
        for (int k = 0; k < 64; ++k) z += (double)bf16r(p[2368 + h * 64 + k]) * h1[k];
        h2[h] = gelu_t(z);
    }
    for (int c = 0; c < 2; ++c) {
        double o = p[6656 + c];
        for (int h = 0; h < 64; ++h) o += (double)bf16r(p[6528 + c * 64 + h]) * h2[h];
        out[c] = o * (double)s->cl_scales[4 + c];
    }
}

/* subgrid force of the stage field: u (spe, physical) and its gradients (ux);
 * with the per-step cadence, stages B and C add stage A's force again */
static void closure_force(bo_state* s, cplx** t) {
    const int ia = s->ia, ja = s->ja;
    const size_t n = s->n;
    double* const O = s->cl_out;
    if (s->cl_per_step && s->cl_stage > 0) {
        OMP_FOR
        for (long mm = 0; mm < (long)n; ++mm) {
            t[0][mm] += O[mm];
            t[1][mm] += O[n + mm];
        }
        return;
    }
    double* F = s->cl_fld;
    for (size_t m = 0; m < n; ++m) {
        F[m] = creal(s->spe[0][m]);
        F[n + m] = creal(s->spe[1][m]);
        F[2 * n + m] = creal(s->ux[0][0][m]) + creal(s->ux[1][1][m]);
        F[3 * n + m] = creal(s->ux[1][0][m]) - creal(s->ux[0][1][m]);
    }
    float inv[4];
    for (int c = 0; c < 4; ++c) inv[c] = 1.0f / s->cl_scales[c];
    OMP_FOR
    for (long mm = 0; mm < (long)n; ++mm) {
        const int i = (int)(mm % ia), j = (int)(mm / ia);
        float x[36];
        for (int pos = 0; pos < 9; ++pos) {
            const int dx = pos % 3 - 1, dy = pos / 3 - 1;
            const size_t q = (size_t)((i + dx + ia) % ia) + (size_t)ia * (size_t)((j + dy + ja) % ja);
            for (int c = 0; c < 4; ++c) x[4 * pos + c] = bf16r((float)F[c * n + q] * inv[c]);
        }
        double o[2];
        closure_point(s, x, o);
        O[mm] = o[0];
        O[n + mm] = o[1];
        t[0][mm] += o[0];
        t[1][mm] += o[1];
    }
}

int bo_set_closure(bo_state* s, const float* params, const float* scales) {
    if (s->nd != 2 || s->ia != s->ja) return fail(BASTR_EINVAL, "bo_set_closure: square 2-D grids only");
    if (!params) {
        free(s->cl_params);
        free(s->cl_fld);
        free(s->cl_out);
        s->cl_params = NULL;
        s->cl_fld = NULL;
        s->cl_out = NULL;
        return BASTR_OK;
    }
    if (!s->cl_params) {
        s->cl_params = (float*)malloc(sizeof(float) * 6658);
        s->cl_fld = (double*)malloc(sizeof(double) * 4 * s->n);
        s->cl_out = (double*)malloc(sizeof(double) * 2 * s->n);
        if (!s->cl_params || !s->cl_fld || !s->cl_out) return fail(BASTR_ENOMEM, "bo_set_closure: allocation failed");
    }
    memcpy(s->cl_params, params, sizeof(float) * 6658);
    for (int c = 0; c < 6; ++c) s->cl_scales[c] = scales[c];
    return BASTR_OK;
}

int bo_set_closure_per_step(bo_state* s, int32_t per_step) {
    if (per_step != 0 && per_step != 1) return fail(BASTR_EINVAL, "bo_set_closure_per_step: 0 or 1");
    s->cl_per_step = per_step;
    return BASTR_OK;
}

/* compute_ut2D (solution.F90:816-863) and compute_ut3D (:865-928) */
static void compute_ut(bo_state* s, cplx** t) {
    const int nd = s->nd;
    const double nu = s->nu;
    OMP_FOR
    for (long mm = 0; mm < (long)s->n; ++mm) {
        const size_t m = (size_t)mm;
        const double kk2 = nd == 2 ? s->kw[0][m] * s->kw[0][m] + s->kw[1][m] * s->kw[1][m]
                                   : s->kw[0][m] * s->kw[0][m] + s->kw[1][m] * s->kw[1][m] + s->kw[2][m] * s->kw[2][m];
        for (int i = 0; i < nd; ++i) {
            for (int j = 0; j < nd; ++j) s->ux[i][j][m] = I * s->spe[i][m] * s->kw[j][m];
            s->uxixi[i][m] = -s->spe[i][m] * kk2;
        }
    }
    for (int i = 0; i < nd; ++i)
        for (int j = 0; j < nd; ++j) ffti(s, s->ux[i][j]);
    for (int i = 0; i < nd; ++i) ffti(s, s->uxixi[i]);
    for (int i = 0; i < nd; ++i) ffti(s, s->spe[i]);
    OMP_FOR
    for (long mm = 0; mm < (long)s->n; ++mm) {
        const size_t m = (size_t)mm;
        if (nd == 2) {
            for (int i = 0; i < 2; ++i)
                t[i][m] = -creal(s->spe[0][m]) * creal(s->ux[i][0][m])
                          - creal(s->spe[1][m]) * creal(s->ux[i][1][m])
                          + nu * creal(s->uxixi[i][m]);
        } else {
            /* the 3-D form multiplies by the complex u_ispe (solution.F90:903-908) */
            for (int i = 0; i < 3; ++i)
                t[i][m] = -s->spe[0][m] * creal(s->ux[i][0][m]) - s->spe[1][m] * creal(s->ux[i][1][m])
                          - s->spe[2][m] * creal(s->ux[i][2][m]) + nu * creal(s->uxixi[i][m]);
        }
    }
    if (s->cl_params) closure_force(s, t);
    for (int i = 0; i < nd; ++i) fftf(s, t[i]);
    for (int i = 0; i < nd; ++i) dealias(s, t[i]);
    if (s->p.lproject) projection(s, t);
    for (int i = 0; i < nd; ++i) ffti(s, t[i]);
}

static void load_spe(bo_state* s) {
    for (int c = 0; c < s->nd; ++c) {
        cplx* sp = s->spe[c];
        const double* u = s->u[c];
        OMP_FOR
        for (long m = 0; m < (long)s->n; ++m) sp[m] = u[m];
        fftf(s, sp);
    }
}

/* RK32D / RK33D (solution.F90:172-203, 233-270) */
static void rk3(bo_state* s) {
    const int nd = s->nd;
    const double dt = s->dt;
    s->cl_stage = 0;
    compute_ut(s, s->tA);
    for (int c = 0; c < nd; ++c) {
        const double* u = s->u[c];
        const cplx* a = s->tA[c];
        cplx* sp = s->spe[c];
        OMP_FOR
        for (long m = 0; m < (long)s->n; ++m) sp[m] = u[m] + dt * creal(a[m]) / 2.0;
        fftf(s, sp);
    }
    s->cl_stage = 1;
    compute_ut(s, s->tB);
    for (int c = 0; c < nd; ++c) {
        const double* u = s->u[c];
        const cplx* a = s->tA[c];
        const cplx* b = s->tB[c];
        cplx* sp = s->spe[c];
        OMP_FOR
        for (long m = 0; m < (long)s->n; ++m) sp[m] = u[m] - dt * creal(a[m]) + 2.0 * dt * creal(b[m]);
        fftf(s, sp);
    }
    s->cl_stage = 2;
    compute_ut(s, s->tC);
    for (int c = 0; c < nd; ++c) {
        double* u = s->u[c];
        const cplx* a = s->tA[c];
        const cplx* b = s->tB[c];
        const cplx* cc = s->tC[c];
        OMP_FOR
        for (long m = 0; m < (long)s->n; ++m)
            u[m] = u[m] + dt * (creal(a[m]) + 4.0 * creal(b[m]) + creal(cc[m])) / 6.0;
    }
}

/* CN2D / CN3D (solution.F90:205-231, 272-301) */
static void cn(bo_state* s) {
    const int nd = s->nd;
    const double dt = s->dt;
    for (int c = 0; c < nd; ++c) memcpy(s->uold[c], s->u[c], sizeof(double) * s->n);
    for (int step = 1; step <= 6; ++step) {
        compute_ut(s, s->tA);
        for (int c = 0; c < nd; ++c) {
            double* u = s->u[c];
            const double* uo = s->uold[c];
            const cplx* a = s->tA[c];
            OMP_FOR
            for (long m = 0; m < (long)s->n; ++m) u[m] = uo[m] + dt * creal(a[m]);
        }
        if (step == 6) break;
        for (int c = 0; c < nd; ++c) {
            const double* u = s->u[c];
            const double* uo = s->uold[c];
            cplx* sp = s->spe[c];
            OMP_FOR
            for (long m = 0; m < (long)s->n; ++m) sp[m] = 0.5 * (u[m] + uo[m]);
            fftf(s, sp);
        }
    }
}

/* Deterministic chunked sum of f(m) over all points: chunk partials in
 * parallel, combined in chunk order (independent of the thread count). */
typedef struct { double v[4]; } sum4;

/* forcing2D / forcing3D, methods 1, 2, 5 (solution.F90:303-540, 542-814) */
static int forcing(bo_state* s, bastr_diag* dg) {
    const int nd = s->nd;
    const int fm = s->p.forcemethod;
    const double N = nd == 2 ? (double)(s->ia * s->ja) : (double)(s->ia * s->ja * s->ka);
    if (fm == 0) return BASTR_OK;
    if (fm == 1) {
        double part[NCHUNK] = {0};
        OMP_FOR
        for (int ch = 0; ch < NCHUNK; ++ch) {
            const size_t lo = s->n * ch / NCHUNK, hi = s->n * (ch + 1) / NCHUNK;
            double e = 0;
            for (size_t m = lo; m < hi; ++m) {
                double q = 0;
                for (int c = 0; c < nd; ++c) q += s->u[c][m] * s->u[c][m];
                if (nd == 2) q = s->u[0][m] * s->u[0][m] + s->u[1][m] * s->u[1][m];
                e = e + q / 2.0;
            }
            part[ch] = e;
        }
        double energy = 0;
        for (int ch = 0; ch < NCHUNK; ++ch) energy += part[ch];
        energy = energy / N;
        const double factor = sqrt(s->p.target_energy / energy);
        for (int c = 0; c < nd; ++c) {
            double* u = s->u[c];
            OMP_FOR
            for (long m = 0; m < (long)s->n; ++m) u[m] = factor * u[m];
        }
        if (dg) { dg->force_factor = factor; dg->force_E = NAN; dg->force_energy = energy; }
        return BASTR_OK;
    }
    if (fm == 2 || fm == 5) {
        const double dk = 0.5;
        const int fk = s->p.forcek;
        load_spe(s);
        OMP_FOR
        for (long mm = 0; mm < (long)s->n; ++mm) {
            const size_t m = (size_t)mm;
            const double kk = kabs(s, m);
            if ((kk - dk) < fk && (kk + dk) > fk) {
                /* the projected variant exists only in forcing2D (solution.F90:505-512) */
                if (fm == 5 && s->p.lproject && nd == 2) {
                    cplx ud = 0;
                    for (int c = 0; c < nd; ++c) ud += s->spe[c][m] * s->kw[c][m] / kk;
                    for (int c = 0; c < nd; ++c) s->force[c][m] = ud * s->kw[c][m] / kk;
                } else {
                    for (int c = 0; c < nd; ++c) s->force[c][m] = s->spe[c][m];
                }
            } else {
                for (int c = 0; c < nd; ++c) s->force[c][m] = 0.0;
            }
        }
        for (int c = 0; c < nd; ++c) ffti(s, s->force[c]);
        double pe[NCHUNK] = {0}, pu[NCHUNK] = {0};
        OMP_FOR
        for (int ch = 0; ch < NCHUNK; ++ch) {
            const size_t lo = s->n * ch / NCHUNK, hi = s->n * (ch + 1) / NCHUNK;
            double E = 0, en = 0;
            for (size_t m = lo; m < hi; ++m) {
                /* E = E + f1**2 + f2**2 [+ f3**2]; energy = energy + (u1**2 + u2**2 [+ u3**2]) */
                double ue = 0;
                for (int c = 0; c < nd; ++c) {
                    E = E + creal(s->force[c][m]) * creal(s->force[c][m]);
                    ue += s->u[c][m] * s->u[c][m];
                }
                en = en + ue;
            }
            pe[ch] = E; pu[ch] = en;
        }
        double E = 0, energy = 0;
        for (int ch = 0; ch < NCHUNK; ++ch) { E += pe[ch]; energy += pu[ch]; }
        E = E / N;
        energy = energy / N;
        double factor;
        if (fm == 2) {
            factor = sqrt(s->p.target_energy / E) - 1.0;
        } else {
            factor = (s->p.target_energy > energy) ? sqrt((s->p.target_energy - energy) / E + 1.0) - 1.0 : 0.0;
        }
        for (int c = 0; c < nd; ++c) {
            double* u = s->u[c];
            const cplx* f = s->force[c];
            OMP_FOR
            for (long m = 0; m < (long)s->n; ++m) u[m] = u[m] + factor * creal(f[m]);
        }
        if (dg) {
            dg->force_factor = factor;
            dg->force_E = E;
            dg->force_energy = fm == 5 ? energy : NAN;
        }
        return BASTR_OK;
    }
    return fail(BASTR_EINVAL, "forcing: forcemethod 3/4 draw from the unseeded Fortran RNG and are not supported");
}

/* kint(k, dk, 2) = nint(k/dk) (tool.F90:126-147); Fortran nint rounds half away */
static int kint2(double k, double dk) { return (int)lround(k / dk); }

/* spectra_compute2D / 3D (solution.F90:930-1016, 1018-1108) */
static void spectra(bo_state* s, bastr_diag* dg, double* spec) {
    const int nd = s->nd, K = s->allkmax;
    const double dk = 1.0;
    const int nb = K + 1;
    double* pEs = (double*)calloc((size_t)NCHUNK * nb, sizeof(double));
    double* pEd = (double*)calloc((size_t)NCHUNK * nb, sizeof(double));
    double* pkn = (double*)calloc((size_t)NCHUNK * nb, sizeof(double));
    int64_t* pcnt = (int64_t*)calloc((size_t)NCHUNK * nb, sizeof(int64_t));
    sum4 tot[NCHUNK];
    OMP_FOR
    for (int ch = 0; ch < NCHUNK; ++ch) {
        const size_t lo = s->n * ch / NCHUNK, hi = s->n * (ch + 1) / NCHUNK;
        double Edspe = 0, Esspe = 0, k2Edspe = 0, kMEdspe = 0;
        double* Es = pEs + (size_t)ch * nb;
        double* Ed = pEd + (size_t)ch * nb;
        double* kn = pkn + (size_t)ch * nb;
        int64_t* Ec = pcnt + (size_t)ch * nb;
        for (size_t m = lo; m < hi; ++m) {
            const double kk = kabs(s, m);
            double es_m = 0, ed_m = 0, es2 = 0, es3 = 0;
            if (kk > dk / 2) {
                if (nd == 2) {
                    const double k1 = s->kw[0][m], k2 = s->kw[1][m];
                    const cplx usspe = s->spe[0][m] * k2 / kk - s->spe[1][m] * k1 / kk;
                    const cplx udspe = s->spe[0][m] * k1 / kk + s->spe[1][m] * k2 / kk;
                    kMEdspe = kMEdspe + creal(udspe * conj(udspe)) / 2 / kk;
                    es_m = creal(usspe * conj(usspe)) / 2;
                    ed_m = creal(udspe * conj(udspe)) / 2;
                } else {
                    const double k1 = s->kw[0][m], k2 = s->kw[1][m], k3 = s->kw[2][m];
                    const cplx udspe = s->spe[0][m] * k1 / kk + s->spe[1][m] * k2 / kk + s->spe[2][m] * k3 / kk;
                    const cplx u1d = udspe * k1 / kk, u2d = udspe * k2 / kk, u3d = udspe * k3 / kk;
                    const cplx u1s = s->spe[0][m] - u1d, u2s = s->spe[1][m] - u2d, u3s = s->spe[2][m] - u3d;
                    kMEdspe = kMEdspe + creal(udspe * conj(udspe)) / 2 / kk;
                    es_m = creal(u1s * conj(u1s)) / 2;
                    es2 = creal(u2s * conj(u2s)) / 2;
                    es3 = creal(u3s * conj(u3s)) / 2;
                    ed_m = creal(udspe * conj(udspe)) / 2;
                }
            }
            const int ko = kint2(kk, dk);
            if (ko <= K) {
                Ec[ko] += 1;
                Es[ko] = Es[ko] + es_m + es2 + es3;
                Ed[ko] = Ed[ko] + ed_m;
                kn[ko] = kn[ko] + kk;
            }
            Edspe = Edspe + ed_m;
            Esspe = Esspe + es_m + es2 + es3;
            k2Edspe = k2Edspe + ed_m * (kk * kk);
        }
        tot[ch].v[0] = Esspe; tot[ch].v[1] = Edspe; tot[ch].v[2] = k2Edspe; tot[ch].v[3] = kMEdspe;
    }
    double T[4] = {0, 0, 0, 0};
    for (int b = 0; b <= K; ++b) { s->Es[b] = 0; s->Ed[b] = 0; s->kn[b] = 0; s->Ecount[b] = 0; }
    for (int ch = 0; ch < NCHUNK; ++ch) {
        for (int q = 0; q < 4; ++q) T[q] += tot[ch].v[q];
        for (int b = 0; b <= K; ++b) {
            s->Es[b] += pEs[(size_t)ch * nb + b];
            s->Ed[b] += pEd[(size_t)ch * nb + b];
            s->kn[b] += pkn[(size_t)ch * nb + b];
            s->Ecount[b] += pcnt[(size_t)ch * nb + b];
        }
    }
    /* do i=1,allkmax: kn(i) = psum(kn(i))/Ecount(i) */
    for (int b = 1; b <= K; ++b) s->kn[b] = s->kn[b] / (double)s->Ecount[b];
    s->Esspe = T[0];
    s->Edspe = T[1];
    if (dg) { dg->Esspe = T[0]; dg->Edspe = T[1]; dg->k2Edspe = T[2]; dg->kMEdspe = T[3]; }
    if (spec) {
        for (int b = 1; b <= K; ++b) {
            spec[b - 1] = s->kn[b];
            spec[K + b - 1] = s->Es[b];
            spec[2 * K + b - 1] = s->Ed[b];
        }
    }
    free(pEs); free(pEd); free(pkn); free(pcnt);
}

/* velgrad_calculate2D / 3D (solution.F90:1110-1186, 1188-1286) */
static void velgrad(bo_state* s, bastr_diag* dg) {
    const int nd = s->nd;
    const double nu = s->nu;
    OMP_FOR
    for (long mm = 0; mm < (long)s->n; ++mm) {
        const size_t m = (size_t)mm;
        double kk = 0;
        if (nd == 2) kk = s->kw[0][m] * s->kw[0][m] + s->kw[1][m] * s->kw[1][m];
        else kk = s->kw[0][m] * s->kw[0][m] + s->kw[1][m] * s->kw[1][m] + s->kw[2][m] * s->kw[2][m];
        for (int i = 0; i < nd; ++i) {
            for (int j = 0; j < nd; ++j) s->ux[i][j][m] = I * s->spe[i][m] * s->kw[j][m];
            s->uxixi[i][m] = -s->spe[i][m] * kk;
        }
        if (nd == 2)
            s->thetaxixi[m] = -I * (s->spe[0][m] * s->kw[0][m] * kk + s->spe[1][m] * s->kw[1][m] * kk);
        else
            s->thetaxixi[m] = -I * (s->spe[0][m] * s->kw[0][m] * kk + s->spe[1][m] * s->kw[1][m] * kk
                                    + s->spe[2][m] * s->kw[2][m] * kk);
    }
    for (int i = 0; i < nd; ++i)
        for (int j = 0; j < nd; ++j) ffti(s, s->ux[i][j]);
    for (int i = 0; i < nd; ++i) ffti(s, s->uxixi[i]);
    ffti(s, s->thetaxixi);

    double part[NCHUNK][8];
    double pmin[NCHUNK];
    const double eta0 = (double)(2.0f * 3.14f); /* eta_min = 2*3.14 in default real */
    OMP_FOR
    for (int ch = 0; ch < NCHUNK; ++ch) {
        const size_t lo = s->n * ch / NCHUNK, hi = s->n * (ch + 1) / NCHUNK;
        double umumtheta2 = 0, umumijji = 0, u2theta = 0, dissp = 0, urms = 0, dudx2 = 0, epsilon = 0;
        double eta_min = eta0;
        for (size_t m = lo; m < hi; ++m) {
            if (nd == 2) {
                const double a11 = creal(s->ux[0][0][m]), a12 = creal(s->ux[0][1][m]);
                const double a21 = creal(s->ux[1][0][m]), a22 = creal(s->ux[1][1][m]);
                const double u1 = s->u[0][m], u2 = s->u[1][m];
                const double div = a11 + a22;
                umumtheta2 = umumtheta2 + div * div * (u1 * u1 + u2 * u2);
                umumijji = umumijji + (a11 * a11 + a22 * a22 + 2.0 * a12 * a21) * (u1 * u1 + u2 * u2);
                u2theta = u2theta + div * (u1 * u1 + u2 * u2);
                dissp = dissp + 2.0 * u1 * div * creal(s->uxixi[0][m]) + 2.0 * u2 * div * creal(s->uxixi[1][m])
                        + (u1 * u1 + u2 * u2) * creal(s->thetaxixi[m]);
                epsilon = epsilon + nu * (a12 - a21) * (a12 - a21) + nu * div * div;
                eta_min = fmin(eta_min, sqrt(nu / fabs(div)));
                urms = urms + u1 * u1 + u2 * u2;
                dudx2 = dudx2 + a11 * a11 + a22 * a22;
            } else {
                double a[3][3];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) a[i][j] = creal(s->ux[i][j][m]);
                const double u1 = s->u[0][m], u2 = s->u[1][m], u3 = s->u[2][m];
                const double uu = (u1 * u1 + u2 * u2 + u3 * u3);
                const double div = a[0][0] + a[1][1] + a[2][2];
                umumtheta2 = umumtheta2 + div * div * uu;
                umumijji = umumijji + (a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2]
                                       + 2.0 * a[0][1] * a[1][0] + 2.0 * a[0][2] * a[2][0]
                                       + 2.0 * a[1][2] * a[2][1]) * uu;
                u2theta = u2theta + div * uu;
                dissp = dissp + 2.0 * u1 * div * creal(s->uxixi[0][m]) + 2.0 * u2 * div * creal(s->uxixi[1][m])
                        + 2.0 * u3 * div * creal(s->uxixi[2][m]) + uu * creal(s->thetaxixi[m]);
                epsilon = epsilon + nu * (a[0][1] - a[1][0]) * (a[0][1] - a[1][0])
                          + nu * (a[1][2] - a[2][1]) * (a[1][2] - a[2][1])
                          + nu * (a[0][2] - a[2][0]) * (a[0][2] - a[2][0]) + nu * div * div;
                eta_min = fmin(eta_min, sqrt(nu / fabs(div)));
                urms = urms + u1 * u1 + u2 * u2 + u3 * u3;
                dudx2 = dudx2 + a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
            }
        }
        part[ch][0] = umumtheta2; part[ch][1] = umumijji; part[ch][2] = u2theta; part[ch][3] = dissp;
        part[ch][4] = urms; part[ch][5] = dudx2; part[ch][6] = epsilon;
        pmin[ch] = eta_min;
    }
    double v[7] = {0};
    double eta_min = eta0;
    for (int ch = 0; ch < NCHUNK; ++ch) {
        for (int q = 0; q < 7; ++q) v[q] += part[ch][q];
        eta_min = fmin(eta_min, pmin[ch]);
    }
    const double N = nd == 2 ? (double)(s->ia * s->ja) : (double)(s->ia * s->ja * s->ka);
    const double umumtheta2 = v[0] / N, umumijji = v[1] / N, u2theta = v[2] / N, dissp = v[3] / N;
    const double urms = sqrt(v[4] / N), dudx2 = v[5] / N, epsilon = v[6] / N;
    const double dim = nd == 2 ? 2.0 : 3.0;
    const double ufluc = urms / sqrt(dim);
    const double Taylength = ufluc / sqrt(dudx2 / dim);
    const double ReTay = ufluc * Taylength / nu;
    const double Kollength = sqrt(sqrt(nu * nu * nu / epsilon));
    if (dg) {
        dg->umumtheta2 = umumtheta2; dg->umumijji = umumijji; dg->u2theta = u2theta; dg->dissp = dissp;
        dg->eta_min = eta_min; dg->Taylength = Taylength; dg->ReTay = ReTay; dg->Kollength = Kollength;
    }
}

/* mainloop2D/3D entry dealias (solution.F90:41-52, 105-122) */
static void prologue(bo_state* s) {
    load_spe(s);
    for (int c = 0; c < s->nd; ++c) dealias(s, s->spe[c]);
    for (int c = 0; c < s->nd; ++c) {
        ffti(s, s->spe[c]);
        double* u = s->u[c];
        const cplx* sp = s->spe[c];
        OMP_FOR
        for (long m = 0; m < (long)s->n; ++m) u[m] = creal(sp[m]);
    }
}

/* ------------------------------------------------------------------------- */
/* public API                                                                 */
/* ------------------------------------------------------------------------- */

int bo_create(const bastr_params* p, bo_state** out) {
    if (!p || !out) return fail(BASTR_EINVAL, "bo_create: NULL argument");
    if (p->ia < 2 || p->ja < 2 || p->ka < 0) return fail(BASTR_EINVAL, "bo_create: bad grid");
    if (p->timemethod != 1 && p->timemethod != 2) return fail(BASTR_EINVAL, "Undefined time method");
    if (p->forcemethod == 3 || p->forcemethod == 4 || p->forcemethod < 0 || p->forcemethod > 5)
        return fail(BASTR_EINVAL, "bo_create: unsupported forcemethod");
    bo_state* s = (bo_state*)calloc(1, sizeof(bo_state));
    s->p = *p;
    s->nd = p->ka == 0 ? 2 : 3;
    s->ia = p->ia; s->ja = p->ja; s->ka = p->ka;
    s->nz = s->nd == 3 ? p->ka : 1;
    s->n = (size_t)s->ia * s->ja * s->nz;
    s->allkmax = (int)(s->ia / 3);
    if (p->kmax < s->allkmax) s->allkmax = p->kmax;
    s->nu = p->reynolds == 0.0 ? 0.0 : 1.0 / p->reynolds;
    s->dt = p->deltat;
    plannd_init(&s->plan, s->ia, s->ja, s->nd == 3 ? s->ka : 0);
    int ok = 1;
    const int nd = s->nd;
    for (int c = 0; c < nd; ++c) {
        s->u[c] = (double*)xcalloc(s->n, sizeof(double), &ok);
        s->kw[c] = (double*)xcalloc(s->n, sizeof(double), &ok);
        s->spe[c] = (cplx*)xcalloc(s->n, sizeof(cplx), &ok);
        s->uxixi[c] = (cplx*)xcalloc(s->n, sizeof(cplx), &ok);
        s->tA[c] = (cplx*)xcalloc(s->n, sizeof(cplx), &ok);
        s->force[c] = (cplx*)xcalloc(s->n, sizeof(cplx), &ok);
        if (p->timemethod == 1) {
            s->tB[c] = (cplx*)xcalloc(s->n, sizeof(cplx), &ok);
            s->tC[c] = (cplx*)xcalloc(s->n, sizeof(cplx), &ok);
        } else {
            s->uold[c] = (double*)xcalloc(s->n, sizeof(double), &ok);
        }
        for (int j = 0; j < nd; ++j) s->ux[c][j] = (cplx*)xcalloc(s->n, sizeof(cplx), &ok);
    }
    s->thetaxixi = (cplx*)xcalloc(s->n, sizeof(cplx), &ok);
    const int nb = s->allkmax + 1;
    s->Es = (double*)xcalloc((size_t)nb, sizeof(double), &ok);
    s->Ed = (double*)xcalloc((size_t)nb, sizeof(double), &ok);
    s->kn = (double*)xcalloc((size_t)nb, sizeof(double), &ok);
    s->Ecount = (int64_t*)xcalloc((size_t)nb, sizeof(int64_t), &ok);
    if (!ok) { bo_destroy(s); return fail(BASTR_ENOMEM, "bo_create: out of memory"); }
    generate_wave(s);
    s->nstep = 0;
    s->time = 0.0;
    *out = s;
    return BASTR_OK;
}

int bo_destroy(bo_state* s) {
    if (!s) return BASTR_OK;
    for (int c = 0; c < 3; ++c) {
        free(s->u[c]); free(s->uold[c]); free(s->kw[c]); free(s->spe[c]); free(s->uxixi[c]);
        free(s->tA[c]); free(s->tB[c]); free(s->tC[c]); free(s->force[c]);
        for (int j = 0; j < 3; ++j) free(s->ux[c][j]);
    }
    free(s->thetaxixi);
    free(s->Es); free(s->Ed); free(s->kn); free(s->Ecount);
    free(s->cl_params); free(s->cl_fld); free(s->cl_out);
    if (s->plan.p[0].tw) plannd_free(&s->plan);
    free(s);
    return BASTR_OK;
}

int bo_allkmax(const bo_state* s) { return s->allkmax; }

int bo_set_fields(bo_state* s, const double* u1, const double* u2, const double* u3) {
    const double* src[3] = {u1, u2, u3};
    for (int c = 0; c < s->nd; ++c) {
        if (!src[c]) return fail(BASTR_EINVAL, "bo_set_fields: missing component");
        memcpy(s->u[c], src[c], sizeof(double) * s->n);
    }
    s->prologue_pending = 1;
    s->have_fields = 1;
    return BASTR_OK;
}

int bo_get_fields(bo_state* s, double* u1, double* u2, double* u3) {
    double* dst[3] = {u1, u2, u3};
    if (!s->have_fields) return fail(BASTR_EINVAL, "bo_get_fields: fields not set");
    /* fields become observable inside mainloop, i.e. after its entry dealias */
    if (s->prologue_pending) { prologue(s); s->prologue_pending = 0; }
    for (int c = 0; c < s->nd; ++c) {
        if (!dst[c]) return fail(BASTR_EINVAL, "bo_get_fields: missing component");
        memcpy(dst[c], s->u[c], sizeof(double) * s->n);
    }
    return BASTR_OK;
}

int bo_set_time(bo_state* s, int64_t nstep, double time) { s->nstep = nstep; s->time = time; return BASTR_OK; }
int bo_get_time(const bo_state* s, int64_t* nstep, double* time) {
    if (nstep) *nstep = s->nstep;
    if (time) *time = s->time;
    return BASTR_OK;
}

int bo_shell_counts(const bo_state* s, int64_t* counts) {
    /* Ecount depends on the grid only: count modes with nint(|k|) = 1..allkmax */
    const int K = s->allkmax;
    for (int b = 0; b < K; ++b) counts[b] = 0;
    for (size_t m = 0; m < s->n; ++m) {
        const int ko = kint2(kabs(s, m), 1.0);
        if (ko >= 1 && ko <= K) counts[ko - 1] += 1;
    }
    return BASTR_OK;
}

/* One iteration of mainloop2D/3D's do-while body (solution.F90:54-97 / 124-169). */
int bo_step(bo_state* s, int32_t nsteps, bastr_diag* diags, double* spectra_out) {
    if (!s->have_fields) return fail(BASTR_EINVAL, "bo_step: fields not set");
    if (s->prologue_pending) { prologue(s); s->prologue_pending = 0; }
    const int K = s->allkmax;
    for (int it = 0; it < nsteps; ++it) {
        bastr_diag dloc;
        memset(&dloc, 0, sizeof dloc);
        bastr_diag* dg = diags ? &diags[it] : &dloc;
        memset(dg, 0, sizeof *dg);
        dg->force_factor = dg->force_E = dg->force_energy = dg->force_Fen = NAN;
        dg->nstep = s->nstep;
        dg->time = s->time;
        load_spe(s);
        spectra(s, dg, spectra_out ? spectra_out + (size_t)it * 3 * K : NULL);
        velgrad(s, dg);
        if (isnan(s->Esspe + s->Edspe)) {
            dg->nan_flag = 1;
            return BASTR_STOP_NAN;
        }
        s->nstep = s->nstep + 1;
        s->time = s->time + s->dt;
        if (s->p.timemethod == 1) rk3(s);
        else cn(s);
        int rc = forcing(s, dg);
        if (rc) return rc;
    }
    return BASTR_OK;
}

static int fft_common(bo_state* s, double* inout, int dir) {
    cplx* a = (cplx*)inout;
    if (dir < 0) fftf(s, a);
    else ffti(s, a);
    return BASTR_OK;
}
int bo_fft_forward(bo_state* s, double* inout) { return fft_common(s, inout, -1); }
int bo_fft_backward(bo_state* s, double* inout) { return fft_common(s, inout, +1); }

/* ------------------------------------------------------------------------- */
/* seeded random-phase initial field                                          */
/* ------------------------------------------------------------------------- */

/* SplitMix64 (Steele, Lea, Flood 2014), the noise source that stands in for
 * the reference's clock-seeded random_number (Bastrpp.F90:626-637). */
static uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

int bo_set_spectrum_law(bo_state* s, int32_t law) {
    if (law != BASTR_LAW_HITGEN && law != BASTR_LAW_PURE) return fail(BASTR_EINVAL, "bo_set_spectrum_law: bad law");
    s->law = law;
    return BASTR_OK;
}

/* create_initial_field_2d / _3d (Bastrpp.F90:601-696, 698-808), statement by
 * statement, with k0 in the role of forcek and `energy` in the role of
 * target_energy.  Loads the field like bo_set_fields. */
int bo_init_random_phase(bo_state* s, double k0, uint64_t seed, double energy) {
    const double PI = 3.14159265358979323846;
    const size_t n = s->n;
    if (!(k0 > 0) || !(energy > 0)) return fail(BASTR_EINVAL, "bo_init_random_phase: k0 and energy must be > 0");
    /* random_angle = random_number(); random_complex = cmplx(random_angle, 0); fft */
    cplx* rc = s->tA[0];
    for (size_t m = 0; m < n; ++m)
        rc[m] = (double)(splitmix64(seed * 0x100000001B3ull + m) >> 11) * 0x1.0p-53;
    fftf(s, rc);
    for (size_t m = 0; m < n; ++m) {
        const double kx = s->kw[0][m], ky = s->kw[1][m], kz = s->nd == 3 ? s->kw[2][m] : 0.0;
        const double kk = sqrt(kx * kx + ky * ky + kz * kz);
        if (kx == 0 && ky == 0 && kz == 0) {
            for (int c = 0; c < s->nd; ++c) s->spe[c][m] = 0.0;
            continue;
        }
        const double q = kk / k0, kk2 = kk * kk;
        double var1;
        if (kk >= k0 || s->law == BASTR_LAW_PURE) var1 = kk2 * kk2 * exp(-2.0 * (q * q));
        else var1 = k0 * k0 * k0 * k0 * k0 * k0 / kk2 * exp(-2.0);
        const double var2 = s->nd == 2 ? sqrt(var1 / 2.0 / PI / kk) : sqrt(var1 / 4.0 / PI / (kk * kk));
        const double th = atan2(cimag(rc[m]), creal(rc[m]));
        const cplx ph = sin(th) + I * cos(th);
        const double kc[3] = {kx, ky, kz};
        for (int c = 0; c < s->nd; ++c) s->spe[c][m] = var2 * kc[c] / kk * ph;
    }
    double e = 0.0;
    for (int c = 0; c < s->nd; ++c) {
        ffti(s, s->spe[c]);
        for (size_t m = 0; m < n; ++m) s->u[c][m] = creal(s->spe[c][m]);
    }
    for (size_t m = 0; m < n; ++m) {
        double q = 0.0;
        for (int c = 0; c < s->nd; ++c) q += s->u[c][m] * s->u[c][m];
        e += q;
    }
    e /= (double)n;
    const double factor = sqrt(energy / e);
    for (int c = 0; c < s->nd; ++c)
        for (size_t m = 0; m < n; ++m) s->u[c][m] *= factor;
    s->prologue_pending = 1;
    s->have_fields = 1;
    s->nstep = 0;
    s->time = 0.0;
    return BASTR_OK;
}

/* ------------------------------------------------------------------------- */
/* post-processing of a field: Bastrpp's spectra2d and transfer2d/3d          */
/* ------------------------------------------------------------------------- */

static cplx* xalloc_c(size_t n) { return (cplx*)calloc(n, sizeof(cplx)); }

/* spectra2d (Bastrpp.F90:299-400): shell sums of |u_s|^2 and |u_d|^2 over the
 * full c2c spectrum of a field as read from a dump (no entry dealias, and no
 * 1/2, unlike spectra_compute2D).  shells[0..K-1] = Es(1..K), shells[K..2K-1] =
 * Ed(1..K); totals = Esspe, Edspe, k2Edspe, kM1Edspe (the module's Esspe and
 * Edspe start from zero in a Bastrpp run). */
int bo_pp_spectra(bo_state* s, const double* u1, const double* u2, const double* u3, double* shells,
                  double* totals) {
    (void)u3;
    if (s->nd != 2) return fail(BASTR_EINVAL, "bo_pp_spectra: spectra2d is 2-D only");
    if (!u1 || !u2 || !shells || !totals) return fail(BASTR_EINVAL, "bo_pp_spectra: NULL argument");
    const size_t n = s->n;
    const int K = s->allkmax;
    cplx* a = xalloc_c(n);
    cplx* b = xalloc_c(n);
    double* Es = (double*)calloc((size_t)K + 1, sizeof(double));
    double* Ed = (double*)calloc((size_t)K + 1, sizeof(double));
    if (!a || !b || !Es || !Ed) { free(a); free(b); free(Es); free(Ed); return fail(BASTR_ENOMEM, "bo_pp_spectra"); }
    for (size_t m = 0; m < n; ++m) { a[m] = u1[m]; b[m] = u2[m]; }
    fftf(s, a);
    fftf(s, b);
    const double dk = 1.0;
    double Esspe = 0, Edspe = 0, k2Edspe = 0, kM1Edspe = 0;
    for (size_t m = 0; m < n; ++m) {
        const double k1 = s->kw[0][m], k2 = s->kw[1][m];
        const double kk = sqrt(k1 * k1 + k2 * k2);
        cplx usspe = 0, udspe = 0;
        if (kk > dk / 2) {
            usspe = a[m] * k2 / kk - b[m] * k1 / kk;
            udspe = a[m] * k1 / kk + b[m] * k2 / kk;
            kM1Edspe = kM1Edspe + creal(udspe * conj(udspe)) / kk;
        }
        const int ko = kint2(kk, dk);
        if (ko <= K) {
            Es[ko] = Es[ko] + creal(usspe * conj(usspe));
            Ed[ko] = Ed[ko] + creal(udspe * conj(udspe));
        }
        Esspe = Esspe + creal(usspe * conj(usspe));
        Edspe = Edspe + creal(udspe * conj(udspe));
        k2Edspe = k2Edspe + creal(udspe * conj(udspe)) * (kk * kk);
    }
    for (int i = 1; i <= K; ++i) { shells[i - 1] = Es[i]; shells[K + i - 1] = Ed[i]; }
    totals[0] = Esspe; totals[1] = Edspe; totals[2] = k2Edspe; totals[3] = kM1Edspe;
    free(a); free(b); free(Es); free(Ed);
    return BASTR_OK;
}

/* transfer2d / transfer3d (Bastrpp.F90:174-297, 402-599): shell sums over the
 * full c2c spectrum of
 *   T  = Re(conj(u_hat_i) fft(-u_j d_j u_i)),
 *   T1 = Im(k_j fft(u_i u_j) conj(u_hat_i)),
 *   T2 = Re(conj(u_hat_i) fft(u_i theta)),  theta = d_j u_j,
 * with the derivatives spectral and the products in physical space, from a
 * field as read from a dump.  out[0..K-1] = T(1..K), out[K..] = T1, out[2K..] = T2. */
int bo_pp_transfer(bo_state* s, const double* u1, const double* u2, const double* u3, double* out) {
    const int nd = s->nd;
    const size_t n = s->n;
    const int K = s->allkmax;
    const double* src[3] = {u1, u2, u3};
    for (int c = 0; c < nd; ++c)
        if (!src[c]) return fail(BASTR_EINVAL, "bo_pp_transfer: missing component");
    if (!out) return fail(BASTR_EINVAL, "bo_pp_transfer: NULL argument");
    cplx* sp[3] = {0, 0, 0};
    cplx* ux[3][3] = {{0}};
    cplx* tA[3] = {0, 0, 0};
    cplx* tC[3] = {0, 0, 0};
    int ok = 1;
    for (int i = 0; i < nd; ++i) {
        if (!(sp[i] = xalloc_c(n)) || !(tA[i] = xalloc_c(n)) || !(tC[i] = xalloc_c(n))) ok = 0;
        for (int j = 0; j < nd; ++j)
            if (!(ux[i][j] = xalloc_c(n))) ok = 0;
    }
    double* T = (double*)calloc((size_t)K + 1, sizeof(double));
    double* T1 = (double*)calloc((size_t)K + 1, sizeof(double));
    double* T2 = (double*)calloc((size_t)K + 1, sizeof(double));
    if (!T || !T1 || !T2) ok = 0;
    if (ok) {
        for (int i = 0; i < nd; ++i) {
            for (size_t m = 0; m < n; ++m) sp[i][m] = src[i][m];
            fftf(s, sp[i]);
        }
        /* uixj = imag * uispe * kj; ifft of all, then of uispe */
        for (int i = 0; i < nd; ++i)
            for (int j = 0; j < nd; ++j) {
                for (size_t m = 0; m < n; ++m) ux[i][j][m] = I * sp[i][m] * s->kw[j][m];
                ffti(s, ux[i][j]);
            }
        for (int i = 0; i < nd; ++i) ffti(s, sp[i]);
        for (size_t m = 0; m < n; ++m) {
            /* theta = dreal(u1x1 + u2x2 [+ u3x3]) */
            cplx dsum = 0;
            for (int j = 0; j < nd; ++j) dsum += ux[j][j][m];
            const double th = creal(dsum);
            for (int i = 0; i < nd; ++i) {
                cplx t = 0;
                for (int j = 0; j < nd; ++j) t = t - sp[j][m] * creal(ux[i][j][m]);
                tA[i][m] = t;
                tC[i][m] = sp[i][m] * th;
            }
            /* the derivative arrays are reused for the products u_i u_j */
            cplx uu[3][3];
            for (int i = 0; i < nd; ++i)
                for (int j = 0; j < nd; ++j) uu[i][j] = sp[i][m] * sp[j][m];
            for (int i = 0; i < nd; ++i)
                for (int j = 0; j < nd; ++j) ux[i][j][m] = uu[i][j];
        }
        for (int i = 0; i < nd; ++i) { fftf(s, tA[i]); }
        for (int i = 0; i < nd; ++i) { fftf(s, sp[i]); }
        for (int i = 0; i < nd; ++i) { fftf(s, tC[i]); }
        for (int i = 0; i < nd; ++i)
            for (int j = 0; j < nd; ++j) fftf(s, ux[i][j]);
        for (size_t m = 0; m < n; ++m) {
            double kk2 = 0;
            for (int j = 0; j < nd; ++j) kk2 += s->kw[j][m] * s->kw[j][m];
            const double kk = sqrt(kk2);
            double trans = 0, trans2 = 0;
            cplx z1 = 0;
            for (int i = 0; i < nd; ++i) {
                trans += creal(conj(sp[i][m]) * tA[i][m]);
                trans2 += creal(conj(sp[i][m]) * tC[i][m]);
                for (int j = 0; j < nd; ++j) z1 += s->kw[j][m] * ux[i][j][m] * conj(sp[i][m]);
            }
            const int ko = kint2(kk, 1.0);
            if (ko <= K) {
                T[ko] = T[ko] + trans;
                T1[ko] = T1[ko] + cimag(z1);
                T2[ko] = T2[ko] + trans2;
            }
        }
        for (int i = 1; i <= K; ++i) {
            out[i - 1] = T[i];
            out[K + i - 1] = T1[i];
            out[2 * K + i - 1] = T2[i];
        }
    }
    for (int i = 0; i < 3; ++i) {
        free(sp[i]); free(tA[i]); free(tC[i]);
        for (int j = 0; j < 3; ++j) free(ux[i][j]);
    }
    free(T); free(T1); free(T2);
    return ok ? BASTR_OK : fail(BASTR_ENOMEM, "bo_pp_transfer: allocation failed");
}
