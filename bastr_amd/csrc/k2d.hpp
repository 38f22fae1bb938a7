// k2d.hpp -- 2-D pseudo-spectral step kernels (gfx950).
//
// Layout in HBM (x = index i of the reference, fastest):
//   physical real field  u[y][x]            (Fortran u(i,j,0) order)
//   half spectrum        S[ky][kx], kx = 0..n0/2, row pitch hp >= n0/2+1
// A half spectrum holds the normalised coefficients u_hat(kx,ky) of the
// reference (fftwlink.F90:333-334); modes with kx < 0 are their Hermitian
// mirrors and are never stored.
//
// Column (y) passes work on CB adjacent kx columns per workgroup (CB*sizeof
// complex >= 64 B contiguous per row); row (x) passes transform one packed
// complex row z = a + i*b per real pair (a, b) and split/merge the half
// spectra on the fly.
#pragma once
#include "fft_block.hpp"
#include "../../include/bastr.h"

namespace k2d {

// Column-pass launch shape for (T, N): CB adjacent columns per workgroup,
// E points per thread (EMAX from N = 256 up, else 8), P = N/E threads per
// column, about 256 threads per workgroup (64 for columns of at most 256
// points: a small grid's pass then spreads over four times the CUs; measured at
// 256^2 FP64, 0.0736 -> 0.0713 ms per iteration).  CB*sizeof(complex) >= 64 B in
// FP64 up to 2048-point columns, 128 B in FP32 up to 1024 (measured: FP32
// columns of 64-byte rows fetched 1.5x their bytes), 32 B otherwise, keeps
// each row access to whole sectors (the neighbouring column blocks, run on
// the same XCD, read the rest of each line from its L2; narrower blocks cost
// extra write-backs of partly written lines); CB <= 16 so
// CB divides every supported row length; LDS is CB*N*sizeof(T) (+1/16
// padding) for the split real/imaginary exchange.  Small workgroups let
// several of them share a CU, so one block's loads overlap another's
// transform.
// CBM = 2 doubles the block width (a wider tile, chosen by the autotuner where
// it fits: at most 1024 threads and 160 KiB of LDS per workgroup).
// The default plan (EMAX = 16) on FP64 columns of 64 to 256 points holds 4
// points per thread in one-wavefront blocks, below the 64-byte row minimum
// (the columns of such a grid sit in the caches): 256^2 FP64 0.0625 -> 0.0615
// ms per iteration.
template <typename T, int N, int EMAX = 16, int CBM = 1>
struct ColCfg {
    static constexpr bool SHORT = sizeof(T) == 8 && EMAX == 16 && N <= 256 && N >= 64;
    static constexpr int E = SHORT ? 4 : (N >= 256 ? EMAX : 8);
    static constexpr int P = N / E;
    static constexpr int CBMIN =
        (sizeof(T) == 8 && N <= 2048 ? 64 : (sizeof(T) == 4 && N <= 1024 ? 128 : 32)) / (int)sizeof(cx<T>);
    static constexpr int TW = (N <= 256 && P <= 64) ? 64 : 256;  // threads a block aims at
    static constexpr int CBW = SHORT ? TW / P : ((TW / P) > CBMIN ? (TW / P) : CBMIN);
    static constexpr int CB = (CBW > 16 ? 16 : CBW) * CBM;
    static constexpr int THREADS = CB * P;
    static constexpr int LDS_BYTES = CB * (N + N / 16) * (int)sizeof(T);
    static constexpr bool FITS = THREADS <= 1024 && LDS_BYTES <= 163840;  // a plan a workgroup can hold
    static_assert(CB >= 1 && N >= E, "column config");
    static_assert(EMAX == 8 || EMAX == 16, "column config");
    static_assert(CBM != 1 || LDS_BYTES <= 81920, "column LDS");
};

// Row-pass launch shape for (T, N0): L rows per workgroup, 8 points per
// thread (measured: FP32 lines of 1024-2048 points at 16 points per thread,
// two LDS exchanges instead of three, ran 1.3x slower in the stage pass).
// About 256 threads per workgroup, but one wavefront where a row takes fewer
// than 64 threads, and in FP64 rows of 64 to 256 points at 4 points per thread
// (a 256-point row is one wavefront): a small grid's pass spreads over more CUs
// and each thread's transform chain is half as long (256^2 FP64: row stage
// 10.5 -> 9.0 us at one wavefront per block, iteration 0.0700 -> 0.0653 ms at 4
// points per thread; in FP32 the 96-member ensemble measured 1.1 % slower at 4
// points, so FP32 keeps 8).
template <typename T, int N>
struct RowCfg {
    static constexpr bool SHORT = sizeof(T) == 8 && N <= 256 && N >= 64;
    static constexpr int E = SHORT ? 4 : 8;
    static constexpr int P = N / E;
    static constexpr int L0 = P >= 256 ? 1 : (P < 64 || SHORT ? 64 / P : 256 / P);
    static constexpr int LMAX = 65536 / (N * (int)sizeof(cx<T>));  // (+1/8 padding)
    static constexpr int L = L0 < LMAX ? L0 : (LMAX < 1 ? 1 : LMAX);
    static constexpr int THREADS = L * P;
    // waves per SIMD asked of the compiler: FP32 stage passes fit 96 registers (5
    // waves); FP64 ones 128 (4 waves, measured: 2048^2 row stage 54.5 -> 49.9 us
    // from 3), but for 1024-point rows, which would spill there
    static constexpr int WPE = sizeof(T) == 4 ? 5 : (N == 1024 ? 1 : 4);
};

// Row-pass shape of the velocity-gradient kernels (8 points per thread, FP32
// included: at 16 the stage-A kernel's four lines and seven FP64 moment sums
// spill; the kernels order their transforms so that at most three lines are
// live).  EE = 0: the 2-D kernel's shape, in FP64 4 points per thread and one
// wavefront per row block for rows of 64 to 256 points as in RowCfg; the 3-D
// kernel asks for 8.
template <typename T, int N, int EE = 0>
struct VgCfg {
    static constexpr bool SHORT = EE == 0 && sizeof(T) == 8 && N <= 256 && N >= 64;
    static constexpr int E = EE ? EE : (SHORT ? 4 : 8);
    static constexpr int P = N / E;
    static constexpr int L0 = P >= 256 ? 1 : (P < 64 || SHORT ? 64 / P : 256 / P);
    static constexpr int LMAX = 65536 / (N * (int)sizeof(cx<T>));  // (+1/8 padding)
    static constexpr int L = L0 < LMAX ? L0 : (LMAX < 1 ? 1 : LMAX);
    static constexpr int THREADS = L * P;
};

template <typename T>
struct Geo {
    int n0, n1;      // ia, ja
    int nh;          // n0/2 + 1
    int hp;          // half-spectrum pitch
    int cut2;        // floor(ia/3)^2: dealias when kx^2+ky^2 >= cut2 (tool.F90:185-191)
    int lproject;
    int kxlim;       // x passes: half-row entries kx >= kxlim are zero (not read, not written)
    double nu, dt, inv_n;
    // ensemble members: blockIdx.z is the member; member m's fields start
    // m*hstride complex (half spectra) or m*rstride real (physical) elements
    // after member 0's
    long hstride, rstride;
    int nmem;        // launch width along z (0 or 1: one member)
};

__host__ __device__ __forceinline__ unsigned zdim(int nmem) { return nmem > 1 ? (unsigned)nmem : 1u; }

// Iteration gate: kernels of the diagnostic phase of iteration `it` run while
// no NaN was found in an earlier iteration; kernels of the step phase also
// stop for the iteration whose diagnostics were NaN (solution.F90:81).  Each
// ensemble member (blockIdx.z) has its own status word.
struct Gate {
    const int* status;  // [member] first iteration index with NaN energy, INT_MAX if none
    int it;
    int step_phase;
    __device__ __forceinline__ bool closed() const {
        const int s = status[blockIdx.z];
        return step_phase ? (s <= it) : (s < it);
    }
};

enum UpdateMode { UP_PROLOGUE = 1, UP_A = 2, UP_B = 3, UP_C = 4, UP_CN_MID = 5, UP_CN_LAST = 6 };
enum RowMode { R_FWD_PACK = 0, R_INV_PACK = 1, R_STAGE = 2, R_C2C_FWD = 3, R_C2C_INV = 4 };


template <typename T>
struct ColFwdArgs {
    Geo<T> g;
    Gate gate;
    const cx<T>* nx[2];  // y-transformed, normalised tendencies (UP_PROLOGUE: the field)
    cx<T>* U[2];         // state (read; written by UP_C / UP_PROLOGUE / UP_CN_LAST)
    cx<T>* UB[2];        // stage-B input (UP_A writes), stage-C input (UP_B writes, UP_C reads)
    cx<T>* KA[2];        // k_A (UP_A writes), k_A + 4 k_B (UP_B writes, UP_C reads)
    int pitch;
    int cn_first;        // UP_CN_MID: this sweep's midpoint is u_hat itself (first sweep)
    int nkx;             // modes kx < nkx of
    int nky;             // nky rows: row r is ky index r, or r + ky_skip from r = ky_skip_at on
    int ky_skip_at, ky_skip;
    // k3d::launch_cf2 (y-forward fused with the update): nkx columns, rows
    // |ky| >= kcut left as they are (kcut 0: every row), nx the x-forward outputs
    int kcut;
    const cx<T>* tw;
};

template <typename T>
struct RowArgs {
    Geo<T> g;
    Gate gate;
    const cx<T>* hin[4];  // half-row inputs
    cx<T>* hout[2];       // half-row outputs
    const T* rin[2];      // real rows in
    T* rout[2];           // real rows out
    const cx<T>* cin;     // full complex rows (c2c)
    cx<T>* cout;
    const cx<T>* tw;
};

template <typename T>
struct VgArgs {
    Geo<T> g;
    Gate gate;
    const cx<T>* w[7];  // u1 u2 (i k2 u1) (i k2 u2) (Lap u1) (Lap u2) (d2 Lap u2), y-inverses
    double* part;       // [8][gridDim.x]
    cx<T>* hout[2];     // non-null: also produce stage A's x-forward nonlinear term
    const cx<T>* tw;
};

template <typename T>
struct SpecArgs {
    Geo<T> g;
    Gate gate;
    const cx<T>* U[2];
    int allkmax;
    double* part_bins;  // [gridDim.x][allkmax+1][2]
    double* part_tot;   // [gridDim.x][4]
    // modes summed: kx < nkx in nky ky rows (row r is ky index r, or r + ky_skip
    // from r = ky_at on); 0, 0, 0, 0: every mode (the others are known zeros)
    int nkx, nky, ky_at, ky_skip;
};

struct FinArgs {
    Gate gate;
    int nblk_spec, nblk_vg, allkmax, ndims;
    double nu, npts;
    const double* part_bins;
    const double* part_tot;
    const double* part_vg;
    double* red;        // [2*allkmax + 12] reduced sums
    bastr_diag* diag;   // slot of this iteration
    double* spec;       // [3][allkmax] slot of this iteration (Es at +allkmax, Ed at +2*allkmax)
    int* status;
    // ensemble members (blockIdx.z): member m's partials, sums, diagnostics
    // and spectra start m*dstride diag slots / m*sstride doubles after member 0's
    int nmem;
    long dstride, sstride;
};

template <typename T>
struct ForceArgs {
    Geo<T> g;
    Gate gate;
    cx<T>* U[2];
    int forcemethod;          // 1, 2 or 5
    int forcek;
    double target_energy;
    const double* part_en;    // [nblk_en] partial sums of |u_hat|^2 over all modes (methods 1, 5)
    int nblk_en;
    double* factor_out;       // method 1: the factor for scale_kernel ([member])
    bastr_diag* diag;
    long dstride;             // diag slots between members
    const int* mforce;        // [member] 0: this member is not forced (null: every member is)
};

template <typename T>
struct EnergyArgs {
    Geo<T> g;
    Gate gate;
    const cx<T>* U[2];
    double* part_en;          // [gridDim.x]
};

template <typename T>
struct ScaleArgs {
    Geo<T> g;
    Gate gate;
    cx<T>* U[2];
    const double* factor;     // [member]
    const int* mforce;        // [member] 0: this member is not forced (null: every member is)
};

// launchers (k2d.hip); n is the transform length along the pass
template <typename T> int launch_update(int mode, const ColFwdArgs<T>& a, hipStream_t s);
template <typename T> int launch_row(int n, int mode, const RowArgs<T>& a, int nrows, hipStream_t s);
// velgrad_calculate2D's moments and stage A's x pass (row_stage_vg_kernel)
template <typename T> int launch_velgrad(int n, const VgArgs<T>& a, int nrows, hipStream_t s, int* nblk);
template <typename T> int launch_spectra(const SpecArgs<T>& a, int nblk, hipStream_t s);
int launch_finalize(const FinArgs& a, hipStream_t s);
template <typename T> int launch_force_band(const ForceArgs<T>& a, hipStream_t s);
template <typename T> int launch_energy(const EnergyArgs<T>& a, int nblk, hipStream_t s);
template <typename T> int launch_scale(const ScaleArgs<T>& a, hipStream_t s);
int col_block_width(int n, int prec);  // CB of the column pass for (n, precision)
int row_block_rows(int n, int prec);   // L of the row pass
int vg_block_rows(int n, int prec, int nd);  // L of the velocity-gradient row pass

}  // namespace k2d
