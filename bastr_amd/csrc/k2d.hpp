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

// Column-pass launch shape for (T, N1): CB adjacent columns per workgroup
// (CB*sizeof(complex) >= 64 B contiguous per row), E points per thread,
// CB*N1/E threads (<= 512 in FP64 so a thread may hold 256 VGPRs without
// spilling), CB*N1*sizeof(T) bytes of LDS (split exchange, <= 64 KiB).
template <typename T, int N>
struct ColCfg {
    static constexpr int CBMIN = 64 / (int)sizeof(cx<T>);
    static constexpr int LDS_CB = 65536 / (N * (int)sizeof(T));  // (+1/16 padding, <= 68 KiB)
    static constexpr int CBW = (256 / (N / 8)) > CBMIN ? (256 / (N / 8)) : CBMIN;
    static constexpr int CB2 = CBW > 16 ? 16 : CBW;
    static constexpr int CB = CB2 < LDS_CB ? CB2 : LDS_CB;
    static constexpr int MAXT = sizeof(T) == 8 ? 512 : 1024;
    static constexpr int E = (CB * (N / 8) <= MAXT) ? 8 : 16;
    static constexpr int P = N / E;
    static constexpr int THREADS = CB * P;
    static_assert(CB >= 1 && THREADS <= MAXT && N >= E, "column config");
};

// Row-pass launch shape for (T, N0): L rows per workgroup.
template <typename T, int N>
struct RowCfg {
    static constexpr int E = 8;
    static constexpr int P = N / E;
    static constexpr int L0 = P >= 256 ? 1 : 256 / P;
    static constexpr int LMAX = 65536 / (N * (int)sizeof(cx<T>));  // (+1/16 padding)
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
    double nu, dt, inv_n;
};

// Iteration gate: kernels of the diagnostic phase of iteration `it` run while
// no NaN was found in an earlier iteration; kernels of the step phase also
// stop for the iteration whose diagnostics were NaN (solution.F90:81).
struct Gate {
    const int* status;  // first iteration index with NaN energy, INT_MAX if none
    int it;
    int step_phase;
    __device__ __forceinline__ bool closed() const {
        const int s = *status;
        return step_phase ? (s <= it) : (s < it);
    }
};

enum ColInvMode { CI_PLAIN = 0, CI_STAGE = 1, CI_STAGE_VG = 2 };
enum ColFwdMode { CF_PLAIN = 0, CF_PROLOGUE = 1, CF_A = 2, CF_B = 3, CF_C = 4, CF_CN_MID = 5, CF_CN_LAST = 6 };
enum RowMode { R_FWD_PACK = 0, R_INV_PACK = 1, R_STAGE = 2, R_C2C_FWD = 3, R_C2C_INV = 4 };

template <typename T>
struct ColInvArgs {
    Geo<T> g;
    Gate gate;
    const cx<T>* in[2];
    cx<T>* out[7];
    int nfield;  // CI_PLAIN: number of fields (in[f] -> out[f])
    int pitch;   // row pitch of in/out (hp, or n0 for full c2c arrays)
    const cx<T>* tw;
};

template <typename T>
struct ColFwdArgs {
    Geo<T> g;
    Gate gate;
    const cx<T>* nx[2];  // row-transformed inputs
    cx<T>* U[2];         // state (read; written by CF_C / CF_PROLOGUE)
    cx<T>* UB[2];        // stage-B input (CF_A writes), stage-C input (CF_B writes, CF_C reads)
    cx<T>* KA[2];        // k_A (CF_A writes), k_A + 4 k_B (CF_B writes, CF_C reads)
    cx<T>* out[2];       // CF_PLAIN output
    int nfield;
    int pitch;
    int cn_first;        // CF_CN_MID: this sweep's midpoint is u_hat itself (first sweep)
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
    const cx<T>* w[7];  // A1 A2 B1 B2 C1 C2 D
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
    double* factor_out;       // method 1: the factor for scale_kernel
    bastr_diag* diag;
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
    const double* factor;
};

// launchers (k2d.hip); n is the transform length along the pass
template <typename T> int launch_col_inv(int n, int mode, const ColInvArgs<T>& a, int ncols, hipStream_t s);
template <typename T> int launch_col_fwd(int n, int mode, const ColFwdArgs<T>& a, int ncols, hipStream_t s);
template <typename T> int launch_row(int n, int mode, const RowArgs<T>& a, int nrows, hipStream_t s);
template <typename T> int launch_velgrad(int n, const VgArgs<T>& a, int nrows, hipStream_t s, int* nblk);
template <typename T> int launch_spectra(const SpecArgs<T>& a, int nblk, hipStream_t s);
int launch_finalize(const FinArgs& a, hipStream_t s);
template <typename T> int launch_force_band(const ForceArgs<T>& a, hipStream_t s);
template <typename T> int launch_energy(const EnergyArgs<T>& a, int nblk, hipStream_t s);
template <typename T> int launch_scale(const ScaleArgs<T>& a, hipStream_t s);
int col_block_width(int n, int prec);  // CB of the column pass for (n, precision)
int row_block_rows(int n, int prec);   // L of the row pass
int velgrad_blocks(int n0, int n1, int prec);

}  // namespace k2d
