// k3d.hpp -- 3-D pseudo-spectral step kernels (gfx950).
//
// Layout in HBM (x = index i of the reference, fastest):
//   physical real field  u[z][y][x]              (Fortran u(i,j,k))
//   half spectrum        S[kz][ky][kx], kx = 0..n0/2, row pitch hp >= n0/2+1
// Passes:
//   x  row kernels over the n1*n2 rows (pack two real fields per complex FFT)
//   y  column passes inside each kz plane (pitch hp, n2 planes)
//   z  column passes whose columns are all (ky, kx) of a plane (pitch hp*n1)
#pragma once
#include "fft_block.hpp"
#include "k2d.hpp"

namespace k3d {

using k2d::ColCfg;
using k2d::Gate;
using k2d::RowCfg;

// Multiplier of a column-pass term, a function of the signed wavenumber k of
// the position along the pass (ko = k with the Nyquist entry zeroed for odd
// powers, the Hermitian part of a c2c real-part inverse).
// Multipliers of a column-pass term, k the pass wavenumber; the LAP forms also
// use the column's own wavenumber kc (a z pass, whose columns sit at fixed ky):
//   M_LAP = -(kc^2 + k^2), M_LAP_IKC = -i (kc^2 + k^2) kc, M_LAP_IK = -i (kc^2 + k^2) k
//   (kc^2 + kx^2 in place of kc^2 where kx_mod is set).
// Odd powers of a wavenumber vanish at that axis's Nyquist index.
enum Mult { M_ONE = 0, M_IK = 1, M_MK2 = 2, M_MIK3 = 3, M_LAP = 4, M_LAP_IKC = 5, M_LAP_IK = 6 };

struct ColJob {
    int nterm;
    int in[4];
    int mult[4];
    int out;
};

template <typename T>
struct ColXArgs {
    Gate gate;
    const cx<T>* in[10];
    cx<T>* out[13];
    ColJob job[13];
    int njob;
    int pitch;           // elements between consecutive positions along the pass
    int ncols;           // columns per plane (multiple of CB)
    long plane_stride;   // elements between planes
    int nplanes;
    double scale;        // applied to every output (1: none)
    int kc_div, kc_n;    // LAP multipliers: kc = ksigned(column offset / kc_div, kc_n) (kc_div 0: unused)
    int kx_mod;          // > 0: LAP multipliers also take kx^2, kx = column offset mod kx_mod (a 3-D
                         // z pass over (kx, ky) columns: the full Laplacian)
    int kcut;            // > 0: entries with |k| >= kcut along the pass are zero: an inverse
                         // pass does not read them, a forward pass does not write them
    int pskip_at, pskip; // plane p >= pskip_at is plane p + pskip (0, 0: planes in order)
    const cx<T>* tw;
    long mstride;        // 2-D ensembles: elements between members (blockIdx.z)
    int nmem;            // launch width along z (0 or 1: one member)
};

template <typename T>
struct Geo3 {
    int n0, n1, n2;
    int nh, hp;
    int cut2;
    int lproject;
    int kxlim;       // x passes: half-row entries kx >= kxlim are zero (not read, not written)
    double nu, dt, inv_n;
};

enum Cf3Mode { C3_PROLOGUE = 1, C3_A = 2, C3_B = 3, C3_C = 4, C3_CN_MID = 5, C3_CN_LAST = 6 };

// z-forward pass + spectral update of the three components
template <typename T>
struct Cf3Args {
    Geo3<T> g;
    Gate gate;
    const cx<T>* nx[3];   // y-transformed inputs
    cx<T>* U[3];
    cx<T>* UB[3];
    cx<T>* KA[3];
    int cn_first;         // C3_CN_MID: this sweep's midpoint is u_hat itself (first sweep)
    // columns: nkx consecutive offsets (kx, or the flat ky*hp + kx when nky = 1)
    // from the start of each of nky ky rows (row r is ky index r, or r + ky_skip
    // from r = ky_skip_at on)
    int nkx, nky, ky_skip_at, ky_skip;
    int kzcut;            // > 0: modes with |kz| >= kzcut are left as they are (the
                          // dealiased state's zeros; not read, not written)
    const cx<T>* tw;
};

enum Row3Mode { R3_FWD_PACK = 0, R3_INV_PACK = 1, R3_STAGE = 2 };

template <typename T>
struct Row3Args {
    Geo3<T> g;
    Gate gate;
    const cx<T>* hin[9];  // R3_STAGE: Y1 Y2 Y3, Y2_1 Y2_2 Y2_3, Y3_1 Y3_2 Y3_3 ; R3_INV_PACK: A1 A2 A3
    cx<T>* hout[3];
    const T* rin[3];
    T* rout[3];
    const cx<T>* tw;
};

template <typename T>
struct Vg3Args {
    Geo3<T> g;
    Gate gate;
    const cx<T>* w[13];   // Y_c(3), Y2_c(3), Y3_c(3), L_c(3), D
    double* part;         // [8][gridDim.x]
    const cx<T>* tw;
};

template <typename T>
struct Spec3Args {
    Geo3<T> g;
    Gate gate;
    const cx<T>* U[3];
    int allkmax;
    double* part_bins;    // [(K+1)*2][gridDim.x]
    double* part_tot;     // [4][gridDim.x]
    // modes summed: kx < nkx, nky ky and nkz kz indices (index r is r, or
    // r + skip from r = at on); 0s: every mode (the others are known zeros)
    int nkx, nky, ky_at, ky_skip, nkz, kz_at, kz_skip;
};

template <typename T>
struct Force3Args {
    Geo3<T> g;
    Gate gate;
    cx<T>* U[3];
    int forcemethod;
    int forcek;
    double target_energy;
    const double* part_en;
    int nblk_en;
    double* factor_out;
    bastr_diag* diag;
};

template <typename T>
struct Energy3Args {
    Geo3<T> g;
    Gate gate;
    const cx<T>* U[3];
    double* part_en;
};

template <typename T>
struct Scale3Args {
    Geo3<T> g;
    Gate gate;
    cx<T>* U[3];
    const double* factor;
};

template <typename T> int launch_colx(int n, bool inverse, const ColXArgs<T>& a, hipStream_t s);

// Plan of the single-term column passes of length n (prec 0 FP64, 1 FP32):
// 16 or 8 points per thread (radix-16 or radix-8 passes) on blocks of the
// default width or twice it (a wider tile, longer row segments).  A
// process-wide table: 16 points, default width unless set (by the autotuner,
// bastr_autotune); colx_plan_epoch counts the changes (captured step graphs
// hold the kernels of the old plan).
int colx_points(int n, int prec);
int colx_width(int n, int prec);  // 1: the default block width, 2: twice it
void set_colx_plan(int n, int prec, int points, int width);
uint64_t colx_plan_epoch();
// whether the plan (points per thread, width multiplier) exists for length n
bool colx_plan_ok(int n, int prec, int points, int width);
int colx_plan_cb(int n, int prec, int points, int width);  // its block width (columns)
template <typename T> int launch_cf3(int n, int mode, const Cf3Args<T>& a, hipStream_t s);
template <typename T> int launch_cf2(int n, int mode, const k2d::ColFwdArgs<T>& a, hipStream_t s);
// whether cf2 keeps the column passes' block width at n rows (else the
// y-forward and the update run as two kernels)
template <typename T> bool cf2_full_rows(int n);
template <typename T> int launch_row3(int n, int mode, const Row3Args<T>& a, int nrows, hipStream_t s);
template <typename T> int launch_vg3(int n, const Vg3Args<T>& a, int nrows, hipStream_t s, int nblk);
template <typename T> int launch_spec3(const Spec3Args<T>& a, int nblk, hipStream_t s);
template <typename T> int launch_force3(const Force3Args<T>& a, hipStream_t s);
template <typename T> int launch_energy3(const Energy3Args<T>& a, int nblk, hipStream_t s);
template <typename T> int launch_scale3(const Scale3Args<T>& a, hipStream_t s);

}  // namespace k3d
