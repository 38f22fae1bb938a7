// 2-D step: precision-independent kernels and launch-shape queries.
#include <atomic>
#include <climits>
#include <math.h>
#include "k2d.hpp"
#include "k3d.hpp"

namespace k2d {

// Reduce the per-workgroup partials of one iteration (spectra_compute2D
// :991-1000, velgrad :1158-1175).  One workgroup per quantity, fixed
// summation order: thread t adds partials t, t+256, ..., then a fixed LDS
// tree.  Quantities: Es[1..K], Ed[1..K], 4 spectral totals, 7 velgrad sums,
// then the eta_min minimum.
// BT threads per quantity: 256 with an LDS tree where a quantity has many
// partials (3-D: one per row block), one wavefront with a shuffle butterfly
// where it has few (at most 512: no barriers, as a small grid's reduction is
// latency-bound; 256^2 FP64 0.0712 -> 0.0708 ms per iteration, while 2048
// partials per quantity at 2048^2 were faster on 256 threads)
template <int BT>
__global__ __launch_bounds__(BT) void reduce_partials_kernel(FinArgs a) {
    if (a.gate.closed()) return;
    __shared__ double sh[BT];
    const int K = a.allkmax;
    {  // ensemble member blockIdx.z: its partials and reduced sums
        const size_t m = blockIdx.z;
        a.part_bins += m * (size_t)a.nblk_spec * (K + 1) * 2;
        a.part_tot += m * (size_t)a.nblk_spec * 4;
        a.part_vg += m * (size_t)a.nblk_vg * 8;
        a.red += m * (2 * (size_t)K + 12);
    }
    const int q = blockIdx.x, tid = threadIdx.x;
    const double* src;
    int n;
    bool is_min = false;
    if (q < 2 * K) {
        const int b = q % K + 1, c = q / K;  // Es bins then Ed bins
        src = a.part_bins + ((size_t)b * 2 + c) * a.nblk_spec;
        n = a.nblk_spec;
    } else if (q < 2 * K + 4) {
        src = a.part_tot + (size_t)(q - 2 * K) * a.nblk_spec;
        n = a.nblk_spec;
    } else {
        const int v = q - 2 * K - 4;
        src = a.part_vg + (size_t)v * a.nblk_vg;
        n = a.nblk_vg;
        is_min = v == 7;
    }
    // each thread folds src[tid], src[tid + 256], ... in order; the loads of
    // eight of them are issued together (a 3-D velgrad moment has one partial
    // per row block: 65536 at 512^3)
    const double none = is_min ? (double)(2.0f * 3.14f) : 0.0;
    double acc = none;
    for (int w0 = tid; w0 < n; w0 += BT * 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = w0 + u * BT < n ? src[w0 + u * BT] : none;
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (w0 + u * BT < n) acc = is_min ? fmin(acc, v[u]) : acc + v[u];
    }
    if constexpr (BT == 64) {
        acc = is_min ? wave_min(acc) : wave_sum(acc);
        if (tid == 0) a.red[q] = acc;
    } else {
        sh[tid] = acc;
        __syncthreads();
        for (int o = BT / 2; o > 0; o >>= 1) {
            if (tid < o) sh[tid] = is_min ? fmin(sh[tid], sh[tid + o]) : sh[tid] + sh[tid + o];
            __syncthreads();
        }
        if (tid == 0) a.red[q] = sh[0];
    }
}

// Diagnostics of the iteration from the reduced sums; NaN stop
// (solution.F90:69-81).
__global__ void finalize_kernel(FinArgs a) {
    if (a.gate.closed()) return;
    const int K = a.allkmax;
    {  // ensemble member blockIdx.z
        const size_t m = blockIdx.z;
        a.red += m * (2 * (size_t)K + 12);
        a.diag += m * (size_t)a.dstride;
        a.spec += m * (size_t)a.sstride;
        a.status += m;
    }
    const double* r = a.red;
    for (int b = threadIdx.x; b < K; b += blockDim.x) {
        a.spec[K + b] = r[b];
        a.spec[2 * K + b] = r[K + b];
    }
    if (threadIdx.x == 0) {
        const double* t = r + 2 * K;
        const double* v = r + 2 * K + 4;
        const double N = a.npts, nu = a.nu;
        const double dim = a.ndims == 2 ? 2.0 : 3.0;
        const double urms = sqrt(v[4] / N), dudx2 = v[5] / N, eps = v[6] / N;
        const double ufluc = urms / sqrt(dim);
        const double tay = ufluc / sqrt(dudx2 / dim);
        bastr_diag* d = a.diag;
        d->Esspe = t[0];
        d->Edspe = t[1];
        d->k2Edspe = t[2];
        d->kMEdspe = t[3];
        d->umumtheta2 = v[0] / N;
        d->umumijji = v[1] / N;
        d->u2theta = v[2] / N;
        d->dissp = v[3] / N;
        d->eta_min = v[7];
        d->Taylength = tay;
        d->ReTay = ufluc * tay / nu;
        d->Kollength = sqrt(sqrt(nu * nu * nu / eps));
        if (isnan(t[0] + t[1])) {
            d->nan_flag = 1;
            if (*a.status > a.gate.it) *a.status = a.gate.it;
        }
    }
}

int launch_finalize(const FinArgs& a, hipStream_t s) {
    const dim3 grid(2 * a.allkmax + 12, 1, zdim(a.nmem));
    if (a.nblk_spec <= 512 && a.nblk_vg <= 512) hipLaunchKernelGGL(reduce_partials_kernel<64>, grid, dim3(64), 0, s, a);
    else hipLaunchKernelGGL(reduce_partials_kernel<256>, grid, dim3(256), 0, s, a);
    hipLaunchKernelGGL(finalize_kernel, dim3(1, 1, zdim(a.nmem)), dim3(256), 0, s, a);
    return 0;
}

// Column-block width that pads the half-spectrum pitch: the default plan's,
// or for FP64 columns of 64 to 256 points (4 points per thread, narrow blocks)
// the 8-point plan's, so the pitch stays a multiple of every plan's width.
template <typename T>
static int cb_of(int n) {
    switch (n) {
#define X(NN) case NN: return ColCfg<T, NN>::SHORT && ColCfg<T, NN, 8>::CB > ColCfg<T, NN>::CB ? ColCfg<T, NN, 8>::CB : ColCfg<T, NN>::CB;
        X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)
#undef X
    }
    return -1;
}
template <typename T>
static int l_of(int n) {
    switch (n) {
#define X(NN) case NN: return RowCfg<T, NN>::L;
        X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)
#undef X
    }
    return -1;
}

int col_block_width(int n, int prec) { return prec == BASTR_PREC_F32 ? cb_of<float>(n) : cb_of<double>(n); }
int row_block_rows(int n, int prec) { return prec == BASTR_PREC_F32 ? l_of<float>(n) : l_of<double>(n); }
template <typename T, int EE>
static int lvg_of(int n) {
    switch (n) {
#define X(NN) case NN: return VgCfg<T, NN, EE>::L;
        X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)
#undef X
    }
    return -1;
}
// rows per workgroup of the velocity-gradient x pass (2-D VgCfg, 3-D Vg3Cfg:
// 8 points per thread)
int vg_block_rows(int n, int prec, int nd) {
    if (nd == 3) return prec == BASTR_PREC_F32 ? lvg_of<float, 8>(n) : lvg_of<double, 8>(n);
    return prec == BASTR_PREC_F32 ? lvg_of<float, 0>(n) : lvg_of<double, 0>(n);
}

}  // namespace k2d

namespace k3d {

namespace {
std::atomic<int> g_colx_plan[2][13];  // [prec][log2 n]: points | width << 8; 0 = the default
std::atomic<uint64_t> g_plan_epoch{0};

int lg2(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

template <typename T, int N>
bool ok_of(int points, int width) {
    if (points == 16 && width == 1) return true;
    if (points == 8 && N < 256) return false;  // below 256 every plan is the 8-point default
    if (points == 8) return width == 1 ? ColCfg<T, N, 8, 1>::FITS : ColCfg<T, N, 8, 2>::FITS;
    return ColCfg<T, N, 16, 2>::FITS;
}

template <typename T, int N>
int cb_of_plan(int points, int width) {
    if (points == 8) return width == 2 ? ColCfg<T, N, 8, 2>::CB : ColCfg<T, N, 8, 1>::CB;
    return width == 2 ? ColCfg<T, N, 16, 2>::CB : ColCfg<T, N, 16, 1>::CB;
}

template <typename T>
int plan_cb_of(int n, int points, int width) {
    switch (n) {
#define X(NN) case NN: return cb_of_plan<T, NN>(points, width);
        X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)
#undef X
    }
    return 0;
}

template <typename T>
bool plan_ok_of(int n, int points, int width) {
    switch (n) {
#define X(NN) case NN: return ok_of<T, NN>(points, width);
        X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)
#undef X
    }
    return false;
}

int plan_of(int n, int prec) {
    const int l = lg2(n);
    if (l < 0 || l > 12 || prec < 0 || prec > 1) return 0;
    return g_colx_plan[prec][l].load(std::memory_order_relaxed);
}
}  // namespace

int colx_points(int n, int prec) { return (plan_of(n, prec) & 0xff) == 8 ? 8 : 16; }
int colx_width(int n, int prec) { return (plan_of(n, prec) >> 8) == 2 ? 2 : 1; }

void set_colx_plan(int n, int prec, int points, int width) {
    const int l = lg2(n);
    if (l < 0 || l > 12 || prec < 0 || prec > 1) return;
    const int v = (points == 16 && width == 1) ? 0 : (points | (width << 8));
    if (g_colx_plan[prec][l].exchange(v) != v) g_plan_epoch.fetch_add(1);
}

uint64_t colx_plan_epoch() { return g_plan_epoch.load(); }

int colx_plan_cb(int n, int prec, int points, int width) {
    return prec == 1 ? plan_cb_of<float>(n, points, width) : plan_cb_of<double>(n, points, width);
}

bool colx_plan_ok(int n, int prec, int points, int width) {
    if ((points != 8 && points != 16) || (width != 1 && width != 2)) return false;
    return prec == 1 ? plan_ok_of<float>(n, points, width) : plan_ok_of<double>(n, points, width);
}

}  // namespace k3d
