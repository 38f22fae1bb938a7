// pp.hip -- post-processing kernels (Bastrpp.F90:174-599): the products of
// transfer2d/3d in the x pass and their shell sums.
#include <climits>

#include "k2d_impl.hpp"
#include "k3d_impl.hpp"
#include "pp.hpp"

namespace pp {

using k2d::ksigned;
using k2d::pack_at;
using k2d::unpack_store;
using k2d::RowCfg;

// 2-D: reference transfer2d (Bastrpp.F90:174-297), physical-space part:
//   u1tA = -u1 d1u1 - u2 d2u1, u2tA = -u1 d1u2 - u2 d2u2,
//   uitC = ui * (d1u1 + d2u2), u1x1 = u1 u1, u1x2 = u1 u2, u2x2 = u2 u2.
template <typename T, int N>
__global__ __launch_bounds__((RowCfg<T, N>::THREADS)) void row2_kernel(RowArgs<T> a) {
    using C = RowCfg<T, N>;
    constexpr int E = C::E, P = C::P, LR = C::L;
    if (a.gate.closed()) return;
    __shared__ cx<T> lds[LR * lds_line_cx(N)];
    const int tid = threadIdx.x;
    const int l = tid / P, j = tid % P;
    const int nrows = a.g.n1;
    const int row0 = blockIdx.x * LR + l;
    const bool valid = row0 < nrows;
    const int row = valid ? row0 : nrows - 1;
    const LdsLine<T, false, (P <= 64)> Ls{lds + l * lds_line_cx(N), 1};
    const size_t ro = (size_t)row * a.g.hp;
    const cx<T>* A1 = a.hin[0] + ro;
    const cx<T>* A2 = a.hin[1] + ro;
    const cx<T>* B1 = a.hin[2] + ro;
    const cx<T>* B2 = a.hin[3] + ro;
    cx<T> z[E], w[E], t[E];
    T th[E];
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const int m = j + k * P;
        z[k] = pack_at<T, N>(A1, A2, m);
        const T kxs = (m == N / 2) ? (T)0 : (T)ksigned(m, N);
        w[k] = cscale(cmuli(z[k]), kxs);
    }
    line_fft<T, N, E, true, false>(z, j, Ls, a.tw);  // (u1, u2)
    line_fft<T, N, E, true, false>(w, j, Ls, a.tw);  // (d1 u1, d1 u2)
#pragma unroll
    for (int k = 0; k < E; ++k) {
        t[k] = {-(z[k].x * w[k].x), -(z[k].x * w[k].y)};
        th[k] = w[k].x;
    }
#pragma unroll
    for (int k = 0; k < E; ++k) w[k] = pack_at<T, N>(B1, B2, j + k * P);
    line_fft<T, N, E, true, false>(w, j, Ls, a.tw);  // (d2 u1, d2 u2)
#pragma unroll
    for (int k = 0; k < E; ++k) {
        t[k].x = t[k].x - z[k].y * w[k].x;
        t[k].y = t[k].y - z[k].y * w[k].y;
        th[k] = th[k] + w[k].y;
    }
    line_fft<T, N, E, false, false>(t, j, Ls, a.tw);
    unpack_store<T, N, E>(t, j, Ls, a.hout[0] + ro, a.hout[1] + ro, valid);
#pragma unroll
    for (int k = 0; k < E; ++k) t[k] = {z[k].x * th[k], z[k].y * th[k]};
    line_fft<T, N, E, false, false>(t, j, Ls, a.tw);
    unpack_store<T, N, E>(t, j, Ls, a.hout[2] + ro, a.hout[3] + ro, valid);
#pragma unroll
    for (int k = 0; k < E; ++k) t[k] = {z[k].x * z[k].x, z[k].x * z[k].y};
    line_fft<T, N, E, false, false>(t, j, Ls, a.tw);
    unpack_store<T, N, E>(t, j, Ls, a.hout[4] + ro, a.hout[5] + ro, valid);
#pragma unroll
    for (int k = 0; k < E; ++k) t[k] = {z[k].y * z[k].y, (T)0};
    line_fft<T, N, E, false, false>(t, j, Ls, a.tw);
    unpack_store<T, N, E>(t, j, Ls, a.hout[6] + ro, a.hout[7] + ro, valid);
}

// 3-D: transfer3d (Bastrpp.F90:402-599), physical-space part, accumulated in
// the reference's term order (j = 1, 2, 3).
template <typename T, int N>
__global__ __launch_bounds__((RowCfg<T, N>::THREADS)) void row3_kernel(RowArgs<T> a) {
    using C = RowCfg<T, N>;
    constexpr int E = C::E, P = C::P, LR = C::L;
    if (a.gate.closed()) return;
    __shared__ cx<T> lds[LR * lds_line_cx(N)];
    const int tid = threadIdx.x;
    const int l = tid / P, j = tid % P;
    const int nrows = a.g.n1 * a.n2;
    const int row0 = blockIdx.x * LR + l;
    const bool valid = row0 < nrows;
    const int row = valid ? row0 : nrows - 1;
    const LdsLine<T, false, (P <= 64)> Ls{lds + l * lds_line_cx(N), 1};
    const size_t ro = (size_t)row * a.g.hp;
    const cx<T>* Y[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) Y[q] = a.hin[q] + ro;
    cx<T> z[E], w[E], t[E];
    T u3[E], tA1[E], tA2[E], tA3[E], th[E];
#pragma unroll
    for (int k = 0; k < E; ++k) z[k] = pack_at<T, N>(Y[0], Y[1], j + k * P);
    line_fft<T, N, E, true, false>(z, j, Ls, a.tw);  // (u1, u2)
#pragma unroll
    for (int k = 0; k < E; ++k) w[k] = k3d::pack_a_ika<T, N>(Y[2], j + k * P);
    line_fft<T, N, E, true, false>(w, j, Ls, a.tw);  // (u3, d1 u3)
#pragma unroll
    for (int k = 0; k < E; ++k) {
        u3[k] = w[k].x;
        tA3[k] = -(z[k].x * w[k].y);
    }
    // products u_i u_j: S11 S12 | S13 S22 | S23 S33
#pragma unroll
    for (int k = 0; k < E; ++k) t[k] = {z[k].x * z[k].x, z[k].x * z[k].y};
    line_fft<T, N, E, false, false>(t, j, Ls, a.tw);
    unpack_store<T, N, E>(t, j, Ls, a.hout[0] + ro, a.hout[1] + ro, valid);
#pragma unroll
    for (int k = 0; k < E; ++k) t[k] = {z[k].x * u3[k], z[k].y * z[k].y};
    line_fft<T, N, E, false, false>(t, j, Ls, a.tw);
    unpack_store<T, N, E>(t, j, Ls, a.hout[2] + ro, a.hout[3] + ro, valid);
#pragma unroll
    for (int k = 0; k < E; ++k) t[k] = {z[k].y * u3[k], u3[k] * u3[k]};
    line_fft<T, N, E, false, false>(t, j, Ls, a.tw);
    unpack_store<T, N, E>(t, j, Ls, a.hout[4] + ro, a.hout[5] + ro, valid);
    // gradient terms
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const int m = j + k * P;
        const T kxs = (m == N / 2) ? (T)0 : (T)ksigned(m, N);
        w[k] = cscale(cmuli(pack_at<T, N>(Y[0], Y[1], m)), kxs);
    }
    line_fft<T, N, E, true, false>(w, j, Ls, a.tw);  // (d1 u1, d1 u2)
#pragma unroll
    for (int k = 0; k < E; ++k) {
        tA1[k] = -(z[k].x * w[k].x);
        tA2[k] = -(z[k].x * w[k].y);
        th[k] = w[k].x;
    }
#pragma unroll
    for (int k = 0; k < E; ++k) w[k] = pack_at<T, N>(Y[3], Y[4], j + k * P);
    line_fft<T, N, E, true, false>(w, j, Ls, a.tw);  // (d2 u1, d2 u2)
#pragma unroll
    for (int k = 0; k < E; ++k) {
        tA1[k] = tA1[k] - z[k].y * w[k].x;
        tA2[k] = tA2[k] - z[k].y * w[k].y;
        th[k] = th[k] + w[k].y;
    }
#pragma unroll
    for (int k = 0; k < E; ++k) w[k] = pack_at<T, N>(Y[6], Y[7], j + k * P);
    line_fft<T, N, E, true, false>(w, j, Ls, a.tw);  // (d3 u1, d3 u2)
#pragma unroll
    for (int k = 0; k < E; ++k) {
        tA1[k] = tA1[k] - u3[k] * w[k].x;
        tA2[k] = tA2[k] - u3[k] * w[k].y;
    }
#pragma unroll
    for (int k = 0; k < E; ++k) w[k] = pack_at<T, N>(Y[5], Y[8], j + k * P);
    line_fft<T, N, E, true, false>(w, j, Ls, a.tw);  // (d2 u3, d3 u3)
#pragma unroll
    for (int k = 0; k < E; ++k) {
        tA3[k] = tA3[k] - z[k].y * w[k].x;
        tA3[k] = tA3[k] - u3[k] * w[k].y;
        th[k] = th[k] + w[k].y;
    }
#pragma unroll
    for (int k = 0; k < E; ++k) t[k] = {tA1[k], tA2[k]};
    line_fft<T, N, E, false, false>(t, j, Ls, a.tw);
    unpack_store<T, N, E>(t, j, Ls, a.hout[6] + ro, a.hout[7] + ro, valid);
#pragma unroll
    for (int k = 0; k < E; ++k) t[k] = {tA3[k], z[k].x * th[k]};
    line_fft<T, N, E, false, false>(t, j, Ls, a.tw);
    unpack_store<T, N, E>(t, j, Ls, a.hout[8] + ro, a.hout[9] + ro, valid);
#pragma unroll
    for (int k = 0; k < E; ++k) t[k] = {z[k].y * th[k], u3[k] * th[k]};
    line_fft<T, N, E, false, false>(t, j, Ls, a.tw);
    unpack_store<T, N, E>(t, j, Ls, a.hout[10] + ro, a.hout[11] + ro, valid);
}

// Shell sums of transfer2d/3d over the half spectrum.  A stored mode with
// 0 < kx < ia/2 stands for the reference's pair (k, -k) of the full spectrum:
// weight 2, except that the reference gives the partner of a mode on a
// Nyquist plane of axis j (j = 2, 3) the same +N/2 wavenumber there, so the
// k_j terms of T1 cancel within the pair.  Modes at kx = 0 and kx = ia/2 are
// stored once each and are taken as they are (weight 1).
template <typename T>
__global__ __launch_bounds__(64) void bins_kernel(BinArgs<T> a) {
    if (a.gate.closed()) return;
    extern __shared__ double sh[];
    const int K = a.allkmax, nb = K + 1;
    const int lane = threadIdx.x;
    double* binsA = sh;           // [nb][2]: T, T1
    double* binsB = sh + 2 * nb;  // [nb][2]: T2, 0
    for (int i = lane; i < 4 * nb; i += 64) sh[i] = 0;
    __syncthreads();
    const int nd = a.nd, n0 = a.n0, n1 = a.n1, n2 = nd == 3 ? a.n2 : 1, nh = a.nh;
    const size_t hp = (size_t)a.hp;
    const int nbx = (nh + 63) / 64;
    const long ntask = (long)n1 * n2 * nbx;
    for (long t = blockIdx.x; t < ntask; t += gridDim.x) {
        const long r = t / nbx;
        const int kx = (int)(t % nbx) * 64 + lane;
        double tt = 0, t1 = 0, t2 = 0;
        int s = INT_MAX;
        if (kx < nh) {
            const int r1 = (int)(r % n1), r2 = (int)(r / n1);
            const int ky = ksigned(r1, n1), kz = nd == 3 ? ksigned(r2, n2) : 0;
            const bool pair = !(kx == 0 || kx == n0 / 2);
            const double w = pair ? 2.0 : 1.0;
            const double kc[3] = {(double)kx, (pair && 2 * r1 == n1) ? 0.0 : (double)ky,
                                  (pair && nd == 3 && 2 * r2 == n2) ? 0.0 : (double)kz};
            const size_t o = (size_t)r * hp + kx;
            double sa = 0, sc = 0, zi = 0;
            for (int i = 0; i < nd; ++i) {
                const cx<T> ui = a.U[i][o];
                const cx<T> fa = a.F[nd == 2 ? i : 6 + i][o];
                const cx<T> fc = a.F[nd == 2 ? 2 + i : 9 + i][o];
                sa += (double)ui.x * fa.x + (double)ui.y * fa.y;  // Re(conj(u_i) tA_i)
                sc += (double)ui.x * fc.x + (double)ui.y * fc.y;  // Re(conj(u_i) tC_i)
                for (int jj = 0; jj < nd; ++jj) {
                    // S_ij is stored once per pair i <= j: 2-D S11 S12 S22 at 4..6,
                    // 3-D S11 S12 S13 S22 S23 S33 at 0..5
                    const int lo = i < jj ? i : jj, hi = i < jj ? jj : i;
                    const int si = nd == 2 ? 4 + lo + hi : (lo == 0 ? hi : (lo == 1 ? 2 + hi : 5));
                    const cx<T> sij = a.F[si][o];
                    zi += kc[jj] * ((double)sij.y * ui.x - (double)sij.x * ui.y);  // Im(k_j S_ij conj(u_i))
                }
            }
            tt = w * sa;
            t1 = w * zi;
            t2 = w * sc;
            const double kk = sqrt((double)kx * kx + (double)ky * ky + (double)kz * kz);
            const int sh_ = (int)(kk + 0.5);
            if (sh_ <= K) s = sh_;
        }
        bin_runs(binsA, s, tt, t1, lane);
        bin_runs(binsB, s, t2, 0.0, lane);
    }
    __syncthreads();
    // partials [shell][quantity][block]: the host sums over blocks in block order
    for (int b = lane; b < nb; b += 64) {
        a.part[((size_t)b * 3 + 0) * gridDim.x + blockIdx.x] = binsA[2 * b];
        a.part[((size_t)b * 3 + 1) * gridDim.x + blockIdx.x] = binsA[2 * b + 1];
        a.part[((size_t)b * 3 + 2) * gridDim.x + blockIdx.x] = binsB[2 * b];
    }
}

#define PP2_SIZES(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)
#define PP3_SIZES(X) X(16) X(32) X(64) X(128) X(256) X(512)

template <typename T>
int launch_row2(int n0, const RowArgs<T>& a, hipStream_t s) {
    switch (n0) {
#define X(NN)                                                                                   \
    case NN: {                                                                                  \
        using C = RowCfg<T, NN>;                                                                \
        hipLaunchKernelGGL((row2_kernel<T, NN>), dim3((a.g.n1 + C::L - 1) / C::L), dim3(C::THREADS), 0, s, a); \
        return 0;                                                                               \
    }
        PP2_SIZES(X)
#undef X
    }
    return -1;
}
template <typename T>
int launch_row3(int n0, const RowArgs<T>& a, hipStream_t s) {
    switch (n0) {
#define X(NN)                                                                                   \
    case NN: {                                                                                  \
        using C = RowCfg<T, NN>;                                                                \
        hipLaunchKernelGGL((row3_kernel<T, NN>), dim3((a.g.n1 * a.n2 + C::L - 1) / C::L), dim3(C::THREADS), 0, s, a); \
        return 0;                                                                               \
    }
        PP3_SIZES(X)
#undef X
    }
    return -1;
}
template <typename T>
int launch_bins(const BinArgs<T>& a, int nblk, hipStream_t s) {
    const size_t shm = sizeof(double) * 4 * (size_t)(a.allkmax + 1);
    hipLaunchKernelGGL((bins_kernel<T>), dim3(nblk), dim3(64), shm, s, a);
    return 0;
}

template int launch_row2<float>(int, const RowArgs<float>&, hipStream_t);
template int launch_row2<double>(int, const RowArgs<double>&, hipStream_t);
template int launch_row3<float>(int, const RowArgs<float>&, hipStream_t);
template int launch_row3<double>(int, const RowArgs<double>&, hipStream_t);
template int launch_bins<float>(const BinArgs<float>&, int, hipStream_t);
template int launch_bins<double>(const BinArgs<double>&, int, hipStream_t);

}  // namespace pp
