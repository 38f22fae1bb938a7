// k2d_impl.hpp -- kernel bodies of the 2-D step (included once per precision).
//
// Reference correspondence (src/solution.F90, src/tool.F90):
//   k3d::colx_kernel (inverse)  spectral derivatives u_hat * i k_y (compute_ut2D :825-830),
//                               y half of the inverse transforms (:832-839); the
//                               same job-table column pass serves 2-D and 3-D
//   row_kernel<R_STAGE>         x half of those inverse transforms, the nonlinear
//                               term -u_j d_j u_i (:841-846), x half of fft2d (:850-851)
//   k3d::colx_kernel (forward)  y half of fft2d, normalisation (fftwlink.F90:334)
//   update_kernel<UP_A/B/C>     + nu*Lap(u) in spectral space, dealiasing (:853-854),
//                               projection (:856-858), Kutta RK3 combination (:183-200)
//   spectra_kernel              spectra_compute2D (:930-1016)
//   colx (velgrad jobs) +
//   row_stage_vg_kernel         velgrad_calculate2D (:1110-1186), all its moments, fused
//                               with stage A's x pass
//   force_band_kernel           forcing2D methods 1, 2, 5 (:303-540)
#pragma once
#include <climits>
#include "k2d.hpp"

namespace k2d {

__device__ __forceinline__ int ksigned(int m, int n) { return m <= n / 2 ? m : m - n; }

// Ensemble members: blockIdx.z selects the member; these move an argument
// set's field pointers (and per-block partial sums, per-member scalars) from
// member 0's to that member's.  Member 0 (every single-field launch) is
// unchanged.
template <typename T>
__device__ __forceinline__ void at_member(ColFwdArgs<T>& a) {
    const size_t o = (size_t)blockIdx.z * (size_t)a.g.hstride;
    for (int c = 0; c < 2; ++c) {
        a.nx[c] += o;
        a.U[c] += o;
        a.UB[c] += o;
        a.KA[c] += o;
    }
}
template <typename T>
__device__ __forceinline__ void at_member(RowArgs<T>& a) {
    const size_t o = (size_t)blockIdx.z * (size_t)a.g.hstride, r = (size_t)blockIdx.z * (size_t)a.g.rstride;
    for (int q = 0; q < 4; ++q) a.hin[q] += o;
    for (int c = 0; c < 2; ++c) {
        a.hout[c] += o;
        a.rin[c] += r;
        a.rout[c] += r;
    }
}
template <typename T>
__device__ __forceinline__ void at_member(VgArgs<T>& a) {
    const size_t o = (size_t)blockIdx.z * (size_t)a.g.hstride;
    for (int q = 0; q < 7; ++q) a.w[q] += o;
    a.hout[0] += o;
    a.hout[1] += o;
    a.part += (size_t)blockIdx.z * 8 * gridDim.x;
}
template <typename T>
__device__ __forceinline__ void at_member(SpecArgs<T>& a) {
    const size_t o = (size_t)blockIdx.z * (size_t)a.g.hstride;
    a.U[0] += o;
    a.U[1] += o;
    a.part_bins += (size_t)blockIdx.z * (size_t)(a.allkmax + 1) * 2 * gridDim.x;
    a.part_tot += (size_t)blockIdx.z * 4 * gridDim.x;
}
template <typename T>
__device__ __forceinline__ void at_member(EnergyArgs<T>& a) {
    const size_t o = (size_t)blockIdx.z * (size_t)a.g.hstride;
    a.U[0] += o;
    a.U[1] += o;
    a.part_en += (size_t)blockIdx.z * gridDim.x;
}
template <typename T>
__device__ __forceinline__ void at_member(ScaleArgs<T>& a) {
    const size_t o = (size_t)blockIdx.z * (size_t)a.g.hstride;
    a.U[0] += o;
    a.U[1] += o;
    a.factor += blockIdx.z;
}
template <typename T>
__device__ __forceinline__ void at_member(ForceArgs<T>& a) {
    const size_t o = (size_t)blockIdx.z * (size_t)a.g.hstride;
    a.U[0] += o;
    a.U[1] += o;
    a.part_en += (size_t)blockIdx.z * a.nblk_en;
    a.factor_out += blockIdx.z;
    a.diag += (size_t)blockIdx.z * a.dstride;
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8), each XCD with its own L2.  Column blocks share 128-byte lines with
// their neighbours, so map consecutive column blocks to the same XCD:
// XCD x handles the contiguous range [x*nb/8, (x+1)*nb/8).
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int x = b & 7, q = b >> 3;
    const int lo = (int)(((long)nb * x) >> 3);
    const int hi = (int)(((long)nb * (x + 1)) >> 3);
    const int v = lo + q;
    return (v < hi) ? v : -1;  // -1: this slot has no block (only when nb % 8 != 0)
}
__device__ __forceinline__ int xcd_grid(int nb) { return ((nb + 7) >> 3) << 3; }

// pack_at from the half-row values a = A[h], b = B[h], h = min(m, N - m)
template <typename T, int N>
__device__ __forceinline__ cx<T> pack_vals(cx<T> a, cx<T> b, int m) {
    if (m == 0 || m == N / 2) return {a.x, b.x};
    if (m < N / 2) return {a.x - b.y, a.y + b.x};
    return {a.x + b.y, b.x - a.y};
}

// Half-row entry X[h], zero from h = lim on.  The load is issued for every
// lane (a lane past lim reads the DC entry, which its row's other lanes read
// too) and the value selected, so a row's loads stay one branch-free group.
template <typename T>
__device__ __forceinline__ cx<T> half_at(const cx<T>* __restrict__ X, int h, int lim) {
    const bool in = h < lim;
    const cx<T> v = X[in ? h : 0];
    return in ? v : cx<T>{(T)0, (T)0};
}

// z = a + i b for packed position m of a length-N row, from half rows A, B
// (kx = 0..N/2).  DC and Nyquist take real parts (the real part of a c2c
// inverse equals the c2r inverse of the Hermitian part).  Entries kx >= lim
// are zero.
template <typename T, int N>
__device__ __forceinline__ cx<T> pack_at(const cx<T>* __restrict__ A, const cx<T>* __restrict__ B, int m,
                                         int lim = N) {
    const int h = m <= N / 2 ? m : N - m;
    return pack_vals<T, N>(half_at(A, h, lim), half_at(B, h, lim), m);
}

// Loads through buffer resources: the hardware returns zero for an offset at
// or past the resource's range, with no memory access.
constexpr int kBufWord3 = 0x00020000;          // gfx9 raw buffer, no format conversion
constexpr unsigned kBufRange = 0x80000000u;    // whole-array resources (every array < 2 GiB)

// off: the lane's byte offset (the whole offset: measured, a wave-uniform part
// passed as the instructions' SGPR offset gave results that differed from run
// to run while other streams' kernels ran beside them).  An off at or past
// the range (kBufRange) reads zero.
template <typename T>
__device__ __forceinline__ cx<T> buf_load(__amdgpu_buffer_rsrc_t r, unsigned off) {
    cx<T> v;
    if constexpr (sizeof(cx<T>) == 16) {
        typedef unsigned int u4 __attribute__((ext_vector_type(4)));
        const u4 w = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
        __builtin_memcpy(&v, &w, 16);
    } else {
        typedef unsigned int u2 __attribute__((ext_vector_type(2)));
        const u2 w = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0);
        __builtin_memcpy(&v, &w, 8);
    }
    return v;
}

// Store through a buffer resource: an off at or past the range is dropped.
template <typename T>
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, unsigned off, cx<T> v) {
    if constexpr (sizeof(cx<T>) == 16) {
        typedef unsigned int u4 __attribute__((ext_vector_type(4)));
        u4 w;
        __builtin_memcpy(&w, &v, 16);
        __builtin_amdgcn_raw_buffer_store_b128(w, r, (int)off, 0, 0);
    } else {
        typedef unsigned int u2 __attribute__((ext_vector_type(2)));
        u2 w;
        __builtin_memcpy(&w, &v, 8);
        __builtin_amdgcn_raw_buffer_store_b64(w, r, (int)off, 0, 0);
    }
}

// Half row ro (element offset of its first entry) of the array at base, its
// entries from lim on read as zero.  BUF (the row offset is wave-uniform):
// a resource whose range ends at the row's entry lim, so no per-lane
// predicate at all; otherwise a resource over the whole array and an offset
// moved out of range for h >= lim (one select per point, shared by every
// array read at that point).
template <typename T, bool BUF>
struct HalfRow;

template <typename T>
struct HalfRow<T, true> {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ HalfRow(const cx<T>* base, size_t ro, int lim)
        : r(__builtin_amdgcn_make_buffer_rsrc((void*)(base + ro), (short)0, lim * (int)sizeof(cx<T>), kBufWord3)) {}
    __device__ __forceinline__ cx<T> operator[](int h) const { return buf_load<T>(r, (unsigned)h * sizeof(cx<T>)); }
};

template <typename T>
struct HalfRow<T, false> {
    __amdgpu_buffer_rsrc_t r;
    unsigned ro;
    int lim;
    __device__ __forceinline__ HalfRow(const cx<T>* base, size_t ro_, int lim_)
        : r(__builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)kBufRange, kBufWord3)),
          ro((unsigned)ro_), lim(lim_) {}
    __device__ __forceinline__ cx<T> operator[](int h) const {
        return buf_load<T>(r, h < lim ? (ro + (unsigned)h) * (unsigned)sizeof(cx<T>) : kBufRange);
    }
};

// pack_at over two HalfRows
template <typename T, int N, typename R>
__device__ __forceinline__ cx<T> pack_rows(const R& A, const R& B, int m) {
    const int h = m <= N / 2 ? m : N - m;
    return pack_vals<T, N>(A[h], B[h], m);
}

// a wave-uniform pointer the compiler cannot see through: loads from it are
// not merged with earlier loads of the same addresses
template <typename P>
__device__ __forceinline__ P launder(P p) {
    asm volatile("" : "+s"(p));
    return p;
}

// the row index of a line, made wave-uniform (SGPR) when the line spans whole wavefronts
template <int P>
__device__ __forceinline__ int line_row(int row) {
    if constexpr (P >= 64) return __builtin_amdgcn_readfirstlane(row);
    return row;
}

// Store the half spectra of a forward-transformed packed row: thread j holds
// X[j + k*P]; writes (X[h] + conj X[N-h])/2 and (X[h] - conj X[N-h])/(2i)
// for h < lim.
template <typename T, int N, int E, bool WAVE>
__device__ __forceinline__ void unpack_store(cx<T>* v, int j, const LdsLine<T, false, WAVE>& L, cx<T>* __restrict__ o1,
                                             cx<T>* __restrict__ o2, bool valid, int lim = N) {
    constexpr int P = N / E;
    lds_sync<WAVE>();
#pragma unroll
    for (int k = 0; k < E; ++k) L.st(j + k * P, v[k]);
    lds_sync<WAVE>();
    const T half = (T)0.5;
#pragma unroll
    for (int k = 0; k < E / 2; ++k) {
        const int h = j + k * P;
        const cx<T> a = L.ld(h);
        const cx<T> bm = L.ld((N - h) & (N - 1));
        const cx<T> b = cconj(bm);
        const cx<T> d = csub(a, b);
        if (valid && h < lim) {
            o1[h] = {(a.x + b.x) * half, (a.y + b.y) * half};
            o2[h] = {d.y * half, -d.x * half};
        }
    }
    if (j == 0 && valid && N / 2 < lim) {
        const cx<T> a = L.ld(N / 2);
        o1[N / 2] = {a.x, (T)0};
        o2[N / 2] = {a.y, (T)0};
    }
}

// ---------------------------------------------------------------------------
// spectral update after the y-forward column pass (k3d::colx_kernel, which
// also applies the 1/(ia*ja) of fft2d): one thread per mode, rows streamed
// contiguously.
//   UP_PROLOGUE  u_hat = dealias(fft2d(u))                        (mainloop2D :41-47)
//   UP_A/B/C     r = P(D(N + nu Lap s)), Kutta RK3 combination    (RK32D :183-200)
//   UP_CN_*      u = uold + dt r, next midpoint 0.5 (u + uold)    (CN2D :215-225)
// s is the stage input (the viscous term of compute_ut2D :841-846).
// ---------------------------------------------------------------------------
template <typename T, int MODE>
__global__ __launch_bounds__(256) void update_kernel(ColFwdArgs<T> a) {
    if (a.gate.closed()) return;
    at_member(a);
    const int kx = blockIdx.x * 256 + threadIdx.x;
    const int m = (int)blockIdx.y >= a.ky_skip_at ? (int)blockIdx.y + a.ky_skip : (int)blockIdx.y;
    if (kx >= a.nkx) return;
    const int N = a.g.n1;
    const int ky = ksigned(m, N);
    const int kk2 = kx * kx + ky * ky;
    const size_t idx = (size_t)m * a.pitch + kx;
    cx<T> t0 = a.nx[0][idx], t1 = a.nx[1][idx];
    if (MODE == UP_PROLOGUE) {
        if (kk2 >= a.g.cut2) { t0 = {0, 0}; t1 = {0, 0}; }
        a.U[0][idx] = t0;
        a.U[1][idx] = t1;
        return;
    }
    const T dt = (T)a.g.dt;
    const cx<T> U0 = a.U[0][idx], U1 = a.U[1][idx];
    cx<T> s0, s1;  // this stage's input spectrum (for the viscous term)
    cx<T> ka0, ka1;
    if (MODE == UP_A) {
        s0 = U0; s1 = U1;
    } else if (MODE == UP_CN_MID || MODE == UP_CN_LAST) {
        if (a.cn_first) { s0 = U0; s1 = U1; }                  // first sweep: u_hat
        else { s0 = a.UB[0][idx]; s1 = a.UB[1][idx]; }         // this sweep's midpoint
    } else if (MODE == UP_B) {
        ka0 = a.KA[0][idx]; ka1 = a.KA[1][idx];
        s0 = cadd(U0, cscale(cscale(ka0, dt), (T)0.5));
        s1 = cadd(U1, cscale(cscale(ka1, dt), (T)0.5));
    } else {
        ka0 = a.KA[0][idx]; ka1 = a.KA[1][idx];
        s0 = a.UB[0][idx]; s1 = a.UB[1][idx];
    }
    // r_i = N_i + nu * Lap(u_i)   (compute_ut2D :841-846)
    const T vis = (T)(-a.g.nu * (double)kk2);
    t0 = cadd(t0, cscale(s0, vis));
    t1 = cadd(t1, cscale(s1, vis));
    if (kk2 >= a.g.cut2) { t0 = {0, 0}; t1 = {0, 0}; }
    if (a.g.lproject) {
        // projection2D (tool.F90:219-243)
        if (kk2 > 0) {
            const T kkr = sqrt((T)kk2);
            const T c1 = (T)kx / kkr, c2 = (T)ky / kkr;
            if (2 * m == N) {
                // y-Nyquist row (kept by the dealias when ja < 2 ia / 3): the reference
                // projects the full c2c spectrum, where the partner of (kx, +N/2) is
                // (-kx, +N/2), then keeps the real part of the inverse, so the two
                // projections' kx*ky cross terms cancel
                t0 = cscale(cscale(t0, c1), c1);
                t1 = cscale(cscale(t1, c2), c2);
            } else {
                const cx<T> ud = cadd(cscale(t0, c1), cscale(t1, c2));
                t0 = cscale(ud, c1);
                t1 = cscale(ud, c2);
            }
        } else {
            t0 = {0, 0}; t1 = {0, 0};
        }
    }
    if (MODE == UP_CN_MID) {
        const cx<T> n0 = cadd(U0, cscale(t0, dt)), n1 = cadd(U1, cscale(t1, dt));
        a.UB[0][idx] = cscale(cadd(n0, U0), (T)0.5);
        a.UB[1][idx] = cscale(cadd(n1, U1), (T)0.5);
    } else if (MODE == UP_CN_LAST) {
        a.U[0][idx] = cadd(U0, cscale(t0, dt));
        a.U[1][idx] = cadd(U1, cscale(t1, dt));
    } else if (MODE == UP_A) {
        a.KA[0][idx] = t0; a.KA[1][idx] = t1;
        a.UB[0][idx] = cadd(U0, cscale(cscale(t0, dt), (T)0.5));
        a.UB[1][idx] = cadd(U1, cscale(cscale(t1, dt), (T)0.5));
    } else if (MODE == UP_B) {
        const T dt2 = (T)2 * dt;
        a.UB[0][idx] = cadd(csub(U0, cscale(ka0, dt)), cscale(t0, dt2));
        a.UB[1][idx] = cadd(csub(U1, cscale(ka1, dt)), cscale(t1, dt2));
        a.KA[0][idx] = cadd(ka0, cscale(t0, (T)4));
        a.KA[1][idx] = cadd(ka1, cscale(t1, (T)4));
    } else {
        const T r6 = (T)(1.0 / 6.0);
        a.U[0][idx] = cadd(U0, cscale(cscale(cadd(ka0, t0), dt), r6));
        a.U[1][idx] = cadd(U1, cscale(cscale(cadd(ka1, t1), dt), r6));
    }
}

// ---------------------------------------------------------------------------
// x-direction passes
// ---------------------------------------------------------------------------
template <typename T, int N, int MODE>
__global__ __launch_bounds__((RowCfg<T, N>::THREADS)) __attribute__((amdgpu_waves_per_eu(RowCfg<T, N>::WPE)))
void row_kernel(RowArgs<T> a) {
    using C = RowCfg<T, N>;
    constexpr int E = C::E, P = C::P, LR = C::L;
    if (a.gate.closed()) return;
    if (MODE != R_C2C_FWD && MODE != R_C2C_INV) at_member(a);
    __shared__ cx<T> lds[LR * lds_line_cx(N)];
    const int tid = threadIdx.x;
    const int l = tid / P, j = tid % P;
    const int row0 = blockIdx.x * LR + l;
    const bool valid = row0 < a.g.n1;
    const int row = valid ? row0 : a.g.n1 - 1;
    const LdsLine<T, false, (P <= 64)> Ls{lds + l * lds_line_cx(N), 1};
    const size_t hp = (size_t)a.g.hp;

    if (MODE == R_C2C_FWD || MODE == R_C2C_INV) {
        cx<T> v[E];
#pragma unroll
        for (int k = 0; k < E; ++k) v[k] = a.cin[(size_t)row * N + j + k * P];
        line_fft<T, N, E, MODE == R_C2C_INV, false>(v, j, Ls, a.tw);
        if (valid) {
#pragma unroll
            for (int k = 0; k < E; ++k) a.cout[(size_t)row * N + j + k * P] = v[k];
        }
        return;
    }
    if (MODE == R_FWD_PACK) {
        cx<T> v[E];
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const size_t o = (size_t)row * N + j + k * P;
            v[k] = {a.rin[0][o], a.rin[1][o]};
        }
        line_fft<T, N, E, false, false>(v, j, Ls, a.tw);
        unpack_store<T, N, E>(v, j, Ls, a.hout[0] + (size_t)row * hp, a.hout[1] + (size_t)row * hp, valid);
        return;
    }
    constexpr bool BUF = P >= 64;
    const size_t ro = (size_t)line_row<P>(row) * hp;
    const int lim = a.g.kxlim;
    const HalfRow<T, BUF> A1(a.hin[0], ro, lim), A2(a.hin[1], ro, lim);
    if (MODE == R_INV_PACK) {
        cx<T> v[E];
#pragma unroll
        for (int k = 0; k < E; ++k) v[k] = pack_rows<T, N>(A1, A2, j + k * P);
        line_fft<T, N, E, true, false>(v, j, Ls, a.tw);
        if (valid) {
#pragma unroll
            for (int k = 0; k < E; ++k) {
                const size_t o = (size_t)row * N + j + k * P;
                a.rout[0][o] = v[k].x;
                a.rout[1][o] = v[k].y;
            }
        }
        return;
    }
    // R_STAGE: u, du/dx, du/dy from half rows, nonlinear term, forward x pass
    cx<T> z[E], zx[E];
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const int m = j + k * P;
        z[k] = pack_rows<T, N>(A1, A2, m);
        const T kxs = (m == N / 2) ? (T)0 : (T)ksigned(m, N);
        zx[k] = cscale(cmuli(z[k]), kxs);
    }
    {
        // the two transforms in lockstep through the row's LDS as two split lines
        T* const tl = (T*)(lds + l * lds_line_cx(N));
        const LdsLine<T, true, (P <= 64)> La{tl, 1}, Lb{tl + lds_line(N), 1};
        line_fft2<T, N, E, true, (P <= 64)>(z, zx, j, La, Lb, a.tw);  // (u1, u2), (d1 u1, d1 u2)
    }
    cx<T> n[E];
#pragma unroll
    for (int k = 0; k < E; ++k) n[k] = {-(z[k].x * zx[k].x), -(z[k].x * zx[k].y)};
    const HalfRow<T, BUF> B1(a.hin[2], ro, lim), B2(a.hin[3], ro, lim);
#pragma unroll
    for (int k = 0; k < E; ++k) {
        // two load groups (as in row_stage_vg_kernel)
        if (k % (E / 2) == 0) __builtin_amdgcn_sched_barrier(0);
        zx[k] = pack_rows<T, N>(B1, B2, j + k * P);
    }
    __builtin_amdgcn_sched_barrier(0);
    line_fft<T, N, E, true, false>(zx, j, Ls, a.tw);  // (d2 u1, d2 u2)
#pragma unroll
    for (int k = 0; k < E; ++k) {
        n[k].x = n[k].x - z[k].y * zx[k].x;
        n[k].y = n[k].y - z[k].y * zx[k].y;
    }
    line_fft<T, N, E, false, false>(n, j, Ls, a.tw);
    unpack_store<T, N, E>(n, j, Ls, a.hout[0] + (size_t)row * hp, a.hout[1] + (size_t)row * hp, valid, lim);
}

// Block reduction of NQ moment partial sums (sums; the last one of the
// velgrad set, eta, is a minimum) into part[(q0 + q) * gridDim.x + block].
template <int THREADS, int NQ>
__device__ __forceinline__ void vg_reduce(double* vals, double (*red)[8], int q0, double* part) {
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
#pragma unroll
    for (int q = 0; q < NQ; ++q) vals[q] = (q0 + q == 7) ? wave_min(vals[q]) : wave_sum(vals[q]);
    __syncthreads();
    if (lane == 0)
        for (int q = 0; q < NQ; ++q) red[wave][q] = vals[q];
    __syncthreads();
    if (tid < NQ) {
        const bool mn = q0 + tid == 7;
        double acc = red[0][tid];
        for (int wv = 1; wv < THREADS / 64; ++wv) acc = mn ? fmin(acc, red[wv][tid]) : acc + red[wv][tid];
        part[(size_t)(q0 + tid) * gridDim.x + blockIdx.x] = acc;
    }
}

// Stage A of RK3 fused with velgrad_calculate2D (solution.F90:1110-1186):
// the derivative fields of the nonlinear term are the velgrad ones, so one
// pass over A1 A2 B1 B2 produces both the x-forward nonlinear term and the
// moments' partial sums (part[q * gridDim.x + block]).  The dissipation
// moment, sum 2 u_i theta Lap(u_i) + |u|^2 Lap(theta), takes u and
// theta = d1 u1 + d2 u2 from the gradient transforms and two more x
// transforms: (Lap u1, Lap u2), whose y-inverses the column pass formed with
// the full multiplier -(k1^2 + k2^2), and Lap(theta) = i k1 Lap u1 + d2 Lap u2.
// Occupancy asked of the compiler: 3 waves per SIMD (168 registers in FP64,
// 128 in FP32; FP64 1024- and 4096-point rows 2, which need more registers).
template <typename T, int N>
__global__ __launch_bounds__((VgCfg<T, N>::THREADS))
__attribute__((amdgpu_waves_per_eu((sizeof(T) == 8 && (N == 1024 || N == 4096)) ? 2 : 3)))
void row_stage_vg_kernel(VgArgs<T> a) {
    using C = VgCfg<T, N>;
    constexpr int E = C::E, P = C::P, LR = C::L;
    if (a.gate.closed()) return;
    at_member(a);
    __shared__ cx<T> lds[LR * lds_line_cx(N)];
    __shared__ double red[C::THREADS / 64][8];
    // theta at each thread's points, held here (each thread reads back only
    // what it wrote) from the moments to the dissipation moment
    __shared__ T thl[LR * N];
    const int tid = threadIdx.x;
    const int l = tid / P, j = tid % P;
    const int row0 = blockIdx.x * LR + l;
    const bool valid = row0 < a.g.n1;
    const int row = valid ? row0 : a.g.n1 - 1;
    const LdsLine<T, false> Ls{lds + l * lds_line_cx(N), 1};
    T* const th = thl + l * N;
    const size_t hp = (size_t)a.g.hp;
    constexpr bool BUF = P >= 64;
    const size_t ro = (size_t)line_row<P>(row) * hp;
    const int lim = a.g.kxlim;
    const HalfRow<T, BUF> A1(a.w[0], ro, lim), A2(a.w[1], ro, lim);
    const HalfRow<T, BUF> B1(a.w[2], ro, lim), B2(a.w[3], ro, lim);
    const double nu = a.g.nu;
    const double eta0 = (double)(2.0f * 3.14f);

    cx<T> z[E], w[E], n[E];
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const int m = j + k * P;
        z[k] = pack_rows<T, N>(A1, A2, m);
        const T kxs = (m == N / 2) ? (T)0 : (T)ksigned(m, N);
        w[k] = cscale(cmuli(z[k]), kxs);
    }
    {
        // the two transforms in lockstep through the row's LDS as two split lines
        T* const tl = (T*)(lds + l * lds_line_cx(N));
        const LdsLine<T, true> La{tl, 1}, Lb{tl + lds_line(N), 1};
        line_fft2<T, N, E, true, false>(z, w, j, La, Lb, a.tw);  // (u1, u2), (d1u1, d1u2)
    }
    // z and w stay live for the moments, so the nonlinear term is formed after
    // the third transform (not held across it)
    cx<T> y[E];
#pragma unroll
    for (int k = 0; k < E; ++k) {
        // two load groups of E/2 points: their 16-byte loads land in registers
        // of their own before pack_rows combines them
        if (k % (E / 2) == 0) __builtin_amdgcn_sched_barrier(0);
        y[k] = pack_rows<T, N>(B1, B2, j + k * P);
    }
    __builtin_amdgcn_sched_barrier(0);
    line_fft<T, N, E, true, false>(y, j, Ls, a.tw);  // d2u1, d2u2
    // eta_min = min sqrt(nu / |div|) = sqrt(nu / max |div|): both functions are
    // monotone, so one division and one square root per thread give the same value
    double s_um2 = 0, s_ijji = 0, s_u2t = 0, s_urms = 0, s_dudx = 0, s_eps = 0, dmax = 0;
    // the nonlinear term goes to the line's LDS as it is formed (each thread
    // reads back only its own entries before the forward transform), so it
    // is not held in registers beside z, w and y
    lds_sync<false>();
#pragma unroll
    for (int k = 0; k < E; ++k) {
        cx<T> nk = {-(z[k].x * w[k].x), -(z[k].x * w[k].y)};
        nk.x = nk.x - z[k].y * y[k].x;
        nk.y = nk.y - z[k].y * y[k].y;
        Ls.st(j + k * P, nk);
        const double u1 = z[k].x, u2 = z[k].y;
        const double b11 = w[k].x, b21 = w[k].y, b12 = y[k].x, b22 = y[k].y;
        const double div = b11 + b22;
        th[j + k * P] = (T)div;
        const double uu = u1 * u1 + u2 * u2;
        s_um2 += div * div * uu;
        s_ijji += (b11 * b11 + b22 * b22 + 2.0 * b12 * b21) * uu;
        s_u2t += div * uu;
        s_eps += nu * (b12 - b21) * (b12 - b21) + nu * div * div;
        dmax = fmax(dmax, fabs(div));
        s_urms += u1 * u1 + u2 * u2;
        s_dudx += b11 * b11 + b22 * b22;
    }
    // the first-derivative moments' partial sums now (their accumulators are
    // not held across the transforms below)
    {
        double eta = fmin(eta0, sqrt(nu / dmax));
        if (!valid) { s_um2 = s_ijji = s_u2t = s_urms = s_dudx = s_eps = 0; eta = eta0; }
        double v0[3] = {s_um2, s_ijji, s_u2t};
        double v1[4] = {s_urms, s_dudx, s_eps, eta};
        vg_reduce<C::THREADS, 3>(v0, red, 0, a.part);
        vg_reduce<C::THREADS, 4>(v1, red, 4, a.part);
    }
    __builtin_amdgcn_sched_barrier(0);
    // nonlinear term: x-forward and split into the two half rows
#pragma unroll
    for (int k = 0; k < E; ++k) n[k] = Ls.ld(j + k * P);
    line_fft<T, N, E, false, false>(n, j, Ls, a.tw);
    unpack_store<T, N, E>(n, j, Ls, a.hout[0] + ro, a.hout[1] + ro, valid, lim);

    // the dissipation moment: (Lap u1, Lap u2), then Lap(theta) = i k1 Lap u1 +
    // d2 Lap u2 (the second transform's imaginary input is zero), from the
    // y-inverses the column pass formed.  Lap u1 is read again (from the L2)
    // through a laundered base pointer: otherwise the compiler merges these
    // loads with the first ones and holds their values across the transform.
    const HalfRow<T, BUF> L1(a.w[4], ro, lim), L2(a.w[5], ro, lim), D2(a.w[6], ro, lim);
    const HalfRow<T, BUF> L1d(launder(a.w[4]), ro, lim);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < E; ++k) {
        if (k % (E / 2) == 0) __builtin_amdgcn_sched_barrier(0);
        w[k] = pack_rows<T, N>(L1, L2, j + k * P);
    }
    __builtin_amdgcn_sched_barrier(0);
    line_fft<T, N, E, true, false>(w, j, Ls, a.tw);  // Lap(u1), Lap(u2)
    double s_dis = 0;
#pragma unroll
    for (int k = 0; k < E; ++k)
        s_dis += 2.0 * (double)z[k].x * (double)th[j + k * P] * (double)w[k].x +
                 2.0 * (double)z[k].y * (double)th[j + k * P] * (double)w[k].y;
#pragma unroll
    for (int k = 0; k < E; ++k) {
        if (k % (E / 2) == 0) __builtin_amdgcn_sched_barrier(0);
        const int m = j + k * P;
        const int h = m <= N / 2 ? m : N - m;
        const T k1o = (h == N / 2) ? (T)0 : (T)h;
        w[k] = pack_vals<T, N>(cadd(cscale(cmuli(L1d[h]), k1o), D2[h]), cx<T>{(T)0, (T)0}, m);
    }
    __builtin_amdgcn_sched_barrier(0);
    line_fft<T, N, E, true, false>(w, j, Ls, a.tw);  // Lap(theta)
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const double uu = (double)z[k].x * z[k].x + (double)z[k].y * z[k].y;
        s_dis += uu * (double)w[k].x;
    }

    if (!valid) s_dis = 0;
    double vd[1] = {s_dis};
    vg_reduce<C::THREADS, 1>(vd, red, 3, a.part);
}

// shell spectra (spectra_compute2D, solution.F90:930-1016); deterministic:
// each wave owns its LDS bins and adds to them in program order.
template <typename T>
__global__ __launch_bounds__(256) void spectra_kernel(SpecArgs<T> a) {
    if (a.gate.closed()) return;
    at_member(a);
    extern __shared__ double sh[];
    const int K = a.allkmax, nb = K + 1;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    double* bins = sh;                // [4][nb][2]
    double* tots = sh + 4 * nb * 2;   // [4][4]
    for (int i = tid; i < 4 * nb * 2; i += 256) bins[i] = 0;
    __syncthreads();
    const int n0 = a.g.n0, n1 = a.g.n1;
    const int nh = a.nkx > 0 ? a.nkx : a.g.nh;  // kx limit of the modes summed
    const int nrow = a.nky > 0 ? a.nky : n1;
    const size_t hp = (size_t)a.g.hp;
    const int nbx = (nh + 63) / 64;
    const int ntask = nrow * nbx;  // < 2^31 (launch_spectra checks)
    // tasks are dealt round-robin over blocks and waves, so the few tasks
    // with many shells <= allkmax (small |k|) are spread over the machine
    const int tstride = 4 * gridDim.x;
    const int t1 = ntask;
    double tot[4] = {0, 0, 0, 0};
    // the member's fields through buffer resources with 32-bit offsets
    // (launch_spectra checks that a field fits 2^31 bytes)
    const __amdgpu_buffer_rsrc_t rU0 = __builtin_amdgcn_make_buffer_rsrc((void*)a.U[0], (short)0, (int)kBufRange, kBufWord3);
    const __amdgpu_buffer_rsrc_t rU1 = __builtin_amdgcn_make_buffer_rsrc((void*)a.U[1], (short)0, (int)kBufRange, kBufWord3);
    constexpr int TB = 4;  // tasks per wave iteration: all loads issued before the arithmetic
    for (int t = blockIdx.x * 4 + wave; t < t1; t += tstride * TB) {
        cx<T> u1[TB], u2[TB];
        int r_[TB], kx_[TB];  // each task's row and kx, formed once (32-bit)
#pragma unroll
        for (int q = 0; q < TB; ++q) {
            const int tq = t + tstride * q;
            const int r0 = tq / nbx;
            const int r = r0 >= a.ky_at ? r0 + a.ky_skip : r0;
            const int kx = (tq - r0 * nbx) * 64 + lane;
            r_[q] = r;
            kx_[q] = kx;
            // lanes past the summed modes read zero through an offset past the
            // buffer's range: no branch around the loads, which then all issue
            // before the first is waited for
            const unsigned o = (tq < t1 && kx < nh) ? (unsigned)(((size_t)r * hp + kx) * sizeof(cx<T>)) : kBufRange;
            u1[q] = buf_load<T>(rU0, o);
            u2[q] = buf_load<T>(rU1, o);
        }
#pragma unroll
        for (int q = 0; q < TB; ++q) {
            const int tq = t + tstride * q;
            if (tq >= t1) break;  // wave-uniform
            const int r = r_[q];
            const int kx = kx_[q];
            double es = 0, ed = 0, w = 0;
            int s = INT_MAX;
            if (kx < nh) {
                const int ky = ksigned(r, n1);
                w = (kx == 0 || kx == n0 / 2) ? 1.0 : 2.0;
                const double kk = sqrt((double)(kx * kx + ky * ky));
                if (kk > 0.5) {
                    const double ik = 1.0 / kk;
                    const double c1 = kx * ik, c2 = ky * ik;
                    if (2 * r == n1 && kx > 0 && 2 * kx < n0) {
                        // y-Nyquist row: this entry stands for the reference's modes
                        // (kx, +N/2) and (-kx, +N/2), whose kx*ky cross terms cancel
                        // (the corner (N/2, N/2) stands alone and keeps them)
                        const double a1 = (double)u1[q].x * (double)u1[q].x + (double)u1[q].y * (double)u1[q].y;
                        const double a2 = (double)u2[q].x * (double)u2[q].x + (double)u2[q].y * (double)u2[q].y;
                        es = (a1 * (c2 * c2) + a2 * (c1 * c1)) * 0.5;
                        ed = (a1 * (c1 * c1) + a2 * (c2 * c2)) * 0.5;
                    } else {
                        const double usx = (double)u1[q].x * c2 - (double)u2[q].x * c1;
                        const double usy = (double)u1[q].y * c2 - (double)u2[q].y * c1;
                        const double udx = (double)u1[q].x * c1 + (double)u2[q].x * c2;
                        const double udy = (double)u1[q].y * c1 + (double)u2[q].y * c2;
                        es = (usx * usx + usy * usy) * 0.5;
                        ed = (udx * udx + udy * udy) * 0.5;
                    }
                    tot[3] += w * (ed * ik);
                }
                tot[0] += w * es;
                tot[1] += w * ed;
                tot[2] += w * (ed * (kk * kk));
                const int sh_ = (int)(kk + 0.5);
                if (sh_ <= K) s = sh_;
            }
            bin_runs(bins + wave * nb * 2, s, w * es, w * ed, lane);
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) tot[q] = wave_sum(tot[q]);
    if (lane == 0)
        for (int q = 0; q < 4; ++q) tots[wave * 4 + q] = tot[q];
    __syncthreads();
    // partials are stored [quantity][block] so the reduction reads them coalesced
    for (int b = tid; b < nb; b += 256) {
        for (int c = 0; c < 2; ++c) {
            double acc = 0;
            for (int wv = 0; wv < 4; ++wv) acc += bins[(wv * nb + b) * 2 + c];
            a.part_bins[((size_t)b * 2 + c) * gridDim.x + blockIdx.x] = acc;
        }
    }
    if (tid < 4) {
        double acc = 0;
        for (int wv = 0; wv < 4; ++wv) acc += tots[wv * 4 + tid];
        a.part_tot[(size_t)tid * gridDim.x + blockIdx.x] = acc;
    }
}

// sum over all modes of w*(|u1_hat|^2 + |u2_hat|^2) = <u1^2 + u2^2> (Parseval,
// normalised coefficients), per block, for forcing methods 1 and 5
template <typename T>
__global__ __launch_bounds__(256) void energy_kernel(EnergyArgs<T> a) {
    if (a.gate.closed()) return;
    at_member(a);
    __shared__ double red[4];
    const int n0 = a.g.n0, nh = a.g.nh;
    const size_t hp = (size_t)a.g.hp;
    const size_t total = (size_t)nh * a.g.n1;
    double e = 0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const int kx = (int)(i % nh), r = (int)(i / nh);
        const double w = (kx == 0 || kx == n0 / 2) ? 1.0 : 2.0;
        const cx<T> u1 = a.U[0][(size_t)r * hp + kx], u2 = a.U[1][(size_t)r * hp + kx];
        e += w * ((double)u1.x * u1.x + (double)u1.y * u1.y + (double)u2.x * u2.x + (double)u2.y * u2.y);
    }
    e = wave_sum(e);
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = e;
    __syncthreads();
    if (threadIdx.x == 0) a.part_en[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// forcing2D (solution.F90:303-540) methods 1, 2, 5 on the half spectrum.
//   m1: energy = <u.u>/2, factor = sqrt(E_t/energy), u *= factor (scale_kernel)
//   m2: f = band(u), E = <f.f> (no 1/2), factor = sqrt(E_t/E) - 1, u += factor f
//   m5: f = band(u) [dilatational part if lproject], E = <f.f>, energy = <u.u>,
//       factor = sqrt((E_t-energy)/E + 1) - 1 if E_t > energy else 0, u += factor f
// band: (|k| - 0.5) < forcek < (|k| + 0.5).
template <typename T>
__global__ __launch_bounds__(256) void force_band_kernel(ForceArgs<T> a) {
    if (a.gate.closed()) return;
    if (a.mforce && !a.mforce[blockIdx.z]) return;  // an unforced ensemble member
    at_member(a);
    __shared__ double red[4];
    __shared__ double fac;
    const int tid = threadIdx.x;
    const int fm = a.forcemethod;
    double energy = 0;
    if (fm == 1 || fm == 5) {
        if (tid == 0) {
            double s = 0;
            for (int b = 0; b < a.nblk_en; ++b) s += a.part_en[b];
            energy = fm == 1 ? s / 2 : s;
        }
    }
    if (fm == 1) {
        if (tid == 0) {
            const double factor = sqrt(a.target_energy / energy);
            *a.factor_out = factor;
            a.diag->force_factor = factor;
            a.diag->force_energy = energy;
        }
        return;
    }
    const int n0 = a.g.n0, n1 = a.g.n1;
    const size_t hp = (size_t)a.g.hp;
    const int fk = a.forcek;
    const int kxm = fk < n0 / 2 ? fk : n0 / 2;
    const int kym = fk < n1 / 2 ? fk : n1 / 2 - 1;
    const int nx = kxm + 1, ny = 2 * kym + 1;
    const bool proj = fm == 5 && a.g.lproject;
    double e = 0;
    for (int t = tid; t < nx * ny; t += 256) {
        const int kx = t % nx, ky = t / nx - kym;
        const int r = ky >= 0 ? ky : n1 + ky;
        const double kk = sqrt((double)(kx * kx + ky * ky));
        if ((kk - 0.5) < fk && (kk + 0.5) > fk) {
            const double w = (kx == 0 || kx == n0 / 2) ? 1.0 : 2.0;
            double f1x = a.U[0][(size_t)r * hp + kx].x, f1y = a.U[0][(size_t)r * hp + kx].y;
            double f2x = a.U[1][(size_t)r * hp + kx].x, f2y = a.U[1][(size_t)r * hp + kx].y;
            if (proj) {
                const double c1 = kx / kk, c2 = ky / kk;
                const double udx = f1x * c1 + f2x * c2, udy = f1y * c1 + f2y * c2;
                f1x = udx * c1; f1y = udy * c1; f2x = udx * c2; f2y = udy * c2;
            }
            e += w * (f1x * f1x + f1y * f1y + f2x * f2x + f2y * f2y);
        }
    }
    e = wave_sum(e);
    if (tid % 64 == 0) red[tid / 64] = e;
    __syncthreads();
    if (tid == 0) {
        const double E = red[0] + red[1] + red[2] + red[3];
        double factor;
        if (fm == 2) factor = sqrt(a.target_energy / E) - 1.0;
        else factor = a.target_energy > energy ? sqrt((a.target_energy - energy) / E + 1.0) - 1.0 : 0.0;
        fac = factor;
        a.diag->force_factor = factor;
        a.diag->force_E = E;
        if (fm == 5) a.diag->force_energy = energy;
    }
    __syncthreads();
    const double f = fac;
    for (int t = tid; t < nx * ny; t += 256) {
        const int kx = t % nx, ky = t / nx - kym;
        const int r = ky >= 0 ? ky : n1 + ky;
        const double kk = sqrt((double)(kx * kx + ky * ky));
        if ((kk - 0.5) < fk && (kk + 0.5) > fk) {
            cx<T>* p1 = a.U[0] + (size_t)r * hp + kx;
            cx<T>* p2 = a.U[1] + (size_t)r * hp + kx;
            const cx<T> u1 = *p1, u2 = *p2;
            cx<T> f1 = u1, f2 = u2;
            if (proj) {
                const T c1 = (T)(kx / kk), c2 = (T)(ky / kk);
                const cx<T> ud = cadd(cscale(u1, c1), cscale(u2, c2));
                f1 = cscale(ud, c1);
                f2 = cscale(ud, c2);
            }
            *p1 = cadd(u1, cscale(f1, (T)f));
            *p2 = cadd(u2, cscale(f2, (T)f));
        }
    }
}

// u_hat *= factor on every mode (forcing method 1)
template <typename T>
__global__ __launch_bounds__(256) void scale_kernel(ScaleArgs<T> a) {
    if (a.gate.closed()) return;
    if (a.mforce && !a.mforce[blockIdx.z]) return;  // an unforced ensemble member
    at_member(a);
    const T f = (T)*a.factor;
    const size_t total = (size_t)a.g.hp * a.g.n1;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        a.U[0][i] = cscale(a.U[0][i], f);
        a.U[1][i] = cscale(a.U[1][i], f);
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline int xcd_grid_h(int nb) { return ((nb + 7) / 8) * 8; }

#define K2D_SIZES(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)

template <typename T, int N>
static int row_n(int mode, const RowArgs<T>& a, int nrows, hipStream_t s) {
    using C = RowCfg<T, N>;
    dim3 grid((nrows + C::L - 1) / C::L, 1, zdim(a.g.nmem)), block(C::THREADS);
    switch (mode) {
        case R_FWD_PACK: hipLaunchKernelGGL((row_kernel<T, N, R_FWD_PACK>), grid, block, 0, s, a); break;
        case R_INV_PACK: hipLaunchKernelGGL((row_kernel<T, N, R_INV_PACK>), grid, block, 0, s, a); break;
        case R_STAGE: hipLaunchKernelGGL((row_kernel<T, N, R_STAGE>), grid, block, 0, s, a); break;
        case R_C2C_FWD: hipLaunchKernelGGL((row_kernel<T, N, R_C2C_FWD>), grid, block, 0, s, a); break;
        case R_C2C_INV: hipLaunchKernelGGL((row_kernel<T, N, R_C2C_INV>), grid, block, 0, s, a); break;
        default: return -1;
    }
    return 0;
}

template <typename T, int N>
static int vg_n(const VgArgs<T>& a, int nrows, hipStream_t s, int* nblk) {
    using C = VgCfg<T, N>;
    const int g = (nrows + C::L - 1) / C::L;
    if (nblk && *nblk != g) return -1;  // partial buffer sized for a different grid
    if (!a.hout[0] || !a.hout[1]) return -1;
    hipLaunchKernelGGL((row_stage_vg_kernel<T, N>), dim3(g, 1, zdim(a.g.nmem)), dim3(C::THREADS), 0, s, a);
    return 0;
}

template <typename T>
int launch_update(int mode, const ColFwdArgs<T>& a, hipStream_t s) {
    const int last_row = a.nky - 1 >= a.ky_skip_at ? a.nky - 1 + a.ky_skip : a.nky - 1;
    if (a.nkx < 1 || a.nkx > a.g.nh || a.nky < 1 || a.ky_skip < 0 || last_row >= a.g.n1) return -1;
    dim3 grid((a.nkx + 255) / 256, a.nky, zdim(a.g.nmem)), block(256);
    switch (mode) {
#define M(MM) case MM: hipLaunchKernelGGL((update_kernel<T, MM>), grid, block, 0, s, a); break;
        M(UP_PROLOGUE) M(UP_A) M(UP_B) M(UP_C) M(UP_CN_MID) M(UP_CN_LAST)
#undef M
        default: return -1;
    }
    return 0;
}
template <typename T>
int launch_row(int n, int mode, const RowArgs<T>& a, int nrows, hipStream_t s) {
    switch (n) {
#define X(NN) case NN: return row_n<T, NN>(mode, a, nrows, s);
        K2D_SIZES(X)
#undef X
    }
    return -1;
}
template <typename T>
int launch_velgrad(int n, const VgArgs<T>& a, int nrows, hipStream_t s, int* nblk) {
    switch (n) {
#define X(NN) case NN: return vg_n<T, NN>(a, nrows, s, nblk);
        K2D_SIZES(X)
#undef X
    }
    return -1;
}
template <typename T>
int launch_spectra(const SpecArgs<T>& a, int nblk, hipStream_t s) {
    {  // the kernel's task index is 32-bit, its field offsets 32-bit bytes
        const long nh = a.nkx > 0 ? a.nkx : a.g.nh, nrow = a.nky > 0 ? a.nky : a.g.n1;
        if (nrow * ((nh + 63) / 64) + 16L * nblk >= INT_MAX) return -1;
        if ((long)a.g.hp * a.g.n1 * (long)sizeof(cx<T>) >= (long)kBufRange) return -1;
    }
    const size_t shm = sizeof(double) * (4 * (size_t)(a.allkmax + 1) * 2 + 16);
    hipLaunchKernelGGL((spectra_kernel<T>), dim3(nblk, 1, zdim(a.g.nmem)), dim3(256), shm, s, a);
    return 0;
}
template <typename T>
int launch_force_band(const ForceArgs<T>& a, hipStream_t s) {
    hipLaunchKernelGGL((force_band_kernel<T>), dim3(1, 1, zdim(a.g.nmem)), dim3(256), 0, s, a);
    return 0;
}
template <typename T>
int launch_energy(const EnergyArgs<T>& a, int nblk, hipStream_t s) {
    hipLaunchKernelGGL((energy_kernel<T>), dim3(nblk, 1, zdim(a.g.nmem)), dim3(256), 0, s, a);
    return 0;
}
template <typename T>
int launch_scale(const ScaleArgs<T>& a, hipStream_t s) {
    hipLaunchKernelGGL((scale_kernel<T>), dim3(1024, 1, zdim(a.g.nmem)), dim3(256), 0, s, a);
    return 0;
}

#define K2D_INSTANTIATE(T)                                                                     \
    template int launch_update<T>(int, const ColFwdArgs<T>&, hipStream_t);                      \
    template int launch_row<T>(int, int, const RowArgs<T>&, int, hipStream_t);                 \
    template int launch_velgrad<T>(int, const VgArgs<T>&, int, hipStream_t, int*);             \
    template int launch_spectra<T>(const SpecArgs<T>&, int, hipStream_t);                      \
    template int launch_force_band<T>(const ForceArgs<T>&, hipStream_t);                      \
    template int launch_energy<T>(const EnergyArgs<T>&, int, hipStream_t);                     \
    template int launch_scale<T>(const ScaleArgs<T>&, hipStream_t);

}  // namespace k2d
