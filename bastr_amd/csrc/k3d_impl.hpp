// k3d_impl.hpp -- kernel bodies of the 3-D step (included once per precision).
//
// Reference correspondence (src/solution.F90, src/tool.F90):
//   colx_kernel (inverse jobs)  spectral derivatives (compute_ut3D :874-885,
//                               velgrad_calculate3D :1224-1237), z and y halves
//                               of the inverse transforms
//   row3_kernel<R3_STAGE>       x half of those transforms, -u_j d_j u_i (:903-908),
//                               x half of fft3d (:912-914)
//   colx_kernel (forward) +
//   cf3_kernel<C3_A/B/C>        y and z halves of fft3d, normalisation, +nu*Lap(u),
//                               dealiasing3D (tool.F90:196-217), projection3D
//                               (:245-270), RK33D (:233-270) / CN3D (:272-301)
//   spec3_kernel                spectra_compute3D (:1018-1108)
//   vg3_kernel                  velgrad_calculate3D (:1188-1286), all its moments
//   force3_kernel               forcing3D methods 1, 2, 5 (:542-814)
#pragma once
#include <climits>
#include "k2d_impl.hpp"
#include "k3d.hpp"

namespace k3d {

using k2d::ksigned;
using k2d::pack_at;
using k2d::pack_vals;
using k2d::HalfRow;
using k2d::pack_rows;
using k2d::unpack_store;
using k2d::xcd_block;

// ---------------------------------------------------------------------------
// generic column pass: out_j = FFT(sum_t mult_t(k) in_t), one job per
// workgroup.  The jobs of one column block are consecutive in the
// XCD-contiguous block order, so they run together on one XCD and the
// inputs they share come from its L2.  MT = 1 serves the single-term jobs
// (most of them) with a lean load phase; MT = 4 the sums.
// ---------------------------------------------------------------------------
// The multipliers as (c0 + c2 k^2) g, times i where `usei`: g is the pass's
// wavenumber ko (zero at its Nyquist index) where `useko`, else a per-column
// constant.  Every product is of integer-valued operands, exact or rounded
// once, so each factor equals the direct formula's:
//   M_ONE 1; M_IK i ko; M_MK2 -k^2; M_MIK3 -i ko^3 = i (-k^2) ko;
//   M_LAP -(kc^2 + k^2); M_LAP_IKC -i (kc^2 + k^2) kco; M_LAP_IK -i (kc^2 + k^2) ko
// (kc^2 + kx^2 for kc^2 with a column kx: the full 3-D Laplacian).
template <typename T>
struct MultCoef {
    T c0, c2, g;
    bool useko, usei;
};
template <typename T>
__device__ __forceinline__ MultCoef<T> mult_coef(int mult, T kc, T kco, T kx = 0) {
    const T kc2 = -(kc * kc + kx * kx);
    switch (mult) {
        case M_IK: return {(T)1, (T)0, (T)1, true, true};
        case M_MK2: return {(T)0, (T)-1, (T)1, false, false};
        case M_MIK3: return {(T)0, (T)-1, (T)1, true, true};
        case M_LAP: return {kc2, (T)-1, (T)1, false, false};
        case M_LAP_IKC: return {kc2, (T)-1, kco, false, true};
        case M_LAP_IK: return {kc2, (T)-1, (T)1, true, true};
        default: return {(T)1, (T)0, (T)1, false, false};
    }
}
// v times the multiplier at position m of a length-N pass
template <typename T, int N>
__device__ __forceinline__ cx<T> apply_mult(cx<T> v, const MultCoef<T>& c, int m) {
    const T k = (T)ksigned(m, N);
    const T ko = (m == N / 2) ? (T)0 : k;
    const T f = (c.c0 + c.c2 * (k * k)) * (c.useko ? ko : c.g);
    const cx<T> w = cscale(v, f);
    return c.usei ? cmuli(w) : w;
}

// the multi-term variant runs 8 points per thread in FP64 (register pressure);
// the single-term one EM points (16 by default, 8 where the autotuner finds
// radix-8 passes faster) on blocks CBM times the default width (colx_plan)
template <typename T, int N, int MT, int EM = 16, int CBM = 1>
using ColxCfg = ColCfg<T, N, MT == 1 ? EM : (sizeof(T) == 4 ? 16 : 8), MT == 1 ? CBM : 1>;

// waves per SIMD asked of the compiler: a workgroup of W waves per SIMD
// (THREADS / 256) only ever adds W, so ask for the largest multiple of W up to
// `want` (asking 3 of a 2-wave workgroup only lowers the register budget)
constexpr int wpe_for(int threads, int want) {
    const int w = threads <= 256 ? 1 : threads / 256;
    return w >= want ? w : want / w * w;
}

// Loads and stores go through a buffer resource at the column block's first
// entry (wave-uniform), with 32-bit offsets ((m pitch + l) entries; colx_n
// checks that N pitch entries fit 2^31 bytes): entries the pass skips
// (|k| >= kcut) take an offset past the range, which reads zero or drops the
// store, so neither phase branches per point.
template <typename T, int N, bool INV, int MT, int EM = 16, int CBM = 1>
__global__ __launch_bounds__((ColxCfg<T, N, MT, EM, CBM>::THREADS))
__attribute__((amdgpu_waves_per_eu(wpe_for(ColxCfg<T, N, MT, EM, CBM>::THREADS, MT == 1 ? 3 : 2))))
void colx_kernel(ColXArgs<T> a) {
    using C = ColxCfg<T, N, MT, EM, CBM>;
    static_assert(C::FITS, "column plan exceeds a workgroup");
    constexpr int E = C::E, P = C::P, CB = C::CB;
    constexpr unsigned ES = sizeof(cx<T>);
    if (a.gate.closed()) return;
    __shared__ T lds[CB * lds_line(N)];
    const int nbpp = a.ncols / CB;
    const int vb = xcd_block(blockIdx.x, a.njob * nbpp * a.nplanes);
    if (vb < 0) return;
    const int q = vb % a.njob, blk = vb / a.njob;
    const int plane0 = blk / nbpp, cb = blk % nbpp;
    const int plane = plane0 >= a.pskip_at ? plane0 + a.pskip : plane0;
    const int tid = threadIdx.x;
    const int l = tid % CB, j = tid / CB;
    // the block's first entry (wave-uniform) and this lane's offset from it
    const size_t blk0 = (size_t)blockIdx.z * (size_t)a.mstride + (size_t)plane * a.plane_stride + (size_t)cb * CB;
    const unsigned pitch = (unsigned)a.pitch;
    const unsigned lo = ((unsigned)j * pitch + (unsigned)l) * ES;
    const LdsLine<T, true> L{lds + l, CB};
    const ColJob& jb = a.job[q];
    // the column's own wavenumber (LAP multipliers only)
    T kc = 0, kco = 0, kx = 0;
    if (a.kc_div > 0) {
        const unsigned co = (unsigned)((size_t)plane * a.plane_stride + (size_t)cb * CB + l);
        const unsigned ic = co / (unsigned)a.kc_div;
        kc = (T)ksigned((int)ic, a.kc_n);
        kco = (2 * (int)ic == a.kc_n) ? (T)0 : kc;
        if (a.kx_mod > 0) kx = (T)(co % (unsigned)a.kx_mod);
    }
    // entries |k| >= kcut along the pass: zero in an inverse pass's inputs,
    // not stored by a forward pass
    const int kcut = a.kcut > 0 ? a.kcut : N;
    auto live = [&](int m) { return (m <= N / 2 ? m : N - m) < kcut; };
    // point k: the lane offset lo plus k P pitch entries, or past the range
    // where the pass skips the point
    auto at = [&](int k) { return lo + (unsigned)(k * P) * pitch * ES; };
    auto off = [&](int k) { return live(j + k * P) ? at(k) : k2d::kBufRange; };
    cx<T> v[E];
    if (MT == 1) {
        const auto r = __builtin_amdgcn_make_buffer_rsrc((void*)(a.in[jb.in[0]] + blk0), (short)0,
                                                         (int)k2d::kBufRange, k2d::kBufWord3);
        const int mult = jb.mult[0];
#pragma unroll
        for (int k = 0; k < E; ++k) v[k] = k2d::buf_load<T>(r, INV ? off(k) : at(k));
        if (INV && mult != M_ONE) {  // (forward jobs take no multiplier: colx_n)
            const MultCoef<T> c = mult_coef<T>(mult, kc, kco, kx);
#pragma unroll
            for (int k = 0; k < E; ++k) v[k] = apply_mult<T, N>(v[k], c, j + k * P);
        }
    } else {
#pragma unroll
        for (int k = 0; k < E; ++k) v[k] = {(T)0, (T)0};
        for (int t = 0; t < jb.nterm; ++t) {
            // one term's loads at a time: keeps the next term's loads from
            // being hoisted over this one's accumulation (register pressure)
            __builtin_amdgcn_sched_barrier(0);
            const auto r = __builtin_amdgcn_make_buffer_rsrc((void*)(a.in[jb.in[t]] + blk0), (short)0,
                                                             (int)k2d::kBufRange, k2d::kBufWord3);
            const MultCoef<T> c = mult_coef<T>(jb.mult[t], kc, kco, kx);
#pragma unroll
            for (int k = 0; k < E; ++k) {
                const cx<T> x = k2d::buf_load<T>(r, INV ? off(k) : at(k));
                v[k] = cadd(v[k], jb.mult[t] == M_ONE ? x : apply_mult<T, N>(x, c, j + k * P));
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    line_fft<T, N, E, INV, true>(v, j, L, a.tw);
    const auto w = __builtin_amdgcn_make_buffer_rsrc((void*)(a.out[jb.out] + blk0), (short)0, (int)k2d::kBufRange,
                                                     k2d::kBufWord3);
    const T sc = (T)a.scale;
#pragma unroll
    for (int k = 0; k < E; ++k)
        k2d::buf_store<T>(w, INV ? at(k) : off(k), (sc == (T)1) ? v[k] : cscale(v[k], sc));
}

// ---------------------------------------------------------------------------
// z-forward pass fused with the spectral update of the three components.
// The projection couples the components, so a workgroup transforms all
// three columns of its block and updates the modes in registers:
//   C3_PROLOGUE  u_hat = dealias(fft3d(u))                       (mainloop3D :111-116)
//   C3_A/B/C     r = P(D(N + nu Lap s)), Kutta RK3 combination   (RK33D :233-270)
//   C3_CN_*      u = uold + dt r, next midpoint 0.5 (u + uold)   (CN3D :272-301)
// with compute_ut3D's viscous term (:903-908), dealiasing3D (tool.F90:196-217)
// and projection3D (:245-270); s is the stage input.
// ---------------------------------------------------------------------------
template <typename T, int N>
using Fwd3Cfg = ColCfg<T, N, sizeof(T) == 8 ? 8 : 16>;

template <typename T, int N, int MODE>
__global__ __launch_bounds__((Fwd3Cfg<T, N>::THREADS)) __attribute__((amdgpu_waves_per_eu(2)))
void cf3_kernel(Cf3Args<T> a) {
    using C = Fwd3Cfg<T, N>;
    constexpr int E = C::E, P = C::P, CB = C::CB;
    if (a.gate.closed()) return;
    __shared__ T lds[CB * lds_line(N)];
    const int ncols = a.g.hp * a.g.n1;
    const int nbr = a.nkx / CB;
    const int blk = xcd_block(blockIdx.x, a.nky * nbr);
    if (blk < 0) return;
    const int tid = threadIdx.x;
    const int l = tid % CB, j = tid / CB;
    const int r0 = blk / nbr;
    const int r = r0 >= a.ky_skip_at ? r0 + a.ky_skip : r0;
    const int col = r * a.g.hp + (blk % nbr) * CB + l;
    const LdsLine<T, true> L{lds + l, CB};
    const size_t pitch = (size_t)ncols;
    const T sc = (T)a.g.inv_n;
    cx<T> v[3][E];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int k = 0; k < E; ++k) v[c][k] = a.nx[c][(size_t)(j + k * P) * pitch + col];
        line_fft<T, N, E, false, true>(v[c], j, L, a.tw);
    }
    const int kx = col % a.g.hp;
    const int ky = ksigned(col / a.g.hp, a.g.n1);
    const T dt = (T)a.g.dt;
#pragma unroll
    for (int k = 0; k < E; ++k) {
        if (k % 2 == 0) __builtin_amdgcn_sched_barrier(0);
        const int m = j + k * P;
        const int kz = ksigned(m, N);
        // beyond the dealias radius along z the state holds zeros that stay
        // zeros; the inverse passes do not read these entries
        if (MODE != C3_PROLOGUE && a.kzcut > 0 && (kz < 0 ? -kz : kz) >= a.kzcut) continue;
        const int kk2 = kx * kx + ky * ky + kz * kz;
        const size_t idx = (size_t)m * pitch + col;
        cx<T> t[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) t[c] = cscale(v[c][k], sc);
        if (MODE == C3_PROLOGUE) {
#pragma unroll
            for (int c = 0; c < 3; ++c) a.U[c][idx] = (kk2 >= a.g.cut2) ? cx<T>{0, 0} : t[c];
            continue;
        }
        cx<T> U[3], s[3], ka[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) U[c] = a.U[c][idx];
        if (MODE == C3_A) {
#pragma unroll
            for (int c = 0; c < 3; ++c) s[c] = U[c];
        } else if (MODE == C3_CN_MID || MODE == C3_CN_LAST) {
#pragma unroll
            for (int c = 0; c < 3; ++c) s[c] = a.cn_first ? U[c] : a.UB[c][idx];
        } else if (MODE == C3_B) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                ka[c] = a.KA[c][idx];
                s[c] = cadd(U[c], cscale(cscale(ka[c], dt), (T)0.5));
            }
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                ka[c] = a.KA[c][idx];
                s[c] = a.UB[c][idx];
            }
        }
        const T vis = (T)(-a.g.nu * (double)kk2);
#pragma unroll
        for (int c = 0; c < 3; ++c) t[c] = cadd(t[c], cscale(s[c], vis));
        if (kk2 >= a.g.cut2) {
#pragma unroll
            for (int c = 0; c < 3; ++c) t[c] = {0, 0};
        }
        if (a.g.lproject) {
            if (kk2 > 0) {
                const T kkr = sqrt((T)kk2);
                const T c1 = (T)kx / kkr, c2 = (T)ky / kkr, c3 = (T)kz / kkr;
                const cx<T> ud = cadd(cadd(cscale(t[0], c1), cscale(t[1], c2)), cscale(t[2], c3));
                t[0] = cscale(ud, c1);
                t[1] = cscale(ud, c2);
                t[2] = cscale(ud, c3);
            } else {
#pragma unroll
                for (int c = 0; c < 3; ++c) t[c] = {0, 0};
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (MODE == C3_A) {
                a.KA[c][idx] = t[c];
                a.UB[c][idx] = cadd(U[c], cscale(cscale(t[c], dt), (T)0.5));
            } else if (MODE == C3_B) {
                const T dt2 = (T)2 * dt;
                a.UB[c][idx] = cadd(csub(U[c], cscale(ka[c], dt)), cscale(t[c], dt2));
                a.KA[c][idx] = cadd(ka[c], cscale(t[c], (T)4));
            } else if (MODE == C3_C) {
                a.U[c][idx] = cadd(U[c], cscale(cscale(cadd(ka[c], t[c]), dt), (T)(1.0 / 6.0)));
            } else if (MODE == C3_CN_MID) {
                const cx<T> nw = cadd(U[c], cscale(t[c], dt));
                a.UB[c][idx] = cscale(cadd(nw, U[c]), (T)0.5);
            } else {  // C3_CN_LAST
                a.U[c][idx] = cadd(U[c], cscale(t[c], dt));
            }
        }
    }
}

// one real field as a packed row (second field zero)
// ---------------------------------------------------------------------------
// 2-D y-forward pass fused with the spectral update (the 2-D counterpart of
// cf3_kernel): a workgroup transforms both components of CB kx columns and
// updates the modes in registers, so the y-transformed tendencies never go
// to HBM.  Stage semantics as k2d::update_kernel (RK32D :183-200, CN2D
// :215-225, compute_ut2D :841-846, dealiasing2D, projection2D); rows with
// |ky| >= kcut hold the dealiased state's zeros and are left as they are.
// ---------------------------------------------------------------------------
// 8 points per thread in FP64 (two components are live across the second
// transform), the column rules' block width capped at 512 threads (so 256
// registers per thread are available)
template <typename T, int N>
struct Fwd2Cfg {
    using Base = ColCfg<T, N, sizeof(T) == 8 ? 8 : 16>;
    // FP64 columns of 64 to 256 points: 4 points per thread and one wavefront
    // per block, so a small grid's fused pass spreads over more CUs with
    // shorter transform chains (its data sit in the caches, where short row
    // segments cost little)
    static constexpr bool NARROW = sizeof(T) == 8 && N <= 256 && N >= 64;
    static constexpr int E = NARROW ? 4 : Base::E, P = N / E;
    static constexpr int CB = NARROW ? 64 / P : (Base::CB * P > 512 ? (P >= 512 ? 1 : 512 / P) : Base::CB);
    static constexpr int THREADS = CB * P;
};

template <typename T, int N, int MODE>
__global__ __launch_bounds__((Fwd2Cfg<T, N>::THREADS)) __attribute__((amdgpu_waves_per_eu(2)))
void cf2_kernel(k2d::ColFwdArgs<T> a) {
    using C = Fwd2Cfg<T, N>;
    constexpr int E = C::E, P = C::P, CB = C::CB;
    if (a.gate.closed()) return;
    __shared__ T lds[CB * lds_line(N)];
    const int blk = xcd_block(blockIdx.x, a.nkx / CB);
    if (blk < 0) return;
    k2d::at_member(a);
    const int tid = threadIdx.x;
    const int l = tid % CB, j = tid / CB;
    const int kx = blk * CB + l;
    const LdsLine<T, true> L{lds + l, CB};
    const T sc = (T)a.g.inv_n;
    const int kcut = a.kcut > 0 ? a.kcut : N;
    // every array through a buffer resource at the block's first column, with
    // 32-bit lane offsets (cf2_n checks that N pitch entries fit 2^31 bytes)
    const int c0 = blk * CB;
    const unsigned pitch = (unsigned)a.pitch;
    auto rs = [&](const cx<T>* p) {
        return __builtin_amdgcn_make_buffer_rsrc((void*)(p + c0), (short)0, (int)k2d::kBufRange, k2d::kBufWord3);
    };
    auto at = [&](int k) { return ((unsigned)(j + k * P) * pitch + (unsigned)l) * (unsigned)sizeof(cx<T>); };
    const __amdgpu_buffer_rsrc_t rNX[2] = {rs(a.nx[0]), rs(a.nx[1])}, rU[2] = {rs(a.U[0]), rs(a.U[1])},
                                 rUB[2] = {rs(a.UB[0]), rs(a.UB[1])}, rKA[2] = {rs(a.KA[0]), rs(a.KA[1])};
    // both components' loads issued before either transform
    cx<T> v[2][E];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int k = 0; k < E; ++k) v[c][k] = k2d::buf_load<T>(rNX[c], at(k));
#pragma unroll
    for (int c = 0; c < 2; ++c) line_fft<T, N, E, false, true>(v[c], j, L, a.tw);
    const T dt = (T)a.g.dt;
#pragma unroll
    for (int k = 0; k < E; ++k) {
        if (k % 2 == 0) __builtin_amdgcn_sched_barrier(0);
        const int m = j + k * P;
        const int ky = ksigned(m, N);
        if ((ky < 0 ? -ky : ky) >= kcut) continue;
        const int kk2 = kx * kx + ky * ky;
        const unsigned o = at(k);
        cx<T> t0 = cscale(v[0][k], sc), t1 = cscale(v[1][k], sc);
        const cx<T> U0 = k2d::buf_load<T>(rU[0], o), U1 = k2d::buf_load<T>(rU[1], o);
        cx<T> s0, s1, ka0, ka1;
        if (MODE == k2d::UP_A) {
            s0 = U0; s1 = U1;
        } else if (MODE == k2d::UP_CN_MID || MODE == k2d::UP_CN_LAST) {
            if (a.cn_first) { s0 = U0; s1 = U1; }
            else { s0 = k2d::buf_load<T>(rUB[0], o); s1 = k2d::buf_load<T>(rUB[1], o); }
        } else if (MODE == k2d::UP_B) {
            ka0 = k2d::buf_load<T>(rKA[0], o); ka1 = k2d::buf_load<T>(rKA[1], o);
            s0 = cadd(U0, cscale(cscale(ka0, dt), (T)0.5));
            s1 = cadd(U1, cscale(cscale(ka1, dt), (T)0.5));
        } else {
            ka0 = k2d::buf_load<T>(rKA[0], o); ka1 = k2d::buf_load<T>(rKA[1], o);
            s0 = k2d::buf_load<T>(rUB[0], o); s1 = k2d::buf_load<T>(rUB[1], o);
        }
        const T vis = (T)(-a.g.nu * (double)kk2);
        t0 = cadd(t0, cscale(s0, vis));
        t1 = cadd(t1, cscale(s1, vis));
        if (kk2 >= a.g.cut2) { t0 = {0, 0}; t1 = {0, 0}; }
        if (a.g.lproject) {
            if (kk2 > 0) {
                const T kkr = sqrt((T)kk2);
                const T c1 = (T)kx / kkr, c2 = (T)ky / kkr;
                if (2 * m == N) {
                    // y-Nyquist row: the partner of (kx, +N/2) in the reference's
                    // c2c projection is (-kx, +N/2); the kx*ky cross terms cancel
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
        if (MODE == k2d::UP_CN_MID) {
            const cx<T> n0 = cadd(U0, cscale(t0, dt)), n1 = cadd(U1, cscale(t1, dt));
            k2d::buf_store<T>(rUB[0], o, cscale(cadd(n0, U0), (T)0.5));
            k2d::buf_store<T>(rUB[1], o, cscale(cadd(n1, U1), (T)0.5));
        } else if (MODE == k2d::UP_CN_LAST) {
            k2d::buf_store<T>(rU[0], o, cadd(U0, cscale(t0, dt)));
            k2d::buf_store<T>(rU[1], o, cadd(U1, cscale(t1, dt)));
        } else if (MODE == k2d::UP_A) {
            k2d::buf_store<T>(rKA[0], o, t0);
            k2d::buf_store<T>(rKA[1], o, t1);
            k2d::buf_store<T>(rUB[0], o, cadd(U0, cscale(cscale(t0, dt), (T)0.5)));
            k2d::buf_store<T>(rUB[1], o, cadd(U1, cscale(cscale(t1, dt), (T)0.5)));
        } else if (MODE == k2d::UP_B) {
            const T dt2 = (T)2 * dt;
            k2d::buf_store<T>(rUB[0], o, cadd(csub(U0, cscale(ka0, dt)), cscale(t0, dt2)));
            k2d::buf_store<T>(rUB[1], o, cadd(csub(U1, cscale(ka1, dt)), cscale(t1, dt2)));
            k2d::buf_store<T>(rKA[0], o, cadd(ka0, cscale(t0, (T)4)));
            k2d::buf_store<T>(rKA[1], o, cadd(ka1, cscale(t1, (T)4)));
        } else {
            const T r6 = (T)(1.0 / 6.0);
            k2d::buf_store<T>(rU[0], o, cadd(U0, cscale(cscale(cadd(ka0, t0), dt), r6)));
            k2d::buf_store<T>(rU[1], o, cadd(U1, cscale(cscale(cadd(ka1, t1), dt), r6)));
        }
    }
}

template <typename T, int N, typename R>
__device__ __forceinline__ cx<T> pack1(const R& A, int m) {
    const int h = m <= N / 2 ? m : N - m;
    const cx<T> a = A[h];
    if (m == 0 || m == N / 2) return {a.x, (T)0};
    if (m < N / 2) return a;
    return cconj(a);
}

// pack (A, i*k1*A) for one half row A: value at m of A + i*(i k1 A)
template <typename T, int N, typename R>
__device__ __forceinline__ cx<T> pack_a_ika(const R& A, int m) {
    const int h = m <= N / 2 ? m : N - m;
    const T ko = (h == N / 2) ? (T)0 : (T)h;
    const cx<T> a = A[h];
    const cx<T> b = cscale(cmuli(a), ko);  // i k1 A at h
    if (m == 0 || m == N / 2) return {a.x, b.x};
    if (m < N / 2) return {a.x - b.y, a.y + b.x};
    return {a.x + b.y, b.x - a.y};
}

// ---------------------------------------------------------------------------
// x passes
// ---------------------------------------------------------------------------
template <typename T, int N, int MODE>
__global__ __launch_bounds__((RowCfg<T, N>::THREADS)) void row3_kernel(Row3Args<T> a) {
    using C = RowCfg<T, N>;
    constexpr int E = C::E, P = C::P, LR = C::L;
    if (a.gate.closed()) return;
    __shared__ cx<T> lds[LR * lds_line_cx(N)];
    // In FP64, R3_STAGE keeps the packed (u1, u2) half-row spectrum here for
    // its second use (the x-derivatives) instead of reading it again; each
    // thread reads back only what it wrote.  (In FP32 the second read is cheaper.)
    constexpr bool KEEP = MODE == R3_STAGE && sizeof(T) == 8;
    __shared__ cx<T> keep[KEEP ? LR * N : 1];
    const int tid = threadIdx.x;
    const int l = tid / P, j = tid % P;
    const int nrows = a.g.n1 * a.g.n2;
    const int row0 = blockIdx.x * LR + l;
    const bool valid = row0 < nrows;
    const int row = valid ? row0 : nrows - 1;
    const LdsLine<T, false, (P <= 64)> Ls{lds + l * lds_line_cx(N), 1};
    const size_t hp = (size_t)a.g.hp;
    // a line of at least 64 threads is wave-uniform: its row offset lives in SGPRs
    const size_t ro = (size_t)(P >= 64 ? __builtin_amdgcn_readfirstlane(row) : row) * hp;

    if (MODE == R3_FWD_PACK) {
        cx<T> v[E];
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const size_t o = (size_t)row * N + j + k * P;
            v[k] = {a.rin[0][o], a.rin[1][o]};
        }
        line_fft<T, N, E, false, false>(v, j, Ls, a.tw);
        unpack_store<T, N, E>(v, j, Ls, a.hout[0] + ro, a.hout[1] + ro, valid);
#pragma unroll
        for (int k = 0; k < E; ++k) v[k] = {a.rin[2][(size_t)row * N + j + k * P], (T)0};
        line_fft<T, N, E, false, false>(v, j, Ls, a.tw);
        if (valid) {
#pragma unroll
            for (int k = 0; k < E; ++k) {
                const int h = j + k * P;
                if (h <= N / 2) a.hout[2][ro + h] = (h == 0 || h == N / 2) ? cx<T>{v[k].x, (T)0} : v[k];
            }
        }
        return;
    }
    constexpr bool BUF = P >= 64;
    const int lim = a.g.kxlim;
    if (MODE == R3_INV_PACK) {
        const HalfRow<T, BUF> H0(a.hin[0], ro, lim), H1(a.hin[1], ro, lim), H2(a.hin[2], ro, lim);
        cx<T> v[E];
#pragma unroll
        for (int k = 0; k < E; ++k) v[k] = pack_rows<T, N>(H0, H1, j + k * P);
        line_fft<T, N, E, true, false>(v, j, Ls, a.tw);
        if (valid) {
#pragma unroll
            for (int k = 0; k < E; ++k) {
                const size_t o = (size_t)row * N + j + k * P;
                a.rout[0][o] = v[k].x;
                a.rout[1][o] = v[k].y;
            }
        }
#pragma unroll
        for (int k = 0; k < E; ++k) v[k] = pack1<T, N>(H2, j + k * P);
        line_fft<T, N, E, true, false>(v, j, Ls, a.tw);
        if (valid) {
#pragma unroll
            for (int k = 0; k < E; ++k) a.rout[2][(size_t)row * N + j + k * P] = v[k].x;
        }
        return;
    }
    // R3_STAGE: u, grad u from nine half rows, nonlinear term, x-forward.
    // n3 first (its gradients (b31, b32, b33) come from the (u3, d1 u3) and
    // (d2 u3, d3 u3) transforms), transformed and stored before n1, n2 are
    // accumulated, so u, one partial term and one transform are live at a time.
    const HalfRow<T, BUF> Y1(a.hin[0], ro, lim), Y2(a.hin[1], ro, lim), Y3(a.hin[2], ro, lim);
    cx<T> z[E], w[E];
    T u3[E];
    cx<T>* kp = keep + (KEEP ? l * N : 0);
#pragma unroll
    for (int k = 0; k < E; ++k) {
        z[k] = pack_rows<T, N>(Y1, Y2, j + k * P);
        if constexpr (KEEP) kp[j + k * P] = z[k];
    }
    {
        T n3[E];
#pragma unroll
        for (int k = 0; k < E; ++k) w[k] = pack_a_ika<T, N>(Y3, j + k * P);
        {
            // the two transforms in lockstep through the row's LDS as two split lines
            T* const tl = (T*)(lds + l * lds_line_cx(N));
            const LdsLine<T, true, (P <= 64)> La{tl, 1}, Lb{tl + lds_line(N), 1};
            line_fft2<T, N, E, true, (P <= 64)>(z, w, j, La, Lb, a.tw);  // (u1, u2), (u3, d1 u3)
        }
#pragma unroll
        for (int k = 0; k < E; ++k) {
            u3[k] = w[k].x;
            n3[k] = -(z[k].x * w[k].y);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < E; ++k) w[k] = pack_rows<T, N>(HalfRow<T, BUF>(a.hin[5], ro, lim), HalfRow<T, BUF>(a.hin[8], ro, lim), j + k * P);
        line_fft<T, N, E, true, false>(w, j, Ls, a.tw);  // (d2 u3, d3 u3)
#pragma unroll
        for (int k = 0; k < E; ++k) w[k] = {n3[k] - z[k].y * w[k].x - u3[k] * w[k].y, (T)0};
    }
    line_fft<T, N, E, false, false>(w, j, Ls, a.tw);
    if (valid) {
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int h = j + k * P;
            if (h <= N / 2 && h < lim) a.hout[2][ro + h] = (h == 0 || h == N / 2) ? cx<T>{w[k].x, (T)0} : w[k];
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    cx<T> n12[E];
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const int m = j + k * P;
        const T kxs = (m == N / 2) ? (T)0 : (T)ksigned(m, N);
        const cx<T> zs = KEEP ? kp[m] : pack_rows<T, N>(Y1, Y2, m);
        w[k] = cscale(cmuli(zs), kxs);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < E; ++k) n12[k] = pack_rows<T, N>(HalfRow<T, BUF>(a.hin[3], ro, lim), HalfRow<T, BUF>(a.hin[4], ro, lim), j + k * P);
    {
        // (d1 u1, d1 u2) and (d2 u1, d2 u2) in lockstep; n12 holds the second
        T* const tl = (T*)(lds + l * lds_line_cx(N));
        const LdsLine<T, true, (P <= 64)> La{tl, 1}, Lb{tl + lds_line(N), 1};
        line_fft2<T, N, E, true, (P <= 64)>(w, n12, j, La, Lb, a.tw);
    }
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const cx<T> d2 = n12[k];
        n12[k] = {-(z[k].x * w[k].x), -(z[k].x * w[k].y)};
        n12[k].x = n12[k].x - z[k].y * d2.x;
        n12[k].y = n12[k].y - z[k].y * d2.y;
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < E; ++k) w[k] = pack_rows<T, N>(HalfRow<T, BUF>(a.hin[6], ro, lim), HalfRow<T, BUF>(a.hin[7], ro, lim), j + k * P);
    line_fft<T, N, E, true, false>(w, j, Ls, a.tw);  // (d3 u1, d3 u2)
#pragma unroll
    for (int k = 0; k < E; ++k) {
        n12[k].x = n12[k].x - u3[k] * w[k].x;
        n12[k].y = n12[k].y - u3[k] * w[k].y;
    }
    line_fft<T, N, E, false, false>(n12, j, Ls, a.tw);
    unpack_store<T, N, E>(n12, j, Ls, a.hout[0] + ro, a.hout[1] + ro, valid, lim);
}

template <typename T, int N>
using Vg3Cfg = k2d::VgCfg<T, N, 8>;

// velocity-gradient moments, 3-D (velgrad_calculate3D, solution.F90:1188-1286).
// The nine gradient components are transformed in pairs chosen so that each
// pair's products enter the moments at once: (b12, b21), (b13, b31),
// (b23, b32) for the symmetric terms, then (b22, b33) completes theta with
// b11 from the (u3, b11) transform.  Only u, b11 and later theta are held
// across transforms, so 8 points per thread fit in registers and a 512-point
// line is one wavefront.  The dissipation moment follows from u, theta and two
// more transforms (Lap u and Lap theta, from rows the column passes formed
// with the full Laplacian), so every moment of the iteration comes from this
// one pass, each input row read once (512^3: 3.20 ms FP64, 2.37 ms FP32,
// against two passes that both read u: 3.72 ms FP64, 2.47 ms FP32).
template <typename T, int N>
__global__ __launch_bounds__((Vg3Cfg<T, N>::THREADS)) __attribute__((amdgpu_waves_per_eu(2)))
void vg3_kernel(Vg3Args<T> a) {
    using C = Vg3Cfg<T, N>;
    constexpr int E = C::E, P = C::P, LR = C::L;
    if (a.gate.closed()) return;
    __shared__ cx<T> lds[LR * lds_line_cx(N)];
    __shared__ double red[C::THREADS / 64][8];
    const int tid = threadIdx.x;
    const int l = tid / P, j = tid % P;
    const int nrows = a.g.n1 * a.g.n2;
    const int row0 = blockIdx.x * LR + l;
    const bool valid = row0 < nrows;
    const int row = valid ? row0 : nrows - 1;
    const LdsLine<T, false, (P <= 64)> Ls{lds + l * lds_line_cx(N), 1};
    // a line of at least 64 threads is wave-uniform: its row offset lives in SGPRs
    const size_t ro = (size_t)(P >= 64 ? __builtin_amdgcn_readfirstlane(row) : row) * a.g.hp;
    constexpr bool BUF = P >= 64;
    const int lim = a.g.kxlim;  // entries from kxlim on are zero
    using Row = HalfRow<T, BUF>;
    auto W = [&](int q) { return Row(a.w[q], ro, lim); };
    const double nu = a.g.nu;
    const double eta0 = (double)(2.0f * 3.14f);
    // W: 0..2 Y_c, 3..5 Y2_c (d2), 6..8 Y3_c (d3), 9..11 L_c, 12 D
    // half-row value of d1 X (odd power of k1: zero at the Nyquist index)
    auto d1 = [](const Row& X, int h) {
        const T ko = (h == N / 2) ? (T)0 : (T)h;
        return cscale(cmuli(X[h]), ko);
    };
    cx<T> z[E], w[E];
    T u3[E], b11[E];
    // eta_min = min sqrt(nu / |div|) = sqrt(nu / max |div|): both functions are
    // monotone, so one division and one square root per thread give the same value
    double s_um2 = 0, s_ijji = 0, s_u2t = 0, s_urms = 0, s_dudx = 0, s_eps = 0, dmax = 0;
#pragma unroll
    for (int k = 0; k < E; ++k) z[k] = pack_rows<T, N>(W(0), W(1), j + k * P);
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const int m = j + k * P;
        const int h = m <= N / 2 ? m : N - m;
        w[k] = pack_vals<T, N>(W(2)[h], d1(W(0), h), m);
    }
    {
        // the two transforms in lockstep through the row's LDS as two split lines
        T* const tl = (T*)(lds + l * lds_line_cx(N));
        const LdsLine<T, true, (P <= 64)> La{tl, 1}, Lb{tl + lds_line(N), 1};
        line_fft2<T, N, E, true, (P <= 64)>(z, w, j, La, Lb, a.tw);  // (u1, u2), (u3, d1 u1)
    }
#pragma unroll
    for (int k = 0; k < E; ++k) { u3[k] = w[k].x; b11[k] = w[k].y; }
    // the symmetric pairs b_ij, b_ji: i j = 1 2, 1 3, 2 3
#pragma unroll
    for (int pr = 0; pr < 3; ++pr) {
        __builtin_amdgcn_sched_barrier(0);
        const Row X = W(pr == 0 ? 3 : (pr == 1 ? 6 : 7));  // d2 u1, d3 u1, d3 u2
        const Row Y = W(pr == 0 ? 1 : (pr == 1 ? 2 : 5));  // (d1 of) u2, u3; d2 u3
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int m = j + k * P;
            const int h = m <= N / 2 ? m : N - m;
            w[k] = pack_vals<T, N>(X[h], pr < 2 ? d1(Y, h) : Y[h], m);
        }
        line_fft<T, N, E, true, false>(w, j, Ls, a.tw);
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const double uu = (double)z[k].x * z[k].x + (double)z[k].y * z[k].y + (double)u3[k] * u3[k];
            const double bij = w[k].x, bji = w[k].y;
            s_ijji += 2.0 * bij * bji * uu;
            s_eps += nu * ((bij - bji) * (bij - bji));
        }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < E; ++k) w[k] = pack_rows<T, N>(W(4), W(8), j + k * P);
    line_fft<T, N, E, true, false>(w, j, Ls, a.tw);  // d2 u2, d3 u3
    T th[E];  // theta at the thread's points, for the dissipation moment
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const double u1 = z[k].x, u2 = z[k].y, uu3 = u3[k];
        const double b11k = b11[k], b22 = w[k].x, b33 = w[k].y;
        const double uu = u1 * u1 + u2 * u2 + uu3 * uu3;
        const double div = b11k + b22 + b33;
        th[k] = (T)div;
        s_um2 += div * div * uu;
        s_ijji += (b11k * b11k + b22 * b22 + b33 * b33) * uu;
        s_u2t += div * uu;
        s_eps += nu * div * div;
        dmax = fmax(dmax, fabs(div));
        s_urms += u1 * u1 + u2 * u2 + uu3 * uu3;
        s_dudx += b11k * b11k + b22 * b22 + b33 * b33;
    }
    {
        double eta = fmin(eta0, sqrt(nu / dmax));
        if (!valid) { s_um2 = s_ijji = s_u2t = s_urms = s_dudx = s_eps = 0; eta = eta0; }
        double v0[3] = {s_um2, s_ijji, s_u2t};
        double v1[4] = {s_urms, s_dudx, s_eps, eta};
        k2d::vg_reduce<C::THREADS, 3>(v0, red, 0, a.part);
        k2d::vg_reduce<C::THREADS, 4>(v1, red, 4, a.part);
    }
    // the dissipation moment, sum 2 theta u_i Lap(u_i) + |u|^2 Lap(theta), from
    // u and theta above and two more x transforms: (Lap u1, Lap u2) and
    // (Lap u3, Lap theta), with Lap(u_c) = L_c (the column passes' full
    // Laplacian) and Lap(theta) = i k1 L_1 + D: rows not read before
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < E; ++k) {
        if (k % (E / 2) == 0) __builtin_amdgcn_sched_barrier(0);
        w[k] = pack_rows<T, N>(W(9), W(10), j + k * P);
    }
    __builtin_amdgcn_sched_barrier(0);
    line_fft<T, N, E, true, false>(w, j, Ls, a.tw);  // Lap(u1), Lap(u2)
    double s_dis = 0;
#pragma unroll
    for (int k = 0; k < E; ++k)
        s_dis += 2.0 * (double)z[k].x * (double)th[k] * (double)w[k].x +
                 2.0 * (double)z[k].y * (double)th[k] * (double)w[k].y;
    auto L1 = Row(k2d::launder(a.w[9]), ro, lim);  // read again, from the L2
#pragma unroll
    for (int k = 0; k < E; ++k) {
        if (k % (E / 2) == 0) __builtin_amdgcn_sched_barrier(0);
        const int m = j + k * P;
        const int h = m <= N / 2 ? m : N - m;
        const T k1o = (h == N / 2) ? (T)0 : (T)h;
        const cx<T> lth = cadd(cscale(cmuli(L1[h]), k1o), W(12)[h]);
        w[k] = pack_vals<T, N>(W(11)[h], lth, m);
    }
    __builtin_amdgcn_sched_barrier(0);
    line_fft<T, N, E, true, false>(w, j, Ls, a.tw);  // Lap(u3), Lap(theta)
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const double u1 = z[k].x, u2 = z[k].y, uu3 = u3[k];
        const double uu = u1 * u1 + u2 * u2 + uu3 * uu3;
        s_dis += 2.0 * uu3 * (double)th[k] * (double)w[k].x + uu * (double)w[k].y;
    }
    if (!valid) s_dis = 0;
    double vd[1] = {s_dis};
    k2d::vg_reduce<C::THREADS, 1>(vd, red, 3, a.part);
}

// shell spectra, 3-D (spectra_compute3D :1018-1108): u_d = (u.k)/|k|,
// u_s = u - u_d k/|k|; same deterministic per-wave binning as in 2-D.
template <typename T>
__global__ __launch_bounds__(256) void spec3_kernel(Spec3Args<T> a) {
    if (a.gate.closed()) return;
    extern __shared__ double sh[];
    const int K = a.allkmax, nb = K + 1;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    double* bins = sh;
    double* tots = sh + 4 * nb * 2;
    for (int i = tid; i < 4 * nb * 2; i += 256) bins[i] = 0;
    __syncthreads();
    const int n0 = a.g.n0, n1 = a.g.n1, n2 = a.g.n2;
    const int nh = a.nkx > 0 ? a.nkx : a.g.nh;  // kx limit of the modes summed
    const int ny = a.nky > 0 ? a.nky : n1, nz = a.nkz > 0 ? a.nkz : n2;
    const size_t hp = (size_t)a.g.hp;
    const int nbx = (nh + 63) / 64;
    const int ntask = ny * nz * nbx;  // < 2^31 (launch_spec3 checks)
    const int tstride = 4 * gridDim.x;
    double tot[4] = {0, 0, 0, 0};
    // the fields through buffer resources with 32-bit offsets (launch_spec3
    // checks that a field fits 2^31 bytes)
    __amdgpu_buffer_rsrc_t rU[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
        rU[c] = __builtin_amdgcn_make_buffer_rsrc((void*)a.U[c], (short)0, (int)k2d::kBufRange, k2d::kBufWord3);
    constexpr int TB = 4;  // tasks per wave iteration: all loads issued before the arithmetic
    for (int t = blockIdx.x * 4 + wave; t < ntask; t += tstride * TB) {
        cx<T> u[TB][3];
        int ky_[TB], kz_[TB], kx_[TB];  // each task's indices, formed once (32-bit)
#pragma unroll
        for (int q = 0; q < TB; ++q) {
            const int tq = t + tstride * q;
            const int r0 = tq / nbx;
            int iy = r0 % ny, iz = r0 / ny;
            if (iy >= a.ky_at) iy += a.ky_skip;
            if (iz >= a.kz_at) iz += a.kz_skip;
            const int kx = (tq - r0 * nbx) * 64 + lane;
            ky_[q] = ksigned(iy, n1);
            kz_[q] = ksigned(iz, n2);
            kx_[q] = kx;
            // lanes past the summed modes read zero through an offset past the
            // buffer's range: no branch around the loads, which then all issue
            // before the first is waited for
            const bool ok = tq < ntask && kx < nh;
            const unsigned o = ok ? (unsigned)((((size_t)iz * n1 + iy) * hp + kx) * sizeof(cx<T>)) : k2d::kBufRange;
#pragma unroll
            for (int c = 0; c < 3; ++c) u[q][c] = k2d::buf_load<T>(rU[c], o);
        }
#pragma unroll
        for (int q = 0; q < TB; ++q) {
            const int tq = t + tstride * q;
            if (tq >= ntask) break;  // wave-uniform
            const int kx = kx_[q];
            double es = 0, ed = 0, w = 0;
            int s = INT_MAX;
            if (kx < nh) {
                const int ky = ky_[q], kz = kz_[q];
                w = (kx == 0 || kx == n0 / 2) ? 1.0 : 2.0;
                const double kk = sqrt((double)(kx * kx + ky * ky + kz * kz));
                if (kk > 0.5) {
                    const cx<T> u1 = u[q][0], u2 = u[q][1], u3 = u[q][2];
                    const double ik = 1.0 / kk;
                    const double c1 = kx * ik, c2 = ky * ik, c3 = kz * ik;
                    const double udx = (double)u1.x * c1 + (double)u2.x * c2 + (double)u3.x * c3;
                    const double udy = (double)u1.y * c1 + (double)u2.y * c2 + (double)u3.y * c3;
                    const double s1x = u1.x - udx * c1, s1y = u1.y - udy * c1;
                    const double s2x = u2.x - udx * c2, s2y = u2.y - udy * c2;
                    const double s3x = u3.x - udx * c3, s3y = u3.y - udy * c3;
                    es = (s1x * s1x + s1y * s1y) * 0.5 + (s2x * s2x + s2y * s2y) * 0.5 + (s3x * s3x + s3y * s3y) * 0.5;
                    ed = (udx * udx + udy * udy) * 0.5;
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

template <typename T>
__global__ __launch_bounds__(256) void energy3_kernel(Energy3Args<T> a) {
    if (a.gate.closed()) return;
    __shared__ double red[4];
    const int n0 = a.g.n0, nh = a.g.nh;
    const size_t hp = (size_t)a.g.hp;
    const size_t total = (size_t)nh * a.g.n1 * a.g.n2;
    double e = 0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const int kx = (int)(i % nh);
        const size_t r = i / nh;
        const double w = (kx == 0 || kx == n0 / 2) ? 1.0 : 2.0;
        double s = 0;
        for (int c = 0; c < 3; ++c) {
            const cx<T> u = a.U[c][r * hp + kx];
            s += (double)u.x * u.x + (double)u.y * u.y;
        }
        e += w * s;
    }
    e = wave_sum(e);
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = e;
    __syncthreads();
    if (threadIdx.x == 0) a.part_en[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// forcing3D methods 1, 2, 5 (solution.F90:554-629, 754-813); the 3-D method 5
// has no projected variant.
template <typename T>
__global__ __launch_bounds__(256) void force3_kernel(Force3Args<T> a) {
    if (a.gate.closed()) return;
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
    const int n0 = a.g.n0, n1 = a.g.n1, n2 = a.g.n2;
    const size_t hp = (size_t)a.g.hp;
    const int fk = a.forcek;
    const int kxm = fk < n0 / 2 ? fk : n0 / 2;
    const int kym = fk < n1 / 2 ? fk : n1 / 2 - 1;
    const int kzm = fk < n2 / 2 ? fk : n2 / 2 - 1;
    const int nx = kxm + 1, ny = 2 * kym + 1, nz = 2 * kzm + 1;
    const int ntot = nx * ny * nz;
    double e = 0;
    for (int t = tid; t < ntot; t += 256) {
        const int kx = t % nx, ky = (t / nx) % ny - kym, kz = t / (nx * ny) - kzm;
        const double kk = sqrt((double)(kx * kx + ky * ky + kz * kz));
        if ((kk - 0.5) < fk && (kk + 0.5) > fk) {
            const size_t o = ((size_t)(kz >= 0 ? kz : n2 + kz) * n1 + (ky >= 0 ? ky : n1 + ky)) * hp + kx;
            const double w = (kx == 0 || kx == n0 / 2) ? 1.0 : 2.0;
            double s = 0;
            for (int c = 0; c < 3; ++c) {
                const cx<T> u = a.U[c][o];
                s += (double)u.x * u.x + (double)u.y * u.y;
            }
            e += w * s;
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
    const T f = (T)fac;
    for (int t = tid; t < ntot; t += 256) {
        const int kx = t % nx, ky = (t / nx) % ny - kym, kz = t / (nx * ny) - kzm;
        const double kk = sqrt((double)(kx * kx + ky * ky + kz * kz));
        if ((kk - 0.5) < fk && (kk + 0.5) > fk) {
            const size_t o = ((size_t)(kz >= 0 ? kz : n2 + kz) * n1 + (ky >= 0 ? ky : n1 + ky)) * hp + kx;
            for (int c = 0; c < 3; ++c) {
                const cx<T> u = a.U[c][o];
                a.U[c][o] = cadd(u, cscale(u, f));
            }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void scale3_kernel(Scale3Args<T> a) {
    if (a.gate.closed()) return;
    const T f = (T)*a.factor;
    const size_t total = (size_t)a.g.hp * a.g.n1 * a.g.n2;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256)
        for (int c = 0; c < 3; ++c) a.U[c][i] = cscale(a.U[c][i], f);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define K3D_SIZES(X) X(16) X(32) X(64) X(128) X(256) X(512)

// one single-term variant: (EM points per thread, CBM times the block width)
template <typename T, int N, int EM, int CBM>
static void colx1_launch(bool inverse, dim3 grid, const ColXArgs<T>& b, hipStream_t s) {
    using C = ColxCfg<T, N, 1, EM, CBM>;
    if constexpr (C::FITS && (EM == 16 || N >= 256)) {
        if (inverse) hipLaunchKernelGGL((colx_kernel<T, N, true, 1, EM, CBM>), grid, dim3(C::THREADS), 0, s, b);
        else hipLaunchKernelGGL((colx_kernel<T, N, false, 1, EM, CBM>), grid, dim3(C::THREADS), 0, s, b);
    }
}

template <typename T, int N>
static int colx_n(bool inverse, const ColXArgs<T>& a, hipStream_t s) {
    using C4 = ColxCfg<T, N, 4>;
    const int prec = sizeof(T) == 4 ? 1 : 0;
    int em = colx_points(N, prec), cbm = colx_width(N, prec);
    if (!colx_plan_ok(N, prec, em, cbm)) em = 16, cbm = 1;
    auto cb_of = [](int e, int m) {
        return e == 8 ? (m == 2 ? ColxCfg<T, N, 1, 8, 2>::CB : ColxCfg<T, N, 1, 8, 1>::CB)
                      : (m == 2 ? ColxCfg<T, N, 1, 16, 2>::CB : ColxCfg<T, N, 1, 16, 1>::CB);
    };
    if (cbm == 2 && a.ncols % cb_of(em, 2)) cbm = 1;  // the wide tile needs whole blocks: the default width
    const int cb1 = cb_of(em, cbm);
    if (a.ncols % cb1 || a.ncols % C4::CB || a.njob < 1 || a.njob > 13) return -1;
    // a block's entries sit within 32-bit byte offsets of its first (colx_kernel)
    if ((long)N * a.pitch * (long)sizeof(cx<T>) >= (long)k2d::kBufRange) return -1;
    // single-term and multi-term jobs go to the two kernel variants
    ColXArgs<T> part[2] = {a, a};
    part[0].njob = part[1].njob = 0;
    for (int q = 0; q < a.njob; ++q) {
        if (a.job[q].nterm < 1 || a.job[q].nterm > 4) return -1;
        if (!inverse && a.job[q].nterm == 1 && a.job[q].mult[0] != M_ONE) return -1;
        ColXArgs<T>& d = part[a.job[q].nterm > 1 ? 1 : 0];
        d.job[d.njob++] = a.job[q];
    }
    for (int v = 0; v < 2; ++v) {
        const ColXArgs<T>& b = part[v];
        if (!b.njob) continue;
        const int cb = v == 0 ? cb1 : C4::CB;
        const long nbl = (long)b.njob * (b.ncols / cb) * b.nplanes;
        if (nbl > INT_MAX - 8) return -1;
        dim3 grid(k2d::xcd_grid_h((int)nbl), 1, k2d::zdim(b.nmem)), block(C4::THREADS);
        if (v == 0) {
            if (em == 8 && cbm == 2) colx1_launch<T, N, 8, 2>(inverse, grid, b, s);
            else if (em == 8) colx1_launch<T, N, 8, 1>(inverse, grid, b, s);
            else if (cbm == 2) colx1_launch<T, N, 16, 2>(inverse, grid, b, s);
            else colx1_launch<T, N, 16, 1>(inverse, grid, b, s);
        } else {
            if (inverse) hipLaunchKernelGGL((colx_kernel<T, N, true, 4>), grid, block, 0, s, b);
            else hipLaunchKernelGGL((colx_kernel<T, N, false, 4>), grid, block, 0, s, b);
        }
    }
    return 0;
}

template <typename T, int N>
static int row3_n(int mode, const Row3Args<T>& a, int nrows, hipStream_t s) {
    using C = RowCfg<T, N>;
    dim3 grid((nrows + C::L - 1) / C::L), block(C::THREADS);
    switch (mode) {
        case R3_FWD_PACK: hipLaunchKernelGGL((row3_kernel<T, N, R3_FWD_PACK>), grid, block, 0, s, a); break;
        case R3_INV_PACK: hipLaunchKernelGGL((row3_kernel<T, N, R3_INV_PACK>), grid, block, 0, s, a); break;
        case R3_STAGE: hipLaunchKernelGGL((row3_kernel<T, N, R3_STAGE>), grid, block, 0, s, a); break;
        default: return -1;
    }
    return 0;
}

template <typename T, int N>
static int vg3_n(const Vg3Args<T>& a, int nrows, hipStream_t s, int nblk) {
    using C = Vg3Cfg<T, N>;
    const int g = (nrows + C::L - 1) / C::L;
    if (g != nblk) return -1;
    hipLaunchKernelGGL((vg3_kernel<T, N>), dim3(g), dim3(C::THREADS), 0, s, a);
    return 0;
}

// the column pass also serves the 2-D step, up to 4096 rows
#define COLX_SIZES(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)
template <typename T>
int launch_colx(int n, bool inverse, const ColXArgs<T>& a, hipStream_t s) {
    switch (n) {
#define X(NN) case NN: return colx_n<T, NN>(inverse, a, s);
        COLX_SIZES(X)
#undef X
    }
    return -1;
}
template <typename T, int N>
static int cf3_n(int mode, const Cf3Args<T>& a, hipStream_t s) {
    using C = Fwd3Cfg<T, N>;
    const int ncols = a.g.hp * a.g.n1;
    const int last_row = a.nky - 1 >= a.ky_skip_at ? a.nky - 1 + a.ky_skip : a.nky - 1;
    if (a.nkx < C::CB || a.nkx % C::CB || a.nky < 1 || a.ky_skip < 0 ||
        (long)last_row * a.g.hp + a.nkx > ncols || (a.nky > 1 && a.nkx > a.g.hp))
        return -1;
    dim3 grid(k2d::xcd_grid_h(a.nky * (a.nkx / C::CB))), block(C::THREADS);
    switch (mode) {
#define M(MM) case MM: hipLaunchKernelGGL((cf3_kernel<T, N, MM>), grid, block, 0, s, a); break;
        M(C3_PROLOGUE) M(C3_A) M(C3_B) M(C3_C) M(C3_CN_MID) M(C3_CN_LAST)
#undef M
        default: return -1;
    }
    return 0;
}
template <typename T, int N>
static int cf2_n(int mode, const k2d::ColFwdArgs<T>& a, hipStream_t s) {
    using C = Fwd2Cfg<T, N>;
    if (a.nkx < C::CB || a.nkx % C::CB || a.nkx > a.pitch || a.g.n1 != N) return -1;
    // a block's entries sit within 32-bit byte offsets of its first (cf2_kernel)
    if ((long)N * a.pitch * (long)sizeof(cx<T>) >= (long)k2d::kBufRange) return -1;
    dim3 grid(k2d::xcd_grid_h(a.nkx / C::CB), 1, k2d::zdim(a.g.nmem)), block(C::THREADS);
    switch (mode) {
#define M(MM) case MM: hipLaunchKernelGGL((cf2_kernel<T, N, MM>), grid, block, 0, s, a); break;
        M(k2d::UP_A) M(k2d::UP_B) M(k2d::UP_C) M(k2d::UP_CN_MID) M(k2d::UP_CN_LAST)
#undef M
        default: return -1;
    }
    return 0;
}
template <typename T>
bool cf2_full_rows(int n) {
    switch (n) {
#define X(NN) case NN: return Fwd2Cfg<T, NN>::CB == Fwd2Cfg<T, NN>::Base::CB || Fwd2Cfg<T, NN>::NARROW;
        COLX_SIZES(X)
#undef X
    }
    return false;
}
template <typename T>
int launch_cf2(int n, int mode, const k2d::ColFwdArgs<T>& a, hipStream_t s) {
    switch (n) {
#define X(NN) case NN: return cf2_n<T, NN>(mode, a, s);
        COLX_SIZES(X)
#undef X
    }
    return -1;
}
template <typename T>
int launch_cf3(int n, int mode, const Cf3Args<T>& a, hipStream_t s) {
    switch (n) {
#define X(NN) case NN: return cf3_n<T, NN>(mode, a, s);
        K3D_SIZES(X)
#undef X
    }
    return -1;
}
template <typename T>
int launch_row3(int n, int mode, const Row3Args<T>& a, int nrows, hipStream_t s) {
    switch (n) {
#define X(NN) case NN: return row3_n<T, NN>(mode, a, nrows, s);
        K3D_SIZES(X)
#undef X
    }
    return -1;
}
template <typename T>
int launch_vg3(int n, const Vg3Args<T>& a, int nrows, hipStream_t s, int nblk) {
    switch (n) {
#define X(NN) case NN: return vg3_n<T, NN>(a, nrows, s, nblk);
        K3D_SIZES(X)
#undef X
    }
    return -1;
}
template <typename T>
int launch_spec3(const Spec3Args<T>& a, int nblk, hipStream_t s) {
    {  // the kernel's task index is 32-bit, its field offsets 32-bit bytes
        const long nh = a.nkx > 0 ? a.nkx : a.g.nh, ny = a.nky > 0 ? a.nky : a.g.n1, nz = a.nkz > 0 ? a.nkz : a.g.n2;
        if (ny * nz * ((nh + 63) / 64) + 4L * nblk * 4 >= INT_MAX) return -1;
        if ((long)a.g.hp * a.g.n1 * a.g.n2 * (long)sizeof(cx<T>) >= (long)k2d::kBufRange) return -1;
    }
    const size_t shm = sizeof(double) * (4 * (size_t)(a.allkmax + 1) * 2 + 16);
    hipLaunchKernelGGL((spec3_kernel<T>), dim3(nblk), dim3(256), shm, s, a);
    return 0;
}
template <typename T>
int launch_force3(const Force3Args<T>& a, hipStream_t s) {
    hipLaunchKernelGGL((force3_kernel<T>), dim3(1), dim3(256), 0, s, a);
    return 0;
}
template <typename T>
int launch_energy3(const Energy3Args<T>& a, int nblk, hipStream_t s) {
    hipLaunchKernelGGL((energy3_kernel<T>), dim3(nblk), dim3(256), 0, s, a);
    return 0;
}
template <typename T>
int launch_scale3(const Scale3Args<T>& a, hipStream_t s) {
    hipLaunchKernelGGL((scale3_kernel<T>), dim3(2048), dim3(256), 0, s, a);
    return 0;
}

#define K3D_INSTANTIATE(T)                                                             \
    template int launch_colx<T>(int, bool, const ColXArgs<T>&, hipStream_t);           \
    template int launch_cf2<T>(int, int, const k2d::ColFwdArgs<T>&, hipStream_t);      \
    template bool cf2_full_rows<T>(int);                                               \
    template int launch_cf3<T>(int, int, const Cf3Args<T>&, hipStream_t);              \
    template int launch_row3<T>(int, int, const Row3Args<T>&, int, hipStream_t);       \
    template int launch_vg3<T>(int, const Vg3Args<T>&, int, hipStream_t, int);         \
    template int launch_spec3<T>(const Spec3Args<T>&, int, hipStream_t);               \
    template int launch_force3<T>(const Force3Args<T>&, hipStream_t);                  \
    template int launch_energy3<T>(const Energy3Args<T>&, int, hipStream_t);           \
    template int launch_scale3<T>(const Scale3Args<T>&, hipStream_t);

}  // namespace k3d
