// fft_block.hpp -- workgroup-level Stockham FFT primitives for gfx950.
//
// A line of N complex points is held by P = N/E threads, E points each, in
// registers: thread j holds x[j + k*P] for k < E ("strided" order).  Each
// radix pass is a butterfly on the thread's own registers; between passes the
// line is exchanged through LDS (optionally in two halves, real then
// imaginary, to halve the LDS footprint).  After the last pass thread j holds
// X[j + k*P] in natural order, so loads before and stores after the transform
// are coalesced along the line.
//
// Pass p of radix R with Ns = product of earlier radices (Stockham autosort):
//   butterfly b reads x[b + r*N/R], twiddles by w_N^{(b mod Ns)*r*N/(Ns*R)},
//   does a radix-R DFT and writes y[(b/Ns)*Ns*R + (b mod Ns) + r*Ns].
// Radices are E for all passes but possibly the last (N = E^a * L, L | E).
// Twiddles come from an FP64-accurate table tw[m] = exp(-2*pi*i*m/N) built on
// the host; the inverse direction conjugates them.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <climits>

template <typename T>
struct cx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return {a.x * s, a.y * s}; }
template <typename T>
__device__ __forceinline__ cx<T> cconj(cx<T> a) { return {a.x, -a.y}; }
// i*a
template <typename T>
__device__ __forceinline__ cx<T> cmuli(cx<T> a) { return {-a.y, a.x}; }

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }

// cos/sin(2*pi*t/16), t = 0..15
__device__ constexpr double kC16[16] = {
    1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
    0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
    -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
    0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
__device__ constexpr double kS16[16] = {
    0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
    1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
    0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
    -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173};

// a * exp(sgn*2*pi*i*t/16) with sgn = -1 (forward) or +1 (inverse); t known at
// compile time after unrolling, so the trivial cases fold away.
template <typename T, bool INV>
__device__ __forceinline__ cx<T> rot16(cx<T> a, int t) {
    t &= 15;
    if (t == 0) return a;
    if (t == 8) return {-a.x, -a.y};
    if (t == 4) return INV ? cx<T>{-a.y, a.x} : cx<T>{a.y, -a.x};
    if (t == 12) return INV ? cx<T>{a.y, -a.x} : cx<T>{-a.y, a.x};
    const T c = (T)kC16[t];
    const T s = INV ? (T)kS16[t] : -(T)kS16[t];
    if (t == 2 || t == 6 || t == 10 || t == 14) {
        // |c| == |s| == sqrt(1/2): (x + iy)(sc + i ss) h with sc, ss = +-1
        const T h = (T)0.70710678118654752440;
        const T sc = c > 0 ? (T)1 : (T)-1, ss = s > 0 ? (T)1 : (T)-1;
        return {(sc * a.x - ss * a.y) * h, (ss * a.x + sc * a.y) * h};
    }
    return {a.x * c - a.y * s, a.x * s + a.y * c};
}

constexpr int bitrev(int x, int bits) {
    int r = 0;
    for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1) << (bits - 1 - i);
    return r;
}

// In-register radix-R DFT (R = 2, 4, 8, 16): X[k] = sum_r v[r] w_R^{+-rk}.
template <typename T, int R, bool INV>
__device__ __forceinline__ void dft(cx<T>* v) {
    constexpr int B = ilog2(R);
    cx<T> a[R];
#pragma unroll
    for (int i = 0; i < R; ++i) a[i] = v[bitrev(i, B)];
#pragma unroll
    for (int len = 2; len <= R; len <<= 1) {
#pragma unroll
        for (int i = 0; i < R; i += len) {
#pragma unroll
            for (int j = 0; j < len / 2; ++j) {
                const cx<T> u = a[i + j];
                const cx<T> w = rot16<T, INV>(a[i + j + len / 2], j * (16 / len));
                a[i + j] = cadd(u, w);
                a[i + j + len / 2] = csub(u, w);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < R; ++i) v[i] = a[i];
}

// Per-line LDS view.  SPLIT: real and imaginary parts exchanged in two rounds
// through one T buffer of the line's length; otherwise one cx<T> buffer.
// WAVE: every lane of the line is in one wavefront (lines of at most 64
// threads, not straddling a wave), so exchanges need no workgroup barrier.
template <typename T, bool SPLIT, bool WAVE = false>
struct LdsLine;

// Ordering of the LDS accesses of one exchange: a workgroup barrier, or for
// wave-private lines only a compiler barrier with wavefront-scope fences (a
// wavefront's LDS operations complete in issue order).
template <bool WAVE>
__device__ __forceinline__ void lds_sync() {
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// Element e of a line lives at padded slot e + e/16 (split lines, 4- or
// 8-byte elements) or e + e/8 (complex 8- or 16-byte elements): the spare
// slots break the power-of-two strides of the Stockham write pattern, which
// otherwise land up to 8 lanes of a pass on the same LDS banks.
__device__ __forceinline__ int lds_slot(int e) { return e + (e >> 4); }
__device__ __forceinline__ int lds_slot_cx(int e) { return e + (e >> 3); }
// LDS elements needed per line of length n
constexpr int lds_line(int n) { return n + n / 16; }
constexpr int lds_line_cx(int n) { return n + n / 8; }

template <typename T, bool WAVE>
struct LdsLine<T, false, WAVE> {
    cx<T>* base;
    int stride;  // element e at base[lds_slot_cx(e)*stride]
    __device__ __forceinline__ void st(int e, cx<T> v) const { base[lds_slot_cx(e) * stride] = v; }
    __device__ __forceinline__ cx<T> ld(int e) const { return base[lds_slot_cx(e) * stride]; }
};

template <typename T, bool WAVE>
struct LdsLine<T, true, WAVE> {
    T* base;
    int stride;
    __device__ __forceinline__ T& at(int e) const { return base[lds_slot(e) * stride]; }
};

// twiddle v[r] *= w_N^{(b mod Ns) * r * N/(Ns*R)}, r = 1..R-1.  w^r is the
// running product with w, refreshed from the table at every 4th power, so
// each factor is at most three roundings away from a table value while only
// two twiddles are live at a time.
template <typename T, int N, int R, bool INV>
__device__ __forceinline__ void twiddle(cx<T>* v, int b, int Ns, const cx<T>* __restrict__ tw) {
    if (Ns == 1) return;
    const int bm = b & (Ns - 1);
    const int step = bm * (N / (Ns * R));
    cx<T> w1 = tw[step];
    if (INV) w1.y = -w1.y;
    cx<T> t = w1;
#pragma unroll
    for (int r = 1; r < R; ++r) {
        if (r % 4 == 0) {
            t = tw[step * r];
            if (INV) t.y = -t.y;
        } else if (r > 1) {
            t = cmul(t, w1);
        }
        v[r] = cmul(v[r], t);
    }
}

// Exchanges with one address register per direction.  An exchange writes
// elements b + d(r) and reads c + d'(r), with b, c known per thread and the
// d(r) known at compile time.  With slot(e) = e + (e >> s) (s = 3 complex,
// 4 split), slot(b + d) = slot(b) + d + (d >> s) whenever
// (b & (2^s - 1)) + (d & (2^s - 1)) < 2^s: then every access of the exchange
// is one base address plus an immediate offset.  lin_write/lin_read check that
// condition at compile time over every thread of the line; where it fails the
// exchange computes each slot.
constexpr int lin_off(int d, int s) { return d + (d >> s); }
constexpr bool lin_fits(int b, int d, int s) { return (b & ((1 << s) - 1)) + (d & ((1 << s) - 1)) < (1 << s); }
// Stockham write of a pass with Ns = ns, radix R: b = (j/ns)*ns*R + j%ns, d = r*ns
constexpr bool lin_write(int P, int ns, int R, int s) {
    for (int j = 0; j < P; ++j)
        for (int r = 0; r < R; ++r)
            if (!lin_fits((j / ns) * ns * R + (j & (ns - 1)), r * ns, s)) return false;
    return true;
}
// read of element j + d(q): d = q*P (full passes), or (q/Lr)*P + (q%Lr)*(N/Lr)
// ahead of a partial last pass of radix Lr
constexpr bool lin_read(int N, int P, int E, int Lr, int s) {
    for (int j = 0; j < P; ++j)
        for (int q = 0; q < E; ++q)
            if (!lin_fits(j, Lr == 1 ? q * P : (q / Lr) * P + (q % Lr) * (N / Lr), s)) return false;
    return true;
}

// exchange of pass-p data: writes b + r*NS, reads j + d(q) as above
template <typename T, int N, int E, int NS, int LR, bool WAVE>
__device__ __forceinline__ void exchange_pass(const LdsLine<T, false, WAVE>& L, cx<T>* v, int b, int j) {
    constexpr int P = N / E, S = 3;
    constexpr bool LW = lin_write(P, NS, E, S), LRD = lin_read(N, P, E, LR, S);
    auto rd = [](int q) { return LR == 1 ? q * P : (q / LR) * P + (q % LR) * (N / LR); };
    lds_sync<WAVE>();
    if constexpr (LW) {
        cx<T>* const w = L.base + lds_slot_cx(b) * L.stride;
#pragma unroll
        for (int r = 0; r < E; ++r) w[lin_off(r * NS, S) * L.stride] = v[r];
    } else {
#pragma unroll
        for (int r = 0; r < E; ++r) L.st(b + r * NS, v[r]);
    }
    lds_sync<WAVE>();
    if constexpr (LRD) {
        const cx<T>* const q0 = L.base + lds_slot_cx(j) * L.stride;
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = q0[lin_off(rd(q), S) * L.stride];
    } else {
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = L.ld(j + rd(q));
    }
}

template <typename T, int N, int E, int NS, int LR, bool WAVE>
__device__ __forceinline__ void exchange_pass(const LdsLine<T, true, WAVE>& L, cx<T>* v, int b, int j) {
    constexpr int P = N / E, S = 4;
    constexpr bool LW = lin_write(P, NS, E, S), LRD = lin_read(N, P, E, LR, S);
    auto rd = [](int q) { return LR == 1 ? q * P : (q / LR) * P + (q % LR) * (N / LR); };
    T* const w = L.base + lds_slot(b) * L.stride;
    T* const q0 = L.base + lds_slot(j) * L.stride;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        lds_sync<WAVE>();
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const T x = h ? v[r].y : v[r].x;
            if constexpr (LW) w[lin_off(r * NS, S) * L.stride] = x;
            else L.at(b + r * NS) = x;
        }
        lds_sync<WAVE>();
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const T x = LRD ? q0[lin_off(rd(q), S) * L.stride] : L.at(j + rd(q));
            if (h) v[q].y = x;
            else v[q].x = x;
        }
    }
}

template <int N, int E>
struct FftShape {
    static_assert((N & (N - 1)) == 0 && N >= 4, "N must be a power of two");
    static_assert(E == 4 || E == 8 || E == 16, "E must be 4, 8 or 16");
    static constexpr int P = N / E;          // threads per line
    static constexpr int NFULL = ilog2(N) / ilog2(E);
    static constexpr int LAST = 1 << (ilog2(N) % ilog2(E));  // 1: no partial pass
    static_assert(NFULL >= 1, "N must be >= E");
};

constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

// Full radix-E passes p, p+1, ... of a line transform (Ns = E^p), each
// followed by its exchange unless it is the last pass.
template <typename T, int N, int E, bool INV, bool SPLIT, bool WAVE, int p>
__device__ __forceinline__ void fft_passes(cx<T>* v, int j, const LdsLine<T, SPLIT, WAVE>& L,
                                           const cx<T>* __restrict__ tw) {
    using S = FftShape<N, E>;
    constexpr int NS = ipow(E, p);
    twiddle<T, N, E, INV>(v, j, NS, tw);
    dft<T, E, INV>(v);
    if constexpr (p < S::NFULL - 1 || S::LAST > 1) {
        const int b = (j / NS) * NS * E + (j & (NS - 1));
        exchange_pass<T, N, E, NS, (p < S::NFULL - 1 ? 1 : S::LAST)>(L, v, b, j);
    }
    if constexpr (p + 1 < S::NFULL) fft_passes<T, N, E, INV, SPLIT, WAVE, p + 1>(v, j, L, tw);
}

// Full line transform.  On entry v[k] = x[j + k*P]; on exit v[k] = X[j + k*P].
// Unnormalised in both directions.
template <typename T, int N, int E, bool INV, bool SPLIT, bool WAVE>
__device__ __forceinline__ void line_fft(cx<T>* v, int j, const LdsLine<T, SPLIT, WAVE>& L,
                                         const cx<T>* __restrict__ tw) {
    using S = FftShape<N, E>;
    constexpr int P = S::P;
    fft_passes<T, N, E, INV, SPLIT, WAVE, 0>(v, j, L, tw);
    if constexpr (S::LAST > 1) {
        constexpr int Ns = ipow(E, S::NFULL);
        constexpr int Lr = S::LAST;
        constexpr int Q = E / Lr;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            twiddle<T, N, Lr, INV>(v + q * Lr, j + q * P, Ns, tw);
            dft<T, Lr, INV>(v + q * Lr);
        }
        cx<T> t[E];
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int r = 0; r < Lr; ++r) t[q + r * Q] = v[q * Lr + r];
#pragma unroll
        for (int k = 0; k < E; ++k) v[k] = t[k];
    }
}

// ---------------------------------------------------------------------------
// Two lines in lockstep (one direction): each pass's twiddle factors are formed
// once for both, and each exchange moves both lines through two T buffers of
// one line's length (real parts, then imaginary parts), the LDS footprint of
// one complex line, with both lines' loads and stores between two barriers.
// The arithmetic per line is line_fft's, so the results are its bit for bit.
// ---------------------------------------------------------------------------
template <typename T, int N, int R, bool INV>
__device__ __forceinline__ void twiddle2(cx<T>* v, cx<T>* u, int b, int Ns, const cx<T>* __restrict__ tw) {
    if (Ns == 1) return;
    const int bm = b & (Ns - 1);
    const int step = bm * (N / (Ns * R));
    cx<T> w1 = tw[step];
    if (INV) w1.y = -w1.y;
    cx<T> t = w1;
#pragma unroll
    for (int r = 1; r < R; ++r) {
        if (r % 4 == 0) {
            t = tw[step * r];
            if (INV) t.y = -t.y;
        } else if (r > 1) {
            t = cmul(t, w1);
        }
        v[r] = cmul(v[r], t);
        u[r] = cmul(u[r], t);
    }
}

template <typename T, int N, int E, int NS, int LR, bool WAVE>
__device__ __forceinline__ void exchange_pass2(const LdsLine<T, true, WAVE>& A, const LdsLine<T, true, WAVE>& B,
                                               cx<T>* v, cx<T>* u, int b, int j) {
    constexpr int P = N / E, S = 4;
    constexpr bool LW = lin_write(P, NS, E, S), LRD = lin_read(N, P, E, LR, S);
    auto rd = [](int q) { return LR == 1 ? q * P : (q / LR) * P + (q % LR) * (N / LR); };
    T* const wa = A.base + lds_slot(b) * A.stride;
    T* const wb = B.base + lds_slot(b) * B.stride;
    T* const qa = A.base + lds_slot(j) * A.stride;
    T* const qb = B.base + lds_slot(j) * B.stride;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        lds_sync<WAVE>();
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const T xa = h ? v[r].y : v[r].x, xb = h ? u[r].y : u[r].x;
            if constexpr (LW) {
                wa[lin_off(r * NS, S) * A.stride] = xa;
                wb[lin_off(r * NS, S) * B.stride] = xb;
            } else {
                A.at(b + r * NS) = xa;
                B.at(b + r * NS) = xb;
            }
        }
        lds_sync<WAVE>();
#pragma unroll
        for (int q = 0; q < E; ++q) {
            const T xa = LRD ? qa[lin_off(rd(q), S) * A.stride] : A.at(j + rd(q));
            const T xb = LRD ? qb[lin_off(rd(q), S) * B.stride] : B.at(j + rd(q));
            if (h) { v[q].y = xa; u[q].y = xb; }
            else { v[q].x = xa; u[q].x = xb; }
        }
    }
}

template <typename T, int N, int E, bool INV, bool WAVE, int p>
__device__ __forceinline__ void fft_passes2(cx<T>* v, cx<T>* u, int j, const LdsLine<T, true, WAVE>& A,
                                            const LdsLine<T, true, WAVE>& B, const cx<T>* __restrict__ tw) {
    using S = FftShape<N, E>;
    constexpr int NS = ipow(E, p);
    twiddle2<T, N, E, INV>(v, u, j, NS, tw);
    dft<T, E, INV>(v);
    dft<T, E, INV>(u);
    if constexpr (p < S::NFULL - 1 || S::LAST > 1) {
        const int b = (j / NS) * NS * E + (j & (NS - 1));
        exchange_pass2<T, N, E, NS, (p < S::NFULL - 1 ? 1 : S::LAST)>(A, B, v, u, b, j);
    }
    if constexpr (p + 1 < S::NFULL) fft_passes2<T, N, E, INV, WAVE, p + 1>(v, u, j, A, B, tw);
}

// line_fft of v and of u, in lockstep through the split buffers A and B
template <typename T, int N, int E, bool INV, bool WAVE>
__device__ __forceinline__ void line_fft2(cx<T>* v, cx<T>* u, int j, const LdsLine<T, true, WAVE>& A,
                                          const LdsLine<T, true, WAVE>& B, const cx<T>* __restrict__ tw) {
    using S = FftShape<N, E>;
    constexpr int P = S::P;
    fft_passes2<T, N, E, INV, WAVE, 0>(v, u, j, A, B, tw);
    if constexpr (S::LAST > 1) {
        constexpr int Ns = ipow(E, S::NFULL);
        constexpr int Lr = S::LAST;
        constexpr int Q = E / Lr;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            twiddle2<T, N, Lr, INV>(v + q * Lr, u + q * Lr, j + q * P, Ns, tw);
            dft<T, Lr, INV>(v + q * Lr);
            dft<T, Lr, INV>(u + q * Lr);
        }
        cx<T> t[E], w[E];
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int r = 0; r < Lr; ++r) {
                t[q + r * Q] = v[q * Lr + r];
                w[q + r * Q] = u[q * Lr + r];
            }
#pragma unroll
        for (int k = 0; k < E; ++k) {
            v[k] = t[k];
            u[k] = w[k];
        }
    }
}

// ---------------------------------------------------------------------------
// wavefront reductions (64 lanes), deterministic butterfly order
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}
__device__ __forceinline__ double wave_min(double x) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x = fmin(x, __shfl_xor(x, o, 64));
    return x;
}
// Deterministic shell binning for one wavefront.  The 64 lanes hold the
// modes kx = kx0 + lane of one row, whose shells s are non-decreasing in the
// lane index (|k| grows with kx >= 0), so every shell is one contiguous run of
// lanes (INT_MAX marks modes beyond the last shell, a run at the end).  A
// segmented inclusive scan (Hillis-Steele, fixed order) sums each run and the
// last lane of the run adds the sums to bins[2*s], bins[2*s + 1].
__device__ __forceinline__ void bin_runs(double* bins, int s, double a, double b, int lane) {
    if (__all(s == INT_MAX)) return;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int so = __shfl_up(s, o, 64);
        const double ao = __shfl_up(a, o, 64);
        const double bo = __shfl_up(b, o, 64);
        if (lane >= o && so == s) {
            a += ao;
            b += bo;
        }
    }
    const int sn = __shfl_down(s, 1, 64);
    if (s != INT_MAX && (lane == 63 || sn != s)) {
        bins[2 * s] += a;
        bins[2 * s + 1] += b;
    }
}
