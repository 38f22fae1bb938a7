// mlp.hip -- inference and training kernels of the per-point MLP closure
// (36 -> 64 -> 64 -> 2, GELU) on v_mfma_f32_32x32x8_bf16.  Layouts and
// arithmetic are described in mlp.hpp.
#include <math.h>

#include <algorithm>
#include <cmath>

#include "mlp.hpp"

namespace mlp {

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

// FP32 -> BF16, round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = (short)f2bf(a);
    r[1] = (short)f2bf(b);
    r[2] = (short)f2bf(c);
    r[3] = (short)f2bf(d);
    return r;
}

// GELU in its tanh form (Hendrycks and Gimpel, 2016):
//   gelu(z) = z (1 + tanh(u)) / 2,  u = sqrt(2/pi) (z + 0.044715 z^3),
//   gelu'(z) = (1 + tanh u) / 2 + z (1 - tanh^2 u) sqrt(2/pi) (1 + 3*0.044715 z^2) / 2,
// tanh u = 1 - 2 / (exp(2u) + 1), branch-free (exp overflows to inf -> tanh = 1)
__device__ __forceinline__ float gelu_tanh(float z) {
    const float u = 0.7978845608028654f * (z + 0.044715f * z * z * z);
    return 1.0f - __fdividef(2.0f, __expf(2.0f * u) + 1.0f);
}
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + gelu_tanh(z)); }
__device__ __forceinline__ float gelu_d(float z) {
    const float t = gelu_tanh(z);
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * 0.7978845608028654f * (1.0f + 0.134145f * z * z);
}

// hidden unit held by accumulator r of block blk in lane half h
__device__ __forceinline__ int arow(int blk, int r, int h) { return 32 * blk + 8 * (r >> 2) + 4 * h + (r & 3); }

__device__ __forceinline__ f32x16 mfma(bf16x4 a, bf16x4 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ bf16x4 ld4(const uint16_t* p) { return *(const bf16x4*)p; }

// The net pointer as an opaque value: loads through it stay where they are
// written instead of being hoisted out of the tile loop (where they would hold
// every bias and operand in registers for the whole loop)
__device__ __forceinline__ const Packed* opaque(const Packed* p) {
    asm volatile("" : "+s"(p));
    return p;
}

// ---------------------------------------------------------------------------
// packed operands from the FP32 master copy
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void pack_kernel(const float* __restrict__ p, Packed* __restrict__ o) {
    for (int i = threadIdx.x; i < 2 * KS1 * 64; i += blockDim.x) {
        const int lane = i & 63, j = (i >> 6) % KS1, blk = i / (64 * KS1);
        const int m = 32 * blk + (lane & 31);
        for (int t = 0; t < 4; ++t) {
            const int f = 8 * j + 4 * (lane >> 5) + t;
            o->A1[i * 4 + t] = f < NF ? f2bf(p[OFF_W1 + m * NF + f]) : 0;
        }
    }
    for (int i = threadIdx.x; i < 2 * KS2 * 64; i += blockDim.x) {
        const int lane = i & 63, j = (i >> 6) % KS2, blk = i / (64 * KS2);
        const int m = 32 * blk + (lane & 31);
        for (int t = 0; t < 4; ++t) {
            const int k = 8 * j + 4 * (lane >> 5) + t;
            o->A2[i * 4 + t] = f2bf(p[OFF_W2 + m * NH + k]);   // W2[h2 = m][h1 = k]
            o->A2T[i * 4 + t] = f2bf(p[OFF_W2 + k * NH + m]);  // W2[h2 = k][h1 = m]
        }
    }
    for (int i = threadIdx.x; i < NO * NH; i += blockDim.x) o->W3[i] = bf2f(f2bf(p[OFF_W3 + i]));
    for (int i = threadIdx.x; i < NH; i += blockDim.x) {
        o->b1[i] = p[OFF_B1 + i];
        o->b2[i] = p[OFF_B2 + i];
    }
    if (threadIdx.x < NO) o->b3[threadIdx.x] = p[OFF_B3 + threadIdx.x];
}

// ---------------------------------------------------------------------------
// forward pass of 32 points per wavefront (lane l%32 = point, lane/32 = half)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void gather(const Fields<T>& F, const Scales& sc, int x, int y, int h, bf16x4* xb) {
    const int msk = F.n - 1;
#pragma unroll
    for (int j = 0; j < KS1; ++j) {
        const int pos = 2 * j + h;  // features 8j + 4h .. +3 are the 4 fields at stencil point pos
        float v[4] = {0.f, 0.f, 0.f, 0.f};
        if (pos < 9) {
            const int dx = pos % 3 - 1, dy = pos / 3 - 1;
            const long idx = (long)((x + dx) & msk) + ((long)((y + dy) & msk) << F.shift);
#pragma unroll
            for (int t = 0; t < 4; ++t) v[t] = (float)F.f[t][idx] * sc.inv_in[t];
        }
        xb[j] = pack4(v[0], v[1], v[2], v[3]);
    }
}

struct Ops {
    bf16x4 a1[2][KS1];
    bf16x4 a2[2][KS2];
};

__device__ __forceinline__ void load_ops(const Packed* __restrict__ P, int lane, Ops& o) {
#pragma unroll
    for (int blk = 0; blk < 2; ++blk) {
#pragma unroll
        for (int j = 0; j < KS1; ++j) o.a1[blk][j] = ld4(&P->A1[((blk * KS1 + j) * 64 + lane) * 4]);
#pragma unroll
        for (int j = 0; j < KS2; ++j) o.a2[blk][j] = ld4(&P->A2[((blk * KS2 + j) * 64 + lane) * 4]);
    }
}

// A operands of the two MFMA layers read from the packed net at each use (the
// training kernel, whose registers hold the gradient accumulators)
struct OpsMem {
    const Packed* P;
    int lane;
    __device__ __forceinline__ bf16x4 a1(int blk, int j) const { return ld4(&P->A1[((blk * KS1 + j) * 64 + lane) * 4]); }
    __device__ __forceinline__ bf16x4 a2(int blk, int j) const { return ld4(&P->A2[((blk * KS2 + j) * 64 + lane) * 4]); }
};
struct OpsReg {
    const Ops& o;
    __device__ __forceinline__ bf16x4 a1(int blk, int j) const { return o.a1[blk][j]; }
    __device__ __forceinline__ bf16x4 a2(int blk, int j) const { return o.a2[blk][j]; }
};

// layer-1 pre-activations z1 = W1 x + b1 (accumulator layout)
template <class O>
__device__ __forceinline__ void layer1(const Packed* __restrict__ P, const O& o, const bf16x4* xb, int h, f32x16* z1) {
#pragma unroll
    for (int blk = 0; blk < 2; ++blk)
#pragma unroll
        for (int r = 0; r < 16; ++r) z1[blk][r] = P->b1[arow(blk, r, h)];
#pragma unroll
    for (int j = 0; j < KS1; ++j)
#pragma unroll
        for (int blk = 0; blk < 2; ++blk) z1[blk] = mfma(o.a1(blk, j), xb[j], z1[blk]);
}

// z2: pre-activations of the second hidden layer (accumulator layout); hb: H1
// as BF16 operands; out: the two outputs (normalised units), in both lane halves
template <class O>
__device__ __forceinline__ void forward(const Packed* __restrict__ P, const O& o, const bf16x4* xb, int lane,
                                        f32x16* z2, bf16x4* hb, float* out) {
    const int h = lane >> 5;
    f32x16 z1[2];
    layer1(P, o, xb, h, z1);
#pragma unroll
    for (int j = 0; j < KS2; ++j) {
        const int blk = j >> 2, q = 4 * (j & 3);
        hb[j] = pack4(gelu(z1[blk][q]), gelu(z1[blk][q + 1]), gelu(z1[blk][q + 2]), gelu(z1[blk][q + 3]));
    }
#pragma unroll
    for (int blk = 0; blk < 2; ++blk)
#pragma unroll
        for (int r = 0; r < 16; ++r) z2[blk][r] = P->b2[arow(blk, r, h)];
#pragma unroll
    for (int j = 0; j < KS2; ++j)
#pragma unroll
        for (int blk = 0; blk < 2; ++blk) z2[blk] = mfma(o.a2(blk, j), hb[j], z2[blk]);
    float o0 = 0.f, o1 = 0.f;
#pragma unroll
    for (int blk = 0; blk < 2; ++blk)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = arow(blk, r, h);
            const float g = gelu(z2[blk][r]);
            o0 += P->W3[row] * g;
            o1 += P->W3[NH + row] * g;
        }
    o0 += __shfl_xor(o0, 32);
    o1 += __shfl_xor(o1, 32);
    out[0] = o0 + P->b3[0];
    out[1] = o1 + P->b3[1];
}

// inference on a whole n x n grid: out_c = output_c * scale_c
template <typename T>
__global__ __launch_bounds__(256) void forward_kernel(const Packed* __restrict__ P, Scales sc, Fields<T> F,
                                                      T* __restrict__ out0, T* __restrict__ out1) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    Ops o;
    load_ops(P, lane, o);
    const long ntile = ((long)F.n << F.shift) >> 5;
    for (long t = (long)blockIdx.x * 4 + wave; t < ntile; t += (long)gridDim.x * 4) {
        const long p = t * 32 + (lane & 31);
        const int x = (int)(p & (F.n - 1)), y = (int)(p >> F.shift);
        bf16x4 xb[KS1], hb[KS2];
        f32x16 z2[2];
        float out[2];
        gather(F, sc, x, y, lane >> 5, xb);
        forward(P, OpsReg{o}, xb, lane, z2, hb, out);
        if (lane < 32) {
            out0[p] = (T)(out[0] * sc.out[0]);
            out1[p] = (T)(out[1] * sc.out[1]);
        }
    }
}

// ---------------------------------------------------------------------------
// training: forward + backward, per-wavefront partial gradients
// ---------------------------------------------------------------------------
__device__ __forceinline__ void tile_coord(const Data& d, const TileSel& s, long t, int& snap, int& y, int& x0) {
    const int tpr = d.n >> 5;  // tiles per row
    if (s.random) {
        uint64_t q = mix64(s.seed * 0x100000001B3ull + (uint64_t)t);
        snap = (int)(q % (uint64_t)d.nsnap);
        q /= (uint64_t)d.nsnap;
        y = (int)(q & (uint64_t)(d.n - 1));
        q >>= d.shift;
        x0 = (int)(q % (uint64_t)tpr) * 32;
    } else {
        const long per = (long)d.n * tpr;
        snap = s.s0 + (int)(t / per);
        const long r = t % per;
        y = (int)(r / tpr);
        x0 = (int)(r % tpr) * 32;
    }
}

constexpr int TRAIN_BLOCKS_MAX = 512;

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void train_kernel(const Packed* __restrict__ Pk, Scales sc, Data d, TileSel sel, float gscale,
                  float* __restrict__ part, float* __restrict__ part_loss) {
    // per wavefront: X[f][pt], H1[h1][pt], dZ2[h2][pt], dZ1[h1][pt] as BF16
    __shared__ uint16_t lds[4][4][64][32];
    // per wavefront, lane-private running sums of the output-layer gradient,
    // [c*32 + k][lane] for the hidden unit of accumulator k in the lane's half
    __shared__ float gw3s[4][64][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, h = lane >> 5, l = lane & 31;
    const long gw = (long)blockIdx.x * 4 + wave, nw = (long)gridDim.x * 4;
    uint16_t(*sX)[32] = lds[wave][0];
    uint16_t(*sH)[32] = lds[wave][1];
    uint16_t(*sD2)[32] = lds[wave][2];
    uint16_t(*sD1)[32] = lds[wave][3];
    for (int r = NF + h; r < 64; r += 2) sX[r][l] = 0;  // padded feature rows stay zero
    f32x16 gW1[2][2], gW2[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) gW1[a][b][r] = gW2[a][b][r] = 0.f;
    float(*gw3)[64] = gw3s[wave];
    for (int k = 0; k < 64; ++k) gw3[k][lane] = 0.f;
    float gb1[2] = {0.f, 0.f}, gb2[2] = {0.f, 0.f}, gb3[2] = {0.f, 0.f}, lossacc = 0.f;
    const long npts = (long)d.n << d.shift;

    for (long t = gw; t < sel.ntiles; t += nw) {
        const Packed* __restrict__ P = opaque(Pk);
        const OpsMem o{P, lane};
        int snap, y, x0;
        tile_coord(d, sel, t, snap, y, x0);
        const int x = x0 + l;
        Fields<float> F;
        for (int c = 0; c < NFLD; ++c) F.f[c] = d.fld + ((long)snap * NFLD + c) * npts;
        F.n = d.n;
        F.shift = d.shift;
        bf16x4 xb[KS1], hb[KS2], db[KS2];
        f32x16 z2[2];
        float out[2];
        gather(F, sc, x, y, h, xb);
        // operands of the weight gradients, [unit][point], as soon as they exist
#pragma unroll
        for (int j = 0; j < KS1; ++j)
#pragma unroll
            for (int tt = 0; tt < 4; ++tt) {
                const int f = 8 * j + 4 * h + tt;
                if (f < NF) sX[f][l] = (uint16_t)xb[j][tt];
            }
        forward(P, o, xb, lane, z2, hb, out);
#pragma unroll
        for (int j = 0; j < KS2; ++j)
#pragma unroll
            for (int tt = 0; tt < 4; ++tt) sH[8 * j + 4 * h + tt][l] = (uint16_t)hb[j][tt];
        __builtin_amdgcn_sched_barrier(0);
        const long p = x + ((long)y << d.shift);
        const float t0 = d.tgt[((long)snap * NO + 0) * npts + p] / sc.out[0];
        const float t1 = d.tgt[((long)snap * NO + 1) * npts + p] / sc.out[1];
        const float e0 = out[0] - t0, e1 = out[1] - t1;
        const float g0 = e0 * gscale, g1 = e1 * gscale;
        if (h == 0) {
            lossacc += e0 * e0 + e1 * e1;
            gb3[0] += g0;
            gb3[1] += g1;
        }
        // output layer: dW3 += g H2, dZ2 = (W3^T g) gelu'(Z2), packed as the
        // chained operand of dH1 (k-step j = accumulators 4(j%4)..+3 of block j/4)
#pragma unroll
        for (int j = 0; j < KS2; ++j) {
            const int blk = j >> 2;
            float dz[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = 4 * (j & 3) + q, row = arow(blk, r, h);
                const float z = z2[blk][r];
                const float a = gelu(z);
                gw3[blk * 16 + r][lane] += g0 * a;
                gw3[32 + blk * 16 + r][lane] += g1 * a;
                dz[q] = (P->W3[row] * g0 + P->W3[NH + row] * g1) * gelu_d(z);
            }
            db[j] = pack4(dz[0], dz[1], dz[2], dz[3]);
#pragma unroll
            for (int tt = 0; tt < 4; ++tt) sD2[8 * j + 4 * h + tt][l] = (uint16_t)db[j][tt];
        }
        __builtin_amdgcn_sched_barrier(0);
        // dH1 = W2^T dZ2 (dZ2 in the chained operand layout), dZ1 = dH1 gelu'(Z1)
        f32x16 dh1[2];
#pragma unroll
        for (int blk = 0; blk < 2; ++blk)
#pragma unroll
            for (int r = 0; r < 16; ++r) dh1[blk][r] = 0.f;
#pragma unroll
        for (int j = 0; j < KS2; ++j)
#pragma unroll
            for (int blk = 0; blk < 2; ++blk)
                dh1[blk] = mfma(ld4(&P->A2T[((blk * KS2 + j) * 64 + lane) * 4]), db[j], dh1[blk]);
        __builtin_amdgcn_sched_barrier(0);
        f32x16 z1[2];  // recomputed: cheaper than holding it through the output layer
        layer1(P, o, xb, h, z1);
#pragma unroll
        for (int blk = 0; blk < 2; ++blk)
#pragma unroll
            for (int r = 0; r < 16; ++r) sD1[arow(blk, r, h)][l] = f2bf(dh1[blk][r] * gelu_d(z1[blk][r]));
        wave_sync();
        __builtin_amdgcn_sched_barrier(0);
        // dW2[h2][h1] += sum_pt dZ2[h2][pt] H1[h1][pt];  dW1[h1][f] += sum_pt dZ1[h1][pt] X[f][pt]
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            bf16x4 A[2], B[2];
#pragma unroll
            for (int mb = 0; mb < 2; ++mb) {
                A[mb] = ld4(&sD2[32 * mb + l][8 * ks + 4 * h]);
                gb2[mb] += (bf2f((uint16_t)A[mb][0]) + bf2f((uint16_t)A[mb][1])) +
                           (bf2f((uint16_t)A[mb][2]) + bf2f((uint16_t)A[mb][3]));
            }
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) B[nb] = ld4(&sH[32 * nb + l][8 * ks + 4 * h]);
#pragma unroll
            for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                for (int nb = 0; nb < 2; ++nb) gW2[mb][nb] = mfma(A[mb], B[nb], gW2[mb][nb]);
#pragma unroll
            for (int mb = 0; mb < 2; ++mb) {
                A[mb] = ld4(&sD1[32 * mb + l][8 * ks + 4 * h]);
                gb1[mb] += (bf2f((uint16_t)A[mb][0]) + bf2f((uint16_t)A[mb][1])) +
                           (bf2f((uint16_t)A[mb][2]) + bf2f((uint16_t)A[mb][3]));
            }
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) B[nb] = ld4(&sX[32 * nb + l][8 * ks + 4 * h]);
#pragma unroll
            for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                for (int nb = 0; nb < 2; ++nb) gW1[mb][nb] = mfma(A[mb], B[nb], gW1[mb][nb]);
        }
        wave_sync();
    }

    // this wavefront's partial gradient
    float* pw = part + gw * NPARAM;
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int u = arow(mb, r, h), c = 32 * nb + l;
                pw[OFF_W2 + u * NH + c] = gW2[mb][nb][r];
                if (c < NF) pw[OFF_W1 + u * NF + c] = gW1[mb][nb][r];
            }
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
        const float v1 = gb1[mb] + __shfl_xor(gb1[mb], 32);
        const float v2 = gb2[mb] + __shfl_xor(gb2[mb], 32);
        if (h == 0) {
            pw[OFF_B1 + 32 * mb + l] = v1;
            pw[OFF_B2 + 32 * mb + l] = v2;
        }
    }
    // dW3[c][row]: sum over the 32 points (lanes of the half that holds row)
    wave_sync();
    for (int c = 0; c < NO; ++c) {
        const int row = lane;
        const int hh = (row >> 2) & 1, k = (row >> 5) * 16 + 4 * ((row & 31) >> 3) + (row & 3);
        float s = 0.f;
        for (int q = 0; q < 32; ++q) s += gw3[c * 32 + k][q + 32 * hh];
        pw[OFF_W3 + c * NH + row] = s;
    }
    float b0 = gb3[0], b1v = gb3[1], ls = lossacc;
    for (int m = 32; m >= 1; m >>= 1) {
        b0 += __shfl_xor(b0, m);
        b1v += __shfl_xor(b1v, m);
        ls += __shfl_xor(ls, m);
    }
    if (lane == 0) {
        pw[OFF_B3 + 0] = b0;
        pw[OFF_B3 + 1] = b1v;
        part_loss[gw] = 0.5f * ls * gscale;
    }
}

// grad[j] = sum_w part[w][j]: 32 parameters per workgroup, eight slices of the
// wavefronts per parameter (slice s sums w = s, s + 8, ... in order, then the
// slices are added in order: a fixed summation order), so a launch of a few
// thousand partials is 209 workgroups of short loops instead of one long loop
// per parameter.  Workgroup 0 also sums the partial losses (in FP64, 256
// strided runs added in a fixed tree).
constexpr int RED_PARAMS = 32, RED_SLICES = 8;

__global__ __launch_bounds__(256) void reduce_kernel(const float* __restrict__ part,
                                                     const float* __restrict__ part_loss, int nw,
                                                     float* __restrict__ grad, double* __restrict__ loss) {
    __shared__ float red[RED_SLICES][RED_PARAMS];
    __shared__ double lred[256];
    const int t = threadIdx.x, q = t % RED_PARAMS, s = t / RED_PARAMS;
    const int j = blockIdx.x * RED_PARAMS + q;
    float acc = 0.f;
    if (j < NPARAM)
        for (int w = s; w < nw; w += RED_SLICES) acc += part[(long)w * NPARAM + j];
    red[s][q] = acc;
    __syncthreads();
    if (s == 0 && j < NPARAM) {
        float g = red[0][q];
        for (int k = 1; k < RED_SLICES; ++k) g += red[k][q];
        grad[j] = g;
    }
    if (blockIdx.x == 0) {
        double ls = 0.0;
        for (int w = t; w < nw; w += 256) ls += (double)part_loss[w];
        lred[t] = ls;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (t < o) lred[t] += lred[t + o];
            __syncthreads();
        }
        if (t == 0) *loss = lred[0];
    }
}

__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ w, float* __restrict__ m,
                                                   float* __restrict__ v, const float* __restrict__ g, float lr,
                                                   float c1, float c2) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= NPARAM) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float gj = g[j];
    const float mj = b1 * m[j] + (1.f - b1) * gj;
    const float vj = b2 * v[j] + (1.f - b2) * gj * gj;
    m[j] = mj;
    v[j] = vj;
    w[j] -= lr * (mj / c1) / (sqrtf(vj / c2) + eps);
}

constexpr int SUMSQ_BLOCKS = 256;

__global__ __launch_bounds__(256) void sumsq_kernel(Data d, double* __restrict__ sums) {
    __shared__ double red[6][256];
    const long npts = (long)d.n << d.shift;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int s = 0; s < d.nsnap; ++s)
        for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < npts; p += (long)gridDim.x * 256) {
            for (int c = 0; c < NFLD; ++c) {
                const double v = d.fld[((long)s * NFLD + c) * npts + p];
                acc[c] += v * v;
            }
            for (int c = 0; c < NO; ++c) {
                const double v = d.tgt[((long)s * NO + c) * npts + p];
                acc[NFLD + c] += v * v;
            }
        }
    for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if (threadIdx.x < w)
            for (int q = 0; q < 6; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 6) sums[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int launch_pack(const float* params, Packed* out, hipStream_t s) {
    hipLaunchKernelGGL(pack_kernel, dim3(1), dim3(256), 0, s, params, out);
    return 0;
}

template <typename T>
int launch_forward(const Packed* net, Scales sc, Fields<T> in, T* out0, T* out1, hipStream_t s) {
    if (in.n < 32 || (1 << in.shift) != in.n) return -1;
    const long ntile = ((long)in.n << in.shift) >> 5;
    const long nb = std::min<long>((ntile + 3) / 4, 2048);
    hipLaunchKernelGGL((forward_kernel<T>), dim3((unsigned)nb), dim3(256), 0, s, net, sc, in, out0, out1);
    return 0;
}

int train_waves(long ntiles) {
    const long nb = std::min<long>((ntiles + 3) / 4, TRAIN_BLOCKS_MAX);
    return (int)(nb < 1 ? 4 : nb * 4);
}

int launch_train(const Packed* net, Scales sc, Data d, TileSel sel, double npts_total, float* part, float* part_loss,
                 hipStream_t s) {
    if (d.n < 32 || (1 << d.shift) != d.n || d.nsnap < 1 || sel.ntiles < 1) return -1;
    if (!sel.random && (sel.s0 < 0 || sel.s0 + sel.ns > d.nsnap || sel.ntiles > (long)sel.ns * d.n * (d.n / 32)))
        return -1;
    const int nw = train_waves(sel.ntiles);
    hipLaunchKernelGGL(train_kernel, dim3(nw / 4), dim3(256), 0, s, net, sc, d, sel, (float)(1.0 / npts_total), part,
                       part_loss);
    return 0;
}

int launch_reduce(const float* part, const float* part_loss, int nw, float* grad, double* loss, hipStream_t s) {
    hipLaunchKernelGGL(reduce_kernel, dim3((NPARAM + RED_PARAMS - 1) / RED_PARAMS), dim3(256), 0, s, part, part_loss, nw,
                       grad, loss);
    return 0;
}

int launch_adam(float* params, float* m, float* v, const float* grad, float lr, int t, hipStream_t s) {
    const double c1 = 1.0 - std::pow(0.9, t), c2 = 1.0 - std::pow(0.999, t);
    hipLaunchKernelGGL(adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, s, params, m, v, grad, lr, (float)c1,
                       (float)c2);
    return 0;
}

int sumsq_blocks() { return SUMSQ_BLOCKS; }

int launch_sumsq(Data d, double* sums, hipStream_t s) {
    hipLaunchKernelGGL(sumsq_kernel, dim3(SUMSQ_BLOCKS), dim3(256), 0, s, d, sums);
    return 0;
}

template int launch_forward<float>(const Packed*, Scales, Fields<float>, float*, float*, hipStream_t);
template int launch_forward<double>(const Packed*, Scales, Fields<double>, double*, double*, hipStream_t);

}  // namespace mlp
