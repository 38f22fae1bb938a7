// closure.hip -- spectral helper kernels of the learned subgrid closure.
#include "closure.hpp"

namespace clo {

__device__ __forceinline__ int kwave(int i, int n) { return i <= n / 2 ? i : i - n; }

template <typename T>
__global__ __launch_bounds__(256) void grad_kernel(Half g, k2d::Gate gate, const cx<T>* __restrict__ u1,
                                                   const cx<T>* __restrict__ u2, cx<T>* __restrict__ th,
                                                   cx<T>* __restrict__ om) {
    if (gate.closed()) return;
    const size_t total = (size_t)g.hp * g.n1;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const int kx = (int)(i % g.hp), r = (int)(i / g.hp);
        const T k1 = (kx < g.nh && kx != g.n0 / 2) ? (T)kx : (T)0;
        const T k2 = r == g.n1 / 2 ? (T)0 : (T)kwave(r, g.n1);
        const cx<T> a = u1[i], b = u2[i];
        // i * (k1 a + k2 b), i * (k1 b - k2 a)
        const cx<T> d = {k1 * a.x + k2 * b.x, k1 * a.y + k2 * b.y};
        const cx<T> w = {k1 * b.x - k2 * a.x, k1 * b.y - k2 * a.y};
        th[i] = {-d.y, d.x};
        om[i] = {-w.y, w.x};
    }
}

template <typename T>
__global__ __launch_bounds__(256) void tau_kernel(Half g, int cut2, const cx<T>* __restrict__ a0,
                                                  const cx<T>* __restrict__ a1, const cx<T>* __restrict__ b0,
                                                  const cx<T>* __restrict__ b1, cx<T>* __restrict__ o0,
                                                  cx<T>* __restrict__ o1) {
    const size_t total = (size_t)g.hp * g.n1;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const int kx = (int)(i % g.hp), r = (int)(i / g.hp);
        const int ky = kwave(r, g.n1);
        const bool keep = kx < g.nh && kx * kx + ky * ky < cut2;
        o0[i] = keep ? csub(a0[i], b0[i]) : cx<T>{(T)0, (T)0};
        o1[i] = keep ? csub(a1[i], b1[i]) : cx<T>{(T)0, (T)0};
    }
}

template <typename T>
__global__ __launch_bounds__(256) void add_kernel(k2d::Gate gate, cx<T>* __restrict__ x0, cx<T>* __restrict__ x1,
                                                  const cx<T>* __restrict__ y0, const cx<T>* __restrict__ y1,
                                                  size_t n) {
    if (gate.closed()) return;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        x0[i] = cadd(x0[i], y0[i]);
        x1[i] = cadd(x1[i], y1[i]);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void to_float_kernel(const T* __restrict__ src, float* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = (float)src[i];
}

static unsigned blocks(size_t n) {
    const size_t b = (n + 255) / 256;
    return (unsigned)(b < 2048 ? (b < 1 ? 1 : b) : 2048);
}

template <typename T>
int launch_grad(Half g, const k2d::Gate& gate, const cx<T>* u1, const cx<T>* u2, cx<T>* th, cx<T>* om, hipStream_t s) {
    hipLaunchKernelGGL((grad_kernel<T>), dim3(blocks((size_t)g.hp * g.n1)), dim3(256), 0, s, g, gate, u1, u2, th, om);
    return 0;
}

// the field pointers travel by value in a small struct (a kernel cannot take host arrays)
template <typename T>
struct Ptrs {
    const cx<T>* src[4];
    cx<T>* dst[4];
};

template <typename T>
__global__ __launch_bounds__(256) void filter_kernel_v(Half f, Half c, int cut2, Ptrs<T> p, int nf) {
    const size_t total = (size_t)c.hp * c.n1;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const int kx = (int)(i % c.hp), r = (int)(i / c.hp);
        const int ky = kwave(r, c.n1);
        const bool keep = kx < c.nh && kx < f.nh && kx * kx + ky * ky < cut2 && 2 * (ky < 0 ? -ky : ky) < f.n1;
        const size_t fi = keep ? (size_t)(ky < 0 ? ky + f.n1 : ky) * f.hp + kx : 0;
        for (int q = 0; q < nf; ++q) p.dst[q][i] = keep ? p.src[q][fi] : cx<T>{(T)0, (T)0};
    }
}

template <typename T>
int launch_filter(Half fine, Half coarse, int cut, const cx<T>* const* src, cx<T>* const* dst, int nf,
                  hipStream_t s) {
    if (nf < 1 || nf > 4) return -1;
    Ptrs<T> p{};
    for (int q = 0; q < nf; ++q) {
        p.src[q] = src[q];
        p.dst[q] = dst[q];
    }
    hipLaunchKernelGGL((filter_kernel_v<T>), dim3(blocks((size_t)coarse.hp * coarse.n1)), dim3(256), 0, s, fine,
                       coarse, cut * cut, p, nf);
    return 0;
}

template <typename T>
int launch_tau(Half g, int cut, const cx<T>* a0, const cx<T>* a1, const cx<T>* b0, const cx<T>* b1, cx<T>* o0,
               cx<T>* o1, hipStream_t s) {
    hipLaunchKernelGGL((tau_kernel<T>), dim3(blocks((size_t)g.hp * g.n1)), dim3(256), 0, s, g, cut * cut, a0, a1, b0,
                       b1, o0, o1);
    return 0;
}

template <typename T>
int launch_add(const k2d::Gate& gate, cx<T>* x0, cx<T>* x1, const cx<T>* y0, const cx<T>* y1, size_t n,
               hipStream_t s) {
    hipLaunchKernelGGL((add_kernel<T>), dim3(blocks(n)), dim3(256), 0, s, gate, x0, x1, y0, y1, n);
    return 0;
}

template <typename T>
int launch_to_float(const T* src, float* dst, size_t n, hipStream_t s) {
    hipLaunchKernelGGL((to_float_kernel<T>), dim3(blocks(n)), dim3(256), 0, s, src, dst, n);
    return 0;
}

#define INST(T)                                                                                                  \
    template int launch_grad<T>(Half, const k2d::Gate&, const cx<T>*, const cx<T>*, cx<T>*, cx<T>*, hipStream_t); \
    template int launch_filter<T>(Half, Half, int, const cx<T>* const*, cx<T>* const*, int, hipStream_t);          \
    template int launch_tau<T>(Half, int, const cx<T>*, const cx<T>*, const cx<T>*, const cx<T>*, cx<T>*, cx<T>*,   \
                               hipStream_t);                                                                       \
    template int launch_add<T>(const k2d::Gate&, cx<T>*, cx<T>*, const cx<T>*, const cx<T>*, size_t, hipStream_t); \
    template int launch_to_float<T>(const T*, float*, size_t, hipStream_t);
INST(float)
INST(double)

}  // namespace clo
