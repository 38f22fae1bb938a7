// Seeded random-phase initial field on the device (the construction of
// Bastrpp's create_initial_field_2d/3d, Bastrpp.F90:601-808).
#include <math.h>
#include <stdint.h>

#include "fft_block.hpp"
#include "init_rp.hpp"

namespace rp {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// uniform noise in [0, 1) indexed by the Fortran-order linear index
template <typename T>
__global__ __launch_bounds__(256) void noise_kernel(T* out, size_t n, uint64_t seed) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        out[i] = (T)((double)(splitmix64(seed * 0x100000001B3ull + i) >> 11) * 0x1.0p-53);
}

template <typename T>
__global__ __launch_bounds__(256) void zero_kernel(T* out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) out[i] = (T)0;
}

// u_hat_c = A(k) k_c/|k| (sin t + i cos t), t = arg(noise_hat), on the half
// spectrum; noise_hat is read from U[0] before it is overwritten.
template <typename T>
__global__ __launch_bounds__(256) void build_kernel(RpArgs a) {
    cx<T>* U[3] = {(cx<T>*)a.U[0], (cx<T>*)a.U[1], (cx<T>*)a.U[2]};
    const size_t total = (size_t)a.hp * a.n1 * a.n2;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const int kx = (int)(i % a.hp);
        const size_t row = i / a.hp;
        const int r1 = (int)(row % a.n1), r2 = (int)(row / a.n1);
        const int ky = r1 <= a.n1 / 2 ? r1 : r1 - a.n1;
        const int kz = a.nd == 3 ? (r2 <= a.n2 / 2 ? r2 : r2 - a.n2) : 0;
        const double kk = sqrt((double)kx * kx + (double)ky * ky + (double)kz * kz);
        const cx<T> nz = U[0][i];
        cx<T> o[3] = {{0, 0}, {0, 0}, {0, 0}};
        if (kx < a.nh && kk > 0) {
            const double t = atan2((double)nz.y, (double)nz.x);
            // hitgen law: E(k) = k^4 exp(-2 (k/k0)^2) from k0 up, k0^6 k^-2 e^-2 below
            // (Bastrpp.F90:654-658); pure law: k^4 exp(-2 (k/k0)^2) at every k
            const double q = kk / a.k0, kk2 = kk * kk, k0 = a.k0;
            const double E = (kk >= k0 || a.law == 1) ? kk2 * kk2 * exp(-2.0 * (q * q))
                                                     : k0 * k0 * k0 * k0 * k0 * k0 / kk2 * exp(-2.0);
            const double A = a.nd == 2 ? sqrt(E / (2.0 * M_PI * kk)) : sqrt(E / (4.0 * M_PI * kk * kk));
            const double st = sin(t), ct = cos(t);
            const double kc[3] = {(double)kx, (double)ky, (double)kz};
            // the field is real(ifft(u_hat)): keep the Hermitian part of u_hat.  Away from
            // the Nyquist index of component c it is u_hat_c itself; at k_c = +n_c/2 the
            // partner mode carries the same k_c and the opposite phase, so it cancels.
            const bool nyq[3] = {kx == a.n0 / 2, r1 == a.n1 / 2, a.nd == 3 && r2 == a.n2 / 2};
            for (int c = 0; c < a.nd; ++c)
                if (!nyq[c]) o[c] = {(T)(A * kc[c] / kk * st), (T)(A * kc[c] / kk * ct)};
        }
        for (int c = 0; c < a.nd; ++c) U[c][i] = o[c];
    }
}

template <typename T>
int launch_noise(void* out, size_t n, uint64_t seed, hipStream_t s) {
    hipLaunchKernelGGL((noise_kernel<T>), dim3(2048), dim3(256), 0, s, (T*)out, n, seed);
    return 0;
}
template <typename T>
int launch_zero(void* out, size_t n, hipStream_t s) {
    hipLaunchKernelGGL((zero_kernel<T>), dim3(2048), dim3(256), 0, s, (T*)out, n);
    return 0;
}
template <typename T>
int launch_build(const RpArgs& a, hipStream_t s) {
    hipLaunchKernelGGL((build_kernel<T>), dim3(2048), dim3(256), 0, s, a);
    return 0;
}

template int launch_noise<float>(void*, size_t, uint64_t, hipStream_t);
template int launch_noise<double>(void*, size_t, uint64_t, hipStream_t);
template int launch_zero<float>(void*, size_t, hipStream_t);
template int launch_zero<double>(void*, size_t, hipStream_t);
template int launch_build<float>(const RpArgs&, hipStream_t);
template int launch_build<double>(const RpArgs&, hipStream_t);

}  // namespace rp
