// Seeded random-phase initial field: kernel launchers (init_rp.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rp {

struct RpArgs {
    void* U[3];       // half spectra; U[0] holds the noise spectrum on entry
    int nd, n0, n1, n2, nh, hp;
    double k0;
    int law;          // BASTR_LAW_HITGEN or BASTR_LAW_PURE (include/bastr.h)
};

template <typename T> int launch_noise(void* out, size_t n, uint64_t seed, hipStream_t s);
template <typename T> int launch_zero(void* out, size_t n, hipStream_t s);
template <typename T> int launch_build(const RpArgs& a, hipStream_t s);

}  // namespace rp
