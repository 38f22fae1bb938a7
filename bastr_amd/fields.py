"""Seeded initial fields with the random-phase construction of Bastrpp's
hitgen2d/hitgen3d (create_initial_field_2d/3d, Bastrpp.F90:601-808).

The reference draws its phases from an RNG seeded by system_clock
(Bastrpp.F90:626-637), so its fields are not reproducible; here the noise is
SplitMix64 of (seed, index), the same source as the device generator
bastr_init_random_phase and the driver's `bastr_run hitgen`.  As in hitgen,
the phase of mode k is the argument of the FFT of uniform noise, and the field
is purely dilatational: u_hat_i = A(|k|) k_i/|k| (sin t + i cos t), u_hat(0) = 0,
  2-D: A = sqrt(E(k) / (2 pi k)),  3-D: A = sqrt(E(k) / (4 pi k^2)),
  E(k) = k^4 exp(-2 (k/k0)^2) for k >= k0, k0^6 k^-2 e^-2 below
  (k0 is the reference's forcek), u = real(ifft(u_hat)), unnormalised.
The field is rescaled so that <sum_i u_i^2> = energy (the reference's
target_energy, Bastrpp.F90:674-684).  Arrays are Fortran-ordered (ia, ja[, ka]).
"""
import numpy as np


def _wavenumbers(n):
    # GenerateWave (tool.F90:31-37): i-1 for i <= n/2+1, else i-n-1
    k = np.arange(n, dtype=np.float64)
    k[n // 2 + 1:] -= n
    return k


def splitmix_noise(shape, seed):
    """Uniform [0, 1) noise: (splitmix64(seed * 0x100000001B3 + m) >> 11) * 2^-53
    for the Fortran-order linear index m (the device generator's source)."""
    n = int(np.prod(shape))
    with np.errstate(over="ignore"):
        x = np.uint64(seed) * np.uint64(0x100000001B3) + np.arange(n, dtype=np.uint64)
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return ((x >> np.uint64(11)).astype(np.float64) * 2.0 ** -53).reshape(shape, order="F")


def spectrum_law(kk, k0, law=0):
    """E(k) of Bastrpp.F90:654-658 (law 0, kk > 0), or k^4 exp(-2 (k/k0)^2) at
    every k (law 1, the pure law)."""
    pure = kk ** 4 * np.exp(-2.0 * (kk / k0) ** 2)
    if law == 1:
        return pure
    return np.where(kk >= k0, pure, k0 ** 6 / kk ** 2 * np.exp(-2.0))


def hitgen(shape, k0=4.0, seed=0, energy=1.0, law=0):
    shape = tuple(int(s) for s in shape)
    nd = len(shape)
    noise = splitmix_noise(shape, seed)
    theta = np.angle(np.fft.fftn(noise))
    ks = np.meshgrid(*[_wavenumbers(n) for n in shape], indexing="ij")
    kk = np.sqrt(sum(k * k for k in ks))
    safe = np.where(kk > 0, kk, 1.0)
    E = spectrum_law(safe, float(k0), law)
    A = np.sqrt(E / (2 * np.pi * safe)) if nd == 2 else np.sqrt(E / (4 * np.pi * safe ** 2))
    A = np.where(kk > 0, A, 0.0)
    phase = np.sin(theta) + 1j * np.cos(theta)
    npts = float(np.prod(shape))
    u = []
    for k in ks:
        uh = A * k / safe * phase
        # backward transform of fftwlink (unnormalised): numpy's ifftn divides by N
        u.append(np.real(np.fft.ifftn(uh)) * npts)
    e = sum(np.mean(x * x) for x in u)
    f = np.sqrt(energy / e)
    return [np.asfortranarray(x * f) for x in u]
