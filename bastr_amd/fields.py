"""Seeded initial fields with the random-phase construction of Bastrpp's
hitgen2d/hitgen3d (Bastrpp.F90:601-808).

The reference draws its phases from an RNG seeded by system_clock
(Bastrpp.F90:626-635), so its fields are not reproducible; here the phases
come from numpy's PCG64 with an explicit seed.  As in hitgen, the phase of
mode k is the argument of the (normalised) FFT of uniform noise, which makes
the spectrum Hermitian so that the field is real, and the field is purely
dilatational: u_hat_i = A(|k|) k_i/|k| (sin t + i cos t).
  2-D: A = sqrt(E(k) / (2 pi k)),  3-D: A = sqrt(E(k) / (4 pi k^2)),
  E(k) = k^4 exp(-2 (k/k0)^2)  (k0 = 4 in the benchmark configurations; the
  reference's hitgen uses k0^6 k^-2 e^-2 below k0 instead).
The field is rescaled so that <sum_i u_i^2> = urms^2 (Bastrpp.F90:674-684).
Arrays are Fortran-ordered (ia, ja[, ka]).
"""
import numpy as np


def _wavenumbers(n):
    # GenerateWave (tool.F90:31-37): i-1 for i <= n/2+1, else i-n-1
    k = np.arange(n, dtype=np.float64)
    k[n // 2 + 1:] -= n
    return k


def hitgen(shape, k0=4.0, seed=0, urms=1.0):
    shape = tuple(int(s) for s in shape)
    nd = len(shape)
    rng = np.random.default_rng(seed)
    noise = rng.random(shape)
    theta = np.angle(np.fft.fftn(noise))
    ks = np.meshgrid(*[_wavenumbers(n) for n in shape], indexing="ij")
    kk = np.sqrt(sum(k * k for k in ks))
    safe = np.where(kk > 0, kk, 1.0)
    E = safe ** 4 * np.exp(-2.0 * (safe / k0) ** 2)
    A = np.sqrt(E / (2 * np.pi * safe)) if nd == 2 else np.sqrt(E / (4 * np.pi * safe ** 2))
    A = np.where(kk > 0, A, 0.0)
    phase = np.sin(theta) + 1j * np.cos(theta)
    npts = float(np.prod(shape))
    u = []
    for k in ks:
        uh = A * k / safe * phase
        # backward transform of fftwlink (unnormalised): numpy's ifftn divides by N
        u.append(np.real(np.fft.ifftn(uh)) * npts)
    energy = sum(np.mean(x * x) for x in u)
    f = urms / np.sqrt(energy)
    return [np.asfortranarray(x * f) for x in u]
