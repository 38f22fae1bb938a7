"""Pinning the closure's CPU checkers (oracle/closure_oracle.py, the closure
hook of oracle/bastr_oracle.c) against what the mathematics fixes: BF16
round-to-nearest-even known answers, the tanh form of GELU against the exact
z Phi(z), the backward pass against central finite differences of the
forward pass (FP64), the stencil orientation of the features on a field
with a known form, the C oracle's force (seen through a step of size
1e-8) against the numpy forward pass, and its once-per-step cadence (one
step of a field too small for the nonlinear term is u + dt D(tau(u)))."""
import math

import numpy as np
import pytest

from bastr_amd.solver import make_params
from oracle import closure_oracle as co


def test_bf16_round_to_nearest_even():
    one = 1.0
    cases = {
        one: one,
        one + 2 ** -8: one,                      # halfway, ties to the even neighbour
        one + 3 * 2 ** -8: one + 2 ** -6,        # halfway, ties to even (k = 2)
        one + 2 ** -8 + 2 ** -20: one + 2 ** -7,  # just above halfway
        -(one + 2 ** -9): -one,
    }
    for x, want in cases.items():
        got = float(co.bf16(np.float32(x)).ravel()[0])
        assert got == float(np.float32(want)), (x, got, want)
    # bit patterns: low half above 0x8000 rounds the high half up, 0x8000 exactly to even
    for bits, want in [(0x7F61B1E6, 0x7F620000), (0x3F808000, 0x3F800000), (0x3F818000, 0x3F820000),
                       (0x3F807FFF, 0x3F800000)]:
        x = np.array([bits], dtype=np.uint32).view(np.float32)
        assert co.bf16(x).view(np.uint32)[0] == want, hex(bits)
    assert np.isnan(co.bf16(np.float32(np.nan)))


def test_gelu_tanh_form_is_close_to_exact_gelu():
    z = np.linspace(-6, 6, 2001)
    exact = np.array([v * 0.5 * (1 + math.erf(v / math.sqrt(2))) for v in z])
    assert np.abs(co.gelu(z) - exact).max() < 1e-3
    # derivative against a central difference of the same form
    h = 1e-6
    assert np.abs(co.gelu_d(z) - (co.gelu(z + h) - co.gelu(z - h)) / (2 * h)).max() < 1e-8


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = rng.standard_normal(co.NPARAM) * 0.3
    X = rng.standard_normal((40, co.NF))
    T = rng.standard_normal((40, co.NO))
    _, g = co.loss_grad(p, X, T, emulate=False)
    idx = np.concatenate([rng.choice(np.arange(lo, hi), min(8, hi - lo), replace=False) for lo, hi in
                          [(co.OFF_W1, co.OFF_B1), (co.OFF_B1, co.OFF_W2), (co.OFF_W2, co.OFF_B2),
                           (co.OFF_B2, co.OFF_W3), (co.OFF_W3, co.OFF_B3), (co.OFF_B3, co.NPARAM)]])
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp, _ = co.loss_grad(q, X, T, emulate=False)
        q[i] -= 2 * h
        lm, _ = co.loss_grad(q, X, T, emulate=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) < 1e-7 * max(1.0, abs(fd)), (i, fd, g[i])


def test_feature_stencil_orientation():
    """f[c](x, y) = c + x + 100 y: feature 4*pos + c of point (x, y) is the field at
    (x + dx, y + dy) mod n with pos = 3 (dy + 1) + (dx + 1)."""
    n = 8
    x = np.arange(n)[:, None]
    y = np.arange(n)[None, :]
    fields = np.stack([c + x + 100.0 * y + 0 * y for c in range(4)])
    X = co.features(fields, [1.0, 1.0, 1.0, 1.0], emulate=False)
    for (px, py) in [(0, 0), (3, 5), (7, 7), (7, 0)]:
        row = X[px + n * py]
        for pos in range(9):
            dx, dy = pos % 3 - 1, pos // 3 - 1
            for c in range(4):
                assert row[4 * pos + c] == c + (px + dx) % n + 100 * ((py + dy) % n)


def spectral_fields(u1, u2):
    n = u1.shape[0]
    k = np.arange(n, dtype=np.float64)
    k[n // 2 + 1:] -= n
    k1, k2 = np.meshgrid(k, k, indexing="ij")

    def d(f, kk):
        return np.real(np.fft.ifft2(1j * kk * np.fft.fft2(f)))

    return np.stack([u1, u2, d(u1, k1) + d(u2, k2), d(u2, k1) - d(u1, k2)])


@pytest.mark.parametrize("n", [32])
def test_c_oracle_closure_force_matches_numpy_forward(orc, n):
    """One step of size 1e-8: (u_closure - u_plain)/dt is the network's force on
    u0 to O(dt); the C oracle's closure against closure_oracle.forward."""
    from bastr_amd.fields import hitgen
    rng = np.random.default_rng(5)
    params = (rng.standard_normal(co.NPARAM) * 0.2).astype(np.float32)
    scales = np.array([0.8, 0.7, 5.0, 6.0, 0.3, 0.4], dtype=np.float32)
    u = hitgen((n, n), k0=3, seed=4)
    p = make_params(n, n, deltat=1e-8, reynolds=100.0)
    a = orc.solver(p)
    b = orc.solver(p)
    a.set_fields(*u)
    b.set_fields(*u)
    u0 = a.get_fields()  # entry-dealiased
    orc.set_closure(b, params, scales)
    a.step(1)
    b.step(1)
    force = [(x - y) / 1e-8 for x, y in zip(b.get_fields(), a.get_fields())]
    X = co.features(spectral_fields(*u0), scales)
    out, _ = co.forward(params, X)
    want = [out[:, c].reshape(n, n, order="F") * scales[4 + c] for c in range(2)]
    # the closure force is dealiased with the tendency: compare dealiased forms
    k = np.arange(n, dtype=np.float64)
    k[n // 2 + 1:] -= n
    k1, k2 = np.meshgrid(k, k, indexing="ij")
    keep = np.sqrt(k1 ** 2 + k2 ** 2) < n // 3
    want = [np.real(np.fft.ifft2(np.fft.fft2(w) * keep)) for w in want]
    err = max(np.abs(f - w).max() for f, w in zip(force, want))
    scale = max(np.abs(w).max() for w in want)
    assert scale > 0.05 and err < 1e-5 * scale, (err, scale)
    # a zero network changes nothing, bit for bit
    c = orc.solver(p)
    c.set_fields(*u)
    orc.set_closure(c, np.zeros(co.NPARAM, np.float32), scales)
    c.step(1)
    assert all(np.array_equal(x, y) for x, y in zip(c.get_fields(), a.get_fields()))


def test_c_oracle_per_step_cadence_holds_stage_a_force(orc):
    """The per-step cadence (bo_set_closure_per_step) adds the force of the
    step's input field at all three RK3 stages, whose weights sum to one.  On a
    field of size 1e-6 with nu ~ 0 the nonlinear and viscous terms vanish to
    1e-9 of the force, so one step is u + dt D(tau(u)), D the tendency's
    dealiasing and tau the numpy forward pass on u: the held force is stage A's
    exactly.  Evaluated at every stage instead, the force sees the stage fields
    u + O(dt tau), which differ from u by far more than u itself."""
    from bastr_amd.fields import hitgen
    n = 32
    rng = np.random.default_rng(7)
    params = (rng.standard_normal(co.NPARAM) * 0.2).astype(np.float32)
    scales = np.array([0.8, 0.7, 5.0, 6.0, 0.3, 0.4], dtype=np.float32)
    u = [x * 1e-6 for x in hitgen((n, n), k0=3, seed=6)]
    dt = 1e-3
    p = make_params(n, n, deltat=dt, reynolds=1e12)
    rate = {}
    for per_step in (True, False):
        s = orc.solver(p)
        s.set_fields(*u)
        u0 = s.get_fields()  # entry-dealiased
        orc.set_closure(s, params, scales)
        orc.set_closure_per_step(s, per_step)
        s.step(1)
        rate[per_step] = [(x - y) / dt for x, y in zip(s.get_fields(), u0)]
    X = co.features(spectral_fields(*u0), scales)
    out, _ = co.forward(params, X)
    want = [out[:, c].reshape(n, n, order="F") * scales[4 + c] for c in range(2)]
    k = np.arange(n, dtype=np.float64)
    k[n // 2 + 1:] -= n
    k1, k2 = np.meshgrid(k, k, indexing="ij")
    keep = np.sqrt(k1 ** 2 + k2 ** 2) < n // 3
    want = [np.real(np.fft.ifft2(np.fft.fft2(w) * keep)) for w in want]
    scale = max(np.abs(w).max() for w in want)
    held = max(np.abs(f - w).max() for f, w in zip(rate[True], want))
    staged = max(np.abs(f - w).max() for f, w in zip(rate[False], want))
    print("held", held, "staged", staged, "scale", scale)
    assert scale > 0.05 and held < 1e-6 * scale, (held, scale)
    assert staged > 2e-4 * scale, (staged, scale)
