"""Pins of the oracle's smoothing kernel (m_particles.f90:1469-1602) against
textbook closed forms, the unit integral, finite differences and symmetry."""
import math

import numpy as np
import pytest

import oracle

H = 0.006
PI_F = float(np.float32(math.pi))


def textbook(skf, q, h):
    """2-D kernels as published (exact pi): Monaghan cubic spline, Gaussian,
    Morris quintic, Wendland C2 (support 2h)."""
    if skf == 1:
        s = 10.0 / (7.0 * math.pi * h * h)
        if q <= 1:
            return s * (1 - 1.5 * q * q + 0.75 * q ** 3)
        if q <= 2:
            return s * 0.25 * (2 - q) ** 3
        return 0.0
    if skf == 2:
        return math.exp(-q * q) / (math.pi * h * h) if q <= 3 else 0.0
    if skf == 3:
        s = 7.0 / (478.0 * math.pi * h * h)
        t = 0.0
        if q <= 3:
            t += (3 - q) ** 5
        if q <= 2:
            t -= 6 * (2 - q) ** 5
        if q <= 1:
            t += 15 * (1 - q) ** 5
        return s * t
    if skf == 4:
        return 7.0 / (4.0 * math.pi * h * h) * (1 - q / 2) ** 4 * (2 * q + 1) if q <= 2 else 0.0


@pytest.mark.parametrize("skf", [1, 2, 3, 4])
def test_values_match_textbook(skf):
    for q in np.linspace(0.0, 3.2, 161):
        w, _ = oracle.kernel(skf, q * H, [q * H, 0.0], H)
        ref = textbook(skf, (q * H) / H, H)  # the q the kernel itself sees
        # The reference's pi (and 2./3.) are FP32 literals: ~3e-8 relative offset.
        assert abs(w - ref) <= 1e-7 * abs(textbook(skf, 0.0, H)) + 1e-300, (skf, q, w, ref)


def test_fp32_literal_offset_is_present():
    # W(0) of the cubic spline is 15/(7 pi h^2) * 2/3 with FP32 pi and FP32 2/3:
    # a relative offset from the exact value of about (pi-pi_f)/pi + (2/3)_f error.
    w, _ = oracle.kernel(1, 0.0, [0.0, 0.0], H)
    exact = 10.0 / (7.0 * math.pi * H * H)
    rel = w / exact - 1.0
    expect = (math.pi / PI_F) * (float(np.float32(2.0) / np.float32(3.0)) / (2.0 / 3.0)) - 1.0
    assert 1e-9 < abs(rel) < 1e-7
    assert abs(rel - expect) < 1e-15


@pytest.mark.parametrize("skf", [1, 2, 3, 4])
def test_unit_integral(skf):
    # integral of W over the plane = 1 (polar quadrature, up to the FP32 pi offset)
    r = np.linspace(0.0, 3.0 * H, 60001)
    w = np.array([oracle.kernel(skf, x, [x, 0.0], H)[0] for x in r])
    integral = np.trapezoid(2.0 * math.pi * r * w, r)
    # the Gaussian is truncated at q = 3: its integral is 1 - exp(-9)
    expect = 1.0 - math.exp(-9.0) if skf == 2 else 1.0
    assert abs(integral - expect) < 2e-6, integral


@pytest.mark.parametrize("skf", [1, 2, 3, 4])
def test_gradient_matches_finite_difference(skf):
    rng = np.random.default_rng(skf)
    for _ in range(200):
        lo = 1.02 if skf == 3 else 0.05   # skf=3, q<=1: see test_quintic_inner_gradient_quirk
        q = rng.uniform(lo, 2.9 if skf in (2, 3) else 1.95)
        ang = rng.uniform(0, 2 * math.pi)
        dx = np.array([math.cos(ang), math.sin(ang)]) * q * H
        _, g = oracle.kernel(skf, q * H, dx, H)
        eps = 1e-7 * H
        for d in range(2):
            e = np.zeros(2)
            e[d] = eps
            xp, xm = dx + e, dx - e
            wp, _ = oracle.kernel(skf, float(np.hypot(*xp)), xp, H)
            wm, _ = oracle.kernel(skf, float(np.hypot(*xm)), xm, H)
            fd = (wp - wm) / (2 * eps)
            scale = abs(textbook(skf, 0.0, H)) / H
            assert abs(g[d] - fd) <= 2e-6 * scale, (skf, q, d, g[d], fd)


@pytest.mark.parametrize("skf", [1, 2, 3, 4])
def test_gradient_is_odd_bit_exact(skf):
    rng = np.random.default_rng(10 + skf)
    for _ in range(500):
        dx = rng.uniform(-2 * H, 2 * H, 2)
        r = float(np.sqrt(dx[0] * dx[0] + dx[1] * dx[1]))
        w1, g1 = oracle.kernel(skf, r, dx, H)
        w2, g2 = oracle.kernel(skf, r, -dx, H)
        assert w1 == w2
        assert np.array_equal(g1, -g2)


def test_quintic_inner_gradient_quirk():
    """For skf=3 and q <= 1 the reference writes grad W with the polynomial
    (-120 + 120 q - 50 q^2) (m_particles.f90:1549-1551), whereas dW/dq of its
    own W gives q (-120 + 120 q^2 - 50 q^3).  The oracle keeps the reference's
    expression; the deviation from the true derivative is exactly
    factor * (120 (q - q^2) - 50 (q^2 - q^3)) / h^2 * dx."""
    rng = np.random.default_rng(33)
    factor = 7.0 / (478.0 * PI_F * H * H)
    for _ in range(100):
        q = rng.uniform(0.05, 0.98)
        dx = np.array([q * H, 0.0])
        _, g = oracle.kernel(3, q * H, dx, H)
        eps = 1e-7 * H
        wp, _ = oracle.kernel(3, q * H + eps, dx + [eps, 0.0], H)
        wm, _ = oracle.kernel(3, q * H - eps, dx - [eps, 0.0], H)
        fd = (wp - wm) / (2 * eps)
        dev = factor * (120 * (q - q * q) - 50 * (q * q - q ** 3)) / (H * H) * dx[0]
        assert abs((g[0] - fd) - dev) <= 1e-5 * abs(factor) / H
