"""The reference's test_line_integral (delta_ray.cpp:1572-1836) as a
known-answer test.  The reference prints four estimates with no asserts; here
they are checked against the exact integrals:

  cosine lobe:  (1/pi) * int_{-10}^{10} y^2 / (x^2 + y^2)^2 dx,  y = 3
  glossy lobe:  int D_ltc(w(x)) * y / |p(x)|^3 dx,  Minv = diag(3, 3, 1) frame(+y)

The CPU test checks those closed forms against brute-force quadrature; the GPU
test runs the device's LTC line sampler (shared with the secondary-edge term)
through dr_test_line_integral and requires the uniform and LTC estimates to
match them within 5 standard errors, and the LTC-sampled cosine estimate, whose
sampling density is proportional to its integrand, to have near-zero
variance."""
import math

import numpy as np
import pytest

Y = 3.0


def cosine_exact():
    # int dx / (x^2 + a^2)^2 = x / (2 a^2 (x^2 + a^2)) + atan(x / a) / (2 a^3)
    a = Y
    F = lambda x: x / (2 * a * a * (x * x + a * a)) + math.atan(x / a) / (2 * a ** 3)
    return a * a / math.pi * (F(10.0) - F(-10.0))


def _glossy_integrand(x):
    p = np.stack([x, np.full_like(x, Y), np.zeros_like(x)], axis=-1)
    r = np.linalg.norm(p, axis=-1)
    d = p / r[:, None]
    # frame(+y) rows: (1, 0, 0), (0, 0, -1), (0, 1, 0); Minv = diag(3, 3, 1) * frame
    td = np.stack([3 * d[:, 0], -3 * d[:, 2], d[:, 1]], axis=-1)
    tl = np.linalg.norm(td, axis=-1)
    brdf = (td[:, 2] / tl) / math.pi * (9.0 / tl ** 3)
    geom = d[:, 1] / r ** 2
    return brdf * geom


def glossy_exact(n=200001):
    x = np.linspace(-10.0, 10.0, n)
    f = _glossy_integrand(x)
    h = x[1] - x[0]
    return h / 3.0 * (f[0] + f[-1] + 4 * f[1:-1:2].sum() + 2 * f[2:-1:2].sum())   # Simpson


def test_closed_forms_match_quadrature():
    x = np.linspace(-10.0, 10.0, 400001)
    cos_f = (1.0 / math.pi) * (Y / np.hypot(x, Y)) * (Y / np.hypot(x, Y) ** 3)
    h = x[1] - x[0]
    simpson = h / 3.0 * (cos_f[0] + cos_f[-1] + 4 * cos_f[1:-1:2].sum() + 2 * cos_f[2:-1:2].sum())
    assert abs(simpson - cosine_exact()) < 1e-9
    assert abs(cosine_exact() - 0.16494488) < 1e-8
    # Minv = diag(3, 3, 1) narrows the lobe around +y, which the segment
    # passes over: more of it is captured than of the cosine lobe
    assert cosine_exact() < glossy_exact() < 1.0
    assert abs(glossy_exact() - glossy_exact(400001)) < 1e-9


@pytest.mark.gpu
def test_line_integral_gpu_known_answers(gpu):
    from delta_ray_amd import delta_ray as dr
    n = 1 << 20
    r = dr.test_line_integral(n, seed=5)
    ce, ge = cosine_exact(), glossy_exact()

    def se(key):
        return math.sqrt(max(r[key], 0.0) / n)
    assert abs(r["Uniform line"] - ce) < 5 * se("Uniform variance") + 1e-7
    assert abs(r["Cosine line"] - ce) < 1e-4 * ce
    assert r["Cosine variance"] < 1e-6 * r["Uniform variance"]
    assert abs(r["Uniform line glossy"] - ge) < 5 * se("Uniform line glossy variance") + 1e-7
    assert abs(r["Cosine line glossy"] - ge) < 5 * se("Cosine line glossy variance") + 1e-7
    assert r["Cosine line glossy variance"] < r["Uniform line glossy variance"]
    # the default sample count of the reference runs too
    r2 = dr.test_line_integral()
    assert abs(r2["Cosine line"] - ce) < 1e-4 * ce
