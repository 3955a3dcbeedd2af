"""Error bounds of device arithmetic that replaces a library call, checked on
the CPU with exact (mpmath) arithmetic."""
import math

import mpmath as mp
import numpy as np
import pytest

mp.mp.dps = 60


def _fma(a, b, c):
    """IEEE fused multiply-add: a*b + c exactly, then one rounding to FP64."""
    return float(mp.mpf(a) * mp.mpf(b) + mp.mpf(c))


def rsq_newton(x, y0, steps=2):
    """The FP64 reciprocal square root of the resident ensemble kernel
    (central-body terms): y <- y + (y/2)(1 - x y^2), the residual formed as
    fma(-(x*y), y, 1) and the update as fma(y/2, e, y), from a hardware
    estimate y0."""
    y = y0
    for _ in range(steps):
        e = _fma(-(x * y), y, 1.0)
        y = _fma(0.5 * y, e, y)
    return y


@pytest.mark.parametrize("steps,bound", [(2, 4 * 2.0 ** -53)])
def test_rsq_two_newton_steps_reach_fp64(steps, bound):
    """v_rsq_f64 is specified to about 2^29 ulp (relative error near 2^-23);
    from any estimate within 2^-21 of 1/sqrt(x), two Newton steps land within
    4 ulp over 60 decades of x, while one step leaves about 1.5 (2^-21)^2."""
    rng = np.random.default_rng(3)
    xs = 10.0 ** rng.uniform(-30, 30, 400)
    worst = 0.0
    for x in xs:
        exact = 1 / mp.sqrt(mp.mpf(x))
        for d in (-2.0 ** -21, -1e-7, 0.0, 3e-8, 2.0 ** -21):
            y0 = float(exact * (1 + mp.mpf(d)))
            y = rsq_newton(float(x), y0, steps)
            worst = max(worst, float(abs(mp.mpf(y) / exact - 1)))
    assert worst <= bound, worst
    # a single step is not enough: the check above would catch a dropped step
    x = 2.0
    exact = 1 / mp.sqrt(mp.mpf(x))
    y1 = rsq_newton(x, float(exact * (1 + mp.mpf(2.0 ** -21))), 1)
    assert float(abs(mp.mpf(y1) / exact - 1)) > 1e-13
