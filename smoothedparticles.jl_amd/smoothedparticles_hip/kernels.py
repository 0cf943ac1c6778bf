"""Wendland quintic kernels on the host (numpy), with the support radius h as
in src/kernels.jl:1.  The device evaluates the same formulas inside the pair
kernels; these are for user code and set-up."""
import numpy as np


def _x(h, r):
    return np.asarray(r, dtype=np.float64) / h


def wendland2(h, r):
    """src/kernels.jl:108-115: 7/pi (1-x)^4 (1+4x) / h^2, 0 for x > 1."""
    x = _x(h, r)
    u = 1.0 - x
    w = 2.228169203286535 * (u * u) * (u * u) * (1.0 + 4.0 * x) / (h * h)
    return np.where(x > 1.0, 0.0, w)


def rDwendland2(h, r):
    """src/kernels.jl:140-147: -140/pi (1-x)^3 / h^4."""
    x = _x(h, r)
    u = 1.0 - x
    w = -44.563384065730695 * (u * u * u) / ((h * h) * (h * h))
    return np.where(x > 1.0, 0.0, w)


def wendland3(h, r):
    """src/kernels.jl:156-163: 21/(2 pi) (1-x)^4 (1+4x) / h^3."""
    x = _x(h, r)
    u = 1.0 - x
    w = 3.3422538049298023 * (u * u) * (u * u) * (1.0 + 4.0 * x) / (h * h * h)
    return np.where(x > 1.0, 0.0, w)


def rDwendland3(h, r):
    """src/kernels.jl:188-195: -210/pi (1-x)^3 / h^5."""
    x = _x(h, r)
    u = 1.0 - x
    w = -66.84507609859604 * (u * u * u) / ((h * h) * (h * h) * h)
    return np.where(x > 1.0, 0.0, w)
