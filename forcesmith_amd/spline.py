"""Host-side makima spline of one tabulated potential, for the model files.

The reference resamples every table when it writes a model
(io::write_native, src/io/output_writer.cpp:23-43): n knots spaced
(hi - lo) / (n - 1) over the table's span, each the spline's value there.
Writing such a file from this package needs the same spline on the host:
  - slopes: SplinePotential::recompute_slopes_ (src/potentials/spline.cpp:
    43-103): modified Akima weights w = |m_a - m_b| + |m_a + m_b| / 2 over the
    secants, quadratic extrapolation of two secants beyond each end, NaN -> 0,
    no slopes for n < 4 (piecewise linear);
  - value: SplinePotential::eval (spline.cpp:215-236): the end secant outside
    [x_0, x_n-1], linear interpolation for n < 4, else the cubic Hermite form
    (1-t)^2 (y0 (1+2t) + s0 (r-x0)) + t^2 (y1 (3-2t) + dx s1 (t-1)) on the
    interval upper_bound(x, r) - 1 clamped to [0, n-2] (spline.cpp:18-30,
    130-135).
The device kernels evaluate the same spline; this module only serves file
output.
"""
from __future__ import annotations

import math

import numpy as np


def makima_slopes(x, y):
    """Knot slopes of the reference's makima spline, or None for n < 4."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    n = len(x)
    if n < 4:
        return None
    sec = [(y[k + 1] - y[k]) / (x[k + 1] - x[k]) for k in range(n - 1)]

    def m(q):  # secant q, extended quadratically beyond both ends
        if 0 <= q <= n - 2:
            return sec[q]
        if q == -1:
            return 2 * sec[0] - sec[1]
        if q == -2:
            return 2 * (2 * sec[0] - sec[1]) - sec[0]
        if q == n - 1:
            return 2 * sec[n - 2] - sec[n - 3]
        return 2 * (2 * sec[n - 2] - sec[n - 3]) - sec[n - 2]

    s = np.zeros(n)
    for k in range(n):
        a, b, c, p = m(k - 1), m(k), m(k + 1), m(k - 2)
        w1 = abs(c - b) + abs(c + b) / 2
        w2 = abs(a - p) + abs(a + p) / 2
        with np.errstate(invalid="ignore", divide="ignore"):
            v = (w1 * a + w2 * b) / (w1 + w2) if (w1 + w2) != 0 else math.nan
        s[k] = 0.0 if math.isnan(v) else v
    return s


def evaluate(x, y, s, r: float) -> float:
    """The spline's value at r (uncached evaluation, end-secant extrapolation)."""
    n = len(x)
    if r <= x[0]:
        return y[0] + (y[1] - y[0]) / (x[1] - x[0]) * (r - x[0])
    if r >= x[n - 1]:
        return y[n - 1] + (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]) * (r - x[n - 1])
    i = int(np.searchsorted(x, r, side="right")) - 1
    i = min(max(i, 0), n - 2)
    if s is None:
        t = (r - x[i]) / (x[i + 1] - x[i])
        return y[i] + t * (y[i + 1] - y[i])
    x0, dx = x[i], x[i + 1] - x[i]
    t = (r - x0) / dx
    omt = 1.0 - t
    return omt * omt * (y[i] * (1.0 + 2.0 * t) + s[i] * (r - x0)) + \
        t * t * (y[i + 1] * (3.0 - 2.0 * t) + dx * s[i + 1] * (t - 1.0))


def resample(x, y, nknots: int):
    """(lo, hi, values) of the spline sampled at nknots points over its span,
    as the reference's model writer samples a table (output_writer.cpp:23-43)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    lo, hi = float(x[0]), float(x[-1])
    step = (hi - lo) / (nknots - 1)
    s = makima_slopes(x, y)
    vals = [evaluate(x, y, s, lo + k * step) for k in range(nknots)]
    if not all(math.isfinite(v) for v in vals):
        raise ValueError("non-finite potential value: cannot tabulate the potential for output")
    return lo, hi, np.array(vals)
