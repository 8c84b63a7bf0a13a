"""Writes tests/golden/makima.npz: interior makima values/derivatives from
scipy.interpolate.Akima1DInterpolator(method="makima"), an independent
library implementation of the modified-Akima slopes the reference restates
(src/potentials/spline.cpp:43-103; both extend the secants by the same
quadratic extrapolation at the ends). Only points inside [x0, xn] are
recorded: the reference extrapolates linearly with the end secant outside,
scipy does not. Run: python tests/golden/gen_makima.py
"""
import os

import numpy as np
from scipy.interpolate import Akima1DInterpolator

rng = np.random.default_rng(20251203)
cases = []
for n in (4, 5, 6, 9, 15, 20):
    x = np.concatenate([[0.5], np.sort(rng.uniform(0.7, 9.3, size=n - 2)), [9.5]])
    y = rng.normal(size=n)
    cases.append((x, y))
# uniform knots, smooth data (the potential-table case)
xu = 1.8 + np.arange(15) * (3.7 / 14)
cases.append((xu, 3.6 * np.exp(-1.26 * xu)))
cases.append((xu, np.sin(xu)))
out = {}
for k, (x, y) in enumerate(cases):
    f = Akima1DInterpolator(x, y, method="makima")
    r = np.concatenate([x, rng.uniform(x[0], x[-1], size=40)])
    out[f"x{k}"] = x
    out[f"y{k}"] = y
    out[f"r{k}"] = r
    out[f"v{k}"] = f(r)
    out[f"d{k}"] = f(r, 1)
out["ncases"] = np.array(len(cases))
np.savez(os.path.join(os.path.dirname(os.path.abspath(__file__)), "makima.npz"), **out)
