"""Writes gamma_ppf.json: scipy.stats.gamma.ppf golden vectors for the
discrete-gamma category rates (b3's update_categories, crates/b3/src/
likelihood/gamma.rs:205-225, which calls computare's Gamma::inverse_cdf).

Run: python tests/golden/make_gamma_ppf.py   (scipy 1.15.3 in this image)
"""
import json
import os

import numpy as np
import scipy
from scipy.stats import gamma

HERE = os.path.dirname(os.path.abspath(__file__))

cases = []
for alpha in [0.05, 0.1, 0.2, 0.5, 1.0, 2.0, 5.0, 10.0, 50.0, 200.0]:
    for k in [1, 2, 4, 8]:
        p = (2 * np.arange(k) + 1) / (2 * k)
        q = gamma.ppf(p, alpha)
        cases.append({"alpha": alpha, "k": k, "quantiles": q.tolist(),
                      "rates": (q / q.mean()).tolist()})
points = []
for alpha in [0.05, 0.3, 1.0, 7.5]:
    for p in [1e-6, 0.01, 0.125, 0.5, 0.875, 0.99]:
        points.append({"alpha": alpha, "p": p, "ppf": float(gamma.ppf(p, alpha)),
                       "cdf_at_1": float(gamma.cdf(1.0, alpha))})
with open(os.path.join(HERE, "gamma_ppf.json"), "w") as f:
    json.dump({"scipy": scipy.__version__, "categories": cases, "points": points}, f, indent=1)
