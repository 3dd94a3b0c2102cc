"""Accuracy and conservation measures reported with every result."""
from __future__ import annotations

import numpy as np


def field_errors(macro_a, macro_b, q_a=None, q_b=None):
    """Relative L-inf and L2 differences of rho, U (both components), T and q
    between two macro fields (ni, nj, 4) holding (rho, U, V, 1/T)."""
    out = {}
    pairs = {"rho": (macro_a[..., 0], macro_b[..., 0]), "U": (macro_a[..., 1:3], macro_b[..., 1:3]),
             "T": (1.0 / macro_a[..., 3], 1.0 / macro_b[..., 3])}
    if q_a is not None:
        pairs["q"] = (q_a, q_b)
    for k, (a, b) in pairs.items():
        d = np.abs(np.asarray(a) - np.asarray(b))
        s = float(np.max(np.abs(b))) or 1.0
        out[k + "_linf"] = float(np.max(d) / s)
        rms = float(np.sqrt(np.mean(np.asarray(b) ** 2))) or 1.0
        out[k + "_l2"] = float(np.sqrt(np.mean(d ** 2)) / rms)
    return out


def mass_drift(m0: float, m1: float) -> float:
    """|M(t) - M0| / M0 with M = bulk + adsorbed mass (ugks_sim.py:476-485)."""
    return abs(m1 - m0) / abs(m0)
