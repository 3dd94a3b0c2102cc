"""Accuracy and conservation measures reported with every result."""
from __future__ import annotations

import numpy as np


def _flat(fields):
    """One (cells, 4) array from a macro field (ni, nj, 4) or a list of them
    (the blocks of one case)."""
    if isinstance(fields, (list, tuple)):
        return np.concatenate([np.asarray(x).reshape(-1, 4) for x in fields])
    return np.asarray(fields).reshape(-1, 4)


def field_errors(macro_a, macro_b, q_a=None, q_b=None):
    """Relative L-inf and L2 differences of rho, U (both components), T and q
    between two macro fields holding (rho, U, V, 1/T): each a (ni, nj, 4)
    array or a list of them covering all blocks of a case.  The scale of a
    quantity is its maximum (L-inf) or RMS (L2) over the whole case, so a
    block the flow has not reached yet does not blow up the ratio."""
    a, b = _flat(macro_a), _flat(macro_b)
    out = {}
    pairs = {"rho": (a[:, 0], b[:, 0]), "U": (a[:, 1:3], b[:, 1:3]), "T": (1.0 / a[:, 3], 1.0 / b[:, 3])}
    if q_a is not None:
        qa = np.concatenate([np.ravel(x) for x in q_a]) if isinstance(q_a, (list, tuple)) else np.ravel(q_a)
        qb = np.concatenate([np.ravel(x) for x in q_b]) if isinstance(q_b, (list, tuple)) else np.ravel(q_b)
        pairs["q"] = (qa, qb)
    for k, (x, y) in pairs.items():
        d = np.abs(x - y)
        s = float(np.max(np.abs(y))) or 1.0
        out[k + "_linf"] = float(np.max(d) / s)
        rms = float(np.sqrt(np.mean(y ** 2))) or 1.0
        out[k + "_l2"] = float(np.sqrt(np.mean(d ** 2)) / rms)
    return out


def mass_drift(m0: float, m1: float) -> float:
    """|M(t) - M0| / M0 with M = bulk + adsorbed mass (ugks_sim.py:476-485)."""
    return abs(m1 - m0) / abs(m0)
