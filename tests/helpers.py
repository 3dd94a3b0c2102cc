"""Comparison helpers shared by the tests."""
import numpy as np

# block buffer name on the GPU host object -> oracle buffer name
BUFFERS = {
    "f_D": "f", "flux_f_S_D": "flux_f_S", "flux_f_W_D": "flux_f_W", "sigma_D": "sigma", "macro_D": "macro",
    "Q_D": "Q", "flux_macro_S_D": "flux_macro_S", "flux_macro_W_D": "flux_macro_W", "prim_D": "prim",
    "aL_D": "aL", "aR_D": "aR", "aT_D": "aT", "faceQ_D": "faceQ", "Mxi_D": "Mxi", "area_D": "area",
    "centre_D": "centre", "side_D": "mid_side", "normal_D": "normal", "length_D": "side_length",
    "wall_prop_D": "wall_prop", "wall_cover_D": "wall_cover", "wall_fluxes_D": "wall_fluxes",
    "wall_dist_D": "wall_dist", "time_step_D": "time_step", "residual_D": "residual",
}


# heat fluxes vanish at equilibrium; their errors are measured against the
# heat-flux scale of the unit state (rho = 1, T = 1) rather than their own size
SCALE_FLOOR = {"Q": 1e-3, "faceQ": 1e-3, "Q_D": 1e-3, "faceQ_D": 1e-3}


def rel_err(a, b, floor=0.0):
    """max|a-b| / max(max|b|, floor) over entries finite in both; NaN/inf positions must agree."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    fa, fb = np.isfinite(a), np.isfinite(b)
    if not np.array_equal(fa, fb):
        return np.inf
    if not fa.any():
        return 0.0
    scale = max(np.max(np.abs(b[fb])), floor)
    d = np.max(np.abs(a[fa] - b[fb]))
    if scale == 0.0:
        return float(d)
    return float(d / scale)


def field_errors(macro_a, macro_b, q_a=None, q_b=None):
    """Relative L-inf and L2 errors of rho, U (vector), T and q."""
    out = {}
    ra, rb = macro_a[..., 0], macro_b[..., 0]
    ua, ub = macro_a[..., 1:3], macro_b[..., 1:3]
    ta, tb = 1.0 / macro_a[..., 3], 1.0 / macro_b[..., 3]
    pairs = {"rho": (ra, rb), "U": (ua, ub), "T": (ta, tb)}
    if q_a is not None:
        pairs["q"] = (q_a, q_b)
    for k, (a, b) in pairs.items():
        d = np.abs(a - b)
        s = np.max(np.abs(b)) or 1.0
        out[k + "_linf"] = float(np.max(d) / s)
        out[k + "_l2"] = float(np.sqrt(np.mean(d ** 2)) / (np.sqrt(np.mean(b ** 2)) or 1.0))
    return out
