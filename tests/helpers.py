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


from pyugks_amd.metrics import field_errors  # noqa: E402,F401
