"""Pins of the oracle's getInternal (clKernel.c:315-372), the diagnostics
written when save_options.internal_data is set: per cell (Tx, Ty, Tz, T) and
(tau_xx, tau_xy, tau_yy, 0) of the current f.

The distributions are sheared, drifting Gaussians sampled on a Gauss-Hermite
velocity grid, whose moments are known in closed form: with covariance S,
mean (U, V) and internal part h = f.y / f.x, Tx = 2 S_xx, Ty = 2 S_yy,
Tz = 2 h, T = (Tx + Ty + Tz) / B, and tau_ab = rho (S_ab - T/2).  The T/2 in
tau_xy (not zero) is the reference's: its stress takes the isotropic part off
every component.  Each cell gets its own parameters, so a swapped axis, a
dropped mean or a wrong cell index fails."""
import ctypes as C

import numpy as np

import oracle as O
from pyugks_amd import cases

H = O.helpers()


def _block(vel=("GH", 24)):
    g = cases.cavity(3, vel, "vanLeer", 0.05, n_cells_j=2)
    return O.OracleSim(g, threads=2).blocks[0]


def _gaussian(q, w, rho, U, V, S):
    """rho times the normal density of mean (U, V), covariance S, at the
    velocities q, times the quadrature weights."""
    c = q - np.array([U, V])
    Si = np.linalg.inv(S)
    e = np.einsum("ka,ab,kb->k", c, Si, c)
    return w * rho * np.exp(-0.5 * e) / (2 * np.pi * np.sqrt(np.linalg.det(S)))


def test_sheared_gaussians_give_their_closed_form_temperatures_and_stress():
    b = _block()
    q, w = b.tables[0], b.tables[1]
    c = b.consts
    G, ni, nj, B = c["ghost"], c["ni"], c["nj"], c["B"]
    rng = np.random.default_rng(5)
    want_T, want_S = np.zeros((ni, nj, 4)), np.zeros((ni, nj, 4))
    f = b["f"]
    f[:] = 1.0
    for i in range(ni):
        for j in range(nj):
            rho, U, V = rng.uniform(0.6, 1.4), rng.uniform(-0.2, 0.2), rng.uniform(-0.2, 0.2)
            sxx, syy = rng.uniform(0.4, 0.6), rng.uniform(0.4, 0.6)
            sxy = rng.uniform(-0.08, 0.08)
            h = rng.uniform(0.3, 0.7)
            S = np.array([[sxx, sxy], [sxy, syy]])
            g = _gaussian(q, w, rho, U, V, S)
            f[i + G, j + G, :, 0] = g
            f[i + G, j + G, :, 1] = h * g
            Tx, Ty, Tz = 2 * sxx, 2 * syy, 2 * h
            T = (Tx + Ty + Tz) / B
            want_T[i, j] = (Tx, Ty, Tz, T)
            want_S[i, j] = (rho * (sxx - T / 2), rho * (sxy - T / 2), rho * (syy - T / 2), 0.0)
    txyz, stress = b.internal()
    assert np.allclose(txyz, want_T, rtol=1e-10, atol=1e-12)
    assert np.allclose(stress, want_S, rtol=1e-10, atol=1e-11)
    # the ghost cells are not read: garbage there changes nothing
    f[:G] = np.nan
    f[:, -G:] = -7.0
    t2, s2 = b.internal()
    assert np.array_equal(t2, txyz) and np.array_equal(s2, stress)


def test_maxwellian_gives_one_temperature():
    """The oracle's own discrete Maxwellian (fM, pinned by its moments):
    Tx = Ty = 1/lambda, Tz = K/lambda exactly (f.y = K/(2 lambda) f.x per
    velocity), T = 1/lambda since B = K + 2, normal stresses zero and
    tau_xy = -rho/(2 lambda)."""
    b = _block()
    q, w = b.tables[0], b.tables[1]
    c = b.consts
    G, ni, nj, K = c["ghost"], c["ni"], c["nj"], c["K"]
    assert abs(c["B"] - (K + 2)) < 1e-12
    P = C.byref(b.block.p)
    f = b["f"]
    prims = {}
    for i in range(ni):
        for j in range(nj):
            prim = (1.0 + 0.1 * i, 0.05 * j - 0.03, 0.02 * i, 0.8 + 0.1 * (i + j))
            prims[i, j] = prim
            for k in range(len(w)):
                r = H.or_fM(P, O.V4(*prim), O.V2(q[k, 0], q[k, 1]), k)
                f[i + G, j + G, k] = (r.x, r.y)
    txyz, stress = b.internal()
    for (i, j), (rho, U, V, lam) in prims.items():
        assert np.allclose(txyz[i, j, [0, 1, 3]], 1.0 / lam, rtol=1e-11)
        assert abs(txyz[i, j, 2] - K / lam) <= 1e-13 * K / lam
        assert np.allclose(stress[i, j, [0, 2]], 0.0, atol=1e-11)
        assert abs(stress[i, j, 1] + rho / (2 * lam)) < 1e-11
        assert stress[i, j, 3] == 0.0
