"""Pins of the oracle's interior flux (iFace, getPLR, initMacroFlux,
calcFaceQ, macroFlux, distFlux: clUGKS.c:110-772) by limits and states in
which the UGKS flux has a closed form that does not use the flux code:

  free-molecular limit (tau/dt -> inf)    every velocity's flux -> L (dt u f_face
                                          - dt^2/2 u^2 sigma), sigma the slope the
                                          reference stores (along the upwind
                                          direction), with an O(dt/tau) error
  uniform heat-flux state (Grad 13)       energy flux L (Mt3 + (1 - Pr) Mt0) q_n,
                                          normal momentum flux L dt p, no mass
                                          flux; the update relaxes q by
                                          (1 + dt (1 - 2 Pr)/2tau) / (1 + dt/2tau)
  shear wave, continuum limit             decay at the Navier-Stokes rate
                                          nu k^2 with nu = tau p / rho
  sound wave, continuum limit             speed sqrt(gamma p / rho) and the
                                          Navier-Stokes/Fourier damping with the
                                          Shakhov conductivity (Pr scaled)

Velocities are Gauss-Hermite where a moment must be exact (the weighted
Maxwellian at U = 0, lambda = 1 is then integrated to rounding).
"""
import copy
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from pyugks_amd import bc as bcd
from pyugks_amd import cases
from pyugks_amd.data import FlowCondition, cartesian_superblock

H = O.helpers()
SOUTH, WEST = 0, 1


def _lib():
    L = O.lib()
    L.or_block_flux.argtypes = [C.c_void_p, C.c_double]
    L.or_block_flux.restype = C.c_int
    L.or_calcMacro.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    return L


def periodic_box(ni, nj, vel, method, Kn, Pr=2.0 / 3.0, fill=None, CFL=0.8):
    g = cases._base(method, Kn, CFL=CFL, Pr=Pr)
    kind, n = vel
    (cases.gauss_hermite if kind == "GH" else cases.newton_cotes)(g, n)
    bcs = [bcd.PeriodicBC(0, bcd.SOUTH), bcd.PeriodicBC(0, bcd.WEST), bcd.PeriodicBC(0, bcd.NORTH),
           bcd.PeriodicBC(0, bcd.EAST)]
    g.blocks = cartesian_superblock(0.0, 1.0, 0.0, 1.0, ni, nj, bcs, fill or FlowCondition(D=1.0, T=1.0))
    return O.OracleSim(g, threads=4)


def velocity_table(o):
    q, w, _, _, _ = o.blocks[0].tables
    return np.asarray(q, dtype=np.float64).reshape(-1, 2), np.asarray(w, dtype=np.float64)


def maxwellian(o, rho=1.0, U=0.0, V=0.0, lam=1.0):
    """The weighted discrete Maxwellian (g, h) the reference uses (fM,
    clFunctions.c:49-58), written out here from its closed form."""
    q, w = velocity_table(o)
    K = o.blocks[0].consts["K"]
    g = w * rho * (lam / math.pi) * np.exp(-lam * ((q[:, 0] - U) ** 2 + (q[:, 1] - V) ** 2))
    return g, g * K / (2.0 * lam)


def cell_centres(o):
    """Centres of every cell, ghosts included, continuing the uniform grid."""
    b = o.blocks[0]
    c = b.consts
    G, NI, NJ, ni, nj = c["ghost"], c["NI"], c["NJ"], c["ni"], c["nj"]
    x = (np.arange(NI) - G + 0.5) / ni
    y = (np.arange(NJ) - G + 0.5) / nj
    return np.meshgrid(x, y, indexing="ij")


# ------------------------------------------------------ free-molecular limit --
def _free_molecular_error(method, Kn):
    o = periodic_box(8, 6, ("NC", 8), method, Kn)
    b = o.blocks[0]
    c = b.consts
    G, ni, nj = c["ghost"], c["ni"], c["nj"]
    gM, hM = maxwellian(o, U=0.2, V=-0.1)
    a, bb = 0.3, -0.2                       # f = M (1 + a x + b y): linear in space
    X, Y = cell_centres(o)
    lin = 1.0 + a * X + bb * Y
    b["f"][..., 0] = lin[..., None] * gM
    b["f"][..., 1] = lin[..., None] * hM
    L = _lib()
    L.or_calcMacro(C.byref(b.block.p), b["f"].ctypes.data, b["macro"].ctypes.data)
    dt = o.dt
    assert L.or_block_flux(C.byref(b.block), dt) == 0
    q, _ = velocity_table(o)
    cx, cy = q[:, 0], q[:, 1]

    def moments(x):          # (mass, x-momentum, y-momentum, energy) of per-velocity (g, h) values
        return np.array([np.sum(x[:, 0]), np.sum(cx * x[:, 0]), np.sum(cy * x[:, 0]),
                         0.5 * np.sum((cx * cx + cy * cy) * x[:, 0] + x[:, 1])])

    err = {"f": 0.0, "macro": 0.0}
    sig_scale = {"f": 0.0, "macro": 0.0}
    main_scale = 0.0
    for face, buf, fm in ((SOUTH, "flux_f_S", "flux_macro_S"), (WEST, "flux_f_W", "flux_macro_W")):
        ii = range(G, G + ni) if face == SOUTH else range(G, G + ni + 1)
        jj = range(G, G + nj + 1) if face == SOUTH else range(G, G + nj)
        for i in ii:
            for j in jj:
                n = b["normal"][i, j, face]
                L_ = b["side_length"][i, j, face]
                mid = b["mid_side"][i, j, face]
                un = q @ n
                fac_face = 1.0 + a * mid[0] + bb * mid[1]
                dfdn = a * n[0] + bb * n[1]                     # slope along +n of (1 + a x + b y)
                # the reference stores the slope along the upwind direction:
                # +n for u_n > 0, -n for u_n <= 0 (the owner cell is upwind)
                sig = np.where(un > 0.0, dfdn, -dfdn)
                want = np.empty((len(un), 2))
                sterm = np.empty((len(un), 2))
                for k, M in enumerate((gM, hM)):
                    sterm[:, k] = -0.5 * dt * dt * un * un * sig * M * L_
                    want[:, k] = L_ * dt * un * fac_face * M + sterm[:, k]
                err["f"] = max(err["f"], float(np.max(np.abs(b[buf][i, j] - want))))
                sig_scale["f"] = max(sig_scale["f"], float(np.max(np.abs(sterm))))
                main_scale = max(main_scale, float(np.max(np.abs(want))))
                # macroFlux: the same limit summed with (1, c_x, c_y, energy) in the global frame
                err["macro"] = max(err["macro"], float(np.max(np.abs(b[fm][i, j] - moments(want)))))
                sig_scale["macro"] = max(sig_scale["macro"], float(np.max(np.abs(moments(sterm)))))
    worst = max(err[k] / sig_scale[k] for k in err)
    return worst, sig_scale["f"] / main_scale, dt


@pytest.mark.parametrize("method", ["vanLeer", "WENO5", "minmod"])
def test_free_molecular_limit_of_the_face_flux(method):
    """tau/dt ~ 1e4 and ~1e5: the equilibrium terms (Mt0, Mt1, Mt2) vanish
    as dt/tau and Mt3 -> dt, Mt4 -> dt^2/2.  On a linear f every
    reconstruction is exact, so each velocity's flux is L (dt u f_face -
    dt^2/2 u^2 sigma).  The error must be small against the slope term and
    fall about tenfold with tau: a dropped or sign-flipped slope term, a
    slope taken along +n for both signs of u_n, or a wrong face value
    fails by O(1)."""
    e3, ratio, dt = _free_molecular_error(method, 100.0)
    e4, _, _ = _free_molecular_error(method, 1000.0)
    assert ratio > 1e-3                 # the slope term is resolved against the transport term
    assert e4 < 2e-3, (e3, e4)
    assert 5.0 < e3 / e4 < 20.0, (e3, e4)


# ------------------------------------------------- uniform heat-flux state --
def _heat_flux_state(Kn, q, Pr=2.0 / 3.0):
    """Periodic box filled with f = M (1 + 2 (c.q)(c^2 - 2)) (g and h alike)
    at rho = 1, U = 0, lambda = 1: on Gauss-Hermite velocities its discrete
    density, momentum and energy are those of M, and its heat flux is q,
    exactly (moments of e^{-c^2} times polynomials of degree <= 7)."""
    o = periodic_box(6, 5, ("GH", 12), "WENO5", Kn, Pr=Pr)
    b = o.blocks[0]
    gM, hM = maxwellian(o)
    qv, _ = velocity_table(o)
    c2 = qv[:, 0] ** 2 + qv[:, 1] ** 2
    phi = 2.0 * (qv @ np.asarray(q)) * (c2 - 2.0)
    b["f"][..., 0] = gM * (1.0 + phi)
    b["f"][..., 1] = hM * (1.0 + phi)
    _lib().or_calcMacro(C.byref(b.block.p), b["f"].ctypes.data, b["macro"].ctypes.data)
    return o, qv, gM, hM


def _heat_flux(f, qv):
    return 0.5 * np.einsum("...v,vk->...k", (qv[:, 0] ** 2 + qv[:, 1] ** 2) * f[..., 0] + f[..., 1], qv)


def test_heat_flux_state_moments_are_exact():
    o, qv, gM, hM = _heat_flux_state(0.01, (0.03, -0.05))
    f = o.blocks[0]["f"][3, 3]
    assert np.allclose(_heat_flux(f, qv), [0.03, -0.05], rtol=0, atol=1e-15)
    m = o.blocks[0]["macro"][2, 2]
    assert np.allclose(m[:3], [1.0, 0.0, 0.0], rtol=0, atol=1e-15)


@pytest.mark.parametrize("Kn", [0.003, 0.05])
def test_uniform_heat_flux_face_fluxes(Kn):
    """Uniform state carrying a heat flux q: the face values are f itself and
    every slope vanishes, so the UGKS flux is L [Mt0 u (g0 + fS(q)) + Mt3 u f]
    with Mt0 + Mt3 = dt.  Its moments are mass 0, momentum L dt p n and
    energy L (Mt3 + (1 - Pr) Mt0) q.n: the Shakhov term of the equilibrium
    carries (1 - Pr) q (calcFaceQ feeding fS), the free transport all of q.
    Per velocity, distFlux minus the two Maxwell and free-transport parts
    (written out here) must have exactly the Shakhov moments.  A dropped
    Shakhov term, a BGK heat flux (Pr = 1), a face heat flux taken against
    the wrong velocity or a wrong sign fails."""
    q = np.array([0.03, -0.05])
    Pr = 2.0 / 3.0
    o, qv, gM, hM = _heat_flux_state(Kn, q, Pr=Pr)
    b = o.blocks[0]
    c = b.consts
    G = c["ghost"]
    f = b["f"].copy()
    dt = o.dt
    assert _lib().or_block_flux(C.byref(b.block), dt) == 0
    prim = b["prim"][G + 2, G + 2]
    assert np.allclose(prim[:3], [1.0, 0.0, 0.0], atol=1e-14)
    tau = H.or_relaxTime(C.byref(b.block.p), O.V4(*prim))
    assert 0.1 < tau / dt < 10.0
    # UGKS time integrals (clUGKS.c:499-503), pinned separately against quadrature
    Mt3 = tau * (1.0 - math.exp(-dt / tau))
    Mt0 = dt - Mt3
    p = prim[0] / (2.0 * prim[3])
    g0, h0 = maxwellian(o, *prim)          # the face Maxwellian
    for face, buf, fm in ((SOUTH, "flux_f_S", "flux_macro_S"), (WEST, "flux_f_W", "flux_macro_W")):
        for (i, j) in ((G + 1, G + 2), (G + 3, G + 1), (G, G)):
            n = b["normal"][i, j, face]
            L_ = b["side_length"][i, j, face]
            want = L_ * np.array([0.0, dt * p * n[0], dt * p * n[1], (Mt3 + (1.0 - Pr) * Mt0) * (q @ n)])
            assert np.allclose(b[fm][i, j], want, rtol=0, atol=1e-10 * L_ * dt), (face, b[fm][i, j], want)
            un = qv @ n
            rest = b[buf][i, j] - L_ * Mt3 * un[:, None] * f[i, j] \
                - L_ * Mt0 * un[:, None] * np.stack([g0, h0], axis=1)
            cx, cy = qv[:, 0], qv[:, 1]
            mom = np.array([np.sum(rest[:, 0]), np.sum(cx * rest[:, 0]), np.sum(cy * rest[:, 0]),
                            0.5 * np.sum((cx * cx + cy * cy) * rest[:, 0] + rest[:, 1])])
            wantS = L_ * Mt0 * np.array([0.0, 0.0, 0.0, (1.0 - Pr) * (q @ n)])
            assert np.allclose(mom, wantS, rtol=0, atol=1e-10 * L_ * dt), (face, mom, wantS)


def test_uniform_heat_flux_relaxes_at_the_shakhov_rate():
    """UGKS_update with no net flux (uniform state, periodic): both
    equilibria carry the old heat flux through fS, so each step maps q to
    q (1 + dt (1 - 2 Pr) / (2 tau)) / (1 + dt / (2 tau)), while density,
    momentum and energy stay put.  BGK relaxation (Pr = 1), the new Q in
    f_eq^{n+1} or a dropped fS term gives another factor."""
    q = np.array([0.04, 0.02])
    Pr = 2.0 / 3.0
    o, qv, gM, hM = _heat_flux_state(0.05, q, Pr=Pr)
    b = o.blocks[0]
    G = b.consts["ghost"]
    m0 = b["macro"].copy()
    tau = H.or_relaxTime(C.byref(b.block.p), O.V4(*m0[0, 0]))
    fac = 1.0
    for _ in range(3):
        dt = o.dt
        o.one_step(True)
        fac *= (1.0 + dt * (1.0 - 2.0 * Pr) / (2.0 * tau)) / (1.0 + dt / (2.0 * tau))
    assert 0.05 < fac < 0.8                      # the heat flux really decays
    got = _heat_flux(b["f"][G:-G, G:-G], qv)
    assert np.allclose(got, fac * q, rtol=0, atol=1e-13), (got[0, 0], fac * q)
    assert np.allclose(b["macro"], m0, rtol=0, atol=1e-13)


# --------------------------------------------------------- continuum limit --
def _periodic_waves(ni, nj, Kn, init, Pr=2.0 / 3.0, dt=0.005):
    """A truly periodic box (orientation HOLD: with the reference's default
    NO_HOLD a periodic join of a rectangle fails the corner-alignment check
    of edgeExchange, clGhostRoutines.c:491-512, and its ghost rows are
    flipped), Gauss-Hermite 12^2, WENO5, a fixed time step."""
    g = cases._base("WENO5", Kn, Pr=Pr)
    cases.gauss_hermite(g, 12)
    bcs = [bcd.PeriodicBC(0, f, orientation=bcd.HOLD) for f in (bcd.SOUTH, bcd.WEST, bcd.NORTH, bcd.EAST)]
    g.blocks = cartesian_superblock(0.0, 1.0, 0.0, 1.0, ni, nj, bcs, FlowCondition(D=1.0, T=1.0), init_macro=init)
    o = O.OracleSim(g, threads=4)
    o.sim.force_time = 1
    o.sim.dt = dt
    return o


def _shear_viscosity(Kn, steps=400, nj=32, A=0.01):
    ni = 4
    y = (np.arange(nj) + 0.5) / nj
    init = np.zeros((ni, nj, 4))
    init[..., 0] = 1.0
    init[..., 1] = A * np.sin(2 * np.pi * y)[None, :]
    init[..., 3] = 1.0
    o = _periodic_waves(ni, nj, Kn, init)
    amp = lambda: 2.0 * np.mean(o.blocks[0]["macro"][:, :, 1] * np.sin(2 * np.pi * y)[None, :])
    a0 = amp()
    o.run(steps)
    t = steps * o.sim.dt
    return -math.log(amp() / a0) / ((2 * np.pi) ** 2 * t)


@pytest.mark.parametrize("Kn", [5e-6, 2e-5])
def test_shear_wave_decays_at_the_navier_stokes_rate(Kn):
    """U = A sin(2 pi y) in a periodic box with tau/dt = 1e-3 and 4e-3: the
    UGKS flux's slope terms (Mt1 u^2 a g0 with a from getPLR/microSlope)
    carry the Navier-Stokes shear stress, so the wave decays as
    exp(-nu k^2 t) with nu = tau p / rho = tau / 2.  A dropped or sign-flipped
    slope term gives no decay or growth; slopes scaled by a wrong distance
    give a wrong rate."""
    nu = _shear_viscosity(Kn)
    assert abs(nu / (0.5 * Kn) - 1.0) < 0.01, nu / (0.5 * Kn)


def _sound(Kn, steps=200, ni=64, eps=1e-3, gam=5.0 / 3.0):
    nj = 4
    x = (np.arange(ni) + 0.5) / ni
    c = math.sqrt(gam * 0.5)
    cs = np.cos(2 * np.pi * x)
    rho = 1.0 + eps * cs
    p = 0.5 * (1.0 + gam * eps * cs)
    init = np.zeros((ni, nj, 4))
    init[..., 0] = rho[:, None]
    init[..., 1] = (eps * c * cs)[:, None]          # right-running acoustic wave
    init[..., 3] = (p / (0.5 * rho))[:, None]       # T (the builder stores 1/T)
    o = _periodic_waves(ni, nj, Kn, init)
    e = np.exp(-2j * np.pi * x)
    amp = lambda: 2.0 * np.mean((o.blocks[0]["macro"][:, 0, 0] - 1.0) * e)
    a, ts = [amp()], [0.0]
    for _ in range(10):
        o.run(steps // 10)
        a.append(amp())
        ts.append(ts[-1] + o.sim.dt * (steps // 10))
    a, ts = np.array(a), np.array(ts)
    alpha = -np.polyfit(ts, np.log(np.abs(a)), 1)[0]
    omega = -np.polyfit(ts, np.unwrap(np.angle(a)), 1)[0]
    return alpha, omega / (2 * np.pi) / c


def test_sound_wave_speed_and_damping():
    """A right-running acoustic wave (1e-3 amplitude, 64 cells per
    wavelength) in the continuum limit: it travels at sqrt(gamma p / rho)
    to 0.1%, and its damping is proportional to tau (the viscous and
    conductive fluxes), with a tau-independent part under a tenth of the
    physical damping at tau/dt = 4e-3.  Dropping the time slope aT
    (initMacroFlux's Mt2 term, which makes the flux second order in time) or
    taking it with the wrong sign gives an O(dt c^2 k^2) damping or growth
    of its own; swapped upwind slopes aL/aR change the damping."""
    a1, c1 = _sound(2e-5)
    a2, c2 = _sound(1e-4)
    slope = (a2 - a1) / (1e-4 - 2e-5)
    a0 = a1 - slope * 2e-5
    assert abs(c1 - 1.0) < 1e-3 and abs(c2 - 1.0) < 1e-3, (c1, c2)
    assert 20.0 < slope < 35.0, slope
    assert abs(a0) < 0.1 * slope * 2e-5, (a0, slope)


# ------------------------------------------------ upwind choice of slopes --
def _bind_face_kernels():
    L = O.lib()
    P = C.c_void_p
    L.or_iFace.argtypes = [P, P, P, P, P, P, P, C.c_int]
    L.or_initMacroFlux.argtypes = [P, P, C.c_int, C.c_double, P, P, P, P, P, C.c_int, C.c_int]
    L.or_macroFlux.argtypes = [P, P, P, P, P, P, C.c_int, C.c_double, P, P, C.c_int, C.c_int]
    L.or_distFlux.argtypes = [P, P, P, P, P, C.c_int, C.c_double, P, P, P, P, P, P, C.c_int, C.c_int]
    return L


def _face_pass_with(b, dt, aL=None, aR=None):
    """Re-run the W-face flux kernels of a block whose face state buffers
    hold a W pass, with the upwind slopes replaced: iFace (face values),
    initMacroFlux (time slope, equilibrium macro flux), macroFlux, distFlux.
    Returns (per-velocity flux, macro flux, time slope)."""
    L = _bind_face_kernels()
    p = C.byref(b.block.p)
    a = lambda n: b[n].ctypes.data
    if aL is not None:
        b["aL"][:] = aL
    if aR is not None:
        b["aR"][:] = aR
    L.or_iFace(p, a("f"), a("flux_f_W"), a("sigma"), a("centre"), a("mid_side"), a("normal"), WEST)
    L.or_initMacroFlux(p, a("flux_macro_W"), WEST, dt, a("prim"), a("aL"), a("aR"), a("aT"), a("Mxi"), 0, 0)
    L.or_macroFlux(p, a("flux_f_W"), a("sigma"), a("flux_macro_W"), a("normal"), a("side_length"), WEST, dt,
                   a("prim"), a("faceQ"), 0, 0)
    L.or_distFlux(p, a("flux_f_W"), a("sigma"), a("normal"), a("side_length"), WEST, dt, a("prim"), a("aL"),
                  a("aR"), a("aT"), a("Mxi"), a("faceQ"), 0, 0)
    return b["flux_f_W"].copy(), b["flux_macro_W"].copy(), b["aT"].copy()


def _half_range_moments(prim, K, n=801, lim=10.0):
    """<psi_i psi_j u_n^k M> over u_n < 0 and over all u_n (k = 0, 1, 2),
    psi = (1, u_n, u_t, (u_n^2 + u_t^2 + xi^2)/2), for the Maxwellian of a
    face state at rest in the face frame: Gauss-Legendre on each half line of
    u_n, a fine trapezoid rule along u_t (spectrally accurate for Gaussians),
    and the xi moments <xi^2> = K/(2 lam), <xi^4> = K (K + 2)/(4 lam^2) in
    closed form."""
    rho, lam = prim[0], prim[3]
    x, wx = np.polynomial.legendre.leggauss(120)
    un = np.concatenate([0.5 * lim * (x - 1.0), 0.5 * lim * (x + 1.0)])     # [-lim, 0] then [0, lim]
    wn = np.concatenate([0.5 * lim * wx, 0.5 * lim * wx])
    ut = np.linspace(-lim, lim, n)
    wt = np.full(n, ut[1] - ut[0])
    Un, Ut = np.meshgrid(un, ut, indexing="ij")
    M = rho * (lam / math.pi) * np.exp(-lam * (Un ** 2 + Ut ** 2)) * np.outer(wn, wt)
    x2, x4 = K / (2 * lam), K * (K + 2) / (4 * lam * lam)
    c2 = Un ** 2 + Ut ** 2
    # psi_i psi_j averaged over xi: the energy entries carry the xi moments
    base = [np.ones_like(Un), Un, Ut]

    def pp(i, j):
        if i < 3 and j < 3:
            return base[i] * base[j]
        if i < 3:
            return base[i] * 0.5 * (c2 + x2)
        if j < 3:
            return base[j] * 0.5 * (c2 + x2)
        return 0.25 * (c2 * c2 + 2 * c2 * x2 + x4)

    neg = np.zeros_like(Un, dtype=bool)
    neg[:len(x)] = True
    out = {}
    for k in (0, 1, 2):
        w = M * Un ** k
        out["full", k] = np.array([[np.sum(w * pp(i, j)) for j in range(4)] for i in range(4)])
        out["neg", k] = np.array([[np.sum(w * neg * pp(i, j)) for j in range(4)] for i in range(4)])
    return out


@pytest.mark.parametrize("side", ["R", "L"])
def test_flux_takes_each_side_slope_from_its_upwind_cells(side):
    """Changing the right-hand slope aR by da (every face) changes distFlux
    only for u_n <= 0, by L Mt1 u_n^2 (da . psi) times the face Maxwellian,
    psi = (1, u_n, u_t, energy); changing aL changes only u_n > 0.  In
    initMacroFlux the change is rho [Mt1 <u_n^2 (da.psi) psi M>_half +
    Mt2 <u_n (dA.psi) psi M>], dA the time-slope response to
    sw = -rho <u_n (da.psi) psi M>_half (moments here by a fine quadrature
    of my own, the linear solve here by numpy).  Slopes from the downwind
    side (aL and aR swapped) fail both."""
    o = periodic_box(6, 5, ("GH", 12), "WENO5", 0.01, fill=FlowCondition(D=1.0, T=1.0))
    b = o.blocks[0]
    c = b.consts
    G, ni, nj, K = c["ghost"], c["ni"], c["nj"], c["K"]
    dt = o.dt
    _lib().or_block_flux(C.byref(b.block), dt)          # fills the W-face state buffers
    a0L, a0R = b["aL"].copy(), b["aR"].copy()
    da = np.array([0.3, -0.2, 0.15, 0.05])
    F1, FM1, T1 = _face_pass_with(b, dt, aL=a0L, aR=a0R)
    if side == "R":
        F2, FM2, T2 = _face_pass_with(b, dt, aL=a0L, aR=a0R + da)
    else:
        F2, FM2, T2 = _face_pass_with(b, dt, aL=a0L + da, aR=a0R)
    q, _ = velocity_table(o)
    checked = 0
    for i in range(G, G + ni + 1):
        for j in range(G, G + nj):
            n = b["normal"][i, j, WEST]
            L_ = b["side_length"][i, j, WEST]
            prim = b["prim"][i, j]
            assert abs(prim[1]) < 1e-12 and abs(prim[2]) < 1e-12      # a gas at rest
            tau = H.or_relaxTime(C.byref(b.block.p), O.V4(*prim))
            Mt = np.zeros(5)
            H.or_timeIntegrals(tau, dt, Mt.ctypes.data_as(C.POINTER(C.c_double)))
            un, ut = q @ n, q @ np.array([-n[1], n[0]])
            g0, h0 = maxwellian(o, prim[0], 0.0, 0.0, prim[3])
            cc = un * un + ut * ut
            # (da . psi) for g: da0 + da1 u_n + da2 u_t + da3 |c|^2/2, and h takes da3 <xi^2>/2 = da3 h/g/2
            pg = da[0] + da[1] * un + da[2] * ut + 0.5 * da[3] * cc
            upwind = (un <= 0.0) if side == "R" else (un > 0.0)
            want_g = L_ * Mt[1] * un * un * (pg * g0 + 0.5 * da[3] * h0) * upwind
            # the time slope's response adds Mt2 u (dA . psi) g0 on every velocity: fitted below
            d = F2[i, j, :, 0] - F1[i, j, :, 0] - want_g
            dA = T2[i, j] - T1[i, j]
            resp = L_ * Mt[2] * un * ((dA[0] + dA[1] * un + dA[2] * ut + 0.5 * dA[3] * cc) * g0 + 0.5 * dA[3] * h0)
            scale = np.max(np.abs(want_g))
            assert np.max(np.abs(d - resp)) < 1e-11 * scale, (i, j)
            # the time slope and the macro flux against independent half-range moments
            mom = _half_range_moments(prim, K)
            half = lambda k: mom["neg", k] if side == "R" else mom["full", k] - mom["neg", k]
            # microSlope: <(dA . psi) psi M> = sw with sw = -<u_n (da . psi) psi M>_half
            sw = -(half(1) @ da)
            dA_want = np.linalg.solve(mom["full", 0], sw)
            assert np.allclose(dA, dA_want, rtol=1e-9, atol=1e-12 * np.max(np.abs(dA_want))), (dA, dA_want)
            dFM = Mt[1] * (half(2) @ da) + Mt[2] * (mom["full", 1] @ dA_want)
            rot = np.array([dFM[0], dFM[1] * n[0] - dFM[2] * n[1], dFM[1] * n[1] + dFM[2] * n[0], dFM[3]]) * L_
            got = FM2[i, j] - FM1[i, j]
            assert np.allclose(got, rot, rtol=1e-8, atol=1e-10 * np.max(np.abs(rot))), (got, rot)
            checked += 1
    assert checked == (ni + 1) * nj
