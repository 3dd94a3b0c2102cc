"""Case builders for the UGKS workloads (lid-driven cavities, adsorbing channel).

Every case is stated directly in non-dimensional units: the reference
values are L_ref = D_ref = T_ref = 1 and R = 1/2, so C_ref = sqrt(2 R T_ref)
= 1 and t_ref = 1 (ugks_data.py:357-371) and the non-dimensionalisation
(ugks_data.py:462-542) leaves D, U, V, T and lengths unchanged.

Settings a case does not name keep the reference defaults: omega = 0.81,
relax_type = 2, delta_dt = 1, dt recomputed every step
(ugks_data.py:170-196).  Pr = 2/3 is used for every cavity (the Shakhov
value the first cavity names).  max_time is infinite: runs stop on a step
count.
"""
from __future__ import annotations

import copy

import numpy as np

from . import bc as bcd
from .data import BlockDef, FlowCondition, UGKSData, cartesian_superblock

SEED = 20261202


def _base(flux_method: str, Kn: float, CFL: float = 0.8, Pr: float = 2.0 / 3.0) -> UGKSData:
    g = UGKSData()
    g.R, g.T_ref, g.D_ref, g.L_ref = 0.5, 1.0, 1.0, 1.0
    g.gamma = 5.0 / 3.0
    g.Pr = Pr
    g.Kn_eff = Kn
    g.flux_method = flux_method
    g.CFL = CFL
    g.max_time = float("inf")
    g.max_step = 1 << 30
    return g


def gauss_hermite(g: UGKSData, n: int):
    """n x n Gauss-Hermite velocities (numpy's hermgauss nodes and weights)."""
    x, w = np.polynomial.hermite.hermgauss(n)
    g.quad_type = "Gauss"
    g.quad = list(x)
    g.weight = list(w)
    g.u_mid = g.v_mid = 0.0


def newton_cotes(g: UGKSData, n: int, lo: float = -6.0, hi: float = 6.0):
    """Newton-Cotes velocities; n is rounded to floor(n/4)*4+1 per axis by init_quad."""
    g.quad_type = "Newton"
    g.u_num = g.v_num = n
    g.u_min = g.v_min = lo
    g.u_max = g.v_max = hi


def cavity(n_cells: int, vel: tuple, flux_method: str, Kn: float, lid: float = 0.15, wall_T: float = 1.0,
           nb: tuple = (1, 1), CFL: float = 0.8, Pr: float = 2.0 / 3.0, rho_perturb: float = 0.0,
           seed: int = SEED, n_cells_j: int | None = None) -> UGKSData:
    """Lid-driven cavity on the unit square: four diffuse walls at wall_T,
    the north wall moving at (lid, 0).  vel = ("GH", n) or ("NC", n)."""
    g = _base(flux_method, Kn, CFL, Pr)
    kind, n = vel
    if kind == "GH":
        gauss_hermite(g, n)
    else:
        newton_cotes(g, n)
    nj = n_cells if n_cells_j is None else n_cells_j
    bcs = [bcd.DiffuseBC(D=1.0, T=wall_T, U=lid, V=0.0), bcd.DiffuseBC(D=1.0, T=wall_T),
           bcd.DiffuseBC(D=1.0, T=wall_T), bcd.DiffuseBC(D=1.0, T=wall_T)]
    fill = FlowCondition(D=1.0, U=0.0, V=0.0, T=1.0)
    init = None
    if rho_perturb:
        rng = np.random.default_rng(seed)
        init = np.zeros((n_cells, nj, 4))
        init[..., 0] = 1.0 * (1.0 + rho_perturb * rng.standard_normal((n_cells, nj)))
        init[..., 3] = 1.0
    g.blocks = cartesian_superblock(0.0, 1.0, 0.0, 1.0, n_cells, nj, bcs, fill, nbi=nb[0], nbj=nb[1],
                                    init_macro=init)
    return g


def curvilinear(g: UGKSData, amp: float = 0.06) -> UGKSData:
    """Bend the grid lines of every block of a unit-square case: vertex
    (x, y) moves by amp * sin(pi x) sin(pi y) * (sin(2 pi y), -sin(2 pi x)).
    The outer boundary stays where it was (straight walls), the interior cells
    become skewed, non-uniform quadrilaterals; block joins stay conforming."""
    for b in g.blocks:
        x, y = np.array(b.x, dtype=np.float64), np.array(b.y, dtype=np.float64)
        bump = amp * np.sin(np.pi * x) * np.sin(np.pi * y)
        b.x = x + bump * np.sin(2 * np.pi * y)
        b.y = y - bump * np.sin(2 * np.pi * x)
    return g


def rotated_pair_cavity(n: int, vel: tuple, flux_method: str, Kn: float, lid: float = 0.15,
                        wall_T: float = 1.0) -> UGKSData:
    """Lid-driven cavity on [0, 2] x [0, 1] as two n x n blocks joined by a
    rotated connection: block 0 covers [0, 1] with i along +x and j along +y;
    block 1 covers [1, 2] with i along +y and j along -x, so its NORTH face
    is the interface (joined to block 0's EAST face) and its EAST face is
    the moving lid.  Block 1 cell (i, j) is the cell (2n - 1 - j, i) of the
    same cavity built as one 2n x n block."""
    g = _base(flux_method, Kn)
    kind, nv = vel
    (gauss_hermite if kind == "GH" else newton_cotes)(g, nv)
    h = 1.0 / n
    k = np.arange(n + 1, dtype=np.float64) * h
    fill = FlowCondition(D=1.0, U=0.0, V=0.0, T=1.0)
    wall = lambda U=0.0: bcd.DiffuseBC(D=1.0, T=wall_T, U=U, V=0.0)
    X0, Y0 = np.meshgrid(k, k, indexing="ij")
    b0 = BlockDef(x=X0, y=Y0, bc_list=[wall(lid), bcd.AdjacentBC(1, bcd.NORTH), wall(), wall()],
                  fill_condition=copy.deepcopy(fill), label="rot-0")
    I, J = np.meshgrid(k, k, indexing="ij")
    b1 = BlockDef(x=2.0 - J, y=I.copy(), bc_list=[bcd.AdjacentBC(0, bcd.EAST), wall(lid), wall(), wall()],
                  fill_condition=copy.deepcopy(fill), label="rot-1")
    g.blocks = [b0, b1]
    return g


def cfg1_cavity_gh(n_cells: int = 64, gh: int = 24) -> UGKSData:
    """64x64 single block, 24x24 Gauss-Hermite, vanLeer, Kn 0.075, lid 0.15."""
    return cavity(n_cells, ("GH", gh), "vanLeer", 0.075)


def cfg2_cavity_2x2(n_per_block: int = 128, nc: int = 32) -> UGKSData:
    """2x2 blocks of 128x128, Newton-Cotes 32 (-> 33) on [-6,6]^2, WENO5, Kn 0.01."""
    return cavity(2 * n_per_block, ("NC", nc), "WENO5", 0.01, nb=(2, 2))


def cfg3_cavity_512(Kn: float = 0.075, n_cells: int = 512, nc: int = 48) -> UGKSData:
    """512x512, Newton-Cotes 48 (-> 49), WENO5, rho perturbed by 1% (seeded)."""
    return cavity(n_cells, ("NC", nc), "WENO5", Kn, rho_perturb=0.01)


def cfg4_batch(ncases: int = 64, seed: int = SEED, sides=(32, 64, 96, 128, 192, 256), vgrids=(16, 24, 32, 40),
               flux_method: str = "WENO5"):
    """64 independent cavities of uneven size (seeded): side, velocity grid,
    Kn log-uniform in [1e-3, 1], lid in [0.05, 0.3], wall T in [0.8, 1.2].
    Velocity grids are Newton-Cotes on [-6,6]^2 (16 -> 17 per axis, etc.)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(ncases):
        side = int(rng.choice(sides))
        nv = int(rng.choice(vgrids))
        Kn = float(10.0 ** rng.uniform(-3.0, 0.0))
        lid = float(rng.uniform(0.05, 0.3))
        wT = float(rng.uniform(0.8, 1.2))
        out.append(cavity(side, ("NC", nv), flux_method, Kn, lid=lid, wall_T=wT))
    return out


def channel_widths(total: int = 1024, nblocks: int = 8, choices=(64, 96, 128, 160, 192), seed: int = SEED):
    """Block widths drawn from ``choices`` summing to ``total`` (seeded rejection sampling)."""
    rng = np.random.default_rng(seed)
    for _ in range(100000):
        w = [int(x) for x in rng.choice(choices, size=nblocks)]
        if sum(w) == total:
            return w
    raise RuntimeError("no width combination found")


# Adsorbing-wall mapping: sticking coefficient s -> gamma_f = S_T k_f = s with
# S_T = 1 (so alpha_p = C_ref D_ref t_ref / S_T = 1) and beta_n = beta_t = 0
# (stickingProbability = 1, clAdsorb.c:6-16).  A desorption rate r maps to the
# constant equilibrium coverage theta_eq with gamma_f (1/theta_eq - 1) = r
# (the desorption law gamma_f (1/theta_eq - 1) * impinging * theta,
# clAdsorb.c:436-438).
def adsorbing_params(sticking: float = 0.3, desorption: float = 0.01):
    theta_eq = sticking / (sticking + desorption)
    return dict(k_f=sticking, S_T=1.0, adsorb=theta_eq, beta_n=0.0, beta_t=0.0, cover_initial=0.0,
                reflect_type="D", dynamic=1)


def cfg5_channel(widths=None, height: int = 256, nc: int = 40, Kn: float = 0.05, flux_method: str = "WENO5",
                 U_in: float = 0.5, length: float = 4.0, reflect_type: str = "D", alpha_n: float = 1.0,
                 alpha_t: float = 1.0) -> UGKSData:
    """Adsorbing-wall channel: blocks of uneven width (total 1024) x 256 cells
    on [0, 4] x [0, 1]; inflow (rho 1, T 1) west, extrapolation east,
    adsorbing south wall (sticking 0.3, desorption 0.01, coverage 0; diffuse
    re-emission, or specular / Cercignani-Lampis with accommodation
    alpha_n, alpha_t), diffuse north wall at T = 1; initial rho 1, U (0.5, 0), T 1."""
    widths = channel_widths() if widths is None else list(widths)
    g = _base(flux_method, Kn)
    newton_cotes(g, nc)
    ap = adsorbing_params()
    ap.update(reflect_type=reflect_type, alpha_n=alpha_n, alpha_t=alpha_t)
    bcs = [bcd.DiffuseBC(D=1.0, T=1.0), bcd.ExtrapolateOutBC(), bcd.AdsorbingBC(D=1.0, T=1.0, **ap),
           bcd.InflowBC(D=1.0, T=1.0)]
    fill = FlowCondition(D=1.0, U=U_in, V=0.0, T=1.0)
    total = sum(widths)
    g.blocks = cartesian_superblock(0.0, length, 0.0, 1.0, total, height, bcs, fill, nbj=1, i_splits=widths)
    return g
