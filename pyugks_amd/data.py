"""Run configuration, velocity quadrature and the per-block constant set.

Counterpart of ugks_data.py (gdata defaults :146-235, init_quad :237-346,
init_ref_values :357-371, newton_coeff :404-417, non_dimensionalise_all
:462-542) and of the constant-emission half of ugks_CL.py:genHeader
(:20-400).  The reference bakes every constant into OpenCL source text, so
a value reaches its kernels after a round trip through a decimal string:

* ``'{}'.format(x)`` / ``str(x)`` under Python 2 keeps 12 significant
  digits (``'%.12g'``, plus ``'.0'`` for integral values): Pr, umax, vmax,
  gam, B, K, OMEGA (ugks_CL.py:86-93) and the WALL_* values (:277-310);
* ``'%0.15e'`` keeps 16 significant digits: Kn_eff, CFL, PI, the adsorbing
  constants, QUAD and WEIGHT (:84-96, :102-156, :318-346);
* ``'%g'`` keeps 6: the constant equilibrium coverage ADSORB_* (:167).

``bake_*`` below reproduce those round trips so the GPU kernels see the
same numbers the reference's kernels see.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import bc as bcd


# ---------------------------------------------------------------- baking ---
def py2_str(x: float) -> str:
    """Python 2.7 ``str(float)``: '%.12g' with '.0' added to integral output."""
    s = "%.12g" % float(x)
    if not any(c in s for c in ".en"):   # 'inf'/'nan' contain 'n'
        s += ".0"
    return s


def bake_str(x: float) -> float:
    return float(py2_str(x))


def bake_e15(x: float) -> float:
    return float("%0.15e" % float(x))


def bake_g(x: float) -> float:
    return float("%g" % float(x))


def isotherm_tables(table) -> tuple:
    """ISO / TRI / NBR of a tabulated Langmuir isotherm (ugks_CL.py:160-248):
    the (P, T, theta) rows baked '%0.15e', and scipy's Delaunay triangulation
    of the unbaked (P, T) -- vertex triples and neighbour triples, neighbour
    k opposite vertex k and -1 on the hull."""
    from scipy.spatial import Delaunay
    tab = np.asarray(table, dtype=np.float64)
    if tab.ndim != 2 or tab.shape[1] != 3:
        raise RuntimeError(f"isotherm table must have three columns (P, T, theta); got shape {tab.shape}")
    d = Delaunay(tab[:, 0:2])
    pts = np.ascontiguousarray([[bake_e15(x) for x in row] for row in tab], dtype=np.float64)
    return (pts, np.ascontiguousarray(d.simplices, dtype=np.int32),
            np.ascontiguousarray(d.neighbors, dtype=np.int32))


def newton_coeff(i: int, n: int) -> float:
    """Composite Boole weights, 1-based index.  ugks_data.py:404-417
    (the end-point weight is 14/15, as the reference has it)."""
    if (i == 1) | (i == n):
        return 14.0 / 15.0
    elif (i - 5) % 4 == 0:
        return 28.0 / 45.0
    elif (i - 3) % 4 == 0:
        return 24.0 / 45.0
    else:
        return 64.0 / 45.0


# ------------------------------------------------------------ conditions ---
@dataclass
class FlowCondition:
    """geom/geom_flow.py:12-66 (the fields the initial state reads)."""
    D: float = 1.0
    U: float = 0.0
    V: float = 0.0
    T: float = 300.0
    qx: float = 0.0
    qy: float = 0.0


@dataclass
class BlockDef:
    """One structured block: vertex coordinates and four BCs (N, E, S, W).

    ``x``/``y`` are (ni+1, nj+1) vertex arrays.  ``init_macro`` optionally
    replaces the uniform fill with per-cell (rho, U, V, T) arrays of shape
    (ni, nj), the counterpart of a UDF fill (ugks_block.py:304-331).
    """
    x: np.ndarray
    y: np.ndarray
    bc_list: list
    fill_condition: FlowCondition
    label: str = ""
    init_macro: np.ndarray | None = None

    @property
    def ni(self) -> int:
        return self.x.shape[0] - 1

    @property
    def nj(self) -> int:
        return self.x.shape[1] - 1


@dataclass
class ResidualOptions:
    """ugks_data.py:74-102: residual monitoring and the stop criteria built on it."""
    global_residual: object = 1e50
    residual_history: list = field(default_factory=list)
    get_residual: bool = False
    residual_count: int = -1
    min_residual: float = 1e-50
    min_slope: float = 1e-3
    slope_start: int = 10
    residual_start: int = 10
    res_drop: float = float("inf")
    non_linear_output: bool = False
    non_linear_output_limit: float = 1e5
    non_linear_save: bool = False
    non_linear_dt: bool = False


@dataclass
class SaveOptions:
    """ugks_data.py:34-68: when and what the step loop writes (HDF5 only; the
    reference's VTK writer and picture output are not rebuilt).  save_name
    names the run's files, h5Name overrides the full path of the global file;
    compression is "gzip" (deflate) or None."""
    save: bool = False
    save_name: str = ""
    h5Name: str = ""
    save_initial_f: bool = False
    save_final_f: bool = False
    save_f_always: bool = False
    save_final_flux: bool = False
    save_final_wall_distribution: bool = False
    internal_data: bool = False      # translational temperatures and stress (getInternal) in each save
    compression: object = "gzip"
    save_count: int = 1
    initial_save_count: object = None
    initial_save_cutoff_time: float = 0.0


@dataclass
class UGKSData:
    """The gdata slots the time step reads, with the reference defaults
    (ugks_data.py:146-235)."""
    flux_method: str = "WENO5"
    CFL: float = 0.3
    dt: float = 0.0
    time: float = 0.0
    step: int = 0
    delta_dt: float = 1.0
    suggest_dt: float = 0.0
    max_time: float = 1.0e-3
    max_step: int = 1
    dt_update_count: int = 1
    # step-loop controls (ugks_data.py:158-178)
    print_count: int = 20
    clock_time_stop: tuple = (1e6, 0)
    exit: bool = False
    runtime_conf: str = ""
    residual_options: ResidualOptions = field(default_factory=ResidualOptions)
    save_options: SaveOptions = field(default_factory=SaveOptions)
    rootName: str = "ugks_run"      # directory the run's HDF5 output goes under (ugks_data.py:180)
    restart: str = ""               # HDF5 global file of the run to restart from (ugks_data.py:183)
    L_ref: float = 1.0
    D_ref: float = 1.0
    T_ref: float = 273.0
    omega: float = 0.81
    R: float = 287.0
    gamma: float = 5.0 / 3.0
    Pr: float = 0.72
    Kn_eff: float = -1.0
    relax_type: int = 2
    quad_type: str | None = None
    quad: object = None
    weight: object = None
    u_num: int = 0
    v_num: int = 0
    u_min: float = 0.0
    v_min: float = 0.0
    u_mid: float = 0.0
    v_mid: float = 0.0
    u_max: float = 0.0
    v_max: float = 0.0
    blocks: list = field(default_factory=list)
    # derived by init_ref_values / init_quad
    Nv: int = 0
    b: float = 0.0
    K: float = 0.0
    C_ref: float = 0.0
    t_ref: float = 0.0
    P_ref: float = 0.0
    umax: float = 0.0
    vmax: float = 0.0
    mirror_NS: np.ndarray | None = None
    mirror_EW: np.ndarray | None = None
    mirror_D: np.ndarray | None = None

    # ugks_data.py:357-371
    def init_ref_values(self):
        self.C_ref = math.sqrt(2 * self.R * self.T_ref)
        self.t_ref = self.L_ref / self.C_ref
        self.P_ref = self.D_ref * self.R * self.T_ref
        self.b = 2.0 / (self.gamma - 1.0)
        self.K = self.b - 2
        self.init_quad()

    # ugks_data.py:237-346
    def init_quad(self):
        if self.quad_type == "Gauss":
            lq = np.asarray(self.quad, dtype=np.float64).ravel()
            lw = np.asarray(self.weight, dtype=np.float64).ravel()
            n = lq.size
            self.Nv = n * n
            quad = np.zeros((self.Nv, 2))
            weight = np.zeros(self.Nv)
            index_array = np.zeros((n, n), dtype=np.int64)
            count = 0
            for i in range(n):
                for j in range(n):
                    index_array[i, j] = count
                    u = lq[i] + self.u_mid
                    v = lq[j] + self.v_mid
                    quad[count] = (u, v)
                    weight[count] = lw[i] * np.exp(u ** 2) * lw[j] * np.exp(v ** 2)
                    count += 1
        elif self.quad_type == "Newton":
            self.u_num = int(self.u_num / 4) * 4 + 1
            self.v_num = int(self.v_num / 4) * 4 + 1
            self.Nv = self.u_num * self.v_num
            quad = np.zeros((self.Nv, 2))
            weight = np.zeros(self.Nv)
            du = (self.u_max - self.u_min) / float(self.u_num - 1)
            dv = (self.v_max - self.v_min) / float(self.v_num - 1)
            index_array = np.zeros((self.u_num, self.v_num), dtype=np.int64)
            count = 0
            for i in range(self.u_num):
                for j in range(self.v_num):
                    index_array[i, j] = count
                    u = self.u_min + i * du
                    v = self.v_min + j * dv
                    quad[count] = (u, v)
                    weight[count] = (newton_coeff(i + 1, self.u_num) * du) * (newton_coeff(j + 1, self.v_num) * dv)
                    count += 1
        elif self.quad_type == "defined":
            quad = np.asarray(self.quad, dtype=np.float64).reshape(-1, 2)
            weight = np.asarray(self.weight, dtype=np.float64).ravel()
            self.Nv = quad.shape[0]
            index_array = None
        else:
            raise ValueError(f"unknown quad_type {self.quad_type!r}")
        self.quad, self.weight = quad, weight
        if index_array is not None:
            self.mirror_NS = np.ravel(np.fliplr(index_array)).astype(np.int64)
            self.mirror_EW = np.ravel(np.flipud(index_array)).astype(np.int64)
            self.mirror_D = np.ravel(np.flipud(np.fliplr(index_array))).astype(np.int64)
        else:
            self.mirror_NS = self.mirror_EW = self.mirror_D = None
        self.umax = abs(np.max(self.quad[:, 0]))
        self.vmax = abs(np.max(self.quad[:, 1]))

    # ugks_data.py:462-542 (grid, BCs, fills, global times)
    def non_dimensionalise_all(self):
        for b in self.blocks:
            fc = b.fill_condition
            fc.D /= self.D_ref
            fc.U /= self.C_ref
            fc.V /= self.C_ref
            fc.T /= self.T_ref
            fc.qx /= (self.D_ref * self.C_ref ** 3)
            fc.qy /= (self.D_ref * self.C_ref ** 3)
            if b.init_macro is not None:
                m = np.array(b.init_macro, dtype=np.float64, copy=True)
                m[..., 0] /= self.D_ref
                m[..., 1:3] /= self.C_ref
                m[..., 3] /= self.T_ref
                b.init_macro = m
            for bc in b.bc_list:
                bc.D /= self.D_ref
                bc.T /= self.T_ref
                bc.U /= self.C_ref
                bc.V /= self.C_ref
                bc.P /= self.P_ref
                if bc.type_of_BC == bcd.ADSORBING:
                    bc.alpha_p = (self.C_ref * self.D_ref * self.t_ref) / bc.S_T
                    bc.gamma_f = bc.S_T * bc.k_f
                    if np.ndim(bc.adsorb) == 0:
                        bc.adsorb = bc.adsorb / bc.S_T
                    else:                      # (pressure, temperature, coverage), ugks_data.py:522-525
                        tab = np.array(bc.adsorb, dtype=np.float64, copy=True)
                        tab[:, 0] /= self.D_ref * self.C_ref ** 2
                        tab[:, 1] /= self.T_ref
                        tab[:, 2] /= bc.S_T
                        bc.adsorb = tab
            b.x = np.asarray(b.x, dtype=np.float64) / self.L_ref
            b.y = np.asarray(b.y, dtype=np.float64) / self.L_ref
        self.dt /= self.t_ref
        self.suggest_dt /= self.t_ref
        self.max_time /= self.t_ref
        self.save_options.initial_save_cutoff_time /= self.t_ref

    def prepare(self):
        """global_preparation minus the job-script exec (ugks_data.py:586-668)."""
        self.init_ref_values()
        self.non_dimensionalise_all()
        return self

    @property
    def ghost(self) -> int:
        """ugks_block.py:88-91"""
        return 3 if self.flux_method == "WENO5" else 2

    @property
    def flux_method_code(self) -> int:
        """ugks_CL.py:33-44"""
        return {"vanLeer": 0, "WENO5": 1, "minmod": 3}[self.flux_method]

    def velocity_tables(self):
        return velocity_tables(self)

    def block_constants(self, b: BlockDef) -> dict:
        return block_constants(self, b)


# -------------------------------------------------------------- geometry ---
def ghost_coordinates(blocks: list, ghost: int) -> list:
    """Vertex arrays with ghost layers, shape (Ni+1, Nj+1, 2), per block.

    Interior vertices come from the block grid (ugks_block.py:212-219).  A
    face joined to another block copies that block's vertices (the plain
    copy branch of xyExchange, clGhostRoutines.c:155-296, valid for the
    axis-aligned E-W / N-S joins built here); every other face extends the
    last cell spacing outward, restating xyExtrapolate
    (clGhostRoutines.c:302-398).  Faces are processed N, E, S, W per block,
    blocks in order (ugks_block.py:718-744)."""
    G = ghost
    out = []
    for b in blocks:
        ni, nj = b.ni, b.nj
        NI, NJ = ni + 2 * G, nj + 2 * G
        xy = np.full((NI + 1, NJ + 1, 2), np.nan)
        xy[G:G + ni + 1, G:G + nj + 1, 0] = b.x
        xy[G:G + ni + 1, G:G + nj + 1, 1] = b.y
        out.append(xy)
    for bi, b in enumerate(blocks):
        xy = out[bi]
        ni, nj = b.ni, b.nj
        NI, NJ = ni + 2 * G, nj + 2 * G
        for face, bcx in enumerate(b.bc_list):
            if bcx.type_of_BC != bcd.ADJACENT:
                _xy_extrapolate(xy, face, ni, nj, G)
                continue
            o = blocks[bcx.other_block]
            oxy = out[bcx.other_block]
            ONI, ONJ = o.ni + 2 * G, o.nj + 2 * G
            # translation taking the other face's first vertex onto this face's
            # (zero for a true join, the period for a periodic one)
            if face == bcd.EAST and bcx.other_face == bcd.WEST:
                off = xy[NI - G, G] - oxy[G, G]
                src = oxy[G + 1:2 * G + 1, G:ONJ - G + 1]
                xy[NI - G + 1:NI + 1, G:NJ - G + 1] = src + off if np.any(off) else src
            elif face == bcd.WEST and bcx.other_face == bcd.EAST:
                off = xy[G, G] - oxy[ONI - G, G]
                src = oxy[ONI - 2 * G:ONI - G, G:ONJ - G + 1]
                xy[0:G, G:NJ - G + 1] = src + off if np.any(off) else src
            elif face == bcd.NORTH and bcx.other_face == bcd.SOUTH:
                off = xy[G, NJ - G] - oxy[G, G]
                src = oxy[G:ONI - G + 1, G + 1:2 * G + 1]
                xy[G:NI - G + 1, NJ - G + 1:NJ + 1] = src + off if np.any(off) else src
            elif face == bcd.SOUTH and bcx.other_face == bcd.NORTH:
                off = xy[G, G] - oxy[G, ONJ - G]
                src = oxy[G:ONI - G + 1, ONJ - 2 * G:ONJ - G]
                xy[G:NI - G + 1, 0:G] = src + off if np.any(off) else src
            else:
                _xy_exchange_rotated(xy, face, oxy, bcx.other_face, G, int(bcx.orientation))
    return out


def _swap_index(this_face, i, j, that_face, NI, NJ, NIB, NJB, G, vertex):
    """getSwapIndex (clGhostRoutines.c:47-138): the index in the other block of
    ghost index (i, j) of this one (NI x NJ cells), for vertices or cells."""
    af, ar = (0, 1) if vertex else (-1, 0)

    def rot90(i, j, n, Ni, Nj):            # n quarter turns counter-clockwise (:5-23)
        for _ in range(n):
            i, j, Ni, Nj = Nj - j - 1, i, Nj, Ni
        return i, j
    N, E, S, W = bcd.NORTH, bcd.EAST, bcd.SOUTH, bcd.WEST
    if this_face == E:
        if that_face == E: return NI + af - i + NIB - 2 * G, j
        if that_face == W: return i - (NI - 2 * G), j
        if that_face == N:
            i, j = rot90(i, j, 3, NI + ar, NJ + ar); return i, j + NJB - 2 * G
        i, j = rot90(i, j, 1, NI + ar, NJ + ar); return i, j - (NI - 2 * G)
    if this_face == W:
        if that_face == W: return NI + af - i - (NI - 2 * G), j
        if that_face == E: return i + NIB - 2 * G, j
        if that_face == N:
            i, j = rot90(i, j, 1, NI + ar, NJ + ar); return i, j + NJB - 2 * G
        i, j = rot90(i, j, 3, NI + ar, NJ + ar); return i, j - (NI - 2 * G)
    if this_face == N:
        if that_face == N: return i, NJ + af - j + NJB - 2 * G
        if that_face == S: return i, j - (NJ - 2 * G)
        if that_face == E:
            i, j = rot90(i, j, 1, NI + ar, NJ + ar); return i + NIB - 2 * G, j
        i, j = rot90(i, j, 3, NI + ar, NJ + ar); return i - (NJ - 2 * G), j
    if that_face == S: return i, NJ + af - j - (NJ - 2 * G)
    if that_face == N: return i, j + NJB - 2 * G
    if that_face == E:
        i, j = rot90(i, j, 3, NI + ar, NJ + ar); return i + NIB - 2 * G, j
    i, j = rot90(i, j, 1, NI + ar, NJ + ar); return i - (NJ - 2 * G), j


def _face_ends(face, NI, NJ, G):
    """First and second vertex of a block face (clGhostRoutines.c:171-236)."""
    return {bcd.NORTH: ((G, NJ - G), (NI - G, NJ - G)), bcd.EAST: ((NI - G, G), (NI - G, NJ - G)),
            bcd.SOUTH: ((G, G), (NI - G, G)), bcd.WEST: ((G, G), (G, NJ - G))}[face]


def _xy_exchange_rotated(xy, face, oxy, other_face, G, ori):
    """xyExchange (clGhostRoutines.c:155-296) for joins other than E-W / N-S:
    every ghost vertex beyond this face takes the other block's vertex at the
    vertex form of getSwapIndex, after the row flip of a misaligned, oriented
    join.  Faces that coincide (end points equal in either order: a
    conforming join) copy the vertices, the reference's NO_TRANSFORM; the
    reference's default TRANSFORM maps the other face's first end point onto
    this face's first, which for a conforming join with its end points in
    opposite order would mirror the ghost vertices back across the join.
    Faces apart (periodic joins) take that similarity (rotation, scale,
    shift) as the reference does."""
    NI, NJ = xy.shape[0] - 1, xy.shape[1] - 1
    ONI, ONJ = oxy.shape[0] - 1, oxy.shape[1] - 1
    ni, nj = NI - 2 * G, NJ - 2 * G
    (a1i, a1j), (a2i, a2j) = _face_ends(face, NI, NJ, G)
    (b1i, b1j), (b2i, b2j) = _face_ends(other_face, ONI, ONJ, G)
    A1, A2, B1, B2 = (complex(*xy[a1i, a1j]), complex(*xy[a2i, a2j]), complex(*oxy[b1i, b1j]),
                      complex(*oxy[b2i, b2j]))
    misalign = abs(A1.real - B1.real) > 1e-12 or abs(A1.imag - B1.imag) > 1e-12 or \
        abs(A2.real - B2.real) > 1e-12 or abs(A2.imag - B2.imag) > 1e-12
    coincide = not misalign or (abs(A1 - B2) <= 1e-12 and abs(A2 - B1) <= 1e-12)
    scale = (A2 - A1) / (B2 - B1)
    if face in (bcd.NORTH, bcd.SOUTH):
        gis, gjs = range(G, G + ni), (range(NJ - G + 1, NJ) if face == bcd.NORTH else range(0, G - 1))
    else:
        gis, gjs = (range(NI - G + 1, NI) if face == bcd.EAST else range(0, G - 1)), range(G, G + nj)
    for gi in gis:
        for gj in gjs:
            for ii in (0, 1):
                for jj in (0, 1):
                    iB, jB = gi + ii, gj + jj
                    if misalign and ori == 1:
                        if face in (bcd.EAST, bcd.WEST):
                            jB = NJ - jB
                        else:
                            iB = NI - iB
                    iB, jB = _swap_index(face, iB, jB, other_face, NI, NJ, ONI, ONJ, G, True)
                    p = oxy[iB, jB]
                    if coincide:
                        xy[gi + ii, gj + jj] = p
                    else:
                        z = A1 + (complex(*p) - B1) * scale
                        xy[gi + ii, gj + jj] = (z.real, z.imag)


def _xy_extrapolate(xy, face, ni, nj, G):
    """xyExtrapolate, clGhostRoutines.c:302-398, one launch (sequential)."""
    NI, NJ = ni + 2 * G, nj + 2 * G
    X = xy
    if face == bcd.NORTH:
        for gi0 in range(ni):
            for gj0 in range(G):
                gi, gj = gi0 + G, gj0 + NJ - G
                ei, ej = gi, NJ - G - 1
                m = float(gj - ej)
                d1 = m * (X[ei, ej + 1] - X[ei, ej])
                d2 = m * (X[ei + 1, ej + 1] - X[ei + 1, ej])
                X[gi, gj] = X[ei, ej] + d1
                X[gi + 1, gj] = X[ei + 1, ej] + d2
                X[gi + 1, gj + 1] = X[ei + 1, ej + 1] + d2
                X[gi, gj + 1] = X[ei, ej + 1] + d1
    elif face == bcd.EAST:
        for gi0 in range(G):
            for gj0 in range(nj):
                gi, gj = gi0 + NI - G, gj0 + G
                ei, ej = NI - G - 1, gj
                m = float(gi - ei)
                d1 = m * (X[ei + 1, ej] - X[ei, ej])
                d2 = m * (X[ei + 1, ej + 1] - X[ei, ej + 1])
                X[gi, gj] = X[ei, ej] + d1
                X[gi + 1, gj] = X[ei + 1, ej] + d1
                X[gi + 1, gj + 1] = X[ei + 1, ej + 1] + d2
                X[gi, gj + 1] = X[ei, ej + 1] + d2
    elif face == bcd.SOUTH:
        for gi0 in range(ni):
            for gj0 in range(G):
                gi, gj = gi0 + G, gj0
                ei, ej = gi, G
                m = float(ej - gj)
                d1 = m * (X[ei, ej] - X[ei, ej + 1])
                d2 = m * (X[ei + 1, ej] - X[ei + 1, ej + 1])
                X[gi, gj] = X[ei, ej] + d1
                X[gi + 1, gj] = X[ei + 1, ej] + d2
                X[gi + 1, gj + 1] = X[ei + 1, ej + 1] + d2
                X[gi, gj + 1] = X[ei, ej + 1] + d1
    else:
        for gi0 in range(G):
            for gj0 in range(nj):
                gi, gj = gi0, gj0 + G
                ei, ej = G, gj
                m = float(ei - gi)
                d1 = m * (X[ei, ej] - X[ei + 1, ej])
                d2 = m * (X[ei, ej + 1] - X[ei + 1, ej + 1])
                X[gi, gj] = X[ei, ej] + d1
                X[gi + 1, gj] = X[ei + 1, ej] + d1
                X[gi + 1, gj + 1] = X[ei + 1, ej + 1] + d2
                X[gi, gj + 1] = X[ei, ej + 1] + d2


def subdivide_vertex_range(ncells: int, nblocks: int):
    """geom/geom_block.py:701-719"""
    out = []
    vmin = 0
    for iblk in range(nblocks):
        n = int(ncells / (nblocks - iblk))
        out.append((vmin, vmin + n))
        vmin += n
        ncells -= n
    return out


def cartesian_superblock(x0, x1, y0, y1, nni, nnj, bc_list, fill_condition,
                         nbi=1, nbj=1, i_splits=None, init_macro=None, label="sblk"):
    """Uniform Cartesian grid on [x0,x1]x[y0,y1] split into nbi x nbj blocks,
    after SuperBlock2D (geom/geom_block.py:721-792): blocks are created with
    i outer, j inner; outer faces take ``bc_list``, inner faces are joined
    E-W and N-S.  ``i_splits`` optionally gives explicit block widths.
    Returns the list of BlockDef in creation order.

    The reference builds the grid with pygeom's transfinite interpolation;
    here vertices are x0 + (x1-x0)*i/nni exactly."""
    xs = x0 + (x1 - x0) * (np.arange(nni + 1, dtype=np.float64) / nni)
    ys = y0 + (y1 - y0) * (np.arange(nnj + 1, dtype=np.float64) / nnj)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    if i_splits is not None:
        assert sum(i_splits) == nni
        si, acc = [], 0
        for w in i_splits:
            si.append((acc, acc + w)); acc += w
        nbi = len(i_splits)
    else:
        si = subdivide_vertex_range(nni, nbi)
    sj = subdivide_vertex_range(nnj, nbj)
    blks = []
    grid = {}
    for i in range(nbi):
        for j in range(nbj):
            imin, imax = si[i]
            jmin, jmax = sj[j]
            new_bcs = [bcd.ReflectBC() for _ in range(4)]
            import copy as _copy
            if i == 0:
                new_bcs[bcd.WEST] = _copy.deepcopy(bc_list[bcd.WEST])
            if i == nbi - 1:
                new_bcs[bcd.EAST] = _copy.deepcopy(bc_list[bcd.EAST])
            if j == 0:
                new_bcs[bcd.SOUTH] = _copy.deepcopy(bc_list[bcd.SOUTH])
            if j == nbj - 1:
                new_bcs[bcd.NORTH] = _copy.deepcopy(bc_list[bcd.NORTH])
            im = None
            if init_macro is not None:
                im = np.array(init_macro[imin:imax, jmin:jmax], copy=True)
            b = BlockDef(x=X[imin:imax + 1, jmin:jmax + 1].copy(), y=Y[imin:imax + 1, jmin:jmax + 1].copy(),
                         bc_list=new_bcs, fill_condition=_copy.deepcopy(fill_condition),
                         label=f"{label}-{i}-{j}", init_macro=im)
            grid[(i, j)] = len(blks)
            blks.append(b)
    # connect_blocks_2D, geom_block.py:410-442 and :783-790
    for i in range(1, nbi):
        for j in range(nbj):
            a, c = grid[(i - 1, j)], grid[(i, j)]
            blks[a].bc_list[bcd.EAST] = bcd.AdjacentBC(c, bcd.WEST)
            blks[c].bc_list[bcd.WEST] = bcd.AdjacentBC(a, bcd.EAST)
    for i in range(nbi):
        for j in range(1, nbj):
            a, c = grid[(i, j - 1)], grid[(i, j)]
            blks[a].bc_list[bcd.NORTH] = bcd.AdjacentBC(c, bcd.SOUTH)
            blks[c].bc_list[bcd.SOUTH] = bcd.AdjacentBC(a, bcd.NORTH)
    return blks


def initial_macro(b: BlockDef):
    """macro_H / Q_H before initFunctions (ugks_block.py:270-331): uniform
    fill, or per-cell values; T is stored as 1/T."""
    m = np.ones((b.ni, b.nj, 4), dtype=np.float64)
    q = np.ones((b.ni, b.nj, 2), dtype=np.float64)
    fc = b.fill_condition
    m[:, :, 0] *= fc.D
    m[:, :, 1] *= fc.U
    m[:, :, 2] *= fc.V
    m[:, :, 3] *= 1.0 / fc.T
    q[:, :, 0] *= fc.qx
    q[:, :, 1] *= fc.qy
    if b.init_macro is not None:
        im = np.asarray(b.init_macro, dtype=np.float64)
        m[:, :, 0] = im[:, :, 0]
        m[:, :, 1] = im[:, :, 1]
        m[:, :, 2] = im[:, :, 2]
        m[:, :, 3] = 1.0 / im[:, :, 3]
    return m, q


def ghost_of(gdata) -> int:
    """ugks_block.py:88-91"""
    return 3 if gdata.flux_method == "WENO5" else 2


def flux_method_code(gdata) -> int:
    """ugks_CL.py:33-44"""
    return {"vanLeer": 0, "WENO5": 1, "minmod": 3}[gdata.flux_method]


def velocity_tables(gdata):
    """QUAD/WEIGHT as the kernels see them ('%0.15e', ugks_CL.py:318-346) and the
    mirror tables.  ``gdata`` is this package's UGKSData or the reference's
    (the same attribute names: quad, weight, Nv, mirror_*)."""
    q = np.array([[bake_e15(a), bake_e15(b)] for a, b in gdata.quad], dtype=np.float64)
    w = np.array([bake_e15(a) for a in gdata.weight], dtype=np.float64)
    nv = gdata.Nv
    zeros = np.zeros(nv, dtype=np.int64)
    mns = gdata.mirror_NS if gdata.mirror_NS is not None else zeros
    mew = gdata.mirror_EW if gdata.mirror_EW is not None else zeros
    md = gdata.mirror_D if gdata.mirror_D is not None else zeros
    return (np.ascontiguousarray(q), np.ascontiguousarray(w), np.ascontiguousarray(mns, dtype=np.int64),
            np.ascontiguousarray(mew, dtype=np.int64), np.ascontiguousarray(md, dtype=np.int64))

def block_constants(gdata, b) -> dict:
    """The #define set genHeader emits for one block (ugks_CL.py:58-310).
    ``gdata`` and ``b`` are this package's UGKSData and BlockDef or the
    reference's gdata and UGKSBlock: only the attribute names they share are
    read (flux_method, Nv, relax_type, Pr, Kn_eff, gamma, b, K, omega, CFL,
    umax, vmax; ni, nj, bc_list and the boundary conditions' fields)."""
    G = ghost_of(gdata)
    c = dict(ni=b.ni, nj=b.nj, NI=b.ni + 2 * G, NJ=b.nj + 2 * G, NV=gdata.Nv, ghost=G,
             flux_method=flux_method_code(gdata), relax_type=int(gdata.relax_type),
             Pr=bake_str(gdata.Pr), Kn_eff=bake_e15(gdata.Kn_eff), gam=bake_str(gdata.gamma),
             B=bake_str(gdata.b), K=bake_str(gdata.K), omega=bake_str(gdata.omega),
             CFL=bake_e15(gdata.CFL), umax=bake_str(gdata.umax), vmax=bake_str(gdata.vmax))
    wD, wU, wV, wT, wP = [], [], [], [], []
    bn, bt, gf, ap, dyn, cad, ads = [], [], [], [], [], [], []
    an, at = [], []
    iso, iso_tri, iso_nbr = [None] * 4, [None] * 4, [None] * 4
    for bcx in b.bc_list:
        if bcx.type_of_BC in (bcd.ADSORBING, bcd.DIFFUSE, bcd.CONSTANT, bcd.INFLOW, bcd.OUTFLOW):
            wD.append(bake_str(bcx.D))
            wU.append(bake_str(bcx.U))
            wV.append(bake_str(bcx.V))
            with np.errstate(divide="ignore"):       # '1.0/(%s)', ugks_CL.py:298 (C: 1/0 = inf)
                wT.append(float(np.float64(1.0) / np.float64(bake_str(bcx.T))))
            wP.append(bake_str(bcx.P) if bcx.type_of_BC == bcd.OUTFLOW else -1.0)
        else:
            wD.append(-1.0); wU.append(-1.0); wV.append(-1.0); wT.append(-1.0); wP.append(-1.0)
        if bcx.type_of_BC == bcd.ADSORBING:
            bn.append(bake_e15(bcx.beta_n)); bt.append(bake_e15(bcx.beta_t))
            gf.append(bake_e15(bcx.gamma_f)); ap.append(bake_e15(bcx.alpha_p))
            an.append(bake_e15(bcx.alpha_n)); at.append(bake_e15(bcx.alpha_t))
            dyn.append(int(bcx.dynamic))
            if np.ndim(bcx.adsorb) != 0:
                k = len(cad)
                iso[k], iso_tri[k], iso_nbr[k] = isotherm_tables(bcx.adsorb)
                cad.append(0); ads.append(0.0)
            else:
                a = float(bcx.adsorb)
                if (a > 1.0) | (a <= 0.0):
                    raise RuntimeError(f"adsorbing wall: equilibrium coverage {a!r} (adsorb / S_T) must lie in (0, 1]")
                cad.append(1); ads.append(bake_g(a))
        else:
            bn.append(-1.0); bt.append(-1.0); gf.append(-1.0); ap.append(-1.0)
            an.append(-1.0); at.append(-1.0)
            dyn.append(-1); cad.append(0); ads.append(0.0)
    c.update(wall_D=wD, wall_U=wU, wall_V=wV, wall_T=wT, wall_P=wP, beta_n=bn, beta_t=bt,
             gamma_f=gf, alpha_p=ap, dynamic_vartheta=dyn, constant_adsorb=cad, adsorb=ads,
             alpha_n=an, alpha_t=at, iso=iso, iso_tri=iso_tri, iso_nbr=iso_nbr)
    # wall values given as expressions, on the walls that carry wall values (ugks_CL.py:278-298)
    from .udf import block_programs, is_time_dependent
    carry = [bcx if bcx.type_of_BC in (bcd.ADSORBING, bcd.DIFFUSE, bcd.CONSTANT, bcd.INFLOW, bcd.OUTFLOW)
             else None for bcx in b.bc_list]
    c["udf"] = block_programs(carry) if any(x is not None and x.has_udf() for x in carry) else None
    c["time_dependent_BC"] = any(is_time_dependent(bcx) for bcx in b.bc_list)
    return c
