"""Boundary-condition codes and containers.

Mirrors geom/geom_bc_defs.py (codes :27-51, BoundaryCondition fields and
defaults :82-170, the convenience subclasses :199-443) and geom/geom_defs.py
(face ids :13-16).  Only the fields the time step reads are kept.
"""
from __future__ import annotations

import copy

# face ids, geom/geom_defs.py:13-16 (GNORTH..GWEST in the kernels, ugks_CL.py:74-77)
NORTH, EAST, SOUTH, WEST = 0, 1, 2, 3
faceName = ["north", "east", "south", "west"]

# BC codes, geom/geom_bc_defs.py:27-37
ADJACENT = 0
COMMON = 0
EXTRAPOLATE_OUT = 1
CONSTANT = 2
REFLECT = 3
DIFFUSE = 4
PERIODIC = 5
INFLOW = 6
OUTFLOW = 7
ADSORBING = 8
BOUNCE_BACK = 9

# distribution flips, geom/geom_bc_defs.py:40-43
NO_FLIP, FLIP_NS, FLIP_D, FLIP_EW = 0, 1, 2, 3
# orientation flags, geom/geom_bc_defs.py:50-51
HOLD, NO_HOLD = 0, 1

REFLECT_TYPES = {"S": 0, "D": 1, "CL": 2}


class BoundaryCondition:
    """geom_bc_defs.py:82-170.  Same field names and defaults."""

    def __init__(self, type_of_BC=REFLECT, D=0.0, T=0.0, U=0.0, V=0.0, P=0.0,
                 adsorb=None, beta_n=0, beta_t=0, alpha_n=1, alpha_t=1, k_f=0,
                 S_T=1.0, cover_initial=0, reflect_type="S", dynamic=1,
                 other_block=-1, other_face=-1, orientation=NO_HOLD,
                 flip_distribution=NO_FLIP, label="", UDF_D=None, UDF_U=None, UDF_V=None, UDF_T=None):
        self.type_of_BC = type_of_BC
        self.D, self.T, self.U, self.V, self.P = D, T, U, V, P
        # wall values as OpenCL C expressions in s, x, y, t (geom_bc_defs.py:115-146)
        self.UDF_D, self.UDF_U, self.UDF_V, self.UDF_T = UDF_D, UDF_U, UDF_V, UDF_T
        self.adsorb = copy.copy(adsorb)
        self.beta_n, self.beta_t = beta_n, beta_t
        self.alpha_n, self.alpha_t = alpha_n, alpha_t
        self.k_f, self.S_T = k_f, S_T
        self.cover_initial = cover_initial
        self.reflect_type = reflect_type
        self.dynamic = dynamic
        self.gamma_f = None
        self.alpha_p = 0
        self.other_block, self.other_face = other_block, other_face
        self.orientation = orientation
        self.flip_distribution = flip_distribution
        self.label = label

    def has_udf(self) -> bool:
        return any(u is not None for u in (self.UDF_D, self.UDF_U, self.UDF_V, self.UDF_T))

    def __repr__(self):
        return f"BC(type={self.type_of_BC}, D={self.D}, U={self.U}, V={self.V}, T={self.T}, label={self.label!r})"


def AdjacentBC(other_block=-1, other_face=-1, orientation=NO_HOLD):
    """geom_bc_defs.py:199-219"""
    return BoundaryCondition(ADJACENT, other_block=other_block, other_face=other_face,
                             orientation=orientation, label="ADJACENT")


def PeriodicBC(other_block=-1, other_face=-1, orientation=NO_HOLD, flip_distribution=NO_FLIP):
    """geom_bc_defs.py:220-239 (type_of_BC is ADJACENT there too)"""
    return BoundaryCondition(ADJACENT, other_block=other_block, other_face=other_face,
                             orientation=orientation, flip_distribution=flip_distribution,
                             label="PERIODIC")


def ExtrapolateOutBC():
    """geom_bc_defs.py:241-255"""
    return BoundaryCondition(EXTRAPOLATE_OUT, label="EXTRAPOLATE OUT")


def ConstantBC(D=0, T=0, U=0, V=0, UDF_D=None, UDF_U=None, UDF_V=None, UDF_T=None):
    """geom_bc_defs.py:257-286"""
    return BoundaryCondition(CONSTANT, D=D, T=T, U=U, V=V, label="CONSTANT", UDF_D=UDF_D, UDF_U=UDF_U,
                             UDF_V=UDF_V, UDF_T=UDF_T)


def ReflectBC():
    """geom_bc_defs.py:288-300"""
    return BoundaryCondition(REFLECT, label="REFLECT")


def BounceBackBC():
    """geom_bc_defs.py:302-314"""
    return BoundaryCondition(BOUNCE_BACK, label="BOUNCE_BACK")


def DiffuseBC(D=1, T=0, U=0, V=0, UDF_D=None, UDF_U=None, UDF_V=None, UDF_T=None):
    """geom_bc_defs.py:316-353"""
    return BoundaryCondition(DIFFUSE, D=D, T=T, U=U, V=V, label="DIFFUSE", UDF_D=UDF_D, UDF_U=UDF_U, UDF_V=UDF_V,
                             UDF_T=UDF_T)


def AdsorbingBC(D=1, T=0, U=0, V=0, adsorb=None, beta_n=0, beta_t=0, alpha_n=1,
                alpha_t=1, k_f=0, S_T=1.0, cover_initial=0, reflect_type="S", dynamic=1,
                UDF_D=None, UDF_U=None, UDF_V=None, UDF_T=None):
    """geom_bc_defs.py:355-414"""
    return BoundaryCondition(ADSORBING, D=D, T=T, U=U, V=V, adsorb=adsorb, beta_n=beta_n,
                             beta_t=beta_t, alpha_n=alpha_n, alpha_t=alpha_t, k_f=k_f,
                             S_T=S_T, cover_initial=cover_initial, reflect_type=reflect_type,
                             dynamic=dynamic, label="ADSORBING", UDF_D=UDF_D, UDF_U=UDF_U, UDF_V=UDF_V,
                             UDF_T=UDF_T)


def InflowBC(D=1, T=1):
    """geom_bc_defs.py:416-429 (no velocity: it is taken from the adjacent cell)"""
    return BoundaryCondition(INFLOW, D=D, T=T, label="INFLOW")


def OutflowBC(P=1, D=0.0):
    """geom_bc_defs.py:431-443.  The reference constructor leaves D at 0, which
    makes WALL_*_P / WALL_*_D infinite in paraBC (clKernel.c:162-163); D is
    exposed here so a usable outflow can be specified."""
    return BoundaryCondition(OUTFLOW, P=P, D=D, label="OUTFLOW")
