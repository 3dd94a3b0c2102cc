"""pyugks_amd -- the pyUGKS time step (UGKS, Shakhov/BGK) on AMD MI355X.

Host counterpart of the reference's block runtime (``UGKSBlock``,
``UGKSim``) over a C-ABI HIP library (``libugks_hip.so``, include/ugks_hip.h).
"""
from . import bc, data, cases  # noqa: F401

__all__ = ["bc", "data", "cases"]
