"""GPU checks of the error flag, the residual monitor and the pyopencl call
form, against the oracle and against numpy computed from the oracle's state.

  isotherm clamp flag (clAdsorb.c:430-434,    a wall state off the tabulated
    polled at ugks_block.py:840-876)          isotherm raises code 4 on the first
                                              step on both device paths, as in
                                              the oracle
  getResidual + host normalisation            sqrt(ni nj sum (prim_new -
    (clUGKS.c:1259-1280, ugks_block.py:        prim_old)^2) / sum |prim_new| per
    1303-1327)                                component, from the oracle's
                                              macro before and after the step
  edgeMassFlow (clUGKS.c:1283-1337)           sum over the open faces of
                                              (n_out . U) rho L from the oracle's
                                              macro and geometry
  prg.<kernel>(queue, global, local, *args)   the reference's call form drives
    (ugks_block.py:1066-1069)                 calcMacro, calcQ, iFace, getFaceCons
                                              and clFindDT to the oracle's values
"""
import copy
import ctypes as C

import numpy as np
import pytest

from helpers import rel_err
from pyugks_amd import bc as bcd
from pyugks_amd import cases

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _off_table_channel():
    """An adsorbing channel whose isotherm table covers P in [0.02, 0.2]:
    the wall state (P = rho T / 2 = 0.5) is off it."""
    rng = np.random.default_rng(3)
    P, T = rng.uniform(0.02, 0.2, 30), rng.uniform(0.6, 1.6, 30)
    table = np.stack([P, T, 0.5 * P / (P + 0.1)], axis=1)
    g = cases.cfg5_channel(widths=[6, 8], height=6, nc=8, Kn=0.05)
    for b in g.blocks:
        b.bc_list[bcd.SOUTH].adsorb = table.copy()
    return g


def test_isotherm_clamp_flag_code_4(hip):
    """A wall state off the tabulated isotherm: adsorbingWall_P1 raises code 4
    and returns early for those wall cells (their coverage is left as it
    was); UGKSBlock.wall_flux then skips the wall's second pass, wall flux
    and wall mass/energy fluxes, but UGKS_flux ignores its return value, so
    the step goes on and UGKS_update clears the flag (clAdsorb.c:430-434,
    ugks_block.py:930-983, :1049-1260, :1287).  The wall faces keep their
    reconstructed values as fluxes, and the next step ends in NaN.  The
    per-kernel GPU path reports the error, matches the oracle after the
    first step and raises the NaN flag on the second, as the oracle does; the
    fused plan raises code 4 on the first step."""
    from oracle import OracleSim
    from pyugks_amd.fused import StepPlan
    from pyugks_amd.sim import UGKSim
    print("[gpu] isotherm clamp flag, reference path and fused plan", flush=True)
    g = _off_table_channel()
    ora = OracleSim(copy.deepcopy(g), threads=4)
    gpu = UGKSim(copy.deepcopy(g))
    cover0 = [b["wall_cover"].copy() for b in ora.blocks]
    ora.one_step(True)
    _, err = gpu.one_step(True)
    assert not err
    for bg, bo in zip(gpu.blocks, ora.blocks):
        assert rel_err(bg.f_D.download(), bo["f"]) <= TOL
        assert rel_err(bg.macro_D.download(), bo["macro"]) <= TOL
        assert rel_err(bg.wall_cover_D.download(), bo["wall_cover"]) <= TOL
    # the off-table cells kept their coverage (P1 returned before updating it)
    assert all(np.array_equal(c0, bo["wall_cover"]) for c0, bo in zip(cover0, ora.blocks))
    assert any(getattr(b, "last_error", "").startswith("pressure / temperature outside") for b in gpu.blocks)
    with pytest.raises(RuntimeError):
        ora.one_step(True)
    _, err = gpu.one_step(True)
    assert err and gpu.last_error == "NaN encountered"
    for precision in ("fp64", "mixed"):
        sim = UGKSim(copy.deepcopy(g))
        plan = StepPlan(sim, precision)
        plan.run(1)
        assert plan.flag() == 4.0, precision
        plan.close()


def test_residual_and_boundary_mass_flow_against_the_oracle_state(hip):
    from oracle import OracleSim
    from pyugks_amd.sim import UGKSim
    print("[gpu] residual monitor and boundary mass flow", flush=True)
    g = cases.cfg5_channel(widths=[6, 8], height=6, nc=8, Kn=0.05)
    gpu = UGKSim(copy.deepcopy(g))
    ora = OracleSim(copy.deepcopy(g), threads=4)
    for step in range(3):
        before = [b["macro"].copy() for b in ora.blocks]
        _, err = gpu.one_step(True, get_res=True)
        assert not err
        ora.one_step(True)
        nopen = 0
        for bg, bo, m0 in zip(gpu.blocks, ora.blocks, before):
            m1 = bo["macro"]
            ni, nj = m1.shape[:2]
            expect = np.sqrt(ni * nj * np.sum((m1 - m0) ** 2, axis=(0, 1))) / np.sum(np.abs(m1), axis=(0, 1))
            assert np.all(expect > 0)
            assert np.allclose(bg.residual, expect, rtol=1e-9, atol=0), (step, bg.residual, expect)
            G = bo.consts["ghost"]
            n, L = bo["normal"], bo["side_length"]
            mdot = 0.0
            for face, bcx in enumerate(bg.bc_list):
                if bcx.type_of_BC not in (bcd.EXTRAPOLATE_OUT, bcd.INFLOW, bcd.OUTFLOW, bcd.CONSTANT):
                    continue
                # the boundary face of each edge cell, its normal turned outward
                if face == bcd.NORTH:
                    nn, ll, cells = n[G:G + ni, G + nj, 0], L[G:G + ni, G + nj, 0], m1[:, -1]
                elif face == bcd.SOUTH:
                    nn, ll, cells = -n[G:G + ni, G, 0], L[G:G + ni, G, 0], m1[:, 0]
                elif face == bcd.EAST:
                    nn, ll, cells = n[G + ni, G:G + nj, 1], L[G + ni, G:G + nj, 1], m1[-1, :]
                else:
                    nn, ll, cells = -n[G, G:G + nj, 1], L[G, G:G + nj, 1], m1[0, :]
                mdot += np.sum((nn[:, 0] * cells[:, 1] + nn[:, 1] * cells[:, 2]) * cells[:, 0] * ll)
                nopen += 1
            assert abs(bg.bc_mdot - mdot) <= 1e-12 * max(abs(mdot), 1e-300), (step, bg.bc_mdot, mdot)
        assert nopen == 2                   # the inflow and the outflow face


def test_reference_call_form_drives_the_kernels(hip):
    """lib.CLCompatProgram takes the pyopencl form prg.<kernel>(queue,
    global_size, local_size, *args) with the reference's buffers, and the
    kernels so driven reproduce the oracle's calcMacro, calcQ, iFace,
    getFaceCons, clFindDT and getInternal on the oracle's initial state."""
    import oracle as O
    from pyugks_amd import lib as L
    print("[gpu] pyopencl call form", flush=True)
    g = cases.cavity(10, ("NC", 12), "WENO5", 0.05, lid=0.2, n_cells_j=8, rho_perturb=0.01)
    o = O.OracleSim(copy.deepcopy(g), threads=2)
    b = o.blocks[0]
    c = b.consts
    gp = copy.deepcopy(g).prepare()
    desc, keep = L.make_desc(gp.block_constants(gp.blocks[0]), gp.velocity_tables())
    prg = L.CLCompatProgram(desc, keep)
    NI, NJ, NV, ni, nj = c["NI"], c["NJ"], c["NV"], c["ni"], c["nj"]
    dev = {k: L.DeviceArray.from_host(b[k]) for k in ("f", "centre", "mid_side", "normal", "xy", "area", "Q")}
    macro_D = L.DeviceArray((ni, nj, 4), np.float64, fill=0)
    Q_D = L.DeviceArray((ni, nj, 2), np.float64, fill=0)
    face_f = L.DeviceArray((NI, NJ, NV, 2), np.float64, fill=0)
    sigma = L.DeviceArray((NI, NJ, NV, 2), np.float64, fill=0)
    prim = L.DeviceArray((NI, NJ, 4), np.float64, fill=0)
    ts = L.DeviceArray((ni, nj), np.float64, fill=0)
    gsz = (ni, nj, 1)
    prg.calcMacro(None, gsz, None, dev["f"], macro_D)
    prg.calcQ(None, gsz, None, dev["f"], macro_D, Q_D)
    prg.iFace(None, gsz, None, dev["f"], face_f, sigma, dev["centre"], dev["mid_side"], dev["normal"],
              np.int32(0))
    prg.getFaceCons(None, gsz, None, face_f, dev["normal"], np.int32(0), prim, np.int32(1), np.int32(1))
    prg.clFindDT(None, gsz, None, dev["xy"], dev["area"], macro_D, ts, np.float64(gp.Kn_eff))
    txyz_D, stress_D = L.DeviceArray((ni, nj, 4), np.float64, fill=0), L.DeviceArray((ni, nj, 4), np.float64, fill=0)
    prg.getInternal(None, gsz, None, dev["f"], txyz_D, stress_D)
    L.synchronize()
    # the oracle's own kernels on the same inputs
    lib = O.lib()
    ptr = lambda a: C.c_void_p(a.ctypes.data)
    p = C.byref(b.block.p)
    om, oq = np.zeros((ni, nj, 4)), np.zeros((ni, nj, 2))
    of, osg, opr, ots = np.zeros((NI, NJ, NV, 2)), np.zeros((NI, NJ, NV, 2)), np.zeros((NI, NJ, 4)), np.zeros((ni, nj))
    lib.or_calcMacro.argtypes = [C.c_void_p] * 3
    lib.or_calcQ.argtypes = [C.c_void_p] * 4
    lib.or_iFace.argtypes = [C.c_void_p] * 7 + [C.c_int]
    lib.or_getFaceCons.argtypes = [C.c_void_p] * 3 + [C.c_int, C.c_void_p, C.c_int, C.c_int]
    lib.or_clFindDT.argtypes = [C.c_void_p] * 5 + [C.c_double]
    lib.or_calcMacro(p, ptr(b["f"]), ptr(om))
    lib.or_calcQ(p, ptr(b["f"]), ptr(om), ptr(oq))
    lib.or_iFace(p, ptr(b["f"]), ptr(of), ptr(osg), ptr(b["centre"]), ptr(b["mid_side"]), ptr(b["normal"]), 0)
    lib.or_getFaceCons(p, ptr(of), ptr(b["normal"]), 0, ptr(opr), 1, 1)
    lib.or_clFindDT(p, ptr(b["xy"]), ptr(b["area"]), ptr(om), ptr(ots), gp.Kn_eff)
    G = c["ghost"]
    inner = (slice(G, G + ni), slice(G, G + nj + 1))
    assert rel_err(macro_D.download(), om) <= TOL
    assert rel_err(Q_D.download(), oq, 1e-3) <= TOL
    assert rel_err(face_f.download()[inner], of[inner]) <= TOL
    assert rel_err(prim.download()[G + 1:G + ni, G + 1:G + nj], opr[G + 1:G + ni, G + 1:G + nj]) <= TOL
    assert rel_err(ts.download(), ots) <= TOL
    otxyz, ostress = b.internal()
    assert rel_err(txyz_D.download(), otxyz) <= TOL and rel_err(stress_D.download(), ostress) <= TOL
    prg.close()
