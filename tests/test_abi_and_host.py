"""CPU checks: the C-ABI library exports what include/ugks_hip.h declares,
the ctypes descriptors match the header, and the host-side case logic."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from pyugks_amd import bc as bcd
from pyugks_amd import cases, lib
from pyugks_amd.data import cartesian_superblock, ghost_coordinates, subdivide_vertex_range, FlowCondition

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "ugks_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|size_t|const char \*)\s*(ugks_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = lib.load()
    names = header_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding covers the header exactly
    assert sorted(set(lib.exported_symbols())) == names


def test_descriptor_layouts():
    # ugks_block_desc: 8 ints, 9 doubles, 5x4 doubles, 4x4 doubles, 2x4 ints, 4 doubles, 5 pointers,
    # 2x4 doubles, 3x4 pointers, 2x4 ints; wall-value programs: 2 pointers, 2 ints, 2x16 ints
    expect = 8 * 4 + 9 * 8 + 20 * 8 + 16 * 8 + 8 * 4 + 4 * 8 + 5 * 8 + 8 * 8 + 12 * 8 + 8 * 4
    expect += 2 * 8 + 2 * 4 + 32 * 4
    assert C.sizeof(lib.BlockDesc) == expect
    assert C.sizeof(lib.BlockBuffers) == 22 * 8
    assert C.sizeof(lib.BCDesc) == 6 * 4
    assert C.sizeof(lib.DtDesc) == 6 * 8 + 4 * 4 + 8


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(lib, "_lib", None)
    with pytest.raises(lib.UGKSError):
        lib.load(str(tmp_path / "nope.so"))
    monkeypatch.setattr(lib, "_lib", None)
    lib.load()


def test_subdivide_and_superblock_connectivity():
    assert subdivide_vertex_range(10, 3) == [(0, 3), (3, 6), (6, 10)]
    blks = cases.cfg2_cavity_2x2(n_per_block=8, nc=8).blocks
    assert [b.label for b in blks] == ["sblk-0-0", "sblk-0-1", "sblk-1-0", "sblk-1-1"]
    assert blks[0].bc_list[bcd.EAST].type_of_BC == bcd.ADJACENT and blks[0].bc_list[bcd.EAST].other_block == 2
    assert blks[0].bc_list[bcd.NORTH].other_block == 1 and blks[0].bc_list[bcd.NORTH].other_face == bcd.SOUTH
    assert blks[3].bc_list[bcd.WEST].other_block == 1 and blks[3].bc_list[bcd.SOUTH].other_block == 2
    assert blks[0].bc_list[bcd.WEST].type_of_BC == bcd.DIFFUSE


def test_ghost_coordinates_extend_uniform_grid_and_join_blocks():
    g = cases.cfg2_cavity_2x2(n_per_block=6, nc=8)
    xys = ghost_coordinates(g.blocks, 3)
    h = 1.0 / 12
    for b, xy in zip(g.blocks, xys):
        x = xy[..., 0]
        # interior + ghost vertices along i at an interior j are uniformly spaced
        row = x[:, 5]
        assert np.allclose(np.diff(row), h, rtol=0, atol=1e-14)
        col = xy[5, :, 1]
        assert np.allclose(np.diff(col), h, rtol=0, atol=1e-14)
    # block (0,0) east ghosts are block (1,0)'s interior vertices, bit for bit
    assert np.array_equal(xys[0][10:13, 3:10], xys[2][4:7, 3:10])


def test_periodic_ghosts_are_translated():
    gg = cases._base("vanLeer", 0.1)
    cases.gauss_hermite(gg, 4)
    bcs = [bcd.PeriodicBC(0, bcd.SOUTH), bcd.PeriodicBC(0, bcd.WEST), bcd.PeriodicBC(0, bcd.NORTH),
           bcd.PeriodicBC(0, bcd.EAST)]
    gg.blocks = cartesian_superblock(0.0, 1.0, 0.0, 1.0, 4, 4, bcs, FlowCondition())
    xy = ghost_coordinates(gg.blocks, 2)[0]
    assert np.allclose(np.diff(xy[:, 3, 0]), 0.25, atol=1e-14)   # continues past x = 1 and below x = 0
    assert xy[0, 3, 0] == -0.5 and xy[-1, 3, 0] == 1.5


def test_channel_widths_and_adsorbing_mapping():
    w = cases.channel_widths()
    assert sum(w) == 1024 and len(w) == 8 and set(w) <= {64, 96, 128, 160, 192}
    ap = cases.adsorbing_params(0.3, 0.01)
    theta = ap["adsorb"]
    assert abs(0.3 * (1 / theta - 1) - 0.01) < 1e-15
    g = cases.cfg5_channel(widths=[4, 4], height=4, nc=8).prepare()
    c = g.block_constants(g.blocks[0])
    assert c["gamma_f"][bcd.SOUTH] == 0.3 and c["alpha_p"][bcd.SOUTH] == 1.0
    assert c["constant_adsorb"][bcd.SOUTH] == 1 and abs(c["adsorb"][bcd.SOUTH] - 0.967742) < 1e-15


def test_batch_cases_are_seeded_and_uneven():
    a = cases.cfg4_batch()
    b = cases.cfg4_batch()
    assert len(a) == 64
    sides = [x.blocks[0].ni for x in a]
    assert sides == [x.blocks[0].ni for x in b]
    assert len(set(sides)) > 3
    kn = [x.Kn_eff for x in a]
    assert min(kn) >= 1e-3 and max(kn) <= 1.0


def _bare_sim(tmp_path=None):
    from pyugks_amd.data import UGKSData
    from pyugks_amd.sim import UGKSim
    s = UGKSim.__new__(UGKSim)
    s.gdata = UGKSData()
    s.step = 0
    return s


def test_runtime_control_file_sets_only_loop_controls(tmp_path):
    """The per-step runtime file (ugks_data.py:373-380) is read as literal
    assignments; anything else is ignored, nothing is executed."""
    s = _bare_sim()
    conf = tmp_path / "run.conf"
    conf.write_text("gdata.max_time = 2.5\nexit = True\nCFL = 0.5  # comment\n"
                    "import os\nR = 3.0\nmax_step = __import__('os').getpid()\n")
    s.gdata.runtime_conf = str(conf)
    with pytest.raises(ValueError):
        s.read_runtime_conf()            # a non-literal value is refused, not evaluated
    conf.write_text("gdata.max_time = 2.5\nexit = True\nCFL = 0.5  # comment\nimport os\nR = 3.0\n")
    s.read_runtime_conf()
    assert s.gdata.max_time == 2.5 and s.gdata.exit is True and s.gdata.CFL == 0.5
    assert s.gdata.R == 287.0


def test_residual_slope_orders_power_laws():
    """analyse_residual (ugks_sim.py:722-777): the filtered log-log slope of
    the residual history.  The reference's short filtfilt padding biases the
    end-point estimate low, so what is checked is what its stop criterion
    relies on: a flat history gives slope 0 (it stops), and a steeper decay
    gives a larger slope, component by component."""
    s = _bare_sim()
    s.gdata.residual_options.slope_start = 10
    N = np.arange(1, 400, 3)
    s.time_history_residual_N = list(N)
    s.time_history_residual = [np.array([n ** -1.5, 2 * n ** -0.5, 0.3]) for n in N]
    s.step = int(N[-1])
    slopes = s.analyse_residual()
    assert slopes[0] > slopes[1] > 10 * s.gdata.residual_options.min_slope
    assert slopes[2] < 1e-12
    s.step = 5
    assert s.analyse_residual() == np.inf


def test_over_wide_block_is_refused_before_any_launch():
    """The face passes address a face's stencil with 32-bit byte offsets:
    (2R-1)*NJ*NV + NV pairs must stay below 2^32 bytes.  The check runs on the
    host (no device call) and plan creation applies it to every block."""
    L = lib.load()
    WENO5, VANLEER = 1, 0
    # the 2x2x128^2 cavity with 33^2 velocities fits easily
    assert L.ugks_step_block_fits(134, 1089, WENO5, lib.UGKS_FP64) == 0
    # largest NJ that fits for FP64 WENO5 with 49^2 velocities: (5 NJ + 1) * 2401 * 16 < 2^32
    nj_max = ((1 << 32) // (2401 * 16) - 1) // 5
    if ((5 * nj_max + 1) * 2401 * 16) >= (1 << 32):
        nj_max -= 1
    assert L.ugks_step_block_fits(nj_max, 2401, WENO5, lib.UGKS_FP64) == 0
    assert L.ugks_step_block_fits(nj_max + 1, 2401, WENO5, lib.UGKS_FP64) == -2
    assert b"32-bit" in L.ugks_last_error()
    # half the bytes per pair in the mixed step, a 3-cell span for vanLeer
    assert L.ugks_step_block_fits(nj_max + 1, 2401, WENO5, lib.UGKS_MIXED) == 0
    assert L.ugks_step_block_fits(5 * nj_max // 3, 2401, VANLEER, lib.UGKS_FP64) == 0
    assert L.ugks_step_block_fits(0, 2401, WENO5, lib.UGKS_FP64) == -1


def test_block_constants_from_objects_with_the_reference_attribute_names():
    """The drop-in binding builds a block's descriptor from the reference's own
    gdata and UGKSBlock: only attribute names those objects have are read."""
    import types

    from pyugks_amd import cases
    from pyugks_amd.data import block_constants, velocity_tables
    g = cases.cavity(6, ("NC", 8), "WENO5", 0.05, n_cells_j=4).prepare()
    ref_g = types.SimpleNamespace(**{k: getattr(g, k) for k in (
        "flux_method", "Nv", "relax_type", "Pr", "Kn_eff", "gamma", "b", "K", "omega", "CFL", "umax", "vmax",
        "quad", "weight", "mirror_NS", "mirror_EW", "mirror_D")})
    blk = g.blocks[0]
    ref_b = types.SimpleNamespace(ni=blk.ni, nj=blk.nj, bc_list=blk.bc_list)
    a, b = block_constants(ref_g, ref_b), g.block_constants(blk)
    assert set(a) == set(b)
    for k in a:
        if isinstance(a[k], list) and a[k] and isinstance(a[k][0], np.ndarray):
            continue
        assert a[k] == b[k] or (a[k] is None and b[k] is None), k
    for x, y in zip(velocity_tables(ref_g), g.velocity_tables()):
        assert np.array_equal(x, y)
    desc, keep = lib.make_desc(a, velocity_tables(ref_g))
    assert desc.NV == g.Nv and desc.ghost == 3


def _xcd_map(h, n, G=32):
    """The fused kernels' workgroup -> work item map (xcd_remap in
    ugks_fused.hip, runs of G items dealt to the 8 XCDs in turn), read from the
    source's own constant."""
    full = n // (8 * G) * (8 * G)
    if h >= full:
        return h
    x, k = h & 7, h >> 3
    return ((k // G) * 8 + x) * G + (k % G)


def test_xcd_map_is_a_bijection_and_balances_the_batch():
    """Every workgroup id maps to one work item and every item is reached, for
    any grid size; and on the 64-case batch (cases of 289 to 1681 velocities
    in random order) each XCD's share of the face-pass work is within 1% of
    the mean, where one contiguous eighth per XCD was 48% over it."""
    import copy
    src = open(os.path.join(ROOT, "pyugks_amd", "csrc", "ugks_fused.hip")).read()
    G = int(re.search(r"#define UGKS_XCD_GROUP (\d+)", src).group(1))
    assert G == 32
    for n in list(range(1, 700)) + [8 * G * 5 + 3, 100003]:
        got = sorted(_xcd_map(h, n, G) for h in range(n))
        assert got == list(range(n)), n
    work = []
    for g in cases.cfg4_batch():
        gp = copy.deepcopy(g).prepare()
        for b in gp.blocks:
            work += [gp.Nv] * ((b.ni * (b.nj - 1) + (b.ni - 1) * b.nj) // 4)
    work = np.asarray(work, dtype=np.float64)
    n = len(work)
    share = np.zeros(8)
    items = np.array([_xcd_map(h, n, G) for h in range(n)])
    np.add.at(share, np.arange(n) & 7, work[items])
    assert share.max() / share.mean() < 1.01
    contiguous = np.array([c.sum() for c in np.array_split(work, 8)])
    assert contiguous.max() / contiguous.mean() > 1.3
