"""Batches of grids (vls_batch_*): every grid of a batch against the
single-grid entry points on the same data, bit for bit (area/volume sums to
rounding: their summation order differs), in FP64 (where the single-grid path
is the CPU oracle's own result) and FP32 bulk; the persistent launch shape
and hipGraph replays do not change any result."""
import importlib

import numpy as np
import pytest
import torch

from golden import cases

vls = importlib.import_module("2018_7_vesicle_amd")
bm = importlib.import_module("2018_7_vesicle_amd.batch")

pytestmark = pytest.mark.gpu

# uneven boxes: rows not a multiple of 64 (partial tiles), distinct spacings per axis
SHAPES = [(20, 24, 18), (33, 17, 21), (64, 12, 9), (70, 26, 15)]


def _grids():
    out = []
    for k, (rows, cols, pges) in enumerate(SHAPES):
        xv = np.linspace(-1.0, 1.0, cols)
        yv = np.linspace(-0.9 - 0.05 * k, 0.9, rows)
        zv = np.linspace(-1.1, 1.0 + 0.1 * k, pges)
        out.append(cases.Grid(xv, yv, zv))
    return out


def _fields(g, k):
    r = np.sqrt((g.X - 0.05 * k) ** 2 + g.Y ** 2 + (g.Z + 0.03 * k) ** 2)
    F = (r - 0.55) * (1 + 0.15 * np.sin(3 * g.X + k) * np.cos(2 * g.Y))
    C = np.sin(2 * g.X) + g.Y * g.Z + 0.1 * k
    u = cases.smooth_velocity(g)
    return F, C, u


def _setup(dtype):
    grids = _grids()
    b = bm.Batch([(g.rows, g.cols, g.pges) for g in grids], [g.spacing for g in grids], dtype=dtype)
    data = [_fields(g, k) for k, g in enumerate(grids)]
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)  # noqa: E731
    F = b.pack([T(d[0]) for d in data])
    C = b.pack([T(d[1]) for d in data])
    u = [b.pack([T(d[2][a]) for d in data]) for a in range(3)]
    Gs = [vls.GD3(g.xv, g.yv, g.zv) for g in grids]
    return b, grids, Gs, F, C, u


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batch_methods_equal_single_grids(dtype):
    b, grids, Gs, F, C, u = _setup(dtype)
    # per-grid dt
    dt = b.cfl_dt(*u, 0.4)
    dt_h = dt.cpu().numpy()
    singles = []
    for gi, G in enumerate(Gs):
        m = vls.SDF3(G, b.grid_view(F, gi).clone())
        ug = [b.grid_view(c, gi).contiguous() for c in u]
        assert dt_h[gi] == 0.4 * G.Dx / m.maxSpeed(*ug)
        singles.append((m, ug))
    # reinitialization (mask and distances of F)
    b.set_distance(F)
    b.reinitialize(F, 4)
    # extension of C off the reinitialized F
    C0 = C.clone()
    b.set_distance(F)
    b.extend(F, C, 5)
    # advection and normal velocity
    Fa = F.clone()
    b.advect(Fa, *u, dt, 2)
    Vn = b.zeros()
    b.normal_velocity(Fa, *u, Vn)
    av = b.area_volume(Fa)
    for gi, (m, ug) in enumerate(singles):
        m.setDistance()
        Fr = m.WENO5RK3Reinitialization(m.F, 4)
        assert torch.equal(Fr, b.grid_view(F, gi)), f"reinit grid {gi}"
        m.F = Fr
        m.setDistance()
        Cg = m.WENO5RK3Extend(b.grid_view(C0, gi).clone(), 5)
        assert torch.equal(Cg, b.grid_view(C, gi)), f"extend grid {gi}"
        Fa_g = m.WENO5RK3Advection(Fr, *ug, float(dt_h[gi]), 2)
        assert torch.equal(Fa_g, b.grid_view(Fa, gi)), f"advect grid {gi}"
        m.F = Fa_g
        vn = torch.zeros_like(Fa_g)
        m.normalVelocity(*ug, vn)
        assert torch.equal(vn, b.grid_view(Vn, gi)), f"normal velocity grid {gi}"
        o = torch.zeros(2, dtype=torch.float64, device="cuda")
        m.areaVolumeDevice(o, 1, 1.5 * Gs[gi].Dx)
        assert torch.allclose(o, av[gi], rtol=1e-13, atol=0), (o, av[gi])


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batch_step_equals_single_grid_sequence(dtype):
    """the fused (graph-captured) batch step against the same sequence of
    single-grid calls, on a reinitialization step"""
    b, grids, Gs, F, C, u = _setup(dtype)
    dt = b.cfl_dt(*u, 0.4)
    b.set_distance(F)
    F0 = F.clone()
    Vn = b.zeros()
    av = torch.zeros((b.ngrids, 2), dtype=torch.float64, device="cuda")
    b.step(F, Vn, *u, dt, 6, 5, av)
    for gi, G in enumerate(Gs):
        ug = [b.grid_view(c, gi).contiguous() for c in u]
        m = vls.SDF3(G, b.grid_view(F0, gi).clone())
        m.F = m.WENO5RK3Advection(m.F, *ug, float(dt[gi]), 1)
        m.setDistance()
        vn = torch.zeros_like(m.F)
        m.normalVelocity(*ug, vn)
        vn = m.WENO5RK3Extend(vn, 6)
        assert torch.equal(vn, b.grid_view(Vn, gi)), f"Vn grid {gi}"
        m.F = m.WENO5RK3Reinitialization(m.F, 5)
        assert torch.equal(m.F, b.grid_view(F, gi)), f"F grid {gi}"
        o = torch.zeros(2, dtype=torch.float64, device="cuda")
        m.areaVolumeDevice(o, 1, 1.5 * G.Dx)
        assert torch.allclose(o, av[gi], rtol=1e-13, atol=0)


def test_persistent_launch_shape_does_not_change_results():
    """persistent block count of the tile-loop kernels and z-march depth of
    the bulk stage kernels"""
    outs = []
    for blocks, depth in ((8, 8), (64, 1), (0, 5), (0, 40)):
        b, grids, Gs, F, C, u = _setup(torch.float32)
        b.blocks = blocks
        b.zmarch_depth = depth
        assert b.zmarch_depth == depth
        b.set_distance(F)
        b.reinitialize(F, 3)
        b.extend(F, C, 3)
        outs.append((F.clone(), C.clone()))
    for F, C in outs[1:]:
        assert torch.equal(F, outs[0][0]) and torch.equal(C, outs[0][1])


def test_graph_replays_equal_eager():
    """a captured batch step replayed three times, with eager work and
    synchronizations between the replays, against the eager result"""
    b, grids, Gs, F, C, u = _setup(torch.float32)
    dt = b.cfl_dt(*u, 0.4)
    b.set_distance(F)
    F0 = F.clone()
    prev = vls.set_graphs(False)
    try:
        Fe, Ve = F0.clone(), b.zeros()
        ave = torch.zeros((b.ngrids, 2), dtype=torch.float64, device="cuda")
        b.step(Fe, Ve, *u, dt, 4, 3, ave)
        torch.cuda.synchronize()
        vls.set_graphs(True)
        Fg, Vg = F0.clone(), b.zeros()
        avg = torch.zeros((b.ngrids, 2), dtype=torch.float64, device="cuda")
        for rep in range(3):
            Fg.copy_(F0)
            Vg.zero_()
            avg.zero_()
            b.step(Fg, Vg, *u, dt, 4, 3, avg)
            junk = (Fg.double() * 2).sum()  # eager work between replays
            torch.cuda.synchronize()
            assert torch.isfinite(junk)
            assert torch.equal(Fg, Fe) and torch.equal(Vg, Ve) and torch.equal(avg, ave), f"replay {rep}"
    finally:
        vls.set_graphs(prev)


def test_single_grid_graph_replays_equal_eager():
    g = cases.linspace_grid(28)
    G = vls.GD3(g.xv, g.yv, g.zv)
    F, C, _ = _fields(g, 0)
    Ft = torch.from_numpy(F).cuda()
    Ct = torch.from_numpy(C).cuda()
    m = vls.SDF3(G, Ft)
    m.setDistance()
    prev = vls.set_graphs(False)
    try:
        ref_c = m.WENO5RK3Extend(Ct, 4)
        ref_f = m.WENO5RK3Reinitialization(Ft, 4)
        vls.set_graphs(True)
        out_c = Ct.clone()
        out_f = Ft.clone()
        for rep in range(3):
            out_c.copy_(Ct)
            out_f.copy_(Ft)
            from oracle import pyoracle  # noqa: F401  (eager host work between replays)
            vls.lib().vls_sdf3_weno5_rk3_extend(m._h, m.F.data_ptr(), out_c.data_ptr(), 4, None)
            vls.lib().vls_sdf3_weno5_rk3_reinit(m._h, out_f.data_ptr(), 4, None)
            torch.cuda.synchronize()
            assert torch.equal(out_c, ref_c) and torch.equal(out_f, ref_f), f"replay {rep}"
    finally:
        vls.set_graphs(prev)


def test_synthetic_inputs_match_numpy():
    cfg = bm.BatchConfig(vesicles=3, edges=(24, 32))
    edges, axes, modes = bm.batch_inputs(cfg)
    b = bm.Batch([(int(e),) * 3 for e in edges], [(2.0 / (e - 1),) * 3 for e in edges], dtype=torch.float64)
    lo = torch.full((3, 3), -1.0, dtype=torch.float64, device="cuda")
    F = b.empty()
    b.ellipsoid(torch.from_numpy(axes).cuda(), lo, F)
    u = [b.empty() for _ in range(3)]
    b.curl_velocity(torch.from_numpy(modes).cuda(), lo, *u)
    for v, e in enumerate(edges):
        e = int(e)
        dx = 2.0 / (e - 1)
        xv = -1.0 + dx * np.arange(e)
        Z, X, Y = np.meshgrid(xv, xv, xv, indexing="ij")
        a, bb, c = axes[v]
        Fn = np.sqrt(X ** 2 / a ** 2 + Y ** 2 / bb ** 2 + Z ** 2 / c ** 2) - 1.0
        assert np.abs(b.grid_view(F, v).cpu().numpy() - Fn).max() < 1e-15
        un = np.zeros((3,) + X.shape)
        for row in modes[v]:
            s = -np.sin(row[0] * X + row[1] * Y + row[2] * Z + row[6])
            for k in range(3):
                un[k] += s * row[3 + k]
        for k in range(3):
            assert np.abs(b.grid_view(u[k], v).cpu().numpy() - un[k]).max() < 1e-14


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_tile_shapes_give_same_results(dtype):
    """single grid and batch: 32 x 8 tiles give the 64 x 4 result bit for bit"""
    g = cases.Grid(np.linspace(-1, 1, 37), np.linspace(-1, 1, 45), np.linspace(-1, 1, 29))
    G = vls.GD3(g.xv, g.yv, g.zv)
    F, C, u = _fields(g, 1)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)  # noqa: E731
    outs = []
    for tr in (64, 32):
        m = vls.SDF3(G, T(F))
        m.setTileRows(tr)
        m.setDistance()
        Fr = m.WENO5RK3Reinitialization(m.F, 3)
        m.F = Fr
        m.setDistance()
        Cn = m.WENO5RK3Extend(T(C), 3)
        Fa = m.WENO5RK3Advection(Fr, *[T(c) for c in u], 0.01, 2)
        outs.append((Fr, Cn, Fa))
    for a, b in zip(*outs):
        assert torch.equal(a, b)
    bouts = []
    for tr in (64, 32):
        b, grids, Gs, Fb, Cb, ub = _setup(dtype)
        b.tile_rows = tr
        assert b.tile_rows == tr
        b.set_distance(Fb)
        b.reinitialize(Fb, 3)
        b.extend(Fb, Cb, 3)
        bouts.append((Fb, Cb))
    for a, c in zip(*bouts):
        assert torch.equal(a, c)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_zmarch_depth_gives_same_results(dtype):
    """single grid: the pages per block of the z-marching bulk kernels (1, 3,
    all 29, and the default) do not change any result"""
    g = cases.Grid(np.linspace(-1, 1, 37), np.linspace(-1, 1, 45), np.linspace(-1, 1, 29))
    G = vls.GD3(g.xv, g.yv, g.zv)
    F, C, u = _fields(g, 2)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)  # noqa: E731
    outs = []
    for tr, depth in [(64, 0), (64, 1), (64, 3), (64, 29), (32, 5)]:
        m = vls.SDF3(G, T(F))
        m.setTileRows(tr)
        m.setZmarchDepth(depth)
        got = m.launch()
        assert got[0] == tr and (depth == 0 or got[1] == depth)
        m.setDistance()
        Fr = m.WENO5RK3Reinitialization(m.F, 3)
        m.F = Fr
        m.setDistance()
        Cn = m.WENO5RK3Extend(T(C), 3)
        Fa = m.WENO5RK3Advection(Fr, *[T(c) for c in u], 0.01, 2)
        outs.append((Fr, Cn, Fa))
    for o in outs[1:]:
        for a, b in zip(o, outs[0]):
            assert torch.equal(a, b)


def test_autotune_keeps_results():
    b, grids, Gs, F, C, u = _setup(torch.float32)
    ref_F, ref_C = F.clone(), C.clone()
    b.set_distance(ref_F)
    b.reinitialize(ref_F, 2)
    b.extend(ref_F, ref_C, 2)
    tr, blocks = b.autotune(F, 1)
    assert tr in (32, 64) and blocks % 8 == 0 and blocks >= 8
    b.set_distance(F)
    b.reinitialize(F, 2)
    b.extend(F, C, 2)
    assert torch.equal(F, ref_F) and torch.equal(C, ref_C)
    g = grids[3]
    m = vls.SDF3(Gs[3], b.grid_view(F, 3).clone())
    assert m.autotune(1) in (32, 64)


def test_reinit_and_extend_need_set_distance_first():
    # they read the interface list that set_distance builds; before it exists
    # they refuse, with no device work, instead of reading uninitialized lists
    b, grids, Gs, F, C, u = _setup(torch.float64)
    with pytest.raises(vls.VLSError, match="set_distance"):
        b.reinitialize(F, 2)
    with pytest.raises(vls.VLSError, match="set_distance"):
        b.extend(F, C, 2)
    b.set_distance(F)
    b.extend(F, C, 1)
    b.reinitialize(F, 1)
    torch.cuda.synchronize()
    G = Gs[0]
    g = grids[0]
    m = vls.SDF3(G, torch.from_numpy(np.ascontiguousarray(_fields(g, 0)[0])).cuda())
    with pytest.raises(vls.VLSError, match="set_distance"):
        m.WENO5RK3Reinitialization(m.F, 2)
    m.setDistance()
    m.WENO5RK3Reinitialization(m.F, 1)
    torch.cuda.synchronize()


def test_batch_variants_equal_single_grids():
    """WENORK3Reinitialization, WENORK3Extend and ENORK2Extend on a batch of
    uneven grids, every grid against the single-grid variant on the same data,
    bit for bit (the single-grid variants are the oracle's results: the
    variant tests); afterwards the WENO5 calls need set_distance again."""
    b, grids, Gs, F, C, u = _setup(torch.float64)
    Fr = F.clone()
    b.wenork3_reinitialize(Fr, 7)
    Cw = C.clone()
    b.wenork3_extend(F, Cw, 6)
    Ce = C.clone()
    b.enork2_extend(F, Ce, 5)
    torch.cuda.synchronize()
    for gi, G in enumerate(Gs):
        m = vls.SDF3(G, b.grid_view(F, gi).clone())
        assert torch.equal(m.WENORK3Reinitialization(m.F, 7), b.grid_view(Fr, gi)), f"WENORK3 reinit grid {gi}"
        c0 = b.grid_view(C, gi).clone()
        assert torch.equal(m.WENORK3Extend(c0, 6), b.grid_view(Cw, gi)), f"WENORK3Extend grid {gi}"
        assert torch.equal(m.ENORK2Extend(c0, 5), b.grid_view(Ce, gi)), f"ENORK2Extend grid {gi}"
    with pytest.raises(vls.VLSError, match="set_distance"):
        b.reinitialize(F, 1)
    b.set_distance(F)
    b.reinitialize(F, 1)
    b32 = bm.Batch([(g.rows, g.cols, g.pges) for g in grids], [g.spacing for g in grids], dtype=torch.float32)
    with pytest.raises(vls.VLSError, match="FP64"):
        b32.wenork3_extend(b32.zeros(), b32.zeros(), 1)
