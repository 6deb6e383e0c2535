"""Batches of level-set grids over the C ABI (vls_batch_*), and the batched
vesicle workload.

A `Batch` holds many independent grids of different sizes.  Its fields are
packed 1-D torch tensors: grid g occupies elements offsets[g]..offsets[g+1]-1
in the column-major layout of a single grid, and `grid_view(field, g)` gives
that grid as a (pges, cols, rows) view.  Every method advances all grids with
one launch per operator stage; each grid gets what the SDF3 method of the
same name gives it alone (bit for bit, the area/volume sums to rounding).
Methods work in place on the tensors they are given.

`BatchPipeline` is the batched vesicle workload: n vesicles, each on a cube
of edge drawn from `edges` over [-1,1]^3, an ellipsoid with semi-axes drawn
from `axes_range`, and its own curl velocity of `modes` Fourier modes (seed
seed + v for vesicle v, drawn as pipeline.curl_velocity draws them).  A step
advects F (TVD-RK3, dt = cfl*dx/max|u| per grid), locates the interface,
extends u.N off it (`ext_iters` WENO5-RK3 iterations), reinitializes F every
`reinit_every` steps (`reinit_iters` iterations) and records the cosine
area and volume of every grid.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from ._lib import D, I, P, VLSError, call
from .sdf3 import _stream


def _p(t):
    return P(t.data_ptr())


class Batch:
    def __init__(self, shapes, spacings, dtype=torch.float32, device="cuda"):
        """shapes: (rows, cols, pges) per grid; spacings: (dx, dy, dz) per grid."""
        if dtype not in (torch.float32, torch.float64):
            raise VLSError(f"dtype must be float32 or float64, got {dtype}")
        self.dtype = dtype
        self.device = torch.device(device)
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        self.spacings = [tuple(float(v) for v in s) for s in spacings]
        n = len(self.shapes)
        if n == 0 or len(self.spacings) != n:
            raise VLSError("one spacing triple per grid, at least one grid")
        arr = lambda vals, ct: (ct * n)(*vals)  # noqa: E731
        self._h = ctypes.c_void_p()
        call("vls_batch_create", ctypes.byref(self._h), n, arr([s[0] for s in self.shapes], ctypes.c_int),
             arr([s[1] for s in self.shapes], ctypes.c_int), arr([s[2] for s in self.shapes], ctypes.c_int),
             arr([s[0] for s in self.spacings], ctypes.c_double), arr([s[1] for s in self.spacings], ctypes.c_double),
             arr([s[2] for s in self.spacings], ctypes.c_double))
        offs = (ctypes.c_int64 * (n + 1))()
        call("vls_batch_layout", self._h, offs, None)
        self.offsets = list(offs)
        self.ngrids = n
        self.total = self.offsets[-1]
        self._sfx = "_f32" if dtype == torch.float32 else ""

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                torch.cuda.synchronize()
                call("vls_batch_destroy", h)
            except Exception:
                pass
            self._h = None

    # -- layout --------------------------------------------------------------
    def empty(self, dtype=None):
        return torch.empty(self.total, dtype=dtype or self.dtype, device=self.device)

    def zeros(self, dtype=None):
        return torch.zeros(self.total, dtype=dtype or self.dtype, device=self.device)

    def grid_view(self, packed, g):
        rows, cols, pges = self.shapes[g]
        return packed[self.offsets[g]:self.offsets[g + 1]].view(pges, cols, rows)

    def pack(self, fields):
        """pack per-grid (pges, cols, rows) tensors into one field"""
        out = self.empty()
        for g, f in enumerate(fields):
            self.grid_view(out, g).copy_(f)
        return out

    @property
    def ntiles(self):
        offs = (ctypes.c_int64 * (self.ngrids + 1))()
        nt = ctypes.c_int()
        call("vls_batch_layout", self._h, offs, ctypes.byref(nt))
        return nt.value

    @property
    def blocks(self):
        b = ctypes.c_int()
        call("vls_batch_get_blocks", self._h, ctypes.byref(b))
        return b.value

    @blocks.setter
    def blocks(self, value):
        call("vls_batch_set_blocks", self._h, int(value))

    @property
    def tile_rows(self):
        r = ctypes.c_int()
        call("vls_batch_get_tile_rows", self._h, ctypes.byref(r))
        return r.value

    @tile_rows.setter
    def tile_rows(self, value):
        call("vls_batch_set_tile_rows", self._h, int(value))

    @property
    def zmarch_depth(self):
        """pages per work item of the z-marching bulk stage kernels"""
        d = ctypes.c_int()
        call("vls_batch_get_zmarch_depth", self._h, ctypes.byref(d))
        return d.value

    @zmarch_depth.setter
    def zmarch_depth(self, value):
        call("vls_batch_set_zmarch_depth", self._h, int(value))

    def autotune(self, F, iterations=2):
        """pick the fastest tile shape and z-march depth for extension stages off
        F's interface (see zmarch_depth); returns (tile_rows, blocks).
        Synchronizes."""
        tr, nb = ctypes.c_int(), ctypes.c_int()
        call("vls_batch_autotune" + self._sfx, self._h, self._f(F), int(iterations), ctypes.byref(tr),
             ctypes.byref(nb), _stream())
        return tr.value, nb.value

    def _f(self, t):
        if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != self.dtype or not t.is_contiguous() \
                or t.numel() != self.total:
            raise VLSError(f"expected a contiguous {self.dtype} CUDA tensor of {self.total} elements")
        return _p(t)

    def _d(self, t, n):
        if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float64 or not t.is_contiguous() \
                or t.numel() < n:
            raise VLSError(f"expected a contiguous float64 CUDA tensor of at least {n} elements")
        return _p(t)

    # -- methods (in place) ----------------------------------------------------
    def set_distance(self, F):
        call("vls_batch_set_distance" + self._sfx, self._h, self._f(F), _stream())

    def reinitialize(self, F, iterations):
        call("vls_batch_weno5_rk3_reinit" + self._sfx, self._h, self._f(F), int(iterations), _stream())

    def extend(self, F, C, iterations):
        call("vls_batch_weno5_rk3_extend" + self._sfx, self._h, self._f(F), self._f(C), int(iterations), _stream())

    # the WENO-RK3 / ENO-RK2 variants of every grid (FP64 batches; each
    # computes its own distances of F, and leaves set_distance to be rerun
    # before a WENO5 reinitialization or extension)
    def _fp64(self, name):
        if self.dtype != torch.float64:
            raise VLSError(f"{name} runs on FP64 batches only")

    def wenork3_reinitialize(self, F, iterations):
        """WENORK3Reinitialization of every grid (+SD/@SDF3/WENORK3Reinitialization.m)"""
        self._fp64("wenork3_reinitialize")
        call("vls_batch_wenork3_reinit", self._h, self._f(F), int(iterations), _stream())

    def wenork3_extend(self, F, C, iterations):
        """WENORK3Extend of C off F on every grid (+SD/@SDF3/WENORK3Extend.m)"""
        self._fp64("wenork3_extend")
        call("vls_batch_wenork3_extend", self._h, self._f(F), self._f(C), int(iterations), _stream())

    def enork2_extend(self, F, C, iterations):
        """ENORK2Extend of C off F on every grid (+SD/@SDF3/ENORK2Extend.m)"""
        self._fp64("enork2_extend")
        call("vls_batch_enork2_extend", self._h, self._f(F), self._f(C), int(iterations), _stream())

    def advect(self, F, vx, vy, vz, dt, steps=1):
        call("vls_batch_advect_rk3" + self._sfx, self._h, self._f(F), self._f(vx), self._f(vy), self._f(vz),
             self._d(dt, self.ngrids), int(steps), _stream())

    def normal_velocity(self, F, vx, vy, vz, out):
        call("vls_batch_normal_velocity" + self._sfx, self._h, self._f(F), self._f(vx), self._f(vy), self._f(vz),
             self._f(out), _stream())
        return out

    def area_volume(self, F, out=None, eps_factor=1.5):
        """(ngrids, 2) float64 device tensor of cosine area and volume"""
        if out is None:
            out = torch.empty((self.ngrids, 2), dtype=torch.float64, device=self.device)
        call("vls_batch_area_volume_cosine" + self._sfx, self._h, self._f(F), float(eps_factor),
             self._d(out, 2 * self.ngrids), _stream())
        return out

    def cfl_dt(self, vx, vy, vz, cfl, out=None):
        if out is None:
            out = torch.empty(self.ngrids, dtype=torch.float64, device=self.device)
        call("vls_batch_cfl_dt" + self._sfx, self._h, self._f(vx), self._f(vy), self._f(vz), float(cfl),
             self._d(out, self.ngrids), _stream())
        return out

    def curl_velocity(self, modes, lo, vx, vy, vz):
        """modes: (ngrids, nmodes, 7) float64 device tensor (k, k x a, psi); lo: (ngrids, 3)"""
        nm = modes.shape[1]
        call("vls_batch_curl_velocity" + self._sfx, self._h, self._d(modes, self.ngrids * nm * 7), int(nm),
             self._d(lo, 3 * self.ngrids), self._f(vx), self._f(vy), self._f(vz), _stream())

    def ellipsoid(self, axes, lo, F):
        call("vls_batch_ellipsoid" + self._sfx, self._h, self._d(axes, 3 * self.ngrids), self._d(lo, 3 * self.ngrids),
             self._f(F), _stream())

    def step(self, F, Vn, vx, vy, vz, dt, ext_iters, reinit_iters, av_row, eps_factor=1.5):
        call("vls_batch_step" + self._sfx, self._h, self._f(F), self._f(Vn), self._f(vx), self._f(vy), self._f(vz),
             self._d(dt, self.ngrids), int(ext_iters), int(reinit_iters), float(eps_factor),
             self._d(av_row, 2 * self.ngrids), _stream())


def set_graphs(enable):
    """turn hipGraph capture of the wrapper-level loops on or off; returns the previous setting"""
    from ._lib import lib

    return bool(lib().vls_set_graphs(1 if enable else 0))


# ---------------------------------------------------------------------------
# the batched vesicle workload
# ---------------------------------------------------------------------------
@dataclass
class BatchConfig:
    vesicles: int = 512
    edges: tuple = (32, 48, 64, 80, 96)
    axes_range: tuple = (0.3, 0.8)
    lo: float = -1.0
    hi: float = 1.0
    modes: int = 8
    amp: float = 0.1
    kmax: int = 4
    seed: int = 2018
    cfl: float = 0.4
    ext_iters: int = 40
    reinit_every: int = 5
    reinit_iters: int = 100
    initial_reinit: int = 20
    eps_factor: float = 1.5
    precision: str = "fp32"
    autotune: bool = True


def vesicle_modes(seed, modes, amp, kmax):
    """(modes, 7) rows k(3), k x a (3), psi of one vesicle's velocity, drawn as
    pipeline.curl_velocity draws them"""
    rng = np.random.default_rng(seed)
    rows = []
    while len(rows) < modes:
        k = rng.integers(-kmax, kmax + 1, size=3)
        if not np.any(k):
            continue
        k = k.astype(np.float64)
        a = rng.normal(0.0, amp, size=3)
        ph = rng.uniform(0.0, 2 * np.pi)
        rows.append(np.concatenate([k, np.cross(k, a), [ph]]))
    return np.array(rows)


def batch_inputs(cfg):
    """edges (n,), semi-axes (n, 3), mode tables (n, modes, 7) of the workload"""
    rng = np.random.default_rng(cfg.seed)
    edges = rng.choice(np.asarray(cfg.edges), size=cfg.vesicles)
    axes = rng.uniform(cfg.axes_range[0], cfg.axes_range[1], size=(cfg.vesicles, 3))
    modes = np.stack([vesicle_modes(cfg.seed + v, cfg.modes, cfg.amp, cfg.kmax) for v in range(cfg.vesicles)])
    return edges, axes, modes


def batch_updates_per_step(cfg, step_index, total_points):
    """WENO operator-stage passes x points of one step: 3 advection stages, the
    upwind normal, 3 per extension iteration, 3 per reinitialization iteration"""
    u = 3 + 1 + 3 * cfg.ext_iters
    if (step_index + 1) % cfg.reinit_every == 0:
        u += 3 * cfg.reinit_iters
    return u * total_points


@dataclass
class BatchPipeline:
    cfg: BatchConfig = field(default_factory=BatchConfig)
    device: str = "cuda"

    def __post_init__(self):
        cfg = self.cfg
        dtype = {"fp32": torch.float32, "fp64": torch.float64}[cfg.precision]
        edges, axes, modes = batch_inputs(cfg)
        self.edges, self.axes, self.modes = edges, axes, modes
        shapes = [(int(e), int(e), int(e)) for e in edges]
        spac = [((cfg.hi - cfg.lo) / (e - 1),) * 3 for e in edges]
        self.batch = b = Batch(shapes, spac, dtype=dtype, device=self.device)
        dev = b.device
        lo = torch.full((b.ngrids, 3), cfg.lo, dtype=torch.float64, device=dev)
        self.F = b.empty()
        b.ellipsoid(torch.from_numpy(axes).to(dev), lo, self.F)
        self.u = [b.empty() for _ in range(3)]
        b.curl_velocity(torch.from_numpy(modes).to(dev), lo, *self.u)
        b.set_distance(self.F)
        b.reinitialize(self.F, cfg.initial_reinit)
        self.tuned = (*b.autotune(self.F), b.zmarch_depth) if cfg.autotune else None
        self.dt = b.cfl_dt(*self.u, cfg.cfl)
        self.Vn = b.zeros()
        self._av = torch.zeros((b.ngrids, 2), dtype=torch.float64, device=b.device)  # the step's output row
        self.steps_done = 0
        self.area_volume = None

    def run(self, steps):
        """advance `steps` steps; returns the WENO updates done.  Every step is one
        captured graph (one for plain steps, one for reinitialization steps:
        the pointers they are keyed by do not change between steps)."""
        cfg, b = self.cfg, self.batch
        av = torch.zeros((steps, b.ngrids, 2), dtype=torch.float64, device=b.device)
        done = 0
        for s in range(steps):
            i = self.steps_done
            reinit = cfg.reinit_iters if (i + 1) % cfg.reinit_every == 0 else 0
            b.step(self.F, self.Vn, *self.u, self.dt, cfg.ext_iters, reinit, self._av, cfg.eps_factor)
            av[s].copy_(self._av)
            done += batch_updates_per_step(cfg, i, b.total)
            self.steps_done += 1
        self.area_volume = av
        return done

    def warm(self):
        """capture both step graphs (plain and reinitialization) and restore the state"""
        F0, steps = self.F.clone(), self.steps_done
        self.steps_done = self.cfg.reinit_every - 1
        self.run(1)
        self.run(1)
        self.F.copy_(F0)
        self.Vn.zero_()
        self.steps_done = steps


# ---------------------------------------------------------------------------
# the plain FP64 path on the same workload, and the comparison
# ---------------------------------------------------------------------------
def _grid_of(cfg, edge):
    from .sdf3 import GD3

    xv = np.linspace(cfg.lo, cfg.hi, int(edge))
    return GD3(xv, xv, xv)


def per_grid_fp64(cfg, steps, vesicles):
    """The plain FP64 GPU path on vesicles `vesicles` of the workload: one SDF3
    object per vesicle, its methods called vesicle after vesicle with eager
    launches (no batching, no graphs), as the reference's drivers hold one
    object per level set.  Returns (updates/s, seconds, final F per vesicle)."""
    import time

    from .pipeline import curl_velocity
    from .sdf3 import SDF3

    edges, axes, modes = batch_inputs(cfg)
    prev = set_graphs(False)
    try:
        objs = []
        for v in vesicles:
            G = _grid_of(cfg, edges[v])
            X, Y, Z = G.meshgrid()
            a, b, c = (float(t) for t in axes[v])
            F0 = torch.sqrt(X ** 2 / a ** 2 + Y ** 2 / b ** 2 + Z ** 2 / c ** 2) - 1.0
            m = SDF3(G, F0.contiguous())
            m.setDistance()
            m.F = m.WENO5RK3Reinitialization(m.F, cfg.initial_reinit)
            Xn, Yn, Zn = (t.cpu().numpy() for t in (X, Y, Z))
            u = curl_velocity(Xn, Yn, Zn, cfg.modes, cfg.amp, cfg.kmax, cfg.seed + v)
            u = [torch.from_numpy(np.ascontiguousarray(t)).to(G.device) for t in u]
            dt = cfg.cfl * G.Dx / m.maxSpeed(*u)
            objs.append((m, u, dt, G.zeros(), G))
        av = torch.zeros((steps, len(objs), 2), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        updates = 0
        for s in range(steps):
            reinit = (s + 1) % cfg.reinit_every == 0
            for k, (m, u, dt, Vn, G) in enumerate(objs):
                m.F = m.WENO5RK3Advection(m.F, *u, dt, 1)
                m.setDistance()
                m.normalVelocity(*u, Vn)
                Vn.copy_(m.WENO5RK3Extend(Vn, cfg.ext_iters))
                if reinit:
                    m.F = m.WENO5RK3Reinitialization(m.F, cfg.reinit_iters)
                m.areaVolumeDevice(av[s, k], 1, cfg.eps_factor * G.Dx)
                updates += batch_updates_per_step(cfg, s, G.NumElt)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    finally:
        set_graphs(prev)
    return updates / wall, wall, [o[0].F for o in objs]


def batch_comparison(cfg, steps=10, sample_every=16, log=None, conditioning=True):
    """Config 4: the batched workload in FP32 bulk (autotuned launch shape) and
    in FP64 (same batching), and the plain FP64 path (one object per vesicle)
    on every `sample_every`-th vesicle; rates in WENO updates/s, and the FP32
    run's area/volume drift and interface displacement against the FP64 run
    on every vesicle, both from the same initial state (the FP64 run's,
    rounded).  conditioning=True adds FP64 from that rounded initial state:
    its departure from the FP64 run is how far the problem itself carries a
    one-rounding change of the data, the floor under any FP32 result.
    log(msg), if given, is called as each part ends."""
    log = log or (lambda msg: None)
    import dataclasses
    import time

    from . import diagnostics

    out = {"vesicles": cfg.vesicles, "steps": steps, "edges": list(cfg.edges)}
    runs = {}
    kinds = [("fp64", "fp64"), ("fp32", "fp32")] + ([("fp64_rounded_start", "fp64")] if conditioning else [])
    for name, prec in kinds:
        c = dataclasses.replace(cfg, precision=prec)
        p = BatchPipeline(c)
        if name != "fp64":  # the FP64 run's initial state (rounded to FP32) and time steps
            p.F.copy_(runs["fp64"]["F0"].float().to(p.F.dtype))
            p.dt.copy_(runs["fp64"]["dt"])
        F0 = p.F.clone()
        p.warm()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        upd = p.run(steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        runs[name] = {"F0": F0, "F": p.F.clone(), "av": p.area_volume.clone(), "batch": p.batch,
                      "dt": p.dt.clone()}
        if name != "fp64_rounded_start":
            out[name + "_batched"] = {"updates_per_s": upd / wall, "wall_s": wall, "points": p.batch.total,
                                      "tile_rows": p.batch.tile_rows, "blocks": p.batch.blocks}
        del p
        log(f"{name} batched run done ({wall:.1f} s)")
    b64 = runs["fp64"]["batch"]
    sample = list(range(0, cfg.vesicles, sample_every))
    rate, wall, _ = per_grid_fp64(cfg, steps, sample)
    log(f"plain FP64 runs of {len(sample)} vesicles done ({wall:.1f} s)")
    out["fp64_per_vesicle"] = {"updates_per_s": rate, "wall_s": wall, "vesicles": len(sample)}
    out["speedup_fp32_batched_vs_fp64_per_vesicle"] = out["fp32_batched"]["updates_per_s"] / rate
    out["speedup_fp32_batched_vs_fp64_batched"] = (out["fp32_batched"]["updates_per_s"] /
                                                   out["fp64_batched"]["updates_per_s"])
    av64 = runs["fp64"]["av"]
    for name in [k for k, _ in kinds if k != "fp64"]:
        rel = ((runs[name]["av"] - av64).abs() / av64.abs()).amax(dim=0).cpu().numpy()
        disp = []
        for v in range(cfg.vesicles):
            G = _grid_of(cfg, b64.shapes[v][0])
            d = diagnostics._displacement_stats(G, b64.grid_view(runs[name]["F"], v).to(torch.float64),
                                                b64.grid_view(runs["fp64"]["F"], v))
            disp.append(d["normal"]["max_dx"])
        disp = np.asarray(disp)
        key = "errors_fp32_vs_fp64" if name == "fp32" else "fp64_rounded_start_vs_fp64"
        out[key] = {
            "area_rel_drift_max": float(rel[:, 0].max()),
            "volume_rel_drift_max": float(rel[:, 1].max()),
            "interface_displacement_normal_max_dx": float(disp.max()),
            "interface_displacement_normal_p99_dx": float(np.percentile(disp, 99)),
            "interface_displacement_normal_median_dx": float(np.median(disp)),
            "displacement_vesicles": len(disp),
        }
        log(f"{key} done")
    return out
