"""The full vesicle-step pipeline at one grid: the workload the benchmark runs.

One step advances an ellipsoidal vesicle by a seeded divergence-free velocity
and runs every operator of the level-set path once in the order the
reference's drivers use them (src/7_0_paper/2_0_Vesicle/p00_test_singlephase_relax.m:60-196):

  1. WENO5 + TVD-RK3 advection of F and of an auxiliary field psi (3 stages each);
  2. setDistance(F), normal velocity u.N and its WENO5-RK3 extension off the
     interface (`ext_iters` iterations);
  3. every `reinit_every` steps: WENO5-RK3 reinitialization of F
     (`reinit_iters` iterations) and closest-point surface redistance of psi
     (`redist_iters` = (iter1, iter2));
  4. cosine-smoothed delta/H area and volume (half-width 1.5*dx), kept on the
     device, one row per step.

Synthetic inputs (all seeded, float64):
  * box [lo,hi]^3 with n points per axis (MATLAB linspace/meshgrid);
  * F0 = sqrt(x^2/a^2 + y^2/b^2 + z^2/c^2) - 1 (the form of
    +SD/+Shape/Ellipsoid.m:45) reinitialized `initial_reinit` times;
  * u = curl A with A = sum_m a_m cos(k_m.x + psi_m): u = sum_m -sin(k_m.x + psi_m) (k_m x a_m),
    k_m integer vectors with components in [-kmax, kmax] (not all 0),
    a_m ~ N(0, amp^2)^3, psi_m ~ U[0, 2pi), drawn by numpy.random.default_rng(seed);
  * psi0 = x - 0.2 cos(2 pi y);
  * dt = cfl * dx / max|u| (the CFL convention of p00_test_singlephase_relax.m:89).

precision "fp32" runs the same step on float32 fields through the FP32-bulk
entry points; the initial reinitialization, max|u| and dt are computed in
FP64 first, so both precisions start from the same state (rounded once).
"""
from dataclasses import dataclass, field

import numpy as np
import torch

from .sdf3 import GD3, SDF3


@dataclass
class PipelineConfig:
    n: int = 128
    lo: float = -1.2
    hi: float = 1.2
    axes: tuple = (0.8, 0.5, 0.5)
    modes: int = 16
    amp: float = 0.1
    kmax: int = 4
    seed: int = 2018
    cfl: float = 0.4
    ext_iters: int = 40
    reinit_every: int = 5
    reinit_iters: int = 100
    redist_iters: tuple = (10, 5)
    initial_reinit: int = 100
    eps_factor: float = 1.5
    precision: str = "fp64"  # or "fp32" (FP32 bulk)
    reinit_fp64_tail: int = 0  # fp32: the last iterations of each reinitialization call run in FP64
    autotune: bool = True  # pick the stage kernels' launch shape for the grid at set-up (results unchanged)


def curl_velocity(X, Y, Z, modes, amp, kmax, seed):
    """u = curl of sum_m a_m cos(k_m.x + psi_m), evaluated per point (numpy, float64)."""
    rng = np.random.default_rng(seed)
    ks, avs, phs = [], [], []
    while len(ks) < modes:
        k = rng.integers(-kmax, kmax + 1, size=3)
        if not np.any(k):
            continue
        ks.append(k.astype(np.float64))
        avs.append(rng.normal(0.0, amp, size=3))
        phs.append(rng.uniform(0.0, 2 * np.pi))
    u = np.zeros_like(X)
    v = np.zeros_like(X)
    w = np.zeros_like(X)
    for k, a, ph in zip(ks, avs, phs):
        s = -np.sin(k[0] * X + k[1] * Y + k[2] * Z + ph)
        kxa = np.cross(k, a)
        u += s * kxa[0]
        v += s * kxa[1]
        w += s * kxa[2]
    return u, v, w


def make_inputs(cfg):
    """Host (numpy) inputs of the workload: grid vectors, F0, psi0, u."""
    xv = np.linspace(cfg.lo, cfg.hi, cfg.n)
    Z, X, Y = np.meshgrid(xv, xv, xv, indexing="ij")
    a, b, c = cfg.axes
    F0 = np.sqrt(X ** 2 / a ** 2 + Y ** 2 / b ** 2 + Z ** 2 / c ** 2) - 1.0
    psi0 = X - 0.2 * np.cos(2 * np.pi * Y)
    u = curl_velocity(X, Y, Z, cfg.modes, cfg.amp, cfg.kmax, cfg.seed)
    return xv, F0, psi0, u


def updates_per_step(cfg, step_index):
    """WENO operator-stage evaluations of step `step_index` (0-based), in units of
    full-box passes: advection stages, upwind normals, extension stages and
    reinitialization stages.  Distance location, interface interpolation,
    normal velocity and the area/volume sums run every step but are not counted."""
    u = 6 + 1 + 3 * cfg.ext_iters
    if (step_index + 1) % cfg.reinit_every == 0:
        i1, i2 = cfg.redist_iters
        u += 3 * cfg.reinit_iters
        u += 1 + i1 * (3 * 3 * i2 + 3) + 3 * i2
    return u


@dataclass
class Pipeline:
    cfg: PipelineConfig = field(default_factory=PipelineConfig)
    device: str = "cuda"

    def __post_init__(self):
        cfg = self.cfg
        xv, F0, psi0, u = make_inputs(cfg)
        dev = torch.device(self.device)
        self.grid = GD3(xv, xv, xv, device=dev)
        T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
        self.u = [T(c) for c in u]
        self.psi = T(psi0)
        self.map = SDF3(self.grid, T(F0))
        self.map.setDistance()
        self.map.F = self.map.WENO5RK3Reinitialization(self.map.F, cfg.initial_reinit)
        self.max_speed = self.map.maxSpeed(*self.u)
        self.dt = cfg.cfl * self.grid.Dx / self.max_speed
        self.eps = cfg.eps_factor * self.grid.Dx
        if cfg.precision == "fp32":
            F = self.map.F.float()
            del self.map
            self.map = SDF3(self.grid, F)
            self.u = [c.float() for c in self.u]
            self.psi = self.psi.float()
            self.map.setReinitFP64Tail(cfg.reinit_fp64_tail)
        elif cfg.precision != "fp64":
            raise ValueError(f"precision must be fp64 or fp32, got {cfg.precision}")
        if cfg.autotune:
            self.map.autotune(iterations=5)  # 5 timed iterations per candidate: a steadier choice
        self.Vn = self.grid.zeros(self.map.dtype)
        self._av = torch.zeros(2, dtype=torch.float64, device=dev)  # the step's area/volume output
        self.area_volume = None
        self.steps_done = 0
        self.ext_events = []  # (start, end) event pairs around the extension, when recording

    def _normal_velocity(self, out):
        self.map.normalVelocity(*self.u, out)

    def _area_volume(self, out_row):
        self.map.areaVolumeDevice(out_row, 1, self.eps)

    def run(self, steps, record_extension=False):
        """Advance `steps` steps; returns the WENO updates done (full-box passes x points)."""
        cfg, m = self.cfg, self.map
        n = self.grid.NumElt
        av = torch.zeros((steps, 2), dtype=torch.float64, device=self.grid.device)
        done = 0
        for s in range(steps):
            i = self.steps_done
            m.advect_(m.F, *self.u, self.dt, 1)
            m.advect_(self.psi, *self.u, self.dt, 1)
            m.setDistance()
            self._normal_velocity(self.Vn)
            if record_extension:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
            m.extend_(self.Vn, cfg.ext_iters)
            if record_extension:
                e1.record()
                self.ext_events.append((e0, e1))
            if (i + 1) % cfg.reinit_every == 0:
                m.reinitialize_(m.F, cfg.reinit_iters)
                m.redistance_(self.psi, *cfg.redist_iters)
            self._area_volume(self._av)
            av[s].copy_(self._av)
            done += updates_per_step(cfg, i) * n
            self.steps_done += 1
        self.area_volume = av
        return done
