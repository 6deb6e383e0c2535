"""The single-phase vesicle relaxation step: the loop body of
src/7_0_paper/2_0_Vesicle/p00_test_singlephase_relax.m:60-196 on the GPU.

One step, in the driver's order:
  1. GPUsetCalculusToolBox of F (mean and Gaussian curvature, Towers delta/H);
  2. MeanCurvature = WENORK3Extend(MeanCurvature, ext_iters)          (:76)
  3. NormalSpeedBend = Kappa*(Lap H + 0.5 H(H^2-4G) + 2 SponC G - 0.5 SponC^2 H),
     then ENORK2Extend(., ext_iters)                                    (:79-84)
  4. area, volume, Dt = CFLNumber*Ds/max|Vb| on |F|<2Ds, the 2x2 tension/
     pressure solve and normalSpeed = (T H - Vb + P)|grad F|           (:67-109)
  5. smoothGMRES(normalSpeed, Dt, 0.5), then ENORK2Extend(., ext_iters) (:110-112)
  6. F = F - Dt*speed; setDistance                                      (:114-115)
  7. every reinit_every steps: WENO5RK3Reinitialization(F, reinit_iters)  (:191-193)
The energy surfaceIntegral((H - SponC)^2) of :118 is kept on the device per
step.  Dt is read back once per step (the smoothing and the update need it
as a scalar), as the driver's gather does.

The toolbox writes into buffers allocated once; every operator is a HIP
kernel of the library (the FFT through rocFFT).

BatchRelax steps several vesicles together: the three extensions, setDistance
and the reinitialization, which take almost all of a step's time, run as one
batched launch sequence over every grid, and the rest (toolbox, bending
speed, tension/pressure, smoothing, update) per grid on views of the packed
fields, with one read-back of all time steps per step.  Each vesicle's
results are bit for bit those of its own VesicleRelax.
"""
from dataclasses import dataclass

import torch

from ._lib import VLSError
from .batch import Batch
from .sdf3 import GD3, SDF3, TOOLBOX_FIELDS, feval_set_calculus_toolbox


@dataclass
class RelaxConfig:
    cfl: float = 0.5          # CFLNumber of the 96^3 run (p00:7)
    kappa: float = 1.0        # bending modulus (:46)
    sponc: float = 0.0        # spontaneous curvature (:12)
    alpha: float = 0.5        # smoothGMRES(normalSpeed, Dt, 0.5) (:110)
    ext_iters: int = 100      # WENORK3Extend / ENORK2Extend iterations (:76, :84, :111)
    reinit_every: int = 1     # ReinitializationRate of the 96^3 run (:7)
    reinit_iters: int = 100   # WENO5RK3Reinitialization(map.F, 100) (:192)
    autotune: bool = True     # pick the stage kernels' tile shape and z-march depth for this grid once


def oblate_ellipsoid(n, reduced_volume=0.6, ratio=0.35):
    """+SD/+Shape/Ellipsoid.m:5-44 for TYPE 'O': xv = linspace(-1,1,n), yv = zv =
    ((1-n)/2 : (n-1)/2)*dx, a = b = ratio*2, c = a*cq with cq interpolated
    (interp1, linear) in the table of reduced volumes of the spheroids with
    half axes (1,1,c), c = 0.01:0.01:0.99 (the oblate area is closed-form).
    Returns (xv, yv, zv, F) with F of shape (n, n, n) = (pges, cols, rows)."""
    import numpy as np

    xv = np.linspace(-1.0, 1.0, n)
    dx = xv[1] - xv[0]
    yv = (np.arange(n) + (1 - n) / 2.0) * dx
    zv = yv.copy()
    cs = np.arange(1, 100) / 100.0
    e = np.sqrt(1 - cs ** 2)
    area = 2 * np.pi * (1 + cs ** 2 / e * np.arctanh(e))
    vol = 4 * np.pi * cs / 3
    rv = (3 * vol / 4 / np.pi) * (4 * np.pi / area) ** 1.5
    cq = float(np.interp(reduced_volume, rv, cs))
    a = ratio * 2.0
    c = a * cq
    Z, X, Y = np.meshgrid(zv, xv, yv, indexing="ij")
    F = np.sqrt(X ** 2 / a ** 2 + Y ** 2 / a ** 2 + Z ** 2 / c ** 2) - 1
    return xv, yv, zv, np.ascontiguousarray(F)


class VesicleRelax:
    """State of the driver: the SDF3 object, the initial area and volume, time."""

    def __init__(self, grid: GD3, F: torch.Tensor, cfg: RelaxConfig = RelaxConfig()):
        if F.dtype != torch.float64:
            raise VLSError("the relaxation step is FP64")
        self.cfg = cfg
        self.G = grid
        self.map = SDF3(grid, F.contiguous())
        if cfg.autotune:
            self.map.autotune(iterations=5)  # leaves the distances of F set, as setDistance does
        else:
            self.map.setDistance()
        self._tool = [grid.zeros() for _ in TOOLBOX_FIELDS]
        self._hp = grid.zeros()
        self._toolbox()
        self.InitialArea, self.InitialVolume = self.map.calAreaVolume()
        self.time = 0.0
        self.iteration = 1
        self.info = None
        self.energy = []

    def _toolbox(self):
        m, g = self.map, self.G
        torch.clamp(m.F, min=0.0, out=self._hp)
        feval_set_calculus_toolbox(self._tool, m.F, self._hp, g.mrows, g.ncols, g.lshts, g.Dx, g.Dy, g.Dz, g.Ds)
        for name, t in zip(TOOLBOX_FIELDS, self._tool):
            setattr(m, name, t)
        m.HPrimal = self._hp

    def step(self):
        cfg, m, g = self.cfg, self.map, self.G
        self.iteration += 1
        self._toolbox()
        H = m.WENORK3Extend(m.MeanCurvature, cfg.ext_iters)
        Vb = m.ENORK2Extend(g.bendingSpeed(H, m.GaussianCurvature, cfg.kappa, cfg.sponc), cfg.ext_iters)
        speed, info = m.tensionPressure(H, Vb, self.InitialArea, self.InitialVolume, CFLNumber=cfg.cfl)
        dt = float(info[2])  # the one read-back of the step
        g.smoothGMRES(speed, dt, cfg.alpha, out=speed)
        speed = m.ENORK2Extend(speed, cfg.ext_iters)
        m.F = m.F - dt * speed
        m.setDistance()
        band = m.F.abs() < 2 * g.Ds
        hs = H - cfg.sponc
        self.energy.append((hs * hs * m.DiracDelta * m.FGradMag)[band].sum() * g.Ds ** 3)
        if cfg.reinit_every > 0 and self.iteration % cfg.reinit_every == 0:
            m.F = m.WENO5RK3Reinitialization(m.F, cfg.reinit_iters)
        self.time += dt
        self.info = info
        return info


class BatchRelax:
    """VesicleRelax for several vesicles on grids of any sizes, stepped together.

    F holds every grid's level set packed (Batch layout); `grid_view(g)` is
    vesicle g's (pges, cols, rows) view of it.  Per step, in the order of
    VesicleRelax.step: the toolbox of every grid, one batched WENORK3Extend
    of the mean curvatures, the bending speeds, one batched ENORK2Extend, the
    tension/pressure solves, one read-back of the time steps, the smoothings,
    one batched ENORK2Extend of the speeds, the updates, one batched
    setDistance and one batched WENO5RK3Reinitialization."""

    def __init__(self, grids, Fs, cfg: RelaxConfig = RelaxConfig()):
        if len(grids) == 0 or len(grids) != len(Fs):
            raise VLSError("one level set per grid, at least one grid")
        if any(F.dtype != torch.float64 for F in Fs):
            raise VLSError("the relaxation step is FP64")
        self.cfg = cfg
        self.grids = list(grids)
        self.B = Batch([(g.mrows, g.ncols, g.lshts) for g in grids], [(g.Dx, g.Dy, g.Dz) for g in grids],
                       dtype=torch.float64, device=grids[0].device)
        self.F = self.B.pack([F.contiguous() for F in Fs])
        if cfg.autotune:
            self.B.autotune(self.F, iterations=5)  # times copies; F is left as it is
        # the tension/pressure reductions and the initial area and volume run
        # on an SDF3 handle; its work memory (about 450 bytes per point) is
        # shared by every vesicle of one grid geometry, pointed at each
        # vesicle's views in turn (its calls are ordered on the stream)
        self._maps = {}
        for g in grids:
            key = (g.mrows, g.ncols, g.lshts, g.Dx, g.Dy, g.Dz)
            if key not in self._maps:
                self._maps[key] = SDF3(g, g.zeros())
        self._tool = [self.B.zeros() for _ in TOOLBOX_FIELDS]
        self._hp = self.B.zeros()
        self._H = self.B.zeros()
        self._Vb = self.B.zeros()
        self._speed = self.B.zeros()
        self._toolbox()
        av = [self._map(i).calAreaVolume() for i in range(len(grids))]
        self.InitialArea = [a for a, _ in av]
        self.InitialVolume = [v for _, v in av]
        self.time = [0.0] * len(grids)
        self.iteration = 1
        self.info = None
        self.energy = []

    def grid_view(self, g, packed=None):
        return self.B.grid_view(self.F if packed is None else packed, g)

    def _map(self, i):
        """the shared SDF3 of vesicle i's geometry, pointed at its level set and toolbox"""
        g = self.grids[i]
        m = self._maps[(g.mrows, g.ncols, g.lshts, g.Dx, g.Dy, g.Dz)]
        m.F = self.grid_view(i)
        for name, t in zip(TOOLBOX_FIELDS, self._tool):
            setattr(m, name, self.grid_view(i, t))
        m.HPrimal = self.grid_view(i, self._hp)
        return m

    def _toolbox(self):
        torch.clamp(self.F, min=0.0, out=self._hp)
        for i, g in enumerate(self.grids):
            tool = [self.grid_view(i, t) for t in self._tool]
            feval_set_calculus_toolbox(tool, self.grid_view(i), self.grid_view(i, self._hp), g.mrows, g.ncols,
                                       g.lshts, g.Dx, g.Dy, g.Dz, g.Ds)

    def step(self):
        """one relaxation step of every vesicle; returns the (ngrids, 10) device
        tensor of their tension/pressure scalars (TENSION_PRESSURE_FIELDS)"""
        cfg, B = self.cfg, self.B
        gauss = TOOLBOX_FIELDS.index("GaussianCurvature")
        delta, gmag = TOOLBOX_FIELDS.index("DiracDelta"), TOOLBOX_FIELDS.index("FGradMag")
        self.iteration += 1
        self._toolbox()
        self._H.copy_(self._tool[TOOLBOX_FIELDS.index("MeanCurvature")])
        B.wenork3_extend(self.F, self._H, cfg.ext_iters)
        for i, g in enumerate(self.grids):
            g.bendingSpeed(self.grid_view(i, self._H), self.grid_view(i, self._tool[gauss]), cfg.kappa, cfg.sponc,
                           out=self.grid_view(i, self._Vb))
        B.enork2_extend(self.F, self._Vb, cfg.ext_iters)
        self._speed.zero_()
        infos = []
        for i in range(len(self.grids)):
            _, info = self._map(i).tensionPressure(self.grid_view(i, self._H), self.grid_view(i, self._Vb),
                                                   self.InitialArea[i], self.InitialVolume[i], CFLNumber=cfg.cfl,
                                                   out=self.grid_view(i, self._speed))
            infos.append(info)
        info = torch.stack(infos)
        dts = info[:, 2].tolist()  # the one read-back of the step
        for i, g in enumerate(self.grids):
            sp = self.grid_view(i, self._speed)
            g.smoothGMRES(sp, dts[i], cfg.alpha, out=sp)
        B.enork2_extend(self.F, self._speed, cfg.ext_iters)
        for i in range(len(self.grids)):
            Fi = self.grid_view(i)
            Fi.copy_(Fi - dts[i] * self.grid_view(i, self._speed))
        B.set_distance(self.F)
        energy = []
        for i, g in enumerate(self.grids):
            band = self.grid_view(i).abs() < 2 * g.Ds
            hs = self.grid_view(i, self._H) - cfg.sponc
            e = hs * hs * self.grid_view(i, self._tool[delta]) * self.grid_view(i, self._tool[gmag])
            energy.append(e[band].sum() * g.Ds ** 3)
        self.energy.append(energy)
        if cfg.reinit_every > 0 and self.iteration % cfg.reinit_every == 0:
            B.reinitialize(self.F, cfg.reinit_iters)
        self.time = [t + dt for t, dt in zip(self.time, dts)]
        self.info = info
        return info
