"""Shape matching through the discrete adjoint: gradient of an interface
mismatch after many advection steps with respect to the velocity's Fourier
coefficients.

Workload (one grid, FP64): n^3 points on [-1,1]^3; F0 = |x| - 0.6; velocity
u_a(x) = sum_m c[a,m] b_m(x) with 64 fixed real Fourier modes
b_m(x) = cos(pi k_m.x + theta_m) (k_m integer vectors in [-2,2]^3 \\ {0},
theta_m ~ U[0,2pi), drawn with numpy.random.default_rng(seed)) and
coefficients c ~ N(0, coef_std^2) (seeded); `steps` WENO5 + TVD-RK3 advection
steps with dt = cfl*dx/max|u| fixed from the initial coefficients; after every
`reinit_every`-th step, setDistance and `reinit_iters` WENO5-RK3
reinitialization iterations.  Loss
    L(c) = sum_i (H_eps(F_T,i) - H_eps(F_target,i))^2 dV,
with H_eps the cosine-smoothed Heaviside of half-width eps = 1.5 dx
(src/0_0_setup/+SD/@SDF3/SDF3.m:126-136) and F_target the exact signed
distance to an ellipsoid with semi-axes (0.7, 0.55, 0.45) in a seeded
orientation (closest-point parameter found by bisection in FP64).

The gradient comes from the hand-written adjoint kernels of the library
(vls_sdf3_adjoint_*): a forward pass stores F every `ckpt_every` steps; the
backward pass recomputes each segment from its checkpoint and sweeps back
through it.  `jvp` gives directional derivatives with the tangent kernels.
"""
from dataclasses import dataclass

import numpy as np
import torch

from ._lib import P, call
from .sdf3 import GD3, SDF3, _stream


def _p(t):
    return P(t.data_ptr())


@dataclass
class AdjointConfig:
    n: int = 96
    steps: int = 200
    reinit_every: int = 10
    reinit_iters: int = 10
    ckpt_every: int = 20
    modes: int = 64
    coef_std: float = 0.05
    cfl: float = 0.4
    seed: int = 2018
    target_axes: tuple = (0.7, 0.55, 0.45)
    radius0: float = 0.6
    eps_factor: float = 1.5


def random_rotation(rng):
    """a rotation matrix from a seeded normal 3x3 (QR with a positive diagonal)"""
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))[None, :]
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def ellipsoid_sdf(X, Y, Z, axes, R, iters=200):
    """signed distance to the ellipsoid {|R^T x / axes| = 1} (torch, FP64):
    closest point y_i a_i^2/(t + a_i^2) with t the root of
    sum_i (a_i y_i/(t + a_i^2))^2 = 1, by bisection"""
    Rt = torch.as_tensor(R, dtype=torch.float64, device=X.device)
    y = [Rt[0, k] * X + Rt[1, k] * Y + Rt[2, k] * Z for k in range(3)]  # ellipsoid frame
    a2 = [float(a) ** 2 for a in axes]
    f0 = sum(yk * yk / ak for yk, ak in zip(y, a2))
    inside = f0 < 1.0
    rmag = torch.sqrt(y[0] ** 2 + y[1] ** 2 + y[2] ** 2)
    amin2 = min(a2)
    lo = torch.where(inside, torch.full_like(rmag, -amin2), torch.zeros_like(rmag))
    hi = torch.where(inside, torch.zeros_like(rmag), rmag * max(axes) + 1e-300)

    def G(t):
        return sum((float(np.sqrt(ak)) * yk / (t + ak)) ** 2 for yk, ak in zip(y, a2)) - 1.0

    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        g = G(mid)
        lo = torch.where(g > 0, mid, lo)
        hi = torch.where(g > 0, hi, mid)
    t = 0.5 * (lo + hi)
    xc = [yk * ak / (t + ak) for yk, ak in zip(y, a2)]
    d = torch.sqrt(sum((yk - xk) ** 2 for yk, xk in zip(y, xc)))
    return torch.where(inside, -d, d)


def heaviside(F, eps):
    """cosine-smoothed Heaviside and its derivative"""
    H = torch.where(F > eps, torch.ones_like(F), torch.zeros_like(F))
    band = F.abs() <= eps
    Hb = (1 + F / eps + torch.sin(np.pi * F / eps) / np.pi) / 2
    dH = torch.where(band, (1 + torch.cos(np.pi * F / eps)) / (2 * eps), torch.zeros_like(F))
    return torch.where(band, Hb, H), dH


class ShapeMatching:
    def __init__(self, cfg: AdjointConfig = AdjointConfig(), device="cuda"):
        self.cfg = cfg
        rng = np.random.default_rng(cfg.seed)
        ks = []
        while len(ks) < cfg.modes:
            k = rng.integers(-2, 3, size=3)
            if np.any(k):
                ks.append(k)
        self.k = np.array(ks, dtype=np.float64)
        self.theta = rng.uniform(0.0, 2 * np.pi, size=cfg.modes)
        self.c0 = rng.normal(0.0, cfg.coef_std, size=(3, cfg.modes))
        self.R = random_rotation(rng)
        xv = np.linspace(-1.0, 1.0, cfg.n)
        self.grid = G = GD3(xv, xv, xv, device=device)
        X, Y, Z = G.meshgrid()
        self.dV = G.Dx * G.Dy * G.Dz
        self.eps = cfg.eps_factor * G.Dx
        # basis (modes, n): b_m(x) = cos(pi k_m.x + theta_m)
        kt = torch.as_tensor(self.k, device=X.device)
        th = torch.as_tensor(self.theta, device=X.device)
        flat = torch.stack([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)])  # (3, n)
        self.B = torch.cos(np.pi * (kt @ flat) + th[:, None]).contiguous()
        self.F0 = (torch.sqrt(X ** 2 + Y ** 2 + Z ** 2) - cfg.radius0).contiguous()
        self.Ft = ellipsoid_sdf(X, Y, Z, cfg.target_axes, self.R).contiguous()
        self.Ht, _ = heaviside(self.Ft, self.eps)
        self.map = SDF3(G, self.F0.clone())
        self.u = [G.zeros() for _ in range(3)]
        self._set_velocity(self.c0)
        self.dt = cfg.cfl * G.Dx / self.map.maxSpeed(*self.u)
        self.F = G.zeros()

    # -- pieces ----------------------------------------------------------------
    def velocity(self, c):
        """u = c @ B as three (pges, cols, rows) fields"""
        c = torch.as_tensor(c, dtype=torch.float64, device=self.B.device)
        U = c @ self.B
        return [U[a].reshape(self.grid.Size).contiguous() for a in range(3)]

    def _set_velocity(self, c):
        for dst, src in zip(self.u, self.velocity(c)):
            dst.copy_(src)

    def loss_and_seed(self, F):
        """L and dL/dF"""
        H, dH = heaviside(F, self.eps)
        r = H - self.Ht
        return float((r * r).sum()) * self.dV, (2.0 * self.dV) * r * dH

    def _reinit_step(self, i):
        return (i + 1) % self.cfg.reinit_every == 0

    def _advance(self, F, i):
        """step i in place: advection, then the reinitialization when due"""
        m = self.map
        m.advect_(F, *self.u, self.dt, 1)
        if self._reinit_step(i):
            m.F = F
            m.setDistance()
            m.reinitialize_(F, self.cfg.reinit_iters)

    # -- forward, gradient, directional derivative ----------------------------------
    def loss(self, c):
        self._set_velocity(c)
        F = self.F
        F.copy_(self.F0)
        for i in range(self.cfg.steps):
            self._advance(F, i)
        return self.loss_and_seed(F)[0]

    def loss_with_signature(self, c):
        """L(c) and the branch signature of the forward: a hash of the sign
        pattern of F after every advection step and after every
        reinitialization.  Every jump of the discrete map sits where that
        pattern changes: a subcell crossing appearing or vanishing, the
        quadratic-ENO / quintic branch of the root solve (a kink is two sign
        changes within a stencil), the inside/outside mask, the uniform /
        nonuniform WENO switch of the reinitialization (a sign change beside
        a node).  Two coefficient vectors with equal signatures lie on one
        smooth piece of the map, up to the sign tests inside the
        reinitialization's RK stages, which follow the same sign pattern."""
        if not hasattr(self, "_hash_w"):
            g = torch.Generator(device="cpu").manual_seed(20181207)
            self._hash_w = torch.randint(0, 1 << 40, self.grid.Size, generator=g, dtype=torch.int64).to(
                self.F0.device)
        self._set_velocity(c)
        F = self.F
        F.copy_(self.F0)
        n = self.cfg.steps + sum(1 for i in range(self.cfg.steps) if self._reinit_step(i))
        sig = torch.zeros(n, dtype=torch.int64, device=F.device)
        k = 0
        m = self.map
        for i in range(self.cfg.steps):
            m.advect_(F, *self.u, self.dt, 1)
            sig[k] = torch.sum(self._hash_w * (F < 0))
            k += 1
            if self._reinit_step(i):
                m.F = F
                m.setDistance()
                m.reinitialize_(F, self.cfg.reinit_iters)
                sig[k] = torch.sum(self._hash_w * (F < 0))
                k += 1
        return self.loss_and_seed(F)[0], sig.cpu().numpy()

    def gradient(self, c):
        """(L, dL/dc (3, modes)) by the adjoint with checkpointing"""
        cfg, m = self.cfg, self.map
        self._set_velocity(c)
        F = self.F
        F.copy_(self.F0)
        ckpt = {}
        for i in range(cfg.steps):
            if i % cfg.ckpt_every == 0:
                ckpt[i] = F.clone()
            self._advance(F, i)
        L, lam = self.loss_and_seed(F)
        lam = lam.contiguous()
        gu = [self.grid.zeros() for _ in range(3)]
        W = self.grid.zeros()
        for s0 in sorted(ckpt, reverse=True):
            s1 = min(s0 + cfg.ckpt_every, cfg.steps)
            # recompute the segment, keeping the field before each advection
            # and before each reinitialization
            W.copy_(ckpt[s0])
            before_adv, before_re = {}, {}
            for i in range(s0, s1):
                before_adv[i] = W.clone()
                m.advect_(W, *self.u, self.dt, 1)
                if self._reinit_step(i):
                    before_re[i] = W.clone()
                    m.F = W
                    m.setDistance()
                    m.reinitialize_(W, cfg.reinit_iters)
            for i in range(s1 - 1, s0 - 1, -1):
                if i in before_re:
                    call("vls_sdf3_adjoint_reinit", m._h, _p(before_re[i]), cfg.reinit_iters, _p(lam), _p(lam),
                         _stream())
                call("vls_sdf3_adjoint_advect_rk3", m._h, _p(before_adv[i]), *[_p(u) for u in self.u],
                     float(self.dt), _p(lam), _p(lam), *[_p(g) for g in gu], _stream())
        G = torch.stack([g.reshape(-1) for g in gu])  # (3, n)
        return L, (G @ self.B.T).cpu().numpy()

    def jvp(self, c, cdot):
        """dL/dc . cdot with the tangent kernels"""
        cfg, m = self.cfg, self.map
        self._set_velocity(c)
        ud = self.velocity(cdot)
        F = self.F0.clone()
        Fd = self.grid.zeros()
        F2, Fd2 = self.grid.zeros(), self.grid.zeros()
        for i in range(cfg.steps):
            call("vls_sdf3_tangent_advect_rk3", m._h, _p(F), _p(Fd), *[_p(u) for u in self.u],
                 *[_p(u) for u in ud], float(self.dt), _p(F2), _p(Fd2), _stream())
            F, F2 = F2, F
            Fd, Fd2 = Fd2, Fd
            if self._reinit_step(i):
                call("vls_sdf3_tangent_reinit", m._h, _p(F), _p(Fd), cfg.reinit_iters, _p(F2), _p(Fd2), _stream())
                F, F2 = F2, F
                Fd, Fd2 = Fd2, Fd
        L, seed = self.loss_and_seed(F)
        return L, float((seed * Fd).sum())


def classified_gradient_check(prob, h=1e-6):
    """The adjoint gradient against FP64 central differences with step h on
    every coefficient, each coefficient classified by the branch signatures
    (ShapeMatching.loss_with_signature) at c - h e_k, c and c + h e_k:
    `crossing` marks the coefficients whose perturbation changes the
    signature, i.e. whose difference quotient spans a jump of the discrete
    map.  Returns the gradient, the central, forward and backward quotients
    and the mask."""
    c0 = prob.c0
    L0, s0 = prob.loss_with_signature(c0)
    L, g = prob.gradient(c0)
    fd = np.zeros_like(g)
    fw = np.zeros_like(g)
    bw = np.zeros_like(g)
    crossing = np.zeros(g.shape, dtype=bool)
    for a in range(g.shape[0]):
        for k in range(g.shape[1]):
            cp, cm = c0.copy(), c0.copy()
            cp[a, k] += h
            cm[a, k] -= h
            Lp, sp = prob.loss_with_signature(cp)
            Lm, sm = prob.loss_with_signature(cm)
            fd[a, k] = (Lp - Lm) / (2 * h)
            fw[a, k] = (Lp - L0) / h
            bw[a, k] = (L0 - Lm) / h
            crossing[a, k] = not (np.array_equal(sp, s0) and np.array_equal(sm, s0))
    return {"loss": L0, "loss_gradient_pass": L, "gradient": g, "fd": fd, "forward": fw, "backward": bw,
            "crossing": crossing}


def check_gradient(prob, hs=(1e-6,), directions=8, seed=7):
    """The adjoint gradient at the problem's coefficients against FP64 central
    differences on every coefficient, for each step in `hs`, and against the
    tangent-mode directional derivative on `directions` random directions;
    with timings.

    The discrete map is piecewise smooth: a node value crossing zero or a
    sign-change test flipping (the uniform/nonuniform WENO switch of the
    reinitialization, a subcell crossing appearing, the ENO/quintic branch of
    the root solve, the inside/outside mask) is a jump.  Central differences
    whose step reaches such a threshold measure the jump, not the
    derivative; `one_sided_mismatch` (the largest relative difference
    between the forward and backward quotients) shows when that happens."""
    import time

    c0 = prob.c0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L = prob.loss(c0)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    L2, g = prob.gradient(c0)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    out = {"loss": L, "loss_gradient_pass": L2, "forward_s": t1 - t0, "gradient_s": t2 - t1,
           "gradient_over_forward": (t2 - t1) / (t1 - t0), "fd": []}
    for h in hs:
        ta = time.perf_counter()
        fd = np.zeros_like(g)
        fw = np.zeros_like(g)
        bw = np.zeros_like(g)
        for a in range(g.shape[0]):
            for k in range(g.shape[1]):
                cp, cm = c0.copy(), c0.copy()
                cp[a, k] += h
                cm[a, k] -= h
                Lp, Lm = prob.loss(cp), prob.loss(cm)
                fd[a, k] = (Lp - Lm) / (2 * h)
                fw[a, k] = (Lp - L) / h
                bw[a, k] = (L - Lm) / h
        big = np.abs(fd) > 1e-3 * np.abs(fd).max()
        out["fd"].append({
            "h": h,
            "evaluations": 2 * g.size,
            "seconds": time.perf_counter() - ta,
            "grad_vs_fd_rel_norm": float(np.linalg.norm(g - fd) / np.linalg.norm(fd)),
            "grad_vs_fd_rel_max": float((np.abs(g - fd)[big] / np.abs(fd)[big]).max()),
            "one_sided_mismatch": float((np.abs(fw - bw)[big] / np.abs(fd)[big]).max()),
        })
    t3 = time.perf_counter()
    rng = np.random.default_rng(seed)
    jrel = []
    for _ in range(directions):
        d = rng.normal(size=g.shape)
        _, dj = prob.jvp(c0, d)
        jrel.append(abs(dj - float((g * d).sum())) / abs(dj))
    out.update({"jvp_directions": directions, "grad_vs_jvp_rel_max": float(max(jrel)),
                "jvp_s": time.perf_counter() - t3})
    return out
