"""CPU oracle for the operators of the vesicle relaxation step that follow the
level-set path.

TEST INFRASTRUCTURE ONLY.  Loaded by tests/ and by nothing in the product; the
HIP library never calls it.

numpy restatement of (paths relative to src/7_0_paper/2_0_Vesicle/):
  - @GD3's sparse periodic operators and wavenumbers (+SD/@GD3/GD3.m:110-207);
  - smoothFFT / smoothDiffusionFFT (GD3.m:453-458, 479-484);
  - smoothGMRES / smoothDiffusionGMRES (GD3.m:435-451, 461-477): MATLAB's
    gmres(Op, S, [], 1e-12, 300, @mfun) -- left-preconditioned GMRES from a
    zero start, no restart, stopping when the preconditioned relative residual
    is at most tol -- restated with modified Gram-Schmidt Arnoldi and Givens
    rotations; and the exact solution of the same circulant system by the DFT;
  - the bending speed, surface integrals and tension/pressure solve of
    p00_test_singlephase_relax.m:67-109 (SDF3.m:73-85 for the integrals).

Fields are numpy arrays of shape (pges, cols, rows): the reference's
column-major [rows, cols, pges] layout, so MATLAB dimension 1 (rows, y) is
numpy axis 2, dimension 2 (cols, x) axis 1 and dimension 3 (pges, z) axis 0.

Pins (tests/test_oracle_relax.py): a Fourier mode through smoothFFT against
its closed form; the wavenumbers against numpy's fftfreq (even sizes) and the
reference's colon-and-fftshift values (odd sizes); the circulant solve and
GMRES against the residual of the sparse operator built as GD3.m builds it;
the 2x2 solve against Cramer's rule.
"""
import math

import numpy as np
import scipy.sparse as sp

def circshift(f, shift_rows, shift_cols, shift_pges):
    """MATLAB circshift(f, [shift_rows shift_cols shift_pges]) on a (pges, cols, rows) array"""
    return np.roll(np.roll(np.roll(f, shift_rows, axis=2), shift_cols, axis=1), shift_pges, axis=0)


def gd3_wavenumbers(n, h):
    """fftshift(2*pi*(-n/2 : n/2-1)/(n*h)) (GD3.m:195-201)"""
    v = -n / 2 + np.arange(n, dtype=np.float64)  # the colon -n/2 : n/2-1 has n entries
    return np.fft.fftshift(2 * math.pi * v / (n * h))


def gd3_k2(shape, dx, dy, dz):
    """kx.^2 + ky.^2 + kz.^2 of meshgrid(kx, ky, kz) (GD3.m:203) on (pges, cols, rows)"""
    pges, cols, rows = shape
    kx = gd3_wavenumbers(cols, dx)[None, :, None]
    ky = gd3_wavenumbers(rows, dy)[None, None, :]
    kz = gd3_wavenumbers(pges, dz)[:, None, None]
    return kx ** 2 + ky ** 2 + kz ** 2


def smooth_fft(rhs, dx, dy, dz, dt, alpha, diffusion=False):
    """smoothFFT (GD3.m:453-458) / smoothDiffusionFFT (:479-484)"""
    k2 = gd3_k2(rhs.shape, dx, dy, dz)
    sym = k2 if diffusion else k2 ** 2
    return np.real(np.fft.ifftn(np.fft.fftn(rhs) / (1 + alpha * dt * sym)))


def gd3_laplacian_matrix(shape, dx, dy, dz):
    """obj.LLaplacian = Lxx + Lyy + Lzz, built as GD3.m:110-188 builds it: sparse
    matrices from the circshift index maps (duplicate entries add)"""
    pges, cols, rows = shape
    n = rows * cols * pges
    ooo = np.arange(n).reshape(shape)
    i = ooo.ravel()

    def second(fwd, bwd, h):
        m = (sp.csr_matrix((np.ones(n), (i, fwd.ravel())), shape=(n, n))
             + sp.csr_matrix((np.full(n, -2.0), (i, i)), shape=(n, n))
             + sp.csr_matrix((np.ones(n), (i, bwd.ravel())), shape=(n, n)))
        return m * (1 / h ** 2)

    lxx = second(circshift(ooo, 0, -1, 0), circshift(ooo, 0, 1, 0), dx)  # oXo, oxo
    lyy = second(circshift(ooo, -1, 0, 0), circshift(ooo, 1, 0, 0), dy)  # Yoo, yoo
    lzz = second(circshift(ooo, 0, 0, -1), circshift(ooo, 0, 0, 1), dz)  # ooZ, ooz
    return (lxx + lyy + lzz).tocsr()


def smooth_matrix(shape, dx, dy, dz, dt, alpha, diffusion=False):
    """Op = Idt + alpha*dt*LBiLaplacian (GD3.m:438) or Idt - alpha*dt*LLaplacian (:464)"""
    L = gd3_laplacian_matrix(shape, dx, dy, dz)
    n = L.shape[0]
    eye = sp.identity(n, format="csr")
    return (eye - alpha * dt * L) if diffusion else (eye + alpha * dt * (L @ L))


def gmres_left(A, b, precond, tol=1e-12, maxit=300):
    """x solving A x = b from x0 = 0 by GMRES on precond(A x) = precond(b), no
    restart, at most maxit iterations, stopping when |precond(b - A x)| <=
    tol*|precond(b)| (gmres(A, b, [], tol, maxit, M) of MATLAB with M^-1 =
    precond).  Returns (x, iterations, preconditioned relative residual)."""
    r0 = precond(b)
    beta = np.linalg.norm(r0)
    n = b.size
    if beta == 0:
        return np.zeros(n), 0, 0.0
    m = min(maxit, n)
    V = np.zeros((m + 1, n))
    Hm = np.zeros((m + 1, m))
    cs = np.zeros(m)
    sn = np.zeros(m)
    g = np.zeros(m + 1)
    g[0] = beta
    V[0] = r0 / beta
    k = 0
    for k in range(m):
        w = precond(A @ V[k])
        for j in range(k + 1):
            Hm[j, k] = V[j] @ w
            w = w - Hm[j, k] * V[j]
        Hm[k + 1, k] = np.linalg.norm(w)
        if Hm[k + 1, k] > 0:
            V[k + 1] = w / Hm[k + 1, k]
        for j in range(k):
            t = cs[j] * Hm[j, k] + sn[j] * Hm[j + 1, k]
            Hm[j + 1, k] = -sn[j] * Hm[j, k] + cs[j] * Hm[j + 1, k]
            Hm[j, k] = t
        den = math.hypot(Hm[k, k], Hm[k + 1, k])
        cs[k], sn[k] = Hm[k, k] / den, Hm[k + 1, k] / den
        Hm[k, k] = den
        Hm[k + 1, k] = 0.0
        g[k + 1] = -sn[k] * g[k]
        g[k] = cs[k] * g[k]
        if abs(g[k + 1]) <= tol * beta:
            break
    kk = k + 1
    y = np.linalg.solve(np.triu(Hm[:kk, :kk]), g[:kk])
    x = V[:kk].T @ y
    return x, kk, abs(g[kk]) / beta


def smooth_gmres(rhs, dx, dy, dz, dt, alpha, diffusion=False, tol=1e-12, maxit=300):
    """smoothGMRES (GD3.m:435-451) / smoothDiffusionGMRES (:461-477) with the
    FFT preconditioner mfun; returns (field, iterations, relres)"""
    shape = rhs.shape
    A = smooth_matrix(shape, dx, dy, dz, dt, alpha, diffusion)
    k2 = gd3_k2(shape, dx, dy, dz)
    sym = 1 + alpha * dt * (k2 if diffusion else k2 ** 2)

    def mfun(s):
        return np.real(np.fft.ifftn(np.fft.fftn(s.reshape(shape)) / sym)).ravel()

    x, it, rr = gmres_left(A, rhs.ravel(), mfun, tol, maxit)
    return x.reshape(shape), it, rr


def fd_eigenvalues(shape, dx, dy, dz):
    """eigenvalues of the periodic LLaplacian at each DFT bin:
    sum over axes of -4 sin^2(pi j/n)/h^2"""
    pges, cols, rows = shape

    def ax(n, h):
        return -4 * np.sin(math.pi * np.arange(n) / n) ** 2 / h ** 2

    return ax(pges, dz)[:, None, None] + ax(cols, dx)[None, :, None] + ax(rows, dy)[None, None, :]


def smooth_exact(rhs, dx, dy, dz, dt, alpha, diffusion=False):
    """the solution of the smoothGMRES system by the DFT that diagonalizes it"""
    lam = fd_eigenvalues(rhs.shape, dx, dy, dz)
    sym = 1 + alpha * dt * (-lam if diffusion else lam * lam)
    return np.real(np.fft.ifftn(np.fft.fftn(rhs) / sym))


def gd3_laplacian(f, dx, dy, dz):
    """GD3.Laplacian (GD3.m:344-346 with Fxx :248-250), periodic"""
    fxx = (circshift(f, 0, -1, 0) - 2 * f + circshift(f, 0, 1, 0)) / (dx ** 2)
    fyy = (circshift(f, -1, 0, 0) - 2 * f + circshift(f, 1, 0, 0)) / (dy ** 2)
    fzz = (circshift(f, 0, 0, -1) - 2 * f + circshift(f, 0, 0, 1)) / (dz ** 2)
    return fxx + fyy + fzz


def bending_speed(H, G, dx, dy, dz, kappa, sponc):
    """p00_test_singlephase_relax.m:79-83"""
    return kappa * (gd3_laplacian(H, dx, dy, dz) + 0.5 * H * (H ** 2 - 4 * G) + 2. * sponc * G
                    - 0.5 * sponc ** 2 * H)


def tension_pressure(F, H, Vb, dirac, gmag, heav, ds, area0, volume0, dt=None, cfl=None):
    """p00_test_singlephase_relax.m:67-109 with calArea/calVolume/surfaceIntegral
    of SDF3.m:73-85 (exactly rounded sums, math.fsum).  Returns (speed, dict)."""
    band = np.abs(F) < 2 * ds
    inside = F < 2 * ds
    ds3 = ds ** 3

    def surf(f):
        return math.fsum((f[band] * dirac[band] * gmag[band]).ravel()) * ds3

    area = math.fsum(dirac[band].ravel()) * ds3
    volume = math.fsum((1 - heav[inside]).ravel()) * ds3
    vmax = float(np.abs(Vb[band]).max())
    if dt is None or dt <= 0:
        dt = cfl * ds / vmax
    c11 = surf(H ** 2)
    c12 = surf(H)
    s1 = surf(H * Vb) - (area0 - area) / dt
    s2 = surf(Vb) + (volume0 - volume) / dt
    T, P = np.linalg.solve(np.array([[c11, c12], [c12, area0]]), np.array([s1, s2]))
    speed = (T * H - Vb + P) * gmag
    return speed, dict(Tension=T, Pressure=P, Dt=dt, c11=c11, c12=c12, s1=s1, s2=s2, area=area, volume=volume,
                       vmax=vmax)


def relax_step(F, area0, volume0, dx, dy, dz, cfl=0.5, kappa=1.0, sponc=0.0, alpha=0.5, ext_iters=100,
               reinit=True, reinit_iters=100):
    """one loop body of p00_test_singlephase_relax.m:60-196 on the CPU oracle:
    toolbox, WENORK3Extend of the mean curvature, bending speed and its
    ENORK2Extend, tension/pressure, smoothGMRES (its exact solution), the
    ENORK2Extend of the smoothed speed, F - Dt*speed, setDistance and the
    WENO5-RK3 reinitialization.  Returns (F, info)."""
    from . import pyoracle as po

    ds = (dx + dy + dz) / 3
    ft = po.set_calculus_toolbox(F, np.maximum(F, 0), dx, dy, dz, ds)
    H = po.variant_extend(ft["MeanCurvature"], ext_iters, F, dx, dy, dz, "wenork3")
    Vb = po.variant_extend(bending_speed(H, ft["GaussianCurvature"], dx, dy, dz, kappa, sponc), ext_iters, F, dx, dy,
                           dz, "enork2")
    speed, info = tension_pressure(F, H, Vb, ft["DiracDelta"], ft["FGradMag"], ft["Heaviside"], ds, area0, volume0,
                                   cfl=cfl)
    dt = info["Dt"]
    sm = smooth_exact(speed, dx, dy, dz, dt, alpha)
    sm = po.variant_extend(sm, ext_iters, F, dx, dy, dz, "enork2")
    F = F - dt * sm
    d6, md = po.set_distance(F, dx, dy, dz)
    if reinit:
        F = po.weno5_rk3_reinit(F, reinit_iters, d6, md, dx, dy, dz)
    return F, info
