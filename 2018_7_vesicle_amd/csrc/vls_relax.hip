// The operators of the vesicle relaxation step that follow the level-set path
// (src/7_0_paper/2_0_Vesicle/p00_test_singlephase_relax.m:72-112):
//   * the smoothers of @GD3 (+SD/@GD3/GD3.m:435-486): (I + a dt L^2)^-1 and
//     (I - a dt L)^-1 of a field, L the periodic 7-point Laplacian of the
//     sparse operators (GD3.m:155-191);
//   * the bending speed, the surface integrals and the 2x2 tension/pressure
//     solve that enforce area and volume, and the normal speed (p00:78-109).
//
// FFTs are rocFFT double-precision complex transforms in place on one buffer
// per grid size and device (plans cached for the life of the process).
//
// smoothFFT / smoothDiffusionFFT divide by the spectral symbol 1 + a dt |k|^4
// (|k|^2) with @GD3's wavenumbers (GD3.m:195-207).  smoothGMRES /
// smoothDiffusionGMRES ask for the solution of the finite-difference system
// (I + a dt L^2) x = s, which MATLAB's gmres iterates towards (tolerance
// 1e-12, FFT preconditioner).  L is circulant on the periodic box, so the DFT
// diagonalizes it exactly: its eigenvalue at bin (r, c, p) is
//   lambda = -4 sin^2(pi c/cols)/dx^2 - 4 sin^2(pi r/rows)/dy^2 - 4 sin^2(pi p/pges)/dz^2,
// and x = IDFT(DFT(s) / (1 + a dt lambda^2)) is the system's solution to
// rounding: the limit of the GMRES iteration, in one pair of transforms.
#include <rocfft/rocfft.h>

#include <map>
#include <cstdint>
#include <mutex>
#include <tuple>

#include "vls_tile.hpp"

namespace vls {

namespace {

inline Box rbox(const Geom& g) { return Box{g.rows, g.cols, g.pges, g.dx, g.dy, g.dz}; }

inline dim3 rgrid(const Geom& g) { return dim3((g.rows + 63) / 64, (g.cols + 3) / 4, g.pges); }

__device__ __forceinline__ int wrapn(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// @GD3's wavenumber of FFT bin j on an axis of n points and spacing h:
// fftshift(2*pi*(-n/2 : n/2-1)/(n*h)) (GD3.m:195-201); fftshift moves the last
// floor(n/2) entries to the front, which for odd n keeps the reference's
// half-integer offsets
__device__ __forceinline__ double gd3_wavenumber(int j, int n, double h)
{
    const int m = j - n / 2;
    const int src = m < 0 ? m + n : m;
    const double v = -n / 2.0 + (double)src;
    return 2 * 3.141592653589793 * v / (n * h);
}

// eigenvalue of one axis of the periodic second difference (Lxx of GD3.m:155-158)
__device__ __forceinline__ double fd_eigen(int j, int n, double h)
{
    const double s = sin(3.141592653589793 * (double)j / (double)n);
    return -4.0 * s * s / (h * h);
}

__global__ void k_load_complex(double2* __restrict__ z, const double* __restrict__ x, idx_t n)
{
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = make_double2(x[i], 0.0);
}

// z /= 1 + c K at every bin; K = |k|^4 or lambda^2 (bilaplacian), |k|^2 or -lambda
// (diffusion)
__global__ void __launch_bounds__(256) k_spectral_divide(double2* __restrict__ z, Box b, double c, int diffusion,
                                                         int exact)
{
    int r, cc, p;
    if (!point_of_thread(b, r, cc, p)) return;
    const idx_t ind = lin(b, r, cc, p);
    double K;
    if (exact) {
        const double lam = fd_eigen(cc, b.cols, b.dx) + fd_eigen(r, b.rows, b.dy) + fd_eigen(p, b.pges, b.dz);
        K = diffusion ? -lam : lam * lam;
    } else {
        const double kx = gd3_wavenumber(cc, b.cols, b.dx);
        const double ky = gd3_wavenumber(r, b.rows, b.dy);
        const double kz = gd3_wavenumber(p, b.pges, b.dz);
        const double k2 = kx * kx + ky * ky + kz * kz;
        K = diffusion ? k2 : k2 * k2;
    }
    const double d = 1 + c * K;
    const double2 v = z[ind];
    z[ind] = make_double2(v.x / d, v.y / d);
}

// real(ifftn(.)): rocFFT's inverse is unnormalized
__global__ void k_store_real(double* __restrict__ out, const double2* __restrict__ z, idx_t n, double nn)
{
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = z[i].x / nn;
}

// periodic 7-point Laplacian at (r, c, p), obj.LLaplacian * x (GD3.m:155-188)
__device__ __forceinline__ double plap(const double* __restrict__ x, const Box& b, int r, int c, int p)
{
    const idx_t o = lin(b, r, c, p);
    const double x0 = x[o];
    const double lx = (x[lin(b, r, wrapn(c + 1, b.cols), p)] - 2 * x0 + x[lin(b, r, wrapn(c - 1, b.cols), p)]) /
                      (b.dx * b.dx);
    const double ly = (x[lin(b, wrapn(r + 1, b.rows), c, p)] - 2 * x0 + x[lin(b, wrapn(r - 1, b.rows), c, p)]) /
                      (b.dy * b.dy);
    const double lz = (x[lin(b, r, c, wrapn(p + 1, b.pges))] - 2 * x0 + x[lin(b, r, c, wrapn(p - 1, b.pges))]) /
                      (b.dz * b.dz);
    return lx + ly + lz;
}

// y = (Idt + c LBiLaplacian) x or (Idt - c LLaplacian) x (GD3.m:438, 464)
__global__ void __launch_bounds__(256) k_smooth_operator(double* __restrict__ y, const double* __restrict__ x, Box b,
                                                         double c, int diffusion)
{
    int r, cc, p;
    if (!point_of_thread(b, r, cc, p)) return;
    const idx_t ind = lin(b, r, cc, p);
    if (diffusion) {
        y[ind] = x[ind] - c * plap(x, b, r, cc, p);
        return;
    }
    const double l0 = plap(x, b, r, cc, p);
    const double lxp = plap(x, b, r, wrapn(cc + 1, b.cols), p), lxm = plap(x, b, r, wrapn(cc - 1, b.cols), p);
    const double lyp = plap(x, b, wrapn(r + 1, b.rows), cc, p), lym = plap(x, b, wrapn(r - 1, b.rows), cc, p);
    const double lzp = plap(x, b, r, cc, wrapn(p + 1, b.pges)), lzm = plap(x, b, r, cc, wrapn(p - 1, b.pges));
    const double ll = (lxp - 2 * l0 + lxm) / (b.dx * b.dx) + (lyp - 2 * l0 + lym) / (b.dy * b.dy) +
                      (lzp - 2 * l0 + lzm) / (b.dz * b.dz);
    y[ind] = x[ind] + c * ll;
}

// NormalSpeedBend = Kappa*(GD3.Laplacian(H) + 0.5 H (H^2 - 4 G) + 2 SponC G
// - 0.5 SponC^2 H) (p00_test_singlephase_relax.m:79-83), periodic Laplacian
// (GD3.m:248-250, 344-346)
__global__ void __launch_bounds__(256) k_bending_speed(double* __restrict__ out, const double* __restrict__ H,
                                                       const double* __restrict__ G, Box b, double kappa,
                                                       double sponc)
{
    int r, c, p;
    if (!point_of_thread(b, r, c, p)) return;
    const idx_t ind = lin(b, r, c, p);
    const double h = H[ind], g = G[ind];
    const double t = plap(H, b, r, c, p) + 0.5 * h * (h * h - 4 * g) + 2. * sponc * g - 0.5 * (sponc * sponc) * h;
    out[ind] = kappa * t;
}

// one pass over the grid: Sum delta (|F|<2Ds), Sum (1-Heaviside) (F<2Ds) and the
// four surface integrals Sum f delta |grad F| (SDF3.m:73-85) of f = H^2, H,
// H*Vb, Vb, with the maximum of |Vb| on the band (p00:85-86)
constexpr int kRelaxSums = 6;

__global__ void __launch_bounds__(256) k_relax_sums(double* __restrict__ partial, const double* __restrict__ F,
                                                    const double* __restrict__ H, const double* __restrict__ Vb,
                                                    const double* __restrict__ dirac, const double* __restrict__ gmag,
                                                    const double* __restrict__ heav, double ds, idx_t n)
{
    KSum k[kRelaxSums];
#pragma unroll
    for (int q = 0; q < kRelaxSums; ++q) k[q] = {0.0, 0.0};
    double vmax = 0.0;
    const idx_t stride = (idx_t)gridDim.x * blockDim.x;
    for (idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double f = F[i];
        if (fabs(f) < 2 * ds) {
            const double d = dirac[i], g = gmag[i], h = H[i], v = Vb[i];
            k[0].add(d);
            k[1].add(h * h * d * g);
            k[2].add(h * d * g);
            k[3].add(h * v * d * g);
            k[4].add(v * d * g);
            vmax = dmax(vmax, fabs(v));
        }
        if (f < 2 * ds) k[5].add(1 - heav[i]);
    }
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) vmax = dmax(vmax, shfl_xor_d(vmax, s));
    __shared__ double lmax[4];
    if ((threadIdx.x & 63) == 0) lmax[threadIdx.x >> 6] = vmax;
    block_reduce_store<kRelaxSums>(k, partial);
    if (threadIdx.x == 0)
        partial[(size_t)gridDim.x * 2 * kRelaxSums + blockIdx.x] =
            dmax(dmax(lmax[0], lmax[1]), dmax(lmax[2], lmax[3]));
}

// merges the partials in a fixed order, then (thread 0) the time step
// Dt = CFL Ds / max|Vb| unless given, and [c11 c12; c12 A0] [T; P] = [s1; s2]
// (p00:88-107) by Gaussian elimination with partial pivoting, as MATLAB's
// backslash does for a 2x2 system.  out = {Tension, Pressure, Dt, c11, c12,
// s1, s2, area, volume, max|Vb|}
__global__ void __launch_bounds__(256) k_relax_final(double* __restrict__ out, const double* __restrict__ partial,
                                                     int nblk, double ds, double a0, double v0, double dt_in,
                                                     double cfl)
{
    KSum k[kRelaxSums];
#pragma unroll
    for (int q = 0; q < kRelaxSums; ++q) k[q] = {0.0, 0.0};
    double vmax = 0.0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
#pragma unroll
        for (int q = 0; q < kRelaxSums; ++q) {
            KSum o{partial[(size_t)i * 2 * kRelaxSums + 2 * q], partial[(size_t)i * 2 * kRelaxSums + 2 * q + 1]};
            k[q].merge(o);
        }
        vmax = dmax(vmax, partial[(size_t)nblk * 2 * kRelaxSums + i]);
    }
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) vmax = dmax(vmax, shfl_xor_d(vmax, s));
    __shared__ double lmax[4];
    if ((threadIdx.x & 63) == 0) lmax[threadIdx.x >> 6] = vmax;
    block_reduce<kRelaxSums>(k);
    if (threadIdx.x != 0) return;
    vmax = dmax(dmax(lmax[0], lmax[1]), dmax(lmax[2], lmax[3]));
    double S[kRelaxSums];
#pragma unroll
    for (int q = 0; q < kRelaxSums; ++q) S[q] = k[q].s + k[q].c;
    const double ds3 = ds * ds * ds;
    const double area = S[0] * ds3, volume = S[5] * ds3;
    const double dt = dt_in > 0 ? dt_in : cfl * ds / vmax;
    const double c11 = S[1] * ds3, c12 = S[2] * ds3, c21 = c12, c22 = a0;
    const double s1 = S[3] * ds3 - (a0 - area) / dt;
    const double s2 = S[4] * ds3 + (v0 - volume) / dt;
    double a11 = c11, a12 = c12, a21 = c21, a22 = c22, b1 = s1, b2 = s2;
    if (fabs(a21) > fabs(a11)) {
        double t;
        t = a11; a11 = a21; a21 = t;
        t = a12; a12 = a22; a22 = t;
        t = b1; b1 = b2; b2 = t;
    }
    const double l = a21 / a11;
    const double u22 = a22 - l * a12;
    const double y2 = b2 - l * b1;
    const double P = y2 / u22;
    const double T = (b1 - a12 * P) / a11;
    out[0] = T; out[1] = P; out[2] = dt; out[3] = c11; out[4] = c12; out[5] = s1; out[6] = s2;
    out[7] = area; out[8] = volume; out[9] = vmax;
}

// normalSpeed = (Tension H - NormalSpeedBend + Pressure) FGradMag (p00:109)
__global__ void k_normal_speed_tp(double* __restrict__ out, const double* __restrict__ H,
                                  const double* __restrict__ Vb, const double* __restrict__ gmag,
                                  const double* __restrict__ tp, idx_t n)
{
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (tp[0] * H[i] - Vb[i] + tp[1]) * gmag[i];
}

// ---------------------------------------------------------------------------
// rocFFT plans, one pair per (device, stream, grid size), with the complex
// buffer the transforms run in and the library's work buffer.  Keying by the
// stream gives calls on different streams buffers of their own, so they never
// overwrite each other's data; calls on one stream are ordered by it.
// ---------------------------------------------------------------------------
struct FftPlan {
    rocfft_plan fwd = nullptr, inv = nullptr;
    rocfft_execution_info info = nullptr;
    void* work = nullptr;
    double2* buf = nullptr;
    ~FftPlan()  // only plans that failed to build are destroyed; the cached ones live with the process
    {
        if (fwd) rocfft_plan_destroy(fwd);
        if (inv) rocfft_plan_destroy(inv);
        if (info) rocfft_execution_info_destroy(info);
        if (work) (void)hipFree(work);
        if (buf) (void)hipFree(buf);
    }
};

std::mutex g_fft_mu;
std::map<std::tuple<int, uintptr_t, int, int, int>, FftPlan*> g_fft;
bool g_rocfft_ready = false;

hipError_t fft_plan(const Geom& g, hipStream_t s, FftPlan*& out)
{
    std::lock_guard<std::mutex> lock(g_fft_mu);
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const auto key = std::make_tuple(dev, (uintptr_t)s, g.rows, g.cols, g.pges);
    auto it = g_fft.find(key);
    if (it != g_fft.end()) { out = it->second; return hipSuccess; }
    if (!g_rocfft_ready) {
        if (rocfft_setup() != rocfft_status_success) return hipErrorNotInitialized;
        g_rocfft_ready = true;
    }
    FftPlan* p = new FftPlan;
    const size_t len[3] = {(size_t)g.rows, (size_t)g.cols, (size_t)g.pges};
    size_t w1 = 0, w2 = 0;
    if (rocfft_plan_create(&p->fwd, rocfft_placement_inplace, rocfft_transform_type_complex_forward,
                           rocfft_precision_double, 3, len, 1, nullptr) != rocfft_status_success ||
        rocfft_plan_create(&p->inv, rocfft_placement_inplace, rocfft_transform_type_complex_inverse,
                           rocfft_precision_double, 3, len, 1, nullptr) != rocfft_status_success ||
        rocfft_plan_get_work_buffer_size(p->fwd, &w1) != rocfft_status_success ||
        rocfft_plan_get_work_buffer_size(p->inv, &w2) != rocfft_status_success ||
        rocfft_execution_info_create(&p->info) != rocfft_status_success) {
        delete p;
        return hipErrorInvalidValue;
    }
    const size_t wb = w1 > w2 ? w1 : w2;
    if ((e = hipMalloc((void**)&p->buf, sizeof(double2) * (size_t)g.n())) != hipSuccess) {
        delete p;
        return e;
    }
    if (wb > 0) {
        if ((e = hipMalloc(&p->work, wb)) != hipSuccess) {
            delete p;
            return e;
        }
        if (rocfft_execution_info_set_work_buffer(p->info, p->work, wb) != rocfft_status_success) {
            delete p;
            return hipErrorInvalidValue;
        }
    }
    g_fft[key] = p;
    out = p;
    return hipSuccess;
}

hipError_t fft_run(FftPlan* p, rocfft_plan plan, hipStream_t s)
{
    std::lock_guard<std::mutex> lock(g_fft_mu);
    void* io[1] = {p->buf};
    if (rocfft_execution_info_set_stream(p->info, s) != rocfft_status_success ||
        rocfft_execute(plan, io, nullptr, p->info) != rocfft_status_success)
        return hipErrorLaunchFailure;
    return hipGetLastError();
}

}  // namespace

hipError_t launch_smooth(double* out, const double* rhs, const Geom& g, double c, int diffusion, int exact,
                         hipStream_t s)
{
    FftPlan* p = nullptr;
    hipError_t e = fft_plan(g, s, p);
    if (e != hipSuccess) return e;
    const idx_t n = (idx_t)g.n();
    const dim3 flat((unsigned)((n + 255) / 256));
    k_load_complex<<<flat, dim3(256), 0, s>>>(p->buf, rhs, n);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = fft_run(p, p->fwd, s)) != hipSuccess) return e;
    k_spectral_divide<<<rgrid(g), dim3(64, 4, 1), 0, s>>>(p->buf, rbox(g), c, diffusion, exact);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = fft_run(p, p->inv, s)) != hipSuccess) return e;
    k_store_real<<<flat, dim3(256), 0, s>>>(out, p->buf, n, (double)n);
    return hipGetLastError();
}

hipError_t launch_smooth_operator(double* y, const double* x, const Geom& g, double c, int diffusion, hipStream_t s)
{
    k_smooth_operator<<<rgrid(g), dim3(64, 4, 1), 0, s>>>(y, x, rbox(g), c, diffusion);
    return hipGetLastError();
}

hipError_t launch_bending_speed(double* out, const double* H, const double* G, const Geom& g, double kappa,
                                double sponc, hipStream_t s)
{
    k_bending_speed<<<rgrid(g), dim3(64, 4, 1), 0, s>>>(out, H, G, rbox(g), kappa, sponc);
    return hipGetLastError();
}

hipError_t launch_tension_pressure(double* out10, double* partial, double* speed, const double* F, const double* H,
                                   const double* Vb, const double* dirac, const double* gmag, const double* heav,
                                   const Geom& g, double ds, double a0, double v0, double dt, double cfl,
                                   hipStream_t s)
{
    const int64_t want = (g.n() + 255) / 256;
    const int nb = (int)(want < reduce_blocks() ? (want < 1 ? 1 : want) : reduce_blocks());
    const idx_t n = (idx_t)g.n();
    k_relax_sums<<<dim3(nb), dim3(256), 0, s>>>(partial, F, H, Vb, dirac, gmag, heav, ds, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_relax_final<<<dim3(1), dim3(256), 0, s>>>(out10, partial, nb, ds, a0, v0, dt, cfl);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_normal_speed_tp<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(speed, H, Vb, gmag, out10, n);
    return hipGetLastError();
}

}  // namespace vls
