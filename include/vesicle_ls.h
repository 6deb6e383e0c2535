/*
 * vesicle_ls.h -- C ABI of the MI355X level-set hot path.
 *
 * Drop-in replacement for the CUDA kernels and MATLAB wrappers of the
 * level-set path of tiankuizhang/2018_7_Vesicle (paths below are relative to
 * src/7_0_paper/2_0_Vesicle/).  The reference binds its kernels through
 * parallel.gpu.CUDAKernel objects (+SD/@SDF3/GPUInitialize.m:16-180) and calls
 * them as feval(kernel, args...); the functions in section 1 keep each
 * kernel's positional arguments and meaning, so a MATLAB mex shim or any other
 * host can swap feval for a call here.  Section 2 replaces the MATLAB wrapper
 * methods of @SDF3 with single calls on a handle that owns the object's
 * device state.
 *
 * Conventions
 *   - fields are double[rows*cols*pges] in device memory, column-major
 *     [rows, cols, pges] (rows = y, cols = x, pges = z):
 *     ind = pge*rows*cols + col*rows + row; stencil neighbours are clamped;
 *   - pointers are caller-owned device pointers; non-const pointers are
 *     outputs written in place (feval's in/out semantics); `mask` is one byte
 *     per point (MATLAB logical);
 *   - `stream` is a hipStream_t, NULL = the default stream; every call is
 *     asynchronous on it except where a host result is returned;
 *   - return value: VLS_OK (0) or a negative VLS_ERR_* code; vls_last_error()
 *     describes the last failure of the calling thread.  No C++ exception
 *     crosses the ABI.
 */
#ifndef VESICLE_LS_H
#define VESICLE_LS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VLS_OK 0
#define VLS_ERR_INVALID (-1) /* bad size, null pointer or option */
#define VLS_ERR_HIP (-2)     /* a HIP runtime call failed */
#define VLS_ERR_ALLOC (-3)   /* device allocation failed */

typedef void *vls_stream_t; /* hipStream_t */

const char *vls_last_error(void);
const char *vls_version(void);

/* ------------------------------------------------------------------------ */
/* 1. kernel-level drop-ins (one per reference __global__ entry point)      */
/* ------------------------------------------------------------------------ */

/* boundary_location: CUDA_Code/1_3_WENO5RK3_Reinitialization/boundary_location.cu:84-85.
 * Sixth-order subcell distances; the six arrays are in/out and must be
 * pre-filled with dx,dx,dy,dy,dz,dz (+SD/@SDF3/setDistance.m:5-10). */
int vls_boundary_location(double *xpr, double *xpl, double *ypf, double *ypb, double *zpu, double *zpd,
                          const double *lsf, int num_ele, int rows, int cols, int pges, double dx, double dy,
                          double dz, vls_stream_t stream);

/* boundary_correction: CUDA_Code/1_0_ENORK2_Reinitialization/boundary_correction.cu:103-104.
 * Quadratic-ENO subcell distances, same in/out convention. */
int vls_boundary_correction(double *xpr, double *xpl, double *ypf, double *ypb, double *zpu, double *zpd,
                            const double *lsf, int num_ele, int rows, int cols, int pges, double dx, double dy,
                            double dz, vls_stream_t stream);

/* re_step (WENO5): CUDA_Code/1_3_WENO5RK3_Reinitialization/re_step.cu:283-284. */
int vls_re_step(double *step, const double *lsf, const uint8_t *mask, const double *deltat, const double *xpr,
                const double *xpl, const double *ypf, const double *ypb, const double *zpu, const double *zpd,
                int rows, int cols, int pges, double dx, double dy, double dz, int num_ele, vls_stream_t stream);

/* re_step (ENO2): CUDA_Code/1_0_ENORK2_Reinitialization/boundary_correction.cu:147-148. */
int vls_eno2_re_step(double *step, const double *lsf, const uint8_t *mask, const double *deltat,
                     const double *xpr, const double *xpl, const double *ypf, const double *ypb,
                     const double *zpu, const double *zpd, int rows, int cols, int pges, double dx, double dy,
                     double dz, int num_ele, vls_stream_t stream);

/* upwind_normal: CUDA_Code/2_2_WENO5RK3_Extend/upwind_normal.cu:14-15. */
int vls_upwind_normal(double *nx, double *ny, double *nz, const double *lsf, const double *xpr, const double *xpl,
                      const double *ypf, const double *ypb, const double *zpu, const double *zpd, int rows,
                      int cols, int pges, double dx, double dy, double dz, int num_ele, vls_stream_t stream);

/* boundary_interpolant: CUDA_Code/2_2_WENO5RK3_Extend/boundary_interpolant.cu:102-103.
 * `lsf` is the field being extended (the reference passes C there). */
int vls_boundary_interpolant(double *cpr, double *cpl, double *cpf, double *cpb, double *cpu, double *cpd,
                             const double *xpr, const double *xpl, const double *ypf, const double *ypb,
                             const double *zpu, const double *zpd, const double *lsf, int rows, int cols,
                             int pges, double dx, double dy, double dz, int num_ele, vls_stream_t stream);

/* extend_step: CUDA_Code/2_2_WENO5RK3_Extend/extend_step.cu:10-11. */
int vls_extend_step(double *step, const double *deltat, const double *lsf, const double *vx, const double *vy,
                    const double *vz, const double *xpr, const double *xpl, const double *ypf, const double *ypb,
                    const double *zpu, const double *zpd, const double *cpr, const double *cpl, const double *cpf,
                    const double *cpb, const double *cpu, const double *cpd, int rows, int cols, int pges,
                    double dx, double dy, double dz, int num_ele, vls_stream_t stream);

/* WENO-RK3 / ENO-RK2 variants.  boundary_correction (cubic):
 * CUDA_Code/1_1_WENORK3_Reinitialization/weno_rk3_reinitialization.cu:251-252 (bound
 * as cubic_boundary_correction, +SD/@SDF3/GPUInitialize.m:32-36).  Cubic
 * subcell distances (Newton on the cubic through 4 nodes, quadratic ENO next to
 * a kink); in/out convention of vls_boundary_location. */
int vls_cubic_boundary_correction(double *xpr, double *xpl, double *ypf, double *ypb, double *zpu, double *zpd,
                                  const double *lsf, int num_ele, int rows, int cols, int pges, double dx,
                                  double dy, double dz, vls_stream_t stream);

/* re_step (WENO-RK3): weno_rk3_reinitialization.cu:295-296 (WENORK3_reinitialization_step,
 * GPUInitialize.m:38-42).  Cubic ENO next to a sign change, WENO5 elsewhere;
 * frozen where a distance is below 1e-6 dx. */
int vls_wenork3_re_step(double *step, const double *lsf, const uint8_t *mask, const double *deltat,
                        const double *xpr, const double *xpl, const double *ypf, const double *ypb,
                        const double *zpu, const double *zpd, int rows, int cols, int pges, double dx, double dy,
                        double dz, int num_ele, vls_stream_t stream);

/* upwind_normal (ENO-RK2 extension): CUDA_Code/2_0_ENORK2_Extend/enork2_extend.cu:102-103
 * (ENORK2_upwind_normal, GPUInitialize.m:64-68). */
int vls_enork2_upwind_normal(double *nx, double *ny, double *nz, const double *lsf, const double *xpr,
                             const double *xpl, const double *ypf, const double *ypb, const double *zpu,
                             const double *zpd, int rows, int cols, int pges, double dx, double dy, double dz,
                             int num_ele, vls_stream_t stream);

/* boundary_interpolate (ENO-RK2): enork2_extend.cu:174-175 (ENORK2_boundary_interpolate,
 * GPUInitialize.m:76-80).  Quadratic ENO values of lsf at the crossings; 0 elsewhere. */
int vls_enork2_boundary_interpolate(double *cpr, double *cpl, double *cpf, double *cpb, double *cpu, double *cpd,
                                    const double *xpr, const double *xpl, const double *ypf, const double *ypb,
                                    const double *zpu, const double *zpd, const double *lsf, int rows, int cols,
                                    int pges, double dx, double dy, double dz, int num_ele, vls_stream_t stream);

/* extend_step (ENO-RK2): enork2_extend.cu:213-214 (ENORK2_extend_step, GPUInitialize.m:70-74). */
int vls_enork2_extend_step(double *step, const double *deltat, const double *lsf, const double *vx, const double *vy,
                           const double *vz, const double *xpr, const double *xpl, const double *ypf,
                           const double *ypb, const double *zpu, const double *zpd, const double *cpr,
                           const double *cpl, const double *cpf, const double *cpb, const double *cpu,
                           const double *cpd, int rows, int cols, int pges, double dx, double dy, double dz,
                           int num_ele, vls_stream_t stream);

/* upwind_normal (WENO-RK3 extension): CUDA_Code/2_1_WENORK3_Extend/wenork3_extend.cu:202-203
 * (WENORK3_upwind_normal, GPUInitialize.m:86-90). */
int vls_wenork3_upwind_normal(double *nx, double *ny, double *nz, const double *lsf, const double *xpr,
                              const double *xpl, const double *ypf, const double *ypb, const double *zpu,
                              const double *zpd, int rows, int cols, int pges, double dx, double dy, double dz,
                              int num_ele, vls_stream_t stream);

/* boundary_interpolate (WENO-RK3): wenork3_extend.cu:331-332 (WENORK3_boundary_interpolate,
 * GPUInitialize.m:98-102).  Cubic values of lsf at the crossings; 0 elsewhere. */
int vls_wenork3_boundary_interpolate(double *cpr, double *cpl, double *cpf, double *cpb, double *cpu, double *cpd,
                                     const double *xpr, const double *xpl, const double *ypf, const double *ypb,
                                     const double *zpu, const double *zpd, const double *lsf, int rows, int cols,
                                     int pges, double dx, double dy, double dz, int num_ele, vls_stream_t stream);

/* extend_step (WENO-RK3): wenork3_extend.cu:372-373 (WENORK3_extend_step, GPUInitialize.m:92-96). */
int vls_wenork3_extend_step(double *step, const double *deltat, const double *lsf, const double *vx,
                            const double *vy, const double *vz, const double *xpr, const double *xpl,
                            const double *ypf, const double *ypb, const double *zpu, const double *zpd,
                            const double *cpr, const double *cpl, const double *cpf, const double *cpb,
                            const double *cpu, const double *cpd, int rows, int cols, int pges, double dx, double dy,
                            double dz, int num_ele, vls_stream_t stream);

/* advection_step: CUDA_Code/4_0_Calculus_ToolBox/advection_step.cu:30-31 (v.grad(lsf), no dt). */
int vls_advection_step(double *astep, const double *vx, const double *vy, const double *vz, const double *lsf,
                       int rows, int cols, int pges, double dx, double dy, double dz, vls_stream_t stream);

/* weno_derivative: CUDA_Code/5_0_Nonoscillating_Reconstruction/nonoscillating_interpolant.cu:59-60. */
int vls_weno_derivative(double *back_x, double *fore_x, double *back_y, double *fore_y, double *back_z,
                        double *fore_z, const double *lsf, int rows, int cols, int pges, double dx, double dy,
                        double dz, vls_stream_t stream);

/* set_calculus_toolbox: CUDA_Code/4_0_Calculus_ToolBox/set_calculus_toolbox.cu:9-10.
 * out18 = {Fx,Fy,Fz,FGradMag,Nx,Ny,Nz,Fxx,Fyy,Fzz,Fxy,Fyz,Fzx,FLaplacian,
 *          MeanCurvature,GaussianCurvature,Heaviside,DiracDelta} (18 device pointers). */
int vls_set_calculus_toolbox(double *const *out18, const double *lsf, const double *hprimal, int rows, int cols,
                             int pges, double dx, double dy, double dz, double ds, int num_ele,
                             vls_stream_t stream);

/* auxi_set_calculus_toolbox: CUDA_Code/4_0_Calculus_ToolBox/auxi_set_calculus_toolbox.cu:9-199,
 * called by +SD/@SDF3/GPUAsetCalculusToolBox.m:40-54.  lsf is the auxiliary
 * field A, ahprimal its primal max(A,0); fin13 = {Fx,Fy,Fz,FGradMag,Nx,Ny,Nz,
 * Fxx,Fyy,Fzz,Fxy,Fyz,Fzx} of F (set_calculus_toolbox's first 13 outputs);
 * out27 = {Ax,Ay,Az,AGradMag,Cx,Cy,Cz,NormCrossAF,tx,ty,tz,nx,ny,nz,Axx,Ayy,Azz,
 * Axy,Ayz,Azx,ALaplacian,GeodesicCurvature,NormalCurvature,GeodesicTorsion,
 * BPerpendicular,AHeaviside,ADiracDelta} (27 device pointers). */
int vls_auxi_set_calculus_toolbox(double *const *out27, const double *lsf, const double *ahprimal,
                                  const double *const *fin13, int rows, int cols, int pges, double dx, double dy,
                                  double dz, double ds, int num_ele, vls_stream_t stream);

/* GPUAsetCalculusToolBox: +SD/@SDF3/GPUAsetCalculusToolBox.m:2-64, the whole
 * method: AHPrimal = max(A,0) (:4), the kernel above (:40-54), then
 * AFDiracDelta = C.(grad Heaviside x grad AHeaviside)/NormCrossAF^2 (:57-61) and
 * ADiracDeltaDn = (Lap AHeaviside - ADiracDelta Lap A)/AGradMag (:63) with
 * @GD3's periodic central differences (+SD/@GD3/GD3.m:112-117,218-250,344-346).
 * ftool14 = fin13 above followed by F's Heaviside; out30 = {AHPrimal, out27,
 * AFDiracDelta, ADiracDeltaDn}. */
int vls_aux_calculus_toolbox(double *const *out30, const double *A, const double *const *ftool14, int rows,
                             int cols, int pges, double dx, double dy, double dz, double ds, vls_stream_t stream);

/* The @GD3 smoothers of the vesicle step (+SD/@GD3/GD3.m:435-486), applied
 * to rhs on the periodic box; out may alias rhs.  diffusion = 0: the
 * bilaplacian smoothers, (Idt + alpha*dt*LBiLaplacian)^-1; diffusion = 1:
 * (Idt - alpha*dt*LLaplacian)^-1.  rocFFT double-precision transforms; the
 * first call for a grid size on a stream builds and keeps plans and a complex
 * buffer for that (grid size, stream) until the process ends, so calls on
 * different streams never share a buffer; calls on one stream are ordered.
 *
 * vls_smooth_fft: smoothFFT (GD3.m:453-458) / smoothDiffusionFFT (:479-484),
 * real(ifftn(fftn(rhs) ./ (1 + alpha*dt*|k|^4 or |k|^2))) with @GD3's
 * wavenumbers (GD3.m:195-207). */
int vls_smooth_fft(double *out, const double *rhs, int rows, int cols, int pges, double dx, double dy, double dz,
                   double dt, double alpha, int diffusion, vls_stream_t stream);

/* vls_smooth_gmres: smoothGMRES (GD3.m:435-451) / smoothDiffusionGMRES
 * (:461-477), which the drivers call on the normal speed
 * (p00_test_singlephase_relax.m:110).  Returns the solution of the same
 * finite-difference system that MATLAB's gmres solves to 1e-12: the periodic
 * operator is circulant, so one pair of FFTs with its exact eigenvalues
 * solves it to rounding. */
int vls_smooth_gmres(double *out, const double *rhs, int rows, int cols, int pges, double dx, double dy,
                     double dz, double dt, double alpha, int diffusion, vls_stream_t stream);

/* y = Op*x with Op = obj.Idt + alpha*dt*obj.LBiLaplacian (GD3.m:438) or
 * obj.Idt - alpha*dt*obj.LLaplacian (:464); y must not alias x.  The residual
 * of a smoothing solve is rhs - Op*out. */
int vls_smooth_operator(double *y, const double *x, int rows, int cols, int pges, double dx, double dy, double dz,
                        double dt, double alpha, int diffusion, vls_stream_t stream);

/* NormalSpeedBend = Kappa*(GD3.Laplacian(H) + 0.5*H.*(H.^2 - 4*G) + 2*SponC*G
 * - 0.5*SponC^2*H) (p00_test_singlephase_relax.m:79-83); H the (extended) mean
 * curvature, G the Gaussian curvature; periodic Laplacian.  out must not alias H. */
int vls_bending_speed(double *out, const double *H, const double *G, int rows, int cols, int pges, double dx,
                      double dy, double dz, double kappa, double sponc, vls_stream_t stream);

/* MATLAB Level 5 MAT-files (the format of the reference's fixtures and run
 * records, e.g. src/7_0_paper/2_0_Vesicle/pear.mat and the CFL*.mat files that
 * p00_test_singlephase_relax.m:197-206 saves with MATLAB's save), for fixtures
 * and restart.  Host memory only; little-endian files; compressed
 * (miCOMPRESSED) and plain elements.  Only numeric arrays are listed; every
 * numeric class reads as double in MATLAB's column-major order.
 *
 * vls_mat_count: number of numeric arrays in the file. */
int vls_mat_count(const char *path, int *count);
/* vls_mat_info: name (NUL-terminated, at most name_cap bytes), rank and sizes
 * (dims8, up to 8) of the index-th numeric array; is_complex may be NULL. */
int vls_mat_info(const char *path, int index, char *name, int name_cap, int *ndims, int64_t *dims8,
                 int *is_complex);
/* vls_mat_read: the real part of the array called name (the first numeric
 * array when name is NULL) into out, which holds exactly count doubles. */
int vls_mat_read(const char *path, const char *name, double *out, int64_t count);
/* vls_mat_write: a new file with narrays double arrays (names 1-63
 * characters, ndims[k] in 2..8, dims8 holding 8 sizes per array); compress =
 * 1 stores each array zlib-compressed as MATLAB's default save does.  Written
 * to path.tmp and renamed. */
int vls_mat_write(const char *path, int narrays, const char *const *names, const double *const *data,
                  const int *ndims, const int64_t *dims8, int compress);

/* ------------------------------------------------------------------------ */
/* 2. the @SDF3 object: one handle, one call per MATLAB method              */
/* ------------------------------------------------------------------------ */
typedef struct vls_sdf3 *vls_sdf3_t;

/* SD.SDF3(grid, x, y, z, F) with grid = SD.GD3 (+SD/@GD3/GD3.m:86-105,
 * +SD/@SDF3/SDF3.m:11-15): allocates the object's device state for a
 * rows x cols x pges box with spacings dx, dy, dz.  No limit on rows (the
 * reference's launch shape needs rows <= 256, GPUInitialize.m:7); a page
 * (rows x cols) holds fewer than 2^26 points. */
int vls_sdf3_create(vls_sdf3_t *out, int rows, int cols, int pges, double dx, double dy, double dz);
int vls_sdf3_destroy(vls_sdf3_t h);
/* tile shape of the fused bulk stage kernels: 64 (64 x 4 points of a page,
 * the default) or 32 (32 x 8); results do not depend on it */
int vls_sdf3_set_tile_rows(vls_sdf3_t h, int tile_rows);
/* pages per block of the z-marching bulk stage kernels: each block walks one
 * tile column over `depth` consecutive pages (0, the default: chosen by the
 * box size, about 8 blocks per CU); results do not depend on it */
int vls_sdf3_set_zmarch_depth(vls_sdf3_t h, int depth);
/* the tile rows and z-march depth the bulk stage kernels launch with */
int vls_sdf3_get_launch(vls_sdf3_t h, int *tile_rows, int *depth);

/* setDistance (+SD/@SDF3/setDistance.m:3-21): subcell distances of F with the
 * sixth-order locator, and min_dist.  Stored in the handle: the points with
 * a subcell distance, with their six distances, in an interface list (what
 * the reinitialization, extension and redistance read); the dense arrays are
 * built from it when vls_sdf3_distances asks for them. */
int vls_sdf3_set_distance(vls_sdf3_t h, const double *F, vls_stream_t stream);

/* device pointers of the handle's xpr,xpl,ypf,ypb,zpu,zpd,min_dist (7 entries,
 * read-only) as setDistance.m leaves them.  The first call after a
 * set_distance builds them (spacings, then the listed points' distances)
 * after a device-wide synchronization, and returns once they are written. */
int vls_sdf3_distances(vls_sdf3_t h, const double **out7);

/* WENO5RK3Reinitialization (+SD/@SDF3/WENO5RK3Reinitialization.m:3-43): F in
 * place; mask = F<0 at entry, distances from the last set_distance. */
int vls_sdf3_weno5_rk3_reinit(vls_sdf3_t h, double *F, int iterations, vls_stream_t stream);

/* ENORK2Reinitialization (+SD/@SDF3/ENORK2Reinitialization.m:3-45): ENO2
 * distances of the argument F, dt = cfl*min_dist.  rk2 = 1 is the reference's
 * Heun RK2 with cfl 0.3; rk2 = 0 is forward Euler. */
int vls_sdf3_eno2_reinit(vls_sdf3_t h, double *F, int iterations, int rk2, double cfl, vls_stream_t stream);

/* WENORK3Reinitialization (+SD/@SDF3/WENORK3Reinitialization.m:3-61): F in
 * place; cubic distances of the argument F, mask = F<0, dt = 0.3*min
 * distance, Shu-Osher RK3 of the WENO-RK3 re_step. */
int vls_sdf3_wenork3_reinit(vls_sdf3_t h, double *F, int iterations, vls_stream_t stream);

/* ENORK2Extend (+SD/@SDF3/ENORK2Extend.m:9-81): extends C (in place) off the
 * zero set of F with ENO2 derivatives, quadratic-ENO distances and interface
 * values of F and C, Heun RK2. */
int vls_sdf3_enork2_extend(vls_sdf3_t h, const double *F, double *C, int iterations, vls_stream_t stream);

/* WENORK3Extend (+SD/@SDF3/WENORK3Extend.m:9-98): extends C (in place) off the
 * zero set of F with cubic distances and interface values, cubic ENO next to
 * the interface and WENO5 elsewhere, Shu-Osher RK3. */
int vls_sdf3_wenork3_extend(vls_sdf3_t h, const double *F, double *C, int iterations, vls_stream_t stream);

/* WENO5RK3Extend (+SD/@SDF3/WENO5RK3Extend.m:9-76): extends C (in place) off
 * the zero set of F, which must be the field of the last set_distance. */
int vls_sdf3_weno5_rk3_extend(vls_sdf3_t h, const double *F, double *C, int iterations, vls_stream_t stream);

/* WENO5RK3ClosetPointSurfaceRedistance
 * (+SD/@SDF3/WENO5RK3ClosetPointSurfaceRedistance.m:4-157): A in place. */
int vls_sdf3_surface_redistance(vls_sdf3_t h, const double *F, double *A, int iter1, int iter2,
                                vls_stream_t stream);

/* TVD-RK3 WENO5 advection of F by (vx,vy,vz) for `steps` steps of dt: the
 * Shu-Osher form of +SD/@SDF3/WENORK3SurfaceConservationLaw.m:10-29 with the
 * operator of advection_step.cu:30-56.  No reference entry point composes the
 * two; this is that composition. */
int vls_sdf3_advect_rk3(vls_sdf3_t h, double *F, const double *vx, const double *vy, const double *vz, double dt,
                        int steps, vls_stream_t stream);

/* max over the box of max(|vx|,|vy|,|vz|), returned to the host (synchronizes) */
int vls_sdf3_max_speed(vls_sdf3_t h, const double *vx, const double *vy, const double *vz, double *out,
                       vls_stream_t stream);

/* calArea/calVolume (+SD/@SDF3/SDF3.m:73-80) with the Towers delta/H of
 * GPUsetCalculusToolBox (+SD/@SDF3/GPUsetCalculusToolBox.m:4,30-40), fused.
 * Compensated FP64 sums; returns to the host (synchronizes). */
int vls_sdf3_area_volume_towers(vls_sdf3_t h, const double *F, double *area, double *volume,
                                vls_stream_t stream);

/* The tension/pressure step of the vesicle driver (p00_test_singlephase_relax.m:
 * 67-109): area = calArea, volume = calVolume (SDF3.m:73-80), the surface
 * integrals c11 = surfaceIntegral(H.^2), c12 = surfaceIntegral(H) (SDF3.m:82-85),
 * Dt = cfl*Ds/max|Vb| over |F|<2Ds when dt <= 0 (else dt), s1 =
 * surfaceIntegral(H.*Vb) - (area0 - area)/Dt, s2 = surfaceIntegral(Vb) +
 * (volume0 - volume)/Dt, [Tension; Pressure] = [c11 c12; c12 area0] \ [s1; s2]
 * (partial pivoting), and speed = (Tension*H - Vb + Pressure).*gradmag.  H is
 * the extended mean curvature, Vb the extended bending speed, dirac/gradmag/
 * heaviside the toolbox fields of F.  out10 (device, 10 doubles) = {Tension,
 * Pressure, Dt, c11, c12, s1, s2, area, volume, max|Vb|}.  Compensated sums in
 * a fixed order; nothing returns to the host, so the call can be captured. */
int vls_sdf3_tension_pressure(vls_sdf3_t h, const double *F, const double *H, const double *Vb, const double *dirac,
                              const double *gradmag, const double *heaviside, double area0, double volume0,
                              double dt, double cfl, double *speed, double *out10, vls_stream_t stream);

/* Cosine-smoothed delta/H of half-width eps (src/0_0_setup/+SD/@SDF3/SDF3.m:59,126-136,
 * 178-187; the reference uses eps = 1.5*dx).  Returns to the host (synchronizes). */
int vls_sdf3_area_volume_cosine(vls_sdf3_t h, const double *F, double eps, double *area, double *volume,
                                vls_stream_t stream);

/* Either area/volume pair written to device memory out2[0..1] without
 * synchronizing: kind 0 = Towers (eps unused), kind 1 = cosine of half-width eps. */
int vls_sdf3_area_volume_device(vls_sdf3_t h, const double *F, int kind, double eps, double *out2,
                                vls_stream_t stream);

/* Normal component u.N of a velocity with N = grad F / max(|grad F|, 1e-14)
 * from central differences (set_calculus_toolbox.cu:31-44): the field the
 * vesicle drivers extend off the interface (V/p00_test_singlephase_relax.m:76-84). */
int vls_sdf3_normal_velocity(vls_sdf3_t h, const double *F, const double *vx, const double *vy, const double *vz,
                             double *out, vls_stream_t stream);

/* ------------------------------------------------------------------------ */
/* 3. FP32-bulk variants of the handle methods                              */
/* ------------------------------------------------------------------------ */
/* The same methods on float fields (same handle, same semantics).  The
 * reference computes everything in FP64; these entry points have no
 * counterpart there.  What stays FP64:
 *   - the subcell crossing distances (the sixth-order locator and its root
 *     solve read F converted to double and store double distances);
 *   - every interface-list point (a point with a subcell distance): its
 *     nonuniform WENO, Godunov term and RK combination run in FP64 on the
 *     float values (a bulk point that meets a sign change during the
 *     reinitialization stays in the FP32 bulk: the reference's stencil there
 *     is the nonuniform one at full spacings, the uniform WENO5 up to
 *     rounding);
 *   - the interface values and velocities of the extension (list order);
 *   - the area/volume and max-speed reductions (compensated FP64 sums).
 * The bulk (every other point) runs its WENO5, upwinding, Godunov term and RK
 * combination in FP32, storing float results. */
int vls_sdf3_set_distance_f32(vls_sdf3_t h, const float *F, vls_stream_t stream);
int vls_sdf3_weno5_rk3_reinit_f32(vls_sdf3_t h, float *F, int iterations, vls_stream_t stream);
/* The last `iterations` iterations of every later vls_sdf3_weno5_rk3_reinit_f32
 * call on this handle run in FP64 (no counterpart in the reference; default
 * 0): F is widened into an FP64 work field of the handle, iterated with the
 * call's mask, flags and interface list, and rounded back to float, so the
 * call stays one call of its iteration count.  Over a 100-iteration call the
 * FP32 iterations' rounding adds up where the distance function converges
 * slowly; finishing in FP64 takes it out.  A count above a call's iteration
 * count makes the whole call FP64.  Negative counts: VLS_ERR_INVALID. */
int vls_sdf3_set_reinit_fp64_tail(vls_sdf3_t h, int iterations);
int vls_sdf3_weno5_rk3_extend_f32(vls_sdf3_t h, const float *F, float *C, int iterations, vls_stream_t stream);
int vls_sdf3_surface_redistance_f32(vls_sdf3_t h, const float *F, float *A, int iter1, int iter2,
                                    vls_stream_t stream);
int vls_sdf3_advect_rk3_f32(vls_sdf3_t h, float *F, const float *vx, const float *vy, const float *vz, double dt,
                            int steps, vls_stream_t stream);
int vls_sdf3_max_speed_f32(vls_sdf3_t h, const float *vx, const float *vy, const float *vz, double *out,
                           vls_stream_t stream);
int vls_sdf3_area_volume_device_f32(vls_sdf3_t h, const float *F, int kind, double eps, double *out2,
                                    vls_stream_t stream);
int vls_sdf3_normal_velocity_f32(vls_sdf3_t h, const float *F, const float *vx, const float *vy, const float *vz,
                                 float *out, vls_stream_t stream);

/* Allocation-free WENO5-RK3 reinitialization: no counterpart in the
 * reference.  It is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same result, bit for bit, as vls_sdf3_create + vls_sdf3_set_distance(F) +
 * vls_sdf3_weno5_rk3_reinit(F, iterations) on a fresh handle: the distances
 * of F, then `iterations` RK3 iterations of F in place, on `stream`.  Its work
 * memory is one static array per GPU (about 81 bytes per point for
 * VLS_STATIC_POINTS points, 81 MiB, held by each GPU that loads the library
 * until the process ends); a box of more points is rejected (VLS_ERR_INVALID).
 * It calls no hipMalloc/hipFree and records no memset.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams or
 * threads) race and must be serialized by the caller. */
#define VLS_STATIC_POINTS (1 << 20)
int vls_weno5_rk3_reinit_static(double *F, int iterations, int rows, int cols, int pges, double dx, double dy,
                                double dz, vls_stream_t stream);

/* Graph capture of the wrapper-level loops (no counterpart in the reference,
 * which launches every kernel from MATLAB): with it on (the default), a
 * reinitialization, extension, surface redistance, advection or batch step
 * call is captured once into a hipGraph, keyed by its device pointers and
 * counts, and replayed after that (8 graphs cached per handle).  Returns the
 * previous setting. */
int vls_set_graphs(int enable);

/* Measurement only, no counterpart in the reference: the vector-ALU issue
 * rate of the device (wave64 instructions per second over all compute units)
 * for independent FP32 (fp64 = 0) or FP64 fused multiply-add chains, timed
 * with HIP events on `stream`; the benchmark reports a stage kernel's VALU
 * instructions per second against it.  Allocates and frees a small buffer. */
int vls_probe_valu_rate(int fp64, double *insts_per_s, vls_stream_t stream);

/* Measurement only, no counterpart in the reference: the HBM rate (bytes per
 * second) of a streaming kernel that reads five float arrays of `points`
 * elements and writes one, the access mix of an FP32-bulk extension stage,
 * timed with HIP events on `stream` (points >= 1024; choose the arrays larger
 * than the last-level cache).  Allocates and frees its buffers. */
int vls_probe_stream_rate(int64_t points, double *bytes_per_s, vls_stream_t stream);


/* ------------------------------------------------------------------------ */
/* 4. batches of grids                                                      */
/* ------------------------------------------------------------------------ */
/* Many independent level sets of different box sizes (one vesicle each),
 * advanced together: every operator stage of every grid is one launch of a
 * persistent kernel over the batch's tile list.  No counterpart in the
 * reference, which holds one SDF3 object per grid; each grid of a batch gets
 * what the vls_sdf3_* method of the same name gives it on its own (bit for
 * bit, except the area/volume sums, whose summation order differs).
 *
 * Fields of a batch are packed: the grids' column-major arrays back to back,
 * grid g at elements offsets[g] .. offsets[g+1]-1 (vls_batch_layout), at most
 * 2^31 - 1 points in all.  Methods come in FP64 and FP32-bulk (_f32) forms,
 * as the vls_sdf3_* ones.  Per-grid scalars are device arrays of doubles. */
typedef struct vls_batch *vls_batch_t;

int vls_batch_create(vls_batch_t *out, int ngrids, const int *rows, const int *cols, const int *pges,
                     const double *dx, const double *dy, const double *dz);
int vls_batch_destroy(vls_batch_t b);
/* offsets: ngrids + 1 host int64; ntiles (may be null): 64x4x1 tiles in all */
int vls_batch_layout(vls_batch_t b, int64_t *offsets, int *ntiles);
/* persistent blocks of a tile-loop launch (a multiple of 8; 0 restores the
 * default, 8 per CU); cached graphs are dropped */
int vls_batch_set_blocks(vls_batch_t b, int blocks);
int vls_batch_get_blocks(vls_batch_t b, int *blocks);
/* tile shape of the bulk kernels: 64 (64 x 4 points of a page) or 32 (32 x 8);
 * the default is the one with fewer tiles; cached graphs are dropped */
int vls_batch_set_tile_rows(vls_batch_t b, int tile_rows);
int vls_batch_get_tile_rows(vls_batch_t b, int *tile_rows);
/* pages per work item of the batch's z-marching bulk stage kernels (each
 * block walks one tile column of one grid over `depth` pages; default 8);
 * results do not depend on it.  Synchronizes the device. */
int vls_batch_set_zmarch_depth(vls_batch_t b, int depth);
int vls_batch_get_zmarch_depth(vls_batch_t b, int *depth);

int vls_batch_set_distance(vls_batch_t b, const double *F, vls_stream_t stream);
int vls_batch_weno5_rk3_reinit(vls_batch_t b, double *F, int iterations, vls_stream_t stream);
int vls_batch_weno5_rk3_extend(vls_batch_t b, const double *F, double *C, int iterations, vls_stream_t stream);
/* dt: one time step per grid */
int vls_batch_advect_rk3(vls_batch_t b, double *F, const double *vx, const double *vy, const double *vz,
                         const double *dt, int steps, vls_stream_t stream);
int vls_batch_normal_velocity(vls_batch_t b, const double *F, const double *vx, const double *vy, const double *vz,
                              double *out, vls_stream_t stream);
/* cosine delta/H of half-width eps_factor*dx per grid: out2n[2g] = area, out2n[2g+1] = volume */
int vls_batch_area_volume_cosine(vls_batch_t b, const double *F, double eps_factor, double *out2n,
                                 vls_stream_t stream);
/* dt[g] = cfl*dx_g/max|u| over grid g */
int vls_batch_cfl_dt(vls_batch_t b, const double *vx, const double *vy, const double *vz, double cfl, double *dt,
                     vls_stream_t stream);
/* synthetic inputs: u = curl of sum_m a_m cos(k_m.x + psi_m) for each grid
 * (modes: ngrids*nmodes*7 device doubles k(3), k x a (3), psi), and the
 * ellipsoid field sqrt(x^2/a^2+y^2/b^2+z^2/c^2)-1 (axes: 3 per grid); lo: the
 * (x, y, z) of each grid's first node, 3 per grid; all device arrays */
int vls_batch_curl_velocity(vls_batch_t b, const double *modes, int nmodes, const double *lo, double *vx,
                            double *vy, double *vz, vls_stream_t stream);
int vls_batch_ellipsoid(vls_batch_t b, const double *axes, const double *lo, double *F, vls_stream_t stream);
/* one vesicle step of every grid: RK3 advection of F by u with dt, set
 * distance, Vn = u.N, extension of Vn (ext_iters), reinitialization of F when
 * reinit_iters > 0, cosine area/volume into av (2 per grid) */
int vls_batch_step(vls_batch_t b, double *F, double *Vn, const double *vx, const double *vy, const double *vz,
                   const double *dt, int ext_iters, int reinit_iters, double eps_factor, double *av,
                   vls_stream_t stream);

int vls_batch_set_distance_f32(vls_batch_t b, const float *F, vls_stream_t stream);
int vls_batch_weno5_rk3_reinit_f32(vls_batch_t b, float *F, int iterations, vls_stream_t stream);
int vls_batch_weno5_rk3_extend_f32(vls_batch_t b, const float *F, float *C, int iterations, vls_stream_t stream);
int vls_batch_advect_rk3_f32(vls_batch_t b, float *F, const float *vx, const float *vy, const float *vz,
                             const double *dt, int steps, vls_stream_t stream);
int vls_batch_normal_velocity_f32(vls_batch_t b, const float *F, const float *vx, const float *vy, const float *vz,
                                  float *out, vls_stream_t stream);
int vls_batch_area_volume_cosine_f32(vls_batch_t b, const float *F, double eps_factor, double *out2n,
                                     vls_stream_t stream);
int vls_batch_cfl_dt_f32(vls_batch_t b, const float *vx, const float *vy, const float *vz, double cfl, double *dt,
                         vls_stream_t stream);
int vls_batch_curl_velocity_f32(vls_batch_t b, const double *modes, int nmodes, const double *lo, float *vx,
                                float *vy, float *vz, vls_stream_t stream);
int vls_batch_ellipsoid_f32(vls_batch_t b, const double *axes, const double *lo, float *F, vls_stream_t stream);
int vls_batch_step_f32(vls_batch_t b, float *F, float *Vn, const float *vx, const float *vy, const float *vz,
                       const double *dt, int ext_iters, int reinit_iters, double eps_factor, double *av,
                       vls_stream_t stream);

/* The WENO-RK3 / ENO-RK2 variants the relaxation driver calls every step
 * (p00_test_singlephase_relax.m:76,84,115), on every grid of a batch at once,
 * FP64: WENORK3Reinitialization (+SD/@SDF3/WENORK3Reinitialization.m:3-61),
 * WENORK3Extend (WENORK3Extend.m:9-98) and ENORK2Extend (ENORK2Extend.m:9-81)
 * of C off F.  Each computes the variant's own distances of F, as its MATLAB
 * method does; every grid gets the single-grid result bit for bit
 * (vls_sdf3_wenork3_reinit, vls_sdf3_wenork3_extend, vls_sdf3_enork2_extend).
 * They leave the batch's interface list holding the variant's distances, so
 * vls_batch_set_distance must run again before a WENO5 reinitialization or
 * extension. */
int vls_batch_wenork3_reinit(vls_batch_t b, double *F, int iterations, vls_stream_t stream);
int vls_batch_wenork3_extend(vls_batch_t b, const double *F, double *C, int iterations, vls_stream_t stream);
int vls_batch_enork2_extend(vls_batch_t b, const double *F, double *C, int iterations, vls_stream_t stream);

/* ------------------------------------------------------------------------ */
/* 5. autotuning of the launch shape (no counterpart in the reference)      */
/* ------------------------------------------------------------------------ */
/* Times `iterations` extension iterations of a scratch copy of F, off F's
 * interface, plus as many reinitialization iterations of another copy, for
 * each candidate launch shape and keeps the one with the smallest sum: tile rows
 * 64 or 32 and the z-march depth (a batch: 4, 8, 16 or 32 pages; a single grid:
 * the depth chosen by the box size, half or twice it; read back with
 * vls_batch_get_zmarch_depth / vls_sdf3_get_launch).  A batch's persistent
 * blocks are left at 8 per CU.  Results of
 * every method are the same for every shape.  Synchronizes.  The handle is
 * left with F's distances (as after set_distance(F)); a batch's cached graphs
 * are dropped.  tile_rows and blocks (may be null) receive the choice. */
int vls_batch_autotune(vls_batch_t b, const double *F, int iterations, int *tile_rows, int *blocks,
                       vls_stream_t stream);
int vls_batch_autotune_f32(vls_batch_t b, const float *F, int iterations, int *tile_rows, int *blocks,
                           vls_stream_t stream);
int vls_sdf3_autotune(vls_sdf3_t h, const double *F, int iterations, int *tile_rows, vls_stream_t stream);
int vls_sdf3_autotune_f32(vls_sdf3_t h, const float *F, int iterations, int *tile_rows, vls_stream_t stream);

/* ------------------------------------------------------------------------ */
/* 6. discrete adjoint and tangent (FP64, one grid; no counterpart in the   */
/*    reference, which has no derivatives)                                  */
/* ------------------------------------------------------------------------ */
/* Exact derivatives of the discrete maps computed by vls_sdf3_advect_rk3
 * (one step) and by set_distance + vls_sdf3_weno5_rk3_reinit (the
 * reinitialization of F0 with F0's own distances and mask): WENO weights
 * differentiated exactly, upwind sides, sign tests, Godunov branches and
 * frozen points fixed as the forward pass takes them, subcell distances
 * differentiated through their quintic root.  Boxes need >= 4 points per axis.
 *
 * advect: lam_in = (dF1/dF)^T lam_out; g{x,y,z} += (dF1/du_{x,y,z})^T lam_out
 * (pointwise, accumulated: zero them first).  lam_in may alias lam_out.
 * Tangent: Fout = F1, Fdot_out = (dF1/dF) Fdot + sum_a (dF1/du_a) udot_a. */
int vls_sdf3_adjoint_advect_rk3(vls_sdf3_t h, const double *F, const double *vx, const double *vy, const double *vz,
                                double dt, const double *lam_out, double *lam_in, double *gx, double *gy, double *gz,
                                vls_stream_t stream);
int vls_sdf3_tangent_advect_rk3(vls_sdf3_t h, const double *F, const double *Fdot, const double *vx,
                                const double *vy, const double *vz, const double *udx, const double *udy,
                                const double *udz, double dt, double *Fout, double *Fdot_out, vls_stream_t stream);
/* reinit: lam_in = (dR/dF0)^T lam_out; tangent: Fout = R(F0), Fdot_out = (dR/dF0) F0dot.
 * Both recompute R and leave the handle with F0's distances and mask. */
int vls_sdf3_adjoint_reinit(vls_sdf3_t h, const double *F0, int iterations, const double *lam_out, double *lam_in,
                            vls_stream_t stream);
int vls_sdf3_tangent_reinit(vls_sdf3_t h, const double *F0, const double *F0dot, int iterations, double *Fout,
                            double *Fdot_out, vls_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* VESICLE_LS_H */
