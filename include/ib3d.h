/*
 * ib3d.h -- C ABI of the MI355X-native IB3D hot path.
 *
 * The reference exposes its hot path as argument-less Fortran subroutines
 * that share the `global` module (modGlobal.f90:1-106).  This library keeps
 * one device-resident copy of that module per process and exports one entry
 * point per subroutine; module scalars the subroutine would read are passed
 * by value, host-visible results come back through pointers.  A drop-in
 * Fortran shim (ib3d-gpu_amd/fortran/ib3d_dropin.f90) binds these with
 * iso_c_binding under the reference's own subroutine names.
 *
 * Conventions
 *   - every function returns 0 on success and -1 on error; the message is
 *     available from ib3d_last_error();
 *   - plain pointers and sizes only; arrays in the reference's Fortran
 *     layout (column-major, padded (nx+2)(ny+2)(nz+2), i fastest);
 *   - one host thread, one HIP stream; a call returns once its host-visible
 *     outputs are valid (device work it queued may still be running until the
 *     next synchronising call, exactly like the fields it leaves in HBM);
 *   - nothing here routes through a CPU implementation: without a GPU the
 *     calls fail.
 */
#ifndef IB3D_H
#define IB3D_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- lifecycle ---------------------------------------------------------- */
/* Allocates the device copy of the module for an nx*ny*nz interior grid
 * (replaces allocateArrays, subAllocateArrays.f90:1-15, plus the OpenACC
 * `declare create` mirrors, modGlobal.f90:67-105). */
int ib3d_init(int64_t nx, int64_t ny, int64_t nz);
int ib3d_finalize(void);
const char *ib3d_last_error(void);
int ib3d_sync(void);

/* Host <-> device copies of a module array by its reference name, e.g.
 * "u", "ut", "p", "pc", "cell", "x1", "deltax", "xp", "Acx",
 * "fluidIndexPtr", "interceptedIndexPtr", "nelp", "pNormDis", "xcent",
 * "Elemcell" ... (replaces `!$acc update host/device`, e.g. main.f90:55).
 * `nbytes` must equal the reference array's size in its reference kind
 * (REAL(8) / INTEGER(8)); conversions to the device layout happen inside. */
int ib3d_upload(const char *name, const void *host, int64_t nbytes);
int ib3d_download(const char *name, void *host, int64_t nbytes);
/* Element count of a named array as the reference allocates it (0 if the
 * array does not exist yet, e.g. lists before cellCount). */
int64_t ib3d_array_count(const char *name);

/* ---- setup (host I/O stays with the caller) ----------------------------- */
/* Grid nodes as read from gridx/y/z.dat (nx+1, ny+1, nz+1 values): builds
 * deltax.., ghost nodes and the staggered coordinates exactly as
 * readInput does (subReadInput.f90:56-129). */
int ib3d_set_grid(const double *gx, const double *gy, const double *gz);
/* Surface mesh as read by readSurfaceMeshGmsh (subReadInput.f90:147-167):
 * node coordinates and 1-based triangle connectivity. */
int ib3d_set_mesh(int64_t n_nodes, const double *xnode, const double *ynode, const double *znode,
                  int64_t n_elems, const int64_t *p1, const int64_t *p2, const int64_t *p3);

/* subSearch.f90:2-39 */
int ib3d_shiftSurfaceNodesInitial(double xshift, double yshift, double zshift);
/* Moving body.  The reference's shiftSurfaceNodesMovingbody (subSearch.f90:
 * 41-76) integrates a free motion; this entry point places the body at a
 * prescribed rigid displacement (dx,dy,dz) from its initial placement:
 * xnode1 = (xnode + xshift) + dx.  The per-step sequence is then
 * computeSurfaceNorm, selectiveRetagging, cellCount, computeNormDistance. */
int ib3d_shiftSurfaceNodesMovingbody(double xshift, double yshift, double zshift, double dx, double dy,
                                     double dz);
/* subSearch.f90:327-512: re-tag the 6x6x6 block around every intercepted
 * cell of the current lists (cells that stop being solid take the wall
 * velocity); counts = {ib, fluid, solid} */
int ib3d_selectiveRetagging(int32_t counts[3]);
/* Wall velocity (used instead of the reference's usurf = vsurf = wsurf = 0,
 * subForcing.f90:144-145, in the ghost-cell fit, the solid cells and the
 * surface shear) and body acceleration (dp/dn = -a.n at the wall instead of
 * 0, subForcing.f90:11 and 20).  Zero keeps the reference's fixed body bit for
 * bit. */
int ib3d_set_wall_motion(const double vel[3], const double acc[3]);
/* Several bodies in one surface mesh: body id (1..n_bodies) of every
 * element, in mesh order.  No counterpart in the reference, whose stressCal
 * sums over all elements. */
int ib3d_set_element_bodies(const int32_t *body, int64_t n_elems, int32_t n_bodies);
/* Per-body coefficients of the last stressCal1/stressCal2 call: for body k,
 * coef[4k..4k+3] = {viscous Cd, pressure Cd, viscous Cl, pressure Cl} with the
 * sums of subStressCal1.f90:327-331 and the normalisation of 441-450
 * restricted to the body's elements and projected areas. */
int ib3d_stress_bodies(double *coef);
/* subSearch.f90:80-131; inor is the normal sign read from inputdata */
int ib3d_computeSurfaceNorm(int64_t inor);
/* subSearch.f90:134-321; window = {i_start,i_end,j_start,j_end,k_start,k_end}
 * (1-based, inclusive); counts = {ibCellCount, fluidCellCount, solidCellCount} */
int ib3d_tagging(const int64_t window[6], int32_t counts[3]);
/* subSearch.f90:515-596; counts = {ib, fluid, solid, red, black} */
int ib3d_cellCount(int32_t counts[5]);
/* subSearch.f90:599-678.  Also starts the forcing plans: their stencils are
 * computed here and the plans are finished by the next velocityForcing1 or
 * pressureForcing1, which also reports a stencil error. */
int ib3d_computeNormDistance(void);
/* subInitialConditions.f90:2-48 */
int ib3d_initialConditions(void);
/* subCoefficientMatrix.f90:1-59 */
int ib3d_coefficientMatrix(void);
/* subNavierStokes.f90:3-619 */
int ib3d_non_uni_coeff(void);

/* ---- per-step kernels --------------------------------------------------- */
/* subNavierStokes.f90:622-1122 */
int ib3d_nsMomentum(double rev, double deltat);
/* subBoundaryConditions.f90:2-195 (the live branch, 137-193) */
int ib3d_velocityBC(void);
/* subBoundaryConditions.f90:199-217 */
int ib3d_solidCellBC(void);
/* subForcing.f90:107-652 */
int ib3d_velocityForcing1(void);
/* subForcing.f90:2-103 */
int ib3d_pressureForcing1(void);
/* subPcorVcor.f90:82-101: b = div(ut) on fluid cells, divmax = max|b| */
int ib3d_computeDiv(double deltat, double *divmax);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ib3d_computeDiv without any allocation.  Same inputs,
 * outputs and kernels; the per-block partial maxima and sums live in a static
 * device array of 2^19 doubles (4 MiB) per GPU, held until the process ends,
 * so grids above about 67 million padded cells are rejected with an error.
 * Calls that overlap on one GPU share that array and race. */
int ib3d_computeDiv_noalloc(double deltat, double *divmax);
/* subPcorVcor.f90:148-212: REDBLACKSOR(epsi, isum, derr, derr2) with the
 * reference rule (omega 1.99, stop when max|pc-pco| < epsi).  Continues from
 * the current pc (poissonSolver zeroes it first). */
int ib3d_REDBLACKSOR(double epsi, int64_t *isum, double *derr, double *derr2);
/* subPcorVcor.f90:105-120 */
int ib3d_correctPressure(void);
/* subPcorVcor.f90:125-142 */
int ib3d_correctVelocity(double deltat);
/* subPcorVcor.f90:2-80: zero pc, computeDiv, SOR, corrections, velocityBC,
 * steady-state check, u=ut.  sor_seconds is the SOR-only time (the
 * reference's cpu_time bracket, subPcorVcor.f90:33-37). */
int ib3d_poissonSolver(double deltat, double eps1, int64_t *nIterPcor, double *divmax,
                       double *derrStdSt, double *sor_seconds);
/* subStressCal1.f90:2-460 / subStressCal2.f90:2-381.
 * coef = {Cd_visc, Cd_press, Cl_visc, Cl_press} */
int ib3d_stressCal1(double re, double coef[4]);
int ib3d_stressCal2(double re, double coef[4]);

/* ---- north-star solver modes (no reference counterpart) ------------------
 * mode 0: the reference rule (above).
 * mode 1: FP64 RB-SOR with the given omega until ||b/dt - A pc||_2 <=
 *         tol*||b/dt||_2, checked every `check` sweeps.
 * mode 2: FP32 RB-SOR sweeps (`check` per outer iteration) on the defect
 *         equation inside an FP64 defect-correction loop with the same
 *         FP64 stopping rule.
 * mode 3: FP64 refinement (GMRES-IR): each outer iteration solves the
 *         defect equation with FP32 flexible GMRES, every Arnoldi step
 *         preconditioned by `check` FP32 RB-SOR sweeps at `omega`; same FP64
 *         stopping rule.
 * mode 4: as mode 3, every Arnoldi step preconditioned by one FP32
 *         geometric-multigrid V-cycle instead (ib3d_set_multigrid); `check`
 *         and `omega` are unused.
 * warm != 0 starts each solve from the previous step's pc instead of 0. */
int ib3d_set_poisson_mode(int mode, double omega, double tol, int check, int warm,
                          int64_t max_sweeps);
/* modes 3 and 4: Arnoldi steps per refinement step, 1..23 (default 20) */
int ib3d_set_krylov(int restart);
/* mode 4 V-cycle: nu1 >= 1 pre- and nu2 >= 0 post-smoothing red-black SOR
 * sweeps at `omega` on every level but the coarsest, `coarse_sweeps` at
 * `coarse_omega` there; max_levels caps the hierarchy (0: halve every
 * interior dimension while each stays even and >= 4).  Defaults 4, 4, 1.55,
 * 4, 1.5, 0. */
int ib3d_set_multigrid(int nu1, int nu2, double omega, int coarse_sweeps, double coarse_omega,
                       int max_levels);
/* modes 3 and 4: a Krylov cycle stops early once its least-squares residual
 * estimate is below target x tol x ||b/dt|| (default 0.3); the FP64 residual
 * of the refinement then decides whether another cycle runs. */
int ib3d_set_krylov_target(double target);
/* mode 4: pre- and post-smoothing sweeps on the levels below the finest
 * (ib3d_set_multigrid sets both the finest and these; call this after it).
 * Those levels are small enough that a sweep costs about a launch, whatever
 * the level's size. */
int ib3d_set_multigrid_coarse(int nu1, int nu2);
/* mode 4: pre- and post-smoothing sweeps on one level (1..7) cycled by
 * launches, overriding the coarse levels' counts there; (-1, -1) removes the
 * override.  ib3d_set_multigrid removes every override.  A sweep of level 1
 * (one eighth of the finest level) is still one launch of real work, so it
 * may pay to sweep it less than the smaller levels. */
int ib3d_set_multigrid_level(int level, int nu1, int nu2);
/* mode 4: pre- and post-smoothing sweeps on the coarsest levels, those whose
 * arrays together fit in 64 KB of LDS and are cycled inside one workgroup
 * (-1, -1: the coarse levels' counts; ib3d_set_multigrid resets it).  Their
 * sweeps cost a barrier each, not a launch, so more of them can pay. */
int ib3d_set_multigrid_small(int nu1, int nu2);
/* mode 4: smooth the first `levels` multigrid levels (2, the default: the
 * finest and the next) with whole red-black sweeps in one launch each, the
 * others with two half-sweep launches (0: none), unless they are tiled
 * (ib3d_set_coarse_tiles).  Both give identical results.  planes: z planes
 * per workgroup on the finest level (0: chosen so that the grid holds about
 * `blocks` workgroups); below it the grid holds about `coarse_blocks`
 * (defaults 0, 480 and 192). */
int ib3d_set_fused_sweeps(int levels, int planes, int blocks, int coarse_blocks);
/* mode 4: smooth the multigrid levels from `from_level` on (2, the
 * default; 0 for none) that are cycled by launches with two temporally
 * blocked launches per smoothing phase: each workgroup sweeps an 8x8x8 cube
 * plus a halo in LDS, ceil(n/2) sweeps in the first launch and floor(n/2) in
 * the second (n = 2..4), instead of 2n half-sweep launches.  Both give
 * identical results. */
int ib3d_set_coarse_tiles(int from_level);
/* mode 4: on the first `levels` multigrid levels (64, the default: every
 * level cycled by launches; 0 for none) form the residual and restrict it to
 * the next level's right-hand side in one z-marching launch, instead of
 * storing the residual (finest level) or gathering each child's stencil per
 * coarse slot (below).  Both give identical results. */
int ib3d_set_fused_restrict(int levels);
/* No counterpart in the reference; needs no GPU.  Host self-check of the
 * residual kernel's guard on an nx x ny x nz grid: with a finite b/dt in
 * every slot (an overwritten NaN sentinel), counts the neighbour and
 * coefficient reads of admitted slots that fall outside the arrays (0 is
 * correct).  drop_interior != 0 checks the guard without its interior
 * test instead, which reads outside. */
int ib3d_selfcheck_resid_guard(int64_t nx, int64_t ny, int64_t nz, int drop_interior, int64_t *outside);
/* nsMomentum's interior columns: on == 0 (the default) reads every stencil
 * value from HBM through 32-bit byte offsets (64-bit indices when the FP64
 * fields span 4 GiB or more); 1, 2, 3 march LDS tiles of 64 x 8, 64 x 4 and
 * 64 x 8 columns (3 and 4 waves per SIMD) in z; 4 reads from HBM with
 * 64-bit indices.  All give identical results. */
int ib3d_set_momentum_tiles(int on);
/* modes 3 and 4: replay each refinement iteration (Krylov cycle, FP64
 * residual and its norm) as one captured HIP graph (on != 0) or launch its
 * kernels one by one (on == 0, the default).  Both run the same kernels in
 * the same order and give identical results.  No counterpart in the
 * reference, which has no GPU launch layer. */
int ib3d_set_graphs(int on);
/* Statistics of the last solve: sweeps (SOR sweeps of either precision;
 * V-cycles in mode 4),
 * outer iterations, final and initial relative residual, SOR ms, whole
 * poissonSolver ms. */
int ib3d_poisson_stats(int64_t *sweeps, int64_t *outer, double *relres, double *relres0,
                       double *sor_ms, double *poisson_ms);
/* FP64 residual of the current pc: ||b/dt - A pc||_2 and ||b/dt||_2 */
int ib3d_residual_norm(double *rnorm, double *bnorm);

/* One whole time step on the device in the reference's order
 * (main.f90:156-174): nsMomentum, velocityBC, solidCellBC, velocityForcing1,
 * velocityBC, poissonSolver, pressureForcing1, stressCal (first != 0 ->
 * stressCal1).  out = {nIterPcor, divmax, derrStdSt, Cd_visc, Cd_press,
 * Cl_visc, Cl_press, sor_seconds}. */
int ib3d_step(int first, double rev, double re, double deltat, double eps1, double out[8]);

/* Kernel micro-benchmark used for the roofline figure: `sweeps` RB-SOR
 * sweeps (precision 64 or 32: two half-sweep launches each; 33: FP32, one
 * launch per sweep, as on the finest multigrid level) on the current
 * right-hand side, timed with HIP events on the library stream.
 * ms_per_sweep = average duration of one red+black sweep. */
int ib3d_bench_sor(int precision, int sweeps, double omega, double *ms_per_sweep);
/* Diagnostics (no reference counterpart): one streaming device copy of
 * `nbytes` (read nbytes + write nbytes), the calibration point for turning
 * hardware fetch/write counters into bytes; ms = its duration. */
int ib3d_probe_copy(int64_t nbytes, double *ms);

/* ---- learned initial guess (no counterpart in the reference) -------------
 * A 3-D CNN (3 -> 16 -> 16 -> 16 -> 1 channels, 3x3x3, ReLU) predicts the
 * start of each L2-mode Poisson solve from the warm start's defect
 * r0 = b/dt - A p_prev (scaled by 1/|A_diag|), the fluid mask and p_prev.
 * The solve keeps the prediction only when its FP64 residual is below the
 * previous-step guess's and iterates to the same tolerance either way.
 * Weights are BF16 on MFMA with FP32 accumulation; the FP32 master copy is
 * trained by Adam in the library's own kernels. */
/* allocate the 15601 parameters (He-normal weights from `seed`, zero biases) */
int ib3d_cnn_init(uint64_t seed);
int ib3d_cnn_release(void);
int64_t ib3d_cnn_param_count(void);
/* parameter vector: per layer W[tap][cin][cout] (tap = kz*9 + ky*3 + kx,
 * offset -1..1 each) then bias[cout]; layers 3->16, 16->16, 16->16, 16->1 */
int ib3d_cnn_set_params(const float *params, int64_t n);
int ib3d_cnn_get_params(float *params, int64_t n);
/* forward on nb volumes (nb x vz x vy x vx x 3 FP32, channels last, zero
 * padding outside each volume) -> nb x vz x vy x vx FP32 */
int ib3d_cnn_forward(const float *x, int64_t nb, int64_t vx, int64_t vy, int64_t vz, float *y);
/* gradient of sum(dy * forward(x)) with respect to the parameters */
int ib3d_cnn_gradient(const float *x, const float *dy, int64_t nb, int64_t vx, int64_t vy, int64_t vz,
                      float *grad);
/* one Adam step with the given gradient */
int ib3d_cnn_adam(const float *grad, double lr, double beta1, double beta2, double eps);
/* record = 1: every L2-mode solve stores (input, converged correction) as a
 * training pair on the device; guess = 1: every L2-mode solve starts from the
 * network's guess when it lowers the residual */
int ib3d_cnn_config(int record, int guess);
/* The guess is kept when its relative residual is below ratio x the
 * previous-step guess's, ratio in (0, 1].  1 (the default) keeps any guess
 * that lowers the residual.  A smaller ratio asks for a margin: with the
 * multigrid solver (mode 4), a guess that lowers the residual norm by less
 * than about 1.5x can make the solve need more Krylov steps, not fewer. */
int ib3d_cnn_set_accept(double ratio);
int64_t ib3d_cnn_pair_count(void);
int ib3d_cnn_clear_pairs(void);
/* Adam over 64^3 tiles (halo 4) of the recorded pairs; mean loss per epoch */
int ib3d_cnn_train(int epochs, int batch, double lr, uint64_t seed, double *loss_per_epoch);
/* solves that used the network, how many kept its guess, and the last
 * solve's relative residuals of the previous-step guess and of the network's
 * guess, and the guess's cost in ms */
int ib3d_cnn_stats(int64_t *solves, int64_t *kept, double *rel_prev, double *rel_guess, double *guess_ms);

#ifdef __cplusplus
}
#endif
#endif
