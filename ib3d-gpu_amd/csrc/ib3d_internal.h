// ib3d_internal.h -- device-resident state of the IB3D hot path on MI355X.
//
// One process-wide context mirrors the reference's `global` module
// (modGlobal.f90:1-106): every array lives in HBM for the whole run and the
// host only sees it through explicit upload/download calls.
//
// Layouts
//   * cell-centred and face fields: Fortran order, (nx+2)(ny+2)(nz+2), i
//     fastest, 1-based (i,j,k) -> (i-1) + sx*((j-1) + sy*(k-1)).
//   * tags: int8 (0 fluid, 1 solid, 2 intercepted).
//   * Poisson unknowns: colour-compacted.  A cell (ii,jj,kk) (0-based) has
//     colour c = (ii+jj+kk)&1 (0 = the reference's "red", i+j+k odd in
//     1-based indices) and lives at  c-array[(jj + sy*kk)*H + (ii>>1)],
//     H = (sx+1)/2 rounded up to even (so cell pairs are 16-byte aligned).  Each half-sweep therefore streams two contiguous arrays.
//     The right-hand side b/dt is stored the same way; non-fluid cells hold a
//     NaN sentinel and are never updated (their pc stays 0, as in the
//     reference where only fluid cells are in the red/black lists).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#define IB3D_CHECK(expr)                                                     \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess) {                                              \
            ib3d::set_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
            return -1;                                                       \
        }                                                                    \
    } while (0)

#define IB3D_REQUIRE(cond, msg)                  \
    do {                                         \
        if (!(cond)) {                           \
            ib3d::set_error(msg);                \
            return -1;                           \
        }                                        \
    } while (0)

namespace ib3d {

void set_error(const std::string &msg);

// A forcing plan: the serial image-point updates of velocityForcing1 /
// pressureForcing1 regrouped into dependency levels.
struct ForcePlan {
    int n_updates = 0;
    int n_levels = 0;
    std::vector<int> level_start;      // host: [n_levels+1] offsets into the level-sorted order
    int *upd = nullptr;                // device: packed per-update record (see ib3d_ib.hip)
    int *srcs = nullptr;               // device: 8 read sources per update (-1: the field)
    double *pos = nullptr;             // device: image point per update
    double *res = nullptr;             // device: value produced by each update
    int *final_loc = nullptr;          // device: write location of each final writer
    int *final_upd = nullptr;          // device: the update that writes it last
    int n_final = 0;
    int final_stride = 0;              // distance between the loc | upd | comp columns of final_loc
    int cap_upd = 0, cap_final = 0;    // allocated sizes (the buffers are reused across rebuilds)
};

struct Ctx {
    bool ready = false;
    int64_t nx = 0, ny = 0, nz = 0, sx = 0, sy = 0, sz = 0;
    int64_t P = 0;        // padded cell count
    int64_t H = 0;        // half-row length of the colour arrays
    int64_t NC = 0;       // cells per colour array = sy*sz*H
    int64_t sxp = 0;      // row stride of the x Laplacian rows (sx rounded up to 8)
    hipStream_t stream = nullptr;
    int device = 0;

    // 1-D grid, host copies (1-based, size n+4) and device mirrors
    std::vector<double> hx1, hy1, hz1, hdx, hdy, hdz, hxu, hyu, hzu, hxv, hyv, hzv, hxw, hyw, hzw,
        hxp, hyp, hzp;
    double *x1 = nullptr, *y1 = nullptr, *z1 = nullptr, *dx = nullptr, *dy = nullptr, *dz = nullptr;
    double *xu = nullptr, *yu = nullptr, *zu = nullptr, *xv = nullptr, *yv = nullptr, *zv = nullptr;
    double *xw = nullptr, *yw = nullptr, *zw = nullptr, *xp = nullptr, *yp = nullptr, *zp = nullptr;
    bool grid_ready = false;

    // convection tables (6 distinct sets x 12 coefficients, 1-based)
    std::vector<double> htab[6];
    double *tab[6] = {};
    int64_t tn[6] = {};
    bool coeff_ready = false;
    // momentum tables per axis (index-only sub-expressions and reciprocals,
    // built by ns_momentum from the grid and the convection tables)
    double *mtab[3] = {};
    double *mrec[3] = {};   // y, z: the same per index, with the convection tables
    int64_t mtn[3] = {};
    bool mom_ready = false, mom_fast = false;
    // GMRES-IR iterations replayed as a captured HIP graph (ib3d_set_graphs)
    bool use_graphs = false;

    // Laplacian weights: Acx(n,3) folded with the boundary conditions, kept
    // host-side in the reference's layout and on the device as per-index
    // rows ax1/ax2/ax3 over 0-based padded index (ghost entries 0)
    std::vector<double> hAcx, hAcy, hAcz;
    double *ax = nullptr, *ay = nullptr, *az = nullptr;   // [3][sxp], [3][sy], [3][sz]
    bool matrix_ready = false;

    // fields
    double *u = nullptr, *v = nullptr, *w = nullptr, *ut = nullptr, *vt = nullptr, *wt = nullptr;
    double *p = nullptr, *resi_u = nullptr, *resi_v = nullptr, *resi_w = nullptr;
    int8_t *tag = nullptr;
    int *minElem = nullptr;

    // Poisson (colour compacted)
    double *pc[2] = {}, *bdt[2] = {};
    float *e32[2] = {}, *r32[2] = {};
    double *slots = nullptr;       // per-iteration max|dpc| (reference rule)
    int64_t n_slots = 0;
    double *partial = nullptr;     // per-block partial sums
    int64_t n_partial = 0;
    double bnorm2 = 0.0;           // ||b/dt||^2 of the last computeDiv
    double *scalars = nullptr;     // small device scalars (64)
    double *hscal = nullptr;       // pinned host mirror

    // surface mesh
    int64_t ibNodes = 0, ibElems = 0;
    double *xnode = nullptr, *ynode = nullptr, *znode = nullptr;
    double *xnode1 = nullptr, *ynode1 = nullptr, *znode1 = nullptr;
    int *elp = nullptr;            // [3][ibElems] 1-based node ids
    double *xcent = nullptr, *ycent = nullptr, *zcent = nullptr;
    double *cosA = nullptr, *cosB = nullptr, *cosG = nullptr;
    double *alpha3 = nullptr, *beta3 = nullptr, *gamma3 = nullptr;
    bool norms_ready = false;

    // lists: int32 SoA [3][count], (k,j,i) order like the reference
    int64_t ibCount = 0, fluidCount = 0, solidCount = 0, redCount = 0, blackCount = 0;
    int *ibList = nullptr, *fluidList = nullptr, *solidList = nullptr;
    int64_t cap_lists[3] = {0, 0, 0};  // allocated entries of fluid, solid, ib lists (grow only)
    int64_t cap_nel = 0;               // allocated IB cells of nel / normdis
    bool lists_ready = false;
    int *nel = nullptr;            // one element per intercepted cell (nelp = nelu1 = ... )
    double *normdis = nullptr;     // [7][ibCount]: p, u1, u2, v1, v2, w1, w2
    bool normdist_ready = false;
    ForcePlan vplan, pplan;
    bool plans_ready = false;
    bool plans_pending = false;    // stencils computed and on their way to the host; plans not built yet

    // stress stencil cache [5][3][ibElems]: Elemcell, ucell, vcell, wcell, pcell
    int *scell = nullptr;
    double *scontrib = nullptr;   // per-element stress contributions (7 x ibElems)
    int *body = nullptr;          // body id (1..nbody) of every element
    int nbody = 0;
    bool stress_ready = false;

    // rigid-body motion (zero for a fixed body): wall velocity used by the
    // forcing, the solid cells and the surface forces; body acceleration
    // for dp/dn at the wall; retagging scratch
    double wall_vel[3] = {0.0, 0.0, 0.0}, wall_acc[3] = {0.0, 0.0, 0.0};
    uint8_t *rt_flag = nullptr;    // P flags: cell queued for retagging
    int *rt_list = nullptr;        // unique queued cells (linear index)
    int *rt_count = nullptr;
    int64_t rt_cap = 0;

    // Poisson configuration
    int mode = 0;                  // 0 reference rule, 1 FP64 L2, 2 mixed FP32/FP64, 3 FP32 GMRES in FP64 refinement, 4 the same with a multigrid preconditioner
    int krylov_m = 20;             // modes 3, 4: Arnoldi steps per refinement step
    double krylov_target = 0.3;    // modes 3, 4: a cycle stops once its residual estimate is below this x tol x ||b/dt||
    int mg_nu1 = 4, mg_nu2 = 4;    // mode 4: pre- and post-smoothing sweeps per level
    int mg_nu1c = 4, mg_nu2c = 4;  // mode 4: the same on the levels below the finest
    // mode 4: per-level overrides of (nu1c, nu2c) for levels 1..7 cycled by
    // launches (-1: none)
    int mg_nu_lv[8][2] = {{-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}};
    // mode 4: sweeps before and after on the levels smoothed inside the
    // one-workgroup kernel (-1: nu1c, nu2c)
    int mg_nu_small[2] = {-1, -1};
    double mg_omega = 1.55;        // mode 4: smoother relaxation factor
    int mg_coarse = 4;             // mode 4: sweeps on the coarsest level
    double mg_coarse_omega = 1.5;  // mode 4: relaxation factor there
    int mg_levels = 0;             // mode 4: level cap (0: coarsen while every dimension halves)
    double omega = 1.99;
    double tol = 1e-6;
    int check = 10;                // sweeps between L2 checks (mode 1) / inner sweeps (mode 2)
    int warm = 0;                  // 1: start from the previous pc
    int64_t max_sweeps = 2000000;
    int batch = 32;                // reference rule: iterations launched per host check
    bool have_prev = false;

    // statistics of the last Poisson solve
    int64_t last_sweeps = 0;
    int64_t last_outer = 0;
    int64_t last_cyc1 = 0;  // Arnoldi steps of the previous solve's first cycle (0: unknown)
    int64_t geom_version = 0;      // bumped whenever cellCount rebuilds the lists (the tags changed)
    double last_relres = 0.0, last_r0 = 0.0;
    double last_sor_ms = 0.0, last_poisson_ms = 0.0;
    hipEvent_t ev[5] = {};         // [0..3] solve timing, [4] forcing stencils copied to the host
};

Ctx &ctx();

// launch helpers
inline unsigned nblocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// Division by a launch-invariant divisor d (2 <= d < 2^31) for numerators
// below 2^31: q = umulhi(n, m) >> sh with m = ceil(2^(31+l)/d), l =
// ceil(log2 d) (Granlund-Montgomery; exact because m*d - 2^(31+l) < d <= 2^l).
struct FastDiv {
    uint32_t m = 0;
    int sh = 0;
    int d = 1;
};

// The colour-compacted layout of an nx x ny x nz interior grid: padded
// extents, slots per colour row (H, even), slots per colour (NC) and the x
// coefficient row stride (sxp).  A thread reads the 4 x-coefficient entries
// 2h..2h+3 of slot pair h (h < H), so a row holds at least 2H + 2 entries;
// a multiple of 8 keeps those 4-vectors aligned.
struct ColourLayout {
    int64_t sx, sy, sz, H, NC, sxp;
};
static inline ColourLayout colour_layout(int64_t nx, int64_t ny, int64_t nz) {
    ColourLayout L;
    L.sx = nx + 2; L.sy = ny + 2; L.sz = nz + 2;
    L.H = (((L.sx + 1) / 2) + 1) & ~int64_t(1);
    L.NC = L.sy * L.sz * L.H;
    L.sxp = (std::max<int64_t>(L.sx, 2 * L.H + 2) + 7) & ~int64_t(7);
    return L;
}

static inline FastDiv make_fastdiv(int d) {
    FastDiv f;
    int l = 0;
    while ((1ll << l) < d) l++;
    const int s = 31 + l;
    f.m = (uint32_t)(((1ull << s) + (uint64_t)d - 1) / (uint64_t)d);
    f.sh = s - 32;
    f.d = d;
    return f;
}

__host__ __device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv &f) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umulhi(n, f.m) >> f.sh;
#else
    return (uint32_t)(((uint64_t)n * f.m) >> (32 + f.sh));
#endif
}

// max that keeps a NaN operand (fmax would drop it)
__device__ inline double nan_max(double a, double b) { return (a != a || b != b) ? __builtin_nan("") : fmax(a, b); }

// max over a 256-thread workgroup into partial[blockIdx.x] (no atomics: one
// contended address per launch serialises tens of thousands of workgroups)
__device__ inline void block_max_partial(double v, double *partial) {
    for (int off = 32; off > 0; off >>= 1) v = nan_max(v, __shfl_xor(v, off, 64));
    __shared__ double sm_max[4];
    if ((threadIdx.x & 63) == 0) sm_max[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        partial[blockIdx.x] = nan_max(nan_max(sm_max[0], sm_max[1]), nan_max(sm_max[2], sm_max[3]));
}
// *out = max of partial[0..n) (NaN if any is NaN), on the library stream
void launch_max_partials(const double *partial, int64_t n, double *out);

// Every transfer and fill goes through the library's stream, which is
// created non-blocking and therefore not ordered with the null stream: a
// blocking hipMemcpy/hipMemset would race with kernels queued on it.  These
// return after the operation has completed.
inline hipError_t h2d(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return hipSuccess;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx().stream);
    return e != hipSuccess ? e : hipStreamSynchronize(ctx().stream);
}
inline hipError_t d2h(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return hipSuccess;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx().stream);
    return e != hipSuccess ? e : hipStreamSynchronize(ctx().stream);
}
inline hipError_t dzero(void *dst, size_t bytes) {
    hipError_t e = hipMemsetAsync(dst, 0, bytes, ctx().stream);
    return e != hipSuccess ? e : hipStreamSynchronize(ctx().stream);
}
int ensure_slots(int64_t n);
int ensure_partial(int64_t n);

// modules
int fields_init(Ctx &c);
int poisson_solve(Ctx &c, double deltat, double eps1, int64_t *niter, double *divmax, double *derr_std);
// device scalar slots 40, 41: ||b/dt|| and the Krylov target of a solve
// queued before the host has read the norm
static constexpr int kNormSlot = 40;
int compute_div(Ctx &c, double deltat, double *divmax, double *bnorm2, bool wait = true);
int norm_target_async(Ctx &c, const double *bnorm2, double *out);
int compute_div_noalloc(Ctx &c, double deltat, double *divmax);
int noalloc_setup(int device);
int set_fused_sweeps(int on, int tz, int blocks, int coarse_blocks);
int set_coarse_tiles(int from_level);
int set_fused_restrict(int levels);
int64_t resid_guard_selfcheck(int64_t nx, int64_t ny, int64_t nz, bool drop_interior);
int set_momentum_tiles(int on);
int redblack_sor_reference(Ctx &c, double epsi, int64_t *isum, double *derr2);
int plans_launch(Ctx &c);
int shift_nodes_moving(Ctx &c, const double base[3], const double disp[3]);
int selective_retagging(Ctx &c, int32_t counts[3]);
int stress_bodies(Ctx &c, double *coef);
int velocity_bc(Ctx &c);

// learned initial guess (ib3d_cnn.hip)
bool cnn_active();
int cnn_before_solve(Ctx &c, const double *rel_prev, double *rel_guess, int (*resid)(Ctx &, double *));
int cnn_after_solve(Ctx &c, const double *rel_prev, const double *rel_guess);
void cnn_release();

}  // namespace ib3d
