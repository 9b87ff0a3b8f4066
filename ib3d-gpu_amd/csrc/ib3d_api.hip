// ib3d_api.hip -- C ABI entry points (include/ib3d.h) and the device copy of
// the reference's `global` module (modGlobal.f90:1-106).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/ib3d.h"
#include "ib3d_internal.h"

namespace ib3d {

static Ctx g_ctx;
static std::string g_err;
static thread_local char g_err_buf[512];

Ctx &ctx() { return g_ctx; }
void set_error(const std::string &msg) { g_err = msg; }

// defined in the other translation units
int ns_momentum(Ctx &c, double rev, double deltat);
int solid_bc(Ctx &c);
int correct(Ctx &c, double dt, int do_p, int do_v);
int steady_copy(Ctx &c, double dt, double *derr_std);
int sor_solve(Ctx &c, double bnorm2, double eps1, const double *dnorm = nullptr);
int residual_norm(Ctx &c, double *rnorm, double *bnorm);
int bench_sor(Ctx &c, int precision, int sweeps, double omega, double *ms);
int probe_copy(Ctx &c, int64_t nbytes, double *ms);
int sor_float_tables(Ctx &c);
void sor_release_tables();
int check_fastdiv(int d, int64_t nmax);
int shift_nodes(Ctx &c, double xs, double ys, double zs);
int surface_norm(Ctx &c, int64_t inor);
int tagging(Ctx &c, const int64_t w[6], int32_t counts[3]);
int cell_count(Ctx &c, int32_t counts[5]);
int norm_distance(Ctx &c);
int velocity_forcing(Ctx &c);
int pressure_forcing(Ctx &c);
int stress(Ctx &c, int first, double re, double coef[4]);
void release_plans(Ctx &c);
void ib_release();

template <typename T>
static int dalloc(T **p, int64_t n) {
    IB3D_CHECK(hipMalloc((void **)p, std::max<int64_t>(1, n) * sizeof(T)));
    IB3D_CHECK(dzero(*p, std::max<int64_t>(1, n) * sizeof(T)));
    return 0;
}
template <typename T>
static void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

__global__ void k_fill_nan64(double *a, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) a[q] = __builtin_nan("");
}
__global__ void k_fill_nan32(float *a, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) a[q] = __builtin_nanf("");
}

// natural <-> colour-compacted copies of pc
__global__ void k_gather_pc(const double *p0, const double *p1, double *nat, int sx, int sy, int sz, int64_t H) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t P = (int64_t)sx * sy * sz;
    if (c >= P) return;
    const int ii = (int)(c % sx);
    const int64_t row = c / sx;
    const int jj = (int)(row % sy), kk = (int)(row / sy);
    const int64_t f = row * H + (ii >> 1);
    nat[c] = ((ii + jj + kk) & 1) ? p1[f] : p0[f];
}
__global__ void k_scatter_pc(double *p0, double *p1, const double *nat, int sx, int sy, int sz, int64_t H) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t P = (int64_t)sx * sy * sz;
    if (c >= P) return;
    const int ii = (int)(c % sx);
    const int64_t row = c / sx;
    const int jj = (int)(row % sy), kk = (int)(row / sy);
    const int64_t f = row * H + (ii >> 1);
    (((ii + jj + kk) & 1) ? p1 : p0)[f] = nat[c];
}

// subInitialConditions.f90:9-42
__global__ void k_initial(double *u, const int8_t *cell, int nx, int ny, int nz) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t P = (int64_t)(nx + 2) * (ny + 2) * (nz + 2);
    if (c >= P) return;
    const int ii = (int)(c % (nx + 2));
    const int64_t row = c / (nx + 2);
    const int jj = (int)(row % (ny + 2)), kk = (int)(row / (ny + 2));
    const bool interior = ii >= 1 && ii <= nx && jj >= 1 && jj <= ny && kk >= 1 && kk <= nz;
    u[c] = (interior && cell[c] == 0) ? 1.0 : 0.0;
}

static void free_all(Ctx &c) {
    double **d1[] = {&c.x1, &c.y1, &c.z1, &c.dx, &c.dy, &c.dz, &c.xu, &c.yu, &c.zu, &c.xv, &c.yv, &c.zv,
                     &c.xw, &c.yw, &c.zw, &c.xp, &c.yp, &c.zp, &c.ax, &c.ay, &c.az, &c.u, &c.v, &c.w,
                     &c.ut, &c.vt, &c.wt, &c.p, &c.resi_u, &c.resi_v, &c.resi_w, &c.pc[0], &c.pc[1],
                     &c.bdt[0], &c.bdt[1], &c.slots, &c.partial, &c.scalars, &c.xnode, &c.ynode, &c.znode,
                     &c.xnode1, &c.ynode1, &c.znode1, &c.xcent, &c.ycent, &c.zcent, &c.cosA, &c.cosB,
                     &c.cosG, &c.alpha3, &c.beta3, &c.gamma3, &c.normdis};
    for (auto pp : d1) dfree(*pp);
    for (int t = 0; t < 6; t++) dfree(c.tab[t]);
    for (int t = 0; t < 3; t++) {
        dfree(c.mtab[t]);
        dfree(c.mrec[t]);
    }
    c.mtn[0] = c.mtn[1] = c.mtn[2] = 0;
    c.mom_ready = false;
    dfree(c.e32[0]); dfree(c.r32[0]);   // colour 1 lives in the same blocks
    c.e32[1] = nullptr;
    c.r32[1] = nullptr;
    dfree(c.tag); dfree(c.minElem); dfree(c.elp);
    dfree(c.ibList); dfree(c.fluidList); dfree(c.solidList);
    c.cap_lists[0] = c.cap_lists[1] = c.cap_lists[2] = 0;
    c.cap_nel = 0; dfree(c.nel); dfree(c.scell); dfree(c.scontrib);
    dfree(c.rt_flag); dfree(c.rt_list); dfree(c.rt_count); dfree(c.body);
    release_plans(c);
    if (c.hscal) (void)hipHostFree(c.hscal);
    c.hscal = nullptr;
    for (auto &e : c.ev)
        if (e) (void)hipEventDestroy(e);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    sor_release_tables();
    ib_release();
    c = Ctx();
}

// --------------------------------------------------------------- grid setup
// subReadInput.f90:56-129
static void grid_axis(int64_t n, const double *g, std::vector<double> &x1, std::vector<double> &d) {
    x1.assign(n + 4, 0.0);
    d.assign(n + 4, 0.0);
    for (int64_t i = 2; i <= n + 2; i++) x1[i] = g[i - 2];
    for (int64_t i = 2; i <= n + 1; i++) d[i] = x1[i + 1] - x1[i];
    d[1] = d[2];
    d[n + 2] = d[n + 1];
    x1[1] = x1[2] - d[1];
    x1[n + 3] = x1[n + 2] + d[n + 2];
}

static int upload_vec(double *dst, const std::vector<double> &v) {
    IB3D_CHECK(h2d(dst, v.data(), v.size() * sizeof(double)));
    return 0;
}

static int push_grid(Ctx &c) {
    const int64_t nx = c.nx, ny = c.ny, nz = c.nz;
    c.hxu.assign(nx + 4, 0.0); c.hyv.assign(ny + 4, 0.0); c.hzw.assign(nz + 4, 0.0);
    c.hyu.assign(ny + 4, 0.0); c.hyw.assign(ny + 4, 0.0); c.hyp.assign(ny + 4, 0.0);
    c.hxv.assign(nx + 4, 0.0); c.hxw.assign(nx + 4, 0.0); c.hxp.assign(nx + 4, 0.0);
    c.hzu.assign(nz + 4, 0.0); c.hzv.assign(nz + 4, 0.0); c.hzp.assign(nz + 4, 0.0);
    for (int64_t i = 1; i <= nx + 3; i++) c.hxu[i] = c.hx1[i];
    for (int64_t i = 1; i <= ny + 3; i++) c.hyv[i] = c.hy1[i];
    for (int64_t i = 1; i <= nz + 3; i++) c.hzw[i] = c.hz1[i];
    for (int64_t i = 1; i <= ny + 2; i++) {
        c.hyu[i] = 0.5 * (c.hy1[i] + c.hy1[i + 1]);
        c.hyw[i] = c.hyu[i];
        c.hyp[i] = c.hyu[i];
    }
    for (int64_t i = 1; i <= nx + 2; i++) {
        c.hxv[i] = 0.5 * (c.hx1[i] + c.hx1[i + 1]);
        c.hxw[i] = c.hxv[i];
        c.hxp[i] = c.hxv[i];
    }
    for (int64_t i = 1; i <= nz + 2; i++) {
        c.hzu[i] = 0.5 * (c.hz1[i] + c.hz1[i + 1]);
        c.hzv[i] = c.hzu[i];
        c.hzp[i] = c.hzu[i];
    }
    if (upload_vec(c.x1, c.hx1) || upload_vec(c.y1, c.hy1) || upload_vec(c.z1, c.hz1) || upload_vec(c.dx, c.hdx) ||
        upload_vec(c.dy, c.hdy) || upload_vec(c.dz, c.hdz) || upload_vec(c.xu, c.hxu) || upload_vec(c.yu, c.hyu) ||
        upload_vec(c.zu, c.hzu) || upload_vec(c.xv, c.hxv) || upload_vec(c.yv, c.hyv) || upload_vec(c.zv, c.hzv) ||
        upload_vec(c.xw, c.hxw) || upload_vec(c.yw, c.hyw) || upload_vec(c.zw, c.hzw) || upload_vec(c.xp, c.hxp) ||
        upload_vec(c.yp, c.hyp) || upload_vec(c.zp, c.hzp))
        return -1;
    c.grid_ready = true;
    return 0;
}

// --------------------------------------------------------- coefficient tables
// Real powers with integral exponents, evaluated as the reference build
// (amdflang, -fdefault-real-8) does: x**2.0 -> x*x, x**3.0 -> x*(x*x),
// x**4.0 -> pow(x, 4.0).
static inline double P2(double x) { return x * x; }
static inline double P3(double x) { return x * (x * x); }
static inline double P4(double x) { return std::pow(x, 4.0); }

// subNavierStokes.f90:36-116 for one spacing triple
static void kk_coeffs(double t1, double t2, double t3, double out[12]) {
    const double f_1 = t3;
    const double f_2 = 2.0 * t3 + P2(t3);
    const double f_3 = (3.0 * t3 + 3.0 * P2(t3)) + P3(t3);
    const double f_4 = ((4.0 * t3 + 6.0 * P2(t3)) + 4.0 * P3(t3)) + P4(t3);
    const double f_5 = 1.0 / (t1 * t2);
    const double f_6 = (2.0 * t1 + 1.0) / (P2(t1) * P2(t2));
    const double f_7 = ((3.0 * P2(t1) + 3.0 * t1) + 1.0) / (P3(t1) * P3(t2));
    const double f_8 = (((4.0 * P3(t1) + 6.0 * P2(t1)) + 4.0 * t1) + 1.0) / (P4(t1) * P4(t2));
    const double s51 = -1.0 / P4(t2), s52 = f_4, s53 = f_4, s54 = s51;
    const double s55 = -(f_1 / P4(t2) + f_4 / t2);
    const double s56 = (f_4 / P2(t2) - f_2 / P4(t2));
    const double s57 = -(f_3 / P4(t2) + f_4 / P3(t2));
    const double s61 = -1.0, s62 = f_8, s63 = f_8, s64 = -1.0;
    const double s65 = (f_5 + f_8), s66 = (f_8 - f_6), s67 = (f_7 + f_8);
    const double s71 = -1.0, s72 = (1.0 + f_4), s73 = f_4, s74 = (f_4 - f_1), s75 = (f_4 - f_2), s76 = (f_4 - f_3);
    const double s81 = -1.0 / P4(t2), s82 = (f_8 + (1.0 / P4(t2))), s83 = f_8;
    const double s84 = (f_5 / P4(t2) - f_8 / t2);
    const double s85 = (f_8 / P2(t2) - f_6 / P4(t2));
    const double s86 = (f_7 / P4(t2) - f_8 / P3(t2));
    const double s91 = -s66 * s51, s92 = s56 * s61, s93 = -(s54 * s66 + s56 * s62);
    const double s94 = (s56 * s64 + s66 * s52), s95 = (s56 * s63 - s66 * s53);
    const double s96 = (s56 * s65 - s66 * s55), s97 = (s56 * s67 - s66 * s57);
    const double s101 = -s85 * s71, s102 = -s85 * s72, s103 = s75 * s81, s104 = s75 * s82;
    const double s105 = (s75 * s83 - s85 * s73), s106 = (s75 * s84 - s85 * s74), s107 = (s75 * s86 - s85 * s76);
    out[0] = (s97 * s101 - s107 * s91);
    out[1] = (s97 * s102 + s107 * s93);
    out[2] = (s95 * s107 - s105 * s97);
    out[3] = (s94 * s107 + s104 * s97);
    out[4] = (s97 * s103 - s107 * s92);
    out[5] = (s97 * s106 - s107 * s96);
    const double ak_1 = ((1.0 + 2.0 * t1) * (t3 + P2(t3))) * t2;
    const double ak_2 = (1.0 + t1) * (2.0 * t3 + P2(t3));
    const double ak_3 = (1.0 + t1);
    const double ak_4 = (1.0 + t1) * P2(1.0 + t3);
    const double ak_5 = (P2(1.0 + t1) * (t3 + P2(t3))) * t2;
    const double ak_6 = (P2(t1) * t2) * (t3 + P2(t3));
    const double ak_7 = ((1.0 + t1) * (t3 + P2(t3))) * t2;
    out[6] = ak_3;
    out[7] = -ak_4;
    out[8] = (ak_1 + ak_2);
    out[9] = -ak_5;
    out[10] = ak_6;
    out[11] = ak_7;
}

static int non_uni_coeff(Ctx &c) {
    IB3D_REQUIRE(c.grid_ready, "non_uni_coeff: no grid");
    const std::vector<double> *d[3] = {&c.hdx, &c.hdy, &c.hdz};
    const int64_t nn[3] = {c.nx, c.ny, c.nz};
    double cf[12];
    for (int a = 0; a < 3; a++) {
        const std::vector<double> &dl = *d[a];
        std::vector<double> &face = c.htab[a], &cent = c.htab[3 + a];
        const int64_t tn = c.tn[a];
        face.assign(12 * tn, 0.0);
        cent.assign(12 * tn, 0.0);
        for (int64_t i = 2; i <= nn[a]; i++) {
            kk_coeffs(dl[i] / dl[i - 1], dl[i + 1] / dl[i], dl[i + 2] / dl[i + 1], cf);
            for (int q = 0; q < 12; q++) face[q * tn + i] = cf[q];
            const double tmp1 = (i == 2) ? dl[i - 1] : 0.5 * (dl[i - 1] + dl[i - 2]);
            const double tmp2 = 0.5 * (dl[i] + dl[i - 1]);
            const double tmp3 = 0.5 * (dl[i] + dl[i + 1]);
            const double tmp4 = 0.5 * (dl[i + 1] + dl[i + 2]);
            kk_coeffs(tmp2 / tmp1, tmp3 / tmp2, tmp4 / tmp3, cf);
            for (int q = 0; q < 12; q++) cent[q * tn + i] = cf[q];
        }
    }
    for (int t = 0; t < 6; t++)
        if (upload_vec(c.tab[t], c.htab[t])) return -1;
    c.coeff_ready = true;
    c.mom_ready = false;
    return 0;
}

// subCoefficientMatrix.f90:8-53
static int coefficient_matrix(Ctx &c) {
    IB3D_REQUIRE(c.grid_ready, "coefficientMatrix: no grid");
    const int64_t nx = c.nx, ny = c.ny, nz = c.nz;
    c.hAcx.assign(nx * 3, 0.0);
    c.hAcy.assign(ny * 3, 0.0);
    c.hAcz.assign(nz * 3, 0.0);
    auto build = [](std::vector<double> &A, int64_t n, const std::vector<double> &cp) {
        for (int64_t i = 2; i <= n + 1; i++) {
            const double r1 = cp[i] - cp[i - 1], r2 = cp[i + 1] - cp[i], rs = r1 + r2;
            A[(i - 2) + 0 * n] = 2.0 / (r1 * rs);
            A[(i - 2) + 1 * n] = -2.0 / (r1 * r2);
            A[(i - 2) + 2 * n] = 2.0 / (r2 * rs);
        }
    };
    build(c.hAcy, ny, c.hyp);
    build(c.hAcx, nx, c.hxp);
    build(c.hAcz, nz, c.hzp);
    auto A = [](std::vector<double> &M, int64_t n, int64_t i, int col) -> double & { return M[(i - 1) + n * (col - 1)]; };
    A(c.hAcx, nx, 1, 2) = A(c.hAcx, nx, 1, 2) + A(c.hAcx, nx, 1, 1);
    A(c.hAcx, nx, 1, 1) = 0.0;
    A(c.hAcx, nx, nx, 2) = A(c.hAcx, nx, nx, 2) - A(c.hAcx, nx, nx, 3);
    A(c.hAcx, nx, nx, 3) = 0.0;
    A(c.hAcy, ny, 1, 2) = A(c.hAcy, ny, 1, 2) + A(c.hAcy, ny, 1, 1);
    A(c.hAcy, ny, 1, 1) = 0.0;
    A(c.hAcy, ny, ny, 2) = A(c.hAcy, ny, ny, 2) + A(c.hAcy, ny, ny, 3);
    A(c.hAcy, ny, ny, 3) = 0.0;
    A(c.hAcz, nz, 1, 2) = A(c.hAcz, nz, 1, 2) + A(c.hAcz, nz, 1, 1);
    A(c.hAcz, nz, 1, 1) = 0.0;
    A(c.hAcz, nz, nz, 2) = A(c.hAcz, nz, nz, 2) + A(c.hAcz, nz, nz, 3);
    A(c.hAcz, nz, nz, 3) = 0.0;
    return 0;
}

// device rows ax[col*s + ii] = Acx(ii, col+1) for ii = 1..n, 0 at the ghosts
// (the x rows use the padded stride sxp so 4-vectors of them stay aligned)
static int push_matrix(Ctx &c) {
    auto rows = [](const std::vector<double> &M, int64_t n, int64_t s) {
        std::vector<double> r(3 * s, 0.0);
        for (int col = 0; col < 3; col++)
            for (int64_t ii = 1; ii <= n; ii++) r[col * s + ii] = M[(ii - 1) + n * col];
        return r;
    };
    if (upload_vec(c.ax, rows(c.hAcx, c.nx, c.sxp)) || upload_vec(c.ay, rows(c.hAcy, c.ny, c.sy)) ||
        upload_vec(c.az, rows(c.hAcz, c.nz, c.sz)))
        return -1;
    sor_release_tables();
    c.matrix_ready = true;
    return 0;
}

// ------------------------------------------------------------ registry
enum Kind { F64_FIELD, PC_FIELD, TAG_FIELD, I32_FIELD, F64_VEC, HOST_VEC, LIST, ELEM_I32, IB_F64, IB_NEL, TAB };

struct Entry {
    Kind kind;
    void *dev;
    int64_t count;      // elements as the reference holds them
    int64_t offset;     // element offset inside a packed device array
    std::vector<double> *host;
    int set;
    int col;
};

static bool lookup(Ctx &c, const std::string &n, Entry &e) {
    e = Entry{F64_FIELD, nullptr, 0, 0, nullptr, 0, 0};
    const int64_t P = c.P;
    struct { const char *n; double *d; } fld[] = {{"u", c.u}, {"v", c.v}, {"w", c.w}, {"ut", c.ut}, {"vt", c.vt},
        {"wt", c.wt}, {"p", c.p}, {"resi_u", c.resi_u}, {"resi_v", c.resi_v}, {"resi_w", c.resi_w}};
    for (auto &f : fld)
        if (n == f.n) { e.kind = F64_FIELD; e.dev = f.d; e.count = P; return true; }
    if (n == "pc") { e.kind = PC_FIELD; e.count = P; return true; }
    if (n == "cell") { e.kind = TAG_FIELD; e.dev = c.tag; e.count = P; return true; }
    if (n == "minElemcell") { e.kind = I32_FIELD; e.dev = c.minElem; e.count = P; return true; }
    struct { const char *n; double *d; std::vector<double> *h; int64_t cnt; } vec[] = {
        {"x1", c.x1, &c.hx1, c.nx + 3}, {"y1", c.y1, &c.hy1, c.ny + 3}, {"z1", c.z1, &c.hz1, c.nz + 3},
        {"deltax", c.dx, &c.hdx, c.nx + 2}, {"deltay", c.dy, &c.hdy, c.ny + 2}, {"deltaz", c.dz, &c.hdz, c.nz + 2},
        {"xu", c.xu, &c.hxu, c.nx + 3}, {"yu", c.yu, &c.hyu, c.ny + 2}, {"zu", c.zu, &c.hzu, c.nz + 2},
        {"xv", c.xv, &c.hxv, c.nx + 2}, {"yv", c.yv, &c.hyv, c.ny + 3}, {"zv", c.zv, &c.hzv, c.nz + 2},
        {"xw", c.xw, &c.hxw, c.nx + 2}, {"yw", c.yw, &c.hyw, c.ny + 2}, {"zw", c.zw, &c.hzw, c.nz + 3},
        {"xp", c.xp, &c.hxp, c.nx + 2}, {"yp", c.yp, &c.hyp, c.ny + 2}, {"zp", c.zp, &c.hzp, c.nz + 2}};
    for (auto &v : vec)
        if (n == v.n) { e.kind = F64_VEC; e.dev = v.d; e.host = v.h; e.count = v.cnt; return true; }
    if (n == "Acx") { e.kind = HOST_VEC; e.host = &c.hAcx; e.count = c.nx * 3; return true; }
    if (n == "Acy") { e.kind = HOST_VEC; e.host = &c.hAcy; e.count = c.ny * 3; return true; }
    if (n == "Acz") { e.kind = HOST_VEC; e.host = &c.hAcz; e.count = c.nz * 3; return true; }
    struct { const char *n; int *d; int64_t cnt; } lst[] = {{"fluidIndexPtr", c.fluidList, c.fluidCount},
        {"solidIndexPtr", c.solidList, c.solidCount}, {"interceptedIndexPtr", c.ibList, c.ibCount},
        {"redCellIndexPtr", nullptr, c.redCount}, {"blackCellIndexPtr", nullptr, c.blackCount}};
    for (auto &l : lst)
        if (n == l.n) {
            if (!c.lists_ready) return false;
            e.kind = LIST; e.dev = l.d; e.count = 3 * l.cnt; e.set = (n[0] == 'r') ? 1 : (n[0] == 'b') ? 2 : 0;
            return true;
        }
    const char *ibf[] = {"pNormDis", "u1NormDis", "u2NormDis", "v1NormDis", "v2NormDis", "w1NormDis", "w2NormDis"};
    for (int q = 0; q < 7; q++)
        if (n == ibf[q]) {
            if (!c.normdist_ready) return false;
            e.kind = IB_F64; e.dev = c.normdis; e.offset = q * c.ibCount; e.count = c.ibCount; return true;
        }
    const char *nels[] = {"nelp", "nelu1", "nelu2", "nelv1", "nelv2", "nelw1", "nelw2"};
    for (auto nm : nels)
        if (n == nm) {
            if (!c.normdist_ready) return false;
            e.kind = IB_NEL; e.dev = c.nel; e.count = c.ibCount; return true;
        }
    struct { const char *n; double *d; int64_t cnt; } el[] = {{"xcent", c.xcent, c.ibElems},
        {"ycent", c.ycent, c.ibElems}, {"zcent", c.zcent, c.ibElems}, {"cosAlpha", c.cosA, c.ibElems},
        {"cosBeta", c.cosB, c.ibElems}, {"cosGamma", c.cosG, c.ibElems}, {"alpha3", c.alpha3, c.ibElems},
        {"beta3", c.beta3, c.ibElems}, {"gamma3", c.gamma3, c.ibElems}, {"xnode", c.xnode, c.ibNodes},
        {"ynode", c.ynode, c.ibNodes}, {"znode", c.znode, c.ibNodes}, {"xnode1", c.xnode1, c.ibNodes},
        {"ynode1", c.ynode1, c.ibNodes}, {"znode1", c.znode1, c.ibNodes}};
    for (auto &x : el)
        if (n == x.n) {
            if (!x.d) return false;
            e.kind = F64_FIELD; e.dev = x.d; e.count = x.cnt; return true;
        }
    const char *sc[] = {"Elemcell", "ucell", "vcell", "wcell", "pcell"};
    for (int q = 0; q < 5; q++)
        if (n == sc[q]) {
            if (!c.stress_ready) return false;
            e.kind = ELEM_I32; e.dev = c.scell; e.offset = (int64_t)q * 3 * c.ibElems; e.count = 3 * c.ibElems;
            return true;
        }
    // convection tables ca1_uu .. ck6_wv (1-based, n+2 entries)
    if (n.size() == 6 && (n.rfind("ca", 0) == 0 || n.rfind("ck", 0) == 0) && n[3] == '_') {
        const int idx = n[2] - '1';
        const std::string s = n.substr(4);
        const char *names[] = {"uu", "vv", "ww", "uv", "vu", "wu", "uw", "vw", "wv"};
        const int map[] = {0, 1, 2, 3, 4, 5, 3, 4, 5};
        for (int q = 0; q < 9; q++)
            if (s == names[q] && idx >= 0 && idx < 6) {
                if (!c.coeff_ready) return false;
                e.kind = TAB; e.set = map[q]; e.col = (n[1] == 'a' ? 0 : 6) + idx;
                e.count = c.tn[e.set] - 1;
                return true;
            }
    }
    return false;
}

}  // namespace ib3d

using namespace ib3d;

extern "C" {

const char *ib3d_last_error(void) {
    std::snprintf(g_err_buf, sizeof g_err_buf, "%s", g_err.c_str());
    return g_err_buf;
}

int ib3d_init(int64_t nx, int64_t ny, int64_t nz) {
    Ctx &c = ctx();
    if (c.ready) free_all(c);
    IB3D_REQUIRE(nx >= 4 && ny >= 4 && nz >= 4, "ib3d_init: each grid dimension must be >= 4");
    int ndev = 0;
    IB3D_CHECK(hipGetDeviceCount(&ndev));
    IB3D_REQUIRE(ndev > 0, "ib3d_init: no GPU visible");
    IB3D_CHECK(hipGetDevice(&c.device));
    if (noalloc_setup(c.device)) return -1;
    c.nx = nx; c.ny = ny; c.nz = nz;
    const ColourLayout lay = colour_layout(nx, ny, nz);
    c.sx = lay.sx; c.sy = lay.sy; c.sz = lay.sz;
    c.P = c.sx * c.sy * c.sz;
    c.H = lay.H;
    c.NC = lay.NC;
    c.sxp = lay.sxp;
    IB3D_REQUIRE(c.P < (int64_t)1 << 31, "ib3d_init: grid too large for 32-bit cell indices");
    if (check_fastdiv((int)c.H, c.NC + 1) || check_fastdiv((int)c.sy, c.NC / c.H + 1)) {
        set_error("ib3d_init: fast division self-check failed");
        return -1;
    }
    IB3D_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    for (auto &e : c.ev) IB3D_CHECK(hipEventCreate(&e));
    const int64_t n1x = nx + 4, n1y = ny + 4, n1z = nz + 4;
    double **xs[] = {&c.x1, &c.dx, &c.xu, &c.xv, &c.xw, &c.xp};
    double **ys[] = {&c.y1, &c.dy, &c.yu, &c.yv, &c.yw, &c.yp};
    double **zs[] = {&c.z1, &c.dz, &c.zu, &c.zv, &c.zw, &c.zp};
    for (auto p : xs) if (dalloc(p, n1x)) return -1;
    for (auto p : ys) if (dalloc(p, n1y)) return -1;
    for (auto p : zs) if (dalloc(p, n1z)) return -1;
    const int64_t nn[6] = {nx, ny, nz, nx, ny, nz};
    for (int t = 0; t < 6; t++) {
        c.tn[t] = nn[t] + 3;
        if (dalloc(&c.tab[t], 12 * c.tn[t])) return -1;
    }
    if (dalloc(&c.ax, 3 * c.sxp) || dalloc(&c.ay, 3 * c.sy) || dalloc(&c.az, 3 * c.sz)) return -1;
    double **f[] = {&c.u, &c.v, &c.w, &c.ut, &c.vt, &c.wt, &c.p, &c.resi_u, &c.resi_v, &c.resi_w};
    for (auto p : f) if (dalloc(p, c.P)) return -1;
    if (dalloc(&c.tag, c.P) || dalloc(&c.minElem, c.P)) return -1;
    // FP32 colour pairs are one block each (colour 1 follows colour 0), so the
    // Krylov vector kernels can treat them as one vector of 2*NC values
    if (dalloc(&c.e32[0], 2 * c.NC) || dalloc(&c.r32[0], 2 * c.NC)) return -1;
    c.e32[1] = c.e32[0] + c.NC;
    c.r32[1] = c.r32[0] + c.NC;
    hipLaunchKernelGGL(k_fill_nan32, dim3(nblocks(2 * c.NC, 256)), dim3(256), 0, c.stream, c.r32[0], 2 * c.NC);
    for (int col = 0; col < 2; col++) {
        if (dalloc(&c.pc[col], c.NC) || dalloc(&c.bdt[col], c.NC)) return -1;
        hipLaunchKernelGGL(k_fill_nan64, dim3(nblocks(c.NC, 256)), dim3(256), 0, c.stream, c.bdt[col], c.NC);
    }
    if (dalloc(&c.scalars, 64)) return -1;
    IB3D_CHECK(hipHostMalloc((void **)&c.hscal, 64 * sizeof(double), hipHostMallocDefault));
    if (ensure_slots(4096)) return -1;
    IB3D_CHECK(hipStreamSynchronize(c.stream));
    c.ready = true;
    return 0;
}

int ib3d_finalize(void) {
    Ctx &c = ctx();
    if (c.ready) {
        (void)hipStreamSynchronize(c.stream);
        free_all(c);
    }
    return 0;
}

int ib3d_sync(void) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "ib3d_sync: not initialised");
    IB3D_CHECK(hipStreamSynchronize(c.stream));
    return 0;
}

int64_t ib3d_array_count(const char *name) {
    Ctx &c = ctx();
    if (!c.ready) return 0;
    Entry e;
    if (!lookup(c, name, e)) return 0;
    return e.count;
}

int ib3d_download(const char *name, void *host, int64_t nbytes) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "ib3d_download: not initialised");
    Entry e;
    IB3D_REQUIRE(lookup(c, name, e), std::string("ib3d_download: unknown or not yet computed array ") + name);
    IB3D_REQUIRE(nbytes == e.count * 8, std::string("ib3d_download: size mismatch for ") + name);
    IB3D_CHECK(hipStreamSynchronize(c.stream));
    switch (e.kind) {
        case F64_FIELD:
            IB3D_CHECK(d2h(host, (double *)e.dev + e.offset, nbytes));
            break;
        case IB_F64:
            IB3D_CHECK(d2h(host, (double *)e.dev + e.offset, nbytes));
            break;
        case PC_FIELD: {
            double *tmp = nullptr;
            IB3D_CHECK(hipMalloc(&tmp, c.P * sizeof(double)));
            hipLaunchKernelGGL(k_gather_pc, dim3(nblocks(c.P, 256)), dim3(256), 0, c.stream, c.pc[0], c.pc[1], tmp,
                               (int)c.sx, (int)c.sy, (int)c.sz, c.H);
            IB3D_CHECK(hipStreamSynchronize(c.stream));
            IB3D_CHECK(d2h(host, tmp, nbytes));
            IB3D_CHECK(hipFree(tmp));
            break;
        }
        case TAG_FIELD: {
            std::vector<int8_t> t(c.P);
            IB3D_CHECK(d2h(t.data(), c.tag, c.P));
            int64_t *o = (int64_t *)host;
            for (int64_t q = 0; q < c.P; q++) o[q] = t[q];
            break;
        }
        case I32_FIELD: case IB_NEL: case ELEM_I32: {
            std::vector<int> t(e.count);
            IB3D_CHECK(d2h(t.data(), (int *)e.dev + e.offset, e.count * sizeof(int)));
            int64_t *o = (int64_t *)host;
            for (int64_t q = 0; q < e.count; q++) o[q] = t[q];
            break;
        }
        case LIST: {
            int64_t *o = (int64_t *)host;
            if (e.set == 0) {
                std::vector<int> t(e.count);
                if (e.count) IB3D_CHECK(d2h(t.data(), e.dev, e.count * sizeof(int)));
                for (int64_t q = 0; q < e.count; q++) o[q] = t[q];
            } else {
                // red/black lists: the fluid list split by mod(i+j+k,2)
                // (subSearch.f90:557-589)
                const int64_t nf = c.fluidCount, cnt = e.count / 3;
                std::vector<int> t(3 * nf);
                if (nf) IB3D_CHECK(d2h(t.data(), c.fluidList, 3 * nf * sizeof(int)));
                int64_t r = 0;
                for (int64_t q = 0; q < nf; q++) {
                    const int odd = (t[q] + t[nf + q] + t[2 * nf + q]) % 2 == 1;
                    if ((e.set == 1) == odd) {
                        o[r] = t[q];
                        o[cnt + r] = t[nf + q];
                        o[2 * cnt + r] = t[2 * nf + q];
                        r++;
                    }
                }
            }
            break;
        }
        case F64_VEC:
            std::memcpy(host, e.host->data() + 1, nbytes);
            break;
        case HOST_VEC:
            std::memcpy(host, e.host->data(), nbytes);
            break;
        case TAB: {
            const std::vector<double> &t = c.htab[e.set];
            const int64_t tn = c.tn[e.set];
            std::memcpy(host, t.data() + e.col * tn + 1, nbytes);
            break;
        }
    }
    return 0;
}

int ib3d_upload(const char *name, const void *host, int64_t nbytes) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "ib3d_upload: not initialised");
    Entry e;
    IB3D_REQUIRE(lookup(c, name, e), std::string("ib3d_upload: unknown array ") + name);
    IB3D_REQUIRE(nbytes == e.count * 8, std::string("ib3d_upload: size mismatch for ") + name);
    IB3D_CHECK(hipStreamSynchronize(c.stream));
    switch (e.kind) {
        case F64_FIELD: {
            IB3D_CHECK(h2d((double *)e.dev + e.offset, host, nbytes));
            // surface arrays: the normal distances, the forcing plans and the
            // stress stencils were found on the old surface; new nodes also
            // leave the normals stale
            const double *surf[] = {c.xcent, c.ycent, c.zcent, c.cosA, c.cosB, c.cosG, c.alpha3, c.beta3, c.gamma3};
            const double *nodes[] = {c.xnode, c.ynode, c.znode, c.xnode1, c.ynode1, c.znode1};
            bool is_surf = false, is_node = false;
            for (auto a : surf) is_surf |= (e.dev == a);
            for (auto a : nodes) is_node |= (e.dev == a);
            if (is_surf || is_node) {
                c.normdist_ready = c.plans_ready = c.plans_pending = false;
                c.stress_ready = false;
            }
            if (is_node) c.norms_ready = false;
            break;
        }
        case PC_FIELD: {
            double *tmp = nullptr;
            IB3D_CHECK(hipMalloc(&tmp, c.P * sizeof(double)));
            IB3D_CHECK(h2d(tmp, host, nbytes));
            hipLaunchKernelGGL(k_scatter_pc, dim3(nblocks(c.P, 256)), dim3(256), 0, c.stream, c.pc[0], c.pc[1], tmp,
                               (int)c.sx, (int)c.sy, (int)c.sz, c.H);
            IB3D_CHECK(hipStreamSynchronize(c.stream));
            IB3D_CHECK(hipFree(tmp));
            c.have_prev = true;
            break;
        }
        case TAG_FIELD: {
            const int64_t *in = (const int64_t *)host;
            std::vector<int8_t> t(c.P);
            for (int64_t q = 0; q < c.P; q++) {
                IB3D_REQUIRE(in[q] >= 0 && in[q] <= 2, "ib3d_upload: cell tags must be 0, 1 or 2");
                t[q] = (int8_t)in[q];
            }
            IB3D_CHECK(h2d(c.tag, t.data(), c.P));
            c.lists_ready = c.normdist_ready = c.plans_ready = false;
            c.plans_pending = false;
            break;
        }
        case F64_VEC: {
            std::memcpy(e.host->data() + 1, host, nbytes);
            IB3D_CHECK(h2d(e.dev, e.host->data(), e.host->size() * sizeof(double)));
            c.mom_ready = false;   // the momentum tables read the spacings
            // image points, stress probes and the multigrid levels (built
            // from the grid lines) belong to the old grid
            c.normdist_ready = c.plans_ready = c.plans_pending = false;
            c.stress_ready = false;
            sor_release_tables();
            break;
        }
        case HOST_VEC: {
            std::memcpy(e.host->data(), host, nbytes);
            if (push_matrix(c)) return -1;
            break;
        }
        default:
            set_error(std::string("ib3d_upload: ") + name + " is computed on the device and cannot be uploaded");
            return -1;
    }
    return 0;
}

int ib3d_set_grid(const double *gx, const double *gy, const double *gz) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "ib3d_set_grid: not initialised");
    grid_axis(c.nx, gx, c.hx1, c.hdx);
    grid_axis(c.ny, gy, c.hy1, c.hdy);
    grid_axis(c.nz, gz, c.hz1, c.hdz);
    for (int64_t i = 2; i <= c.nx + 1; i++) IB3D_REQUIRE(c.hdx[i] > 0, "ib3d_set_grid: x nodes must increase");
    for (int64_t i = 2; i <= c.ny + 1; i++) IB3D_REQUIRE(c.hdy[i] > 0, "ib3d_set_grid: y nodes must increase");
    for (int64_t i = 2; i <= c.nz + 1; i++) IB3D_REQUIRE(c.hdz[i] > 0, "ib3d_set_grid: z nodes must increase");
    c.coeff_ready = c.matrix_ready = false;
    return push_grid(c);
}

int ib3d_set_mesh(int64_t n_nodes, const double *xn, const double *yn, const double *zn, int64_t n_elems,
                  const int64_t *p1, const int64_t *p2, const int64_t *p3) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "ib3d_set_mesh: not initialised");
    IB3D_REQUIRE(n_nodes > 0 && n_elems > 0, "ib3d_set_mesh: empty mesh");
    std::vector<int> el(3 * n_elems);
    for (int64_t e = 0; e < n_elems; e++) {
        const int64_t a = p1[e], b = p2[e], d = p3[e];
        IB3D_REQUIRE(a >= 1 && b >= 1 && d >= 1 && a <= n_nodes && b <= n_nodes && d <= n_nodes,
                     "ib3d_set_mesh: connectivity out of range");
        el[e] = (int)a;
        el[n_elems + e] = (int)b;
        el[2 * n_elems + e] = (int)d;
    }
    double **nd[] = {&c.xnode, &c.ynode, &c.znode, &c.xnode1, &c.ynode1, &c.znode1};
    for (auto p : nd) dfree(*p);
    double **ed[] = {&c.xcent, &c.ycent, &c.zcent, &c.cosA, &c.cosB, &c.cosG, &c.alpha3, &c.beta3, &c.gamma3};
    for (auto p : ed) dfree(*p);
    dfree(c.elp);
    dfree(c.scell); dfree(c.scontrib);
    dfree(c.body);
    c.nbody = 0;
    for (auto p : nd) if (dalloc(p, n_nodes)) return -1;
    for (auto p : ed) if (dalloc(p, n_elems)) return -1;
    if (dalloc(&c.elp, 3 * n_elems)) return -1;
    IB3D_CHECK(h2d(c.xnode, xn, n_nodes * 8));
    IB3D_CHECK(h2d(c.ynode, yn, n_nodes * 8));
    IB3D_CHECK(h2d(c.znode, zn, n_nodes * 8));
    IB3D_CHECK(h2d(c.elp, el.data(), el.size() * sizeof(int)));
    c.ibNodes = n_nodes;
    c.ibElems = n_elems;
    c.norms_ready = c.stress_ready = false;
    return 0;
}

int ib3d_shiftSurfaceNodesInitial(double xs, double ys, double zs) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "not initialised");
    return shift_nodes(c, xs, ys, zs);
}

int ib3d_shiftSurfaceNodesMovingbody(double xshift, double yshift, double zshift, double dx, double dy,
                                     double dz) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "shiftSurfaceNodesMovingbody: not initialised");
    const double base[3] = {xshift, yshift, zshift}, d[3] = {dx, dy, dz};
    return shift_nodes_moving(c, base, d);
}

int ib3d_selectiveRetagging(int32_t counts[3]) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "selectiveRetagging: not initialised");
    return selective_retagging(c, counts);
}

int ib3d_set_element_bodies(const int32_t *body, int64_t n_elems, int32_t n_bodies) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "set_element_bodies: not initialised");
    IB3D_REQUIRE(n_elems == c.ibElems && c.ibElems > 0, "set_element_bodies: one body id per surface element");
    IB3D_REQUIRE(n_bodies >= 1 && n_bodies <= 65535, "set_element_bodies: 1..65535 bodies");
    for (int64_t e = 0; e < n_elems; e++)
        IB3D_REQUIRE(body[e] >= 1 && body[e] <= n_bodies, "set_element_bodies: body ids must lie in 1..n_bodies");
    dfree(c.body);
    if (dalloc(&c.body, n_elems)) return -1;
    IB3D_CHECK(h2d(c.body, body, n_elems * sizeof(int32_t)));
    c.nbody = n_bodies;
    return 0;
}

int ib3d_stress_bodies(double *coef) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "stress_bodies: not initialised");
    return stress_bodies(c, coef);
}

int ib3d_set_wall_motion(const double vel[3], const double acc[3]) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "set_wall_motion: not initialised");
    for (int q = 0; q < 3; q++) {
        IB3D_REQUIRE(std::isfinite(vel[q]) && std::isfinite(acc[q]), "set_wall_motion: non-finite value");
        c.wall_vel[q] = vel[q];
        c.wall_acc[q] = acc[q];
    }
    return 0;
}

int ib3d_computeSurfaceNorm(int64_t inor) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "not initialised");
    return surface_norm(c, inor);
}

int ib3d_tagging(const int64_t window[6], int32_t counts[3]) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready && c.grid_ready, "tagging: no grid");
    return tagging(c, window, counts);
}

int ib3d_cellCount(int32_t counts[5]) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "not initialised");
    return cell_count(c, counts);
}

int ib3d_computeNormDistance(void) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "not initialised");
    return norm_distance(c);
}

int ib3d_initialConditions(void) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready && c.lists_ready, "initialConditions: cellCount has not run");
    double *z[] = {c.ut, c.v, c.vt, c.w, c.wt, c.p, c.resi_u, c.resi_v, c.resi_w};
    for (auto p : z) IB3D_CHECK(hipMemsetAsync(p, 0, c.P * sizeof(double), c.stream));
    hipLaunchKernelGGL(k_initial, dim3(nblocks(c.P, 256)), dim3(256), 0, c.stream, c.u, c.tag, (int)c.nx, (int)c.ny,
                       (int)c.nz);
    IB3D_CHECK(hipMemsetAsync(c.pc[0], 0, c.NC * sizeof(double), c.stream));
    IB3D_CHECK(hipMemsetAsync(c.pc[1], 0, c.NC * sizeof(double), c.stream));
    IB3D_CHECK(hipGetLastError());
    IB3D_CHECK(hipStreamSynchronize(c.stream));
    c.have_prev = false;
    return 0;
}

int ib3d_coefficientMatrix(void) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "not initialised");
    if (coefficient_matrix(c)) return -1;
    return push_matrix(c);
}

int ib3d_non_uni_coeff(void) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "not initialised");
    return non_uni_coeff(c);
}

int ib3d_nsMomentum(double rev, double deltat) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "not initialised");
    return ns_momentum(c, rev, deltat);
}

int ib3d_velocityBC(void) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "not initialised");
    return velocity_bc(c);
}

int ib3d_solidCellBC(void) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "not initialised");
    return solid_bc(c);
}

int ib3d_velocityForcing1(void) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "not initialised");
    return velocity_forcing(c);
}

int ib3d_pressureForcing1(void) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "not initialised");
    return pressure_forcing(c);
}

int ib3d_computeDiv(double deltat, double *divmax) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready && c.lists_ready, "computeDiv: cellCount has not run");
    return compute_div(c, deltat, divmax, nullptr);
}

int ib3d_computeDiv_noalloc(double deltat, double *divmax) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready && c.lists_ready, "computeDiv_noalloc: cellCount has not run");
    return compute_div_noalloc(c, deltat, divmax);
}

int ib3d_REDBLACKSOR(double epsi, int64_t *isum, double *derr, double *derr2) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready && c.matrix_ready, "REDBLACKSOR: coefficientMatrix has not run");
    *derr = 0.0;
    return redblack_sor_reference(c, epsi, isum, derr2);
}

int ib3d_correctPressure(void) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "not initialised");
    return correct(c, 0.0, 1, 0);
}

int ib3d_correctVelocity(double deltat) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "not initialised");
    return correct(c, deltat, 0, 1);
}

// The Poisson stage (subPcorVcor.f90): divergence, solve, correction, wall
// conditions and the steady-state measure.  deferred: the steady-state
// measure and the stage times are left to arrive with the caller's next
// synchronisation (poisson_collect reads them), so the whole step waits on
// the host fewer times.
static int poisson_stage(Ctx &c, double deltat, double eps1, int64_t *nIterPcor, double *divmax,
                         double *derrStdSt, double *sor_seconds, bool deferred);
static void poisson_collect(Ctx &c, double *divmax, double *derrStdSt, double *sor_seconds) {
    if (divmax) {
        *divmax = c.hscal[0];
        c.bnorm2 = c.hscal[1];
    }
    float t_sor = 0, t_all = 0;
    (void)hipEventElapsedTime(&t_sor, c.ev[1], c.ev[2]);
    (void)hipEventElapsedTime(&t_all, c.ev[0], c.ev[3]);
    c.last_sor_ms = t_sor;
    c.last_poisson_ms = t_all;
    if (sor_seconds) *sor_seconds = t_sor * 1e-3;
    if (derrStdSt) *derrStdSt = c.hscal[3];
}

int ib3d_poissonSolver(double deltat, double eps1, int64_t *nIterPcor, double *divmax, double *derrStdSt,
                       double *sor_seconds) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready && c.matrix_ready && c.lists_ready, "poissonSolver: setup incomplete");
    return poisson_stage(c, deltat, eps1, nIterPcor, divmax, derrStdSt, sor_seconds, false);
}

static int poisson_stage(Ctx &c, double deltat, double eps1, int64_t *nIterPcor, double *divmax,
                         double *derrStdSt, double *sor_seconds, bool deferred) {
    IB3D_CHECK(hipEventRecord(c.ev[0], c.stream));
    if (c.mode == 0 || !c.warm || !c.have_prev) {
        IB3D_CHECK(hipMemsetAsync(c.pc[0], 0, c.NC * sizeof(double), c.stream));
        IB3D_CHECK(hipMemsetAsync(c.pc[1], 0, c.NC * sizeof(double), c.stream));
    }
    double bnorm2 = 0.0;
    // the Krylov modes can take ||b/dt|| from device memory: then nothing
    // waits for the host between the divergence and the solve
    const bool dev_norm = deferred && (c.mode == 3 || c.mode == 4) && !c.use_graphs && !cnn_active();
    double *dnorm = c.scalars + kNormSlot;
    if (compute_div(c, deltat, divmax, &bnorm2, !dev_norm)) return -1;
    if (dev_norm && norm_target_async(c, c.scalars + 1, dnorm)) return -1;
    IB3D_CHECK(hipEventRecord(c.ev[1], c.stream));
    if (sor_solve(c, bnorm2, eps1, dev_norm ? dnorm : nullptr)) return -1;
    IB3D_CHECK(hipEventRecord(c.ev[2], c.stream));
    if (correct(c, deltat, 1, 1)) return -1;
    if (velocity_bc(c)) return -1;
    if (steady_copy(c, deltat, nullptr)) return -1;
    IB3D_CHECK(hipEventRecord(c.ev[3], c.stream));
    *nIterPcor = c.last_sweeps;
    c.have_prev = true;
    if (deferred) return 0;
    IB3D_CHECK(hipEventSynchronize(c.ev[3]));
    poisson_collect(c, nullptr, derrStdSt, sor_seconds);
    return 0;
}

int ib3d_stressCal1(double re, double coef[4]) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "not initialised");
    return stress(c, 1, re, coef);
}

int ib3d_stressCal2(double re, double coef[4]) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "not initialised");
    return stress(c, 0, re, coef);
}

int ib3d_set_poisson_mode(int mode, double omega, double tol, int check, int warm, int64_t max_sweeps) {
    Ctx &c = ctx();
    IB3D_REQUIRE(mode >= 0 && mode <= 4, "set_poisson_mode: mode must be 0, 1, 2, 3 or 4");
    IB3D_REQUIRE(omega > 0.0 && omega < 2.0, "set_poisson_mode: omega must lie in (0, 2)");
    IB3D_REQUIRE(check >= 1, "set_poisson_mode: check must be >= 1");
    IB3D_REQUIRE(tol > 0.0, "set_poisson_mode: tol must be positive");
    c.mode = mode;
    c.omega = mode == 0 ? 1.99 : omega;
    c.tol = tol;
    c.check = check;
    c.warm = warm;
    c.max_sweeps = max_sweeps > 0 ? max_sweeps : 2000000;
    return 0;
}

int ib3d_set_krylov(int restart) {
    Ctx &c = ctx();
    IB3D_REQUIRE(restart >= 1 && restart <= 23, "set_krylov: restart length must be 1..23");
    c.krylov_m = restart;
    return 0;
}

int ib3d_set_multigrid(int nu1, int nu2, double omega, int coarse_sweeps, double coarse_omega, int max_levels) {
    Ctx &c = ctx();
    IB3D_REQUIRE(nu1 >= 1 && nu1 <= 64, "set_multigrid: pre-smoothing sweeps must be 1..64");
    IB3D_REQUIRE(nu2 >= 0 && nu2 <= 64, "set_multigrid: post-smoothing sweeps must be 0..64");
    IB3D_REQUIRE(coarse_sweeps >= 1 && coarse_sweeps <= 1024, "set_multigrid: coarse sweeps must be 1..1024");
    IB3D_REQUIRE(omega > 0.0 && omega < 2.0 && coarse_omega > 0.0 && coarse_omega < 2.0,
                 "set_multigrid: relaxation factors must lie in (0, 2)");
    IB3D_REQUIRE(max_levels == 0 || max_levels >= 2, "set_multigrid: max_levels must be 0 (no cap) or >= 2");
    c.mg_nu1 = c.mg_nu1c = nu1;
    c.mg_nu2 = c.mg_nu2c = nu2;
    for (auto &lv : c.mg_nu_lv) lv[0] = lv[1] = -1;
    c.mg_nu_small[0] = c.mg_nu_small[1] = -1;
    c.mg_omega = omega;
    c.mg_coarse = coarse_sweeps;
    c.mg_coarse_omega = coarse_omega;
    c.mg_levels = max_levels;
    return 0;
}

int ib3d_set_krylov_target(double target) {
    Ctx &c = ctx();
    IB3D_REQUIRE(target > 0.0 && target <= 1.0, "set_krylov_target: the target must lie in (0, 1]");
    c.krylov_target = target;
    return 0;
}

int ib3d_set_multigrid_coarse(int nu1, int nu2) {
    Ctx &c = ctx();
    IB3D_REQUIRE(nu1 >= 1 && nu1 <= 64, "set_multigrid_coarse: pre-smoothing sweeps must be 1..64");
    IB3D_REQUIRE(nu2 >= 0 && nu2 <= 64, "set_multigrid_coarse: post-smoothing sweeps must be 0..64");
    c.mg_nu1c = nu1;
    c.mg_nu2c = nu2;
    return 0;
}

int ib3d_set_multigrid_level(int level, int nu1, int nu2) {
    Ctx &c = ctx();
    IB3D_REQUIRE(level >= 1 && level <= 7, "set_multigrid_level: level must be 1..7");
    IB3D_REQUIRE((nu1 == -1 && nu2 == -1) || (nu1 >= 1 && nu1 <= 64 && nu2 >= 0 && nu2 <= 64),
                 "set_multigrid_level: sweeps must be 1..64 before and 0..64 after (-1, -1: no override)");
    c.mg_nu_lv[level][0] = nu1;
    c.mg_nu_lv[level][1] = nu2;
    return 0;
}

int ib3d_set_multigrid_small(int nu1, int nu2) {
    Ctx &c = ctx();
    IB3D_REQUIRE((nu1 == -1 && nu2 == -1) || (nu1 >= 1 && nu1 <= 64 && nu2 >= 0 && nu2 <= 64),
                 "set_multigrid_small: sweeps must be 1..64 before and 0..64 after (-1, -1: the coarse counts)");
    c.mg_nu_small[0] = nu1;
    c.mg_nu_small[1] = nu2;
    return 0;
}

int ib3d_set_fused_sweeps(int levels, int planes, int blocks, int coarse_blocks) {
    return set_fused_sweeps(levels, planes, blocks, coarse_blocks);
}
int ib3d_set_coarse_tiles(int from_level) { return set_coarse_tiles(from_level); }
int ib3d_set_fused_restrict(int levels) { return set_fused_restrict(levels); }

int ib3d_selfcheck_resid_guard(int64_t nx, int64_t ny, int64_t nz, int drop_interior, int64_t *outside) {
    IB3D_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && outside, "selfcheck_resid_guard: bad grid or output");
    *outside = resid_guard_selfcheck(nx, ny, nz, drop_interior != 0);
    return 0;
}

int ib3d_set_momentum_tiles(int on) { return set_momentum_tiles(on); }

int ib3d_set_graphs(int on) {
    ctx().use_graphs = on != 0;
    return 0;
}

int ib3d_poisson_stats(int64_t *sweeps, int64_t *outer, double *relres, double *relres0, double *sor_ms,
                       double *poisson_ms) {
    Ctx &c = ctx();
    *sweeps = c.last_sweeps;
    *outer = c.last_outer;
    *relres = c.last_relres;
    *relres0 = c.last_r0;
    *sor_ms = c.last_sor_ms;
    *poisson_ms = c.last_poisson_ms;
    return 0;
}

int ib3d_residual_norm(double *rnorm, double *bnorm) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready && c.matrix_ready, "residual_norm: setup incomplete");
    return residual_norm(c, rnorm, bnorm);
}

int ib3d_step(int first, double rev, double re, double deltat, double eps1, double out[8]) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "not initialised");
    if (ns_momentum(c, rev, deltat) || velocity_bc(c) || solid_bc(c) || velocity_forcing(c) || velocity_bc(c))
        return -1;
    int64_t it = 0;
    double dm = 0, ds = 0, sor_s = 0;
    IB3D_REQUIRE(c.matrix_ready && c.lists_ready, "step: setup incomplete");
    if (poisson_stage(c, deltat, eps1, &it, &dm, &ds, &sor_s, true)) return -1;
    if (pressure_forcing(c)) return -1;
    double coef[4];
    if (stress(c, first, re, coef)) return -1;   // waits for the stream
    poisson_collect(c, &dm, &ds, &sor_s);
    out[0] = (double)it;
    out[1] = dm;
    out[2] = ds;
    for (int q = 0; q < 4; q++) out[3 + q] = coef[q];
    out[7] = sor_s;
    return 0;
}

int ib3d_bench_sor(int precision, int sweeps, double omega, double *ms_per_sweep) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready && c.matrix_ready, "bench_sor: setup incomplete");
    IB3D_REQUIRE(precision == 64 || precision == 32 || precision == 33,
                 "bench_sor: precision must be 64, 32 or 33 (FP32, one launch per sweep)");
    IB3D_REQUIRE(sweeps >= 1, "bench_sor: sweeps must be >= 1");
    return bench_sor(c, precision, sweeps, omega, ms_per_sweep);
}

int ib3d_probe_copy(int64_t nbytes, double *ms) {
    Ctx &c = ctx();
    IB3D_REQUIRE(c.ready, "probe_copy: not initialised");
    IB3D_REQUIRE(nbytes >= 32, "probe_copy: nbytes must be >= 32");
    return probe_copy(c, nbytes, ms);
}

}  // extern "C"
