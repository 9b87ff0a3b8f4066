// ib3d_fields.hip -- momentum predictor, boundary conditions, projection
// corrections and the steady-state check (subNavierStokes.f90:622-1122,
// subBoundaryConditions.f90:137-217, subPcorVcor.f90:2-142).
//
// The reference iterates gathered fluid lists; each fluid cell's update is
// independent of the others, so these kernels walk the padded grid in
// Fortran order (coalesced, i fastest) and test the int8 tag instead.
// Arithmetic is written operand for operand like the reference (built with
// -ffp-contract=off, correctly rounded division), so results are bitwise.
#include "ib3d_internal.h"

namespace ib3d {

struct Tabs {
    const double *t[6];
    int tn[6];
};

enum { UU = 0, VV = 1, WW = 2, UV = 3, VU = 4, WU = 5 };

#define TB(set, c, i) (T.t[set][(c) * T.tn[set] + (i)])

// 3rd-order Kawamura-Kuwahara convection term (subNavierStokes.f90:812-831)
__device__ __forceinline__ double kk3(const Tabs &T, int set, int i, double vel, double f2, double f1, double f0,
                                      double fm1, double fm2, double dA, double dB) {
    const double sa = (((TB(set, 0, i) * f2 + TB(set, 1, i) * f1) + TB(set, 2, i) * f0) + TB(set, 3, i) * fm1) +
                      TB(set, 4, i) * fm2;
    const double sk = (((TB(set, 6, i) * f2 + TB(set, 7, i) * f1) + TB(set, 8, i) * f0) + TB(set, 9, i) * fm1) +
                      TB(set, 10, i) * fm2;
    return (vel * sa) / (TB(set, 5, i) * dA) + (fabs(vel) * sk) / ((2.0 * TB(set, 11, i)) * dB);
}

// non-uniform 1st-order upwind term (subNavierStokes.f90:847-866), al = 1
__device__ __forceinline__ double up1(double vel, double den, double r, double fm, double f0, double fp) {
    const double t1 = (((-1.0 / (r * (r + 1.0))) * fm - (1.0 - 1.0 / r) * f0) + (r / (r + 1.0)) * fp);
    const double t2 = ((2.0 / (r * (r + 1.0))) * fm - (2.0 / r) * f0) + (2.0 / (r + 1.0)) * fp;
    return (vel / den) * t1 - ((1.0 * fabs(vel)) * (1.0 / (2.0 * den))) * t2;
}

__global__ void __launch_bounds__(256)
k_momentum(const double *__restrict__ u, const double *__restrict__ v, const double *__restrict__ w,
           const double *__restrict__ p, const int8_t *__restrict__ cell, double *__restrict__ ut,
           double *__restrict__ vt, double *__restrict__ wt, double *__restrict__ ru, double *__restrict__ rv,
           double *__restrict__ rw, const double *__restrict__ dxa, const double *__restrict__ dya,
           const double *__restrict__ dza, Tabs T, int nx, int ny, int nz, double rev, double deltat) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 2;
    const int k = blockIdx.z + 2;
    if (i > nx + 1 || j > ny + 1) return;
    const int64_t sx = nx + 2, sxy = sx * (ny + 2);
    const int64_t c = (i - 1) + sx * ((j - 1) + (int64_t)(ny + 2) * (k - 1));
    if (cell[c] != 0) return;
#define U(a, b, d) u[c + (a) + (b) * sx + (d) * sxy]
#define V(a, b, d) v[c + (a) + (b) * sx + (d) * sxy]
#define W(a, b, d) w[c + (a) + (b) * sx + (d) * sxy]
#define P(a, b, d) p[c + (a) + (b) * sx + (d) * sxy]
#define CL(a, b, d) cell[c + (a) + (b) * sx + (d) * sxy]
    const double dxr = dxa[i + 1], dx = dxa[i], dxl = dxa[i - 1];
    const double dye = dya[j + 1], dy = dya[j], dyw = dya[j - 1];
    const double dzt = dza[k + 1], dz = dza[k], dzb = dza[k - 1];
    const double u22 = U(-1, 0, 0) - U(0, 0, 0), u4 = U(0, 0, 0) - U(1, 0, 0);
    const double u6 = U(0, -1, 0) - U(0, 0, 0), u8 = U(0, 0, 0) - U(0, 1, 0);
    const double u10 = U(0, 0, -1) - U(0, 0, 0), u12 = U(0, 0, 0) - U(0, 0, 1);
    const double v4 = V(-1, 0, 0) - V(0, 0, 0), v6 = V(0, 0, 0) - V(1, 0, 0);
    const double v8 = V(0, -1, 0) - V(0, 0, 0), v10 = V(0, 0, 0) - V(0, 1, 0);
    const double v12 = V(0, 0, -1) - V(0, 0, 0), v14 = V(0, 0, 0) - V(0, 0, 1);
    const double w6 = W(-1, 0, 0) - W(0, 0, 0), w8 = W(0, 0, 0) - W(1, 0, 0);
    const double w10 = W(0, -1, 0) - W(0, 0, 0), w12 = W(0, 0, 0) - W(0, 1, 0);
    const double w14 = W(0, 0, -1) - W(0, 0, 0), w16 = W(0, 0, 0) - W(0, 0, 1);
    const double dpdx = (P(0, 0, 0) - P(1, 0, 0)) / (0.5 * (dxr + dx));
    const double dpdy = (P(0, 0, 0) - P(0, 1, 0)) / (0.5 * (dye + dy));
    const double dpdz = (P(0, 0, 0) - P(0, 0, 1)) / (0.5 * (dzt + dz));
    const double dx2xr = dx + dxr, dx2xl = dx + dxl, dy2ye = dy + dye, dy2yw = dy + dyw;
    const double dz2zt = dz + dzt, dz2zb = dz + dzb;
    const bool no_ib = CL(1, 0, 0) != 2 && CL(-1, 0, 0) != 2 && CL(0, 1, 0) != 2 && CL(0, -1, 0) != 2 &&
                       CL(0, 0, 1) != 2 && CL(0, 0, -1) != 2;

    // U momentum (subNavierStokes.f90:792-897)
    {
        const double vu_e = V(1, 0, 0) + (dxr / dx2xr) * (V(0, 0, 0) - V(1, 0, 0));
        const double vu_w = V(1, -1, 0) + (dxr / dx2xr) * (V(0, -1, 0) - V(1, -1, 0));
        const double v_in_um = 0.5 * (vu_e + vu_w);
        const double wu_n = W(1, 0, 0) + (dxr / dx2xr) * (W(0, 0, 0) - W(1, 0, 0));
        const double wu_s = W(1, 0, -1) + (dxr / dx2xr) * (W(0, 0, -1) - W(1, 0, -1));
        const double w_in_um = 0.5 * (wu_n + wu_s);
        double a, b, d;
        if (i != 2 && i < nx && j != 2 && j < ny + 1 && k != 2 && k < nz + 1 && no_ib) {
            const double ddy = 0.5 * (dya[j] + dya[j - 1]), ddye = 0.5 * (dya[j] + dya[j + 1]);
            const double ddz = 0.5 * (dza[k] + dza[k - 1]), ddzr = 0.5 * (dza[k] + dza[k + 1]);
            const double uc = U(0, 0, 0);
            a = kk3(T, UU, i, uc, U(2, 0, 0), U(1, 0, 0), uc, U(-1, 0, 0), U(-2, 0, 0), dxa[i + 1], dxa[i]);
            b = kk3(T, VU, j, v_in_um, U(0, 2, 0), U(0, 1, 0), uc, U(0, -1, 0), U(0, -2, 0), ddye, ddy);
            d = kk3(T, WU, k, w_in_um, U(0, 0, 2), U(0, 0, 1), uc, U(0, 0, -1), U(0, 0, -2), ddzr, ddz);
        } else {
            const double r1x = dxa[i] / dxa[i + 1];
            const double r1yn = 0.5 * (dya[j] + dya[j - 1]), r1yd = 0.5 * (dya[j] + dya[j + 1]);
            const double r1y = r1yn / r1yd;
            const double r1zn = 0.5 * (dza[k] + dza[k - 1]), r1zd = 0.5 * (dza[k] + dza[k + 1]);
            const double r1z = r1zn / r1zd;
            a = up1(U(0, 0, 0), dxa[i + 1], r1x, U(-1, 0, 0), U(0, 0, 0), U(1, 0, 0));
            b = up1(v_in_um, r1yd, r1y, U(0, -1, 0), U(0, 0, 0), U(0, 1, 0));
            d = up1(w_in_um, r1zd, r1z, U(0, 0, -1), U(0, 0, 0), U(0, 0, 1));
        }
        const double d2x = (2.0 / dx2xr) * ((-u4 / dxr) + (u22 / dx));
        const double d2y = (2.0 / (0.5 * (dy2ye + dy2yw))) * ((-u8 / (0.5 * dy2ye)) + (u6 / (0.5 * dy2yw)));
        const double d2z = (2.0 / (0.5 * (dz2zt + dz2zb))) * ((-u12 / (0.5 * dz2zt)) + (u10 / (0.5 * dz2zb)));
        const double xtt2 = rev * ((d2x + d2y) + d2z);
        const double res = ((-a - b) - d) + xtt2;
        ut[c] = U(0, 0, 0) + deltat * (res + dpdx);
        ru[c] = res;
    }
    // V momentum (subNavierStokes.f90:901-1006)
    {
        const double uv_e = U(0, 1, 0) + (dye / dy2ye) * (U(0, 0, 0) - U(0, 1, 0));
        const double uv_w = U(-1, 1, 0) + (dye / dy2ye) * (U(-1, 0, 0) - U(-1, 1, 0));
        const double u_in_vm = 0.5 * (uv_e + uv_w);
        const double wv_n = W(0, 1, 0) + (dye / dy2ye) * (W(0, 0, 0) - W(0, 1, 0));
        const double wv_s = W(0, 1, -1) + (dye / dy2ye) * (W(0, 0, -1) - W(0, 1, -1));
        const double w_in_vm = 0.5 * (wv_n + wv_s);
        double a, b, d;
        if (i != 2 && i < nx + 1 && j != 2 && j < ny && k != 2 && k < nz + 1 && no_ib) {
            const double ddx = 0.5 * (dxa[i] + dxa[i - 1]), ddxr = 0.5 * (dxa[i] + dxa[i + 1]);
            const double ddz = 0.5 * (dza[k] + dza[k - 1]), ddzr = 0.5 * (dza[k] + dza[k + 1]);
            const double vc = V(0, 0, 0);
            a = kk3(T, UV, i, u_in_vm, V(2, 0, 0), V(1, 0, 0), vc, V(-1, 0, 0), V(-2, 0, 0), ddxr, ddx);
            b = kk3(T, VV, j, vc, V(0, 2, 0), V(0, 1, 0), vc, V(0, -1, 0), V(0, -2, 0), dya[j + 1], dya[j]);
            d = kk3(T, WU, k, w_in_vm, V(0, 0, 2), V(0, 0, 1), vc, V(0, 0, -1), V(0, 0, -2), ddzr, ddz);
        } else {
            const double r1xn = 0.5 * (dxa[i] + dxa[i - 1]), r1xd = 0.5 * (dxa[i] + dxa[i + 1]);
            const double r1x = r1xn / r1xd;
            const double r1y = dya[j] / dya[j + 1];
            const double r1zn = 0.5 * (dza[k] + dza[k - 1]), r1zd = 0.5 * (dza[k] + dza[k + 1]);
            const double r1z = r1zn / r1zd;
            a = up1(u_in_vm, r1xd, r1x, V(-1, 0, 0), V(0, 0, 0), V(1, 0, 0));
            b = up1(V(0, 0, 0), dya[j + 1], r1y, V(0, -1, 0), V(0, 0, 0), V(0, 1, 0));
            d = up1(w_in_vm, r1zd, r1z, V(0, 0, -1), V(0, 0, 0), V(0, 0, 1));
        }
        const double d2x = (2.0 / (0.5 * (dx2xr + dx2xl))) * ((-v6 / (0.5 * dx2xr)) + (v4 / (0.5 * dx2xl)));
        const double d2y = (2.0 / dy2ye) * ((-v10 / dye) + (v8 / dy));
        const double d2z = (2.0 / (0.5 * (dz2zt + dz2zb))) * ((-v14 / (0.5 * dz2zt)) + (v12 / (0.5 * dz2zb)));
        const double ytt2 = rev * ((d2x + d2y) + d2z);
        const double res = ((-a - b) - d) + ytt2;
        vt[c] = V(0, 0, 0) + deltat * (res + dpdy);
        rv[c] = res;
    }
    // W momentum (subNavierStokes.f90:1010-1116)
    {
        const double uw_e = U(0, 0, 1) + (dzt / dz2zt) * (U(0, 0, 0) - U(0, 0, 1));
        const double uw_w = U(-1, 0, 1) + (dzt / dz2zt) * (U(-1, 0, 0) - U(-1, 0, 1));
        const double u_in_wm = 0.5 * (uw_e + uw_w);
        const double vw_n = V(0, 0, 1) + (dzt / dz2zt) * (V(0, 0, 0) - V(0, 0, 1));
        const double vw_s = V(0, -1, 1) + (dzt / dz2zt) * (V(0, -1, 0) - V(0, -1, 1));
        const double v_in_wm = 0.5 * (vw_n + vw_s);
        double a, b, d;
        if (i != 2 && i < nx + 1 && j != 2 && j < ny + 1 && k != 2 && k < nz && no_ib) {
            const double ddx = 0.5 * (dxa[i] + dxa[i - 1]), ddxr = 0.5 * (dxa[i] + dxa[i + 1]);
            const double ddy = 0.5 * (dya[j] + dya[j - 1]), ddye = 0.5 * (dya[j] + dya[j + 1]);
            const double wc = W(0, 0, 0);
            a = kk3(T, UV, i, u_in_wm, W(2, 0, 0), W(1, 0, 0), wc, W(-1, 0, 0), W(-2, 0, 0), ddxr, ddx);
            b = kk3(T, VU, j, v_in_wm, W(0, 2, 0), W(0, 1, 0), wc, W(0, -1, 0), W(0, -2, 0), ddye, ddy);
            d = kk3(T, WW, k, wc, W(0, 0, 2), W(0, 0, 1), wc, W(0, 0, -1), W(0, 0, -2), dza[k + 1], dza[k]);
        } else {
            const double r1xn = 0.5 * (dxa[i] + dxa[i - 1]), r1xd = 0.5 * (dxa[i] + dxa[i + 1]);
            const double r1x = r1xn / r1xd;
            const double r1yn = 0.5 * (dya[j] + dya[j - 1]), r1yd = 0.5 * (dya[j] + dya[j + 1]);
            const double r1y = r1yn / r1yd;
            const double r1z = dza[k] / dza[k + 1];
            a = up1(u_in_wm, r1xd, r1x, W(-1, 0, 0), W(0, 0, 0), W(1, 0, 0));
            b = up1(v_in_wm, r1yd, r1y, W(0, -1, 0), W(0, 0, 0), W(0, 1, 0));
            d = up1(W(0, 0, 0), dza[k + 1], r1z, W(0, 0, -1), W(0, 0, 0), W(0, 0, 1));
        }
        const double d2x = (2.0 / (0.5 * (dx2xr + dx2xl))) * ((-w8 / (0.5 * dx2xr)) + (w6 / (0.5 * dx2xl)));
        const double d2y = (2.0 / (0.5 * (dy2ye + dy2yw))) * ((-w12 / (0.5 * dy2ye)) + (w10 / (0.5 * dy2yw)));
        const double d2z = (2.0 / dz2zt) * ((-w16 / dzt) + (w14 / dz));
        const double ztt2 = rev * ((d2x + d2y) + d2z);
        const double res = ((-a - b) - d) + ztt2;
        wt[c] = W(0, 0, 0) + deltat * (res + dpdz);
        rw[c] = res;
    }
#undef U
#undef V
#undef W
#undef P
#undef CL
}

// subBoundaryConditions.f90:141-156: inlet u = 1, reflected v and w; outlet
// zero gradient
__global__ void k_bc_x(double *ut, double *vt, double *wt, int nx, int ny, int nz) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int k = blockIdx.y + 2;
    if (j > ny + 1) return;
    const int64_t sx = nx + 2;
    const int64_t r = sx * ((j - 1) + (int64_t)(ny + 2) * (k - 1));  // index of (1,j,k)
    ut[r] = 1.0;
    vt[r] = -vt[r + 1];
    wt[r] = -wt[r + 1];
    ut[r + nx] = ut[r + nx - 1];
    vt[r + nx + 1] = vt[r + nx];
    wt[r + nx + 1] = wt[r + nx];
}

// subBoundaryConditions.f90:161-174: free slip at the y walls
__global__ void k_bc_y(double *ut, double *vt, double *wt, int nx, int ny, int nz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int k = blockIdx.y + 2;
    if (i > nx + 1) return;
    const int64_t sx = nx + 2, sxy = sx * (ny + 2);
    const int64_t b = (i - 1) + sxy * (k - 1);  // (i,1,k)
    const int64_t t = b + sx * (ny + 1);        // (i,ny+2,k)
    ut[b] = ut[b + sx];
    wt[b] = wt[b + sx];
    vt[b] = 0.0;
    ut[t] = ut[t - sx];
    wt[t] = wt[t - sx];
    vt[t - sx] = 0.0;
}

// subBoundaryConditions.f90:179-192: free slip at the z walls
__global__ void k_bc_z(double *ut, double *vt, double *wt, int nx, int ny, int nz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int j = blockIdx.y + 2;
    if (i > nx + 1) return;
    const int64_t sx = nx + 2, sxy = sx * (ny + 2);
    const int64_t b = (i - 1) + sx * (j - 1);  // (i,j,1)
    const int64_t t = b + sxy * (nz + 1);      // (i,j,nz+2)
    ut[b] = ut[b + sxy];
    vt[b] = vt[b + sxy];
    wt[b] = 0.0;
    ut[t] = ut[t - sxy];
    vt[t] = vt[t - sxy];
    wt[t - sxy] = 0.0;
}

// subBoundaryConditions.f90:207-215
__global__ void k_solid_bc(const int *list, int64_t n, double *ut, double *vt, double *wt, double *p, int nx,
                           int ny, double wu, double wv, double ww) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int i = list[q], j = list[n + q], k = list[2 * n + q];
    const int64_t c = (i - 1) + (int64_t)(nx + 2) * ((j - 1) + (int64_t)(ny + 2) * (k - 1));
    ut[c] = wu;   // 0 in the reference (a body at rest); the body's velocity when it moves
    vt[c] = wv;
    wt[c] = ww;
    p[c] = 0.0;
}

__device__ __forceinline__ double pc_at(const double *p0, const double *p1, int64_t H, int sy, int ii, int jj,
                                        int kk) {
    const int64_t f = ((int64_t)jj + (int64_t)sy * kk) * H + (ii >> 1);
    return ((ii + jj + kk) & 1) ? p1[f] : p0[f];
}

// subPcorVcor.f90:113-118 and 132-140, fused: p += pc and the velocity
// correction on the east/north/top faces of each fluid cell
__global__ void __launch_bounds__(256)
k_correct(double *__restrict__ p, double *__restrict__ ut, double *__restrict__ vt, double *__restrict__ wt,
          const double *__restrict__ pc0, const double *__restrict__ pc1, const int8_t *__restrict__ cell,
          const double *__restrict__ dxa, const double *__restrict__ dya, const double *__restrict__ dza, int nx,
          int ny, int nz, int64_t H, double dt, int do_p, int do_v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 2;
    const int k = blockIdx.z + 2;
    if (i > nx + 1 || j > ny + 1) return;
    const int sy = ny + 2;
    const int64_t c = (i - 1) + (int64_t)(nx + 2) * ((j - 1) + (int64_t)sy * (k - 1));
    if (cell[c] != 0) return;
    const int ii = i - 1, jj = j - 1, kk = k - 1;
    const double pcc = pc_at(pc0, pc1, H, sy, ii, jj, kk);
    if (do_p) p[c] = p[c] + pcc;
    if (do_v) {
        ut[c] = ut[c] - (dt / (0.5 * (dxa[i + 1] + dxa[i]))) * (pc_at(pc0, pc1, H, sy, ii + 1, jj, kk) - pcc);
        vt[c] = vt[c] - (dt / (0.5 * (dya[j + 1] + dya[j]))) * (pc_at(pc0, pc1, H, sy, ii, jj + 1, kk) - pcc);
        wt[c] = wt[c] - (dt / (0.5 * (dza[k + 1] + dza[k]))) * (pc_at(pc0, pc1, H, sy, ii, jj, kk + 1) - pcc);
    }
}

// subPcorVcor.f90:45-77: derrStdSt = max |ut-u|/dt over fluid cells, then
// u = ut over the whole padded grid
__global__ void __launch_bounds__(256)
k_steady_copy(double *__restrict__ u, double *__restrict__ v, double *__restrict__ w, const double *__restrict__ ut,
              const double *__restrict__ vt, const double *__restrict__ wt, const int8_t *__restrict__ cell, int nx,
              int ny, int nz, double dt, double *partial) {
    const int P = (nx + 2) * (ny + 2) * (nz + 2);  // < 2^31
    const int c = (int)blockIdx.x * (int)blockDim.x + (int)threadIdx.x;
    double m = 0.0;
    if (c < P) {
        const double a = ut[c], b = vt[c], d = wt[c];
        if (cell[c] == 0) {
            const int row = c / (nx + 2);
            const int ii = c - row * (nx + 2);
            const int kk = row / (ny + 2);
            const int jj = row - kk * (ny + 2);
            if (ii >= 1 && ii <= nx && jj >= 1 && jj <= ny && kk >= 1 && kk <= nz) {
                const double du = fabs(a - u[c]) / dt, dv = fabs(b - v[c]) / dt, dw = fabs(d - w[c]) / dt;
                m = fmax(fmax(du, dv), dw);
                if (du != du || dv != dv || dw != dw) m = __builtin_nan("");
            }
        }
        u[c] = a;
        v[c] = b;
        w[c] = d;
    }
    block_max_partial(m, partial);
}

// ------------------------------------------------------------------ host
static Tabs tabs(const Ctx &c) {
    Tabs T;
    for (int s = 0; s < 6; s++) {
        T.t[s] = c.tab[s];
        T.tn[s] = (int)c.tn[s];
    }
    return T;
}

int ns_momentum(Ctx &c, double rev, double deltat) {
    IB3D_REQUIRE(c.coeff_ready, "nsMomentum: non_uni_coeff has not run");
    IB3D_REQUIRE(c.lists_ready, "nsMomentum: cellCount has not run");
    dim3 blk(64, 4, 1), grd(nblocks(c.nx, 64), nblocks(c.ny, 4), (unsigned)c.nz);
    hipLaunchKernelGGL(k_momentum, grd, blk, 0, c.stream, c.u, c.v, c.w, c.p, c.tag, c.ut, c.vt, c.wt, c.resi_u,
                       c.resi_v, c.resi_w, c.dx, c.dy, c.dz, tabs(c), (int)c.nx, (int)c.ny, (int)c.nz, rev, deltat);
    IB3D_CHECK(hipGetLastError());
    return 0;
}

int velocity_bc(Ctx &c) {
    const int nx = (int)c.nx, ny = (int)c.ny, nz = (int)c.nz;
    hipLaunchKernelGGL(k_bc_x, dim3(nblocks(ny, 64), nz), dim3(64), 0, c.stream, c.ut, c.vt, c.wt, nx, ny, nz);
    hipLaunchKernelGGL(k_bc_y, dim3(nblocks(nx, 64), nz), dim3(64), 0, c.stream, c.ut, c.vt, c.wt, nx, ny, nz);
    hipLaunchKernelGGL(k_bc_z, dim3(nblocks(nx, 64), ny), dim3(64), 0, c.stream, c.ut, c.vt, c.wt, nx, ny, nz);
    IB3D_CHECK(hipGetLastError());
    return 0;
}

int solid_bc(Ctx &c) {
    IB3D_REQUIRE(c.lists_ready, "solidCellBC: cellCount has not run");
    if (c.solidCount == 0) return 0;
    hipLaunchKernelGGL(k_solid_bc, dim3(nblocks(c.solidCount, 256)), dim3(256), 0, c.stream, c.solidList,
                       c.solidCount, c.ut, c.vt, c.wt, c.p, (int)c.nx, (int)c.ny, c.wall_vel[0], c.wall_vel[1],
                       c.wall_vel[2]);
    IB3D_CHECK(hipGetLastError());
    return 0;
}

int correct(Ctx &c, double dt, int do_p, int do_v) {
    dim3 blk(64, 4, 1), grd(nblocks(c.nx, 64), nblocks(c.ny, 4), (unsigned)c.nz);
    hipLaunchKernelGGL(k_correct, grd, blk, 0, c.stream, c.p, c.ut, c.vt, c.wt, c.pc[0], c.pc[1], c.tag, c.dx, c.dy,
                       c.dz, (int)c.nx, (int)c.ny, (int)c.nz, c.H, dt, do_p, do_v);
    IB3D_CHECK(hipGetLastError());
    return 0;
}

int steady_copy(Ctx &c, double dt, double *derr_std) {
    const unsigned nb = nblocks(c.P, 256);
    if (ensure_partial(nb)) return -1;
    hipLaunchKernelGGL(k_steady_copy, dim3(nb), dim3(256), 0, c.stream, c.u, c.v, c.w, c.ut, c.vt,
                       c.wt, c.tag, (int)c.nx, (int)c.ny, (int)c.nz, dt, c.partial);
    launch_max_partials(c.partial, (int64_t)nb, c.scalars + 3);
    IB3D_CHECK(hipGetLastError());
    IB3D_CHECK(hipMemcpyAsync(c.hscal + 3, c.scalars + 3, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    IB3D_CHECK(hipStreamSynchronize(c.stream));
    *derr_std = c.hscal[3];
    return 0;
}

}  // namespace ib3d
