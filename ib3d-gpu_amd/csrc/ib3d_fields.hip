// ib3d_fields.hip -- momentum predictor, boundary conditions, projection
// corrections and the steady-state check (subNavierStokes.f90:622-1122,
// subBoundaryConditions.f90:137-217, subPcorVcor.f90:2-142).
//
// The reference iterates gathered fluid lists; each fluid cell's update is
// independent of the others, so these kernels walk the padded grid in
// Fortran order (coalesced, i fastest) and test the int8 tag instead.
// Arithmetic is written operand for operand like the reference (built with
// -ffp-contract=off, correctly rounded division), so results are bitwise.
// The predictor takes its index-only sub-expressions from per-axis tables
// computed on the host in the same operations, and divides field values
// through correctly rounded reciprocals (dv below), which is also exact.
#include "ib3d_internal.h"

#include <type_traits>

namespace ib3d {

struct Tabs {
    const double *t[6];
    int tn[6];
};

enum { UU = 0, VV = 1, WW = 2, UV = 3, VU = 4, WU = 5 };


// Per-axis momentum tables, quantity-major with n+3 entries per quantity.
// Each entry is a sub-expression of the momentum formula that depends on
// one grid index only, evaluated once on the host in the same IEEE
// operations as the formula (so reading it changes no bit), together with
// the correctly rounded reciprocal of every denominator that divides a
// field value.  "S" quantities serve the velocity component staggered along
// the axis (u along x), "C" the two components across it.
enum MomQ {
    M_AS,        // d(i+1) / (d(i) + d(i+1)): face interpolation weight
    M_CS,        // 2 / (d(i) + d(i+1)): second-difference weight
    M_D1, M_RD1, // d(i+1) and 1/d(i+1)
    M_D0, M_RD0, // d(i) and 1/d(i)
    M_KS1, M_RKS1, M_KS2, M_RKS2,  // face table: t5*d(i+1), (2 t11)*d(i), reciprocals
    M_US,        // 7 upwind factors at den = d(i+1), r = d(i)/d(i+1)
    M_CC = M_US + 7,               // 2 / (0.5 ((d(i)+d(i+1)) + (d(i)+d(i-1))))
    M_H1, M_RH1, // h1 = 0.5 (d(i) + d(i+1)) and 1/h1
    M_H0, M_RH0, // h0 = 0.5 (d(i) + d(i-1)) and 1/h0
    M_KC1, M_RKC1, M_KC2, M_RKC2,  // centre table: t5*h1, (2 t11)*h0, reciprocals
    M_UC,        // 7 upwind factors at den = h1, r = h0/h1
    M_NQ = M_UC + 7
};

typedef __attribute__((address_space(4))) double CDouble;

struct MomTabs {
    const double *m[3];    // quantity-major tables (x axis: lanes read consecutive entries)
    int mn[3];
    const double *rec[3];  // y, z axes: one kMomRec-double record per index (see below)
};

// y and z records: the M_NQ quantities, then the face (staggered) and the
// centre (across) convection coefficients of that axis.  A wave shares one
// j and one k, so a record arrives through a few wide scalar loads.
static constexpr int kMomRec = 64, kRecS = M_NQ, kRecC = M_NQ + 12;
static_assert(kRecC + 12 <= kMomRec, "momentum record overflow");

// a / b, correctly rounded, from y = RN(1/b) and two FMA corrections of
// a*y.  The first brings the quotient within one ulp; by Markstein's
// theorem (y within half an ulp of 1/b, quotient within one ulp) the
// remainder a - b*q is then exact and the second correction rounds
// correctly.  The host enables this only when every denominator lies in
// [2^-64, 2^64]; numerators outside [2^-900, 2^900] (zeros, NaN and
// infinities among them) use the division instruction, so no intermediate
// leaves the normal range and signed zeros come out as from a / b.
template <bool FAST>
__device__ __forceinline__ double dv(double a, double b, double y) {
    if (!FAST) return a / b;
    const double m = fabs(a);
    if (!(m >= 0x1p-900 && m <= 0x1p900)) return a / b;
    double q = a * y;
    double r = fma(-b, q, a);
    q = fma(r, y, q);
    r = fma(-b, q, a);
    return fma(r, y, q);
}

// 3rd-order Kawamura-Kuwahara convection term (subNavierStokes.f90:812-831);
// dA = t5*dA', dB = (2 t11)*dB' come from the axis table with their reciprocals
// (coefficient c of the stencil at tc[c * ts])
template <bool FAST, class PT>
__device__ __forceinline__ double kk3(PT tc, int ts, double vel, double f2, double f1, double f0,
                                      double fm1, double fm2, double dA, double rA, double dB, double rB) {
#define TC(c) tc[(c) * ts]
    const double sa = (((TC(0) * f2 + TC(1) * f1) + TC(2) * f0) + TC(3) * fm1) + TC(4) * fm2;
    const double sk = (((TC(6) * f2 + TC(7) * f1) + TC(8) * f0) + TC(9) * fm1) + TC(10) * fm2;
#undef TC
    return dv<FAST>(vel * sa, dA, rA) + dv<FAST>(fabs(vel) * sk, dB, rB);
}

// non-uniform 1st-order upwind term (subNavierStokes.f90:847-866), al = 1;
// g[0..6*n] = -1/(r(r+1)), 1-1/r, r/(r+1), 2/(r(r+1)), 2/r, 2/(r+1), 1/(2 den)
template <bool FAST, class PT>
__device__ __forceinline__ double up1(PT g, int n, double vel, double den, double rden, double fm,
                                      double f0, double fp) {
    const double t1 = ((g[0] * fm - g[n] * f0) + g[2 * n] * fp);
    const double t2 = (g[3 * n] * fm - g[4 * n] * f0) + g[5 * n] * fp;
    return dv<FAST>(vel, den, rden) * t1 - (fabs(vel) * g[6 * n]) * t2;
}

// Where the predictor reads u, v, w at offset (a, b, d) from its cell, and
// how it addresses the other arrays of the padded grid (at<a, b, d>(base):
// the element at that offset).
// GlobalFields: straight from the padded fields in HBM, 64-bit indices.
struct GlobalFields {
    const double *__restrict__ u, *__restrict__ v, *__restrict__ w;
    int64_t c, sx, sxy;
    template <int a, int b, int d> __device__ __forceinline__ double U() const { return u[c + a + b * sx + d * sxy]; }
    template <int a, int b, int d> __device__ __forceinline__ double V() const { return v[c + a + b * sx + d * sxy]; }
    template <int a, int b, int d> __device__ __forceinline__ double W() const { return w[c + a + b * sx + d * sxy]; }
    template <int a, int b, int d, class T> __device__ __forceinline__ T &at(T *base) const {
        return base[c + a + b * sx + d * sxy];
    }
    __device__ __forceinline__ void split() const {}
};

// GlobalFields32: the same values through 32-bit byte offsets from the
// uniform array bases, one offset register per address instead of a 64-bit
// pair (fewer registers, one more wave per SIMD).  For grids whose FP64
// fields span less than 4 GiB.
struct GlobalFields32 {
    const double *__restrict__ u, *__restrict__ v, *__restrict__ w;
    uint32_t c, sx, sxy;  // element offsets; c * 8 < 2^32
    template <int a, int b, int d, int S> __device__ __forceinline__ uint32_t off() const {
        return (c + (uint32_t)a + (uint32_t)b * sx + (uint32_t)d * sxy) * (uint32_t)S;
    }
    template <int a, int b, int d> __device__ __forceinline__ double U() const { return at<a, b, d>(u); }
    template <int a, int b, int d> __device__ __forceinline__ double V() const { return at<a, b, d>(v); }
    template <int a, int b, int d> __device__ __forceinline__ double W() const { return at<a, b, d>(w); }
    template <int a, int b, int d, class T> __device__ __forceinline__ T &at(T *base) const {
        typedef std::conditional_t<std::is_const_v<T>, const char, char> B;
        return *reinterpret_cast<T *>(reinterpret_cast<B *>(base) + off<a, b, d, (int)sizeof(T)>());
    }
    __device__ __forceinline__ void split() const {}
};

// TiledFields: the cell's own plane from an LDS tile (row stride kTileS),
// its own column's planes k-2..k+2 from registers, and the four values
// off both the plane and the column from HBM.
static constexpr int kMomTX = 64, kMomTZ = 8;
static constexpr int kTileS = kMomTX + 4;
struct TiledFields {
    const double *tu, *tv, *tw;          // LDS, at the cell
    double qu[5], qv[5], qw[5];          // registers, planes k-2..k+2
    const double *__restrict__ u, *__restrict__ v, *__restrict__ w;
    int64_t c, sx, sxy;
    template <int a, int b, int d> __device__ __forceinline__ double U() const {
        if constexpr (d == 0) return tu[a + b * kTileS];
        else if constexpr (a == 0 && b == 0) return qu[d + 2];
        else return u[c + a + b * sx + d * sxy];
    }
    template <int a, int b, int d> __device__ __forceinline__ double V() const {
        if constexpr (d == 0) return tv[a + b * kTileS];
        else if constexpr (a == 0 && b == 0) return qv[d + 2];
        else return v[c + a + b * sx + d * sxy];
    }
    template <int a, int b, int d> __device__ __forceinline__ double W() const {
        if constexpr (d == 0) return tw[a + b * kTileS];
        else if constexpr (a == 0 && b == 0) return qw[d + 2];
        else return w[c + a + b * sx + d * sxy];
    }
    template <int a, int b, int d, class T> __device__ __forceinline__ T &at(T *base) const {
        return base[c + a + b * sx + d * sxy];
    }
    // the three components are scheduled one after the other: their values
    // come from LDS, so loading them late costs little, and keeping the
    // components apart keeps the register count down
    __device__ __forceinline__ void split() const { __builtin_amdgcn_sched_barrier(0); }
};

// One fluid cell of the explicit predictor (subNavierStokes.f90:622-1122).
// The Kawamura-Kuwahara stencil applies away from the walls and from
// intercepted cells, the upwind stencil elsewhere.  F supplies the velocity
// values (GlobalFields or TiledFields: the same values, so the same bits).
template <bool FAST, class F>
__device__ __forceinline__ void momentum_cell(
    const F &A, int i, int j, int k, const double *__restrict__ p, const int8_t *__restrict__ cell,
    double *__restrict__ ut, double *__restrict__ vt, double *__restrict__ wt, double *__restrict__ ru,
    double *__restrict__ rv, double *__restrict__ rw, const Tabs &T, const MomTabs &M, int nx, int ny, int nz,
    double rev, double deltat) {
    if (A.template at<0, 0, 0>(cell) != 0) return;
#define U(a, b, d) A.template U<a, b, d>()
#define V(a, b, d) A.template V<a, b, d>()
#define W(a, b, d) A.template W<a, b, d>()
#define P(a, b, d) A.template at<a, b, d>(p)
#define CL(a, b, d) A.template at<a, b, d>(cell)
#define AT(f) A.template at<0, 0, 0>(f)
    // the y and z records go through the constant address space: the kernel
    // never writes them, so uniform reads of them stay scalar loads even after
    // the stores of an earlier cell in the same thread
    const CDouble *ry = (const CDouble *)(M.rec[1] + (int64_t)j * kMomRec);
    const CDouble *rz = (const CDouble *)(M.rec[2] + (int64_t)k * kMomRec);
    const double *tuu = T.t[UU] + i, *tuv = T.t[UV] + i;
    const int nuu = T.tn[UU], nuv = T.tn[UV];
#define MX(q) M.m[0][(q) * M.mn[0] + i]
#define MY(q) ry[q]
#define MZ(q) rz[q]
    // each component's differences and pressure gradient are formed inside
    // its own block, so that none of them stays live across another
    // component
    const bool no_ib = CL(1, 0, 0) != 2 && CL(-1, 0, 0) != 2 && CL(0, 1, 0) != 2 && CL(0, -1, 0) != 2 &&
                       CL(0, 0, 1) != 2 && CL(0, 0, -1) != 2;
    const int mx = M.mn[0], my = 1, mz = 1;

    // U momentum (subNavierStokes.f90:792-897)
    {
        const double u22 = U(-1, 0, 0) - U(0, 0, 0), u4 = U(0, 0, 0) - U(1, 0, 0);
        const double u6 = U(0, -1, 0) - U(0, 0, 0), u8 = U(0, 0, 0) - U(0, 1, 0);
        const double u10 = U(0, 0, -1) - U(0, 0, 0), u12 = U(0, 0, 0) - U(0, 0, 1);
        const double dpdx = dv<FAST>(P(0, 0, 0) - P(1, 0, 0), MX(M_H1), MX(M_RH1));
        const double ax = MX(M_AS);
        const double vu_e = V(1, 0, 0) + ax * (V(0, 0, 0) - V(1, 0, 0));
        const double vu_w = V(1, -1, 0) + ax * (V(0, -1, 0) - V(1, -1, 0));
        const double v_in_um = 0.5 * (vu_e + vu_w);
        const double wu_n = W(1, 0, 0) + ax * (W(0, 0, 0) - W(1, 0, 0));
        const double wu_s = W(1, 0, -1) + ax * (W(0, 0, -1) - W(1, 0, -1));
        const double w_in_um = 0.5 * (wu_n + wu_s);
        double a, b, d;
        if (i != 2 && i < nx && j != 2 && j < ny + 1 && k != 2 && k < nz + 1 && no_ib) {
            const double uc = U(0, 0, 0);
            a = kk3<FAST>(tuu, nuu, uc, U(2, 0, 0), U(1, 0, 0), uc, U(-1, 0, 0), U(-2, 0, 0), MX(M_KS1), MX(M_RKS1),
                          MX(M_KS2), MX(M_RKS2));
            b = kk3<FAST>(ry + kRecC, 1, v_in_um, U(0, 2, 0), U(0, 1, 0), uc, U(0, -1, 0), U(0, -2, 0), MY(M_KC1),
                          MY(M_RKC1), MY(M_KC2), MY(M_RKC2));
            d = kk3<FAST>(rz + kRecC, 1, w_in_um, U(0, 0, 2), U(0, 0, 1), uc, U(0, 0, -1), U(0, 0, -2), MZ(M_KC1),
                          MZ(M_RKC1), MZ(M_KC2), MZ(M_RKC2));
        } else {
            a = up1<FAST>(&MX(M_US), mx, U(0, 0, 0), MX(M_D1), MX(M_RD1), U(-1, 0, 0), U(0, 0, 0), U(1, 0, 0));
            b = up1<FAST>(&MY(M_UC), my, v_in_um, MY(M_H1), MY(M_RH1), U(0, -1, 0), U(0, 0, 0), U(0, 1, 0));
            d = up1<FAST>(&MZ(M_UC), mz, w_in_um, MZ(M_H1), MZ(M_RH1), U(0, 0, -1), U(0, 0, 0), U(0, 0, 1));
        }
        const double d2x = MX(M_CS) * (dv<FAST>(-u4, MX(M_D1), MX(M_RD1)) + dv<FAST>(u22, MX(M_D0), MX(M_RD0)));
        const double d2y = MY(M_CC) * (dv<FAST>(-u8, MY(M_H1), MY(M_RH1)) + dv<FAST>(u6, MY(M_H0), MY(M_RH0)));
        const double d2z = MZ(M_CC) * (dv<FAST>(-u12, MZ(M_H1), MZ(M_RH1)) + dv<FAST>(u10, MZ(M_H0), MZ(M_RH0)));
        const double xtt2 = rev * ((d2x + d2y) + d2z);
        const double res = ((-a - b) - d) + xtt2;
        AT(ut) = U(0, 0, 0) + deltat * (res + dpdx);
        AT(ru) = res;
    }
    A.split();
    // V momentum (subNavierStokes.f90:901-1006)
    {
        const double v4 = V(-1, 0, 0) - V(0, 0, 0), v6 = V(0, 0, 0) - V(1, 0, 0);
        const double v8 = V(0, -1, 0) - V(0, 0, 0), v10 = V(0, 0, 0) - V(0, 1, 0);
        const double v12 = V(0, 0, -1) - V(0, 0, 0), v14 = V(0, 0, 0) - V(0, 0, 1);
        const double dpdy = dv<FAST>(P(0, 0, 0) - P(0, 1, 0), MY(M_H1), MY(M_RH1));
        const double ay = MY(M_AS);
        const double uv_e = U(0, 1, 0) + ay * (U(0, 0, 0) - U(0, 1, 0));
        const double uv_w = U(-1, 1, 0) + ay * (U(-1, 0, 0) - U(-1, 1, 0));
        const double u_in_vm = 0.5 * (uv_e + uv_w);
        const double wv_n = W(0, 1, 0) + ay * (W(0, 0, 0) - W(0, 1, 0));
        const double wv_s = W(0, 1, -1) + ay * (W(0, 0, -1) - W(0, 1, -1));
        const double w_in_vm = 0.5 * (wv_n + wv_s);
        double a, b, d;
        if (i != 2 && i < nx + 1 && j != 2 && j < ny && k != 2 && k < nz + 1 && no_ib) {
            const double vc = V(0, 0, 0);
            a = kk3<FAST>(tuv, nuv, u_in_vm, V(2, 0, 0), V(1, 0, 0), vc, V(-1, 0, 0), V(-2, 0, 0), MX(M_KC1),
                          MX(M_RKC1), MX(M_KC2), MX(M_RKC2));
            b = kk3<FAST>(ry + kRecS, 1, vc, V(0, 2, 0), V(0, 1, 0), vc, V(0, -1, 0), V(0, -2, 0), MY(M_KS1), MY(M_RKS1),
                          MY(M_KS2), MY(M_RKS2));
            d = kk3<FAST>(rz + kRecC, 1, w_in_vm, V(0, 0, 2), V(0, 0, 1), vc, V(0, 0, -1), V(0, 0, -2), MZ(M_KC1),
                          MZ(M_RKC1), MZ(M_KC2), MZ(M_RKC2));
        } else {
            a = up1<FAST>(&MX(M_UC), mx, u_in_vm, MX(M_H1), MX(M_RH1), V(-1, 0, 0), V(0, 0, 0), V(1, 0, 0));
            b = up1<FAST>(&MY(M_US), my, V(0, 0, 0), MY(M_D1), MY(M_RD1), V(0, -1, 0), V(0, 0, 0), V(0, 1, 0));
            d = up1<FAST>(&MZ(M_UC), mz, w_in_vm, MZ(M_H1), MZ(M_RH1), V(0, 0, -1), V(0, 0, 0), V(0, 0, 1));
        }
        const double d2x = MX(M_CC) * (dv<FAST>(-v6, MX(M_H1), MX(M_RH1)) + dv<FAST>(v4, MX(M_H0), MX(M_RH0)));
        const double d2y = MY(M_CS) * (dv<FAST>(-v10, MY(M_D1), MY(M_RD1)) + dv<FAST>(v8, MY(M_D0), MY(M_RD0)));
        const double d2z = MZ(M_CC) * (dv<FAST>(-v14, MZ(M_H1), MZ(M_RH1)) + dv<FAST>(v12, MZ(M_H0), MZ(M_RH0)));
        const double ytt2 = rev * ((d2x + d2y) + d2z);
        const double res = ((-a - b) - d) + ytt2;
        AT(vt) = V(0, 0, 0) + deltat * (res + dpdy);
        AT(rv) = res;
    }
    A.split();
    // W momentum (subNavierStokes.f90:1010-1116)
    {
        const double w6 = W(-1, 0, 0) - W(0, 0, 0), w8 = W(0, 0, 0) - W(1, 0, 0);
        const double w10 = W(0, -1, 0) - W(0, 0, 0), w12 = W(0, 0, 0) - W(0, 1, 0);
        const double w14 = W(0, 0, -1) - W(0, 0, 0), w16 = W(0, 0, 0) - W(0, 0, 1);
        const double dpdz = dv<FAST>(P(0, 0, 0) - P(0, 0, 1), MZ(M_H1), MZ(M_RH1));
        const double az = MZ(M_AS);
        const double uw_e = U(0, 0, 1) + az * (U(0, 0, 0) - U(0, 0, 1));
        const double uw_w = U(-1, 0, 1) + az * (U(-1, 0, 0) - U(-1, 0, 1));
        const double u_in_wm = 0.5 * (uw_e + uw_w);
        const double vw_n = V(0, 0, 1) + az * (V(0, 0, 0) - V(0, 0, 1));
        const double vw_s = V(0, -1, 1) + az * (V(0, -1, 0) - V(0, -1, 1));
        const double v_in_wm = 0.5 * (vw_n + vw_s);
        double a, b, d;
        if (i != 2 && i < nx + 1 && j != 2 && j < ny + 1 && k != 2 && k < nz && no_ib) {
            const double wc = W(0, 0, 0);
            a = kk3<FAST>(tuv, nuv, u_in_wm, W(2, 0, 0), W(1, 0, 0), wc, W(-1, 0, 0), W(-2, 0, 0), MX(M_KC1),
                          MX(M_RKC1), MX(M_KC2), MX(M_RKC2));
            b = kk3<FAST>(ry + kRecC, 1, v_in_wm, W(0, 2, 0), W(0, 1, 0), wc, W(0, -1, 0), W(0, -2, 0), MY(M_KC1),
                          MY(M_RKC1), MY(M_KC2), MY(M_RKC2));
            d = kk3<FAST>(rz + kRecS, 1, wc, W(0, 0, 2), W(0, 0, 1), wc, W(0, 0, -1), W(0, 0, -2), MZ(M_KS1), MZ(M_RKS1),
                          MZ(M_KS2), MZ(M_RKS2));
        } else {
            a = up1<FAST>(&MX(M_UC), mx, u_in_wm, MX(M_H1), MX(M_RH1), W(-1, 0, 0), W(0, 0, 0), W(1, 0, 0));
            b = up1<FAST>(&MY(M_UC), my, v_in_wm, MY(M_H1), MY(M_RH1), W(0, -1, 0), W(0, 0, 0), W(0, 1, 0));
            d = up1<FAST>(&MZ(M_US), mz, W(0, 0, 0), MZ(M_D1), MZ(M_RD1), W(0, 0, -1), W(0, 0, 0), W(0, 0, 1));
        }
        const double d2x = MX(M_CC) * (dv<FAST>(-w8, MX(M_H1), MX(M_RH1)) + dv<FAST>(w6, MX(M_H0), MX(M_RH0)));
        const double d2y = MY(M_CC) * (dv<FAST>(-w12, MY(M_H1), MY(M_RH1)) + dv<FAST>(w10, MY(M_H0), MY(M_RH0)));
        const double d2z = MZ(M_CS) * (dv<FAST>(-w16, MZ(M_D1), MZ(M_RD1)) + dv<FAST>(w14, MZ(M_D0), MZ(M_RD0)));
        const double ztt2 = rev * ((d2x + d2y) + d2z);
        const double res = ((-a - b) - d) + ztt2;
        AT(wt) = W(0, 0, 0) + deltat * (res + dpdz);
        AT(rw) = res;
    }
#undef U
#undef V
#undef W
#undef P
#undef CL
#undef AT
#undef MX
#undef MY
#undef MZ
}

// Interior columns i in [3, nx-1], where the stencil choice depends on j, k
// and the tags only.  A block is 64 x 4 with one wave per (j, k) row, so j
// and k are wave-uniform: their table entries come through scalar loads and
// the wall tests are scalar branches.
// With 32-bit offsets and the fast divisions the kernel fits 96 registers:
// five waves per SIMD instead of four (the other forms need more, and
// keep the compiler's choice)
template <bool FAST, bool OFF32>
__global__ void __launch_bounds__(256, (FAST && OFF32) ? 5 : 1)
k_momentum(const double *__restrict__ u, const double *__restrict__ v, const double *__restrict__ w,
           const double *__restrict__ p, const int8_t *__restrict__ cell, double *__restrict__ ut,
           double *__restrict__ vt, double *__restrict__ wt, double *__restrict__ ru, double *__restrict__ rv,
           double *__restrict__ rw, Tabs T, MomTabs M, int nx, int ny, int nz, double rev, double deltat) {
    const int i = blockIdx.x * 64 + threadIdx.x + 3;
    const int j = __builtin_amdgcn_readfirstlane((int)(blockIdx.y * 4 + threadIdx.y)) + 2;
    const int k = (int)blockIdx.z + 2;
    if (j > ny + 1 || i > nx - 1) return;
    const int64_t sx = nx + 2, sxy = sx * (ny + 2);
    const int64_t c = (i - 1) + sx * ((j - 1) + (int64_t)(ny + 2) * (k - 1));
    if constexpr (OFF32) {
        const GlobalFields32 A{u, v, w, (uint32_t)c, (uint32_t)sx, (uint32_t)sxy};
        momentum_cell<FAST>(A, i, j, k, p, cell, ut, vt, wt, ru, rv, rw, T, M, nx, ny, nz, rev, deltat);
    } else {
        const GlobalFields A{u, v, w, c, sx, sxy};
        momentum_cell<FAST>(A, i, j, k, p, cell, ut, vt, wt, ru, rv, rw, T, M, nx, ny, nz, rev, deltat);
    }
}

// The same interior columns, a 64 x 8 tile of (i, j) marching through a chunk
// of planes.  Every plane's u, v, w over the tile and a two-cell ring (the
// corners the cross terms reach included) are staged once in LDS, each
// thread keeps its own column's planes k-2..k+2 in registers, so a value
// leaves HBM or L2 about once per plane instead of once per stencil point.
// Two LDS buffers alternate between planes: one barrier per plane.
template <bool FAST, int kMomTY, int OCC>
__global__ void __launch_bounds__(kMomTX * kMomTY, OCC)
k_momentum_lds(const double *__restrict__ u, const double *__restrict__ v, const double *__restrict__ w,
               const double *__restrict__ p, const int8_t *__restrict__ cell, double *__restrict__ ut,
               double *__restrict__ vt, double *__restrict__ wt, double *__restrict__ ru, double *__restrict__ rv,
               double *__restrict__ rw, Tabs T, MomTabs M, int nx, int ny, int nz, double rev, double deltat) {
    constexpr int kTileR = kMomTY + 4;
    __shared__ double tile[2][3][kTileR * kTileS];
    const int tx = (int)threadIdx.x, ty = (int)threadIdx.y, tid = ty * kMomTX + tx;
    const int i0 = 3 + (int)blockIdx.x * kMomTX, j0 = 2 + (int)blockIdx.y * kMomTY;
    const int i = i0 + tx;
    const int j = __builtin_amdgcn_readfirstlane(j0 + ty);
    const int kb = 2 + (int)blockIdx.z * kMomTZ, ke = min(kb + kMomTZ, nz + 2);
    const int64_t sx = nx + 2, sxy = sx * (ny + 2);
    const bool mine_in = i <= nx + 2 && j <= ny + 2;   // (i, j) inside the padded grid
    const bool compute = i <= nx - 1 && j <= ny + 1;
    const int64_t col = (int64_t)(i - 1) + sx * (j - 1);
    // own column, planes k-2..k+2 (0 outside the padded range)
    double qu[5], qv[5], qw[5];
#pragma unroll
    for (int d = -2; d <= 2; d++) {
        const int kk = kb + d;
        const bool in = mine_in && kk >= 1 && kk <= nz + 2;
        const int64_t at = col + sxy * (kk - 1);
        qu[d + 2] = in ? u[at] : 0.0;
        qv[d + 2] = in ? v[at] : 0.0;
        qw[d + 2] = in ? w[at] : 0.0;
    }
    // ring cells of a tile: rows j0..j0+TY-1 at x offsets -2, -1, 64, 65 (4 TY
    // cells), rows -2, -1, TY, TY+1 over the 64 columns (256 cells), and the
    // corners (64, -1) and (-1, TY) that V(1,-1,0) and U(-1,1,0) reach
    constexpr int nside = 4 * kMomTY, nrow = 4 * kMomTX, nring = nside + nrow + 2;
#pragma unroll 1
    for (int k = kb; k < ke; k++) {
        double *t = &tile[(k - kb) & 1][0][0];
        const int64_t pk = sxy * (k - 1);
        const int at = (ty + 2) * kTileS + tx + 2;
        t[at] = qu[2];
        t[kTileR * kTileS + at] = qv[2];
        t[2 * kTileR * kTileS + at] = qw[2];
#pragma unroll 1
        for (int r = tid; r < 3 * nring; r += kMomTX * kMomTY) {
            const int f = r / nring, e = r - f * nring;
            const double *src = f == 0 ? u : (f == 1 ? v : w);
            int ox, oy;
            if (e < nside) {
                const int s4 = e & 3;
                ox = s4 < 2 ? s4 - 2 : kMomTX + s4 - 2;
                oy = e >> 2;
            } else if (e < nside + nrow) {
                const int e2 = e - nside, s4 = e2 / kMomTX;
                ox = e2 - s4 * kMomTX;
                oy = s4 < 2 ? s4 - 2 : kMomTY + s4 - 2;
            } else if (e == nside + nrow) {
                ox = kMomTX;
                oy = -1;
            } else {
                ox = -1;
                oy = kMomTY;
            }
            const int gi = i0 + ox, gj = j0 + oy;
            const bool in = gi >= 1 && gi <= nx + 2 && gj >= 1 && gj <= ny + 2;
            t[f * kTileR * kTileS + (oy + 2) * kTileS + ox + 2] =
                in ? src[(int64_t)(gi - 1) + sx * (gj - 1) + pk] : 0.0;
        }
        __syncthreads();
        if (compute) {
            // keep the x-axis tables' loads inside the loop: hoisted out of it,
            // their ~40 per-lane values would occupy registers for all planes
            MomTabs Ml = M;
            Tabs Tl = T;
            asm volatile("" : "+s"(Ml.m[0]));
            asm volatile("" : "+s"(Tl.t[UU]));
            asm volatile("" : "+s"(Tl.t[UV]));
            const TiledFields A{t + at, t + kTileR * kTileS + at, t + 2 * kTileR * kTileS + at,
                                {qu[0], qu[1], qu[2], qu[3], qu[4]}, {qv[0], qv[1], qv[2], qv[3], qv[4]},
                                {qw[0], qw[1], qw[2], qw[3], qw[4]}, u, v, w, col + pk, sx, sxy};
            momentum_cell<FAST>(A, i, j, k, p, cell, ut, vt, wt, ru, rv, rw, Tl, Ml, nx, ny, nz, rev, deltat);
        }
#pragma unroll
        for (int d = 0; d < 4; d++) {
            qu[d] = qu[d + 1];
            qv[d] = qv[d + 1];
            qw[d] = qw[d + 1];
        }
        const int kk = k + 3;
        const bool in = mine_in && kk <= nz + 2;
        const int64_t at3 = col + sxy * (kk - 1);
        qu[4] = in ? u[at3] : 0.0;
        qv[4] = in ? v[at3] : 0.0;
        qw[4] = in ? w[at3] : 0.0;
    }
}

// The wall columns i = 2, nx, nx+1 of every (j, k) row, kept out of the
// interior waves so that those run one stencil branch.
template <bool FAST>
__global__ void __launch_bounds__(256)
k_momentum_edge(const double *__restrict__ u, const double *__restrict__ v, const double *__restrict__ w,
                const double *__restrict__ p, const int8_t *__restrict__ cell, double *__restrict__ ut,
                double *__restrict__ vt, double *__restrict__ wt, double *__restrict__ ru, double *__restrict__ rv,
                double *__restrict__ rw, Tabs T, MomTabs M, int nx, int ny, int nz, double rev, double deltat) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * ny * nz) return;
    const int s = t % 3, r = t / 3;
    if (s == 1 && nx < 3) return;   // i = nx is then column 2 again
    const int i = s == 0 ? 2 : nx + s - 1;
    const int j = r % ny + 2, k = r / ny + 2;
    const int64_t sx = nx + 2, sxy = sx * (ny + 2);
    const GlobalFields A{u, v, w, (i - 1) + sx * ((j - 1) + (int64_t)(ny + 2) * (k - 1)), sx, sxy};
    momentum_cell<FAST>(A, i, j, k, p, cell, ut, vt, wt, ru, rv, rw, T, M, nx, ny, nz, rev, deltat);
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

// The momentum formula's index-only sub-expressions per axis, evaluated in
// the formula's own operations (host IEEE double, no contraction), and the
// reciprocals of the denominators that divide field values.  The fast
// division is enabled only if every such denominator lies in [2^-64, 2^64].
static int build_mom_tables(Ctx &c) {
    const std::vector<double> *dl[3] = {&c.hdx, &c.hdy, &c.hdz};
    const int64_t nn[3] = {c.nx, c.ny, c.nz};
    bool fast = true;
    auto in_range = [](double d) {
        const double m = std::fabs(d);
        return m >= 0x1p-64 && m <= 0x1p64;
    };
    for (int a = 0; a < 3; a++) {
        const int64_t n = nn[a], tn = n + 3, ftn = c.tn[a];
        const std::vector<double> &d = *dl[a], &face = c.htab[a], &cent = c.htab[3 + a];
        std::vector<double> h(M_NQ * tn, 0.0);
        auto put = [&](int q, int64_t i, double x) { h[q * tn + i] = x; };
        auto upwind = [&](int q, int64_t i, double den, double r) {
            put(q, i, -1.0 / (r * (r + 1.0)));
            put(q + 1, i, 1.0 - 1.0 / r);
            put(q + 2, i, r / (r + 1.0));
            put(q + 3, i, 2.0 / (r * (r + 1.0)));
            put(q + 4, i, 2.0 / r);
            put(q + 5, i, 2.0 / (r + 1.0));
            put(q + 6, i, 1.0 / (2.0 * den));
        };
        for (int64_t i = 2; i <= n + 1; i++) {
            const double d0 = d[i], d1 = d[i + 1], dm = d[i - 1];
            const double h1 = 0.5 * (d0 + d1), h0 = 0.5 * (d0 + dm);
            put(M_AS, i, d1 / (d0 + d1));
            put(M_CS, i, 2.0 / (d0 + d1));
            put(M_D1, i, d1);
            put(M_RD1, i, 1.0 / d1);
            put(M_D0, i, d0);
            put(M_RD0, i, 1.0 / d0);
            upwind(M_US, i, d1, d0 / d1);
            put(M_CC, i, 2.0 / (0.5 * ((d0 + d1) + (d0 + dm))));
            put(M_H1, i, h1);
            put(M_RH1, i, 1.0 / h1);
            put(M_H0, i, h0);
            put(M_RH0, i, 1.0 / h0);
            upwind(M_UC, i, h1, h0 / h1);
            fast = fast && in_range(d0) && in_range(d1) && in_range(h0) && in_range(h1);
            if (i >= 3 && i <= n) {   // where the Kawamura-Kuwahara stencil can apply
                const double ks1 = face[5 * ftn + i] * d1, ks2 = (2.0 * face[11 * ftn + i]) * d0;
                const double kc1 = cent[5 * ftn + i] * h1, kc2 = (2.0 * cent[11 * ftn + i]) * h0;
                put(M_KS1, i, ks1);
                put(M_RKS1, i, 1.0 / ks1);
                put(M_KS2, i, ks2);
                put(M_RKS2, i, 1.0 / ks2);
                put(M_KC1, i, kc1);
                put(M_RKC1, i, 1.0 / kc1);
                put(M_KC2, i, kc2);
                put(M_RKC2, i, 1.0 / kc2);
                // the staggered component uses the face table for i <= n-1 only
                fast = fast && in_range(kc1) && in_range(kc2) && (i == n || (in_range(ks1) && in_range(ks2)));
            }
        }
        // y and z: the same quantities and this axis's two convection tables, per index
        std::vector<double> r(a ? kMomRec * tn : 0, 0.0);
        for (int64_t i = 0; a && i < tn; i++) {
            for (int q = 0; q < M_NQ; q++) r[i * kMomRec + q] = h[q * tn + i];
            for (int q = 0; q < 12 && i < ftn; q++) {
                r[i * kMomRec + kRecS + q] = face[q * ftn + i];
                r[i * kMomRec + kRecC + q] = cent[q * ftn + i];
            }
        }
        if (c.mtn[a] != tn) {
            if (c.mtab[a]) IB3D_CHECK(hipFree(c.mtab[a]));
            if (c.mrec[a]) IB3D_CHECK(hipFree(c.mrec[a]));
            c.mtab[a] = c.mrec[a] = nullptr;
            c.mtn[a] = 0;
            IB3D_CHECK(hipMalloc((void **)&c.mtab[a], h.size() * sizeof(double)));
            if (a) IB3D_CHECK(hipMalloc((void **)&c.mrec[a], r.size() * sizeof(double)));
            c.mtn[a] = tn;
        }
        IB3D_CHECK(h2d(c.mtab[a], h.data(), h.size() * sizeof(double)));
        if (a) IB3D_CHECK(h2d(c.mrec[a], r.data(), r.size() * sizeof(double)));
    }
    c.mom_fast = fast;
    c.mom_ready = true;
    return 0;
}

// the interior predictor from LDS tiles (default) or straight from HBM
static int g_mom_lds = 0;
int set_momentum_tiles(int on) {
    IB3D_REQUIRE(on >= 0 && on <= 4, "set_momentum_tiles: variant must be 0..4");
    g_mom_lds = on;
    return 0;
}

int ns_momentum(Ctx &c, double rev, double deltat) {
    IB3D_REQUIRE(c.coeff_ready, "nsMomentum: non_uni_coeff has not run");
    IB3D_REQUIRE(c.lists_ready, "nsMomentum: cellCount has not run");
    if (!c.mom_ready && build_mom_tables(c)) return -1;
    MomTabs M;
    for (int a = 0; a < 3; a++) {
        M.m[a] = c.mtab[a];
        M.mn[a] = (int)c.mtn[a];
        M.rec[a] = c.mrec[a];
    }
    const Tabs T = tabs(c);
    const int nx = (int)c.nx, ny = (int)c.ny, nz = (int)c.nz;
    const dim3 blk(64, 4, 1);
    const int nxi = std::max(0, nx - 3);   // interior columns 3..nx-1
    const int nedge = 3 * ny * nz;
    // 32-bit byte offsets need the FP64 fields to span less than 4 GiB
    const bool off32 = c.P * (int64_t)sizeof(double) < ((int64_t)1 << 32);
    auto launch = [&](auto fast) {
        constexpr bool F = decltype(fast)::value;
        if (nxi > 0 && g_mom_lds == 1)
            hipLaunchKernelGGL((k_momentum_lds<F, 8, 4>), dim3(nblocks(nxi, kMomTX), nblocks(ny, 8), nblocks(nz, kMomTZ)),
                               dim3(kMomTX, 8, 1), 0, c.stream, c.u, c.v, c.w, c.p, c.tag, c.ut, c.vt, c.wt,
                               c.resi_u, c.resi_v, c.resi_w, T, M, nx, ny, nz, rev, deltat);
        else if (nxi > 0 && g_mom_lds == 2)
            hipLaunchKernelGGL((k_momentum_lds<F, 4, 3>), dim3(nblocks(nxi, kMomTX), nblocks(ny, 4), nblocks(nz, kMomTZ)),
                               dim3(kMomTX, 4, 1), 0, c.stream, c.u, c.v, c.w, c.p, c.tag, c.ut, c.vt, c.wt,
                               c.resi_u, c.resi_v, c.resi_w, T, M, nx, ny, nz, rev, deltat);
        else if (nxi > 0 && g_mom_lds == 3)
            hipLaunchKernelGGL((k_momentum_lds<F, 8, 3>), dim3(nblocks(nxi, kMomTX), nblocks(ny, 8), nblocks(nz, kMomTZ)),
                               dim3(kMomTX, 8, 1), 0, c.stream, c.u, c.v, c.w, c.p, c.tag, c.ut, c.vt, c.wt,
                               c.resi_u, c.resi_v, c.resi_w, T, M, nx, ny, nz, rev, deltat);
        else if (nxi > 0 && g_mom_lds == 0 && off32)
            hipLaunchKernelGGL((k_momentum<F, true>), dim3(nblocks(nxi, 64), nblocks(ny, 4), (unsigned)nz), blk, 0,
                               c.stream, c.u, c.v, c.w, c.p, c.tag, c.ut, c.vt, c.wt, c.resi_u, c.resi_v, c.resi_w, T,
                               M, nx, ny, nz, rev, deltat);
        else if (nxi > 0)
            hipLaunchKernelGGL((k_momentum<F, false>), dim3(nblocks(nxi, 64), nblocks(ny, 4), (unsigned)nz), blk, 0,
                               c.stream, c.u, c.v, c.w, c.p, c.tag, c.ut, c.vt, c.wt, c.resi_u, c.resi_v, c.resi_w, T,
                               M, nx, ny, nz, rev, deltat);
        hipLaunchKernelGGL(k_momentum_edge<F>, dim3(nblocks(nedge, 256)), dim3(256), 0, c.stream, c.u, c.v, c.w,
                           c.p, c.tag, c.ut, c.vt, c.wt, c.resi_u, c.resi_v, c.resi_w, T, M, nx, ny, nz, rev, deltat);
    };
    if (c.mom_fast)
        launch(std::true_type{});
    else
        launch(std::false_type{});
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

// derr_std == nullptr: the result is left to arrive in c.hscal[3] at the
// caller's next synchronisation (one fewer host round trip per step)
int steady_copy(Ctx &c, double dt, double *derr_std) {
    const unsigned nb = nblocks(c.P, 256);
    if (ensure_partial(nb)) return -1;
    hipLaunchKernelGGL(k_steady_copy, dim3(nb), dim3(256), 0, c.stream, c.u, c.v, c.w, c.ut, c.vt,
                       c.wt, c.tag, (int)c.nx, (int)c.ny, (int)c.nz, dt, c.partial);
    launch_max_partials(c.partial, (int64_t)nb, c.scalars + 3);
    IB3D_CHECK(hipGetLastError());
    IB3D_CHECK(hipMemcpyAsync(c.hscal + 3, c.scalars + 3, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (!derr_std) return 0;
    IB3D_CHECK(hipStreamSynchronize(c.stream));
    *derr_std = c.hscal[3];
    return 0;
}

}  // namespace ib3d
