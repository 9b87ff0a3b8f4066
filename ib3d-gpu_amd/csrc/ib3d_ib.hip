// ib3d_ib.hip -- ghost-cell search, list compaction, IB forcing and surface
// forces (subSearch.f90, subForcing.f90, subStressCal1.f90/subStressCal2.f90).
//
// Tags, lists and stencil indices are integer functions of FP64 comparisons;
// they are evaluated with the reference's expressions in the reference's
// order (correctly rounded sqrt/division, no contraction) and are therefore
// identical to the reference's.
//
// The reference's forcing loops run serially: a face shared by two
// intercepted cells takes the value of the later list entry, and an image
// stencil may read a face written earlier in the same loop.  Here the updates
// are regrouped into dependency levels once per geometry: every update of a
// level reads either the field as it was when the loop started or the
// result of an update of an earlier level, and the last writer of each face
// stores its value at the end.  That reproduces the serial result exactly.
#include "ib3d_internal.h"

#include <algorithm>
#include <cstring>
#include <cmath>
#include <rocprim/rocprim.hpp>
#include <unordered_map>

namespace ib3d {

__device__ __forceinline__ int64_t lin(int i, int j, int k, int sx, int sy) {
    return (int64_t)(i - 1) + (int64_t)sx * ((int64_t)(j - 1) + (int64_t)sy * (k - 1));
}

// Passes x through an empty asm so that work on it stays inside the branch
// that needs it: in the retagging's nearest-element loop the update branch
// is rarely taken, and a square root hoisted out of it costs ~20 FP64
// instructions on every element.
__device__ __forceinline__ double unhoisted(double x) {
    asm volatile("" : "+v"(x));
    return x;
}

// subSearch.f90:28-35
__global__ void k_shift_nodes(const double *x, const double *y, const double *z, double *x1, double *y1,
                              double *z1, int64_t n, double xs, double ys, double zs) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    x1[q] = x[q] + xs;
    y1[q] = y[q] + ys;
    z1[q] = z[q] + zs;
}

// prescribed rigid displacement of the initially placed body (replaces the
// free motion of subSearch.f90:41-76): x1 = (x + shift) + d
__global__ void k_shift_nodes_moving(const double *x, const double *y, const double *z, double *x1, double *y1,
                                     double *z1, int64_t n, double xs, double ys, double zs, double dx, double dy,
                                     double dz) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    x1[q] = (x[q] + xs) + dx;
    y1[q] = (y[q] + ys) + dy;
    z1[q] = (z[q] + zs) + dz;
}

// subSearch.f90:95-125 (node ids are 1-based)
__global__ void k_surface_norm(const double *xn, const double *yn, const double *zn, const int *elp, int64_t ne,
                               double inor, double *xc, double *yc, double *zc, double *ca, double *cb, double *cg,
                               double *a3, double *b3, double *g3) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= ne) return;
    const int e1 = elp[n] - 1, e2 = elp[ne + n] - 1, e3 = elp[2 * ne + n] - 1;
    const double p1x = xn[e1], p1y = yn[e1], p1z = zn[e1];
    const double p2x = xn[e2], p2y = yn[e2], p2z = zn[e2];
    const double p3x = xn[e3], p3y = yn[e3], p3z = zn[e3];
    xc[n] = ((p2x + p1x) + p3x) / 3.0;
    yc[n] = ((p2y + p1y) + p3y) / 3.0;
    zc[n] = ((p2z + p1z) + p3z) / 3.0;
    const double a = (p2y - p1y) * (p3z - p1z) - (p3y - p1y) * (p2z - p1z);
    const double b = (p2z - p1z) * (p3x - p1x) - (p3z - p1z) * (p2x - p1x);
    const double g = (p2x - p1x) * (p3y - p1y) - (p3x - p1x) * (p2y - p1y);
    a3[n] = a;
    b3[n] = b;
    g3[n] = g;
    const double len = sqrt(((a * a) + (b * b)) + (g * g));
    ca[n] = (a / len) * inor;
    cb[n] = (b / len) * inor;
    cg[n] = (g / len) * inor;
}

// subSearch.f90:166-247: nearest centroid (first index wins ties of the
// sqrt distance) and the 8-corner plane test.  sqrt is only evaluated when
// the squared distance is below the current minimum's: sqrt is monotone, so
// a larger square can never give a strictly smaller distance.
static constexpr int kTagTile = 256;
__global__ void __launch_bounds__(kTagTile)
k_tagging(int i0, int i1, int j0, int j1, int k0, int k1, int sx, int sy, const double *__restrict__ xp,
          const double *__restrict__ yp, const double *__restrict__ zp, const double *__restrict__ x1,
          const double *__restrict__ y1, const double *__restrict__ z1, const double *__restrict__ xc,
          const double *__restrict__ yc, const double *__restrict__ zc, const double *__restrict__ ca,
          const double *__restrict__ cb, const double *__restrict__ cg, int ne, int8_t *tag, int *minElem,
          int *err) {
    __shared__ double sx_[kTagTile], sy_[kTagTile], sz_[kTagTile];
    const int wi = i1 - i0 + 1, wj = j1 - j0 + 1, wk = k1 - k0 + 1;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = q < (int64_t)wi * wj * wk;
    int i = 0, j = 0, k = 0;
    if (active) {
        i = i0 + (int)(q % wi);
        const int64_t r = q / wi;
        j = j0 + (int)(r % wj);
        k = k0 + (int)(r / wj);
    }
    const double n1x = active ? xp[i] : 0.0, n1y = active ? yp[j] : 0.0, n1z = active ? zp[k] : 0.0;
    double minDis = 1e14, smin = __builtin_huge_val();
    int nel = 0;
    for (int base = 0; base < ne; base += kTagTile) {
        const int m = base + threadIdx.x;
        __syncthreads();
        if (m < ne) {
            sx_[threadIdx.x] = xc[m];
            sy_[threadIdx.x] = yc[m];
            sz_[threadIdx.x] = zc[m];
        }
        __syncthreads();
        const int cnt = min(kTagTile, ne - base);
        for (int t = 0; t < cnt; t++) {
            const double dy = n1y - sy_[t], dx = n1x - sx_[t], dz = n1z - sz_[t];
            const double s = ((dy * dy) + (dx * dx)) + (dz * dz);
            if (s < smin) {
                const double dis = sqrt(s);
                if (dis < minDis) {
                    minDis = dis;
                    smin = s;
                    nel = base + t + 1;
                }
            }
        }
    }
    if (!active) return;
    if (nel < 1) {
        atomicOr(err, 1);
        return;
    }
    const int e = nel - 1;
    const double cx = xc[e], cy = yc[e], cz = zc[e], a = ca[e], b = cb[e], g = cg[e];
    const double n2x = x1[i], n3x = x1[i + 1], n2y = y1[j], n3y = y1[j + 1], n2z = z1[k], n3z = z1[k + 1];
    double d[8];
    d[0] = ((n2x - cx) * a + (n2y - cy) * b) + (n2z - cz) * g;
    d[1] = ((n2x - cx) * a + (n3y - cy) * b) + (n2z - cz) * g;
    d[2] = ((n3x - cx) * a + (n2y - cy) * b) + (n2z - cz) * g;
    d[3] = ((n3x - cx) * a + (n3y - cy) * b) + (n2z - cz) * g;
    d[4] = ((n2x - cx) * a + (n2y - cy) * b) + (n3z - cz) * g;
    d[5] = ((n2x - cx) * a + (n3y - cy) * b) + (n3z - cz) * g;
    d[6] = ((n3x - cx) * a + (n2y - cy) * b) + (n3z - cz) * g;
    d[7] = ((n3x - cx) * a + (n3y - cy) * b) + (n3z - cz) * g;
    bool all_in = true, all_out = true;
#pragma unroll
    for (int t = 0; t < 8; t++) {
        all_in = all_in && (d[t] <= -1e-16);
        all_out = all_out && (d[t] > -1e-16);
    }
    const int64_t c = lin(i, j, k, sx, sy);
    tag[c] = all_in ? 1 : (all_out ? 0 : 2);
    minElem[c] = nel;
}

// subSearch.f90:360-363: queue the 6x6x6 block (i1-2..i1+3, ...) around every
// intercepted cell; each cell enters the list once (the reference tags
// overlapping blocks repeatedly, always with the same result).  Cells
// outside the padded index range 1..n+2 are not queued.
__global__ void __launch_bounds__(1024)
k_retag_mark(const int *ib, int nb, int nx, int ny, int nz, uint8_t *flag, int *list, int *count) {
    // one atomicAdd on the list counter per block: a cell's slot = block base
    // + cells newly queued by lower waves + set ballot bits below its lane
    __shared__ int s_wave[16];
    __shared__ int s_base;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    bool add = false;
    int64_t c = 0;
    if (q < (int64_t)nb * 216) {
        const int n = (int)(q / 216), o = (int)(q % 216);
        const int i = ib[n] - 2 + o % 6, j = ib[nb + n] - 2 + (o / 6) % 6, k = ib[2 * nb + n] - 2 + o / 36;
        if (i >= 1 && i <= nx + 2 && j >= 1 && j <= ny + 2 && k >= 1 && k <= nz + 2) {
            c = lin(i, j, k, nx + 2, ny + 2);
            unsigned int *word = reinterpret_cast<unsigned int *>(flag + (c & ~(int64_t)3));
            const unsigned int bit = 1u << (8 * (c & 3));
            // overlapping blocks queue the same cells many times: a relaxed
            // load that already sees the bit skips the atomic (only the
            // atomic that sets the bit queues the cell)
            if (!(__hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bit))
                add = !(atomicOr(word, bit) & bit);
        }
    }
    const unsigned long long bal = __ballot(add);
    if (lane == 0) s_wave[wv] = __popcll(bal);
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int t = 0; t < (int)(blockDim.x >> 6); t++) {
            const int v = s_wave[t];
            s_wave[t] = run;
            run += v;
        }
        s_base = run > 0 ? atomicAdd(count, run) : 0;
    }
    __syncthreads();
    if (add) list[s_base + s_wave[wv] + __popcll(bal & ((1ull << lane) - 1))] = (int)c;
}

// One step of the first-nearest-element scan: sqrt only when the square is
// below the current minimum's (sqrt is monotone, so a larger square can
// never give a strictly smaller distance); ties keep the earlier element.
__device__ __forceinline__ void nearest_update(double s, int elem1, double &smin, double &minDis, int &nel) {
    if (s < smin) {
        const double dis = sqrt(unhoisted(s));
        if (dis < minDis) {
            minDis = dis;
            smin = s;
            nel = elem1;
        }
    }
}

// Tag of one queued cell from its nearest element nel (1-based, 0 = none).
__device__ __forceinline__ void retag_cell(int64_t c, int i, int j, int k, int nel, const double *__restrict__ x1,
                                           const double *__restrict__ y1, const double *__restrict__ z1,
                                           const double *__restrict__ xc, const double *__restrict__ yc,
                                           const double *__restrict__ zc, const double *__restrict__ ca,
                                           const double *__restrict__ cb, const double *__restrict__ cg, int8_t *tag,
                                           int *minElem, uint8_t *flag, double *u, double *v, double *w, double wu,
                                           double wv, double ww, int *err) {
    flag[c] = 0;
    if (nel < 1) {
        atomicOr(err, 1);
        return;
    }
    const int e = nel - 1;
    const double cx = xc[e], cy = yc[e], cz = zc[e], a = ca[e], b = cb[e], g = cg[e];
    const double n2x = x1[i], n3x = x1[i + 1], n2y = y1[j], n3y = y1[j + 1], n2z = z1[k], n3z = z1[k + 1];
    double d[8];
    d[0] = ((n2x - cx) * a + (n2y - cy) * b) + (n2z - cz) * g;
    d[1] = ((n2x - cx) * a + (n3y - cy) * b) + (n2z - cz) * g;
    d[2] = ((n3x - cx) * a + (n2y - cy) * b) + (n2z - cz) * g;
    d[3] = ((n3x - cx) * a + (n3y - cy) * b) + (n2z - cz) * g;
    d[4] = ((n2x - cx) * a + (n2y - cy) * b) + (n3z - cz) * g;
    d[5] = ((n2x - cx) * a + (n3y - cy) * b) + (n3z - cz) * g;
    d[6] = ((n3x - cx) * a + (n2y - cy) * b) + (n3z - cz) * g;
    d[7] = ((n3x - cx) * a + (n3y - cy) * b) + (n3z - cz) * g;
    bool all_in = true, all_out = true;
#pragma unroll
    for (int t = 0; t < 8; t++) {
        all_in = all_in && (d[t] <= -1e-16);
        all_out = all_out && (d[t] > -1e-16);
    }
    const int8_t nt = all_in ? 1 : (all_out ? 0 : 2);
    if (tag[c] == 1 && nt != 1) {
        u[c] = wu;
        v[c] = wv;
        w[c] = ww;
    }
    tag[c] = nt;
    minElem[c] = nel;
}

// subSearch.f90:364-439 for the queued cells: nearest centroid and the
// 8-corner test, the same arithmetic as k_tagging.  A cell that was solid
// and is no longer takes the wall velocity (the reference's body never
// moves, so it has no rule for such cells).  Clears its queue flag.
// A block takes 64 queued cells at a time; its 16 waves each scan one
// contiguous slice of the elements (uniform loads), and wave 0 merges the
// slices in element order with a strict '<', which gives the same first
// nearest element as one pass over all of them.
static constexpr int kRtSlices = 16;
__global__ void __launch_bounds__(64 * kRtSlices)
k_retag_list(const int *list, const int *count, int sx, int sy, const double *__restrict__ xp,
             const double *__restrict__ yp, const double *__restrict__ zp, const double *__restrict__ x1,
             const double *__restrict__ y1, const double *__restrict__ z1, const double *__restrict__ xc,
             const double *__restrict__ yc, const double *__restrict__ zc, const double *__restrict__ ca,
             const double *__restrict__ cb, const double *__restrict__ cg, int ne, int8_t *tag, int *minElem,
             uint8_t *flag, double *u, double *v, double *w, double wu, double wv, double ww, int *err) {
    __shared__ double s_dis[kRtSlices][64];
    __shared__ int s_nel[kRtSlices][64];
    const int nq = *count;
    const int lane = threadIdx.x & 63;
    const int sl = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int chunk = (ne + kRtSlices - 1) / kRtSlices;
    const int m0 = min(ne, sl * chunk), m1 = min(ne, m0 + chunk);
    for (int q0 = blockIdx.x * 64; q0 < nq; q0 += gridDim.x * 64) {
        const int q = q0 + lane;
        const bool active = q < nq;
        int i = 0, j = 0, k = 0;
        int64_t c = 0;
        if (active) {
            c = list[q];
            i = (int)(c % sx) + 1;
            const int64_t r = c / sx;
            j = (int)(r % sy) + 1;
            k = (int)(r / sy) + 1;
        }
        const double n1x = active ? xp[i] : 0.0, n1y = active ? yp[j] : 0.0, n1z = active ? zp[k] : 0.0;
        double minDis = 1e14, smin = __builtin_huge_val();
        int nel = 0;
        // four squared distances per step (their loads batch), then the
        // updates in element order
        int m = m0;
        for (; m + 4 <= m1; m += 4) {
            double s4[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const double dy = n1y - yc[m + t], dx = n1x - xc[m + t], dz = n1z - zc[m + t];
                s4[t] = ((dy * dy) + (dx * dx)) + (dz * dz);
            }
#pragma unroll
            for (int t = 0; t < 4; t++) nearest_update(s4[t], m + t + 1, smin, minDis, nel);
        }
        for (; m < m1; m++) {
            const double dy = n1y - yc[m], dx = n1x - xc[m], dz = n1z - zc[m];
            nearest_update(((dy * dy) + (dx * dx)) + (dz * dz), m + 1, smin, minDis, nel);
        }
        s_dis[sl][lane] = minDis;
        s_nel[sl][lane] = nel;
        __syncthreads();
        if (sl == 0 && active) {
            for (int t = 1; t < kRtSlices; t++) {
                const double d = s_dis[t][lane];
                if (d < minDis) {
                    minDis = d;
                    nel = s_nel[t][lane];
                }
            }
            retag_cell(c, i, j, k, nel, x1, y1, z1, xc, yc, zc, ca, cb, cg, tag, minElem, flag, u, v, w, wu, wv,
                       ww, err);
        }
        __syncthreads();  // s_dis and s_nel are rewritten by the next group
    }
}

// per interior row (j,k): counts of fluid, solid, intercepted, red-fluid
// Per-row counts of fluid (tag 0), solid (1), intercepted (2) and red
// fluid cells; one wave per row, 64 consecutive cells per ballot.
__global__ void __launch_bounds__(256) k_row_counts(const int8_t *tag, int nx, int ny, int nz, int *counts) {
    const int lane = threadIdx.x & 63;
    const int r = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    const int nrows = ny * nz;
    if (r >= nrows) return;  // uniform per wave
    const int j = 2 + r % ny, k = 2 + r / ny;
    const int sx = nx + 2, sy = ny + 2;
    const int8_t *row = tag + lin(1, j, k, sx, sy);
    int f = 0, s = 0, b = 0, red = 0;
    for (int i0 = 2; i0 <= nx + 1; i0 += 64) {
        const int i = i0 + lane;
        const int t = i <= nx + 1 ? (int)row[i - 1] : -1;
        f += __popcll(__ballot(t == 0));
        s += __popcll(__ballot(t == 1));
        b += __popcll(__ballot(t == 2));
        red += __popcll(__ballot(t == 0 && ((i + j + k) & 1)));
    }
    if (lane == 0) {
        counts[r] = f;
        counts[nrows + r] = s;
        counts[2 * nrows + r] = b;
        counts[3 * nrows + r] = red;
    }
}

// Exclusive scan, in place, of the 4 per-row count arrays in tiles of
// kScanTile rows: k_tile_sums sums every tile, then k_scan_counts scans each
// tile from its offset (the sum of the tiles before it).  Grids are
// (tiles, 4); 1024 threads take 8 consecutive rows each.  totals[a] = sum
// of array a, written by the last tile's block.
static constexpr int kScanTile = 8192;
__device__ __forceinline__ void load_rows8(const int *cnt, int lo, int nrows, int v[8], int &s) {
    s = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        v[q] = lo + q < nrows ? cnt[lo + q] : 0;
        s += v[q];
    }
}

__global__ void __launch_bounds__(1024) k_tile_sums(const int *counts, int nrows, int *tsum) {
    __shared__ int wsum[16];
    const int *cnt = counts + (int64_t)blockIdx.y * nrows;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int v[8], s;
    load_rows8(cnt, (int)blockIdx.x * kScanTile + (int)threadIdx.x * 8, nrows, v, s);
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) s += __shfl_xor(s, d, 64);
    if (lane == 0) wsum[wv] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < 16; w++) t += wsum[w];
        tsum[blockIdx.y * gridDim.x + blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(1024) k_scan_counts(int *counts, int nrows, const int *tsum, int *totals) {
    __shared__ int wsum[16];
    __shared__ int base;
    int *cnt = counts + (int64_t)blockIdx.y * nrows;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (threadIdx.x == 0) {
        int o = 0;
        for (int t = 0; t < (int)blockIdx.x; t++) o += tsum[blockIdx.y * gridDim.x + t];
        base = o;
    }
    const int lo = (int)blockIdx.x * kScanTile + (int)threadIdx.x * 8;
    int v[8], s;
    load_rows8(cnt, lo, nrows, v, s);
    int x = s;  // inclusive scan over the wave
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    if (wv == 0) {
        const int w = lane < 16 ? wsum[lane] : 0;
        int z = w;
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) {
            const int y = __shfl_up(z, d, 64);
            if (lane >= d) z += y;
        }
        if (lane < 16) wsum[lane] = z - w;
        if (lane == 15 && blockIdx.x == gridDim.x - 1) totals[blockIdx.y] = base + z;
    }
    __syncthreads();
    int run = base + wsum[wv] + x - s;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        if (lo + q < nrows) cnt[lo + q] = run;
        run += v[q];
    }
}

// order-preserving scatter into (count,3) lists, (k,j,i) order; one wave
// per row, a cell's slot = row offset + set ballot bits below its lane
__global__ void __launch_bounds__(256) k_row_scatter(const int8_t *tag, int nx, int ny, int nz, const int *offs,
                                                     int *fl, int nf, int *sl, int ns, int *bl, int nb) {
    const int lane = threadIdx.x & 63;
    const int r = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    const int nrows = ny * nz;
    if (r >= nrows) return;  // uniform per wave
    const int j = 2 + r % ny, k = 2 + r / ny;
    const int sx = nx + 2, sy = ny + 2;
    const int8_t *row = tag + lin(1, j, k, sx, sy);
    const unsigned long long below = (1ull << lane) - 1ull;
    int f = offs[r], s = offs[nrows + r], b = offs[2 * nrows + r];
    for (int i0 = 2; i0 <= nx + 1; i0 += 64) {
        const int i = i0 + lane;
        const int t = i <= nx + 1 ? (int)row[i - 1] : -1;
        const unsigned long long mf = __ballot(t == 0), ms = __ballot(t == 1), mb = __ballot(t == 2);
        if (t == 0) {
            const int q = f + __popcll(mf & below);
            fl[q] = i; fl[nf + q] = j; fl[2 * nf + q] = k;
        } else if (t == 1) {
            const int q = s + __popcll(ms & below);
            sl[q] = i; sl[ns + q] = j; sl[2 * ns + q] = k;
        } else if (t == 2) {
            const int q = b + __popcll(mb & below);
            bl[q] = i; bl[nb + q] = j; bl[2 * nb + q] = k;
        }
        f += __popcll(mf);
        s += __popcll(ms);
        b += __popcll(mb);
    }
}

// subSearch.f90:621-672
__global__ void k_norm_distance(const int *ib, int nb, const int *minElem, int sx, int sy, const double *xp,
                                const double *yp, const double *zp, const double *x1, const double *y1,
                                const double *z1, const double *xc, const double *yc, const double *zc,
                                const double *ca, const double *cb, const double *cg, int *nel, double *nd, int *err) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nb) return;
    const int i = ib[n], j = ib[nb + n], k = ib[2 * nb + n];
    const int e1 = minElem[lin(i, j, k, sx, sy)];
    if (e1 < 1) {
        atomicOr(err, 2);
        return;
    }
    nel[n] = e1;
    const int e = e1 - 1;
    const double n1x = xp[i], n2x = x1[i], n3x = x1[i + 1];
    const double n1y = yp[j], n2y = y1[j], n3y = y1[j + 1];
    const double n1z = zp[k], n2z = z1[k], n3z = z1[k + 1];
    const double cx = xc[e], cy = yc[e], cz = zc[e], a = ca[e], b = cb[e], g = cg[e];
    nd[n] = ((n1x - cx) * a + (n1y - cy) * b) + (n1z - cz) * g;            // p
    nd[nb + n] = ((n2x - cx) * a + (n1y - cy) * b) + (n1z - cz) * g;       // u1
    nd[2 * nb + n] = ((n3x - cx) * a + (n1y - cy) * b) + (n1z - cz) * g;   // u2
    nd[3 * nb + n] = ((n1x - cx) * a + (n2y - cy) * b) + (n1z - cz) * g;   // v1
    nd[4 * nb + n] = ((n1x - cx) * a + (n3y - cy) * b) + (n1z - cz) * g;   // v2
    nd[5 * nb + n] = ((n1x - cx) * a + (n1y - cy) * b) + (n2z - cz) * g;   // w1
    nd[6 * nb + n] = ((n1x - cx) * a + (n1y - cy) * b) + (n3z - cz) * g;   // w2
}

// ------------------------------------------------------------------------
// image point stencils
// ------------------------------------------------------------------------
// largest l in [lo,hi] with c[l] <= pos and pos < c[l+1]; -1 if none.  For
// strictly increasing c this equals the reference's scan that keeps the
// last matching line (subForcing.f90:156-166).
__device__ __forceinline__ int bracket(const double *c, int lo, int hi, double pos) {
    if (!(pos >= c[lo])) return -1;
    int a = lo, b = hi;
    while (a < b) {
        const int m = (a + b + 1) >> 1;
        if (c[m] <= pos) a = m;
        else b = m - 1;
    }
    return (pos < c[a + 1]) ? a : -1;
}

struct Coords {
    const double *x[4], *y[4], *z[4];  // per component u, v, w, p
};

__device__ __forceinline__ double lerp_ref(double a, double b, double r, double d) { return a + ((b - a) * r) / d; }
__device__ __forceinline__ double d3(double h1, double h2, double f1, double f2, double fp) {
    return (((h1 * h1) * f2 - (h2 * h2) * f1) + ((h2 * h2) - (h1 * h1)) * fp) / (((h1 * h2) * (h1 + h2)) + 1e-16);
}

// trilinear value and normal derivative at the image point from the 8
// stencil values f[a + 2b + 4c] (subForcing.f90:175-217)
__device__ void image_values(const double f[8], const double *cx, const double *cy, const double *cz, int ix,
                             int iy, int iz, double px, double py, double pz, double na, double nb, double ng,
                             double *fpos, double *dfdn) {
    const double f000 = f[0], f100 = f[1], f010 = f[2], f110 = f[3];
    const double f001 = f[4], f101 = f[5], f011 = f[6], f111 = f[7];
    const double rx = px - cx[ix], gx = cx[ix + 1] - cx[ix];
    const double ry = py - cy[iy], gy = cy[iy + 1] - cy[iy];
    const double rz = pz - cz[iz], gz = cz[iz + 1] - cz[iz];
    const double x1z1 = lerp_ref(f000, f100, rx, gx), x2z1 = lerp_ref(f010, f110, rx, gx);
    const double x1z2 = lerp_ref(f001, f101, rx, gx), x2z2 = lerp_ref(f011, f111, rx, gx);
    const double vz1 = lerp_ref(x1z1, x2z1, ry, gy), vz2 = lerp_ref(x1z2, x2z2, ry, gy);
    const double vy1 = lerp_ref(x1z1, x1z2, rz, gz), vy2 = lerp_ref(x2z1, x2z2, rz, gz);
    const double z1x1 = lerp_ref(f000, f001, rz, gz), z2x1 = lerp_ref(f010, f011, rz, gz);
    const double z1x2 = lerp_ref(f100, f101, rz, gz), z2x2 = lerp_ref(f110, f111, rz, gz);
    const double vx1 = lerp_ref(z1x1, z2x1, ry, gy), vx2 = lerp_ref(z1x2, z2x2, ry, gy);
    const double vp = lerp_ref(vx1, vx2, rx, gx);
    double h2 = fabs(cx[ix + 1] - px), h1 = fabs(cx[ix] - px);
    const double ddx = d3(h1, h2, vx1, vx2, vp);
    h2 = fabs(cy[iy + 1] - py);
    h1 = fabs(cy[iy] - py);
    const double ddy = d3(h1, h2, vy1, vy2, vp);
    h2 = fabs(cz[iz + 1] - pz);
    h1 = fabs(cz[iz] - pz);
    const double ddz = d3(h1, h2, vz1, vz2, vp);
    *fpos = vp;
    *dfdn = (ddx * na + ddy * nb) + ddz * ng;
}

// update kinds, in the reference's per-cell order: u2 u1 v2 v1 w2 w1, then p
__constant__ int c_comp[7] = {0, 0, 1, 1, 2, 2, 3};

struct StencilArgs {
    const int *ib;
    int nb;
    const int *nel;
    const double *nd;  // [7][nb]: p u1 u2 v1 v2 w1 w2
    const double *dx, *dy, *dz;
    const double *ca, *cb, *cg;
    Coords C;
    int nx, ny, nz;
};

// One thread per (cell, kind): image point, bracketing grid lines with the
// reference's search ranges and clamps (subForcing.f90:148-173 and the five
// sibling blocks; pressure: 28-46).  Output: idx[3], pos[3], pt1, sur, nel,
// write location.
__global__ void k_force_stencil(StencilArgs A, int *oi, double *od, int *err) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 7 * A.nb) return;
    const int n = q / 7, kind = q % 7;
    const int i = A.ib[n], j = A.ib[A.nb + n], k = A.ib[2 * A.nb + n];
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    // normal-distance slot of each kind and the face node
    const int ndslot[7] = {2, 1, 4, 3, 6, 5, 0};
    const double sur = A.nd[ndslot[kind] * A.nb + n];
    const int e = A.nel[n] - 1;
    const double na = A.ca[e], nb_ = A.cb[e], ng = A.cg[e];
    const int comp = c_comp[kind];
    const double *cx = A.C.x[comp], *cy = A.C.y[comp], *cz = A.C.z[comp];
    const double pt1 = 1.5 * sqrt(((A.dx[i] * A.dx[i]) + (A.dy[j] * A.dy[j])) + (A.dz[k] * A.dz[k])) +
                       (fabs(sur) - sur) * 0.5;
    int fi = i, fj = j, fk = k;           // face node index on its own grid
    int hx = nx + 1, hy = ny + 1, hz = nz + 1;
    int clamp_axis = -1, from = 0, to = 0;
    int wi = i, wj = j, wk = k;           // written location
    switch (kind) {
        case 0: fi = i + 1; hx = nx + 2; clamp_axis = 0; from = nx + 2; to = nx + 1; break;          // u2
        case 1: fi = i; clamp_axis = 0; from = 1; to = 2; wi = i - 1; break;                          // u1
        case 2: fj = j + 1; hy = ny + 2; clamp_axis = 1; from = ny + 2; to = ny + 1; break;          // v2
        case 3: fj = j; clamp_axis = 1; from = 1; to = 2; wj = j - 1; break;                          // v1
        case 4: fk = k + 1; hz = nz + 2; clamp_axis = 2; from = nz + 2; to = nz + 1; break;          // w2
        case 5: fk = k; clamp_axis = 2; from = 1; to = 2; wk = k - 1; break;                          // w1
        default: break;                                                                               // p
    }
    const double px = cx[fi] + pt1 * na, py = cy[fj] + pt1 * nb_, pz = cz[fk] + pt1 * ng;
    int idx[3] = {bracket(cx, 1, hx, px), bracket(cy, 1, hy, py), bracket(cz, 1, hz, pz)};
    if (idx[0] < 0 || idx[1] < 0 || idx[2] < 0) atomicOr(err, 4);
    if (clamp_axis >= 0 && idx[clamp_axis] == from) idx[clamp_axis] = to;
    // every stencil corner must lie inside the padded grid (the reference
    // would read outside its arrays otherwise); such a geometry is rejected
    const int ax = idx[0] - (comp == 0), ay = idx[1] - (comp == 1), az = idx[2] - (comp == 2);
    if (ax < 1 || ax + 1 > nx + 2 || ay < 1 || ay + 1 > ny + 2 || az < 1 || az + 1 > nz + 2) {
        atomicOr(err, 16);
        idx[0] = idx[1] = idx[2] = 1;
    }
    int *o = oi + 8 * (int64_t)q;
    o[0] = comp;
    o[1] = (int)lin(wi, wj, wk, nx + 2, ny + 2);
    o[2] = idx[0];
    o[3] = idx[1];
    o[4] = idx[2];
    o[5] = e;
    o[6] = kind;
    o[7] = n;
    double *d = od + 5 * (int64_t)q;
    d[0] = px;
    d[1] = py;
    d[2] = pz;
    d[3] = pt1;
    d[4] = sur;
}

struct LevelArgs {
    const int *oi;       // [8] per update, level-sorted
    const double *od;    // [5] per update
    const int *srcs;     // [8] per update: -1 = field, else level-sorted update id
    double *res;
    double *F[4];        // ut, vt, wt, p
    const double *ca, *cb, *cg;
    Coords C;
    int sx, sy;
    double wall[3], acc[3];
};

// dp/dn at the wall from the body acceleration a: -(a . n); 0 for a body at
// rest (subForcing.f90:11 and the disabled moving-body form at line 20)
__device__ __forceinline__ double wall_dpdn(const double acc[3], double na, double nb, double ng) {
    if (acc[0] == 0.0 && acc[1] == 0.0 && acc[2] == 0.0) return 0.0;
    return -(((acc[0] * na) + (acc[1] * nb)) + (acc[2] * ng));
}

// subForcing.f90:219-224 (velocity; u_wall = 0 in the reference, the body's
// velocity when it moves) and 92-98 (pressure; dp/dn = 0, or the body's
// acceleration) for the updates [lo, hi) of one dependency level
__global__ void k_force_level(LevelArgs A, int lo, int hi) {
    const int q = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= hi) return;
    const int *o = A.oi + 8 * (int64_t)q;
    const double *d = A.od + 5 * (int64_t)q;
    const int comp = o[0], ix = o[2], iy = o[3], iz = o[4], e = o[5];
    const int ox = comp == 0, oy = comp == 1, oz = comp == 2;
    const int ax = ix - ox, ay = iy - oy, az = iz - oz;
    const double *Fd = A.F[comp];
    double f[8];
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const int s = A.srcs[8 * (int64_t)q + t];
        f[t] = s >= 0 ? A.res[s] : Fd[lin(ax + (t & 1), ay + ((t >> 1) & 1), az + (t >> 2), A.sx, A.sy)];
    }
    const double na = A.ca[e], nb = A.cb[e], ng = A.cg[e];
    double fpos, dn;
    image_values(f, A.C.x[comp], A.C.y[comp], A.C.z[comp], ix, iy, iz, d[0], d[1], d[2], na, nb, ng, &fpos, &dn);
    const double pt1 = d[3], sur = d[4];
    const double n1 = pt1 + sur;
    double val;
    if (comp < 3) {
        const double surf = A.wall[comp];
        const double cval = surf;
        const double bval = (2.0 / n1) * (fpos - surf) - dn;
        const double aval = dn / n1 - (fpos - surf) / (n1 * n1);
        val = ((aval * (sur * sur)) + bval * sur) + cval;
    } else {
        const double dpdn = wall_dpdn(A.acc, na, nb, ng);
        const double bval = dpdn;
        const double aval = (dn - dpdn) / (2.0 * n1);
        const double cval = fpos - ((dn + dpdn) * n1) * 0.5;
        val = ((aval * (sur * sur)) + bval * sur) + cval;
    }
    A.res[q] = val;
}

__global__ void k_force_final(const int *loc, const int *upd, const int *comp, int n, const double *res,
                              double *f0, double *f1, double *f2, double *f3) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double *F = comp[q] == 0 ? f0 : comp[q] == 1 ? f1 : comp[q] == 2 ? f2 : f3;
    F[loc[q]] = res[upd[q]];
}

// ------------------------------------------------------------------------
// surface forces (subStressCal1.f90:85-437, subStressCal2.f90:78-358)
// ------------------------------------------------------------------------
struct StressArgs {
    const double *xc, *yc, *zc, *ca, *cb, *cg, *a3, *b3, *g3;
    const double *x1, *y1, *z1;
    Coords C;
    const double *F[4];  // ut, vt, wt, p
    int ne, nx, ny, nz;
    double re;
    int *cache;          // [5][3][ne]
    int find;            // 1: locate the stencil cells (stressCal1)
    double *contrib;     // [7][ne]
    int *err;
    double wall[3], acc[3];
};

__global__ void k_stress(StressArgs A) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= A.ne) return;
    const int ne = A.ne, nx = A.nx, ny = A.ny, nz = A.nz;
    const int sx = nx + 2, sy = ny + 2;
    int *cell = A.cache;
    int ic, jc, kc;
    if (A.find) {
        ic = bracket(A.x1, 2, nx + 1, A.xc[e]);
        jc = bracket(A.y1, 2, ny + 1, A.yc[e]);
        kc = bracket(A.z1, 2, nz + 1, A.zc[e]);
        if (ic < 0 || jc < 0 || kc < 0) atomicOr(A.err, 8);
        cell[e] = ic; cell[ne + e] = jc; cell[2 * ne + e] = kc;
    } else {
        ic = cell[e]; jc = cell[ne + e]; kc = cell[2 * ne + e];
    }
    if (ic < 0 || jc < 0 || kc < 0) return;
    const double xs = A.xc[e], ys = A.yc[e], zs = A.zc[e];
    const double dX = A.x1[ic + 1] - A.x1[ic], dY = A.y1[jc + 1] - A.y1[jc], dZ = A.z1[kc + 1] - A.z1[kc];
    const double diagdis = sqrt(((dX * dX) + (dY * dY)) + (dZ * dZ));
    const double normdis = diagdis;
    const double na = A.ca[e], nb = A.cb[e], ng = A.cg[e];
    const double px = xs + normdis * na, py = ys + normdis * nb, pz = zs + normdis * ng;
    double dns[3];
    for (int q = 0; q < 4; q++) {
        const double *cx = A.C.x[q], *cy = A.C.y[q], *cz = A.C.z[q];
        int *cc = cell + (int64_t)(q + 1) * 3 * ne;
        int ix, iy, iz;
        if (A.find) {
            ix = bracket(cx, 2, nx + 1, px);
            iy = bracket(cy, 2, ny + 1, py);
            iz = bracket(cz, 2, nz + 1, pz);
            if (ix < 0 || iy < 0 || iz < 0) atomicOr(A.err, 8);
            cc[e] = ix; cc[ne + e] = iy; cc[2 * ne + e] = iz;
        } else {
            ix = cc[e]; iy = cc[ne + e]; iz = cc[2 * ne + e];
        }
        if (ix < 0 || iy < 0 || iz < 0) return;
        const int ax = ix - (q == 0), ay = iy - (q == 1), az = iz - (q == 2);
        double f[8];
#pragma unroll
        for (int t = 0; t < 8; t++) f[t] = A.F[q][lin(ax + (t & 1), ay + ((t >> 1) & 1), az + (t >> 2), sx, sy)];
        double fpos, dn;
        image_values(f, cx, cy, cz, ix, iy, iz, px, py, pz, na, nb, ng, &fpos, &dn);
        if (q < 3) {
            const double surf = A.wall[q];
            dns[q] = (2.0 / normdis) * (fpos - surf) - dn;
        } else {
            const double dpdn = wall_dpdn(A.acc, na, nb, ng);
            const double cval = fpos - ((dn + dpdn) * diagdis) * 0.5;
            const double alen = sqrt(((A.a3[e] * A.a3[e]) + (A.b3[e] * A.b3[e])) + (A.g3[e] * A.g3[e]));
            const double area = alen / 2.0;
            const double area_yz = 0.5 * fabs(A.a3[e]);
            const double area_xz = 0.5 * fabs(A.b3[e]);
            const double proj = (dns[0] * na + dns[1] * nb) + dns[2] * ng;
            const double stx1 = dns[0] - proj * na;
            const double sty1 = dns[1] - proj * nb;
            const double f_surf = cval * area;
            double *o = A.contrib;
            o[e] = ((1.0 / A.re) * stx1) * area;       // viscous drag
            o[ne + e] = -f_surf * na;                  // pressure drag
            o[2 * ne + e] = ((1.0 / A.re) * sty1) * area;  // viscous lift
            o[3 * ne + e] = -f_surf * nb;              // pressure lift
            o[4 * ne + e] = area;
            o[5 * ne + e] = area_xz;
            o[6 * ne + e] = area_yz;
        }
    }
}

// one block per quantity, fixed summation tree (deterministic)
__global__ void k_sum_rows(const double *v, int n, double *out) {
    __shared__ double sm[1024];
    const double *row = v + (int64_t)blockIdx.x * n;
    double s = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) s += row[q];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sm[0];
}

// per-body sums: block (row, body) adds the elements of that body, fixed tree
__global__ void k_sum_rows_bodies(const double *v, int n, const int *body, double *out) {
    __shared__ double sm[1024];
    const double *row = v + (int64_t)blockIdx.x * n;
    const int k = (int)blockIdx.y + 1;
    double s = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x)
        if (body[q] == k) s += row[q];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = sm[0];
}

// ------------------------------------------------------------------ host
static Coords coords(const Ctx &c) {
    Coords C;
    C.x[0] = c.xu; C.y[0] = c.yu; C.z[0] = c.zu;
    C.x[1] = c.xv; C.y[1] = c.yv; C.z[1] = c.zv;
    C.x[2] = c.xw; C.y[2] = c.yw; C.z[2] = c.zw;
    C.x[3] = c.xp; C.y[3] = c.yp; C.z[3] = c.zp;
    return C;
}

static int read_err(Ctx &c, int *dev_err) {
    int h = 0;
    IB3D_CHECK(hipMemcpyAsync(&h, dev_err, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    IB3D_CHECK(hipStreamSynchronize(c.stream));
    return h;
}

int shift_nodes(Ctx &c, double xs, double ys, double zs) {
    IB3D_REQUIRE(c.ibNodes > 0, "shiftSurfaceNodesInitial: no surface mesh");
    hipLaunchKernelGGL(k_shift_nodes, dim3(nblocks(c.ibNodes, 256)), dim3(256), 0, c.stream, c.xnode, c.ynode,
                       c.znode, c.xnode1, c.ynode1, c.znode1, c.ibNodes, xs, ys, zs);
    IB3D_CHECK(hipGetLastError());
    c.norms_ready = false;
    return 0;
}

int surface_norm(Ctx &c, int64_t inor) {
    IB3D_REQUIRE(c.ibElems > 0, "computeSurfaceNorm: no surface mesh");
    hipLaunchKernelGGL(k_surface_norm, dim3(nblocks(c.ibElems, 256)), dim3(256), 0, c.stream, c.xnode1, c.ynode1,
                       c.znode1, c.elp, c.ibElems, (double)inor, c.xcent, c.ycent, c.zcent, c.cosA, c.cosB, c.cosG,
                       c.alpha3, c.beta3, c.gamma3);
    IB3D_CHECK(hipGetLastError());
    c.norms_ready = true;
    return 0;
}

int shift_nodes_moving(Ctx &c, const double base[3], const double disp[3]) {
    IB3D_REQUIRE(c.ibNodes > 0, "shiftSurfaceNodesMovingbody: no surface mesh");
    hipLaunchKernelGGL(k_shift_nodes_moving, dim3(nblocks(c.ibNodes, 256)), dim3(256), 0, c.stream, c.xnode,
                       c.ynode, c.znode, c.xnode1, c.ynode1, c.znode1, c.ibNodes, base[0], base[1], base[2], disp[0],
                       disp[1], disp[2]);
    IB3D_CHECK(hipGetLastError());
    c.norms_ready = false;
    return 0;
}

static int *g_err = nullptr;
static int dev_err(Ctx &c) {
    if (!g_err) IB3D_CHECK(hipMalloc(&g_err, sizeof(int)));
    IB3D_CHECK(hipMemsetAsync(g_err, 0, sizeof(int), c.stream));
    return 0;
}
void count_tags_release();
void ib_release() {
    if (g_err) (void)hipFree(g_err);
    g_err = nullptr;
    count_tags_release();
}

int count_tags(Ctx &c, int32_t out[4], bool scatter);

// subSearch.f90:327-463 selectiveRetagging
int selective_retagging(Ctx &c, int32_t counts[3]) {
    IB3D_REQUIRE(c.norms_ready, "selectiveRetagging: computeSurfaceNorm has not run");
    IB3D_REQUIRE(c.lists_ready, "selectiveRetagging: cellCount has not run");
    const int64_t nb = c.ibCount;
    if (!c.rt_flag) {
        IB3D_CHECK(hipMalloc(&c.rt_flag, (c.P + 4) * sizeof(uint8_t)));
        IB3D_CHECK(hipMemsetAsync(c.rt_flag, 0, (c.P + 4) * sizeof(uint8_t), c.stream));
        IB3D_CHECK(hipMalloc(&c.rt_count, sizeof(int)));
    }
    const int64_t cap = std::min<int64_t>(c.P, std::max<int64_t>(1, 216 * nb));
    if (cap > c.rt_cap) {
        IB3D_CHECK(hipStreamSynchronize(c.stream));
        if (c.rt_list) IB3D_CHECK(hipFree(c.rt_list));
        IB3D_CHECK(hipMalloc(&c.rt_list, cap * sizeof(int)));
        c.rt_cap = cap;
    }
    IB3D_CHECK(hipMemsetAsync(c.minElem, 0, c.P * sizeof(int), c.stream));
    IB3D_CHECK(hipMemsetAsync(c.rt_count, 0, sizeof(int), c.stream));
    if (dev_err(c)) return -1;
    if (nb > 0) {
        hipLaunchKernelGGL(k_retag_mark, dim3(nblocks(216 * nb, 1024)), dim3(1024), 0, c.stream, c.ibList, (int)nb,
                           (int)c.nx, (int)c.ny, (int)c.nz, c.rt_flag, c.rt_list, c.rt_count);
        const unsigned grid = (unsigned)std::min<int64_t>(nblocks(cap, 64), 1024);
        hipLaunchKernelGGL(k_retag_list, dim3(grid), dim3(64 * kRtSlices), 0, c.stream, c.rt_list,
                           c.rt_count, (int)c.sx, (int)c.sy, c.xp, c.yp, c.zp, c.x1, c.y1, c.z1, c.xcent, c.ycent,
                           c.zcent, c.cosA, c.cosB, c.cosG, (int)c.ibElems, c.tag, c.minElem, c.rt_flag, c.u, c.v,
                           c.w, c.wall_vel[0], c.wall_vel[1], c.wall_vel[2], g_err);
    }
    IB3D_CHECK(hipGetLastError());
    IB3D_REQUIRE(read_err(c, g_err) == 0, "selectiveRetagging: a cell found no element closer than 1e14");
    int32_t t4[4];
    if (count_tags(c, t4, false)) return -1;
    counts[0] = t4[2];
    counts[1] = t4[0];
    counts[2] = t4[1];
    c.lists_ready = c.normdist_ready = c.plans_ready = false;
    c.plans_pending = false;
    c.stress_ready = false;
    return 0;
}

int tagging(Ctx &c, const int64_t w[6], int32_t counts[3]) {
    IB3D_REQUIRE(c.norms_ready, "tagging: computeSurfaceNorm has not run");
    IB3D_REQUIRE(w[0] >= 1 && w[1] <= c.nx + 1 && w[2] >= 1 && w[3] <= c.ny + 1 && w[4] >= 1 && w[5] <= c.nz + 1 &&
                     w[0] <= w[1] && w[2] <= w[3] && w[4] <= w[5],
                 "tagging: search window outside the grid");
    IB3D_CHECK(hipMemsetAsync(c.tag, 0, c.P, c.stream));
    IB3D_CHECK(hipMemsetAsync(c.minElem, 0, c.P * sizeof(int), c.stream));
    if (dev_err(c)) return -1;
    const int64_t nw = (w[1] - w[0] + 1) * (w[3] - w[2] + 1) * (w[5] - w[4] + 1);
    hipLaunchKernelGGL(k_tagging, dim3(nblocks(nw, kTagTile)), dim3(kTagTile), 0, c.stream, (int)w[0], (int)w[1],
                       (int)w[2], (int)w[3], (int)w[4], (int)w[5], (int)c.sx, (int)c.sy, c.xp, c.yp, c.zp, c.x1,
                       c.y1, c.z1, c.xcent, c.ycent, c.zcent, c.cosA, c.cosB, c.cosG, (int)c.ibElems, c.tag,
                       c.minElem, g_err);
    IB3D_CHECK(hipGetLastError());
    IB3D_REQUIRE(read_err(c, g_err) == 0, "tagging: a window cell found no element closer than 1e14");
    int32_t t4[4];
    if (count_tags(c, t4, false)) return -1;
    counts[0] = t4[2];
    counts[1] = t4[0];
    counts[2] = t4[1];
    c.lists_ready = false;
    c.normdist_ready = false;
    c.plans_ready = false;
    c.plans_pending = false;
    return 0;
}

// counts of fluid, solid, intercepted, red-fluid; optionally build the lists
static int *g_cnt = nullptr, *g_tot = nullptr;  // per-row counts (4 x rows) and totals, reused
static int g_cnt_rows = 0;

void count_tags_release() {
    if (g_cnt) (void)hipFree(g_cnt);
    if (g_tot) (void)hipFree(g_tot);
    g_cnt = g_tot = nullptr;
    g_cnt_rows = 0;
}

// grow-only device list of 3*n ints
static int grow_list(int *&p, int64_t &cap, int64_t n) {
    n = std::max<int64_t>(1, n);
    if (n <= cap) return 0;
    if (p) IB3D_CHECK(hipFree(p));
    p = nullptr;
    const int64_t c2 = n + n / 8;
    IB3D_CHECK(hipMalloc(&p, 3 * c2 * sizeof(int)));
    cap = c2;
    return 0;
}

int count_tags(Ctx &c, int32_t out[4], bool scatter) {
    const int nrows = (int)(c.ny * c.nz);
    const int ntile = (nrows + kScanTile - 1) / kScanTile;
    if (nrows > g_cnt_rows) {
        if (g_cnt) IB3D_CHECK(hipFree(g_cnt));
        if (!g_tot) IB3D_CHECK(hipMalloc(&g_tot, 4 * sizeof(int)));
        // 4 count arrays, then 4 x ntile tile sums
        IB3D_CHECK(hipMalloc(&g_cnt, (4 * (size_t)nrows + 4 * (size_t)ntile) * sizeof(int)));
        g_cnt_rows = nrows;
    }
    int *cnt = g_cnt, *tot = g_tot, *tsum = g_cnt + 4 * (size_t)nrows;
    hipLaunchKernelGGL(k_row_counts, dim3(nblocks(nrows, 4)), dim3(256), 0, c.stream, c.tag, (int)c.nx, (int)c.ny,
                       (int)c.nz, cnt);
    hipLaunchKernelGGL(k_tile_sums, dim3(ntile, 4), dim3(1024), 0, c.stream, cnt, nrows, tsum);
    hipLaunchKernelGGL(k_scan_counts, dim3(ntile, 4), dim3(1024), 0, c.stream, cnt, nrows, tsum, tot);
    IB3D_CHECK(hipGetLastError());
    int h[4];
    IB3D_CHECK(hipMemcpyAsync(h, tot, sizeof h, hipMemcpyDeviceToHost, c.stream));
    IB3D_CHECK(hipStreamSynchronize(c.stream));
    for (int a = 0; a < 4; a++) out[a] = h[a];
    if (scatter) {
        c.fluidCount = h[0];
        c.solidCount = h[1];
        c.ibCount = h[2];
        c.redCount = h[3];
        c.blackCount = h[0] - h[3];
        if (grow_list(c.fluidList, c.cap_lists[0], c.fluidCount) ||
            grow_list(c.solidList, c.cap_lists[1], c.solidCount) || grow_list(c.ibList, c.cap_lists[2], c.ibCount))
            return -1;
        hipLaunchKernelGGL(k_row_scatter, dim3(nblocks(nrows, 4)), dim3(256), 0, c.stream, c.tag, (int)c.nx,
                           (int)c.ny, (int)c.nz, cnt, c.fluidList, (int)c.fluidCount, c.solidList, (int)c.solidCount,
                           c.ibList, (int)c.ibCount);
        IB3D_CHECK(hipGetLastError());
        IB3D_CHECK(hipStreamSynchronize(c.stream));
    }
    return 0;
}

int cell_count(Ctx &c, int32_t counts[5]) {
    int32_t t4[4];
    if (count_tags(c, t4, true)) return -1;
    c.geom_version++;
    counts[0] = (int32_t)c.ibCount;
    counts[1] = (int32_t)c.fluidCount;
    counts[2] = (int32_t)c.solidCount;
    counts[3] = (int32_t)c.redCount;
    counts[4] = (int32_t)c.blackCount;
    c.lists_ready = true;
    c.normdist_ready = false;
    c.plans_ready = false;
    c.plans_pending = false;
    return 0;
}

int norm_distance(Ctx &c) {
    IB3D_REQUIRE(c.lists_ready, "computeNormDistance: cellCount has not run");
    const int64_t nb = std::max<int64_t>(1, c.ibCount);
    if (nb > c.cap_nel) {
        if (c.nel) IB3D_CHECK(hipFree(c.nel));
        if (c.normdis) IB3D_CHECK(hipFree(c.normdis));
        c.nel = nullptr;
        c.normdis = nullptr;
        const int64_t cap = nb + nb / 8;
        IB3D_CHECK(hipMalloc(&c.nel, cap * sizeof(int)));
        IB3D_CHECK(hipMalloc(&c.normdis, 7 * cap * sizeof(double)));
        c.cap_nel = cap;
    }
    if (dev_err(c)) return -1;
    if (c.ibCount > 0)
        hipLaunchKernelGGL(k_norm_distance, dim3(nblocks(c.ibCount, 256)), dim3(256), 0, c.stream, c.ibList,
                           (int)c.ibCount, c.minElem, (int)c.sx, (int)c.sy, c.xp, c.yp, c.zp, c.x1, c.y1, c.z1,
                           c.xcent, c.ycent, c.zcent, c.cosA, c.cosB, c.cosG, c.nel, c.normdis, g_err);
    IB3D_CHECK(hipGetLastError());
    IB3D_REQUIRE(read_err(c, g_err) == 0, "computeNormDistance: intercepted cell outside the search window");
    c.normdist_ready = true;
    return plans_launch(c);
}

static void free_plan(ForcePlan &p) {
    if (p.upd) (void)hipFree(p.upd);
    if (p.srcs) (void)hipFree(p.srcs);
    if (p.pos) (void)hipFree(p.pos);
    if (p.res) (void)hipFree(p.res);
    if (p.final_loc) (void)hipFree(p.final_loc);
    if (p.final_upd) (void)hipFree(p.final_upd);
    p = ForcePlan();
}

// ------------------------------------------------------------------------
// Forcing plans, built on the device.  The serial loop of velocityForcing1
// (pressureForcing1) visits the updates of one plan in order; update q reads
// the 8 corners of its image stencil and writes one ghost face (cell).  A
// corner that an earlier update of the loop has written must take the value
// of the latest such update, and the last writer of a face decides its final
// value (subForcing.f90:137-652).  Here:
//   1. the write keys (location*4 + component) of the plan are sorted with
//      their serial positions (a stable radix sort keeps equal keys in serial
//      order);
//   2. every corner finds its latest earlier writer by binary search in that
//      order (none: the field value);
//   3. the level of an update is one more than the highest level among its
//      sources, found by relaxing all updates in place until nothing changes
//      (levels only grow and are bounded by the longest chain);
//   4. a stable radix sort by level gives the launch order (serial order
//      within a level, as the reference's loop visits them), and the last
//      entry of every key segment of step 1 is that face's final writer.
// Only the level bounds, the number of final writers and the stencil error
// flag come back to the host, with one small copy.
// ------------------------------------------------------------------------
static int *g_doi = nullptr;        // device: stencil records of the 7 updates per IB cell
static double *g_dod = nullptr;
static int64_t g_dcap = 0;

// Work arrays of the device builder (shared by both plans, grow only).
struct PlanBuild {
    int *wk = nullptr, *wq = nullptr;        // write keys and serial positions
    int *sk = nullptr, *sq = nullptr;        // the same, sorted by key
    int *src = nullptr;                      // [8][m] sources (serial positions, -1: field)
    int *lev = nullptr, *slev = nullptr;     // levels, sorted levels
    int *ord = nullptr;                      // serial positions in launch order
    int *pos_of = nullptr;                   // launch position of every serial position
    int *fflag = nullptr, *fidx = nullptr;   // final-writer flags and their exclusive scan
    int *info = nullptr;                     // [0] error, [1] n_final, [2] n_levels, [3..] level starts,
                                             // then per relaxation sweep a "changed" flag
    int *hinfo = nullptr;                    // pinned mirror of info
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    int64_t cap = 0;
};
static PlanBuild g_pb;
static constexpr int kMaxLevels = 256, kRelaxBatch = 16, kInfoInts = 4 + kMaxLevels + 1 + 1024;

static void free_build() {
    PlanBuild &B = g_pb;
    for (int *q : {B.wk, B.wq, B.sk, B.sq, B.src, B.lev, B.slev, B.ord, B.pos_of, B.fflag, B.fidx, B.info})
        if (q) (void)hipFree(q);
    if (B.hinfo) (void)hipHostFree(B.hinfo);
    if (B.tmp) (void)hipFree(B.tmp);
    B = PlanBuild();
}

void release_plans(Ctx &c) {
    free_plan(c.vplan);
    free_plan(c.pplan);
    if (g_doi) (void)hipFree(g_doi);
    if (g_dod) (void)hipFree(g_dod);
    g_doi = nullptr;
    g_dod = nullptr;
    g_dcap = 0;
    free_build();
}

// serial position q of a plan -> update id (velocity: per cell u2 u1 v2 v1
// w2 w1, ids 7n..7n+5; pressure: id 7n+6)
__device__ __forceinline__ int plan_uid(int q, int vel) { return vel ? 7 * (q / 6) + q % 6 : 7 * q + 6; }

__global__ void k_plan_keys(const int *__restrict__ oi, int m, int vel, int *wk, int *wq) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    const int *o = oi + 8 * (int64_t)plan_uid(q, vel);
    wk[q] = o[1] * 4 + o[0];
    wq[q] = q;
}

// latest writer before q of every stencil corner, by binary search in the
// key-sorted writes (positions ascending within a key)
__global__ void k_plan_sources(const int *__restrict__ oi, int m, int vel, int sx, int sy, const int *__restrict__ sk,
                               const int *__restrict__ sq, int *src, int *lev) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    const int *o = oi + 8 * (int64_t)plan_uid(q, vel);
    const int comp = o[0];
    const int ax = o[2] - (comp == 0) - 1, ay = o[3] - (comp == 1) - 1, az = o[4] - (comp == 2) - 1;
    for (int t = 0; t < 8; t++) {
        const int loc = (ax + (t & 1)) + sx * ((ay + ((t >> 1) & 1)) + sy * (az + (t >> 2)));
        const int key = loc * 4 + comp;
        int lo = 0, hi = m;                  // first index with sk >= key
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (sk[mid] < key) lo = mid + 1; else hi = mid;
        }
        int l2 = lo, h2 = m;                 // first index in the segment with sq >= q
        while (l2 < h2) {
            const int mid = (l2 + h2) >> 1;
            if (sk[mid] == key && sq[mid] < q) l2 = mid + 1; else h2 = mid;
        }
        src[t * (int64_t)m + q] = (l2 > lo) ? sq[l2 - 1] : -1;
    }
    lev[q] = 0;
}

// one in-place relaxation sweep; changed[0] = 1 if any level grew.  Skipped
// once the previous sweep changed nothing.
__global__ void k_plan_relax(int m, const int *__restrict__ src, int *lev, int *changed, int first) {
    if (!first && changed[-1] == 0) return;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    int l = 0;
    for (int t = 0; t < 8; t++) {
        const int s = src[t * (int64_t)m + q];
        if (s >= 0) l = max(l, __hip_atomic_load(&lev[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1);
    }
    if (l > lev[q]) {
        __hip_atomic_store(&lev[q], l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        changed[0] = 1;
    }
}

__global__ void k_plan_fill(int *a, int n, int v) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) a[q] = v;
}

// level bounds from the level-sorted array: info[3 + l] = first position of
// level l, info[2] = number of levels (levels are 0..max without gaps: a
// level-l update has a source at level l-1)
__global__ void k_plan_bounds(int m, const int *__restrict__ slev, const int *__restrict__ ord, int *pos_of,
                              int *info) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m) return;
    pos_of[ord[p]] = p;
    const int l = slev[p];
    if (l >= kMaxLevels) {
        info[0] = 2;
        return;
    }
    if (p == 0 || slev[p - 1] != l) info[3 + l] = p;
    if (p == m - 1) {
        info[2] = l + 1;
        info[3 + l + 1] = m;
    }
}

// the plan in launch order: stencil record, image point, sources as launch
// positions; final writers: the last write of every key segment
__global__ void k_plan_emit(const int *__restrict__ oi, const double *__restrict__ od, int m, int vel,
                            const int *__restrict__ ord, const int *__restrict__ pos_of, const int *__restrict__ src,
                            int *upd, int *srcs, double *pos) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m) return;
    const int q = ord[p];
    const int u = plan_uid(q, vel);
    for (int t = 0; t < 8; t++) {
        upd[8 * (int64_t)p + t] = oi[8 * (int64_t)u + t];
        const int sv = src[t * (int64_t)m + q];
        srcs[8 * (int64_t)p + t] = sv >= 0 ? pos_of[sv] : -1;
    }
    for (int t = 0; t < 5; t++) pos[5 * (int64_t)p + t] = od[5 * (int64_t)u + t];
}

__global__ void k_plan_final_flags(int m, const int *__restrict__ sk, int *fflag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    fflag[i] = (i == m - 1 || sk[i + 1] != sk[i]) ? 1 : 0;
}

__global__ void k_plan_finals(int m, const int *__restrict__ sk, const int *__restrict__ sq,
                              const int *__restrict__ fflag, const int *__restrict__ fidx,
                              const int *__restrict__ pos_of, int *fin, int nf_cap, int *info) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    if (i == m - 1) info[1] = fidx[i] + fflag[i];
    if (!fflag[i]) return;
    const int f = fidx[i];
    fin[f] = sk[i] >> 2;
    fin[nf_cap + f] = pos_of[sq[i]];
    fin[2 * nf_cap + f] = sk[i] & 3;
}

static int build_grow(Ctx &c, int64_t m) {
    PlanBuild &B = g_pb;
    if (!B.info) {
        IB3D_CHECK(hipMalloc(&B.info, kInfoInts * sizeof(int)));
        IB3D_CHECK(hipHostMalloc((void **)&B.hinfo, kInfoInts * sizeof(int), hipHostMallocDefault));
    }
    if (m <= B.cap) return 0;
    IB3D_CHECK(hipStreamSynchronize(c.stream));
    for (int **q : {&B.wk, &B.wq, &B.sk, &B.sq, &B.src, &B.lev, &B.slev, &B.ord, &B.pos_of, &B.fflag, &B.fidx}) {
        if (*q) IB3D_CHECK(hipFree(*q));
        *q = nullptr;
    }
    const int64_t cap = m + m / 4;
    for (int **q : {&B.wk, &B.wq, &B.sk, &B.sq, &B.lev, &B.slev, &B.ord, &B.pos_of, &B.fflag, &B.fidx})
        IB3D_CHECK(hipMalloc(q, (size_t)cap * sizeof(int)));
    IB3D_CHECK(hipMalloc(&B.src, 8 * (size_t)cap * sizeof(int)));
    // radix sort and scan scratch for the largest size
    size_t b1 = 0, b2 = 0;
    IB3D_CHECK(rocprim::radix_sort_pairs(nullptr, b1, B.wk, B.sk, B.wq, B.sq, (unsigned)cap, 0, 32, c.stream));
    IB3D_CHECK(rocprim::exclusive_scan(nullptr, b2, B.fflag, B.fidx, 0, (size_t)cap, rocprim::plus<int>(), c.stream));
    const size_t need = std::max(b1, b2);
    if (need > B.tmp_bytes) {
        if (B.tmp) IB3D_CHECK(hipFree(B.tmp));
        IB3D_CHECK(hipMalloc(&B.tmp, need));
        B.tmp_bytes = need;
    }
    B.cap = cap;
    return 0;
}

static int plan_capacity(ForcePlan &P, int m, int nf_cap) {
    if (m > P.cap_upd) {
        if (P.upd) { (void)hipFree(P.upd); (void)hipFree(P.srcs); (void)hipFree(P.pos); (void)hipFree(P.res); }
        const int cap = m + m / 4;
        IB3D_CHECK(hipMalloc(&P.upd, 8 * (size_t)cap * sizeof(int)));
        IB3D_CHECK(hipMalloc(&P.srcs, 8 * (size_t)cap * sizeof(int)));
        IB3D_CHECK(hipMalloc(&P.pos, 5 * (size_t)cap * sizeof(double)));
        IB3D_CHECK(hipMalloc(&P.res, (size_t)cap * sizeof(double)));
        P.cap_upd = cap;
    }
    if (3 * nf_cap > P.cap_final) {
        if (P.final_loc) (void)hipFree(P.final_loc);
        IB3D_CHECK(hipMalloc(&P.final_loc, 3 * (size_t)nf_cap * sizeof(int)));
        P.cap_final = 3 * nf_cap;
    }
    return 0;
}

// Build one plan on the device; its level bounds and final-writer count come
// back with the copy of `info` (at *host_info once the stream reaches it).
static int plan_device(Ctx &c, ForcePlan &P, int vel, int nb) {
    PlanBuild &B = g_pb;
    const int m = vel ? 6 * nb : nb;
    // every update writes one face, so there are at most m final writers;
    // the final table is sized m and its three columns are m apart
    if (build_grow(c, m) || plan_capacity(P, m, m)) return -1;
    const dim3 tb(256), gb(nblocks(m, 256));
    int bits = 1;
    while (bits < 32 && ((int64_t)1 << bits) <= 4 * c.P) bits++;
    hipLaunchKernelGGL(k_plan_keys, gb, tb, 0, c.stream, g_doi, m, vel, B.wk, B.wq);
    size_t tb_bytes = B.tmp_bytes;
    IB3D_CHECK(rocprim::radix_sort_pairs(B.tmp, tb_bytes, B.wk, B.sk, B.wq, B.sq, (unsigned)m, 0, bits, c.stream));
    hipLaunchKernelGGL(k_plan_sources, gb, tb, 0, c.stream, g_doi, m, vel, (int)c.sx, (int)c.sy, B.sk, B.sq, B.src,
                       B.lev);
    // relaxation sweeps in batches, each gated on the one before
    int *chg = B.info + 4 + kMaxLevels + 1;
    int done = 0;
    for (;;) {
        IB3D_REQUIRE(done + kRelaxBatch <= 1024, "forcing: the plan's dependency chain is too long");
        hipLaunchKernelGGL(k_plan_fill, dim3(1), dim3(kRelaxBatch), 0, c.stream, chg + done, kRelaxBatch, 0);
        for (int r = 0; r < kRelaxBatch; r++)
            hipLaunchKernelGGL(k_plan_relax, gb, tb, 0, c.stream, m, B.src, B.lev, chg + done + r,
                               (done + r == 0) ? 1 : 0);
        done += kRelaxBatch;
        IB3D_CHECK(hipMemcpyAsync(&B.hinfo[0], chg + done - 1, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        IB3D_CHECK(hipStreamSynchronize(c.stream));
        if (!B.hinfo[0]) break;
    }
    tb_bytes = B.tmp_bytes;
    IB3D_CHECK(rocprim::radix_sort_pairs(B.tmp, tb_bytes, B.lev, B.slev, B.wq, B.ord, (unsigned)m, 0, 9, c.stream));
    hipLaunchKernelGGL(k_plan_fill, dim3(1), dim3(64), 0, c.stream, B.info, 4, 0);
    hipLaunchKernelGGL(k_plan_bounds, gb, tb, 0, c.stream, m, B.slev, B.ord, B.pos_of, B.info);
    hipLaunchKernelGGL(k_plan_emit, gb, tb, 0, c.stream, g_doi, g_dod, m, vel, B.ord, B.pos_of, B.src, P.upd, P.srcs,
                       P.pos);
    hipLaunchKernelGGL(k_plan_final_flags, gb, tb, 0, c.stream, m, B.sk, B.fflag);
    tb_bytes = B.tmp_bytes;
    IB3D_CHECK(rocprim::exclusive_scan(B.tmp, tb_bytes, B.fflag, B.fidx, 0, (size_t)m, rocprim::plus<int>(),
                                       c.stream));
    hipLaunchKernelGGL(k_plan_finals, gb, tb, 0, c.stream, m, B.sk, B.sq, B.fflag, B.fidx, B.pos_of, P.final_loc, m,
                       B.info);
    IB3D_CHECK(hipGetLastError());
    IB3D_CHECK(hipMemcpyAsync(B.hinfo, B.info, (4 + kMaxLevels + 1) * sizeof(int), hipMemcpyDeviceToHost, c.stream));
    IB3D_CHECK(hipStreamSynchronize(c.stream));
    IB3D_REQUIRE(B.hinfo[0] == 0, "forcing: the plan has more than 256 dependency levels");
    P.n_updates = m;
    P.n_levels = B.hinfo[2];
    P.n_final = B.hinfo[1];
    P.final_stride = m;
    P.level_start.assign(B.hinfo + 3, B.hinfo + 3 + P.n_levels + 1);
    return 0;
}

// The forcing plans are built in two halves.  plans_launch (at the end of
// computeNormDistance) queues the stencil kernel; plans_finish, at the first
// forcing call, checks the stencil error flag and builds both plans on the
// device (their kernels queue behind the predictor's).
int plans_launch(Ctx &c) {
    c.plans_ready = false;
    c.plans_pending = false;
    c.vplan.n_updates = c.pplan.n_updates = 0;
    c.vplan.n_final = c.pplan.n_final = 0;
    const int nb = (int)c.ibCount;
    if (nb == 0) {
        c.vplan.level_start.assign(1, 0);
        c.vplan.n_levels = 0;
        c.pplan.level_start.assign(1, 0);
        c.pplan.n_levels = 0;
        c.plans_ready = true;
        return 0;
    }
    IB3D_REQUIRE(4 * c.P < ((int64_t)1 << 31), "forcing: grid too large for 32-bit face keys");
    const int64_t nu7 = 7 * (int64_t)nb;
    if (nu7 > g_dcap) {
        if (g_doi) (void)hipFree(g_doi);
        if (g_dod) (void)hipFree(g_dod);
        const int64_t cap = nu7 + nu7 / 4;
        IB3D_CHECK(hipMalloc(&g_doi, 8 * (size_t)cap * sizeof(int)));
        IB3D_CHECK(hipMalloc(&g_dod, 5 * (size_t)cap * sizeof(double)));
        g_dcap = cap;
    }
    if (dev_err(c)) return -1;
    StencilArgs A;
    A.ib = c.ibList; A.nb = nb; A.nel = c.nel; A.nd = c.normdis;
    A.dx = c.dx; A.dy = c.dy; A.dz = c.dz; A.ca = c.cosA; A.cb = c.cosB; A.cg = c.cosG;
    A.C = coords(c);
    A.nx = (int)c.nx; A.ny = (int)c.ny; A.nz = (int)c.nz;
    hipLaunchKernelGGL(k_force_stencil, dim3(nblocks(nu7, 256)), dim3(256), 0, c.stream, A, g_doi, g_dod, g_err);
    IB3D_CHECK(hipGetLastError());
    c.plans_pending = true;
    return 0;
}

static int plans_finish(Ctx &c) {
    c.plans_pending = false;
    const int nb = (int)c.ibCount;  // unchanged since plans_launch: a new list clears the pending flag
    IB3D_REQUIRE(read_err(c, g_err) == 0, "forcing: an image point has no bracketing grid lines");
    if (plan_device(c, c.vplan, 1, nb) || plan_device(c, c.pplan, 0, nb)) return -1;
    c.plans_ready = true;
    return 0;
}

static int run_plan(Ctx &c, ForcePlan &P) {
    if (P.n_updates == 0) return 0;
    LevelArgs A;
    A.oi = P.upd; A.od = P.pos; A.srcs = P.srcs; A.res = P.res;
    A.F[0] = c.ut; A.F[1] = c.vt; A.F[2] = c.wt; A.F[3] = c.p;
    A.ca = c.cosA; A.cb = c.cosB; A.cg = c.cosG;
    A.C = coords(c);
    A.sx = (int)c.sx; A.sy = (int)c.sy;
    for (int q = 0; q < 3; q++) {
        A.wall[q] = c.wall_vel[q];
        A.acc[q] = c.wall_acc[q];
    }
    for (int l = 0; l < P.n_levels; l++) {
        const int lo = P.level_start[l], hi = P.level_start[l + 1];
        if (hi > lo)
            hipLaunchKernelGGL(k_force_level, dim3(nblocks(hi - lo, 128)), dim3(128), 0, c.stream, A, lo, hi);
    }
    const int nf = P.n_final, fs = P.final_stride;
    hipLaunchKernelGGL(k_force_final, dim3(nblocks(nf, 256)), dim3(256), 0, c.stream, P.final_loc,
                       P.final_loc + fs, P.final_loc + 2 * fs, nf, P.res, c.ut, c.vt, c.wt, c.p);
    IB3D_CHECK(hipGetLastError());
    return 0;
}

int velocity_forcing(Ctx &c) {
    if (c.plans_pending && plans_finish(c)) return -1;
    IB3D_REQUIRE(c.plans_ready, "velocityForcing1: computeNormDistance has not run");
    return run_plan(c, c.vplan);
}

int pressure_forcing(Ctx &c) {
    if (c.plans_pending && plans_finish(c)) return -1;
    IB3D_REQUIRE(c.plans_ready, "pressureForcing1: computeNormDistance has not run");
    return run_plan(c, c.pplan);
}

int stress(Ctx &c, int first, double re, double coef[4]) {
    IB3D_REQUIRE(c.norms_ready, "stressCal: computeSurfaceNorm has not run");
    const int ne = (int)c.ibElems;
    if (!c.scell) IB3D_CHECK(hipMalloc(&c.scell, 15 * (size_t)ne * sizeof(int)));
    if (!c.scontrib) IB3D_CHECK(hipMalloc(&c.scontrib, 7 * (size_t)ne * sizeof(double)));
    double *contrib = c.scontrib;
    if (dev_err(c)) return -1;
    StressArgs A;
    A.xc = c.xcent; A.yc = c.ycent; A.zc = c.zcent; A.ca = c.cosA; A.cb = c.cosB; A.cg = c.cosG;
    A.a3 = c.alpha3; A.b3 = c.beta3; A.g3 = c.gamma3;
    A.x1 = c.x1; A.y1 = c.y1; A.z1 = c.z1;
    A.C = coords(c);
    A.F[0] = c.ut; A.F[1] = c.vt; A.F[2] = c.wt; A.F[3] = c.p;
    A.ne = ne; A.nx = (int)c.nx; A.ny = (int)c.ny; A.nz = (int)c.nz;
    A.re = re;
    A.cache = c.scell;
    A.find = (first || !c.stress_ready) ? 1 : 0;
    A.contrib = contrib;
    A.err = g_err;
    for (int q = 0; q < 3; q++) {
        A.wall[q] = c.wall_vel[q];
        A.acc[q] = c.wall_acc[q];
    }
    hipLaunchKernelGGL(k_stress, dim3(nblocks(ne, 128)), dim3(128), 0, c.stream, A);
    hipLaunchKernelGGL(k_sum_rows, dim3(7), dim3(1024), 0, c.stream, contrib, ne, c.scalars + 8);
    IB3D_CHECK(hipGetLastError());
    IB3D_CHECK(hipMemcpyAsync(c.hscal + 8, c.scalars + 8, 7 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    const int err = read_err(c, g_err);
    IB3D_REQUIRE(err == 0, "stressCal: a probe point lies outside the grid");
    c.stress_ready = true;
    const double *s = c.hscal + 8;
    double area_Sx = 0.5 * s[5], area_Sy = 0.5 * s[6];
    coef[0] = 2.0 * (s[0] / area_Sx);
    coef[1] = 2.0 * (s[1] / area_Sy);
    coef[2] = 2.0 * (s[2] / area_Sy);
    coef[3] = 2.0 * (s[3] / area_Sx);
    return 0;
}

}  // namespace ib3d

namespace ib3d {

// Per-body coefficients from the element contributions of the last stress
// call: the sums of subStressCal1.f90:327-331 and the normalisation of
// 441-450 restricted to each body.
int stress_bodies(Ctx &c, double *coef) {
    IB3D_REQUIRE(c.body && c.nbody > 0, "stress_bodies: element bodies not set");
    IB3D_REQUIRE(c.stress_ready, "stress_bodies: no stressCal has run");
    const int ne = (int)c.ibElems, nb = c.nbody;
    double *dsum = nullptr;
    IB3D_CHECK(hipMalloc(&dsum, 7 * (size_t)nb * sizeof(double)));
    hipLaunchKernelGGL(k_sum_rows_bodies, dim3(7, nb), dim3(1024), 0, c.stream, c.scontrib, ne, c.body, dsum);
    IB3D_CHECK(hipGetLastError());
    std::vector<double> h(7 * (size_t)nb);
    IB3D_CHECK(d2h(h.data(), dsum, h.size() * sizeof(double)));
    IB3D_CHECK(hipFree(dsum));
    for (int k = 0; k < nb; k++) {
        const double *sk = h.data() + 7 * (size_t)k;
        const double area_Sx = 0.5 * sk[5], area_Sy = 0.5 * sk[6];
        coef[4 * k] = 2.0 * (sk[0] / area_Sx);
        coef[4 * k + 1] = 2.0 * (sk[1] / area_Sy);
        coef[4 * k + 2] = 2.0 * (sk[2] / area_Sy);
        coef[4 * k + 3] = 2.0 * (sk[3] / area_Sx);
    }
    return 0;
}

}  // namespace ib3d
