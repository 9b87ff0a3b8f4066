// ib3d_sor.hip -- pressure-correction Poisson solve on colour-compacted
// arrays (reference: subPcorVcor.f90:2-212).
//
// The reference keeps pc and pco and sweeps red cells from pco, black cells
// from pc, then copies pc to pco (subPcorVcor.f90:164-201).  Red neighbours
// are black and vice versa, so one in-place array per colour reproduces that
// arithmetic bit for bit: a red update reads the black values of the previous
// iteration and its own previous value; a black update reads the new red
// values.  max|pc-pco| is the max over all updates of |new-old|, fused into
// the sweeps.
//
// Algorithmic HBM traffic per interior cell and red+black sweep: own value
// read + write, the other colour read once as neighbours, b/dt read once =
// 32 B (FP64) or 16 B (FP32).
//
// Kernels
//   k_sor_pz  the half-sweep (FP64 and FP32).  A thread owns two neighbouring
//             cells of one colour row and walks kZch planes in z, keeping the
//             other colour's values below / at / above in registers (the "at"
//             pair also supplies the x neighbours); colour rows have even
//             length so every pair is an aligned vector access, and the
//             coefficients are hoisted out of the z loop.
//   k_sor     one cell per thread: the smoother on the coarse multigrid
//             levels, where a z-marching thread would have too few peers.
//   k_resid_pz  FP64 residual b/dt - A p over both colours (optionally of
//             p + e), written as FP32 for the next inner solve.
//   k_matvec32_pz, k_dots, k_orth, k_givens, k_lsq, k_update_pc
//             FP32 flexible GMRES inside the FP64 refinement (modes 3, 4).
//   k_mg_*    the FP32 multigrid V-cycle that preconditions it in mode 4.
//
// Host loops never wait on every iteration: sweeps are queued in batches and
// each kernel first reads a device-side convergence flag of the previous
// iteration, so a batch that overshoots does no work and changes nothing.
#include "ib3d_internal.h"

#include <cmath>
#include <cstring>
#include <type_traits>

namespace ib3d {

static constexpr int kSorBlock = 256;
static constexpr int kZch = 8;

// Workgroups are dispatched round-robin over the 8 XCDs; give XCD x a
// contiguous range of logical blocks.
__device__ __forceinline__ int xcd_block() {
    const unsigned nb = gridDim.x, b = blockIdx.x;
    const unsigned nfull = nb & ~7u;
    if (b >= nfull) return (int)b;
    return (int)((b & 7u) * (nb >> 3) + (b >> 3));
}

// The same for the 2-D (row block, z chunk) grids of the z-marching kernels:
// the hardware deals linear block ids (x fastest) round-robin to the XCDs, so
// consecutive ids would put neighbouring row blocks, whose y-neighbour rows
// and chunk-boundary planes overlap, in different L2 caches.  Each XCD gets a
// contiguous range of logical ids (row blocks of a few consecutive z chunks).
__device__ __forceinline__ void xcd_block2(int &bx, int &by) {
    const unsigned gx = gridDim.x, nb = gx * gridDim.y, b = blockIdx.y * gx + blockIdx.x;
    const unsigned nfull = nb & ~7u;
    const unsigned l = b >= nfull ? b : (b & 7u) * (nb >> 3) + (b >> 3);
    by = (int)(l / gx);
    bx = (int)(l - (unsigned)by * gx);
}

__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// block max of non-negative doubles -> atomicMax on the bit pattern (IEEE
// order of non-negative doubles equals unsigned integer order; a NaN maps
// above +inf and therefore propagates).
__device__ __forceinline__ void block_max_to(double v, double *dst) {
    __shared__ double sm[kSorBlock / 64];
    v = wave_max(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sm[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = sm[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) m = fmax(m, sm[q]);
        if (m > 0.0 || m != m)
            atomicMax(reinterpret_cast<unsigned long long *>(dst), (unsigned long long)__double_as_longlong(m));
    }
}

// deterministic block sum -> partial[slot]
__device__ __forceinline__ void block_sum_to(double v, double *partial, int64_t slot) {
    __shared__ double sm[kSorBlock / 64];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sm[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) s += sm[q];
        partial[slot] = s;
    }
}

// skip this launch when the previous iteration already converged
__device__ __forceinline__ bool gated(const double *gate, double thr) { return gate && *gate < thr; }

// One 1024-thread workgroup reduces n partials (tens of thousands after a
// full-grid kernel): each thread takes q = t, t + 1024, ... in order, with
// eight loads in flight ahead of its running value, then a tree.
__global__ void __launch_bounds__(1024) k_max_partials(const double *partial, int64_t n, double *out) {
    __shared__ double sm[1024];
    double m = 0.0;
    const int64_t st = blockDim.x;
    int64_t q = threadIdx.x;
    for (; q + 7 * st < n; q += 8 * st) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = partial[q + u * st];
#pragma unroll
        for (int u = 0; u < 8; u++) m = nan_max(m, v[u]);
    }
    for (; q < n; q += st) m = nan_max(m, partial[q]);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] = nan_max(sm[threadIdx.x], sm[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sm[0];
}

void launch_max_partials(const double *partial, int64_t n, double *out) {
    hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(1024), 0, ctx().stream, partial, n, out);
}

// the same for a sum; the additions keep the plain loop's order
__global__ void __launch_bounds__(1024) k_sum_partials(const double *partial, int64_t n, double *out) {
    __shared__ double sm[1024];
    double s = 0.0;
    const int64_t st = blockDim.x;
    int64_t q = threadIdx.x;
    for (; q + 7 * st < n; q += 8 * st) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = partial[q + u * st];
#pragma unroll
        for (int u = 0; u < 8; u++) s += v[u];
    }
    for (; q < n; q += st) s += partial[q];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sm[0];
}

// ||r|| from block partials; rel[o] = ||r||/||b/dt||, conv[o] = 1 while
// rel > tol (the loop continues), 0 once converged.  When gated the previous
// iteration's values are carried forward.
__global__ void k_finish_norm(const double *partial, int64_t n, double bn, double tol, double *rel, double *conv,
                              const double *gate, const double *bnp) {
    __shared__ double sm[1024];
    if (gate && *gate < 0.5) {
        if (threadIdx.x == 0) {
            rel[0] = rel[-1];
            conv[0] = 0.0;
        }
        return;
    }
    double s = 0.0;
    for (int64_t q = threadIdx.x; q < n; q += blockDim.x) s += partial[q];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (bnp) bn = *bnp;   // (a captured cycle reads the solve's norm from device memory)
        const double r = bn > 0.0 ? sqrt(sm[0]) / bn : 0.0;
        rel[0] = r;
        conv[0] = (r > tol || r != r) ? 1.0 : 0.0;
    }
}

struct SorGeom {
    int H, NC, sy, sx, sz, plane;  // plane = H*sy
    int sxp;                       // row stride of the x coefficient table (multiple of 8)
    FastDiv dH, dSy;
};

__host__ __device__ __forceinline__ void decode(const SorGeom &g, int f, int c, int &ii, int &jj, int &kk) {
    const int row = (int)fdiv((uint32_t)f, g.dH);
    const int h = f - row * g.H;
    kk = (int)fdiv((uint32_t)row, g.dSy);
    jj = row - kk * g.sy;
    ii = 2 * h + ((c + jj + kk) & 1);
}

// A row of the residual kernels: an interior cell.  With the finite b/dt
// of a live cell this is what k_resid tests; the interior test keeps every
// neighbour access inside the arrays whatever the NaN sentinel holds.
__host__ __device__ __forceinline__ bool interior_cell(const SorGeom &g, int ii, int jj, int kk) {
    return ii >= 1 && ii <= g.sx - 2 && jj >= 1 && jj <= g.sy - 2 && kk >= 1 && kk <= g.sz - 2;
}

// FP64 keeps the reference's correctly rounded division; the FP32 smoother
// of the defect-correction path multiplies by the hardware reciprocal (its
// rounding only perturbs the inner iterate; the FP64 outer residual decides).
__device__ __forceinline__ double sor_div(double a, double b) { return a / b; }
__device__ __forceinline__ float sor_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// One colour half-sweep, one cell per thread.
// FIRST: the unknown starts from zero (own old value and, for the red
// colour, the other colour are taken as 0 without reading them).
// APPLY (with FIRST): the previous inner solve's increment still held in
// `mine` is first added to the FP64 unknown `acc` (defect correction).
template <typename T, bool TRACK, bool FIRST, bool APPLY>
__global__ void __launch_bounds__(kSorBlock)
k_sor(T *__restrict__ mine, const T *__restrict__ other, const T *__restrict__ rhs,
      const T *__restrict__ axc, const T *__restrict__ ayc, const T *__restrict__ azc, SorGeom g,
      int colour, T omega, T omm, double *slot, const double *gate, double thr, double *__restrict__ acc) {
    if (gated(gate, thr)) return;
    const int f = xcd_block() * kSorBlock + (int)threadIdx.x;
    double d = 0.0;
    if (f < g.NC) {
        const T r = rhs[f];
        if (r == r) {  // NaN sentinel marks cells that are not updated
            if (APPLY) acc[f] += (double)mine[f];
            int ii, jj, kk;
            decode(g, f, colour, ii, jj, kk);
            const int q = (colour + jj + kk) & 1;
            const T a1x = axc[ii], a2x = axc[g.sxp + ii], a3x = axc[2 * g.sxp + ii];
            const T a1y = ayc[jj], a2y = ayc[g.sy + jj], a3y = ayc[2 * g.sy + jj];
            const T a1z = azc[kk], a2z = azc[g.sz + kk], a3z = azc[2 * g.sz + kk];
            T pw, pe, ps, pn, pb, pt, pold;
            if (FIRST && colour == 0) {
                pw = pe = ps = pn = pb = pt = T(0);
            } else {
                pw = other[f - 1 + q];
                pe = other[f + q];
                ps = other[f - g.H];
                pn = other[f + g.H];
                pb = other[f - g.plane];
                pt = other[f + g.plane];
            }
            pold = FIRST ? T(0) : mine[f];
            const T gs = sor_div(((((((r - a1x * pw) - a3x * pe) - a1y * ps) - a3y * pn) - a1z * pb) - a3z * pt),
                                 ((a2x + a2y) + a2z));
            const T pnew = omm * pold + omega * gs;
            mine[f] = pnew;
            if (TRACK) d = (double)fabs(pnew - pold);
        }
    }
    if (TRACK) block_max_to(d, slot);
}

// Half-sweep on cell pairs marching in z (see the file comment).  Cell 0 of
// a pair sits at ii = 2h+q, cell 1 at ii = 2h+2+q (q = row parity); their x
// neighbours are the other colour's pair at the same h plus one value on the
// west (q = 0) or east (q = 1) side.  The x coefficients of both parities are
// loaded once (h is even, so 2h..2h+3 is one aligned 4-vector), the y
// coefficients once per thread and the z coefficients are uniform per plane.
// TRACK: max |new - old| (reference stopping rule).  FIRST: the unknown starts
// from zero.  APPLY (with FIRST): the previous increment, still in `mine`, is
// first added to the FP64 unknown `acc` (defect correction).
typedef double d2v __attribute__((ext_vector_type(2)));
typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
template <typename T> struct Vec;
template <> struct Vec<double> { typedef d2v v2; typedef d4v v4; };
template <> struct Vec<float> { typedef f2v v2; typedef f4v v4; };

template <typename T, bool TRACK, bool FIRST, bool APPLY>
__global__ void __launch_bounds__(kSorBlock)
k_sor_pz(T *__restrict__ mine, const T *__restrict__ other, const T *__restrict__ rhs, const T *__restrict__ axc,
         const T *__restrict__ ayc, const T *__restrict__ azc, SorGeom g, int colour, T omega, T omm,
         double *slot, const double *gate, double thr, double *__restrict__ acc) {
    typedef typename Vec<T>::v2 V2;
    typedef typename Vec<T>::v4 V4;
    if (gated(gate, thr)) return;
    int bx, by;
    xcd_block2(bx, by);
    const int p = 2 * (int)(bx * kSorBlock + threadIdx.x);
    double d = 0.0;
    const int jj = (int)fdiv((uint32_t)p, g.dH);
    const int h = p - jj * g.H;
    if (p < g.plane && jj >= 1 && jj <= g.sy - 2) {
        const int k0 = 1 + by * kZch;
        const int k1 = min(k0 + kZch, g.sz - 1);
        const V4 c1 = *reinterpret_cast<const V4 *>(axc + 2 * h);
        const V4 c2 = *reinterpret_cast<const V4 *>(axc + g.sxp + 2 * h);
        const V4 c3 = *reinterpret_cast<const V4 *>(axc + 2 * g.sxp + 2 * h);
        const T a1y = ayc[jj], a2y = ayc[g.sy + jj], a3y = ayc[2 * g.sy + jj];
        int f = p + k0 * g.plane;
        const V2 zero = V2(T(0));
        const bool no_other = FIRST && colour == 0;
        V2 below = no_other ? zero : *reinterpret_cast<const V2 *>(other + f - g.plane);
        V2 cur = no_other ? zero : *reinterpret_cast<const V2 *>(other + f);
        for (int kk = k0; kk < k1; kk++, f += g.plane) {
            const V2 above = no_other ? zero : *reinterpret_cast<const V2 *>(other + f + g.plane);
            const V2 r2 = *reinterpret_cast<const V2 *>(rhs + f);
            const bool ok0 = r2.x == r2.x, ok1 = r2.y == r2.y;
            if (ok0 || ok1) {
                const int q = (colour + jj + kk) & 1;
                const T side = no_other ? T(0) : (q ? other[f + 2] : other[f - 1]);
                const V2 ps = no_other ? zero : *reinterpret_cast<const V2 *>(other + f - g.H);
                const V2 pn = no_other ? zero : *reinterpret_cast<const V2 *>(other + f + g.H);
                const V2 po = (FIRST && !APPLY) ? zero : *reinterpret_cast<const V2 *>(mine + f);
                if (APPLY) {
                    if (ok0) acc[f] += (double)po.x;
                    if (ok1) acc[f + 1] += (double)po.y;
                }
                const V2 pold = FIRST ? zero : po;
                const T a1z = azc[kk], a2z = azc[g.sz + kk], a3z = azc[2 * g.sz + kk];
                const T a1x0 = q ? c1.y : c1.x, a2x0 = q ? c2.y : c2.x, a3x0 = q ? c3.y : c3.x;
                const T a1x1 = q ? c1.w : c1.z, a2x1 = q ? c2.w : c2.z, a3x1 = q ? c3.w : c3.z;
                V2 out = po;
                if (ok0) {
                    const T pw = q ? cur.x : side, pe = q ? cur.y : cur.x;
                    const T gs = sor_div(((((((r2.x - a1x0 * pw) - a3x0 * pe) - a1y * ps.x) - a3y * pn.x) -
                                           a1z * below.x) - a3z * above.x),
                                         ((a2x0 + a2y) + a2z));
                    out.x = omm * pold.x + omega * gs;
                    if (TRACK) d = fmax(d, (double)fabs(out.x - pold.x));
                }
                if (ok1) {
                    const T pw = q ? cur.y : cur.x, pe = q ? side : cur.y;
                    const T gs = sor_div(((((((r2.y - a1x1 * pw) - a3x1 * pe) - a1y * ps.y) - a3y * pn.y) -
                                           a1z * below.y) - a3z * above.y),
                                         ((a2x1 + a2y) + a2z));
                    out.y = omm * pold.y + omega * gs;
                    if (TRACK) d = fmax(d, (double)fabs(out.y - pold.y));
                }
                *reinterpret_cast<V2 *>(mine + f) = out;
            }
            below = cur;
            cur = above;
        }
    }
    if (TRACK) block_max_to(d, slot);
}

// One whole FP32 red-black sweep in one launch (the finest multigrid level).
// Red is updated from the old black values, black from the new red values,
// exactly as two half-sweeps in place, but the result goes to a second
// buffer: red_new/black_new are written to dst while src keeps the old values,
// so no workgroup ever reads a value another one is writing.
//
// A workgroup owns S consecutive slots of a plane (a colour row is H slots)
// and a chunk of planes [k0, k1).  It marches in z, one plane kk at a time: it
// updates red at kk over its slots plus H halo slots on either side (the y
// neighbours of its black cells), then black at the plane it left behind,
// whose red neighbours are now known.  Red of the two planes just outside the
// chunk is recomputed as a halo and not stored.  Even chunks march up and odd
// chunks down, so two neighbouring chunks reach their common boundary at the
// same time (both at the start or both at the end of their march): the halo
// planes one recomputes are the planes the other is reading then, and its
// inputs come from L2 instead of HBM.  Red values of the current plane go
// through a ring of three LDS planes, so one barrier per plane separates
// their writes from the reads of two planes later.  A thread owns one
// aligned pair of slots.  FIRST: the sweep starts from zero (no old values
// are read).  Every operation matches k_sor_pz, operand for operand, so the
// result is bit for bit that of the two half-sweeps.
static constexpr int kSweepThreads = 1024;
template <bool FIRST>
__global__ void __launch_bounds__(kSweepThreads) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_sweep2_pz(const float *__restrict__ src0, const float *__restrict__ src1, float *__restrict__ dst0,
            float *__restrict__ dst1, const float *__restrict__ rhs0, const float *__restrict__ rhs1,
            const float *__restrict__ axc, const float *__restrict__ ayc, const float *__restrict__ azc, SorGeom g,
            int S, int tz, float omega, float omm, const double *gate, double thr) {
    typedef f2v V2;
    __shared__ float ring[3][2 * kSweepThreads];
    __shared__ float xt[3][2 * kSweepThreads];
    if (gated(gate, thr)) return;
    // x coefficients of the row (2H+2 per array, fewer than 2 x threads
    // since the red range of S+2H slots fits the block): read per update
    // from LDS instead of being held in registers
    for (int i = (int)threadIdx.x; i < 2 * g.H + 2; i += kSweepThreads) {
        xt[0][i] = axc[i];
        xt[1][i] = axc[g.sxp + i];
        xt[2][i] = axc[2 * g.sxp + i];
    }
    __syncthreads();
    int bx, by;
    xcd_block2(bx, by);
    const int lo = 2 * (int)threadIdx.x;         // pair offset inside the red range
    const int R0 = bx * S - g.H;                 // first slot of the red range
    const int s = R0 + lo;                       // my slot pair (s, s+1)
    const bool in_plane = s >= 0 && s + 1 < g.plane && lo < S + 2 * g.H;
    const int jj = in_plane ? (int)fdiv((uint32_t)s, g.dH) : 0;
    const int h = s - jj * g.H;
    const bool row_ok = in_plane && jj >= 1 && jj <= g.sy - 2;
    const bool own = row_ok && s >= bx * S && s < bx * S + S;
    const int k0 = 1 + by * tz;
    const int k1 = min(k0 + tz, g.sz - 1);
    const int np = k1 - k0 + 2;
    const V2 zero = V2(0.0f);
    float a1y = 0.0f, a2y = 0.0f, a3y = 0.0f;
    if (row_ok) {
        a1y = ayc[jj]; a2y = ayc[g.sy + jj]; a3y = ayc[2 * g.sy + jj];
    }
    // march direction d (a compile-time constant in each copy of the body);
    // red planes k0-1 .. k1 in order, starting at ks
    auto march = [&](auto dir) {
        constexpr bool up = decltype(dir)::value;
        constexpr int d = up ? 1 : -1;
        const int ks = up ? k0 - 1 : k1;
        // black old at the plane behind kk (kk-d) and at kk; red new at kk-2d,
        // kk-d and kk
        V2 bo_b = zero, bo_c = zero, rn_b2 = zero, rn_b1 = zero;
        if (!FIRST && row_ok) {
            if (ks - d >= 0 && ks - d <= g.sz - 1) bo_b = *reinterpret_cast<const V2 *>(src1 + s + (ks - d) * g.plane);
            bo_c = *reinterpret_cast<const V2 *>(src1 + s + ks * g.plane);
        }
        // Global inputs of the next plane are loaded one plane ahead, after this
        // plane's red update and before the barrier, so their latency overlaps
        // the barrier and the black update: the red inputs (rhs, old red, old
        // black side/south/north), old black one plane further ahead and the
        // black rhs.  The loads are unconditional, from addresses clamped into
        // the arrays (a lane outside the rows, or a plane outside the interior,
        // reads a valid cell whose value the guards below never use), so no load
        // sits behind a branch and the counters of outstanding loads stay exact.
        // Two register sets alternate (the loop is unrolled by two), so the
        // loaded registers are never copied and no plane waits for the loads it
        // issued itself.  The z coefficients are uniform: scalar loads.
        typedef __attribute__((address_space(4))) float CFloat;
        const CFloat *az = (const CFloat *)azc;
        const int s_ld = row_ok ? s : g.H;
        struct In {
            V2 r2, po, ps, pn, bo_a, rb;
            float side;
        };
        In A{zero, zero, zero, zero, zero, zero, 0.0f}, B = A;
        // 32-bit byte offsets from the uniform array bases (every colour array
        // holds fewer than 2^30 slots), so each load takes one offset register
        auto at2 = [](const float *base, uint32_t i) {
            return *reinterpret_cast<const V2 *>(reinterpret_cast<const char *>(base) + i * 4u);
        };
        auto at1 = [](const float *base, uint32_t i) {
            return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + i * 4u);
        };
        auto put2 = [](float *base, uint32_t i, V2 v) {
            *reinterpret_cast<V2 *>(reinterpret_cast<char *>(base) + i * 4u) = v;
        };
        auto plane_of = [&](int k) { return (uint32_t)(s_ld + min(max(k, 0), g.sz - 1) * g.plane); };
        // inputs of red at plane kr: old black ahead of it (kr+d), and the rhs of
        // the black plane updated after it (kr-d)
        auto load_next = [&](int kr, In &n) {
            const uint32_t fr = (uint32_t)(s_ld + min(max(kr, 1), g.sz - 2) * g.plane);
            n.r2 = at2(rhs0, fr);
            if (!FIRST) {
                n.po = at2(src0, fr);
                n.side = at1(src1, ((jj + kr) & 1) ? fr + 2u : fr - 1u);
                n.ps = at2(src1, fr - (uint32_t)g.H);
                n.pn = at2(src1, fr + (uint32_t)g.H);
                n.bo_a = at2(src1, plane_of(kr + d));
            }
            n.rb = at2(rhs1, plane_of(kr - d));
        };
        auto plane = [&](int t, const In &c, In &n, bool next) {
            const int kk = ks + d * t;
            const int f = s + kk * g.plane;
            // ---- red at plane kk
            V2 rn_c = zero;
            if (row_ok && kk >= 1 && kk <= g.sz - 2) {
                rn_c = c.po;
                const bool ok0 = c.r2.x == c.r2.x, ok1 = c.r2.y == c.r2.y;
                if (ok0 || ok1) {
                    const int q = (jj + kk) & 1;                  // colour 0
                    const V2 bo_dn = up ? bo_b : c.bo_a, bo_up = up ? c.bo_a : bo_b;
                    const float a1z = az[kk], a2z = az[g.sz + kk], a3z = az[2 * g.sz + kk];
                    const int xo = 2 * h + q;
                    const float a1x0 = xt[0][xo], a2x0 = xt[1][xo], a3x0 = xt[2][xo];
                    const float a1x1 = xt[0][xo + 2], a2x1 = xt[1][xo + 2], a3x1 = xt[2][xo + 2];
                    if (ok0) {
                        const float pw = q ? bo_c.x : c.side, pe = q ? bo_c.y : bo_c.x;
                        const float gs = sor_div(((((((c.r2.x - a1x0 * pw) - a3x0 * pe) - a1y * c.ps.x) - a3y * c.pn.x) -
                                                   a1z * bo_dn.x) - a3z * bo_up.x),
                                                 ((a2x0 + a2y) + a2z));
                        rn_c.x = omm * c.po.x + omega * gs;
                    }
                    if (ok1) {
                        const float pw = q ? bo_c.y : bo_c.x, pe = q ? c.side : bo_c.y;
                        const float gs = sor_div(((((((c.r2.y - a1x1 * pw) - a3x1 * pe) - a1y * c.ps.y) - a3y * c.pn.y) -
                                                   a1z * bo_dn.y) - a3z * bo_up.y),
                                                 ((a2x1 + a2y) + a2z));
                        rn_c.y = omm * c.po.y + omega * gs;
                    }
                }
                if (own && kk >= k0 && kk < k1) put2(dst0, (uint32_t)f, rn_c);
            }
            if (next) load_next(kk + d, n);
            float *cur = ring[t % 3];
            cur[lo] = rn_c.x;
            cur[lo + 1] = rn_c.y;
            __syncthreads();
            // ---- black at the plane behind, kb = kk-d
            const int kb = kk - d;
            if (own && t >= 1 && kb >= k0 && kb < k1) {
                const int fb = s + kb * g.plane;
                const float *prev = ring[(t + 2) % 3];
                const V2 po_b = FIRST ? zero : bo_b;
                V2 out = po_b;
                const bool ok0 = c.rb.x == c.rb.x, ok1 = c.rb.y == c.rb.y;
                if (ok0 || ok1) {
                    const int q = (1 + jj + kb) & 1;              // colour 1
                    const V2 rn_dn = up ? rn_b2 : rn_c, rn_up = up ? rn_c : rn_b2;
                    const float sideb = q ? prev[lo + 2] : prev[lo - 1];
                    const V2 cr = V2{prev[lo], prev[lo + 1]};
                    const V2 psb = V2{prev[lo - g.H], prev[lo - g.H + 1]};
                    const V2 pnb = V2{prev[lo + g.H], prev[lo + g.H + 1]};
                    const float a1z = az[kb], a2z = az[g.sz + kb], a3z = az[2 * g.sz + kb];
                    const int xo = 2 * h + q;
                    const float a1x0 = xt[0][xo], a2x0 = xt[1][xo], a3x0 = xt[2][xo];
                    const float a1x1 = xt[0][xo + 2], a2x1 = xt[1][xo + 2], a3x1 = xt[2][xo + 2];
                    if (ok0) {
                        const float pw = q ? cr.x : sideb, pe = q ? cr.y : cr.x;
                        const float gs = sor_div(((((((c.rb.x - a1x0 * pw) - a3x0 * pe) - a1y * psb.x) - a3y * pnb.x) -
                                                   a1z * rn_dn.x) - a3z * rn_up.x),
                                                 ((a2x0 + a2y) + a2z));
                        out.x = omm * po_b.x + omega * gs;
                    }
                    if (ok1) {
                        const float pw = q ? cr.y : cr.x, pe = q ? sideb : cr.y;
                        const float gs = sor_div(((((((c.rb.y - a1x1 * pw) - a3x1 * pe) - a1y * psb.y) - a3y * pnb.y) -
                                                   a1z * rn_dn.y) - a3z * rn_up.y),
                                                 ((a2x1 + a2y) + a2z));
                        out.y = omm * po_b.y + omega * gs;
                    }
                }
                put2(dst1, (uint32_t)fb, out);
            }
            bo_b = bo_c;
            bo_c = c.bo_a;
            rn_b2 = rn_b1;
            rn_b1 = rn_c;
        };
        // every plane but the last loads the next one
        load_next(ks, A);
        int t = 0;
        for (; t + 2 < np; t += 2) {
            plane(t, A, B, true);
            plane(t + 1, B, A, true);
        }
        if (t + 1 < np) {
            plane(t, A, B, true);
            plane(t + 1, B, A, false);
        } else {
            plane(t, A, B, false);
        }
    };
    if ((by & 1) == 0)
        march(std::true_type{});
    else
        march(std::false_type{});
}

// Several red-black sweeps of a coarse multigrid level in one launch
// (temporal blocking).  A workgroup owns a CX x CY x CZ block of cells (8 x 8
// x 8 as launched) and loads it into LDS with a halo of 2M cells on every
// side (M sweeps: each half-sweep moves the dependency one cell), in natural
// (i, j, k) order, both colours interleaved.  Half-sweep m (1..2M) updates
// the cells of its colour that lie at most 2M-m cells outside the block, so
// the values it reads were exact after half-sweep m-1 and the block is exact
// after the last one.  The block's
// interior cells (both colours: updated cells and, copied through, the cells
// outside the solve) go to dst, never to src, which the neighbouring
// workgroups are still reading; walls are never written (they stay as in
// src).  Positions outside the arrays hold 0 and a NaN right-hand side, so
// they are never updated and only wall cells read them.  FIRST: the first
// red half-sweep takes every neighbour and the old value as 0, and the first
// black half-sweep the old value, as k_sor does from zero.  Every update is
// k_sor's, operand for operand: the result is bit for bit that of 2M
// half-sweep launches.
static constexpr int kTileCube = 8;       // the cube of levels 2 and below
static constexpr int kTileMaxSweeps = 2;
static constexpr int kTileThreads = 1024;
template <bool FIRST, int M, int CX, int CY, int CZ>
__global__ void __launch_bounds__(kTileThreads)
k_smooth_tile(const float *__restrict__ src, float *__restrict__ dst, const float *__restrict__ b,
              const float *__restrict__ axc, const float *__restrict__ ayc, const float *__restrict__ azc, SorGeom g,
              float omega, float omm, const double *gate, double thr) {
    static_assert(M >= 1 && M <= kTileMaxSweeps, "sweeps per launch");
    constexpr int R = 2 * M, EX = CX + 2 * R, EY = CY + 2 * R, EZ = CZ + 2 * R;
    constexpr int E2 = EX * EY, E3 = E2 * EZ;
    __shared__ float xs[E3];
    __shared__ float bs[E3];
    __shared__ float cx[3][EX], cy[3][EY], cz[3][EZ];
    if (gated(gate, thr)) return;
    // cuboid origin (first interior cell) minus the halo: tile position 0
    const int i0 = 1 + CX * (int)blockIdx.x - R;
    const int j0 = 1 + CY * (int)blockIdx.y - R;
    const int k0 = 1 + CZ * (int)blockIdx.z - R;
    const int tid = (int)threadIdx.x;
    const float nan = __builtin_nanf("");
    // coefficient rows of the tile's columns (clamped: positions outside
    // the arrays are never updated, so their rows are never used)
    const int ax_ = min(tid / EX, 2), tx_ = tid - ax_ * EX;
    const int ay_ = min(tid / EY, 2), ty_ = tid - ay_ * EY;
    const int az_ = min(tid / EZ, 2), tz_ = tid - az_ * EZ;
    float cxv = 0.0f, cyv = 0.0f, czv = 0.0f;
    if (tid < 3 * EX) cxv = axc[ax_ * g.sxp + min(max(i0 + tx_, 0), g.sx - 1)];
    if (tid < 3 * EY) cyv = ayc[ay_ * g.sy + min(max(j0 + ty_, 0), g.sy - 1)];
    if (tid < 3 * EZ) czv = azc[az_ * g.sz + min(max(k0 + tz_, 0), g.sz - 1)];
    // every thread issues all of its loads (at most kTileLoads cells each)
    // before it stores any, so their latencies overlap; positions outside the
    // arrays load a clamped (valid) cell and store 0 and NaN instead
    constexpr int kTileLoads = (E3 + kTileThreads - 1) / kTileThreads;
    float xv[kTileLoads], bv[kTileLoads];
    bool in[kTileLoads];
#pragma unroll
    for (int q = 0; q < kTileLoads; q++) {
        const int n = min(tid + q * kTileThreads, E3 - 1);
        const int lk = n / E2, r = n - lk * E2, lj = r / EX, li = r - lj * EX;
        const int ii = i0 + li, jj = j0 + lj, kk = k0 + lk;
        in[q] = ii >= 0 && ii < g.sx && jj >= 0 && jj < g.sy && kk >= 0 && kk < g.sz;
        const int ic = min(max(ii, 0), g.sx - 1), jc = min(max(jj, 0), g.sy - 1), kc = min(max(kk, 0), g.sz - 1);
        const uint32_t f = (uint32_t)((kc * g.sy + jc) * g.H + (ic >> 1) + ((ic + jc + kc) & 1) * g.NC);
        bv[q] = *reinterpret_cast<const float *>(reinterpret_cast<const char *>(b) + f * 4u);
        xv[q] = *reinterpret_cast<const float *>(reinterpret_cast<const char *>(src) + f * 4u);
    }
#pragma unroll
    for (int q = 0; q < kTileLoads; q++) {
        const int n = tid + q * kTileThreads;
        if (n < E3) {
            const float bq = in[q] ? bv[q] : nan;
            xs[n] = in[q] ? ((FIRST && bq == bq) ? 0.0f : xv[q]) : 0.0f;
            bs[n] = bq;
        }
    }
    if (tid < 3 * EX) cx[ax_][tx_] = cxv;
    if (tid < 3 * EY) cy[ay_][ty_] = cyv;
    if (tid < 3 * EZ) cz[az_][tz_] = czv;
    __syncthreads();
#pragma unroll
    for (int m = 1; m <= 2 * M; m++) {
        const int col = (m - 1) & 1;
        // region [m, E-m) per axis; cells of one colour per row <= HW
        const int WX = EX - 2 * m, WY = EY - 2 * m, WZ = EZ - 2 * m, HW = (WX + 1) >> 1;
#pragma unroll
        for (int idx = tid; idx < HW * WY * WZ; idx += kTileThreads) {
            const int row = idx / HW, p = idx - row * HW;
            const int lk = m + row / WY, lj = m + (row - (row / WY) * WY);
            const int li = m + (((col - (i0 + m + j0 + lj + k0 + lk)) & 1)) + 2 * p;
            if (li >= EX - m) continue;
            const int n = li + EX * (lj + EY * lk);
            const float r = bs[n];
            if (!(r == r)) continue;
            float pw, pe, ps, pn, pb, pt;
            if (FIRST && m == 1) {
                pw = pe = ps = pn = pb = pt = 0.0f;
            } else {
                pw = xs[n - 1]; pe = xs[n + 1];
                ps = xs[n - EX]; pn = xs[n + EX];
                pb = xs[n - E2]; pt = xs[n + E2];
            }
            const float pold = xs[n];
            const float a1x = cx[0][li], a2x = cx[1][li], a3x = cx[2][li];
            const float a1y = cy[0][lj], a2y = cy[1][lj], a3y = cy[2][lj];
            const float a1z = cz[0][lk], a2z = cz[1][lk], a3z = cz[2][lk];
            const float gs = sor_div(((((((r - a1x * pw) - a3x * pe) - a1y * ps) - a3y * pn) - a1z * pb) - a3z * pt),
                                     ((a2x + a2y) + a2z));
            xs[n] = omm * pold + omega * gs;
        }
        __syncthreads();
    }
    for (int n = tid; n < CX * CY * CZ; n += kTileThreads) {
        const int ck = n / (CX * CY), r = n - ck * CX * CY, cj = r / CX, ci = r - cj * CX;
        const int li = R + ci, lj = R + cj, lk = R + ck;
        const int ii = i0 + li, jj = j0 + lj, kk = k0 + lk;
        if (ii > g.sx - 2 || jj > g.sy - 2 || kk > g.sz - 2) continue;
        const int f = (kk * g.sy + jj) * g.H + (ii >> 1) + ((ii + jj + kk) & 1) * g.NC;
        dst[f] = xs[li + EX * (lj + EY * lk)];
    }
}

// FP64 residual r = b/dt - A p over both colours.  ADD: of p + e (e is the
// FP32 defect-correction increment not yet added).  Writes r as FP32 (NaN
// sentinel elsewhere) when r0 != nullptr, and per-block partial sums of r^2.
template <bool ADD>
__global__ void __launch_bounds__(kSorBlock)
k_resid(const double *__restrict__ p0, const double *__restrict__ p1, const float *__restrict__ e0,
        const float *__restrict__ e1, const double *__restrict__ b0, const double *__restrict__ b1,
        float *__restrict__ r0, float *__restrict__ r1, const double *__restrict__ axc,
        const double *__restrict__ ayc, const double *__restrict__ azc, SorGeom g, double *partial,
        const double *gate, double thr) {
    if (gated(gate, thr)) return;
    const int lb = xcd_block();
    const int f2 = lb * kSorBlock + (int)threadIdx.x;  // over 2*NC
    double rr = 0.0;
    if (f2 < 2 * g.NC) {
        const int colour = f2 >= g.NC;
        const int f = colour ? f2 - g.NC : f2;
        const double *pm = colour ? p1 : p0;
        const double *po = colour ? p0 : p1;
        const float *em = colour ? e1 : e0;
        const float *eo = colour ? e0 : e1;
        const double bd = (colour ? b1 : b0)[f];
        float *rout = r0 ? (colour ? r1 : r0) : nullptr;
        int ii, jj, kk;
        decode(g, f, colour, ii, jj, kk);
        // live cells carry a finite b/dt; the explicit interior test keeps
        // every neighbour access inside the arrays whatever the sentinel holds
        if (bd == bd && interior_cell(g, ii, jj, kk)) {
            const int q = (colour + jj + kk) & 1;
            const int sx = g.sxp, sy = g.sy, nzp = g.sz;
            double pw = po[f - 1 + q], pe = po[f + q], ps = po[f - g.H], pn = po[f + g.H];
            double pb = po[f - g.plane], pt = po[f + g.plane], ps0 = pm[f];
            if (ADD) {
                pw += (double)eo[f - 1 + q];
                pe += (double)eo[f + q];
                ps += (double)eo[f - g.H];
                pn += (double)eo[f + g.H];
                pb += (double)eo[f - g.plane];
                pt += (double)eo[f + g.plane];
                ps0 += (double)em[f];
            }
            const double ap = (((((axc[ii] * pw + axc[2 * sx + ii] * pe) + ayc[jj] * ps) + ayc[2 * sy + jj] * pn) +
                                azc[kk] * pb) + azc[2 * nzp + kk] * pt) +
                              ((axc[sx + ii] + ayc[sy + jj]) + azc[nzp + kk]) * ps0;
            const double r = bd - ap;
            rr = r * r;
            if (rout) rout[f] = (float)r;
        } else if (rout) {
            rout[f] = __builtin_nanf("");
        }
    }
    block_sum_to(rr, partial, lb);
}

// FP64 residual on cell pairs marching in z, both colours per thread (the
// same neighbour pattern as k_sor_pz).  ADD: of p + e.  WRITE_R: the residual
// as FP32 (NaN where b/dt is NaN) for the next inner solve.  One partial sum
// of r^2 per block.  Five waves per SIMD (a few registers spill; measured
// faster than four without spills).
template <bool ADD, bool WRITE_R>
__global__ void __launch_bounds__(kSorBlock, 5)
k_resid_pz(const double *__restrict__ p0, const double *__restrict__ p1, const float *__restrict__ e0,
           const float *__restrict__ e1, const double *__restrict__ b0, const double *__restrict__ b1,
           float *__restrict__ r0, float *__restrict__ r1, const double *__restrict__ axc,
           const double *__restrict__ ayc, const double *__restrict__ azc, SorGeom g, double *partial,
           const double *gate, double thr) {
    if (gated(gate, thr)) return;
    // the x coefficients of a row (3 x (2H+2) doubles, dynamic LDS) are read
    // per update from LDS rather than held in 24 registers for the march
    extern __shared__ double xtab[];
    const int nxt = 2 * g.H + 2;
    for (int i = (int)threadIdx.x; i < 3 * nxt; i += kSorBlock) {
        const int r = i / nxt;
        xtab[i] = axc[r * g.sxp + (i - r * nxt)];
    }
    __syncthreads();
    int bx, by;
    xcd_block2(bx, by);
    const int p = 2 * (int)(bx * kSorBlock + threadIdx.x);
    const int jj = (int)fdiv((uint32_t)p, g.dH);
    const int h = p - jj * g.H;
    double rr = 0.0;
    if (p < g.plane && jj >= 1 && jj <= g.sy - 2) {
        const int k0 = 1 + by * kZch;
        const int k1 = min(k0 + kZch, g.sz - 1);
        const double *x1t = xtab + 2 * h, *x2t = xtab + nxt + 2 * h, *x3t = xtab + 2 * nxt + 2 * h;
        const double a1y = ayc[jj], a2y = ayc[g.sy + jj], a3y = ayc[2 * g.sy + jj];
        const double *pc_[2] = {p0, p1};
        const float *ec_[2] = {e0, e1};
        const double *bc_[2] = {b0, b1};
        float *rc_[2] = {r0, r1};
        auto ld2 = [&](int c, int idx) -> d2v {
            d2v v = *reinterpret_cast<const d2v *>(pc_[c] + idx);
            if (ADD) {
                const f2v e = *reinterpret_cast<const f2v *>(ec_[c] + idx);
                v.x += (double)e.x;
                v.y += (double)e.y;
            }
            return v;
        };
        auto ld1 = [&](int c, int idx) -> double {
            double v = pc_[c][idx];
            if (ADD) v += (double)ec_[c][idx];
            return v;
        };
        int f = p + k0 * g.plane;
        d2v below[2] = {ld2(0, f - g.plane), ld2(1, f - g.plane)};
        d2v cur[2] = {ld2(0, f), ld2(1, f)};
        for (int kk = k0; kk < k1; kk++, f += g.plane) {
            d2v above[2] = {ld2(0, f + g.plane), ld2(1, f + g.plane)};
            const double a1z = azc[kk], a2z = azc[g.sz + kk], a3z = azc[2 * g.sz + kk];
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const int o = 1 - c;
                const d2v bd = *reinterpret_cast<const d2v *>(bc_[c] + f);
                const bool ok0 = bd.x == bd.x, ok1 = bd.y == bd.y;
                f2v rout = {__builtin_nanf(""), __builtin_nanf("")};
                if (ok0 || ok1) {
                    const int q = (c + jj + kk) & 1;
                    const double side = q ? ld1(o, f + 2) : ld1(o, f - 1);
                    const d2v ps = ld2(o, f - g.H), pn = ld2(o, f + g.H);
                    const d2v co = cur[o], own = cur[c];
                    const double a1x0 = x1t[q], a2x0 = x2t[q], a3x0 = x3t[q];
                    const double a1x1 = x1t[q + 2], a2x1 = x2t[q + 2], a3x1 = x3t[q + 2];
                    if (ok0) {
                        const double pw = q ? co.x : side, pe = q ? co.y : co.x;
                        const double ap = (((((a1x0 * pw + a3x0 * pe) + a1y * ps.x) + a3y * pn.x) +
                                            a1z * below[o].x) + a3z * above[o].x) +
                                          ((a2x0 + a2y) + a2z) * own.x;
                        const double r = bd.x - ap;
                        rr += r * r;
                        rout.x = (float)r;
                    }
                    if (ok1) {
                        const double pw = q ? co.y : co.x, pe = q ? side : co.y;
                        const double ap = (((((a1x1 * pw + a3x1 * pe) + a1y * ps.y) + a3y * pn.y) +
                                            a1z * below[o].y) + a3z * above[o].y) +
                                          ((a2x1 + a2y) + a2z) * own.y;
                        const double r = bd.y - ap;
                        rr += r * r;
                        rout.y = (float)r;
                    }
                }
                if (WRITE_R) *reinterpret_cast<f2v *>(rc_[c] + f) = rout;
            }
            below[0] = cur[0];
            below[1] = cur[1];
            cur[0] = above[0];
            cur[1] = above[1];
        }
    }
    block_sum_to(rr, partial, (int64_t)by * gridDim.x + bx);
}

// The cells the solve updates, one bit per slot: word w of colour c holds
// slots 32w..32w+31 (bit s & 31 of word s >> 5), bits[c * nw + w], nw even
// words per colour.  Built from the NaN sentinel of b/dt once per geometry;
// every vector of the solve is NaN on exactly the other slots, so a kernel
// that only tests for NaN can read the bits instead of a whole vector.
__global__ void __launch_bounds__(kSorBlock)
k_active_bits(const double *__restrict__ b0, const double *__restrict__ b1, int64_t NC, uint32_t *__restrict__ bits,
              int64_t nw) {
    const int64_t per = nw * 32;  // slots covered per colour (a multiple of 64)
    const int64_t q = (int64_t)blockIdx.x * kSorBlock + threadIdx.x;
    if (q >= 2 * per) return;      // whole waves: 2*per is a multiple of 64
    const int c = q >= per ? 1 : 0;
    const int64_t s = q - c * per;
    bool on = false;
    if (s < NC) {
        const double v = (c ? b1 : b0)[s];
        on = v == v;
    }
    const uint64_t m = __ballot(on);
    if ((threadIdx.x & 63) == 0) {
        bits[c * nw + (s >> 5)] = (uint32_t)m;
        bits[c * nw + (s >> 5) + 1] = (uint32_t)(m >> 32);
    }
}

// the two flags of the aligned slot pair (f, f+1) of colour c
__device__ __forceinline__ void active_pair(const uint32_t *bits, int64_t nw, int c, int f, bool &ok0, bool &ok1) {
    const uint32_t w = bits[c * nw + (f >> 5)];
    ok0 = (w >> (f & 31)) & 1u;
    ok1 = (w >> ((f & 31) + 1)) & 1u;
}

// Cells the solve does not update (b/dt is NaN there) must hold pc = 0 and
// e = 0: they are the Dirichlet values of their fluid neighbours.  A warm
// start keeps the previous solution, so cells that stopped being fluid
// (a moving body) are cleared before every solve.
__global__ void __launch_bounds__(kSorBlock)
k_clear_dead(double *p0, double *p1, float *e0, float *e1, const double *b0, const double *b1, int64_t NC) {
    const int64_t f2 = (int64_t)blockIdx.x * kSorBlock + threadIdx.x;
    if (f2 >= 2 * NC) return;
    const int colour = f2 >= NC;
    const int64_t f = colour ? f2 - NC : f2;
    const double bd = (colour ? b1 : b0)[f];
    if (bd != bd) {
        (colour ? p1 : p0)[f] = 0.0;
        (colour ? e1 : e0)[f] = 0.0f;
    }
}

// The same from the slot flags: one thread per 32-slot word, visiting only
// its clear bits (a few percent of the slots).  zero_slot (if given) is set
// to 0 as well.
__global__ void __launch_bounds__(kSorBlock)
k_clear_dead_bits(double *p0, double *p1, float *e0, float *e1, const uint32_t *__restrict__ bits, int64_t nw,
                  int64_t NC, double *zero_slot) {
    const int64_t q = (int64_t)blockIdx.x * kSorBlock + threadIdx.x;
    if (zero_slot && q == 0) *zero_slot = 0.0;
    if (q >= 2 * nw) return;
    const int colour = q >= nw ? 1 : 0;
    const int64_t base = (q - colour * nw) * 32;
    uint32_t dead = ~bits[q];
    while (dead) {
        const int b = __builtin_ctz(dead);
        dead &= dead - 1u;
        const int64_t f = base + b;
        if (f >= NC) break;  // bits run upward: the rest lie past the array too
        (colour ? p1 : p0)[f] = 0.0;
        (colour ? e1 : e0)[f] = 0.0f;
    }
}

// p <- p + e on both colours (fluid cells only)
__global__ void __launch_bounds__(kSorBlock)
k_add_e(double *p0, double *p1, const float *e0, const float *e1, const double *b0, const double *b1, int64_t NC) {
    const int64_t f2 = (int64_t)blockIdx.x * kSorBlock + threadIdx.x;
    if (f2 >= 2 * NC) return;
    const int colour = f2 >= NC;
    const int64_t f = colour ? f2 - NC : f2;
    const double bd = (colour ? b1 : b0)[f];
    if (bd == bd) (colour ? p1 : p0)[f] += (double)(colour ? e1 : e0)[f];
}

// subPcorVcor.f90:89-99 computeDiv, writing b/dt into the colour arrays
// (NaN elsewhere), per-block partial maxima of |b| (pmax) and partial sums
// of (b/dt)^2 (psum).  P < 2^31, so the index splits in 32 bits.
__global__ void __launch_bounds__(kSorBlock)
k_div(const double *__restrict__ ut, const double *__restrict__ vt, const double *__restrict__ wt,
      const int8_t *__restrict__ tag, const double *__restrict__ dx, const double *__restrict__ dy,
      const double *__restrict__ dz, double *__restrict__ b0, double *__restrict__ b1, int sx, int sy,
      int sz, int H, FastDiv dsx, FastDiv dsy, double deltat, double *pmax, double *psum) {
    const int P = sx * sy * sz;
    const int idx = (int)blockIdx.x * kSorBlock + (int)threadIdx.x;
    double am = 0.0, bb = 0.0;
    if (idx < P) {
        const int row = (int)fdiv((uint32_t)idx, dsx);
        const int ii = idx - row * sx;
        const int kk = (int)fdiv((uint32_t)row, dsy);
        const int jj = row - kk * sy;
        const int colour = (ii + jj + kk) & 1;
        const int64_t f = (int64_t)row * H + (ii >> 1);
        double val = __builtin_nan("");
        const bool interior = ii >= 1 && ii <= sx - 2 && jj >= 1 && jj <= sy - 2 && kk >= 1 && kk <= sz - 2;
        if (interior && tag[idx] == 0) {
            const int i = ii + 1, j = jj + 1, k = kk + 1;
            const double b = ((ut[idx] - ut[idx - 1]) / dx[i] + (vt[idx] - vt[idx - sx]) / dy[j]) +
                             (wt[idx] - wt[idx - (int64_t)sx * sy]) / dz[k];
            am = fabs(b);
            val = b / deltat;
            bb = val * val;
        }
        (colour ? b1 : b0)[f] = val;
    }
    block_max_partial(am, pmax);
    block_sum_to(bb, psum, blockIdx.x);
}

// plain streaming copy of known size: the reference point for converting
// hardware fetch/write counters into bytes
__global__ void __launch_bounds__(kSorBlock) k_copy_probe(const double4 *__restrict__ a, double4 *__restrict__ b,
                                                          int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * kSorBlock + threadIdx.x;
    if (q < n) b[q] = a[q];
}

// ------------------------------------------------------------ GMRES-IR
// FP32 flexible GMRES(m) inside the FP64 refinement loop: every outer
// iteration solves A d = r (r the FP64 residual, rounded to FP32) with m
// Arnoldi steps whose preconditioner is k FP32 red-black SOR sweeps from
// zero, then adds d to the FP64 pc.  Vectors are two colour arrays back to
// back (2*NC floats, NaN on cells the solve does not update); H, beta and y
// live on the device, so a whole cycle runs without the host.

// y = A z (FP32) on cell pairs marching in z, both colours per thread;
// NaN where the mask array is NaN.  RESID: y = m - A z instead (the mask
// array is then the right-hand side).  bits (not RESID): the slot flags of
// the mask array, read instead of it.
template <bool RESID>
__global__ void __launch_bounds__(kSorBlock)
k_matvec32_pz(const float *__restrict__ z0, const float *__restrict__ z1, const float *__restrict__ m0,
              const float *__restrict__ m1, float *__restrict__ y0, float *__restrict__ y1,
              const float *__restrict__ axc, const float *__restrict__ ayc, const float *__restrict__ azc, SorGeom g,
              const double *gate, double thr, const uint32_t *__restrict__ bits = nullptr, int64_t nw = 0) {
    if (gated(gate, thr)) return;
    int bx, by;
    xcd_block2(bx, by);
    const int p = 2 * (int)(bx * kSorBlock + threadIdx.x);
    const int jj = (int)fdiv((uint32_t)p, g.dH);
    const int h = p - jj * g.H;
    if (!(p < g.plane && jj >= 1 && jj <= g.sy - 2)) return;
    const int k0 = 1 + by * kZch;
    const int k1 = min(k0 + kZch, g.sz - 1);
    const f4v c1 = *reinterpret_cast<const f4v *>(axc + 2 * h);
    const f4v c2 = *reinterpret_cast<const f4v *>(axc + g.sxp + 2 * h);
    const f4v c3 = *reinterpret_cast<const f4v *>(axc + 2 * g.sxp + 2 * h);
    const float a1y = ayc[jj], a2y = ayc[g.sy + jj], a3y = ayc[2 * g.sy + jj];
    const float *zc_[2] = {z0, z1};
    const float *mc_[2] = {m0, m1};
    float *yc_[2] = {y0, y1};
    int f = p + k0 * g.plane;
    f2v below[2] = {*reinterpret_cast<const f2v *>(z0 + f - g.plane), *reinterpret_cast<const f2v *>(z1 + f - g.plane)};
    f2v cur[2] = {*reinterpret_cast<const f2v *>(z0 + f), *reinterpret_cast<const f2v *>(z1 + f)};
    for (int kk = k0; kk < k1; kk++, f += g.plane) {
        f2v above[2] = {*reinterpret_cast<const f2v *>(z0 + f + g.plane), *reinterpret_cast<const f2v *>(z1 + f + g.plane)};
        const float a1z = azc[kk], a2z = azc[g.sz + kk], a3z = azc[2 * g.sz + kk];
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const int o = 1 - c;
            f2v mk;
            bool ok0, ok1;
            if constexpr (RESID) {
                mk = *reinterpret_cast<const f2v *>(mc_[c] + f);
                ok0 = mk.x == mk.x;
                ok1 = mk.y == mk.y;
            } else {
                active_pair(bits, nw, c, f, ok0, ok1);
            }
            f2v out = {__builtin_nanf(""), __builtin_nanf("")};
            if (ok0 || ok1) {
                const int q = (c + jj + kk) & 1;
                const float side = q ? zc_[o][f + 2] : zc_[o][f - 1];
                const f2v ps = *reinterpret_cast<const f2v *>(zc_[o] + f - g.H);
                const f2v pn = *reinterpret_cast<const f2v *>(zc_[o] + f + g.H);
                const f2v co = cur[o], own = cur[c];
                const float a1x0 = q ? c1.y : c1.x, a2x0 = q ? c2.y : c2.x, a3x0 = q ? c3.y : c3.x;
                const float a1x1 = q ? c1.w : c1.z, a2x1 = q ? c2.w : c2.z, a3x1 = q ? c3.w : c3.z;
                if (ok0) {
                    const float pw = q ? co.x : side, pe = q ? co.y : co.x;
                    out.x = (((((a1x0 * pw + a3x0 * pe) + a1y * ps.x) + a3y * pn.x) + a1z * below[o].x) +
                             a3z * above[o].x) + ((a2x0 + a2y) + a2z) * own.x;
                    if (RESID) out.x = mk.x - out.x;
                }
                if (ok1) {
                    const float pw = q ? co.y : co.x, pe = q ? side : co.y;
                    out.y = (((((a1x1 * pw + a3x1 * pe) + a1y * ps.y) + a3y * pn.y) + a1z * below[o].y) +
                             a3z * above[o].y) + ((a2x1 + a2y) + a2z) * own.y;
                    if (RESID) out.y = mk.y - out.y;
                }
            }
            *reinterpret_cast<f2v *>(yc_[c] + f) = out;
        }
        below[0] = cur[0];
        below[1] = cur[1];
        cur[0] = above[0];
        cur[1] = above[1];
    }
}

static constexpr int kMaxKrylov = 24;
static constexpr int kVecBlocks = 2048;

// dot products of w with v_0..v_{NV-1} (vectors of n4 float4 groups, NaN
// entries of w skipped), per-block partials [i][block]; NV is a template
// parameter so the accumulators and loads unroll exactly
template <int NV>
__global__ void __launch_bounds__(kSorBlock)
k_dots(const float *__restrict__ w, const float *__restrict__ V, int64_t stride, int64_t n4,
       double *__restrict__ partial, const double *gate, double thr) {
    if (gated(gate, thr)) return;
    __shared__ double sm[NV][kSorBlock / 64];
    double s[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) s[i] = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * kSorBlock + threadIdx.x; q < n4; q += (int64_t)gridDim.x * kSorBlock) {
        const f4v a = reinterpret_cast<const f4v *>(w)[q];
        const bool ok0 = a.x == a.x, ok1 = a.y == a.y, ok2 = a.z == a.z, ok3 = a.w == a.w;
#pragma unroll
        for (int i = 0; i < NV; i++) {
            const f4v b = reinterpret_cast<const f4v *>(V + i * stride)[q];
            if (ok0) s[i] += (double)a.x * (double)b.x;
            if (ok1) s[i] += (double)a.y * (double)b.y;
            if (ok2) s[i] += (double)a.z * (double)b.z;
            if (ok3) s[i] += (double)a.w * (double)b.w;
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++) {
        const double v = wave_sum(s[i]);
        if (lane == 0) sm[i][wid] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < NV)
        partial[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] =
            ((sm[threadIdx.x][0] + sm[threadIdx.x][1]) + sm[threadIdx.x][2]) + sm[threadIdx.x][3];
}

// block i: out[i] (+)= sum of partial[i][0..nblk)
__global__ void k_rows_sum(const double *__restrict__ partial, int nblk, double *out, int accumulate,
                           const double *gate, double thr) {
    if (gated(gate, thr)) return;
    __shared__ double sm[256];
    double s = 0.0;
    for (int q = threadIdx.x; q < nblk; q += 256) s += partial[(int64_t)blockIdx.x * nblk + q];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = accumulate ? out[blockIdx.x] + sm[0] : sm[0];
}

// w -= sum_i h[i] v_i (FP64 accumulation), then per-block partial sums of
// w^2 (NaN skipped); float4 groups, NV unrolled
template <int NV>
__global__ void __launch_bounds__(kSorBlock)
k_orth(float *__restrict__ w, const float *__restrict__ V, int64_t stride, const double *__restrict__ h, int64_t n4,
       double *__restrict__ partial, const double *gate, double thr) {
    if (gated(gate, thr)) return;
    double hv[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) hv[i] = h[i];
    double s = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * kSorBlock + threadIdx.x; q < n4; q += (int64_t)gridDim.x * kSorBlock) {
        f4v a = reinterpret_cast<f4v *>(w)[q];
        double ax = a.x, ay = a.y, az = a.z, aw = a.w;
#pragma unroll
        for (int i = 0; i < NV; i++) {
            const f4v b = reinterpret_cast<const f4v *>(V + i * stride)[q];
            ax -= hv[i] * (double)b.x;
            ay -= hv[i] * (double)b.y;
            az -= hv[i] * (double)b.z;
            aw -= hv[i] * (double)b.w;
        }
        a.x = (float)ax;
        a.y = (float)ay;
        a.z = (float)az;
        a.w = (float)aw;
        reinterpret_cast<f4v *>(w)[q] = a;
        if (a.x == a.x) s += (double)a.x * (double)a.x;
        if (a.y == a.y) s += (double)a.y * (double)a.y;
        if (a.z == a.z) s += (double)a.z * (double)a.z;
        if (a.w == a.w) s += (double)a.w * (double)a.w;
    }
    block_sum_to(s, partial, blockIdx.x);
}

// launch k_dots<nv> / k_orth<nv> for nv = 1..kMaxKrylov-1
#define IB3D_GS_CASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) \
    X(16) X(17) X(18) X(19) X(20) X(21) X(22) X(23)
static void launch_dots(int nv, dim3 g, dim3 b, hipStream_t st, const float *w, const float *V, int64_t stride,
                        int64_t n4, double *partial, const double *gate) {
    switch (nv) {
#define IB3D_DOTS(N) \
    case N: hipLaunchKernelGGL(k_dots<N>, g, b, 0, st, w, V, stride, n4, partial, gate, 0.5); break;
        IB3D_GS_CASES(IB3D_DOTS)
#undef IB3D_DOTS
        default: break;
    }
}
static void launch_orth(int nv, dim3 g, dim3 b, hipStream_t st, float *w, const float *V, int64_t stride,
                        const double *h, int64_t n4, double *partial, const double *gate) {
    switch (nv) {
#define IB3D_ORTH(N) \
    case N: hipLaunchKernelGGL(k_orth<N>, g, b, 0, st, w, V, stride, h, n4, partial, gate, 0.5); break;
        IB3D_GS_CASES(IB3D_ORTH)
#undef IB3D_ORTH
        default: break;
    }
}
#undef IB3D_GS_CASES

// partial sums of x^2 (NaN skipped)
__global__ void __launch_bounds__(kSorBlock)
k_norm2(const float *__restrict__ x, int64_t n2, double *__restrict__ partial, const double *gate, double thr) {
    if (gated(gate, thr)) return;
    double s = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * kSorBlock + threadIdx.x; q < n2; q += (int64_t)gridDim.x * kSorBlock) {
        const f2v a = reinterpret_cast<const f2v *>(x)[q];
        if (a.x == a.x) s += (double)a.x * (double)a.x;
        if (a.y == a.y) s += (double)a.y * (double)a.y;
    }
    block_sum_to(s, partial, blockIdx.x);
}

// *out = sqrt(sum of partials)
__global__ void k_sqrt_sum(const double *__restrict__ partial, int nblk, double *out, const double *gate,
                           double thr) {
    if (gated(gate, thr)) return;
    __shared__ double sm[256];
    double s = 0.0;
    for (int q = threadIdx.x; q < nblk; q += 256) s += partial[q];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sqrt(sm[0]);
}

// dst = src / *den (NaN kept); a zero or non-finite norm leaves zeros
__global__ void __launch_bounds__(kSorBlock)
k_scale_to(float *__restrict__ dst, const float *__restrict__ src, const double *den, int64_t n2,
           const double *gate, double thr) {
    if (gated(gate, thr)) return;
    const double d = *den;
    const float inv = (d > 0.0 && d < 1e300) ? (float)(1.0 / d) : 0.0f;
    for (int64_t q = (int64_t)blockIdx.x * kSorBlock + threadIdx.x; q < n2; q += (int64_t)gridDim.x * kSorBlock) {
        f2v a = reinterpret_cast<const f2v *>(src)[q];
        a.x = a.x == a.x ? a.x * inv : a.x;
        a.y = a.y == a.y ? a.y * inv : a.y;
        reinterpret_cast<f2v *>(dst)[q] = a;
    }
}

// Arnoldi bookkeeping on the device.  act[j] >= 0.5 while Arnoldi step j
// is to run: act[0] follows the outer gate, and after step j k_givens turns
// the new column of H into a column of R (the rotations so far plus a new
// one, Saad's GMRES with Givens rotations), rotates the right-hand side g and
// clears act[j+1] once |g[j+1]| (the residual norm of the least-squares
// solution after j+1 steps) is at most `target`; the remaining steps of the
// cycle then exit at once.
__global__ void k_act_init(double *act, int m, const double *gate, double thr) {
    const int t = threadIdx.x;
    if (t <= m) act[t] = (t == 0 && !gated(gate, thr)) ? 1.0 : 0.0;
}

// The step's new norm h[j+1] = ||w|| comes first, from the orthogonalisation's
// nblk block partials of w^2: 256 threads, each summing partial[t],
// partial[t+256], ... in order, then a pairwise tree, then the square root.
// It is also kept in nrm[j] for the scaling of the new basis vector, which
// runs after the rotation has cleared h[j+1].
__global__ void __launch_bounds__(256)
k_givens(double *__restrict__ H, int m, int j, double *cs, double *sn, double *gv, const double *beta,
         double *act, double target, const double *targetp, const double *__restrict__ partial, int nblk,
         double *nrm) {
    if (act[j] < 0.5) return;
    double *h = H + (int64_t)(m + 1) * j;
    {
        __shared__ double sm[256];
        double s = 0.0;
        for (int q = threadIdx.x; q < nblk; q += 256) s += partial[q];
        sm[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x != 0) return;
        h[j + 1] = sqrt(sm[0]);
        nrm[j] = h[j + 1];
    }
    if (targetp) target = *targetp;
    if (j == 0) gv[0] = *beta;
    for (int i = 0; i < j; i++) {
        const double t = cs[i] * h[i] + sn[i] * h[i + 1];
        h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
        h[i] = t;
    }
    const double a = h[j], b = h[j + 1];
    const double r = sqrt(a * a + b * b);
    const double c = r > 0.0 ? a / r : 1.0, s = r > 0.0 ? b / r : 0.0;
    cs[j] = c;
    sn[j] = s;
    h[j] = r;
    h[j + 1] = 0.0;
    gv[j + 1] = -s * gv[j];
    gv[j] = c * gv[j];
    act[j + 1] = (r > 0.0 && fabs(gv[j + 1]) > target) ? 1.0 : 0.0;
}

// y from R y = g over the steps that ran (act[j] >= 0.5); y = 0 beyond;
// *steps counts the steps run
__global__ void k_lsq(const double *__restrict__ H, int m, const double *gv, const double *act, double *y,
                      double *steps) {
    if (act[0] < 0.5 || threadIdx.x != 0) return;
    int k = 0;
    while (k < m && act[k] >= 0.5) k++;
    *steps += (double)k;
    for (int j = m - 1; j >= k; j--) y[j] = 0.0;
    for (int j = k - 1; j >= 0; j--) {
        double s = gv[j];
        for (int i = j + 1; i < k; i++) s -= H[j + (int64_t)(m + 1) * i] * y[i];
        const double d = H[j + (int64_t)(m + 1) * j];
        y[j] = d != 0.0 ? s / d : 0.0;
    }
}

// pc += sum_j y_j z_j on cells the solve updates (slot flag set); one pair of
// cells of one colour per thread (NC is even, so pairs never straddle the
// two colour arrays); y_j = 0 (steps that did not run) skips z_j
__global__ void __launch_bounds__(kSorBlock)
k_update_pc(double *__restrict__ p0, double *__restrict__ p1, const float *__restrict__ Z, int64_t stride, int m,
            const double *__restrict__ y, const uint32_t *__restrict__ bits, int64_t nw, int64_t NC,
            const double *gate, double thr) {
    if (gated(gate, thr)) return;
    int nz = 0;  // steps after the last non-zero y add nothing
    for (int j = 0; j < m; j++)
        if (y[j] != 0.0) nz = j + 1;
    const int64_t npair = NC;  // 2*NC cells = NC pairs
    for (int64_t q2 = (int64_t)blockIdx.x * kSorBlock + threadIdx.x; q2 < npair;
         q2 += (int64_t)gridDim.x * kSorBlock) {
        const int64_t q = 2 * q2;
        const int c = q >= NC ? 1 : 0;
        bool ok0, ok1;
        active_pair(bits, nw, c, (int)(q - c * NC), ok0, ok1);
        if (!(ok0 || ok1)) continue;
        double d0 = 0.0, d1 = 0.0;
        for (int j = 0; j < nz; j++) {
            const double yj = y[j];
            if (yj == 0.0) continue;
            const f2v z = *reinterpret_cast<const f2v *>(Z + j * stride + q);
            d0 += yj * (double)z.x;
            d1 += yj * (double)z.y;
        }
        double *p = q < NC ? p0 + q : p1 + (q - NC);
        if (ok0) p[0] += d0;
        if (ok1) p[1] += d1;
    }
}

// ---------------------------------------------------------------- host side
static SorGeom geom(const Ctx &c) {
    SorGeom g;
    g.H = (int)c.H;
    g.NC = (int)c.NC;
    g.sy = (int)c.sy;
    g.sx = (int)c.sx;
    g.sz = (int)c.sz;
    g.plane = (int)(c.H * c.sy);
    g.sxp = (int)c.sxp;
    g.dH = make_fastdiv(g.H);
    g.dSy = make_fastdiv(g.sy);
    return g;
}

// Host self-check of k_resid's guard on an nx x ny x nz grid: every slot of
// both colours is taken as live (a finite b/dt everywhere, the case of a
// sentinel that was overwritten); a slot the guard admits reads its six
// neighbours and its coefficient entries, which must all lie inside the
// arrays.  Returns the number of reads outside.  drop_interior leaves the
// interior test out (the guard before it was added), so a caller can see
// the check catch that form.
int64_t resid_guard_selfcheck(int64_t nx, int64_t ny, int64_t nz, bool drop_interior) {
    const ColourLayout lay = colour_layout(nx, ny, nz);
    SorGeom g;
    g.H = (int)lay.H; g.NC = (int)lay.NC; g.sy = (int)lay.sy; g.sx = (int)lay.sx; g.sz = (int)lay.sz;
    g.plane = (int)(lay.H * lay.sy); g.sxp = (int)lay.sxp;
    g.dH = make_fastdiv(g.H);
    g.dSy = make_fastdiv(g.sy);
    int64_t outside = 0;
    auto in = [&](int64_t i, int64_t n) { return i >= 0 && i < n; };
    for (int colour = 0; colour < 2; colour++) {
        for (int f = 0; f < g.NC; f++) {
            int ii, jj, kk;
            decode(g, f, colour, ii, jj, kk);
            if (!drop_interior && !interior_cell(g, ii, jj, kk)) continue;
            const int q = (colour + jj + kk) & 1;
            for (int64_t idx : {(int64_t)f - 1 + q, (int64_t)f + q, (int64_t)f - g.H, (int64_t)f + g.H,
                                (int64_t)f - g.plane, (int64_t)f + g.plane})
                outside += !in(idx, g.NC);
            outside += !in(ii, g.sxp) + !in(2 * (int64_t)g.sxp + ii, 3 * (int64_t)g.sxp);
            outside += !in(jj, g.sy) + !in(kk, g.sz);
        }
    }
    return outside;
}

// host-side self check of the division helper for this grid's divisors
int check_fastdiv(int d, int64_t nmax) {
    FastDiv f = make_fastdiv(d);
    for (int64_t n = 0; n < nmax; n += 1 + n / 97) {
        if ((int64_t)fdiv((uint32_t)n, f) != n / d) return -1;
    }
    for (int64_t n = std::max<int64_t>(0, nmax - 4096); n < nmax; n++)
        if ((int64_t)fdiv((uint32_t)n, f) != n / d) return -1;
    return 0;
}

// grow the slot array, keeping its contents (a batch may gate on the last
// slot of the previous one)
int ensure_slots(int64_t n) {
    Ctx &c = ctx();
    if (n <= c.n_slots) return 0;
    const int64_t nn = std::max<int64_t>(2 * n, 4096);
    double *fresh = nullptr;
    IB3D_CHECK(hipMalloc(&fresh, nn * sizeof(double)));
    IB3D_CHECK(hipMemsetAsync(fresh, 0, nn * sizeof(double), c.stream));
    if (c.slots) {
        IB3D_CHECK(hipMemcpyAsync(fresh, c.slots, c.n_slots * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
        IB3D_CHECK(hipStreamSynchronize(c.stream));
        IB3D_CHECK(hipFree(c.slots));
    }
    IB3D_CHECK(hipStreamSynchronize(c.stream));
    c.slots = fresh;
    c.n_slots = nn;
    return 0;
}

int ensure_partial(int64_t n) {
    Ctx &c = ctx();
    if (n <= c.n_partial) return 0;
    IB3D_CHECK(hipStreamSynchronize(c.stream));
    if (c.partial) IB3D_CHECK(hipFree(c.partial));
    c.n_partial = n;
    IB3D_CHECK(hipMalloc(&c.partial, n * sizeof(double)));
    return 0;
}

// Float copies of the Laplacian rows for the FP32 smoother, laid out as
// [3][s] per axis like the double rows.
static float *g_axf = nullptr, *g_ayf = nullptr, *g_azf = nullptr;
static int64_t g_f_sx = 0, g_f_sy = 0, g_f_sz = 0;

int sor_float_tables(Ctx &c) {
    if (g_axf && g_f_sx == c.sx && g_f_sy == c.sy && g_f_sz == c.sz) return 0;
    if (g_axf) { (void)hipFree(g_axf); (void)hipFree(g_ayf); (void)hipFree(g_azf); }
    std::vector<double> hx(3 * c.sxp), hy(3 * c.sy), hz(3 * c.sz);
    IB3D_CHECK(d2h(hx.data(), c.ax, hx.size() * 8));
    IB3D_CHECK(d2h(hy.data(), c.ay, hy.size() * 8));
    IB3D_CHECK(d2h(hz.data(), c.az, hz.size() * 8));
    std::vector<float> fx(hx.begin(), hx.end()), fy(hy.begin(), hy.end()), fz(hz.begin(), hz.end());
    IB3D_CHECK(hipMalloc(&g_axf, fx.size() * 4));
    IB3D_CHECK(hipMalloc(&g_ayf, fy.size() * 4));
    IB3D_CHECK(hipMalloc(&g_azf, fz.size() * 4));
    IB3D_CHECK(h2d(g_axf, fx.data(), fx.size() * 4));
    IB3D_CHECK(h2d(g_ayf, fy.data(), fy.size() * 4));
    IB3D_CHECK(h2d(g_azf, fz.data(), fz.size() * 4));
    g_f_sx = c.sx; g_f_sy = c.sy; g_f_sz = c.sz;
    return 0;
}

void gmres_release();
void mg_release();
void cycle_graph_release();
void sor_release_tables() {
    cycle_graph_release();
    mg_release();
    gmres_release();
    if (g_axf) { (void)hipFree(g_axf); (void)hipFree(g_ayf); (void)hipFree(g_azf); }
    g_axf = g_ayf = g_azf = nullptr;
    g_f_sx = g_f_sy = g_f_sz = 0;
}

// Allocation-free computeDiv: the same kernels, with the per-block partial
// maxima and sums of squares and the two reduced values in a static device
// array instead of memory the library allocates on demand.  Each GPU holds
// its own copy of `scratch` for the life of the process; its address comes
// from hipGetSymbolAddress once per device (noalloc_setup, called by
// ib3d_init), because in host code the array's name is not a device address.
// Every call on a GPU shares that GPU's array and leaves its contents behind:
// the kernels overwrite every entry they read, and calls that overlap on one
// GPU race.
static constexpr int64_t kScratchDoubles = (int64_t)1 << 19;  // 4 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];
static constexpr int kMaxDevices = 64;
static double *g_scratch[kMaxDevices] = {};

int noalloc_setup(int device) {
    IB3D_REQUIRE(device >= 0 && device < kMaxDevices, "noalloc_setup: device index out of range");
    if (g_scratch[device]) return 0;
    void *ptr = nullptr;
    IB3D_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch[device] = static_cast<double *>(ptr);
    return 0;
}

int compute_div_noalloc(Ctx &c, double deltat, double *divmax) {
    const int64_t nb = nblocks(c.P, kSorBlock);
    // partial maxima [0, nb), partial sums [nb, 2nb), results [2nb, 2nb+2)
    if (2 * nb + 2 > kScratchDoubles) {
        set_error("computeDiv_noalloc: this grid needs " + std::to_string(2 * nb + 2) +
                  " scratch doubles; the static scratch holds " + std::to_string(kScratchDoubles) +
                  " (at most about 67 million padded cells)");
        return -1;
    }
    IB3D_REQUIRE(c.device >= 0 && c.device < kMaxDevices && g_scratch[c.device],
                 "computeDiv_noalloc: the scratch address of this device has not been looked up (ib3d_init)");
    double *sc = g_scratch[c.device];
    hipLaunchKernelGGL(k_div, dim3((unsigned)nb), dim3(kSorBlock), 0, c.stream, c.ut, c.vt, c.wt, c.tag, c.dx,
                       c.dy, c.dz, c.bdt[0], c.bdt[1], (int)c.sx, (int)c.sy, (int)c.sz, (int)c.H,
                       make_fastdiv((int)c.sx), make_fastdiv((int)c.sy), deltat, sc, sc + nb);
    launch_max_partials(sc, nb, sc + 2 * nb);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, c.stream, sc + nb, nb, sc + 2 * nb + 1);
    IB3D_CHECK(hipGetLastError());
    IB3D_CHECK(hipMemcpyAsync(c.hscal, sc + 2 * nb, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    IB3D_CHECK(hipStreamSynchronize(c.stream));
    *divmax = c.hscal[0];
    c.bnorm2 = c.hscal[1];
    return 0;
}

// wait = false: max|div| and ||b/dt||^2 are left to arrive in c.hscal[0..1]
// at the caller's next synchronisation (the solve then reads the norm from
// device memory)
int compute_div(Ctx &c, double deltat, double *divmax, double *bnorm2, bool wait) {
    const unsigned nb = nblocks(c.P, kSorBlock);
    if (ensure_partial(std::max<int64_t>(2 * (int64_t)nb, nblocks(2 * c.NC, kSorBlock)))) return -1;
    IB3D_CHECK(hipMemsetAsync(c.scalars, 0, 8 * sizeof(double), c.stream));
    hipLaunchKernelGGL(k_div, dim3(nb), dim3(kSorBlock), 0, c.stream, c.ut, c.vt, c.wt, c.tag, c.dx, c.dy,
                       c.dz, c.bdt[0], c.bdt[1], (int)c.sx, (int)c.sy, (int)c.sz, (int)c.H, make_fastdiv((int)c.sx),
                       make_fastdiv((int)c.sy), deltat, c.partial + nb, c.partial);
    launch_max_partials(c.partial + nb, (int64_t)nb, c.scalars);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, c.stream, c.partial, (int64_t)nb, c.scalars + 1);
    IB3D_CHECK(hipGetLastError());
    IB3D_CHECK(hipMemcpyAsync(c.hscal, c.scalars, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (!wait) return 0;
    IB3D_CHECK(hipStreamSynchronize(c.stream));
    *divmax = c.hscal[0];
    if (bnorm2) *bnorm2 = c.hscal[1];
    c.bnorm2 = c.hscal[1];
    return 0;
}

// ||b/dt|| and the Krylov target kt_tol x ||b/dt|| from ||b/dt||^2 on the
// device, for a solve queued before the host has read the norm
__global__ void k_norm_target(const double *bnorm2, double *out, double kt_tol) {
    if (threadIdx.x != 0) return;
    const double bn = sqrt(*bnorm2);
    out[0] = bn;
    out[1] = kt_tol * bn;
}

int norm_target_async(Ctx &c, const double *bnorm2, double *out) {
    hipLaunchKernelGGL(k_norm_target, dim3(1), dim3(64), 0, c.stream, bnorm2, out, c.krylov_target * c.tol);
    IB3D_CHECK(hipGetLastError());
    return 0;
}

static dim3 pz_grid(const Ctx &c) {
    return dim3(nblocks((c.H * c.sy) / 2, kSorBlock), (unsigned)((c.sz - 2 + kZch - 1) / kZch));
}

// FP64 sweep (both colours) gated on *gate < thr
static void sweep64(Ctx &c, double omega, bool track, double *slot, const double *gate, double thr) {
    const SorGeom g = geom(c);
    for (int col = 0; col < 2; col++) {
        if (track)
            hipLaunchKernelGGL((k_sor_pz<double, true, false, false>), pz_grid(c), dim3(kSorBlock), 0, c.stream,
                               c.pc[col], c.pc[1 - col], c.bdt[col], c.ax, c.ay, c.az, g, col, omega, 1.0 - omega,
                               slot, gate, thr, nullptr);
        else
            hipLaunchKernelGGL((k_sor_pz<double, false, false, false>), pz_grid(c), dim3(kSorBlock), 0, c.stream,
                               c.pc[col], c.pc[1 - col], c.bdt[col], c.ax, c.ay, c.az, g, col, omega, 1.0 - omega,
                               slot, gate, thr, nullptr);
    }
}

// FP32 sweep (both colours) on the increment e; first: start from zero,
// apply: first add the previous increment to the FP64 pc
static void sweep32(Ctx &c, double omega, bool first, bool apply, const double *gate, double thr) {
    const SorGeom g = geom(c);
    const float om = (float)omega, omm = (float)(1.0 - omega);
    for (int col = 0; col < 2; col++) {
        float *mine = c.e32[col];
        const float *other = c.e32[1 - col];
        const float *rhs = c.r32[col];
        if (first && apply)
            hipLaunchKernelGGL((k_sor_pz<float, false, true, true>), pz_grid(c), dim3(kSorBlock), 0, c.stream,
                               mine, other, rhs, g_axf, g_ayf, g_azf, g, col, om, omm, nullptr, gate, thr,
                               c.pc[col]);
        else if (first)
            hipLaunchKernelGGL((k_sor_pz<float, false, true, false>), pz_grid(c), dim3(kSorBlock), 0, c.stream,
                               mine, other, rhs, g_axf, g_ayf, g_azf, g, col, om, omm, nullptr, gate, thr, nullptr);
        else
            hipLaunchKernelGGL((k_sor_pz<float, false, false, false>), pz_grid(c), dim3(kSorBlock), 0, c.stream,
                               mine, other, rhs, g_axf, g_ayf, g_azf, g, col, om, omm, nullptr, gate, thr, nullptr);
    }
}

// subPcorVcor.f90:148-212 with the reference stopping rule.  Iterations are
// queued in batches; each kernel first checks whether the previous
// iteration already converged, so the result (and the count) equals the
// one-iteration-at-a-time loop while the host synchronises once per batch.
int redblack_sor_reference(Ctx &c, double epsi, int64_t *isum, double *derr2) {
    const double omega = 1.99;
    int64_t done = 0;
    int batch = std::max(1, c.batch);
    for (;;) {
        if (ensure_slots(done + batch + 1)) return -1;
        IB3D_CHECK(hipMemsetAsync(c.slots + done, 0, batch * sizeof(double), c.stream));
        for (int q = 0; q < batch; q++) {
            const int64_t m = done + q;
            const double *gate = m > 0 ? c.slots + m - 1 : nullptr;
            sweep64(c, omega, true, c.slots + m, gate, epsi);
        }
        IB3D_CHECK(hipGetLastError());
        std::vector<double> h(batch);
        IB3D_CHECK(hipMemcpyAsync(h.data(), c.slots + done, batch * sizeof(double), hipMemcpyDeviceToHost,
                                  c.stream));
        IB3D_CHECK(hipStreamSynchronize(c.stream));
        for (int q = 0; q < batch; q++) {
            if (!(h[q] >= epsi)) {  // the reference loops while derr2 >= epsi
                *isum = done + q + 1;
                *derr2 = h[q];
                c.last_sweeps = *isum;
                return 0;
            }
        }
        done += batch;
        if (done > c.max_sweeps) {
            set_error("REDBLACKSOR: no convergence within the sweep cap");
            return -1;
        }
        batch = std::min(batch * 2, 256);
    }
}

// residual of the current p (+e when add) into r32 (when write_r32) and
// partial sums; then rel[o] and conv[o], gated on conv[o-1]
static void residual_async(Ctx &c, bool add, bool write_r32, double bn, double *rel, double *conv,
                           const double *gate, const double *bnp = nullptr) {
    const dim3 grid = pz_grid(c);
    const int64_t nb = (int64_t)grid.x * grid.y;
    const SorGeom g = geom(c);
    float *r0 = c.r32[0], *r1 = c.r32[1];
    const size_t xlds = (size_t)3 * (2 * c.H + 2) * sizeof(double);  // the kernel's x-coefficient table
#define IB3D_RESID(A, W)                                                                                         \
    hipLaunchKernelGGL((k_resid_pz<A, W>), grid, dim3(kSorBlock), xlds, c.stream, c.pc[0], c.pc[1], c.e32[0],      \
                       c.e32[1], c.bdt[0], c.bdt[1], r0, r1, c.ax, c.ay, c.az, g, c.partial, gate, 0.5)
    if (add && write_r32) IB3D_RESID(true, true);
    else if (add) IB3D_RESID(true, false);
    else if (write_r32) IB3D_RESID(false, true);
    else IB3D_RESID(false, false);
#undef IB3D_RESID
    hipLaunchKernelGGL(k_finish_norm, dim3(1), dim3(1024), 0, c.stream, c.partial, nb, bn, c.tol, rel, conv, gate,
                       bnp);
}

static void add_increment(Ctx &c) {
    hipLaunchKernelGGL(k_add_e, dim3(nblocks(2 * c.NC, kSorBlock)), dim3(kSorBlock), 0, c.stream, c.pc[0], c.pc[1],
                       c.e32[0], c.e32[1], c.bdt[0], c.bdt[1], c.NC);
}

// FP64 or mixed L2 solve.  Outer iteration o (1, 2, ...): `check` sweeps
// (FP64 on pc, or FP32 on the increment e from zero, the first of which also
// adds the previous increment to pc), then the FP64 residual.  Everything of
// iteration o is gated on conv[o-1]; the host reads rel/conv once per batch
// of outer iterations.  Index 0 holds the initial residual.
static double g_bn = 0.0;
static int resid_rel(Ctx &c, double *rel) {
    residual_async(c, false, false, g_bn, rel, c.scalars + 19, nullptr);
    IB3D_CHECK(hipGetLastError());
    return 0;
}

static int sor_l2(Ctx &c, bool mixed, double bn, int64_t *sweeps, int64_t *outer, double *rel, double *rel0) {
    if (mixed && sor_float_tables(c)) return -1;
    if (ensure_partial(nblocks(2 * c.NC, kSorBlock))) return -1;
    const int64_t max_outer = std::max<int64_t>(1, c.max_sweeps / c.check);
    if (ensure_slots(2 * (max_outer + 2))) return -1;
    double *relv = c.slots, *conv = c.slots + c.n_slots / 2;
    hipLaunchKernelGGL(k_clear_dead, dim3(nblocks(2 * c.NC, kSorBlock)), dim3(kSorBlock), 0, c.stream, c.pc[0],
                       c.pc[1], c.e32[0], c.e32[1], c.bdt[0], c.bdt[1], c.NC);
    // learned initial guess / training-pair recording: residual of p_prev
    // (FP32 copy in r32) into scalars[16], the guess's into scalars[18]
    const bool cnn = cnn_active();
    if (cnn) {
        g_bn = bn;
        residual_async(c, false, true, bn, c.scalars + 16, c.scalars + 17, nullptr);
        IB3D_CHECK(hipGetLastError());
        if (cnn_before_solve(c, c.scalars + 16, c.scalars + 18, resid_rel)) return -1;
    }
    residual_async(c, false, mixed, bn, relv, conv, nullptr);
    int64_t o = 0;     // last queued outer iteration
    int64_t seen = 0;  // first index not yet inspected on the host
    // the previous solve's count predicts this one's (warm starts, slowly
    // varying flow): queue a little less than that first, then grow
    int batch = (int)std::min<int64_t>(64, std::max<int64_t>(4, c.last_outer - 2));
    std::vector<double> hr, hc;
    for (;;) {
        for (int q = 0; q < batch && o < max_outer; q++) {
            o++;
            const double *gate = conv + (o - 1);
            for (int s = 0; s < c.check; s++) {
                if (mixed)
                    sweep32(c, c.omega, s == 0, s == 0 && o > 1, gate, 0.5);
                else
                    sweep64(c, c.omega, false, nullptr, gate, 0.5);
            }
            residual_async(c, mixed, mixed, bn, relv + o, conv + o, gate);
        }
        IB3D_CHECK(hipGetLastError());
        const int64_t n = o - seen + 1;
        hr.resize(n);
        hc.resize(n);
        IB3D_CHECK(hipMemcpyAsync(hr.data(), relv + seen, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        IB3D_CHECK(hipMemcpyAsync(hc.data(), conv + seen, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        IB3D_CHECK(hipStreamSynchronize(c.stream));
        if (seen == 0) *rel0 = hr[0];
        for (int64_t q = 0; q < n; q++) {
            if (!(hr[q] == hr[q]) || !std::isfinite(bn)) {
                set_error("poissonSolver: the residual is not finite");
                return -1;
            }
            if (hc[q] < 0.5) {
                const int64_t oc = seen + q;  // converged after oc outer iterations
                if (mixed && oc > 0) add_increment(c);
                IB3D_CHECK(hipGetLastError());
                if (cnn && cnn_after_solve(c, c.scalars + 16, c.scalars + 18)) return -1;
                *sweeps = oc * c.check;
                *outer = oc;
                *rel = hr[q];
                return 0;
            }
        }
        if (o >= max_outer) {
            if (mixed && o > 0) add_increment(c);
            IB3D_CHECK(hipGetLastError());
            if (cnn && cnn_after_solve(c, c.scalars + 16, c.scalars + 18)) return -1;
            *sweeps = o * c.check;
            *outer = o;
            *rel = hr[n - 1];
            return 0;
        }
        seen = o + 1;
        batch = std::min(std::max(batch, 4) * 2, 64);
    }
}

// GMRES-IR buffers: V (m+1 vectors) and Z (m vectors) of 2*NC floats, H,
// beta, y and reduction partials
static float *g_kv = nullptr, *g_kz = nullptr;
static int64_t g_z_geom = -1;
static double *g_kh = nullptr, *g_kpart = nullptr, *g_kpart2 = nullptr;
static uint32_t *g_actbits = nullptr;  // the finest level's slot flags (k_active_bits)
static int64_t g_actbits_nw = 0;       // words per colour
static int64_t g_k_nc = 0;
static int g_k_m = 0;

__global__ void k_fill_nan_f32(float *a, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * kSorBlock + threadIdx.x;
    if (q < n) a[q] = __builtin_nanf("");
}

// ------------------------------------------------------------ multigrid
// FP32 geometric multigrid V-cycle, the FGMRES preconditioner of Poisson
// mode 4 (in place of k SOR sweeps).  Level l+1 halves every interior
// dimension of level l: a coarse cell is the union of 2x2x2 fine cells and
// its faces are every other fine face, so a stretched grid is coarsened
// along its own lines.  The coarse Laplacian rows come from the coarse cell
// centres with the same construction and boundary folds as
// coefficientMatrix (subCoefficientMatrix.f90:8-53).  A coarse cell takes
// part when any of its children does (its right-hand side is then the mean
// of the active children's residuals, else NaN); cells that do not take part
// hold 0, as solid and intercepted cells do on the fine grid.  Smoother:
// red-black SOR sweeps (k_sor_pz), the first from zero; prolongation:
// every active fine cell adds the trilinear interpolant of the coarse
// correction at its centre (coarse cells outside the solve count as 0; at the
// domain ends the ghost value follows the boundary fold: equal to the end
// cell's at the Neumann ends, its negative at the x outlet).

// Interpolation along one axis from a fine cell's centre: coarse cells i0,
// i1 (padded indices of the coarse level) with weights w0, w1.
struct MgAxis {
    int i0, i1;
    float w0, w1;
};

// One level as the kernels see it: two colour arrays per vector (2*NC floats).
struct MgView {
    float *x, *b;                   // correction, right-hand side (NaN where the cell is not solved for)
    float *r;                       // stored residual b - A x (finest level only)
    const float *ax, *ay, *az;      // Laplacian rows
    const MgAxis *tx, *ty, *tz;     // prolongation from the next level
    int nx, ny, nz, H, NC, sy, sz, sxp;
    FastDiv dH, dSy;                // division by H and sy (slot decoding)
};

__device__ __forceinline__ void view_decode(const MgView &L, int f, int c, int &ii, int &jj, int &kk) {
    const int row = (int)fdiv((uint32_t)f, L.dH);
    const int h = f - row * L.H;
    kk = (int)fdiv((uint32_t)row, L.dSy);
    jj = row - kk * L.sy;
    ii = 2 * h + ((c + jj + kk) & 1);
}

// b - A x at an active cell (colour c, slot f, padded position ii, jj, kk);
// the same operation order as the FP32 matrix-vector product
__device__ __forceinline__ float cell_resid(const MgView &L, float rb, int c, int f, int ii, int jj, int kk) {
    const float *o = L.x + (int64_t)(1 - c) * L.NC;
    const int s = ii & 1, plane = L.H * L.sy;
    const float a1x = L.ax[ii], a2x = L.ax[L.sxp + ii], a3x = L.ax[2 * L.sxp + ii];
    const float a1y = L.ay[jj], a2y = L.ay[L.sy + jj], a3y = L.ay[2 * L.sy + jj];
    const float a1z = L.az[kk], a2z = L.az[L.sz + kk], a3z = L.az[2 * L.sz + kk];
    const float ax = (((((a1x * o[f - 1 + s] + a3x * o[f + s]) + a1y * o[f - L.H]) + a3y * o[f + L.H]) +
                       a1z * o[f - plane]) + a3z * o[f + plane]) +
                     ((a2x + a2y) + a2z) * L.x[(int64_t)c * L.NC + f];
    return rb - ax;
}

// coarse right-hand side at coarse slot q (colour cc, slot fc): the mean of
// the active children's residuals b - A x (STORED: read from F.r); NaN on
// coarse cells without an active child and on ghost / padding slots
template <bool STORED>
__device__ __forceinline__ float restrict_at(const MgView &F, const MgView &C, int cc, int fc, int I, int J, int K) {
    float out = __builtin_nanf("");
    if (I >= 1 && I <= C.nx && J >= 1 && J <= C.ny && K >= 1 && K <= C.nz) {
        float s = 0.0f;
        int n = 0;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int ii = 2 * I - 1 + (t & 1), jj = 2 * J - 1 + ((t >> 1) & 1), kk = 2 * K - 1 + (t >> 2);
            const int c = (ii + jj + kk) & 1;
            const int f = (jj + F.sy * kk) * F.H + (ii >> 1);
            if (STORED) {
                const float v = F.r[(int64_t)c * F.NC + f];
                if (v == v) {
                    s += v;
                    n++;
                }
            } else {
                const float rb = F.b[(int64_t)c * F.NC + f];
                if (rb == rb) {
                    s += cell_resid(F, rb, c, f, ii, jj, kk);
                    n++;
                }
            }
        }
        if (n) out = s / (float)n;
    }
    return out;
}

// trilinear interpolant of the coarse correction at an active fine cell
__device__ __forceinline__ float prolong_at(const MgView &F, const MgView &C, int ii, int jj, int kk) {
    const MgAxis ax = F.tx[ii], ay = F.ty[jj], az = F.tz[kk];
    const int I[2] = {ax.i0, ax.i1}, J[2] = {ay.i0, ay.i1}, K[2] = {az.i0, az.i1};
    const float wx[2] = {ax.w0, ax.w1}, wy[2] = {ay.w0, ay.w1}, wz[2] = {az.w0, az.w1};
    float s = 0.0f;
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const int a = t & 1, b = (t >> 1) & 1, d = t >> 2;
        const int cc = (I[a] + J[b] + K[d]) & 1;
        s += (wx[a] * wy[b] * wz[d]) * C.x[(int64_t)cc * C.NC + (J[b] + C.sy * K[d]) * C.H + (I[a] >> 1)];
    }
    return s;
}

// residual of the fine level restricted to the coarse right-hand side, one
// thread per coarse slot; STORED: the finest level's residual was stored by
// the z-marching residual kernel, else it is formed here
template <bool STORED>
__global__ void __launch_bounds__(kSorBlock)
k_mg_resid_restrict(MgView F, MgView C, SorGeom gc, const double *gate, double thr) {
    if (gated(gate, thr)) return;
    const int64_t q = (int64_t)xcd_block() * kSorBlock + threadIdx.x;
    if (q >= 2 * (int64_t)C.NC) return;
    const int cc = q >= C.NC ? 1 : 0;
    const int fc = (int)(q - (int64_t)cc * C.NC);
    int I, J, K;
    decode(gc, fc, cc, I, J, K);
    C.b[q] = restrict_at<STORED>(F, C, cc, fc, I, J, K);
}

// The same restriction in one pass over the fine level, with no stored
// residual: a thread per coarse column (I, J), marching a chunk of coarse
// planes K.  Per fine plane it reads the rows of its four children
// (ii = 2I-1, 2I; jj = 2J-1, 2J), their x and y neighbours and b: across a
// wave every such read is one colour array at consecutive slots (the colour
// of ii = 2I-1 + a does not depend on I), so the rows coalesce.  The
// children's x values on the plane below the pair and on the pair's first
// plane come from the previous step in registers.  The residuals use
// cell_resid's operations and the children are summed in restrict_at's order,
// so the coarse right-hand side is bit for bit the two-pass one.  Every
// coarse slot is written (NaN outside the interior), as k_mg_resid_restrict
// does.  kch: coarse planes per chunk; dW divides by 2*C.H, the coarse cells
// of a padded row.
__global__ void __launch_bounds__(kSorBlock)
k_resid_restrict_pz(MgView F, MgView C, FastDiv dW, int kch, const double *gate, double thr) {
    if (gated(gate, thr)) return;
    int bx, by;
    xcd_block2(bx, by);
    const int W = 2 * C.H;
    const int t = bx * kSorBlock + (int)threadIdx.x;
    if (t >= W * C.sy) return;
    const int J = (int)fdiv((uint32_t)t, dW);
    const int I = t - J * W;
    const int K0 = 1 + by * kch, K1 = min(K0 + kch, C.sz - 1);
    auto cslot = [&](int K) { return ((I + J + K) & 1) * C.NC + (J + C.sy * K) * C.H + (I >> 1); };
    const float nan = __builtin_nanf("");
    if (K0 == 1) C.b[cslot(0)] = nan;
    if (K1 == C.sz - 1) C.b[cslot(C.sz - 1)] = nan;
    if (!(I >= 1 && I <= C.nx && J >= 1 && J <= C.ny)) {
        for (int K = K0; K < K1; K++) C.b[cslot(K)] = nan;
        return;
    }
    typedef __attribute__((address_space(4))) float CFloat;
    const CFloat *az = (const CFloat *)F.az;
    // x of fine cell (2I-1+a, jj, kk), a = -1..2: colour e = (jj+kk)&1 holds
    // a = -1 and 1 at slots I-1 and I, colour 1-e holds a = 0 and 2
    auto rowbase = [&](int jj, int kk) { return (jj + F.sy * kk) * F.H + I; };  // slot of ii = 2I
    struct Pair {
        float c0, c1;  // ii = 2I-1, 2I (the children), or 2I-2, 2I+1 (their x neighbours)
    };
    auto load_we = [&](int jj, int kk) {
        const int f = rowbase(jj, kk), e = (jj + kk) & 1;
        return Pair{F.x[(int64_t)e * F.NC + f - 1], F.x[(int64_t)(1 - e) * F.NC + f]};
    };
    auto load_pair = [&](const float *a, int jj, int kk) {
        const int f = rowbase(jj, kk), e = (jj + kk) & 1;
        return Pair{a[(int64_t)(1 - e) * F.NC + f - 1], a[(int64_t)e * F.NC + f]};
    };
    const int i0 = 2 * I - 1, j0 = 2 * J - 1;
    const float a1x0 = F.ax[i0], a2x0 = F.ax[F.sxp + i0], a3x0 = F.ax[2 * F.sxp + i0];
    const float a1x1 = F.ax[i0 + 1], a2x1 = F.ax[F.sxp + i0 + 1], a3x1 = F.ax[2 * F.sxp + i0 + 1];
    const float a1y0 = F.ay[j0], a2y0 = F.ay[F.sy + j0], a3y0 = F.ay[2 * F.sy + j0];
    const float a1y1 = F.ay[j0 + 1], a2y1 = F.ay[F.sy + j0 + 1], a3y1 = F.ay[2 * F.sy + j0 + 1];
    // cell_resid, operand for operand
    auto resid = [](float rb, float a1x, float a2x, float a3x, float a1y, float a2y, float a3y, float a1z, float a2z,
                    float a3z, float w, float e, float s, float n, float b, float a, float own) {
        const float ax = (((((a1x * w + a3x * e) + a1y * s) + a3y * n) + a1z * b) + a3z * a) + ((a2x + a2y) + a2z) * own;
        return rb - ax;
    };
    // the children's x (rows j0 and j0 + 1) on plane ka - 1 (m) and on plane
    // ka (a) of the current pair of planes (ka = 2K-1, kb = 2K)
    Pair m0 = load_pair(F.x, j0, 2 * K0 - 2), m1 = load_pair(F.x, j0 + 1, 2 * K0 - 2);
    Pair a0 = load_pair(F.x, j0, 2 * K0 - 1), a1 = load_pair(F.x, j0 + 1, 2 * K0 - 1);
    for (int K = K0; K < K1; K++) {
        const int ka = 2 * K - 1, kb = 2 * K;
        const Pair wa0 = load_we(j0, ka), wa1 = load_we(j0 + 1, ka);
        const Pair wb0 = load_we(j0, kb), wb1 = load_we(j0 + 1, kb);
        const Pair b0 = load_pair(F.x, j0, kb), b1 = load_pair(F.x, j0 + 1, kb);
        const Pair sa = load_pair(F.x, j0 - 1, ka), na = load_pair(F.x, j0 + 2, ka);
        const Pair sb = load_pair(F.x, j0 - 1, kb), nb = load_pair(F.x, j0 + 2, kb);
        const Pair c0 = load_pair(F.x, j0, kb + 1), c1 = load_pair(F.x, j0 + 1, kb + 1);
        const Pair ba0 = load_pair(F.b, j0, ka), ba1 = load_pair(F.b, j0 + 1, ka);
        const Pair bb0 = load_pair(F.b, j0, kb), bb1 = load_pair(F.b, j0 + 1, kb);
        const float a1za = az[ka], a2za = az[F.sz + ka], a3za = az[2 * F.sz + ka];
        const float a1zb = az[kb], a2zb = az[F.sz + kb], a3zb = az[2 * F.sz + kb];
        float s = 0.0f;
        int n = 0;
        auto add = [&](float rb, float r) {
            if (rb == rb) {
                s += r;
                n++;
            }
        };
        // plane ka (children a0, a1): below m, above b; west/east: the other
        // child of the pair or the x neighbour; south/north: the other row or
        // the y neighbour
        add(ba0.c0, resid(ba0.c0, a1x0, a2x0, a3x0, a1y0, a2y0, a3y0, a1za, a2za, a3za, wa0.c0, a0.c1, sa.c0, a1.c0,
                          m0.c0, b0.c0, a0.c0));
        add(ba0.c1, resid(ba0.c1, a1x1, a2x1, a3x1, a1y0, a2y0, a3y0, a1za, a2za, a3za, a0.c0, wa0.c1, sa.c1, a1.c1,
                          m0.c1, b0.c1, a0.c1));
        add(ba1.c0, resid(ba1.c0, a1x0, a2x0, a3x0, a1y1, a2y1, a3y1, a1za, a2za, a3za, wa1.c0, a1.c1, a0.c0, na.c0,
                          m1.c0, b1.c0, a1.c0));
        add(ba1.c1, resid(ba1.c1, a1x1, a2x1, a3x1, a1y1, a2y1, a3y1, a1za, a2za, a3za, a1.c0, wa1.c1, a0.c1, na.c1,
                          m1.c1, b1.c1, a1.c1));
        // plane kb (children b0, b1): below a, above c
        add(bb0.c0, resid(bb0.c0, a1x0, a2x0, a3x0, a1y0, a2y0, a3y0, a1zb, a2zb, a3zb, wb0.c0, b0.c1, sb.c0, b1.c0,
                          a0.c0, c0.c0, b0.c0));
        add(bb0.c1, resid(bb0.c1, a1x1, a2x1, a3x1, a1y0, a2y0, a3y0, a1zb, a2zb, a3zb, b0.c0, wb0.c1, sb.c1, b1.c1,
                          a0.c1, c0.c1, b0.c1));
        add(bb1.c0, resid(bb1.c0, a1x0, a2x0, a3x0, a1y1, a2y1, a3y1, a1zb, a2zb, a3zb, wb1.c0, b1.c1, b0.c0, nb.c0,
                          a1.c0, c1.c0, b1.c0));
        add(bb1.c1, resid(bb1.c1, a1x1, a2x1, a3x1, a1y1, a2y1, a3y1, a1zb, a2zb, a3zb, b1.c0, wb1.c1, b0.c1, nb.c1,
                          a1.c1, c1.c1, b1.c1));
        C.b[cslot(K)] = n ? s / (float)n : nan;
        m0 = b0;
        m1 = b1;
        a0 = c0;
        a1 = c1;
    }
}

// x_fine += trilinear interpolant of x_coarse, on active fine cells
__global__ void __launch_bounds__(kSorBlock)
k_mg_prolong(MgView F, MgView C, SorGeom gf, const double *gate, double thr) {
    if (gated(gate, thr)) return;
    const int64_t q = (int64_t)xcd_block() * kSorBlock + threadIdx.x;
    if (q >= 2 * (int64_t)F.NC) return;
    if (!(F.b[q] == F.b[q])) return;
    const int c = q >= F.NC ? 1 : 0;
    int ii, jj, kk;
    decode(gf, (int)(q - (int64_t)c * F.NC), c, ii, jj, kk);
    F.x[q] += prolong_at(F, C, ii, jj, kk);
}

// The same on the finest level, four consecutive fine cells of a row (two of
// each colour, ii = 2h..2h+3) per thread: their trilinear stencils only use
// the coarse x-columns h-1..h+3 (clamped to the padded range), so the y-z
// interpolant is formed once per column (5 x 4 gathers instead of 4 x 8).
__global__ void __launch_bounds__(kSorBlock)
k_mg_prolong_pz(MgView F, MgView C, SorGeom gf, const uint32_t *__restrict__ bits, int64_t nw, const double *gate,
                double thr) {
    if (gated(gate, thr)) return;
    int bx, by;
    xcd_block2(bx, by);
    const int p = 2 * (bx * kSorBlock + (int)threadIdx.x);
    const int jj = (int)fdiv((uint32_t)p, gf.dH);
    const int h = p - jj * gf.H;
    if (!(p < gf.plane && jj >= 1 && jj <= gf.sy - 2)) return;
    const int kk = 1 + by;
    const int f = p + kk * gf.plane;
    // the active cells: the slot flags of F.b (bits), or F.b itself
    bool a0x, a0y, a1x, a1y;
    if (bits) {
        active_pair(bits, nw, 0, f, a0x, a0y);
        active_pair(bits, nw, 1, f, a1x, a1y);
    } else {
        const f2v b0 = *reinterpret_cast<const f2v *>(F.b + f), b1 = *reinterpret_cast<const f2v *>(F.b + F.NC + f);
        a0x = b0.x == b0.x;
        a0y = b0.y == b0.y;
        a1x = b1.x == b1.x;
        a1y = b1.y == b1.y;
    }
    if (!(a0x || a0y || a1x || a1y)) return;
    const MgAxis ay = F.ty[jj], az = F.tz[kk];
    const int J[2] = {ay.i0, ay.i1}, K[2] = {az.i0, az.i1};
    const float wy[2] = {ay.w0, ay.w1}, wz[2] = {az.w0, az.w1};
    float col[5];
#pragma unroll
    for (int a = 0; a < 5; a++) {
        const int I = min(max(h - 1 + a, 0), C.nx + 1);
        float s = 0.0f;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int b = t & 1, d = t >> 1;
            const int cc = (I + J[b] + K[d]) & 1;
            s += (wy[b] * wz[d]) * C.x[(int64_t)cc * C.NC + (J[b] + C.sy * K[d]) * C.H + (I >> 1)];
        }
        col[a] = s;
    }
    // cell at ii = 2h + e: colour (ii + jj + kk) & 1, slot f + (e >> 1).  The
    // table has 2H entries (padding cells: weights 0), and the column index
    // of an active cell lies in 0..4; it is clamped for the others.
    auto pick = [&](int i) {
        const int q = min(max(i - (h - 1), 0), 4);
        return q == 0 ? col[0] : q == 1 ? col[1] : q == 2 ? col[2] : q == 3 ? col[3] : col[4];
    };
    float add[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const MgAxis ax = F.tx[2 * h + e];
        add[e] = ax.w0 * pick(ax.i0) + ax.w1 * pick(ax.i1);
    }
    const int q0 = (jj + kk) & 1;  // colour-0 cells sit at ii = 2h + q0 and 2h + 2 + q0
    f2v x0 = *reinterpret_cast<const f2v *>(F.x + f), x1 = *reinterpret_cast<const f2v *>(F.x + F.NC + f);
    if (a0x) x0.x += add[q0];
    if (a0y) x0.y += add[2 + q0];
    if (a1x) x1.x += add[1 - q0];
    if (a1y) x1.y += add[3 - q0];
    *reinterpret_cast<f2v *>(F.x + f) = x0;
    *reinterpret_cast<f2v *>(F.x + F.NC + f) = x1;
}

// ---- the smallest levels in one workgroup, held in LDS
static constexpr int kMgSmallMax = 8;
static constexpr int kMgSmallThreads = 1024;
static constexpr int64_t kMgSmallLds = 64 * 1024;
struct MgSmall {
    MgView L[kMgSmallMax];   // global arrays of the levels (x, b used for level 0 only)
    int off[kMgSmallMax];    // LDS offset (floats) of each level's block: x, b (2*NC each), the Laplacian
                             // rows (3*sxp, 3*sy, 3*sz) and, but for the coarsest, the prolongation
                             // tables (sx + sy + sz entries, 16-byte aligned)
    int n, nu1, nu2, ncoarse;
    float om, omc;
};

// floats of one level's LDS block (a multiple of 4)
__host__ __device__ inline int mg_lds_floats(int NC, int sx, int sy, int sz, int sxp, bool tables) {
    const int rows = (3 * (sxp + sy + sz) + 3) & ~3;
    return 4 * NC + rows + (tables ? 4 * (sx + sy + sz) : 0);
}

__device__ __forceinline__ MgView lds_view(const MgSmall &S, int l, float *lds) {
    MgView v = S.L[l];
    float *p = lds + S.off[l];
    v.x = p;
    v.b = p + 2 * v.NC;
    v.r = nullptr;
    float *rows = p + 4 * v.NC;
    v.ax = rows;
    v.ay = rows + 3 * v.sxp;
    v.az = rows + 3 * (v.sxp + v.sy);
    if (l + 1 < S.n) {
        MgAxis *t = reinterpret_cast<MgAxis *>(rows + ((3 * (v.sxp + v.sy + v.sz) + 3) & ~3));
        v.tx = t;
        v.ty = t + (v.nx + 2);
        v.tz = t + (v.nx + 2) + v.sy;
    }
    return v;
}

// one colour half-sweep of red-black SOR over the whole level, then a barrier
__device__ void blk_sweep(const MgView &L, int c, bool first, float om) {
    float *mine = L.x + (int64_t)c * L.NC;
    const float *o = L.x + (int64_t)(1 - c) * L.NC, *rhs = L.b + (int64_t)c * L.NC;
    const int plane = L.H * L.sy;
    const bool no_other = first && c == 0;
    for (int f = threadIdx.x; f < L.NC; f += blockDim.x) {
        const float rr = rhs[f];
        if (!(rr == rr)) continue;
        int ii, jj, kk;
        view_decode(L, f, c, ii, jj, kk);
        const int s = ii & 1;
        const float a1x = L.ax[ii], a2x = L.ax[L.sxp + ii], a3x = L.ax[2 * L.sxp + ii];
        const float a1y = L.ay[jj], a2y = L.ay[L.sy + jj], a3y = L.ay[2 * L.sy + jj];
        const float a1z = L.az[kk], a2z = L.az[L.sz + kk], a3z = L.az[2 * L.sz + kk];
        float pw = 0.0f, pe = 0.0f, ps = 0.0f, pn = 0.0f, pb = 0.0f, pt = 0.0f;
        if (!no_other) {
            pw = o[f - 1 + s];
            pe = o[f + s];
            ps = o[f - L.H];
            pn = o[f + L.H];
            pb = o[f - plane];
            pt = o[f + plane];
        }
        const float pold = first ? 0.0f : mine[f];
        const float gs = sor_div(((((((rr - a1x * pw) - a3x * pe) - a1y * ps) - a3y * pn) - a1z * pb) - a3z * pt),
                                 ((a2x + a2y) + a2z));
        mine[f] = (1.0f - om) * pold + om * gs;
    }
    __syncthreads();
}

// The same sweeps with each thread's share of the level held in registers:
// a thread owns slots f = tid + 1024 t (t < kMgSlots) of each colour, and
// what does not change while the level is smoothed (the slot, its
// right-hand side, the neighbour offset, the six off-diagonal coefficients
// and the reciprocal of the diagonal) is loaded once per visit of the level
// instead of once per half-sweep.  The arithmetic is blk_sweep's, operand
// for operand.
static constexpr int kMgSlots = 2;   // slots per colour and thread: NC <= 2048 (larger levels: blk_sweep)
struct MgSlot {
    int f, s;                        // slot (-1: none or not a cell), ii & 1
    float rr, a1x, a3x, a1y, a3y, a1z, a3z, rd;  // rd = rcp((a2x + a2y) + a2z)
};
__device__ __forceinline__ void blk_cache(const MgView &L, MgSlot (&sl)[2][kMgSlots]) {
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int t = 0; t < kMgSlots; t++) {
            MgSlot &q = sl[c][t];
            const int f = (int)threadIdx.x + t * kMgSmallThreads;
            q.f = -1;
            if (f >= L.NC) continue;
            const float rr = L.b[(int64_t)c * L.NC + f];
            if (!(rr == rr)) continue;
            int ii, jj, kk;
            view_decode(L, f, c, ii, jj, kk);
            q.f = f;
            q.s = ii & 1;
            q.rr = rr;
            q.a1x = L.ax[ii];
            q.a3x = L.ax[2 * L.sxp + ii];
            q.a1y = L.ay[jj];
            q.a3y = L.ay[2 * L.sy + jj];
            q.a1z = L.az[kk];
            q.a3z = L.az[2 * L.sz + kk];
            q.rd = __builtin_amdgcn_rcpf((L.ax[L.sxp + ii] + L.ay[L.sy + jj]) + L.az[L.sz + kk]);
        }
}
__device__ __forceinline__ void blk_sweep_cached(const MgView &L, const MgSlot (&sl)[kMgSlots], int c, bool first,
                                                 float om) {
    float *mine = L.x + (int64_t)c * L.NC;
    const float *o = L.x + (int64_t)(1 - c) * L.NC;
    const int plane = L.H * L.sy;
    const bool no_other = first && c == 0;
#pragma unroll
    for (int t = 0; t < kMgSlots; t++) {
        const MgSlot &q = sl[t];
        if (q.f < 0) continue;
        const int f = q.f;
        float pw = 0.0f, pe = 0.0f, ps = 0.0f, pn = 0.0f, pb = 0.0f, pt = 0.0f;
        if (!no_other) {
            pw = o[f - 1 + q.s];
            pe = o[f + q.s];
            ps = o[f - L.H];
            pn = o[f + L.H];
            pb = o[f - plane];
            pt = o[f + plane];
        }
        const float pold = first ? 0.0f : mine[f];
        const float gs =
            ((((((q.rr - q.a1x * pw) - q.a3x * pe) - q.a1y * ps) - q.a3y * pn) - q.a1z * pb) - q.a3z * pt) * q.rd;
        mine[f] = (1.0f - om) * pold + om * gs;
    }
    __syncthreads();
}

__device__ void blk_resid_restrict(const MgView &F, const MgView &C) {
    for (int q = threadIdx.x; q < 2 * C.NC; q += blockDim.x) {
        const int cc = q >= C.NC ? 1 : 0;
        const int fc = q - cc * C.NC;
        int I, J, K;
        view_decode(C, fc, cc, I, J, K);
        C.b[q] = restrict_at<false>(F, C, cc, fc, I, J, K);
    }
    __syncthreads();
}

__device__ void blk_prolong(const MgView &F, const MgView &C) {
    for (int q = threadIdx.x; q < 2 * F.NC; q += blockDim.x) {
        if (!(F.b[q] == F.b[q])) continue;
        const int c = q >= F.NC ? 1 : 0;
        int ii, jj, kk;
        view_decode(F, q - c * F.NC, c, ii, jj, kk);
        F.x[q] += prolong_at(F, C, ii, jj, kk);
    }
    __syncthreads();
}

// The constant tables of a level as one run of 32-bit words: the Laplacian
// rows (3 x sxp, 3 x sy, 3 x sz) and, with `tables`, the prolongation tables
// (4 words per entry: nx+2, sy, sz entries).  Word q's global source and LDS
// destination (false past the end).
static constexpr int kMgGather = 8;
__device__ __forceinline__ int mg_table_words(const MgView &G, bool tables) {
    return 3 * (G.sxp + G.sy + G.sz) + (tables ? 4 * (G.nx + 2 + G.sy + G.sz) : 0);
}
__device__ __forceinline__ bool mg_table_word(const MgView &G, const MgView &V, bool tables, int q, const float *&src,
                                              float *&dst) {
    const int r = 3 * (G.sxp + G.sy + G.sz);
    if (q < r) {
        // the three row tables are contiguous in LDS (lds_view) and each is
        // its own global array
        const int ay = 3 * G.sxp, az = ay + 3 * G.sy;
        src = q < ay ? G.ax + q : q < az ? G.ay + (q - ay) : G.az + (q - az);
        dst = const_cast<float *>(V.ax) + q;
        return true;
    }
    if (!tables) return false;
    q -= r;
    const int ty = 4 * (G.nx + 2), tz = ty + 4 * G.sy;
    if (q >= tz + 4 * G.sz) return false;
    src = q < ty ? reinterpret_cast<const float *>(G.tx) + q
                 : q < tz ? reinterpret_cast<const float *>(G.ty) + (q - ty) : reinterpret_cast<const float *>(G.tz) + (q - tz);
    dst = reinterpret_cast<float *>(const_cast<MgAxis *>(V.tx)) + q;
    return true;
}

// V-cycle over levels S.L[0..n): x of level 0 from its b.  x and b of every
// level live in dynamic LDS for the duration (2 x 2*NC floats per level);
// S.L[0]'s global b is read at the start and its global x written at the end.
__global__ void __launch_bounds__(kMgSmallThreads) k_mg_small(MgSmall S, const double *gate, double thr) {
    if (gated(gate, thr)) return;
    extern __shared__ float lds[];
    // Everything the cycle reads from global memory (level 0's b, every
    // level's Laplacian rows and prolongation tables) is gathered in passes
    // in which each thread issues all of its loads before it stores any, so
    // one pass costs one memory latency: x is zeroed (cells outside the
    // solve hold 0) while the loads are in flight.
    const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
    {
        const MgView V = lds_view(S, 0, lds), G = S.L[0];
        for (int q0 = 0; q0 < 2 * V.NC; q0 += kMgGather * nt) {
            float v[kMgGather];
#pragma unroll
            for (int t = 0; t < kMgGather; t++) {
                const int q = min(q0 + tid + t * nt, 2 * V.NC - 1);
                v[t] = G.b[q];
            }
#pragma unroll
            for (int t = 0; t < kMgGather; t++) {
                const int q = q0 + tid + t * nt;
                if (q < 2 * V.NC) V.b[q] = v[t];
            }
        }
    }
    for (int l = 0; l < S.n; l++) {
        const MgView V = lds_view(S, l, lds);
        for (int q = tid; q < 2 * V.NC; q += nt) V.x[q] = 0.0f;
    }
    int words = 0;
    for (int l = 0; l < S.n; l++) words += mg_table_words(S.L[l], l + 1 < S.n);
    for (int w0 = 0; w0 < words; w0 += kMgGather * nt) {
        float v[kMgGather];
        const float *a[kMgGather];
        float *d[kMgGather];
        // word w of the levels' tables in order (clamped to the last): its
        // source and destination first, then all the loads, then the stores
#pragma unroll
        for (int t = 0; t < kMgGather; t++) {
            const int w = min(w0 + tid + t * nt, words - 1);
            a[t] = S.L[0].ax;
            d[t] = nullptr;
            for (int l = 0, base = 0; l < S.n; l++) {
                const int n = mg_table_words(S.L[l], l + 1 < S.n);
                if (w >= base && w < base + n)
                    mg_table_word(S.L[l], lds_view(S, l, lds), l + 1 < S.n, w - base, a[t], d[t]);
                base += n;
            }
        }
#pragma unroll
        for (int t = 0; t < kMgGather; t++) v[t] = *a[t];
#pragma unroll
        for (int t = 0; t < kMgGather; t++)
            if (w0 + tid + t * nt < words) *d[t] = v[t];
    }
    __syncthreads();
    // n sweeps on level L (the first from zero when `first`), from registers
    // when the level's slots fit kMgSlots per thread
    MgSlot sl[2][kMgSlots];
    auto smooth = [&](const MgView &L, int n, bool first, float om) {
        if (L.NC <= kMgSlots * kMgSmallThreads) {
            blk_cache(L, sl);
            for (int s = 0; s < n; s++) {
                blk_sweep_cached(L, sl[0], 0, first && s == 0, om);
                blk_sweep_cached(L, sl[1], 1, first && s == 0, om);
            }
        } else {
            for (int s = 0; s < n; s++) {
                blk_sweep(L, 0, first && s == 0, om);
                blk_sweep(L, 1, first && s == 0, om);
            }
        }
    };
    for (int l = 0; l + 1 < S.n; l++) {
        const MgView F = lds_view(S, l, lds);
        smooth(F, S.nu1, true, S.om);
        blk_resid_restrict(F, lds_view(S, l + 1, lds));
    }
    smooth(lds_view(S, S.n - 1, lds), S.ncoarse, true, S.omc);
    for (int l = S.n - 2; l >= 0; l--) {
        const MgView F = lds_view(S, l, lds);
        blk_prolong(F, lds_view(S, l + 1, lds));
        smooth(F, S.nu2, false, S.om);
    }
    const MgView V0 = lds_view(S, 0, lds);
    for (int q = threadIdx.x; q < 2 * V0.NC; q += blockDim.x) S.L[0].x[q] = V0.x[q];
}

struct MgLevel {
    int64_t nx = 0, ny = 0, nz = 0, sx = 0, sy = 0, sz = 0, H = 0, NC = 0, sxp = 0;
    float *ax = nullptr, *ay = nullptr, *az = nullptr;  // Laplacian rows (level 0: the FP32 fine tables)
    float *x = nullptr, *b = nullptr;                    // correction, right-hand side (2*NC each)
    float *r = nullptr;                                  // residual (finest level only)
    MgAxis *tx = nullptr, *ty = nullptr, *tz = nullptr;   // prolongation from the next level (sx, sy, sz entries)
    float *x2 = nullptr;                                 // second buffer of the fused sweeps
    int level = 0;
    bool owns_tables = false, owns_xb = false;
};
static std::vector<MgLevel> g_mg;
static int g_mg_small = 0;  // first level cycled inside the single-workgroup kernel (= size: none)
static int g_mg_lds = 0;    // its dynamic LDS bytes
static int64_t g_mg_key[4] = {0, 0, 0, 0};

static SorGeom level_geom(const MgLevel &L) {
    SorGeom g;
    g.H = (int)L.H;
    g.NC = (int)L.NC;
    g.sy = (int)L.sy;
    g.sx = (int)L.sx;
    g.sz = (int)L.sz;
    g.plane = (int)(L.H * L.sy);
    g.sxp = (int)L.sxp;
    g.dH = make_fastdiv(g.H);
    g.dSy = make_fastdiv(g.sy);
    return g;
}

void mg_release() {
    for (MgLevel &L : g_mg) {
        if (L.owns_tables) { (void)hipFree(L.ax); (void)hipFree(L.ay); (void)hipFree(L.az); }
        if (L.owns_xb) { (void)hipFree(L.x); (void)hipFree(L.b); }
        if (L.tx) { (void)hipFree(L.tx); (void)hipFree(L.ty); (void)hipFree(L.tz); }
        if (L.r) (void)hipFree(L.r);
        if (L.x2) (void)hipFree(L.x2);
    }
    g_mg.clear();
    g_mg_key[0] = g_mg_key[1] = g_mg_key[2] = g_mg_key[3] = 0;
}

// Laplacian rows (3 x stride floats, row col holds A(ii, col+1) at ii = 1..n)
// of an axis whose interior faces are `faces` (n+1 values): ghost faces are
// extrapolated by the end widths, centres are face midpoints, and the end
// rows are folded as coefficientMatrix does: Neumann at the low end of every
// axis and at the high end of y and z, the antisymmetric outlet fold at the
// high end of x.
static std::vector<float> mg_rows(const std::vector<double> &faces, bool x_axis, int64_t stride) {
    const int64_t n = (int64_t)faces.size() - 1;
    std::vector<double> x1(n + 4, 0.0), xp(n + 4, 0.0), A(3 * n, 0.0);
    for (int64_t i = 2; i <= n + 2; i++) x1[i] = faces[i - 2];
    x1[1] = x1[2] - (x1[3] - x1[2]);
    x1[n + 3] = x1[n + 2] + (x1[n + 2] - x1[n + 1]);
    for (int64_t i = 1; i <= n + 2; i++) xp[i] = 0.5 * (x1[i] + x1[i + 1]);
    for (int64_t i = 2; i <= n + 1; i++) {
        const double r1 = xp[i] - xp[i - 1], r2 = xp[i + 1] - xp[i], rs = r1 + r2;
        A[(i - 2) + 0 * n] = 2.0 / (r1 * rs);
        A[(i - 2) + 1 * n] = -2.0 / (r1 * r2);
        A[(i - 2) + 2 * n] = 2.0 / (r2 * rs);
    }
    A[0 + n] += A[0];
    A[0] = 0.0;
    if (x_axis)
        A[(n - 1) + n] -= A[(n - 1) + 2 * n];
    else
        A[(n - 1) + n] += A[(n - 1) + 2 * n];
    A[(n - 1) + 2 * n] = 0.0;
    std::vector<float> r(3 * stride, 0.0f);
    for (int col = 0; col < 3; col++)
        for (int64_t ii = 1; ii <= n; ii++) r[col * stride + ii] = (float)A[(ii - 1) + n * col];
    return r;
}

static bool mg_can_halve(int64_t n) { return n % 2 == 0 && n / 2 >= 2; }

// interior faces of an axis at level l (every 2^l-th fine face; x1 holds the
// fine faces at 1-based positions 2..n+2)
static std::vector<double> mg_faces(const std::vector<double> &x1, int64_t n, int l) {
    const int64_t f = int64_t(1) << l;
    std::vector<double> g(n / f + 1);
    for (int64_t j = 0; j <= n / f; j++) g[j] = x1[2 + f * j];
    return g;
}

// centres of padded cells 0..n+1 (ghost widths repeat the end widths)
static std::vector<double> mg_centres(const std::vector<double> &faces) {
    const int64_t n = (int64_t)faces.size() - 1;
    std::vector<double> c(n + 2);
    for (int64_t i = 1; i <= n; i++) c[i] = 0.5 * (faces[i - 1] + faces[i]);
    c[0] = faces[0] - 0.5 * (faces[1] - faces[0]);
    c[n + 1] = faces[n] + 0.5 * (faces[n] - faces[n - 1]);
    return c;
}

// linear interpolation weights from coarse centres to the fine centres of
// one axis; `outlet`: the high-end ghost holds minus the end cell's value
static std::vector<MgAxis> mg_axis(const std::vector<double> &ffaces, const std::vector<double> &cfaces,
                                   bool outlet, int64_t size) {
    const int64_t nf = (int64_t)ffaces.size() - 1, nc = (int64_t)cfaces.size() - 1;
    const std::vector<double> xf = mg_centres(ffaces), xc = mg_centres(cfaces);
    std::vector<MgAxis> t(size, MgAxis{1, 1, 0.0f, 0.0f});
    for (int64_t i = 1; i <= nf; i++) {
        const int64_t I = (i + 1) >> 1;
        const int64_t N = xf[i] < xc[I] ? I - 1 : I + 1;  // the other coarse centre bracketing xf[i]
        const double w = (xf[i] - xc[N]) / (xc[I] - xc[N]); // weight of I
        MgAxis a;
        a.i0 = (int)I;
        a.w0 = (float)w;
        if (N >= 1 && N <= nc) {
            a.i1 = (int)N;
            a.w1 = (float)(1.0 - w);
        } else if (N > nc && outlet) {
            a.i1 = (int)I;
            a.w1 = (float)(-(1.0 - w));
        } else {
            a.i1 = (int)I;
            a.w1 = (float)(1.0 - w);
        }
        t[i] = a;
    }
    return t;
}

static int mg_setup(Ctx &c) {
    const int64_t key[4] = {c.nx, c.ny, c.nz, (int64_t)c.mg_levels};
    if (!g_mg.empty() && key[0] == g_mg_key[0] && key[1] == g_mg_key[1] && key[2] == g_mg_key[2] &&
        key[3] == g_mg_key[3])
        return 0;
    IB3D_CHECK(hipStreamSynchronize(c.stream));
    mg_release();
    if (sor_float_tables(c)) return -1;
    MgLevel L0;
    L0.nx = c.nx; L0.ny = c.ny; L0.nz = c.nz;
    L0.sx = c.sx; L0.sy = c.sy; L0.sz = c.sz;
    L0.H = c.H; L0.NC = c.NC; L0.sxp = c.sxp;
    L0.ax = g_axf; L0.ay = g_ayf; L0.az = g_azf;
    g_mg.push_back(L0);
    const int max_levels = c.mg_levels > 0 ? c.mg_levels : 32;
    while ((int)g_mg.size() < max_levels) {
        const MgLevel &F = g_mg.back();
        if (!(mg_can_halve(F.nx) && mg_can_halve(F.ny) && mg_can_halve(F.nz))) break;
        MgLevel L;
        L.nx = F.nx / 2; L.ny = F.ny / 2; L.nz = F.nz / 2;
        L.sx = L.nx + 2; L.sy = L.ny + 2; L.sz = L.nz + 2;
        L.H = (((L.sx + 1) / 2) + 1) & ~int64_t(1);
        L.NC = L.sy * L.sz * L.H;
        L.sxp = (std::max<int64_t>(L.sx, 2 * L.H + 2) + 7) & ~int64_t(7);  // as the finest level's rows
        g_mg.push_back(L);
    }
    IB3D_REQUIRE(g_mg.size() >= 2, "poissonSolver: the grid cannot be coarsened for multigrid (each interior "
                                   "dimension must be even and at least 4)");
    const int nl = (int)g_mg.size();
    // the coarsest levels whose x and b fit the LDS budget (at most kMgSmallMax of them) are cycled by
    // one workgroup; the finest level never is
    int lds_max = 0;
    IB3D_CHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c.device));
    const int64_t budget = std::min<int64_t>(lds_max, kMgSmallLds);
    g_mg_small = nl;
    int64_t bytes = 0;
    while (g_mg_small > 1 && nl - (g_mg_small - 1) <= kMgSmallMax) {
        const MgLevel &L = g_mg[g_mg_small - 1];
        const int64_t b = 4 * (int64_t)mg_lds_floats((int)L.NC, (int)L.sx, (int)L.sy, (int)L.sz, (int)L.sxp,
                                                     g_mg_small - 1 < nl - 1);
        if (bytes + b > budget) break;
        g_mg_small--;
        bytes += b;
    }
    g_mg_lds = (int)bytes;
    for (int l = 0; l < nl; l++) {
        MgLevel &L = g_mg[l];
        if (check_fastdiv((int)L.H, L.NC + 1) || check_fastdiv((int)L.sy, L.NC / L.H + 1)) {
            set_error("poissonSolver: fast division self-check failed on a multigrid level");
            return -1;
        }
        if (l == 0) {  // the finest residual is stored by the z-marching kernel (rows it skips stay NaN)
            IB3D_CHECK(hipMalloc(&L.r, 2 * L.NC * sizeof(float)));
            hipLaunchKernelGGL(k_fill_nan_f32, dim3(nblocks(2 * L.NC, kSorBlock)), dim3(kSorBlock), 0, c.stream,
                               L.r, 2 * L.NC);
        }
        if (l < nl - 1) {  // the next level's correction comes back through these
            // 2H >= sx x entries: the finest level's prolongation reads whole slot pairs
            const std::vector<MgAxis> tx =
                mg_axis(mg_faces(c.hx1, c.nx, l), mg_faces(c.hx1, c.nx, l + 1), true, std::max(L.sx, 2 * L.H));
            const std::vector<MgAxis> ty = mg_axis(mg_faces(c.hy1, c.ny, l), mg_faces(c.hy1, c.ny, l + 1), false, L.sy);
            const std::vector<MgAxis> tz = mg_axis(mg_faces(c.hz1, c.nz, l), mg_faces(c.hz1, c.nz, l + 1), false, L.sz);
            IB3D_CHECK(hipMalloc(&L.tx, tx.size() * sizeof(MgAxis)));
            IB3D_CHECK(hipMalloc(&L.ty, ty.size() * sizeof(MgAxis)));
            IB3D_CHECK(hipMalloc(&L.tz, tz.size() * sizeof(MgAxis)));
            IB3D_CHECK(h2d(L.tx, tx.data(), tx.size() * sizeof(MgAxis)));
            IB3D_CHECK(h2d(L.ty, ty.data(), ty.size() * sizeof(MgAxis)));
            IB3D_CHECK(h2d(L.tz, tz.data(), tz.size() * sizeof(MgAxis)));
        }
        if (l == 0) continue;
        const std::vector<float> rx = mg_rows(mg_faces(c.hx1, c.nx, l), true, L.sxp);
        const std::vector<float> ry = mg_rows(mg_faces(c.hy1, c.ny, l), false, L.sy);
        const std::vector<float> rz = mg_rows(mg_faces(c.hz1, c.nz, l), false, L.sz);
        IB3D_CHECK(hipMalloc(&L.ax, rx.size() * 4));
        IB3D_CHECK(hipMalloc(&L.ay, ry.size() * 4));
        IB3D_CHECK(hipMalloc(&L.az, rz.size() * 4));
        L.owns_tables = true;
        IB3D_CHECK(h2d(L.ax, rx.data(), rx.size() * 4));
        IB3D_CHECK(h2d(L.ay, ry.data(), ry.size() * 4));
        IB3D_CHECK(h2d(L.az, rz.data(), rz.size() * 4));
        IB3D_CHECK(hipMalloc(&L.x, 2 * L.NC * sizeof(float)));
        IB3D_CHECK(hipMalloc(&L.b, 2 * L.NC * sizeof(float)));
        L.owns_xb = true;
        IB3D_CHECK(dzero(L.x, 2 * L.NC * sizeof(float)));   // cells outside the solve stay 0
        hipLaunchKernelGGL(k_fill_nan_f32, dim3(nblocks(2 * L.NC, kSorBlock)), dim3(kSorBlock), 0, c.stream,
                           L.b, 2 * L.NC);
    }
    // the fused sweeps' second buffer on every level smoothed by launches:
    // cells they never write (walls, and whatever lies outside the solve)
    // stay 0, as in x
    for (int l = 0; l < (int)g_mg.size(); l++) g_mg[l].level = l;
    for (int l = 0; l < (int)g_mg.size(); l++) {
        MgLevel &L = g_mg[l];
        if (l > 0 && l >= g_mg_small) break;
        IB3D_CHECK(hipMalloc(&L.x2, 2 * L.NC * sizeof(float)));
        IB3D_CHECK(dzero(L.x2, 2 * L.NC * sizeof(float)));
    }
    IB3D_CHECK(hipGetLastError());
    for (int q = 0; q < 4; q++) g_mg_key[q] = key[q];
    return 0;
}

static dim3 level_grid(const MgLevel &L) {
    return dim3(nblocks((L.H * L.sy) / 2, kSorBlock), (unsigned)((L.sz - 2 + kZch - 1) / kZch));
}

// n red-black SOR sweeps on level L (the first from zero when `first`);
// z-marching cell pairs on the finest level, one cell per thread below
// Launch shape of the fused sweep: S own slots per workgroup (the red range
// adds H on either side and must fit 2 x kSweepThreads slots), planes per
// chunk so that the grid holds about g_sweep_blocks workgroups.
static int g_fused = 2;          // levels (from the finest) smoothed by whole sweeps in one launch
static int g_coarse_tiles = 2;   // first level smoothed by two temporally blocked launches per phase (0: none)
static int g_fused_restrict = 64;  // levels (from the finest) whose residual is restricted in one pass
static int g_prolong_pz = 2;       // levels (from the finest) prolonged four cells of a row per thread
static int g_sweep_tz = 0;        // planes per chunk on the finest level (0: from g_sweep_blocks)
static int g_sweep_blocks = 480;
static int g_sweep_blocks_coarse = 192;  // workgroup target below the finest level

static bool sweep2_shape(const MgLevel &L, int &S, int &tz, dim3 &grid) {
    const int64_t plane = L.H * L.sy, smax = 2 * kSweepThreads - 2 * L.H;
    if (smax < 64) return false;
    const int64_t nbp = (plane + smax - 1) / smax;
    S = (int)(((plane + nbp - 1) / nbp + 1) & ~int64_t(1));
    const int64_t nz = L.sz - 2;
    const int64_t target = L.level == 0 ? g_sweep_blocks : g_sweep_blocks_coarse;
    int64_t nch = (g_sweep_tz > 0 && L.level == 0) ? (nz + g_sweep_tz - 1) / g_sweep_tz
                                                   : std::min<int64_t>(nz, std::max<int64_t>(1, (target + nbp - 1) / nbp));
    tz = (int)((nz + nch - 1) / nch);
    nch = (nz + tz - 1) / tz;
    grid = dim3((unsigned)nbp, (unsigned)nch);
    return S + 2 * L.H <= 2 * kSweepThreads;
}

static void mg_smooth(Ctx &c, const MgLevel &L, bool finest, int n, bool first, double omega, const double *gate) {
    const SorGeom g = level_geom(L);
    const float om = (float)omega, omm = (float)(1.0 - omega);
    const dim3 cg(nblocks(L.NC, kSorBlock));
    int S = 0, tz = 0;
    dim3 sg;
    if (!finest && g_coarse_tiles > 0 && L.level >= g_coarse_tiles && L.x2 && n >= 2 &&
        (n + 1) / 2 <= kTileMaxSweeps) {
        // two launches of ceil(n/2) and floor(n/2) sweeps: x -> x2 -> x, on
        // 8 x 8 x 8 blocks
        constexpr int T = kTileCube;
        const dim3 tg((unsigned)((L.sx - 2 + T - 1) / T), (unsigned)((L.sy - 2 + T - 1) / T),
                      (unsigned)((L.sz - 2 + T - 1) / T));
        auto launch = [&](int m, bool fz, const float *from, float *to) {
#define IB3D_TILE(F, MM)                                                                                            \
    hipLaunchKernelGGL((k_smooth_tile<F, MM, T, T, T>), tg, dim3(kTileThreads), 0, c.stream, from, to, L.b, L.ax,    \
                       L.ay, L.az, g, om, omm, gate, 0.5)
            if (m == 2) { if (fz) IB3D_TILE(true, 2); else IB3D_TILE(false, 2); }
            else { if (fz) IB3D_TILE(true, 1); else IB3D_TILE(false, 1); }
#undef IB3D_TILE
        };
        launch((n + 1) / 2, first, L.x, L.x2);
        launch(n / 2, false, L.x2, L.x);
        return;
    }
    if (L.level < g_fused && L.x2 && n >= 2 && sweep2_shape(L, S, tz, sg)) {
        // an odd count does its first sweep in place, so that the pairs of
        // fused sweeps end in x
        if (n % 2) {
            mg_smooth(c, L, finest, 1, first, omega, gate);
            n -= 1;
            first = false;
        }
        float *a = L.x, *b = L.x2;
        for (int q = 0; q < n; q++) {
            if (first && q == 0)
                hipLaunchKernelGGL(k_sweep2_pz<true>, sg, dim3(kSweepThreads), 0, c.stream, a, a + L.NC, b,
                                   b + L.NC, L.b, L.b + L.NC, L.ax, L.ay, L.az, g, S, tz, om, omm, gate, 0.5);
            else
                hipLaunchKernelGGL(k_sweep2_pz<false>, sg, dim3(kSweepThreads), 0, c.stream, a, a + L.NC, b,
                                   b + L.NC, L.b, L.b + L.NC, L.ax, L.ay, L.az, g, S, tz, om, omm, gate, 0.5);
            std::swap(a, b);
        }
        return;
    }
    for (int s = 0; s < n; s++) {
        for (int col = 0; col < 2; col++) {
            float *mine = L.x + col * L.NC;
            const float *other = L.x + (1 - col) * L.NC, *rhs = L.b + col * L.NC;
            if (!finest) {
                if (first && s == 0)
                    hipLaunchKernelGGL((k_sor<float, false, true, false>), cg, dim3(kSorBlock), 0, c.stream, mine,
                                       other, rhs, L.ax, L.ay, L.az, g, col, om, omm, nullptr, gate, 0.5, nullptr);
                else
                    hipLaunchKernelGGL((k_sor<float, false, false, false>), cg, dim3(kSorBlock), 0, c.stream, mine,
                                       other, rhs, L.ax, L.ay, L.az, g, col, om, omm, nullptr, gate, 0.5, nullptr);
            } else if (first && s == 0)
                hipLaunchKernelGGL((k_sor_pz<float, false, true, false>), level_grid(L), dim3(kSorBlock), 0, c.stream,
                                   mine, other, rhs, L.ax, L.ay, L.az, g, col, om, omm, nullptr, gate, 0.5, nullptr);
            else
                hipLaunchKernelGGL((k_sor_pz<float, false, false, false>), level_grid(L), dim3(kSorBlock), 0,
                                   c.stream, mine, other, rhs, L.ax, L.ay, L.az, g, col, om, omm, nullptr, gate, 0.5,
                                   nullptr);
        }
    }
}

static MgView level_view(const MgLevel &L) {
    MgView V;
    V.x = L.x; V.b = L.b; V.r = L.r;
    V.ax = L.ax; V.ay = L.ay; V.az = L.az;
    V.tx = L.tx; V.ty = L.ty; V.tz = L.tz;
    V.nx = (int)L.nx; V.ny = (int)L.ny; V.nz = (int)L.nz;
    V.H = (int)L.H; V.NC = (int)L.NC; V.sy = (int)L.sy; V.sz = (int)L.sz; V.sxp = (int)L.sxp;
    V.dH = make_fastdiv(V.H);
    V.dSy = make_fastdiv(V.sy);
    return V;
}

// pre- (post = 0) or post-smoothing (post = 1) sweeps on level l cycled by
// launches: the finest level's, a per-level override, or the coarse levels'
static int level_sweeps(const Ctx &c, int l, int post) {
    if (l == 0) return post ? c.mg_nu2 : c.mg_nu1;
    if (l < 8 && c.mg_nu_lv[l][post] >= 0) return c.mg_nu_lv[l][post];
    return post ? c.mg_nu2c : c.mg_nu1c;
}

// z = M^-1 v: one V-cycle from a zero guess (v and z are 2*NC fine vectors).
// Levels from g_mg_small on are cycled by one single-workgroup launch.
static void mg_vcycle(Ctx &c, const float *v, float *z, const double *gate) {
    const int nl = (int)g_mg.size(), ns = g_mg_small;
    g_mg[0].b = const_cast<float *>(v);
    g_mg[0].x = z;
    const dim3 tb(kSorBlock);
    const int top = std::min(ns, nl - 1);  // levels [0, top) restrict through launches
    for (int l = 0; l < top; l++) {
        const MgLevel &F = g_mg[l], &C = g_mg[l + 1];
        mg_smooth(c, F, l == 0, level_sweeps(c, l, 0), true, c.mg_omega, gate);
        if (l < g_fused_restrict) {
            const int W = (int)(2 * C.H), nbx = nblocks(W * C.sy, kSorBlock), nk = (int)C.sz - 2;
            // coarse planes per chunk: about 1536 workgroups in all
            const int kch = std::max(1, std::min(nk, (nk * nbx + 1535) / 1536));
            hipLaunchKernelGGL(k_resid_restrict_pz, dim3(nbx, (nk + kch - 1) / kch), tb, 0, c.stream, level_view(F),
                               level_view(C), make_fastdiv(W), kch, gate, 0.5);
        } else if (l == 0) {
            hipLaunchKernelGGL(k_matvec32_pz<true>, level_grid(F), tb, 0, c.stream, F.x, F.x + F.NC, F.b,
                               F.b + F.NC, F.r, F.r + F.NC, F.ax, F.ay, F.az, level_geom(F), gate, 0.5);
            hipLaunchKernelGGL(k_mg_resid_restrict<true>, dim3(nblocks(2 * C.NC, kSorBlock)), tb, 0, c.stream,
                               level_view(F), level_view(C), level_geom(C), gate, 0.5);
        } else {
            hipLaunchKernelGGL(k_mg_resid_restrict<false>, dim3(nblocks(2 * C.NC, kSorBlock)), tb, 0, c.stream,
                               level_view(F), level_view(C), level_geom(C), gate, 0.5);
        }
    }
    if (ns < nl) {
        MgSmall S;
        S.n = nl - ns;
        S.nu1 = c.mg_nu_small[0] >= 0 ? c.mg_nu_small[0] : c.mg_nu1c;
        S.nu2 = c.mg_nu_small[1] >= 0 ? c.mg_nu_small[1] : c.mg_nu2c;
        S.ncoarse = c.mg_coarse;
        S.om = (float)c.mg_omega;
        S.omc = (float)c.mg_coarse_omega;
        int off = 0;
        for (int l = ns; l < nl; l++) {
            const MgLevel &L = g_mg[l];
            S.L[l - ns] = level_view(L);
            S.off[l - ns] = off;
            off += mg_lds_floats((int)L.NC, (int)L.sx, (int)L.sy, (int)L.sz, (int)L.sxp, l < nl - 1);
        }
        hipLaunchKernelGGL(k_mg_small, dim3(1), dim3(kMgSmallThreads), g_mg_lds, c.stream, S, gate, 0.5);
    } else {
        mg_smooth(c, g_mg[nl - 1], false, c.mg_coarse, true, c.mg_coarse_omega, gate);
    }
    for (int l = top - 1; l >= 0; l--) {
        const MgLevel &F = g_mg[l], &C = g_mg[l + 1];
        if (l < g_prolong_pz)
            hipLaunchKernelGGL(k_mg_prolong_pz, dim3(nblocks((F.H * F.sy) / 2, kSorBlock), (unsigned)(F.sz - 2)), tb,
                               0, c.stream, level_view(F), level_view(C), level_geom(F),
                               l == 0 ? (const uint32_t *)g_actbits : nullptr, g_actbits_nw, gate, 0.5);
        else
            hipLaunchKernelGGL(k_mg_prolong, dim3(nblocks(2 * F.NC, kSorBlock)), tb, 0, c.stream, level_view(F),
                               level_view(C), level_geom(F), gate, 0.5);
        mg_smooth(c, F, l == 0, level_sweeps(c, l, 1), false, c.mg_omega, gate);
    }
}


// g_kh: H ((m+1) x m, column-major), beta, y[m], cs[m], sn[m], g[m+1], act[m+1], nrm[m]
static int64_t kh_size(int m) { return (int64_t)(m + 1) * m + 1 + 3 * m + 2 * (m + 1) + m; }

static int gmres_buffers(Ctx &c, int m) {
    if (g_kv && g_k_nc == c.NC && g_k_m == m) return 0;
    IB3D_CHECK(hipStreamSynchronize(c.stream));
    if (g_kv) { (void)hipFree(g_kv); (void)hipFree(g_kz); (void)hipFree(g_kh); (void)hipFree(g_kpart); (void)hipFree(g_kpart2); }
    const int64_t stride = 2 * c.NC;
    IB3D_CHECK(hipMalloc(&g_kv, (int64_t)(m + 1) * stride * sizeof(float)));
    // the matrix-vector kernel writes interior rows only: every other entry
    // of a basis vector must read as "not a cell" (NaN) in the vector kernels
    hipLaunchKernelGGL(k_fill_nan_f32, dim3(nblocks((int64_t)(m + 1) * stride, kSorBlock)), dim3(kSorBlock), 0,
                       c.stream, g_kv, (int64_t)(m + 1) * stride);
    IB3D_CHECK(hipMalloc(&g_kz, (int64_t)m * stride * sizeof(float)));
    IB3D_CHECK(dzero(g_kz, (int64_t)m * stride * sizeof(float)));   // cells the sweeps never write stay 0
    IB3D_CHECK(hipMalloc(&g_kh, kh_size(m) * sizeof(double)));
    IB3D_CHECK(dzero(g_kh, kh_size(m) * sizeof(double)));
    IB3D_CHECK(hipMalloc(&g_kpart, (int64_t)kMaxKrylov * kVecBlocks * sizeof(double)));
    IB3D_CHECK(hipMalloc(&g_kpart2, (int64_t)kVecBlocks * sizeof(double)));
    g_k_nc = c.NC;
    g_k_m = m;
    g_z_geom = -1;
    return 0;
}

void gmres_release() {
    if (g_kv) { (void)hipFree(g_kv); (void)hipFree(g_kz); (void)hipFree(g_kh); (void)hipFree(g_kpart); (void)hipFree(g_kpart2); }
    if (g_actbits) (void)hipFree(g_actbits);
    g_actbits = nullptr;
    g_actbits_nw = 0;
    g_kv = g_kz = nullptr;
    g_kh = g_kpart = g_kpart2 = nullptr;
    g_k_nc = 0;
    g_k_m = 0;
}

// One refinement step (outer iteration): FP64 residual r (as FP32 in r32),
// then up to m flexible-GMRES steps in FP32, each preconditioned by k SOR
// sweeps (mode 3) or one multigrid V-cycle (mode 4), stopping early once the
// least-squares residual is at most `target`, then pc += Z y in FP64.  All
// gated on the previous step's convergence flag.
//
// check_from: after each Arnoldi step j >= check_from but the last, the host
// reads act[j+1] and stops queueing the cycle once it is clear.  The steps
// after convergence would exit at once, but each still costs the host the
// launches of a V-cycle (about 40, several microseconds each); one round trip
// costs far less.  The rest of the cycle (k_lsq, the update) is queued as
// always and reads act[] itself.
static constexpr int kActMirror = 62;  // c.hscal slot of the checked flag
static int gmres_cycle(Ctx &c, int m, int k, const double *gate, double target, const double *targetp = nullptr,
                       int check_from = 1 << 30) {
    const SorGeom g = geom(c);
    const int64_t NC = c.NC, stride = 2 * NC, n2 = NC;  // n2: float pairs per vector
    double *H = g_kh, *beta = H + (int64_t)(m + 1) * m, *y = beta + 1, *cs = y + m, *sn = cs + m, *gv = sn + m,
           *act = gv + (m + 1), *nrm = act + (m + 1);
    const dim3 vb(kVecBlocks), tb(kSorBlock);
    const float om = (float)c.omega, omm = (float)(1.0 - c.omega);
    float *V0 = g_kv;
    hipLaunchKernelGGL(k_act_init, dim3(1), dim3(64), 0, c.stream, act, m, gate, 0.5);
    hipLaunchKernelGGL(k_norm2, vb, tb, 0, c.stream, c.r32[0], n2, g_kpart2, act, 0.5);
    hipLaunchKernelGGL(k_sqrt_sum, dim3(1), dim3(256), 0, c.stream, g_kpart2, kVecBlocks, beta, act, 0.5);
    hipLaunchKernelGGL(k_scale_to, vb, tb, 0, c.stream, V0, c.r32[0], beta, n2, act, 0.5);
    for (int j = 0; j < m; j++) {
        const double *aj = act + j;
        float *Z = g_kz + j * stride, *Vj = g_kv + j * stride, *Vn = g_kv + (j + 1) * stride;
        if (c.mode == 4) {
            mg_vcycle(c, Vj, Z, aj);
        } else {
            for (int sw = 0; sw < k; sw++) {
                for (int col = 0; col < 2; col++) {
                    if (sw == 0)
                        hipLaunchKernelGGL((k_sor_pz<float, false, true, false>), pz_grid(c), tb, 0, c.stream,
                                           Z + col * NC, Z + (1 - col) * NC, Vj + col * NC, g_axf, g_ayf, g_azf, g,
                                           col, om, omm, nullptr, aj, 0.5, nullptr);
                    else
                        hipLaunchKernelGGL((k_sor_pz<float, false, false, false>), pz_grid(c), tb, 0, c.stream,
                                           Z + col * NC, Z + (1 - col) * NC, Vj + col * NC, g_axf, g_ayf, g_azf, g,
                                           col, om, omm, nullptr, aj, 0.5, nullptr);
                }
            }
        }
        hipLaunchKernelGGL(k_matvec32_pz<false>, pz_grid(c), tb, 0, c.stream, Z, Z + NC, V0, V0 + NC, Vn, Vn + NC,
                           g_axf, g_ayf, g_azf, g, aj, 0.5, (const uint32_t *)g_actbits, g_actbits_nw);
        double *hj = H + (int64_t)(m + 1) * j;
        launch_dots(j + 1, vb, tb, c.stream, Vn, g_kv, stride, n2 / 2, g_kpart, aj);
        hipLaunchKernelGGL(k_rows_sum, dim3(j + 1), dim3(256), 0, c.stream, g_kpart, kVecBlocks, hj, 0, aj, 0.5);
        launch_orth(j + 1, vb, tb, c.stream, Vn, g_kv, stride, hj, n2 / 2, g_kpart2, aj);
        // the new norm (from orth's partials) and the rotation in one launch
        hipLaunchKernelGGL(k_givens, dim3(1), dim3(256), 0, c.stream, H, m, j, cs, sn, gv, beta, act, target,
                           targetp, (const double *)g_kpart2, kVecBlocks, nrm);
        if (j + 1 < m) hipLaunchKernelGGL(k_scale_to, vb, tb, 0, c.stream, Vn, Vn, nrm + j, n2, aj, 0.5);
        if (j >= check_from && j + 1 < m) {
            IB3D_CHECK(hipMemcpyAsync(c.hscal + kActMirror, act + j + 1, sizeof(double), hipMemcpyDeviceToHost,
                                      c.stream));
            IB3D_CHECK(hipStreamSynchronize(c.stream));
            if (!(c.hscal[kActMirror] >= 0.5)) break;
        }
    }
    hipLaunchKernelGGL(k_lsq, dim3(1), dim3(64), 0, c.stream, H, m, gv, act, y, c.scalars + 24);
    hipLaunchKernelGGL(k_update_pc, vb, tb, 0, c.stream, c.pc[0], c.pc[1], g_kz, stride, m, y,
                       (const uint32_t *)g_actbits, g_actbits_nw, NC, act, 0.5);
    return 0;
}

// Cells outside the solve must read 0 in the preconditioned vectors Z and
// in the multigrid corrections: the sweeps never write a pair of such cells,
// so after the geometry changes (a moving body) values left from the old
// geometry would act as non-zero boundary values.  Clear them once per
// geometry (g_z_geom: the geometry they were cleared for).
static int clear_for_geometry(Ctx &c, int m) {
    if (g_z_geom == c.geom_version) return 0;
    IB3D_CHECK(dzero(g_kz, (int64_t)m * 2 * c.NC * sizeof(float)));
    for (size_t l = 1; l < g_mg.size(); l++)
        if (g_mg[l].x) IB3D_CHECK(dzero(g_mg[l].x, 2 * g_mg[l].NC * sizeof(float)));
    // the slot flags of this geometry, from the NaN sentinel of b/dt (set
    // by computeDiv from the tags)
    const int64_t nw = (c.NC + 63) / 64 * 2;
    if (nw != g_actbits_nw) {
        if (g_actbits) (void)hipFree(g_actbits);
        g_actbits = nullptr;
        g_actbits_nw = 0;
        IB3D_CHECK(hipMalloc(&g_actbits, 2 * nw * sizeof(uint32_t)));
        g_actbits_nw = nw;
    }
    hipLaunchKernelGGL(k_active_bits, dim3(nblocks(2 * nw * 32, kSorBlock)), dim3(kSorBlock), 0, c.stream, c.bdt[0],
                       c.bdt[1], c.NC, g_actbits, nw);
    IB3D_CHECK(hipGetLastError());
    g_z_geom = c.geom_version;
    return 0;
}

// One outer iteration of GMRES-IR (the Krylov cycle, the FP64 residual and
// its norm) as a captured HIP graph, replayed once per iteration in place of
// its about 330 launches (ib3d_set_graphs; off by default: at 256x192x192
// the step time is the same either way).  The graph is static, so everything that
// varies from one iteration or solve to the next lives in device slots
// S = scalars[32..38]: S[0] the previous iteration's relative residual,
// S[1], S[2] this iteration's (rel, conv), S[3] the gate (conv of the
// previous iteration), S[4] the iteration index o, S[5] ||b/dt||, S[6] the
// Krylov target.  k_cycle_end files (rel, conv) under o in the solve's
// arrays, as the launches of the uncaptured loop do, and moves the gate on.
static constexpr int kCycSlot = 32;
__global__ void k_cycle_init(double *S, const double *relv, const double *conv, double bn, double target) {
    if (threadIdx.x != 0) return;
    S[0] = relv[0];
    S[3] = conv[0];
    S[4] = 1.0;
    S[5] = bn;
    S[6] = target;
}
__global__ void k_cycle_end(double *S, double *relv, double *conv) {
    if (threadIdx.x != 0) return;
    const int64_t o = (int64_t)S[4];
    relv[o] = S[1];
    conv[o] = S[2];
    S[0] = S[1];
    S[3] = S[2];
    S[4] = (double)(o + 1);
}

static hipGraphExec_t g_cyc_exec = nullptr;
static std::vector<uint64_t> g_cyc_key;

void cycle_graph_release() {
    if (g_cyc_exec) (void)hipGraphExecDestroy(g_cyc_exec);
    g_cyc_exec = nullptr;
    g_cyc_key.clear();
}

// (re)captures the iteration graph when any buffer or parameter it bakes in
// has changed
static int cycle_graph(Ctx &c, int m, int k, double *relv, double *conv) {
    std::vector<uint64_t> key;
    auto P = [&](const void *q) { key.push_back((uint64_t)(uintptr_t)q); };
    auto D = [&](double v) {
        uint64_t u;
        std::memcpy(&u, &v, 8);
        key.push_back(u);
    };
    for (const void *q : {(const void *)c.pc[0], (const void *)c.pc[1], (const void *)c.r32[0],
                          (const void *)c.r32[1], (const void *)c.e32[0], (const void *)c.e32[1],
                          (const void *)c.bdt[0], (const void *)c.bdt[1], (const void *)c.ax, (const void *)c.ay,
                          (const void *)c.az, (const void *)c.partial, (const void *)c.scalars, (const void *)relv,
                          (const void *)conv, (const void *)g_axf, (const void *)g_ayf, (const void *)g_azf,
                          (const void *)g_kv, (const void *)g_kz, (const void *)g_kh, (const void *)g_kpart,
                          (const void *)g_kpart2, (const void *)g_actbits})
        P(q);
    for (const MgLevel &L : g_mg)
        for (const void *q : {(const void *)L.x, (const void *)L.b, (const void *)L.r, (const void *)L.ax,
                              (const void *)L.ay, (const void *)L.az, (const void *)L.tx, (const void *)L.ty,
                              (const void *)L.tz})
            P(q);
    key.push_back((uint64_t)m);
    key.push_back((uint64_t)k);
    key.push_back((uint64_t)c.mode);
    key.push_back((uint64_t)c.NC);
    key.push_back((uint64_t)c.mg_nu1);
    key.push_back((uint64_t)c.mg_nu2);
    key.push_back((uint64_t)c.mg_nu1c);
    key.push_back((uint64_t)c.mg_nu2c);
    for (const auto &lv : c.mg_nu_lv) {
        key.push_back((uint64_t)(int64_t)lv[0]);
        key.push_back((uint64_t)(int64_t)lv[1]);
    }
    key.push_back((uint64_t)(int64_t)c.mg_nu_small[0]);
    key.push_back((uint64_t)(int64_t)c.mg_nu_small[1]);
    key.push_back((uint64_t)g_fused);
    key.push_back((uint64_t)g_coarse_tiles);
    key.push_back((uint64_t)g_fused_restrict);
    key.push_back((uint64_t)g_sweep_tz);
    key.push_back((uint64_t)g_sweep_blocks);
    key.push_back((uint64_t)g_sweep_blocks_coarse);
    key.push_back((uint64_t)c.mg_coarse);
    key.push_back((uint64_t)g_mg_small);
    key.push_back((uint64_t)g_mg_lds);
    D(c.omega);
    D(c.tol);
    D(c.mg_omega);
    D(c.mg_coarse_omega);
    if (g_cyc_exec && key == g_cyc_key) return 0;
    cycle_graph_release();
    double *S = c.scalars + kCycSlot;
    IB3D_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    if (gmres_cycle(c, m, k, S + 3, 0.0, S + 6)) return -1;
    residual_async(c, false, true, 0.0, S + 1, S + 2, S + 3, S + 5);
    hipLaunchKernelGGL(k_cycle_end, dim3(1), dim3(64), 0, c.stream, S, relv, conv);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(c.stream, &graph);
    IB3D_CHECK(ec);
    const hipError_t ei = hipGraphInstantiate(&g_cyc_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) g_cyc_exec = nullptr;
    IB3D_CHECK(ei);
    g_cyc_key = key;
    return 0;
}

// batch results that fit the pinned host mirror (c.hscal[16 .. 16 + 2 x 14 + 2])
static constexpr int64_t kBatchMirror = 14;

// dnorm (device, or null): ||b/dt|| and the Krylov target, when the host has
// not read the norm (bn is then unused)
static int sor_gmres(Ctx &c, double bn, int64_t *sweeps, int64_t *outer, double *rel, double *rel0,
                     const double *dnorm) {
    const int m = c.krylov_m, k = c.mode == 4 ? 1 : c.check;
    if (sor_float_tables(c)) return -1;
    if (gmres_buffers(c, m)) return -1;
    if (c.mode == 4 && mg_setup(c)) return -1;
    if (clear_for_geometry(c, m)) return -1;
    if (ensure_partial(nblocks(2 * c.NC, kSorBlock))) return -1;
    const int64_t max_outer = std::max<int64_t>(1, c.max_sweeps / ((int64_t)m * k));
    if (ensure_slots(2 * (max_outer + 2))) return -1;
    double *relv = c.slots, *conv = c.slots + c.n_slots / 2;
    // (c.scalars[24]: the Arnoldi steps run in this solve, cleared here)
    hipLaunchKernelGGL(k_clear_dead_bits, dim3(nblocks(2 * g_actbits_nw, kSorBlock)), dim3(kSorBlock), 0, c.stream,
                       c.pc[0], c.pc[1], c.e32[0], c.e32[1], (const uint32_t *)g_actbits, g_actbits_nw, c.NC,
                       c.scalars + 24);
    const bool cnn = cnn_active();
    if (cnn) {
        g_bn = bn;
        residual_async(c, false, true, bn, c.scalars + 16, c.scalars + 17, nullptr);
        IB3D_CHECK(hipGetLastError());
        if (cnn_before_solve(c, c.scalars + 16, c.scalars + 18, resid_rel)) return -1;
    }
    residual_async(c, false, true, bn, relv, conv, nullptr, dnorm);
    int64_t o = 0, seen = 0;
    // A cycle is hundreds of launches; one queued after convergence still
    // costs their launch overhead.  Queue one cycle less than the previous
    // solve needed, then one at a time (a host check costs far less than a
    // cycle).
    int batch = (int)std::min<int64_t>(16, std::max<int64_t>(1, c.last_outer - 1));
    std::vector<double> hr;
    if (c.use_graphs) {
        if (cycle_graph(c, m, k, relv, conv)) return -1;
        hipLaunchKernelGGL(k_cycle_init, dim3(1), dim3(64), 0, c.stream, c.scalars + kCycSlot, relv, conv, bn,
                           c.krylov_target * c.tol * bn);
    }
    for (;;) {
        for (int q = 0; q < batch && o < max_outer; q++) {
            o++;
            if (c.use_graphs) {
                IB3D_CHECK(hipGraphLaunch(g_cyc_exec, c.stream));
                continue;
            }
            const double *gate = conv + (o - 1);
            // the first cycle is checked from the step before the one at
            // which the previous solve's first cycle ended (if it ended
            // early), later cycles (usually a step or two) after every step
            const int from = o > 1 ? 0 : c.last_cyc1 >= 1 && c.last_cyc1 < m ? (int)c.last_cyc1 - 1 : 1 << 30;
            if (gmres_cycle(c, m, k, gate, c.krylov_target * c.tol * bn, dnorm ? dnorm + 1 : nullptr, from))
                return -1;
            residual_async(c, false, true, bn, relv + o, conv + o, gate, dnorm);
        }
        IB3D_CHECK(hipGetLastError());
        // the batch's residuals and flags and the Arnoldi step count, in one
        // host round trip (into the pinned mirror when they fit)
        const int64_t n = o - seen + 1;
        double *pr, *pc, *ps;
        if (n <= kBatchMirror) {
            pr = c.hscal + 16;
            pc = pr + kBatchMirror;
            ps = pc + kBatchMirror;
        } else {
            hr.resize(2 * n + 2);
            pr = hr.data();
            pc = pr + n;
            ps = pc + n;
        }
        IB3D_CHECK(hipMemcpyAsync(pr, relv + seen, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        IB3D_CHECK(hipMemcpyAsync(pc, conv + seen, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        IB3D_CHECK(hipMemcpyAsync(ps, c.scalars + 24, sizeof(double), hipMemcpyDeviceToHost, c.stream));
        if (dnorm) IB3D_CHECK(hipMemcpyAsync(ps + 1, dnorm, sizeof(double), hipMemcpyDeviceToHost, c.stream));
        IB3D_CHECK(hipStreamSynchronize(c.stream));
        if (dnorm) bn = ps[1];
        if (seen == 0) *rel0 = pr[0];
        for (int64_t q = 0; q < n; q++) {
            if (!(pr[q] == pr[q]) || !std::isfinite(bn)) {
                set_error("poissonSolver: the residual is not finite");
                return -1;
            }
            if (pc[q] < 0.5 || seen + q >= max_outer) {
                const int64_t oc = seen + q;
                // the steps counted so far are all of the solve's: the
                // cycles after convergence were gated off
                const double steps = *ps;
                *sweeps = (int64_t)steps * k;
                *outer = oc;
                c.last_cyc1 = oc == 1 ? (int64_t)steps : oc > 1 ? m : 0;
                *rel = pr[q];
                if (cnn && cnn_after_solve(c, c.scalars + 16, c.scalars + 18)) return -1;
                return 0;
            }
        }
        seen = o + 1;
        batch = 1;
    }
}

int sor_solve(Ctx &c, double bnorm2, double eps1, const double *dnorm) {
    const double bn = dnorm ? 0.0 : std::sqrt(bnorm2);
    if (c.mode == 0) {
        int64_t isum;
        double d2;
        if (redblack_sor_reference(c, eps1, &isum, &d2)) return -1;
        c.last_outer = 0;
        c.last_relres = d2;
        c.last_r0 = 0.0;
        return 0;
    }
    int64_t sw = 0, out = 0;
    double rel = 0, rel0 = 0;
    if (c.mode == 3 || c.mode == 4) {
        if (sor_gmres(c, bn, &sw, &out, &rel, &rel0, dnorm)) return -1;
    } else if (sor_l2(c, c.mode == 2, bn, &sw, &out, &rel, &rel0)) {
        return -1;
    }
    c.last_sweeps = sw;
    c.last_outer = out;
    c.last_relres = rel;
    c.last_r0 = rel0;
    // the sweep cap ends the solve like the reference rule's: as an error
    if (!(rel <= c.tol)) {
        set_error("poissonSolver: ||b/dt - A pc|| / ||b/dt|| = " + std::to_string(rel) + " above the tolerance " +
                  std::to_string(c.tol) + " after " + std::to_string(sw) + " sweeps (the sweep cap)");
        return -1;
    }
    return 0;
}

int set_fused_sweeps(int on, int tz, int blocks, int coarse_blocks) {
    IB3D_REQUIRE(on >= 0 && on <= 64 && tz >= 0 && tz <= 4096 && blocks >= 1 && blocks <= 1 << 20 &&
                     coarse_blocks >= 1 && coarse_blocks <= 1 << 20,
                 "set_fused_sweeps: bad level count or launch shape");
    g_fused = on;
    g_sweep_tz = tz;
    g_sweep_blocks = blocks;
    g_sweep_blocks_coarse = coarse_blocks;
    return 0;
}

int set_coarse_tiles(int from_level) {
    IB3D_REQUIRE(from_level >= 0 && from_level <= 64, "set_coarse_tiles: bad level");
    g_coarse_tiles = from_level;
    return 0;
}

int set_fused_restrict(int levels) {
    IB3D_REQUIRE(levels >= 0 && levels <= 64, "set_fused_restrict: bad level count");
    g_fused_restrict = levels;
    return 0;
}

int residual_norm(Ctx &c, double *rnorm, double *bnorm) {
    const unsigned nb = nblocks(2 * c.NC, kSorBlock);
    if (ensure_partial(nb)) return -1;
    hipLaunchKernelGGL(k_resid<false>, dim3(nb), dim3(kSorBlock), 0, c.stream, c.pc[0], c.pc[1], c.e32[0],
                       c.e32[1], c.bdt[0], c.bdt[1], nullptr, nullptr, c.ax, c.ay, c.az, geom(c), c.partial,
                       nullptr, 0.0);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, c.stream, c.partial, (int64_t)nb, c.scalars + 2);
    IB3D_CHECK(hipGetLastError());
    IB3D_CHECK(hipMemcpyAsync(c.hscal + 2, c.scalars + 2, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    IB3D_CHECK(hipStreamSynchronize(c.stream));
    *rnorm = std::sqrt(c.hscal[2]);
    *bnorm = std::sqrt(c.bnorm2);
    return 0;
}

int probe_copy(Ctx &c, int64_t nbytes, double *ms) {
    const int64_t n = nbytes / 32;
    double4 *a = nullptr, *b = nullptr;
    IB3D_CHECK(hipMalloc(&a, n * 32));
    IB3D_CHECK(hipMalloc(&b, n * 32));
    IB3D_CHECK(hipMemsetAsync(a, 0, n * 32, c.stream));
    hipEvent_t e0, e1;
    IB3D_CHECK(hipEventCreate(&e0));
    IB3D_CHECK(hipEventCreate(&e1));
    hipLaunchKernelGGL(k_copy_probe, dim3(nblocks(n, kSorBlock)), dim3(kSorBlock), 0, c.stream, a, b, n);
    IB3D_CHECK(hipEventRecord(e0, c.stream));
    hipLaunchKernelGGL(k_copy_probe, dim3(nblocks(n, kSorBlock)), dim3(kSorBlock), 0, c.stream, a, b, n);
    IB3D_CHECK(hipEventRecord(e1, c.stream));
    IB3D_CHECK(hipEventSynchronize(e1));
    float t = 0;
    IB3D_CHECK(hipEventElapsedTime(&t, e0, e1));
    *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    IB3D_CHECK(hipFree(a));
    IB3D_CHECK(hipFree(b));
    return 0;
}

int bench_sor(Ctx &c, int precision, int sweeps, double omega, double *ms) {
    hipEvent_t a, b;
    IB3D_CHECK(hipEventCreate(&a));
    IB3D_CHECK(hipEventCreate(&b));
    if (precision == 32 || precision == 33) {
        if (sor_float_tables(c)) return -1;
        const unsigned nb = nblocks(2 * c.NC, kSorBlock);
        if (ensure_partial(nb)) return -1;
        hipLaunchKernelGGL(k_resid<false>, dim3(nb), dim3(kSorBlock), 0, c.stream, c.pc[0], c.pc[1], c.e32[0],
                           c.e32[1], c.bdt[0], c.bdt[1], c.r32[0], c.r32[1], c.ax, c.ay, c.az, geom(c), c.partial,
                           nullptr, 0.0);
    }
    // 33: the fused FP32 sweep (one launch per red-black sweep) between e32
    // and a second buffer, as on the finest multigrid level
    // After a mode-4 solve it sweeps the solver's own buffers (the first
    // Krylov vector z from the first basis vector v, and the multigrid's
    // second buffer), so the probe reads and writes the memory the solver's
    // sweeps do; the solver overwrites z and v at its next solve and keeps 0
    // in their cells outside the solve (the sweep only copies those).
    MgLevel L;
    float *x2 = nullptr, *x0 = c.e32[0];
    const float *rhs0 = c.r32[0], *rhs1 = c.r32[1];
    bool own_x2 = false;
    int S = 0, tz = 0;
    dim3 sg;
    if (precision == 33) {
        if (sor_float_tables(c)) return -1;
        L.sx = c.sx; L.sy = c.sy; L.sz = c.sz; L.H = c.H; L.NC = c.NC; L.sxp = c.sxp;
        IB3D_REQUIRE(sweep2_shape(L, S, tz, sg), "bench_sor: the fused sweep does not fit this grid");
        if (!g_mg.empty() && g_mg[0].x2 && g_mg[0].NC == c.NC) {
            x2 = g_mg[0].x2;
            if (g_kv && g_kz && g_k_nc == c.NC) {
                x0 = g_kz;
                rhs0 = g_kv;
                rhs1 = g_kv + c.NC;
            }
        } else {
            IB3D_CHECK(hipMalloc(&x2, 2 * c.NC * sizeof(float)));
            IB3D_CHECK(dzero(x2, 2 * c.NC * sizeof(float)));
            own_x2 = true;
        }
    }
    const SorGeom g = geom(c);
    const float om = (float)omega, omm = (float)(1.0 - omega);
    float *xa = x0, *xb = x2;
    IB3D_CHECK(hipEventRecord(a, c.stream));
    for (int q = 0; q < sweeps; q++) {
        if (precision == 33) {
            hipLaunchKernelGGL(k_sweep2_pz<false>, sg, dim3(kSweepThreads), 0, c.stream, xa, xa + c.NC, xb,
                               xb + c.NC, rhs0, rhs1, g_axf, g_ayf, g_azf, g, S, tz, om, omm, nullptr, 0.0);
            std::swap(xa, xb);
        } else if (precision == 32) {
            sweep32(c, omega, false, false, nullptr, 0.0);
        } else {
            sweep64(c, omega, false, nullptr, nullptr, 0.0);
        }
    }
    IB3D_CHECK(hipEventRecord(b, c.stream));
    IB3D_CHECK(hipEventSynchronize(b));
    float t = 0;
    IB3D_CHECK(hipEventElapsedTime(&t, a, b));
    *ms = (double)t / sweeps;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (x2) {
        if (xa == x2) IB3D_CHECK(hipMemcpyAsync(x0, x2, 2 * c.NC * sizeof(float), hipMemcpyDeviceToDevice,
                                                c.stream));
        IB3D_CHECK(hipStreamSynchronize(c.stream));
        if (own_x2) IB3D_CHECK(hipFree(x2));
    }
    return 0;
}

}  // namespace ib3d
