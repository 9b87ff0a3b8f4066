// ib3d_cnn.hip -- learned initial guess for the pressure-correction solve.
//
// A 3-D CNN (3 -> 16 -> 16 -> 16 -> 1 channels, 3x3x3 kernels, zero "same"
// padding, ReLU after the first three layers) maps, per cell,
//   ch0  r0 / |A_diag| / s_r   (r0 = b/dt - A p_prev: the divergence right-hand
//                               side seen through the previous correction)
//   ch1  1 on cells the solver updates (fluid), 0 on solid / intercepted /
//        wall / padding cells
//   ch2  p_prev / s_p
// to delta / s_r, and the guess is p_prev + delta on fluid cells.  s_r and
// s_p are RMS values over fluid cells, so the network sees scale-free data
// on any grid.  The guess is kept only if its FP64 residual is smaller than
// the previous-step guess's; otherwise the solve starts from p_prev.  SOR
// then iterates to the same tolerance either way.
//
// Everything runs in hand-written kernels:
//   k_conv_ring the 8-channel input layer: output planes in turn from a
//              three-plane LDS ring, taps in order.
//   k_conv_zm  forward implicit-GEMM convolution on MFMA
//              (v_mfma_f32_16x16x32_bf16): D[16 out-ch][16 voxels] =
//              W[16 x 27*Cin] * im2col[27*Cin x 16], BF16 weights and
//              activations, FP32 accumulation; input planes streamed along z
//              through LDS, each plane feeding all three z-taps.
//   k_conv_lds the same product over a tile staged in LDS; with FLIP and
//              transposed weights it is the backward-data convolution.
//   k_conv_dw  weight gradient on MFMA: D[16 out][16 in] per tap =
//              dZ^T[16 x 32 voxels] * X_tap[32 voxels x 16], both operands
//              staged transposed in LDS; deterministic per-block partials.
//   k_adam     Adam on the FP32 master weights, then repacking to BF16.
// Activation layout: channels-last (voxel-major, 8 or 16 BF16 channels per
// voxel), so one lane's eight K values are one 16-byte load.
//
// MFMA operand layout (16x16x32 BF16): lane l supplies A[l%16][8(l/16)+j]
// and B[8(l/16)+j][l%16], j = 0..7, and holds D[4(l/16)+i][l%16], i = 0..3.
#include "ib3d_internal.h"

#include <algorithm>
#include <cmath>
#include <random>
#include <type_traits>
#include <vector>

namespace ib3d {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

static constexpr int kNParam = 15601;
// layer table: real in/out channels, stored input channels, K chunks of 32
static constexpr int kCin[4] = {3, 16, 16, 16};
static constexpr int kCout[4] = {16, 16, 16, 1};
static constexpr int kWoff[4] = {0, 1312, 8240, 15168};
static constexpr int kBoff[4] = {1296, 8224, 15152, 15600};
static constexpr int kNkc[4] = {7, 15, 15, 15};   // forward K chunks (layers 1-3: taps grouped by z plane)
static constexpr int kTile = 64, kHalo = 4, kTileIn = kTile + 2 * kHalo;

struct Vol {
    int vx, vy, vz, nb;
};

__device__ __forceinline__ int64_t vox_index(const Vol &v, int b, int z, int y, int x) {
    return (((int64_t)b * v.vz + z) * v.vy + y) * v.vx + x;
}

// ------------------------------------------------------------------ packing
// forward pack: row m = output channel, k = (tap, input channel).  grouped:
// chunk kc = dz*cpz + c holds in-plane taps 2c, 2c+1 (16 stored channels)
// or 4c .. 4c+3 (8 stored channels) of z-tap dz, padded with zeros past tap
// 8 (k_conv_zm); otherwise chunk kc holds taps 2kc, 2kc+1 or 4kc .. 4kc+3
// in order, zeros past tap 26 (k_conv_ring)
__global__ void k_pack_fwd(const float *__restrict__ prm, int woff, int boff, int cin, int cout, int cinst, int nkc,
                           int grouped, __bf16 *__restrict__ out, float *__restrict__ bias16) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < 16) bias16[idx] = idx < cout ? prm[boff + idx] : 0.0f;
    if (idx >= nkc * 512) return;
    const int j = idx & 7, l = (idx >> 3) & 63, kc = idx >> 9;
    const int m = l & 15, g = l >> 4;
    int t, ci;
    bool ok;
    if (grouped) {
        const int cpz = nkc / 3, dz = kc / cpz, c = kc - dz * cpz;
        const int ip = cinst == 16 ? 2 * c + (g >> 1) : 4 * c + g;
        ci = cinst == 16 ? 8 * (g & 1) + j : j;
        t = 9 * dz + ip;
        ok = ip < 9;
    } else {   // taps in order: 2 (16 channels) or 4 (8 channels) per chunk
        t = cinst == 16 ? 2 * kc + (g >> 1) : 4 * kc + g;
        ci = cinst == 16 ? 8 * (g & 1) + j : j;
        ok = t < 27;
    }
    const float w = (ok && ci < cin && m < cout) ? prm[woff + (t * cin + ci) * cout + m] : 0.0f;
    out[idx] = (__bf16)w;
}

// transposed pack for the backward-data convolution (16 stored channels of
// the layer's output gradient): row m = layer input channel, k = (tap, out)
__global__ void k_pack_bwd(const float *__restrict__ prm, int woff, int cin, int cout, __bf16 *__restrict__ out) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 14 * 512) return;
    const int j = idx & 7, l = (idx >> 3) & 63, kc = idx >> 9;
    const int m = l & 15, g = l >> 4;
    const int t = 2 * kc + (g >> 1), co = 8 * (g & 1) + j;
    const float w = (t < 27 && m < cin && co < cout) ? prm[woff + (t * cin + m) * cout + co] : 0.0f;
    out[idx] = (__bf16)w;
}

// ------------------------------------------------------------ convolution
// Tile-staged form, used for the backward-data pass (MODE 2: times
// ReLU'(gate), 16 BF16 channels out; FLIP: source offset -d instead of +d).
// MODE 0 (+ bias, ReLU) and MODE 1 (+ bias, channel 0 as FP32) are the
// forward epilogues; the forward pass itself runs in k_conv_zm below.
//
// A 256-thread workgroup computes 16 (x) x 4 (y) x 4 (z) output voxels.  It
// loads their 18 x 6 x 6 input neighbourhood once into LDS (zero outside the
// volume: the "same" padding), then every wave runs, for 4 of the 16 output
// rows of 16 voxels, D[16 out-ch][16 voxels] = W[16 x 27*Cin] *
// im2col[27*Cin x 16] as 7 (Cin 8) or 14 (Cin 16) MFMAs of K = 32, its B
// operands read from LDS; the 27-fold reuse of every input value is served
// by LDS instead of L1/L2.
static constexpr int kCtX = 16, kCtY = 4, kCtZ = 4;
static constexpr int kChX = kCtX + 2, kChY = kCtY + 2, kChZ = kCtZ + 2;

template <int CIN, bool FLIP, int MODE>
__global__ void __launch_bounds__(256)
k_conv_lds(const __bf16 *__restrict__ X, const bf16x8 *__restrict__ W, const float *__restrict__ bias, Vol v,
           __bf16 *__restrict__ Y16, float *__restrict__ Y32, const __bf16 *__restrict__ gate, int ntz) {
    constexpr int NKC = CIN == 16 ? 14 : 7;
    constexpr int VU = CIN / 8;  // 16-byte units per voxel
    __shared__ __attribute__((aligned(16))) __bf16 tile[kChZ * kChY * kChX * CIN];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 15, g = lane >> 4;
    const int x0 = (int)blockIdx.x * kCtX, y0 = (int)blockIdx.y * kCtY;
    const int b = (int)blockIdx.z / ntz, z0 = ((int)blockIdx.z - b * ntz) * kCtZ;
    // stage the input neighbourhood
    for (int u = threadIdx.x; u < kChZ * kChY * kChX * VU; u += 256) {
        const int vox = u / VU, part = u - vox * VU;
        const int hx = vox % kChX, hy = (vox / kChX) % kChY, hz = vox / (kChX * kChY);
        const int xs = x0 - 1 + hx, ys = y0 - 1 + hy, zs = z0 - 1 + hz;
        bf16x8 val = {};
        if (xs >= 0 && xs < v.vx && ys >= 0 && ys < v.vy && zs >= 0 && zs < v.vz)
            val = *reinterpret_cast<const bf16x8 *>(X + vox_index(v, b, zs, ys, xs) * CIN + 8 * part);
        *reinterpret_cast<bf16x8 *>(tile + vox * CIN + 8 * part) = val;
    }
    bf16x8 wr[NKC];
#pragma unroll
    for (int kc = 0; kc < NKC; kc++) wr[kc] = W[kc * 64 + lane];
    int toff[NKC];
    bool tap_ok[NKC];
#pragma unroll
    for (int kc = 0; kc < NKC; kc++) {
        int t, c0;
        if (CIN == 16) {
            t = 2 * kc + (g >> 1);
            c0 = 8 * (g & 1);
        } else {
            t = 4 * kc + g;
            c0 = 0;
        }
        tap_ok[kc] = t < 27;
        const int tt = t < 27 ? t : 13;
        int dx = tt % 3 - 1, dy = (tt / 3) % 3 - 1, dz = tt / 9 - 1;
        if (FLIP) {
            dx = -dx;
            dy = -dy;
            dz = -dz;
        }
        // element offset of the tap's source relative to the output voxel's tile slot
        toff[kc] = ((dz * kChY + dy) * kChX + dx) * CIN + c0;
    }
    __syncthreads();
    for (int row = wave; row < kCtY * kCtZ; row += 4) {
        const int ty = row % kCtY, tz = row / kCtY;
        const int x = x0 + n, y = y0 + ty, z = z0 + tz;
        const int base = (((tz + 1) * kChY + (ty + 1)) * kChX + (n + 1)) * CIN;
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kc = 0; kc < NKC; kc++) {
            bf16x8 xv = {};
            if (tap_ok[kc]) xv = *reinterpret_cast<const bf16x8 *>(tile + base + toff[kc]);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wr[kc], xv, acc, 0, 0, 0);
        }
        if (x < v.vx && y < v.vy && z < v.vz) {
            const int64_t vo = vox_index(v, b, z, y, x);
            if (MODE == 1) {
                if (g == 0) Y32[vo] = acc[0] + bias[0];
            } else {
                __bf16 o[4];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    float val;
                    if (MODE == 0) {
                        val = acc[i] + bias[4 * g + i];
                        val = val > 0.0f ? val : 0.0f;
                    } else {
                        val = (float)gate[vo * 16 + 4 * g + i] > 0.0f ? acc[i] : 0.0f;
                    }
                    o[i] = (__bf16)val;
                }
                typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
                bf16x4 ov = {o[0], o[1], o[2], o[3]};
                *reinterpret_cast<bf16x4 *>(Y16 + vo * 16 + 4 * g) = ov;
            }
        }
    }
}

// Forward convolution of the 8-channel input layer, one output plane per
// step (MODE 0: + bias, ReLU, 16 BF16 channels out).  With 8 channels a K
// chunk of 32 holds four taps, so grouping the taps by z plane (as
// k_conv_zm does) would pad 9 in-plane taps to 12; here the 27 taps run in
// order (7 chunks) and each output plane reads its three input planes from
// an LDS ring instead.  A 256-thread workgroup owns a 16 (x) x 16 (y)
// column of output voxels over kRgZ consecutive planes; the ring holds the
// three input planes with a one-voxel x-y halo (18 x 18 voxels).  Each step
// loads the next plane into registers before the MFMAs of the current one
// and writes it over the oldest slot afterwards.  Wave w computes output
// rows y0+4w .. y0+4w+3 with four independent accumulators.
static constexpr int kRgX = 16, kRgY = 16, kRgZ = 16;
static constexpr int kRpX = kRgX + 2, kRpV = kRpX * (kRgY + 2);

template <int CIN, int MODE>
__global__ void __launch_bounds__(256)
k_conv_ring(const __bf16 *__restrict__ X, const bf16x8 *__restrict__ W, const float *__restrict__ bias, Vol v,
          __bf16 *__restrict__ Y16, float *__restrict__ Y32, int nzc) {
    constexpr int NKC = CIN == 16 ? 14 : 7;
    constexpr int VU = CIN / 8;        // 16-byte units per voxel
    constexpr int PE = kRpV * CIN;     // BF16 elements per staged plane
    constexpr int PU = kRpV * VU;      // 16-byte units per staged plane
    constexpr int UPT = (PU + 255) / 256;
    __shared__ __attribute__((aligned(16))) __bf16 ring[3 * PE];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 15, g = lane >> 4;
    const int x0 = (int)blockIdx.x * kRgX, y0 = (int)blockIdx.y * kRgY;
    const int b = (int)blockIdx.z / nzc, z0 = ((int)blockIdx.z - b * nzc) * kRgZ;
    const int z1 = min(z0 + kRgZ, v.vz);
    const int64_t pstride = (int64_t)v.vy * v.vx * CIN;
    const __bf16 *Xb = X + (int64_t)b * v.vz * pstride;

    // this thread's staging units, the same in every plane
    int soff[UPT];
    int64_t goff[UPT];   // -1: beyond the plane or outside the volume in x / y
#pragma unroll
    for (int u = 0; u < UPT; u++) {
        const int unit = (int)threadIdx.x + 256 * u;
        const int vox = unit / VU, part = unit - vox * VU;
        const int xs = x0 - 1 + vox % kRpX, ys = y0 - 1 + vox / kRpX;
        soff[u] = unit < PU ? part * (kRpV * 8) + vox * 8 : -1;
        goff[u] = (unit < PU && xs >= 0 && xs < v.vx && ys >= 0 && ys < v.vy)
                      ? ((int64_t)ys * v.vx + xs) * CIN + 8 * part
                      : -1;
    }
    bf16x8 r[UPT];
    auto load_plane = [&](int zs) {
#pragma unroll
        for (int u = 0; u < UPT; u++) {
            bf16x8 val = {};
            if (goff[u] >= 0 && zs >= 0 && zs < v.vz)
                val = *reinterpret_cast<const bf16x8 *>(Xb + zs * pstride + goff[u]);
            r[u] = val;
        }
    };
    auto store_plane = [&](int slot) {
#pragma unroll
        for (int u = 0; u < UPT; u++)
            if (soff[u] >= 0) *reinterpret_cast<bf16x8 *>(ring + slot * PE + soff[u]) = r[u];
    };

    bf16x8 wr[NKC];
    int toff[NKC], tdz[NKC];
#pragma unroll
    for (int kc = 0; kc < NKC; kc++) {
        wr[kc] = W[kc * 64 + lane];
        int t, part;
        if (CIN == 16) {
            t = 2 * kc + (g >> 1);
            part = g & 1;
        } else {
            t = 4 * kc + g;
            part = 0;
        }
        // padded K slots (t = 27) carry zero weights; point them at the centre tap
        const int tt = t < 27 ? t : 13;
        const int dx = tt % 3 - 1, dy = (tt / 3) % 3 - 1;
        tdz[kc] = tt / 9;
        toff[kc] = part * (kRpV * 8) + (dy * kRpX + dx) * 8;
    }
    float bv[4];
#pragma unroll
    for (int i = 0; i < 4; i++) bv[i] = MODE == 0 ? bias[4 * g + i] : bias[0];
    int rowoff[4];
#pragma unroll
    for (int rr = 0; rr < 4; rr++) rowoff[rr] = ((4 * wave + rr + 1) * kRpX + n + 1) * 8;

    load_plane(z0 - 1);
    store_plane(0);
    load_plane(z0);
    store_plane(1);
    load_plane(z0 + 1);
    store_plane(2);
    __syncthreads();
    int s0 = 0;   // slot of plane z-1; planes z and z+1 follow cyclically
    for (int z = z0; z < z1; z++) {
        const bool more = z + 2 <= z1;   // plane z+2 feeds output plane z+1
        if (more) load_plane(z + 2);
        f32x4 acc[4];
#pragma unroll
        for (int rr = 0; rr < 4; rr++) acc[rr] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kc = 0; kc < NKC; kc++) {
            int sl = s0 + tdz[kc];
            sl = sl >= 3 ? sl - 3 : sl;
            const __bf16 *src = ring + sl * PE + toff[kc];
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                const bf16x8 xv = *reinterpret_cast<const bf16x8 *>(src + rowoff[rr]);
                acc[rr] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wr[kc], xv, acc[rr], 0, 0, 0);
            }
        }
        const int x = x0 + n;
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
            const int y = y0 + 4 * wave + rr;
            if (x < v.vx && y < v.vy) {
                const int64_t vo = vox_index(v, b, z, y, x);
                if (MODE == 1) {
                    if (g == 0) Y32[vo] = acc[rr][0] + bv[0];
                } else {
                    typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
                    bf16x4 ov;
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const float val = acc[rr][i] + bv[i];
                        ov[i] = (__bf16)(val > 0.0f ? val : 0.0f);
                    }
                    *reinterpret_cast<bf16x4 *>(Y16 + vo * 16 + 4 * g) = ov;
                }
            }
        }
        __syncthreads();
        if (more) store_plane(s0);   // plane z+2 replaces plane z-1
        s0 = s0 == 2 ? 0 : s0 + 1;
        __syncthreads();
    }
}

// Forward convolution marching along z (MODE 0: + bias, ReLU, 16 BF16
// channels out; MODE 1: + bias, channel 0 as FP32).  A 256-thread workgroup
// owns a 16 (x) x 16 (y) column of output voxels over kZmZ consecutive
// planes and streams the input planes z0-1 .. z1 through a two-slot LDS
// ring (18 x 18 voxels: the one-voxel x-y halo; zero outside the volume).
// Each input plane is consumed once: the B operand of every in-plane tap is
// read from LDS once and multiplied by the weights of all three z-taps, into
// the accumulators of the output planes p-1, p and p+1 (so LDS delivers a
// third of the operand bytes a per-output-plane im2col would).  The K
// chunks group the taps by plane: chunk dz*CPZ + c holds in-plane taps
// 2c, 2c+1 x 16 channels (4c .. 4c+3 x 8 channels for the 8-channel input),
// the tenth (twelfth) slot padded with zero weights.  While a plane's MFMAs
// run, the next two planes are in flight in registers (a plane's loads have
// two planes of work to arrive); the nearer one goes to the other slot
// afterwards, behind one barrier per plane.  Wave w computes output rows
// y0+4w .. y0+4w+3.  A staged plane is split by channel half ([half][voxel]
// [8 channels]) so the 16 lanes reading one row's voxels touch consecutive
// 16-byte words.
static constexpr int kZmX = 16, kZmY = 16, kZmZ = 24;
static constexpr int kZpX = kZmX + 2, kZpV = kZpX * (kZmY + 2);

template <int CIN, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_conv_zm(const __bf16 *__restrict__ X, const bf16x8 *__restrict__ W, const float *__restrict__ bias, Vol v,
          __bf16 *__restrict__ Y16, float *__restrict__ Y32, int nzc) {
    constexpr int CPZ = CIN == 16 ? 5 : 3;   // K chunks per z-tap (9 in-plane taps)
    constexpr int NKC = 3 * CPZ;
    constexpr int VU = CIN / 8;              // 16-byte units per voxel
    constexpr int PE = kZpV * CIN;           // BF16 elements per staged plane
    constexpr int PU = kZpV * VU;            // 16-byte units per staged plane
    constexpr int UPT = (PU + 255) / 256;
    __shared__ __attribute__((aligned(16))) __bf16 ring[2 * PE];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 15, g = lane >> 4;
    const int x0 = (int)blockIdx.x * kZmX, y0 = (int)blockIdx.y * kZmY;
    const int b = (int)blockIdx.z / nzc, z0 = ((int)blockIdx.z - b * nzc) * kZmZ;
    const int z1 = min(z0 + kZmZ, v.vz);
    const int64_t pstride = (int64_t)v.vy * v.vx * CIN;
    const __bf16 *Xb = X + (int64_t)b * v.vz * pstride;

    // this thread's staging units, the same in every plane
    int soff[UPT];
    int64_t goff[UPT];   // clamped into the plane; gok: inside the volume in x / y
    bool gok[UPT];
#pragma unroll
    for (int u = 0; u < UPT; u++) {
        const int unit = (int)threadIdx.x + 256 * u;
        const int vox = unit / VU, part = unit - vox * VU;
        const int xs = x0 - 1 + vox % kZpX, ys = y0 - 1 + vox / kZpX;
        soff[u] = unit < PU ? part * (kZpV * 8) + vox * 8 : -1;
        gok[u] = unit < PU && xs >= 0 && xs < v.vx && ys >= 0 && ys < v.vy;
        goff[u] = ((int64_t)min(max(ys, 0), v.vy - 1) * v.vx + min(max(xs, 0), v.vx - 1)) * CIN + 8 * part;
    }
    // planes p+1 and p+2 are in flight in rg[(p+1)&1] and rg[p&1] while
    // plane p is consumed (the parity is a template argument, so no copy
    // between the sets waits for a load)
    bf16x8 rg[2][UPT];
    // loads without control flow (clamped address, zero selected after), so
    // that the wait for a plane leaves the later planes' loads in flight
    auto load_plane = [&](int zs, bf16x8 *r) {
        const bool zok = zs >= 0 && zs < v.vz;
        const __bf16 *pl = Xb + min(max(zs, 0), v.vz - 1) * pstride;
#pragma unroll
        for (int u = 0; u < UPT; u++) {
            const bf16x8 val = *reinterpret_cast<const bf16x8 *>(pl + goff[u]);
            r[u] = (gok[u] && zok) ? val : bf16x8{};
        }
    };
    auto store_plane = [&](int slot, const bf16x8 *r) {
#pragma unroll
        for (int u = 0; u < UPT; u++)
            if (soff[u] >= 0) *reinterpret_cast<bf16x8 *>(ring + slot * PE + soff[u]) = r[u];
    };

    bf16x8 wr[NKC];
#pragma unroll
    for (int kc = 0; kc < NKC; kc++) wr[kc] = W[kc * 64 + lane];
    int toff[CPZ];
#pragma unroll
    for (int c = 0; c < CPZ; c++) {
        const int ip = CIN == 16 ? 2 * c + (g >> 1) : 4 * c + g;
        const int part = CIN == 16 ? (g & 1) : 0;
        const int ipp = ip < 9 ? ip : 4;   // padded slots (zero weights) read the centre
        toff[c] = part * (kZpV * 8) + ((ipp / 3 - 1) * kZpX + (ipp % 3 - 1)) * 8;
    }
    int rowoff[4];
#pragma unroll
    for (int rr = 0; rr < 4; rr++) rowoff[rr] = ((4 * wave + rr + 1) * kZpX + n + 1) * 8;
    float bv[4];
#pragma unroll
    for (int i = 0; i < 4; i++) bv[i] = MODE == 0 ? bias[4 * g + i] : bias[0];

    // Three accumulator sets hold the output planes p-1, p, p+1 while input
    // plane p is consumed; the plane loop is unrolled by three so that their
    // roles rotate by renaming.  Every plane runs all three z-taps: the
    // results for planes outside z0 .. z1-1 (at the two ends of the column)
    // are computed and dropped, which keeps the matrix work free of branches.
    f32x4 acc[3][4];
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
        for (int rr = 0; rr < 4; rr++) acc[q][rr] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    int cur = 0;
    auto plane = [&](int p, auto role, auto parity) {
        constexpr int M = decltype(role)::value, Z = (M + 1) % 3, P = (M + 2) % 3;
        constexpr int Q = decltype(parity)::value;
        load_plane(p + 2, rg[Q]);   // past z1 (unused) when p + 2 > z1
        const __bf16 *src = ring + cur * PE;
        // the B operands of chunk c+1 are read while chunk c's 12 MFMAs run
        bf16x8 xb[2][4];
#pragma unroll
        for (int rr = 0; rr < 4; rr++) xb[0][rr] = *reinterpret_cast<const bf16x8 *>(src + toff[0] + rowoff[rr]);
#pragma unroll
        for (int c = 0; c < CPZ; c++) {
            if (c + 1 < CPZ) {
#pragma unroll
                for (int rr = 0; rr < 4; rr++)
                    xb[(c + 1) & 1][rr] = *reinterpret_cast<const bf16x8 *>(src + toff[c + 1] + rowoff[rr]);
            }
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                const bf16x8 xv = xb[c & 1][rr];
                acc[M][rr] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wr[2 * CPZ + c], xv, acc[M][rr], 0, 0, 0);
                acc[Z][rr] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wr[CPZ + c], xv, acc[Z][rr], 0, 0, 0);
                acc[P][rr] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wr[c], xv, acc[P][rr], 0, 0, 0);
            }
        }
        if (p - 1 >= z0) {   // output plane p-1 has all three z-taps
            const int x = x0 + n, z = p - 1;
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                const int y = y0 + 4 * wave + rr;
                if (x < v.vx && y < v.vy) {
                    const int64_t vo = vox_index(v, b, z, y, x);
                    if (MODE == 1) {
                        if (g == 0) Y32[vo] = acc[M][rr][0] + bv[0];
                    } else {
                        typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
                        bf16x4 ov;
#pragma unroll
                        for (int i = 0; i < 4; i++) {
                            const float val = acc[M][rr][i] + bv[i];
                            ov[i] = (__bf16)(val > 0.0f ? val : 0.0f);
                        }
                        *reinterpret_cast<bf16x4 *>(Y16 + vo * 16 + 4 * g) = ov;
                    }
                }
            }
        }
#pragma unroll
        for (int rr = 0; rr < 4; rr++) acc[M][rr] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};   // becomes plane p+2
        // plane p+1 goes to the other slot, whose plane p-1 every wave has
        // finished reading before the previous barrier
        if (p < z1) store_plane(cur ^ 1, rg[Q ^ 1]);
        cur ^= 1;
        __syncthreads();
    };

    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    load_plane(z0 - 1, rg[0]);
    store_plane(0, rg[0]);
    load_plane(z0, rg[1]);
    __syncthreads();
    for (int p = z0 - 1; p <= z1; p += 6) {   // roles cycle by 3, register sets by 2
        plane(p, I0{}, I0{});
        if (p + 1 > z1) break;
        plane(p + 1, I1{}, I1{});
        if (p + 2 > z1) break;
        plane(p + 2, I2{}, I0{});
        if (p + 3 > z1) break;
        plane(p + 3, I0{}, I1{});
        if (p + 4 > z1) break;
        plane(p + 4, I1{}, I0{});
        if (p + 5 > z1) break;
        plane(p + 5, I2{}, I1{});
    }
}

// weight and bias gradient: partial[block][t][out][in] for taps t < 27 and
// the bias as tap 27 (B = ones in column 0)
template <int CIN>
__global__ void __launch_bounds__(256)
k_conv_dw(const __bf16 *__restrict__ X, const __bf16 *__restrict__ dZ, Vol v, float *__restrict__ partial,
          int chunks_per_block, int64_t nchunks) {
    __shared__ __attribute__((aligned(16))) __bf16 sdz[16 * 32];
    __shared__ __attribute__((aligned(16))) __bf16 sx[27 * 16 * 32];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 15, g = lane >> 4;
    if (CIN == 8) {  // rows 8..15 of every tap stay zero
        for (int q = threadIdx.x; q < 27 * 8 * 32; q += 256) {
            const int t = q / 256, rem = q % 256;
            sx[(t * 16 + 8) * 32 + rem] = (__bf16)0.0f;
        }
    }
    f32x4 acc[7];
#pragma unroll
    for (int s = 0; s < 7; s++) acc[s] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    bf16x8 ones = {}, zeros = {};
#pragma unroll
    for (int j = 0; j < 8; j++) ones[j] = (__bf16)(n == 0 ? 1.0f : 0.0f);
    const int ncx = (v.vx + 31) >> 5;
    for (int q = 0; q < chunks_per_block; q++) {
        const int64_t ch = (int64_t)blockIdx.x * chunks_per_block + q;
        if (ch >= nchunks) break;
        const int xc = (int)(ch % ncx);
        int64_t rest = ch / ncx;
        const int y = (int)(rest % v.vy);
        rest /= v.vy;
        const int z = (int)(rest % v.vz);
        const int b = (int)(rest / v.vz);
        const int x0 = xc * 32;
        __syncthreads();
        if (threadIdx.x < 64) {
            const int vi = threadIdx.x >> 1, h = threadIdx.x & 1, x = x0 + vi;
            bf16x8 val = {};
            if (x < v.vx) val = *reinterpret_cast<const bf16x8 *>(dZ + vox_index(v, b, z, y, x) * 16 + 8 * h);
#pragma unroll
            for (int j = 0; j < 8; j++) sdz[(8 * h + j) * 32 + vi] = val[j];
        }
        constexpr int HALVES = CIN / 8;
        for (int task = threadIdx.x; task < 9 * 34 * HALVES; task += 256) {
            const int h = task % HALVES, u = (task / HALVES) % 34, rr = task / (HALVES * 34);
            const int dy = rr % 3 - 1, dz = rr / 3 - 1;
            const int xs = x0 + u - 1, ys = y + dy, zs = z + dz;
            bf16x8 val = {};
            if (xs >= 0 && xs < v.vx && ys >= 0 && ys < v.vy && zs >= 0 && zs < v.vz)
                val = *reinterpret_cast<const bf16x8 *>(X + vox_index(v, b, zs, ys, xs) * CIN + 8 * h);
#pragma unroll
            for (int dx = -1; dx <= 1; dx++) {
                const int vi = u - 1 - dx;
                if (vi < 0 || vi >= 32) continue;
                const int t = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
#pragma unroll
                for (int j = 0; j < 8; j++) sx[(t * 16 + 8 * h + j) * 32 + vi] = val[j];
            }
        }
        __syncthreads();
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&sdz[n * 32 + 8 * g]);
#pragma unroll
        for (int s = 0; s < 7; s++) {
            const int t = wave + 4 * s;
            if (t > 27) break;
            const bf16x8 bv = t < 27 ? *reinterpret_cast<const bf16x8 *>(&sx[(t * 16 + n) * 32 + 8 * g]) : ones;
            acc[s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[s], 0, 0, 0);
        }
    }
    (void)zeros;
#pragma unroll
    for (int s = 0; s < 7; s++) {
        const int t = wave + 4 * s;
        if (t > 27) break;
#pragma unroll
        for (int i = 0; i < 4; i++)
            partial[((int64_t)blockIdx.x * 28 + t) * 256 + (4 * g + i) * 16 + n] = acc[s][i];
    }
}

// sum the block partials into the parameter gradient of one layer
__global__ void k_dw_reduce(const float *__restrict__ partial, int nblk, int cin, int cout, int woff, int boff,
                            float *__restrict__ grad) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int nw = 27 * cin * cout;
    if (q >= nw + cout) return;
    int t, ci, co;
    if (q < nw) {
        t = q / (cin * cout);
        ci = (q / cout) % cin;
        co = q % cout;
    } else {
        t = 27;
        ci = 0;
        co = q - nw;
    }
    float s = 0.0f;
    for (int b = 0; b < nblk; b++) s += partial[((int64_t)b * 28 + t) * 256 + co * 16 + ci];
    grad[q < nw ? woff + q : boff + co] = s;
}

// ------------------------------------------------------------- loss, Adam
// MSE over voxels with a finite target: dZ (16 channels, channel 0 only)
// and per-block partial sums of the squared error
__global__ void __launch_bounds__(256)
k_loss(const float *__restrict__ P, const float *__restrict__ T, int64_t n, float scale, __bf16 *__restrict__ dZ,
       double *__restrict__ partial) {
    __shared__ double sm[4];
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double e2 = 0.0;
    if (q < n) {
        const float t = T[q];
        float d = 0.0f;
        if (t == t) {
            const float e = P[q] - t;
            e2 = (double)e * e;
            d = 2.0f * e * scale;
        }
        bf16x8 z = {};
        z[0] = (__bf16)d;
        *reinterpret_cast<bf16x8 *>(dZ + q * 16) = z;
        *reinterpret_cast<bf16x8 *>(dZ + q * 16 + 8) = bf16x8{};
    }
    for (int off = 32; off > 0; off >>= 1) e2 += __shfl_xor(e2, off, 64);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = e2;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = ((sm[0] + sm[1]) + sm[2]) + sm[3];
}

__global__ void k_sum_dbl(const double *__restrict__ partial, int64_t n, double *out) {
    __shared__ double sm[256];
    double s = 0.0;
    for (int64_t q = threadIdx.x; q < n; q += 256) s += partial[q];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out += sm[0];
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float gq = g[q];
    const float mq = b1 * m[q] + (1.0f - b1) * gq;
    const float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    const float mh = mq / c1, vh = vq / c2;
    p[q] -= lr * mh / (sqrtf(vh) + eps);
}

// --------------------------------------------------------- data movement
// host-format input (nb x vz x vy x vx x 3 FP32) -> 8-channel BF16
__global__ void k_host_to_x8(const float *__restrict__ in, int64_t nvox, __bf16 *__restrict__ X) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= nvox) return;
    bf16x8 o = {};
    o[0] = (__bf16)in[3 * q];
    o[1] = (__bf16)in[3 * q + 1];
    o[2] = (__bf16)in[3 * q + 2];
    *reinterpret_cast<bf16x8 *>(X + q * 8) = o;
}

// output-gradient of one channel (FP32) -> 16-channel BF16 (channel 0)
__global__ void k_dy_to_dz(const float *__restrict__ dy, int64_t nvox, __bf16 *__restrict__ dZ) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= nvox) return;
    bf16x8 z = {};
    z[0] = (__bf16)dy[q];
    *reinterpret_cast<bf16x8 *>(dZ + q * 16) = z;
    *reinterpret_cast<bf16x8 *>(dZ + q * 16 + 8) = bf16x8{};
}

// training crops: item i = (pair's grid, tile origin); 72^3 input with zero
// padding outside the grid, target on the central 64^3 fluid voxels
// (NaN elsewhere)
__global__ void k_crop(const __bf16 *const *__restrict__ gx, const float *const *__restrict__ gt,
                       const int *__restrict__ items, int sx, int sy, int sz, Vol v, __bf16 *__restrict__ X,
                       float *__restrict__ T) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t per = (int64_t)v.vx * v.vy * v.vz;
    if (q >= per * v.nb) return;
    const int b = (int)(q / per);
    int64_t r = q % per;
    const int xt = (int)(r % v.vx);
    r /= v.vx;
    const int yt = (int)(r % v.vy), zt = (int)(r / v.vy);
    const int pair = items[4 * b], ox = items[4 * b + 1], oy = items[4 * b + 2], oz = items[4 * b + 3];
    const int ii = ox - kHalo + xt, jj = oy - kHalo + yt, kk = oz - kHalo + zt;
    bf16x8 val = {};
    float tv = __builtin_nanf("");
    if (ii >= 0 && ii < sx && jj >= 0 && jj < sy && kk >= 0 && kk < sz) {
        const int64_t cell = ii + (int64_t)sx * (jj + (int64_t)sy * kk);
        val = *reinterpret_cast<const bf16x8 *>(gx[pair] + cell * 8);
        const bool central = xt >= kHalo && xt < kHalo + kTile && yt >= kHalo && yt < kHalo + kTile &&
                             zt >= kHalo && zt < kHalo + kTile;
        if (central && (float)val[1] > 0.5f) tv = gt[pair][cell];
    }
    *reinterpret_cast<bf16x8 *>(X + q * 8) = val;
    T[q] = tv;
}

// ------------------------------------------------- solver-side kernels
// partial sums over fluid cells: (r/diag)^2, p^2, count
__global__ void __launch_bounds__(256)
k_cnn_scales(const float *__restrict__ r0, const float *__restrict__ r1, const double *__restrict__ p0,
             const double *__restrict__ p1, const double *__restrict__ b0, const double *__restrict__ b1,
             const double *__restrict__ ax, const double *__restrict__ ay, const double *__restrict__ az, int sx,
             int sy, int sz, int sxp, int H, FastDiv dsx, FastDiv dsy, double *__restrict__ partial) {
    __shared__ double sm[3][4];
    const int P = sx * sy * sz;
    double a = 0.0, bb = 0.0, cnt = 0.0;
    for (int idx = (int)blockIdx.x * 256 + (int)threadIdx.x; idx < P; idx += (int)gridDim.x * 256) {
        const int row = (int)fdiv((uint32_t)idx, dsx), ii = idx - row * sx;
        const int kk = (int)fdiv((uint32_t)row, dsy), jj = row - kk * sy;
        const int colour = (ii + jj + kk) & 1;
        const int64_t f = (int64_t)row * H + (ii >> 1);
        const double bd = (colour ? b1 : b0)[f];
        if (bd == bd) {
            const double diag = -((ax[sxp + ii] + ay[sy + jj]) + az[sz + kk]);
            const double rr = (double)(colour ? r1 : r0)[f] / diag;
            const double pp = (colour ? p1 : p0)[f];
            a += rr * rr;
            bb += pp * pp;
            cnt += 1.0;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        bb += __shfl_xor(bb, off, 64);
        cnt += __shfl_xor(cnt, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sm[0][threadIdx.x >> 6] = a;
        sm[1][threadIdx.x >> 6] = bb;
        sm[2][threadIdx.x >> 6] = cnt;
    }
    __syncthreads();
    if (threadIdx.x < 3)
        partial[3 * blockIdx.x + threadIdx.x] =
            ((sm[threadIdx.x][0] + sm[threadIdx.x][1]) + sm[threadIdx.x][2]) + sm[threadIdx.x][3];
}

// scal = {s_r, 1/s_r, s_p, 1/s_p} from the partials (0 where a sum is 0)
__global__ void k_cnn_scales_finish(const double *__restrict__ partial, int64_t nblk, double *__restrict__ scal) {
    __shared__ double sm[3][256];
    double s[3] = {0.0, 0.0, 0.0};
    for (int64_t q = threadIdx.x; q < nblk; q += 256)
        for (int c = 0; c < 3; c++) s[c] += partial[3 * q + c];
    for (int c = 0; c < 3; c++) sm[c][threadIdx.x] = s[c];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 3; c++) sm[c][threadIdx.x] += sm[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double cnt = sm[2][0] > 0.0 ? sm[2][0] : 1.0;
        const double sr = sqrt(sm[0][0] / cnt), sp = sqrt(sm[1][0] / cnt);
        scal[0] = sr;
        scal[1] = sr > 0.0 ? 1.0 / sr : 0.0;
        scal[2] = sp;
        scal[3] = sp > 0.0 ? 1.0 / sp : 0.0;
    }
}

// network input on the padded grid (channels-last, 8 BF16 channels)
__global__ void __launch_bounds__(256)
k_cnn_input(const float *__restrict__ r0, const float *__restrict__ r1, const double *__restrict__ p0,
            const double *__restrict__ p1, const double *__restrict__ b0, const double *__restrict__ b1,
            const double *__restrict__ ax, const double *__restrict__ ay, const double *__restrict__ az, int sx,
            int sy, int sz, int sxp, int H, FastDiv dsx, FastDiv dsy, const double *__restrict__ scal,
            __bf16 *__restrict__ X) {
    const int P = sx * sy * sz;
    const int idx = (int)blockIdx.x * 256 + (int)threadIdx.x;
    if (idx >= P) return;
    const int row = (int)fdiv((uint32_t)idx, dsx), ii = idx - row * sx;
    const int kk = (int)fdiv((uint32_t)row, dsy), jj = row - kk * sy;
    const int colour = (ii + jj + kk) & 1;
    const int64_t f = (int64_t)row * H + (ii >> 1);
    const double bd = (colour ? b1 : b0)[f];
    bf16x8 o = {};
    if (bd == bd) {
        const double diag = -((ax[sxp + ii] + ay[sy + jj]) + az[sz + kk]);
        o[0] = (__bf16)(float)((double)(colour ? r1 : r0)[f] / diag * scal[1]);
        o[1] = (__bf16)1.0f;
        o[2] = (__bf16)(float)((colour ? p1 : p0)[f] * scal[3]);
    }
    *reinterpret_cast<bf16x8 *>(X + (int64_t)idx * 8) = o;
}

// guess = p_prev + s_r * net on fluid cells (both colours)
__global__ void k_cnn_apply(const float *__restrict__ net, const double *__restrict__ pp0,
                            const double *__restrict__ pp1, const double *__restrict__ b0,
                            const double *__restrict__ b1, double *__restrict__ pc0, double *__restrict__ pc1,
                            int sx, int sy, int sz, int H, FastDiv dsx, FastDiv dsy,
                            const double *__restrict__ scal) {
    const int P = sx * sy * sz;
    const int idx = (int)blockIdx.x * 256 + (int)threadIdx.x;
    if (idx >= P) return;
    const int row = (int)fdiv((uint32_t)idx, dsx), ii = idx - row * sx;
    const int kk = (int)fdiv((uint32_t)row, dsy), jj = row - kk * sy;
    const int colour = (ii + jj + kk) & 1;
    const int64_t f = (int64_t)row * H + (ii >> 1);
    const double bd = (colour ? b1 : b0)[f];
    if (bd == bd) (colour ? pc1 : pc0)[f] = (colour ? pp1 : pp0)[f] + scal[0] * (double)net[idx];
}

// keep the guess only if its residual is below accept x the previous-step
// guess's;
// flag[0] = 1 when the guess is kept
__global__ void k_cnn_select(const double *__restrict__ rel_prev, const double *__restrict__ rel_guess,
                             const double *__restrict__ pp0, const double *__restrict__ pp1, double *__restrict__ pc0,
                             double *__restrict__ pc1, int64_t NC, double *__restrict__ flag, double accept) {
    const bool keep = *rel_guess < accept * *rel_prev;
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q == 0) flag[0] = keep ? 1.0 : 0.0;
    if (keep || q >= NC) return;
    pc0[q] = pp0[q];
    pc1[q] = pp1[q];
}

// training target (p_final - p_prev) / s_r on the padded grid (0 off fluid)
__global__ void k_cnn_target(const double *__restrict__ pc0, const double *__restrict__ pc1,
                             const double *__restrict__ pp0, const double *__restrict__ pp1,
                             const double *__restrict__ b0, const double *__restrict__ b1, int sx, int sy, int sz,
                             int64_t H, const double *__restrict__ scal, float *__restrict__ T) {
    const int64_t P = (int64_t)sx * sy * sz;
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= P) return;
    const int ii = (int)(idx % sx);
    const int64_t row = idx / sx;
    const int jj = (int)(row % sy), kk = (int)(row / sy);
    const int colour = (ii + jj + kk) & 1;
    const int64_t f = row * H + (ii >> 1);
    const double bd = (colour ? b1 : b0)[f];
    T[idx] = bd == bd ? (float)(((colour ? pc1 : pc0)[f] - (colour ? pp1 : pp0)[f]) * scal[1]) : 0.0f;
}

// ------------------------------------------------------------ host state
struct Net {
    bool ready = false;
    float *prm = nullptr, *grad = nullptr, *am = nullptr, *av = nullptr;
    int64_t adam_t = 0;
    __bf16 *wf[4] = {}, *wb[4] = {};
    float *bias = nullptr;  // [4][16]
    // work buffers for nb volumes of nvox voxels
    int64_t cap_vox = 0;
    __bf16 *x8 = nullptr, *a1 = nullptr, *a2 = nullptr, *a3 = nullptr, *za = nullptr, *zb = nullptr;
    float *out = nullptr, *tgt = nullptr, *part = nullptr;
    int64_t cap_part = 0;
    double *lpart = nullptr, *dsc = nullptr;
    int64_t cap_lpart = 0;
    // training set: per recorded solve, 8-channel input and FP32 target grids
    std::vector<__bf16 *> set_x;
    std::vector<float *> set_t;
    int64_t set_sx = 0, set_sy = 0, set_sz = 0, set_nx = 0, set_ny = 0, set_nz = 0;
    // solver integration
    int record = 0, guess = 0;
    double accept = 1.0;  // keep the guess when rel_guess < accept * rel_prev
    double *pprev[2] = {};
    int64_t pprev_nc = 0;
    int64_t solves = 0, kept = 0;
    double last_rel_prev = 0.0, last_rel_guess = 0.0, last_kept = 0.0, last_ms = 0.0;
};
static Net g_net;

static int net_alloc(int64_t nvox) {
    Net &N = g_net;
    if (nvox <= N.cap_vox) return 0;
    IB3D_CHECK(hipStreamSynchronize(ctx().stream));
    __bf16 **b16[] = {&N.a1, &N.a2, &N.a3, &N.za, &N.zb};
    for (auto p : b16) {
        if (*p) IB3D_CHECK(hipFree(*p));
        IB3D_CHECK(hipMalloc(p, nvox * 16 * sizeof(__bf16)));
    }
    if (N.x8) IB3D_CHECK(hipFree(N.x8));
    IB3D_CHECK(hipMalloc(&N.x8, nvox * 8 * sizeof(__bf16)));
    if (N.out) IB3D_CHECK(hipFree(N.out));
    IB3D_CHECK(hipMalloc(&N.out, nvox * sizeof(float)));
    if (N.tgt) IB3D_CHECK(hipFree(N.tgt));
    IB3D_CHECK(hipMalloc(&N.tgt, nvox * sizeof(float)));
    N.cap_vox = nvox;
    return 0;
}

static int pack_all() {
    Net &N = g_net;
    hipStream_t s = ctx().stream;
    for (int L = 0; L < 4; L++) {
        const int cinst = L == 0 ? 8 : 16;
        hipLaunchKernelGGL(k_pack_fwd, dim3(nblocks(kNkc[L] * 512, 256)), dim3(256), 0, s, N.prm, kWoff[L],
                           kBoff[L], kCin[L], kCout[L], cinst, kNkc[L], L > 0 ? 1 : 0, N.wf[L], N.bias + 16 * L);
        if (L > 0)
            hipLaunchKernelGGL(k_pack_bwd, dim3(nblocks(14 * 512, 256)), dim3(256), 0, s, N.prm, kWoff[L], kCin[L],
                               kCout[L], N.wb[L]);
    }
    IB3D_CHECK(hipGetLastError());
    return 0;
}

void cnn_release() {
    Net &N = g_net;
    float **f[] = {&N.prm, &N.grad, &N.am, &N.av, &N.bias, &N.out, &N.tgt, &N.part};
    for (auto p : f)
        if (*p) (void)hipFree(*p);
    __bf16 **b[] = {&N.x8, &N.a1, &N.a2, &N.a3, &N.za, &N.zb};
    for (auto p : b)
        if (*p) (void)hipFree(*p);
    for (int L = 0; L < 4; L++) {
        if (N.wf[L]) (void)hipFree(N.wf[L]);
        if (N.wb[L]) (void)hipFree(N.wb[L]);
    }
    if (N.lpart) (void)hipFree(N.lpart);
    if (N.dsc) (void)hipFree(N.dsc);
    for (auto p : N.set_x) (void)hipFree(p);
    for (auto p : N.set_t) (void)hipFree(p);
    for (auto p : N.pprev)
        if (p) (void)hipFree(p);
    N = Net();
}

static int net_init(uint64_t seed) {
    Net &N = g_net;
    if (!N.prm) {
        float **f[] = {&N.prm, &N.grad, &N.am, &N.av};
        for (auto p : f) IB3D_CHECK(hipMalloc(p, kNParam * sizeof(float)));
        IB3D_CHECK(hipMalloc(&N.bias, 64 * sizeof(float)));
        for (int L = 0; L < 4; L++) {
            IB3D_CHECK(hipMalloc(&N.wf[L], kNkc[L] * 512 * sizeof(__bf16)));
            if (L > 0) IB3D_CHECK(hipMalloc(&N.wb[L], 14 * 512 * sizeof(__bf16)));
        }
        IB3D_CHECK(hipMalloc(&N.dsc, 16 * sizeof(double)));
    }
    // He-normal weights, zero biases
    std::vector<float> h(kNParam, 0.0f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (int L = 0; L < 4; L++) {
        const double sd = std::sqrt(2.0 / (27.0 * kCin[L]));
        for (int q = 0; q < 27 * kCin[L] * kCout[L]; q++) h[kWoff[L] + q] = (float)(sd * nd(rng));
    }
    IB3D_CHECK(h2d(N.prm, h.data(), kNParam * sizeof(float)));
    IB3D_CHECK(dzero(N.am, kNParam * sizeof(float)));
    IB3D_CHECK(dzero(N.av, kNParam * sizeof(float)));
    N.adam_t = 0;
    N.ready = true;
    return pack_all();
}


static dim3 conv_grid(const Vol &v, int *ntz) {
    *ntz = (v.vz + kCtZ - 1) / kCtZ;
    return dim3((unsigned)((v.vx + kCtX - 1) / kCtX), (unsigned)((v.vy + kCtY - 1) / kCtY), (unsigned)(*ntz * v.nb));
}

static void launch_fwd(int L, const Vol &v, const __bf16 *in, __bf16 *out16, float *out32) {
    Net &N = g_net;
    hipStream_t s = ctx().stream;
    const int nzc = (v.vz + kZmZ - 1) / kZmZ;
    const dim3 grid((unsigned)((v.vx + kZmX - 1) / kZmX), (unsigned)((v.vy + kZmY - 1) / kZmY),
                    (unsigned)(nzc * v.nb)),
        blk(256);
    const bf16x8 *W = reinterpret_cast<const bf16x8 *>(N.wf[L]);
    if (L == 0) {
        const int nrc = (v.vz + kRgZ - 1) / kRgZ;
        const dim3 rgrid((unsigned)((v.vx + kRgX - 1) / kRgX), (unsigned)((v.vy + kRgY - 1) / kRgY),
                         (unsigned)(nrc * v.nb));
        hipLaunchKernelGGL((k_conv_ring<8, 0>), rgrid, blk, 0, s, in, W, N.bias, v, out16, nullptr, nrc);
    }
    else if (L < 3)
        hipLaunchKernelGGL((k_conv_zm<16, 0>), grid, blk, 0, s, in, W, N.bias + 16 * L, v, out16, nullptr, nzc);
    else
        hipLaunchKernelGGL((k_conv_zm<16, 1>), grid, blk, 0, s, in, W, N.bias + 48, v, nullptr, out32, nzc);
}

static void launch_bwd_data(int L, const Vol &v, const __bf16 *dz, const __bf16 *gate, __bf16 *out) {
    Net &N = g_net;
    int ntz;
    const dim3 grid = conv_grid(v, &ntz);
    hipLaunchKernelGGL((k_conv_lds<16, true, 2>), grid, dim3(256), 0, ctx().stream, dz,
                       reinterpret_cast<const bf16x8 *>(N.wb[L]), nullptr, v, out, nullptr, gate, ntz);
}

static int launch_dw(int L, const Vol &v, const __bf16 *x, const __bf16 *dz) {
    Net &N = g_net;
    hipStream_t s = ctx().stream;
    const int64_t nch = (int64_t)v.nb * v.vz * v.vy * ((v.vx + 31) / 32);
    const int nblk = (int)std::min<int64_t>(1024, std::max<int64_t>(1, nch / 8));
    const int cpb = (int)((nch + nblk - 1) / nblk);
    const int64_t need = (int64_t)nblk * 28 * 256;
    if (need > N.cap_part) {
        IB3D_CHECK(hipStreamSynchronize(s));
        if (N.part) IB3D_CHECK(hipFree(N.part));
        IB3D_CHECK(hipMalloc(&N.part, need * sizeof(float)));
        N.cap_part = need;
    }
    if (L == 0)
        hipLaunchKernelGGL(k_conv_dw<8>, dim3(nblk), dim3(256), 0, s, x, dz, v, N.part, cpb, nch);
    else
        hipLaunchKernelGGL(k_conv_dw<16>, dim3(nblk), dim3(256), 0, s, x, dz, v, N.part, cpb, nch);
    const int nout = 27 * kCin[L] * kCout[L] + kCout[L];
    hipLaunchKernelGGL(k_dw_reduce, dim3(nblocks(nout, 256)), dim3(256), 0, s, N.part, nblk, kCin[L], kCout[L],
                       kWoff[L], kBoff[L], N.grad);
    IB3D_CHECK(hipGetLastError());
    return 0;
}

// forward of all layers on g_net.x8 -> g_net.out (activations kept)
static void forward(const Vol &v) {
    Net &N = g_net;
    launch_fwd(0, v, N.x8, N.a1, nullptr);
    launch_fwd(1, v, N.a1, N.a2, nullptr);
    launch_fwd(2, v, N.a2, N.a3, nullptr);
    launch_fwd(3, v, N.a3, nullptr, N.out);
}

// backward from dZ4 in g_net.za -> g_net.grad
static int backward(const Vol &v) {
    Net &N = g_net;
    if (launch_dw(3, v, N.a3, N.za)) return -1;
    launch_bwd_data(3, v, N.za, N.a3, N.zb);
    if (launch_dw(2, v, N.a2, N.zb)) return -1;
    launch_bwd_data(2, v, N.zb, N.a2, N.za);
    if (launch_dw(1, v, N.a1, N.za)) return -1;
    launch_bwd_data(1, v, N.za, N.a1, N.zb);
    if (launch_dw(0, v, N.x8, N.zb)) return -1;
    IB3D_CHECK(hipGetLastError());
    return 0;
}

static int adam_step(double lr, double b1, double b2, double eps) {
    Net &N = g_net;
    N.adam_t++;
    const double c1 = 1.0 - std::pow(b1, (double)N.adam_t), c2 = 1.0 - std::pow(b2, (double)N.adam_t);
    hipLaunchKernelGGL(k_adam, dim3(nblocks(kNParam, 256)), dim3(256), 0, ctx().stream, N.prm, N.grad, N.am, N.av,
                       kNParam, (float)lr, (float)b1, (float)b2, (float)eps, (float)c1, (float)c2);
    IB3D_CHECK(hipGetLastError());
    return pack_all();
}

// ------------------------------------------------ solver integration
static int ensure_pprev(Ctx &c) {
    Net &N = g_net;
    if (N.pprev_nc == c.NC) return 0;
    for (auto &p : N.pprev) {
        if (p) IB3D_CHECK(hipFree(p));
        IB3D_CHECK(hipMalloc(&p, c.NC * sizeof(double)));
    }
    N.pprev_nc = c.NC;
    return 0;
}

// r32 holds the FP32 residual of the current pc (= p_prev); builds the
// network input for this solve into `dst` (8-channel grid) and the scales
// into g_net.dsc[0..3]
static int build_input(Ctx &c, __bf16 *dst) {
    Net &N = g_net;
    IB3D_REQUIRE(c.P < (int64_t)1 << 31, "cnn: the grid must hold fewer than 2^31 cells");
    const FastDiv dsx = make_fastdiv((int)c.sx), dsy = make_fastdiv((int)c.sy);
    const int64_t nb = nblocks(c.P, 256), nr = std::min<int64_t>(nb, 1024);
    if (ensure_partial(3 * nr)) return -1;
    hipLaunchKernelGGL(k_cnn_scales, dim3(nr), dim3(256), 0, c.stream, c.r32[0], c.r32[1], c.pc[0], c.pc[1],
                       c.bdt[0], c.bdt[1], c.ax, c.ay, c.az, (int)c.sx, (int)c.sy, (int)c.sz, (int)c.sxp, (int)c.H,
                       dsx, dsy, c.partial);
    hipLaunchKernelGGL(k_cnn_scales_finish, dim3(1), dim3(256), 0, c.stream, c.partial, nr, N.dsc);
    hipLaunchKernelGGL(k_cnn_input, dim3(nb), dim3(256), 0, c.stream, c.r32[0], c.r32[1], c.pc[0], c.pc[1], c.bdt[0],
                       c.bdt[1], c.ax, c.ay, c.az, (int)c.sx, (int)c.sy, (int)c.sz, (int)c.sxp, (int)c.H, dsx, dsy,
                       N.dsc, dst);
    IB3D_CHECK(hipGetLastError());
    return 0;
}

bool cnn_active() { return g_net.ready && (g_net.record || g_net.guess); }

// Called by the L2 solve after the residual of p_prev has been written to
// r32 and its relative norm to *rel_prev.  Keeps a copy of p_prev; records
// the input (record mode) and/or replaces pc by the network's guess when
// that lowers the residual (guess mode).  `resid` evaluates the relative
// residual of the current pc into the given slot.
int cnn_before_solve(Ctx &c, const double *rel_prev, double *rel_guess, int (*resid)(Ctx &, double *)) {
    Net &N = g_net;
    if (ensure_pprev(c)) return -1;
    IB3D_CHECK(hipMemcpyAsync(N.pprev[0], c.pc[0], c.NC * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    IB3D_CHECK(hipMemcpyAsync(N.pprev[1], c.pc[1], c.NC * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    if (N.record) {
        if (N.set_x.empty()) {
            N.set_sx = c.sx; N.set_sy = c.sy; N.set_sz = c.sz;
            N.set_nx = c.nx; N.set_ny = c.ny; N.set_nz = c.nz;
        }
        IB3D_REQUIRE(N.set_sx == c.sx && N.set_sy == c.sy && N.set_sz == c.sz,
                     "cnn: recorded pairs must come from one grid");
        __bf16 *x = nullptr;
        float *t = nullptr;
        IB3D_CHECK(hipMalloc(&x, c.P * 8 * sizeof(__bf16)));
        IB3D_CHECK(hipMalloc(&t, c.P * sizeof(float)));
        N.set_x.push_back(x);
        N.set_t.push_back(t);
        if (build_input(c, x)) return -1;
    }
    if (N.guess) {
        hipEvent_t e0, e1;
        IB3D_CHECK(hipEventCreate(&e0));
        IB3D_CHECK(hipEventCreate(&e1));
        IB3D_CHECK(hipEventRecord(e0, c.stream));
        if (net_alloc(c.P)) return -1;
        if (build_input(c, N.x8)) return -1;
        const Vol v{(int)c.sx, (int)c.sy, (int)c.sz, 1};
        forward(v);
        hipLaunchKernelGGL(k_cnn_apply, dim3(nblocks(c.P, 256)), dim3(256), 0, c.stream, N.out, N.pprev[0],
                           N.pprev[1], c.bdt[0], c.bdt[1], c.pc[0], c.pc[1], (int)c.sx, (int)c.sy, (int)c.sz,
                           (int)c.H, make_fastdiv((int)c.sx), make_fastdiv((int)c.sy), N.dsc);
        if (resid(c, rel_guess)) return -1;
        hipLaunchKernelGGL(k_cnn_select, dim3(nblocks(c.NC, 256)), dim3(256), 0, c.stream, rel_prev, rel_guess,
                           N.pprev[0], N.pprev[1], c.pc[0], c.pc[1], c.NC, N.dsc + 4, N.accept);
        IB3D_CHECK(hipEventRecord(e1, c.stream));
        IB3D_CHECK(hipGetLastError());
        IB3D_CHECK(hipEventSynchronize(e1));
        float ms = 0.0f;
        IB3D_CHECK(hipEventElapsedTime(&ms, e0, e1));
        N.last_ms = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return 0;
}

// after convergence: record the target and the guess statistics
int cnn_after_solve(Ctx &c, const double *rel_prev, const double *rel_guess) {
    Net &N = g_net;
    if (N.record && !N.set_t.empty()) {
        hipLaunchKernelGGL(k_cnn_target, dim3(nblocks(c.P, 256)), dim3(256), 0, c.stream, c.pc[0], c.pc[1],
                           N.pprev[0], N.pprev[1], c.bdt[0], c.bdt[1], (int)c.sx, (int)c.sy, (int)c.sz, c.H, N.dsc,
                           N.set_t.back());
        IB3D_CHECK(hipGetLastError());
    }
    if (N.guess) {
        double h[3];
        IB3D_CHECK(d2h(&h[0], rel_prev, sizeof(double)));
        IB3D_CHECK(d2h(&h[1], rel_guess, sizeof(double)));
        IB3D_CHECK(d2h(&h[2], N.dsc + 4, sizeof(double)));
        N.last_rel_prev = h[0];
        N.last_rel_guess = h[1];
        N.last_kept = h[2];
        N.solves++;
        N.kept += h[2] > 0.5 ? 1 : 0;
    }
    return 0;
}

// ------------------------------------------------------------ training
static int train(int epochs, int batch, double lr, uint64_t seed, double *loss_out) {
    Net &N = g_net;
    Ctx &c = ctx();
    IB3D_REQUIRE(N.ready, "cnn_train: network not initialised");
    IB3D_REQUIRE(!N.set_x.empty(), "cnn_train: no recorded pairs");
    IB3D_REQUIRE(batch >= 1 && batch <= 64, "cnn_train: batch must be 1..64");
    // tiles: 64^3 output blocks covering the interior cells of the grid
    std::vector<int> items;
    for (size_t s = 0; s < N.set_x.size(); s++)
        for (int64_t oz = 1; oz <= N.set_nz; oz += kTile)
            for (int64_t oy = 1; oy <= N.set_ny; oy += kTile)
                for (int64_t ox = 1; ox <= N.set_nx; ox += kTile) {
                    items.push_back((int)s);
                    items.push_back((int)ox);
                    items.push_back((int)oy);
                    items.push_back((int)oz);
                }
    const int npairs = (int)(items.size() / 4);
    const Vol v{kTileIn, kTileIn, kTileIn, batch};
    const int64_t nvox = (int64_t)batch * kTileIn * kTileIn * kTileIn;
    if (net_alloc(nvox)) return -1;
    const int64_t nlb = nblocks(nvox, 256);
    if (nlb > N.cap_lpart) {
        if (N.lpart) IB3D_CHECK(hipFree(N.lpart));
        IB3D_CHECK(hipMalloc(&N.lpart, nlb * sizeof(double)));
        N.cap_lpart = nlb;
    }
    __bf16 **gx = nullptr;
    float **gt = nullptr;
    int *ditems = nullptr;
    IB3D_CHECK(hipMalloc(&gx, N.set_x.size() * sizeof(void *)));
    IB3D_CHECK(hipMalloc(&gt, N.set_t.size() * sizeof(void *)));
    IB3D_CHECK(hipMalloc(&ditems, 4 * batch * sizeof(int)));
    IB3D_CHECK(h2d(gx, N.set_x.data(), N.set_x.size() * sizeof(void *)));
    IB3D_CHECK(h2d(gt, N.set_t.data(), N.set_t.size() * sizeof(void *)));
    std::mt19937_64 rng(seed);
    std::vector<int> order(npairs);
    const float scale = 1.0f / (float)((double)batch * kTile * kTile * kTile);
    for (int ep = 0; ep < epochs; ep++) {
        for (int q = 0; q < npairs; q++) order[q] = q;
        std::shuffle(order.begin(), order.end(), rng);
        IB3D_CHECK(hipMemsetAsync(N.dsc + 8, 0, sizeof(double), c.stream));
        int nbatches = 0;
        for (int q0 = 0; q0 + batch <= npairs; q0 += batch) {
            std::vector<int> bi(4 * batch);
            for (int b = 0; b < batch; b++)
                for (int k = 0; k < 4; k++) bi[4 * b + k] = items[4 * order[q0 + b] + k];
            IB3D_CHECK(h2d(ditems, bi.data(), bi.size() * sizeof(int)));
            hipLaunchKernelGGL(k_crop, dim3(nblocks(nvox, 256)), dim3(256), 0, c.stream, gx, gt, ditems,
                               (int)N.set_sx, (int)N.set_sy, (int)N.set_sz, v, N.x8, N.tgt);
            forward(v);
            hipLaunchKernelGGL(k_loss, dim3(nlb), dim3(256), 0, c.stream, N.out, N.tgt, nvox, scale, N.za, N.lpart);
            hipLaunchKernelGGL(k_sum_dbl, dim3(1), dim3(256), 0, c.stream, N.lpart, nlb, N.dsc + 8);
            if (backward(v)) return -1;
            if (adam_step(lr, 0.9, 0.999, 1e-8)) return -1;
            nbatches++;
        }
        double tot = 0.0;
        IB3D_CHECK(d2h(&tot, N.dsc + 8, sizeof(double)));
        if (loss_out) loss_out[ep] = nbatches ? tot * scale / nbatches : 0.0;
    }
    IB3D_CHECK(hipFree(gx));
    IB3D_CHECK(hipFree(gt));
    IB3D_CHECK(hipFree(ditems));
    return 0;
}

}  // namespace ib3d

using namespace ib3d;

extern "C" {

int ib3d_cnn_init(uint64_t seed) {
    IB3D_REQUIRE(ctx().ready, "cnn_init: ib3d_init has not run");
    return net_init(seed);
}

int64_t ib3d_cnn_param_count(void) { return kNParam; }

int ib3d_cnn_set_params(const float *p, int64_t n) {
    IB3D_REQUIRE(g_net.ready, "cnn_set_params: network not initialised");
    IB3D_REQUIRE(n == kNParam, "cnn_set_params: expected 15601 parameters");
    IB3D_CHECK(h2d(g_net.prm, p, n * sizeof(float)));
    IB3D_CHECK(dzero(g_net.am, kNParam * sizeof(float)));
    IB3D_CHECK(dzero(g_net.av, kNParam * sizeof(float)));
    g_net.adam_t = 0;
    return pack_all();
}

int ib3d_cnn_get_params(float *p, int64_t n) {
    IB3D_REQUIRE(g_net.ready, "cnn_get_params: network not initialised");
    IB3D_REQUIRE(n == kNParam, "cnn_get_params: expected 15601 parameters");
    IB3D_CHECK(d2h(p, g_net.prm, n * sizeof(float)));
    return 0;
}

int ib3d_cnn_forward(const float *x, int64_t nb, int64_t vx, int64_t vy, int64_t vz, float *y) {
    Net &N = g_net;
    IB3D_REQUIRE(N.ready, "cnn_forward: network not initialised");
    IB3D_REQUIRE(nb >= 1 && vx >= 1 && vy >= 1 && vz >= 1, "cnn_forward: bad shape");
    const int64_t nvox = nb * vx * vy * vz;
    if (net_alloc(nvox)) return -1;
    float *tmp = nullptr;
    IB3D_CHECK(hipMalloc(&tmp, nvox * 3 * sizeof(float)));
    IB3D_CHECK(h2d(tmp, x, nvox * 3 * sizeof(float)));
    hipLaunchKernelGGL(k_host_to_x8, dim3(nblocks(nvox, 256)), dim3(256), 0, ctx().stream, tmp, nvox, N.x8);
    forward(Vol{(int)vx, (int)vy, (int)vz, (int)nb});
    IB3D_CHECK(hipGetLastError());
    IB3D_CHECK(d2h(y, N.out, nvox * sizeof(float)));
    IB3D_CHECK(hipFree(tmp));
    return 0;
}

int ib3d_cnn_gradient(const float *x, const float *dy, int64_t nb, int64_t vx, int64_t vy, int64_t vz,
                      float *grad) {
    Net &N = g_net;
    IB3D_REQUIRE(N.ready, "cnn_gradient: network not initialised");
    IB3D_REQUIRE(nb >= 1 && vx >= 1 && vy >= 1 && vz >= 1, "cnn_gradient: bad shape");
    const int64_t nvox = nb * vx * vy * vz;
    if (net_alloc(nvox)) return -1;
    float *tmp = nullptr;
    IB3D_CHECK(hipMalloc(&tmp, nvox * 3 * sizeof(float)));
    IB3D_CHECK(h2d(tmp, x, nvox * 3 * sizeof(float)));
    hipLaunchKernelGGL(k_host_to_x8, dim3(nblocks(nvox, 256)), dim3(256), 0, ctx().stream, tmp, nvox, N.x8);
    const Vol v{(int)vx, (int)vy, (int)vz, (int)nb};
    forward(v);
    IB3D_CHECK(h2d(tmp, dy, nvox * sizeof(float)));
    hipLaunchKernelGGL(k_dy_to_dz, dim3(nblocks(nvox, 256)), dim3(256), 0, ctx().stream, tmp, nvox, N.za);
    if (backward(v)) return -1;
    IB3D_CHECK(d2h(grad, N.grad, kNParam * sizeof(float)));
    IB3D_CHECK(hipFree(tmp));
    return 0;
}

int ib3d_cnn_adam(const float *grad, double lr, double beta1, double beta2, double eps) {
    Net &N = g_net;
    IB3D_REQUIRE(N.ready, "cnn_adam: network not initialised");
    IB3D_CHECK(h2d(N.grad, grad, kNParam * sizeof(float)));
    return adam_step(lr, beta1, beta2, eps);
}

int ib3d_cnn_set_accept(double ratio) {
    IB3D_REQUIRE(ratio > 0.0 && ratio <= 1.0, "cnn_set_accept: ratio must lie in (0, 1]");
    g_net.accept = ratio;
    return 0;
}

int ib3d_cnn_config(int record, int guess) {
    IB3D_REQUIRE(g_net.ready || (!record && !guess), "cnn_config: network not initialised");
    g_net.record = record;
    g_net.guess = guess;
    return 0;
}

int64_t ib3d_cnn_pair_count(void) { return (int64_t)g_net.set_x.size(); }

int ib3d_cnn_clear_pairs(void) {
    Net &N = g_net;
    IB3D_CHECK(hipStreamSynchronize(ctx().stream));
    for (auto p : N.set_x) IB3D_CHECK(hipFree(p));
    for (auto p : N.set_t) IB3D_CHECK(hipFree(p));
    N.set_x.clear();
    N.set_t.clear();
    return 0;
}

int ib3d_cnn_train(int epochs, int batch, double lr, uint64_t seed, double *loss_per_epoch) {
    IB3D_REQUIRE(epochs >= 1, "cnn_train: epochs must be >= 1");
    return train(epochs, batch, lr, seed, loss_per_epoch);
}

int ib3d_cnn_release(void) {
    if (ctx().stream) IB3D_CHECK(hipStreamSynchronize(ctx().stream));
    cnn_release();
    return 0;
}

int ib3d_cnn_stats(int64_t *solves, int64_t *kept, double *rel_prev, double *rel_guess, double *guess_ms) {
    const Net &N = g_net;
    *solves = N.solves;
    *kept = N.kept;
    *rel_prev = N.last_rel_prev;
    *rel_guess = N.last_rel_guess;
    *guess_ms = N.last_ms;
    return 0;
}

}  // extern "C"
