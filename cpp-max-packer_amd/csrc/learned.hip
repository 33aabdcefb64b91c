// Learned orientation/region proposal (no counterpart in the reference): a
// small BF16 3D CNN whose convolutions run on gfx950 MFMA, with hand-written
// forward and training (backward + Adam) kernels.
//
// Network (activations NDHWC, 32 channels innermost, BF16; FP32 accumulate):
//   x0  64^3 x 1    occupancy fraction q/255 of the environment (q: uint8)
//   a1  32^3 x 32   ReLU(conv3x3x3 stride 2 (x0) + b1)
//   a2  16^3 x 32   ReLU(conv stride 2 (a1) + b2 + Wf . descriptor)
//   a3  16^3 x 32   ReLU(conv stride 1 (a2) + b3)
//   a4  16^3 x 32   ReLU(conv stride 1 (a3) + b4)
//   heat[v]   = a4[v] . wh + bh                       (16^3 region logits)
//   g         = mean over v of a4[v]                  (32)
//   f         = sum over v of softmax(heat)[v] a4[v]  (32, pooled where the heat points)
//   hid       = ReLU(Wa [g, f, descriptor] + ba)      (64)
//   logits    = Wo hid + bo                           (96 orientation logits)
// Convolutions are implicit GEMMs on v_mfma_f32_32x32x16_bf16: rows are
// output voxels (forward) or input voxels (input gradient), columns the 32
// channels, and the reduction runs over 27 taps x 32 channels.  The weight
// gradient reduces over voxels with operands transposed through LDS.
// Every reduction has a fixed order, so results are deterministic.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "session.hpp"

namespace sphip {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
using u16 = unsigned short;
using u8 = unsigned char;

constexpr int kIn = 64;                  // input grid per axis
constexpr int kV0 = kIn * kIn * kIn;     // 262144
constexpr int kCh = 32;
constexpr int kV1 = 32 * 32 * 32;        // after layer 1
constexpr int kV2 = 16 * 16 * 16;        // after layers 2-4 (regions)
constexpr int kTaps = 27;
constexpr int kK1 = 32;                  // layer-1 reduction (27 taps padded)
constexpr int kK = kTaps * kCh;          // 864
constexpr int kDesc = 32;
constexpr int kHidden = 64;
constexpr int kOrient = 96;
constexpr int kHeadIn = 2 * 32 + 32;     // [g, f, descriptor]
constexpr int kPoolBlocks = 16;          // pooling blocks per sample (256 voxels each)
constexpr int kPoolPart = 2 + 2 * 32;    // per block: max, sum of exp, exp-weighted and plain channel sums

// Flat FP32 parameter layout.
constexpr int P_W1 = 0, P_B1 = P_W1 + kCh * kTaps;
constexpr int P_W2 = P_B1 + kCh, P_B2 = P_W2 + kCh * kK;
constexpr int P_W3 = P_B2 + kCh, P_B3 = P_W3 + kCh * kK;
constexpr int P_W4 = P_B3 + kCh, P_B4 = P_W4 + kCh * kK;
constexpr int P_WF = P_B4 + kCh;
constexpr int P_WH = P_WF + kCh * kDesc, P_BH = P_WH + kCh;
constexpr int P_WA = P_BH + 1, P_BA = P_WA + kHidden * kHeadIn;
constexpr int P_WO = P_BA + kHidden, P_BO = P_WO + kOrient * kHidden;
constexpr int P_TOTAL = P_BO + kOrient;
constexpr int kHeadParams = P_TOTAL - P_WH;  // wh .. bo, contiguous

constexpr int kRowsPerChunk = 256;  // weight-gradient rows per block
constexpr int kSumSegments = 16;    // channel-sum segments per sample

__device__ __forceinline__ u16 f2bf(float f)
{
    // round to nearest even (finite inputs)
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u16(u >> 16);
}
__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float(unsigned(h) << 16); }

__device__ __forceinline__ bf16x8 load8(const u16* p)
{
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, v);
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, uint4 { 0, 0, 0, 0 }); }
__device__ __forceinline__ bf16x8 pack8(const u16 (&h)[8])
{
    uint4 v;
    v.x = unsigned(h[0]) | (unsigned(h[1]) << 16);
    v.y = unsigned(h[2]) | (unsigned(h[3]) << 16);
    v.z = unsigned(h[4]) | (unsigned(h[5]) << 16);
    v.w = unsigned(h[6]) | (unsigned(h[7]) << 16);
    return __builtin_bit_cast(bf16x8, v);
}
__device__ __forceinline__ u16 input_value(u8 q) { return f2bf(float(q) * (1.0f / 255.0f)); }

// Accumulator element r of lane l holds D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
__device__ __forceinline__ int acc_row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// ---- input: occupancy -> 64^3 fractions ------------------------------------
// Bin b of an axis of extent e covers cells [floor(b e / 64), floor((b+1) e / 64)).
__global__ void k_downsample(const u8* __restrict__ env, int ex, int ey, int ez, u8* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kV0) return;
    const int bx = i % kIn, by = (i / kIn) % kIn, bz = i / (kIn * kIn);
    const int x0 = bx * ex / kIn, x1 = (bx + 1) * ex / kIn;
    const int y0 = by * ey / kIn, y1 = (by + 1) * ey / kIn;
    const int z0 = bz * ez / kIn, z1 = (bz + 1) * ez / kIn;
    unsigned count = 0;
    for (int z = z0; z < z1; ++z)
        for (int y = y0; y < y1; ++y) {
            const u8* row = env + (long long)ex * (y + (long long)ey * z);
            for (int x = x0; x < x1; ++x) count += row[x] ? 1u : 0u;
        }
    const unsigned cells = unsigned(x1 - x0) * unsigned(y1 - y0) * unsigned(z1 - z0);
    out[i] = u8((255u * count + cells / 2) / cells);
}

// ---- convolution as implicit GEMM on MFMA -----------------------------------
// MODE 0 (forward): rows = output voxels at resolution RES, source = input at
//   RES * S; A[row][t, ci] = x[S v + d - 1][ci], B = W[co][t][ci].
// MODE 1 (input gradient, CIN = 32): rows = input voxels at RES, source = dz at
//   RES / S; A[p][t, co] = dz[v][co] where S v = p + 1 - d, B = W[co][t][ci]
//   read from the transposed copy Wt[ci][t][co]; output gated by act[p] > 0.
// Layer 1 (CIN = 1, 2 MFMA steps per tile): one 32-row tile per wave, 4 per
// block.  CIN = 32: one 32-row tile per block, its 27 taps split over the 4
// waves (tap t on wave t % 4) and the partial tiles summed through LDS, which
// keeps the batch-1 inference path spread over 4x more CUs.
template <int MODE, int S, int RES>
__device__ __forceinline__ bool conv_source(int rx, int ry, int rz, int t, int& sx, int& sy, int& sz)
{
    constexpr int SRES = MODE == 0 ? RES * S : RES / S;
    const int dx = t % 3, dy = (t / 3) % 3, dz = t / 9;
    if constexpr (MODE == 0) {
        sx = S * rx + dx - 1;
        sy = S * ry + dy - 1;
        sz = S * rz + dz - 1;
        return sx >= 0 && sy >= 0 && sz >= 0 && sx < SRES && sy < SRES && sz < SRES;
    }
    else {
        const int qx = rx + 1 - dx, qy = ry + 1 - dy, qz = rz + 1 - dz;
        bool ok = qx >= 0 && qy >= 0 && qz >= 0 && (qx % S) == 0 && (qy % S) == 0 && (qz % S) == 0;
        sx = qx / S;
        sy = qy / S;
        sz = qz / S;
        return ok && sx < SRES && sy < SRES && sz < SRES;
    }
}

template <int MODE, int CIN, int S, int RES>
__global__ __launch_bounds__(256) void k_conv(const void* __restrict__ src_v, const unsigned* __restrict__ sample,
                                              const u16* __restrict__ w, const float* __restrict__ bias,
                                              const float* __restrict__ film, const u16* __restrict__ gate,
                                              u16* __restrict__ out)
{
    constexpr int ROWS = RES * RES * RES;
    constexpr int SRES = MODE == 0 ? RES * S : RES / S;
    constexpr int SVOX = SRES * SRES * SRES;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, hi = lane >> 5, c32 = lane & 31;
    const int b = blockIdx.y;
    const long long base = (long long)b * ROWS;
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;
    if constexpr (CIN == 1) {
        const int row0 = (blockIdx.x * 4 + wave) * 32;
        const int r = row0 + c32;
        const int rx = r % RES, ry = (r / RES) % RES, rz = r / (RES * RES);
        const u8* x = static_cast<const u8*>(src_v) + (long long)sample[b] * kV0;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            u16 h[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int t = c * 16 + hi * 8 + j;
                h[j] = 0;
                int sx, sy, sz;
                if (t < kTaps && conv_source<0, S, RES>(rx, ry, rz, t, sx, sy, sz))
                    h[j] = input_value(x[sx + SRES * (sy + SRES * sz)]);
            }
            const bf16x8 a = pack8(h);
            const bf16x8 bv = load8(w + c32 * kK1 + c * 16 + hi * 8);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int row = row0 + acc_row(i, lane);
            const long long at = (base + row) * kCh + c32;
            float v = acc[i] + bias[c32];
            v = v > 0.f ? v : 0.f;
            out[at] = f2bf(v);
        }
    }
    else {
        __shared__ float red[4][32 * 32];
        const int row0 = blockIdx.x * 32;
        const int r = row0 + c32;
        const int rx = r % RES, ry = (r / RES) % RES, rz = r / (RES * RES);
        const u16* x = static_cast<const u16*>(src_v) + (long long)b * SVOX * kCh;
        for (int t = wave; t < kTaps; t += 4) {
            int sx, sy, sz;
            const bool ok = conv_source<MODE, S, RES>(rx, ry, rz, t, sx, sy, sz);
            const u16* srow = x + (long long)(ok ? sx + SRES * (sy + SRES * sz) : 0) * kCh;
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int k0 = c * 16 + hi * 8;
                const bf16x8 a = ok ? load8(srow + k0) : zero8();
                const bf16x8 bv = load8(w + c32 * kK + t * kCh + k0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bv, acc, 0, 0, 0);
            }
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) red[wave][acc_row(i, lane) * 32 + c32] = acc[i];
        __syncthreads();
        for (int e = threadIdx.x; e < 32 * 32; e += 256) {
            const int row = row0 + e / 32, col = e % 32;
            const long long at = (base + row) * kCh + col;
            float v = ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
            if constexpr (MODE == 0) {
                v += bias[col];
                if (film) v += film[b * kCh + col];
                v = v > 0.f ? v : 0.f;
            }
            else {
                v = bf2f(gate[at]) > 0.f ? v : 0.f;
            }
            out[at] = f2bf(v);
        }
    }
}

// Weight gradient: partial[chunk][t][co][ci] = sum over the chunk's rows m of
// dz[m][co] * A[m][t, ci] (A the forward gather of tap t).  Layer 1 has one
// tap block whose 32 columns are the taps (27 used).
template <int CIN, int S, int RES>
__global__ __launch_bounds__(256) void k_conv_wgrad(const void* __restrict__ src_v, const unsigned* __restrict__ sample,
                                                    const u16* __restrict__ dz, float* __restrict__ partial)
{
    constexpr int ROWS = RES * RES * RES;  // output voxels per sample
    constexpr int SRES = RES * S;
    constexpr int SVOX = SRES * SRES * SRES;
    constexpr int TILE = 64, PITCH = TILE + 8;
    __shared__ __attribute__((aligned(16))) u16 dzT[kCh * PITCH];
    __shared__ __attribute__((aligned(16))) u16 xT[kCh * PITCH];
    __shared__ float red[4][kCh * kCh];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, hi = lane >> 5, c32 = lane & 31;
    const int t = blockIdx.x;
    const long long m_base = (long long)blockIdx.y * kRowsPerChunk;
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;
    const int srow = threadIdx.x >> 2, part = threadIdx.x & 3;
    for (int it = 0; it < kRowsPerChunk / TILE; ++it) {
        const long long m = m_base + it * TILE + srow;
        const int b = int(m / ROWS), v = int(m % ROWS);
        const int vx = v % RES, vy = (v / RES) % RES, vz = v / (RES * RES);
        {
            const uint4 q = *reinterpret_cast<const uint4*>(dz + m * kCh + part * 8);
            const u16* h = reinterpret_cast<const u16*>(&q);
#pragma unroll
            for (int j = 0; j < 8; ++j) dzT[(part * 8 + j) * PITCH + srow] = h[j];
        }
        if constexpr (CIN == 1) {
            const u8* x = static_cast<const u8*>(src_v) + (long long)sample[b] * kV0;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int tap = part * 8 + j;
                u16 val = 0;
                if (tap < kTaps) {
                    const int px = S * vx + tap % 3 - 1, py = S * vy + (tap / 3) % 3 - 1, pz = S * vz + tap / 9 - 1;
                    if (px >= 0 && py >= 0 && pz >= 0 && px < SRES && py < SRES && pz < SRES)
                        val = input_value(x[px + SRES * (py + SRES * pz)]);
                }
                xT[tap * PITCH + srow] = val;
            }
        }
        else {
            const int px = S * vx + t % 3 - 1, py = S * vy + (t / 3) % 3 - 1, pz = S * vz + t / 9 - 1;
            const bool ok = px >= 0 && py >= 0 && pz >= 0 && px < SRES && py < SRES && pz < SRES;
            uint4 q { 0, 0, 0, 0 };
            if (ok)
                q = *reinterpret_cast<const uint4*>(static_cast<const u16*>(src_v) +
                                                    ((long long)b * SVOX + px + SRES * (py + SRES * pz)) * kCh +
                                                    part * 8);
            const u16* h = reinterpret_cast<const u16*>(&q);
#pragma unroll
            for (int j = 0; j < 8; ++j) xT[(part * 8 + j) * PITCH + srow] = h[j];
        }
        __syncthreads();
        const int kk = wave * 16 + hi * 8;
        const bf16x8 a = load8(dzT + c32 * PITCH + kk);
        const bf16x8 bv = load8(xT + c32 * PITCH + kk);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bv, acc, 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) red[wave][acc_row(i, lane) * kCh + c32] = acc[i];
    __syncthreads();
    float* dst = partial + ((long long)blockIdx.y * gridDim.x + t) * (kCh * kCh);
    for (int e = threadIdx.x; e < kCh * kCh; e += 256) dst[e] = ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
}

// dW from the partials: conv layers [co][t][ci] (taps = 27) or layer 1 [co][t].
__global__ void k_reduce_wgrad(const float* __restrict__ partial, int chunks, int taps, int layer1, float* __restrict__ dw)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = kCh * kCh;
    if (e >= taps * per) return;
    const int t = e / per, co = (e % per) / kCh, col = e % kCh;
    if (layer1 && col >= kTaps) return;
    float s = 0.f;
    for (int c = 0; c < chunks; ++c) s += partial[((long long)c * taps + t) * per + co * kCh + col];
    if (layer1) dw[co * kTaps + col] = s;
    else dw[co * kK + t * kCh + col] = s;
}

// Per-sample, per-segment channel sums of dz (bias and descriptor gradients).
__global__ void k_channel_sums(const u16* __restrict__ dz, int rows, float* __restrict__ sums)
{
    __shared__ float part[8][kCh];
    const int b = blockIdx.x, seg = blockIdx.y;
    const int c = threadIdx.x & 31, g = threadIdx.x >> 5;
    const int per = rows / kSumSegments;
    const u16* base = dz + ((long long)b * rows + (long long)seg * per) * kCh;
    float s = 0.f;
    for (int v = g; v < per; v += 8) s += bf2f(base[(long long)v * kCh + c]);
    part[g][c] = s;
    __syncthreads();
    if (threadIdx.x < kCh) {
        float t = 0.f;
        for (int i = 0; i < 8; ++i) t += part[i][threadIdx.x];
        sums[(b * kSumSegments + seg) * kCh + threadIdx.x] = t;
    }
}

__global__ void k_film(const float* __restrict__ params, const float* __restrict__ desc, int batch, float* __restrict__ film)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= batch * kCh) return;
    const int b = i / kCh, co = i % kCh;
    float s = 0.f;
    for (int j = 0; j < kDesc; ++j) s += params[P_WF + co * kDesc + j] * desc[b * kDesc + j];
    film[i] = s;
}

__device__ float block_sum(float v, float* sh)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    float t = 0.f;
    for (int i = 0; i < int(blockDim.x >> 6); ++i) t += sh[i];
    return t;
}

__device__ float block_max(float v, float* sh)
{
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_down(v, off, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    float t = sh[0];
    for (int i = 1; i < int(blockDim.x >> 6); ++i) t = fmaxf(t, sh[i]);
    return t;
}

// Heads, first pass: grid (kPoolBlocks, batch), one voxel per thread.  Writes
// the heat logits and per block: the block max m of the heat, sum of
// exp(heat - m), the exp-weighted channel sums and the plain channel sums.
__global__ __launch_bounds__(256) void k_heads_pool(const u16* __restrict__ a4, const float* __restrict__ params,
                                                    float* __restrict__ heat, float* __restrict__ part)
{
    __shared__ float tile[kCh][257];
    __shared__ float red[8][kCh];
    __shared__ float sh[8];
    const int b = blockIdx.y, v = blockIdx.x * 256 + threadIdx.x;
    const u16* a = a4 + ((long long)b * kV2 + v) * kCh;
    float av[kCh];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint4 w = *reinterpret_cast<const uint4*>(a + 8 * q);
        const u16* h = reinterpret_cast<const u16*>(&w);
#pragma unroll
        for (int j = 0; j < 8; ++j) av[8 * q + j] = bf2f(h[j]);
    }
    float hv = 0.f;
#pragma unroll
    for (int c = 0; c < kCh; ++c) hv += av[c] * params[P_WH + c];
    hv += params[P_BH];
    heat[(long long)b * kV2 + v] = hv;
    const float m = block_max(hv, sh);
    const float e = expf(hv - m);
    const float se = block_sum(e, sh);
    float* out = part + ((long long)b * kPoolBlocks + blockIdx.x) * kPoolPart;
    const int c = threadIdx.x & 31, g = threadIdx.x >> 5;
    for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
        for (int k = 0; k < kCh; ++k) tile[k][threadIdx.x] = pass == 0 ? e * av[k] : av[k];
        __syncthreads();
        float t = 0.f;
        for (int r = 0; r < 32; ++r) t += tile[c][g * 32 + r];
        red[g][c] = t;
        __syncthreads();
        if (threadIdx.x < kCh) {
            float u = 0.f;
            for (int i = 0; i < 8; ++i) u += red[i][threadIdx.x];
            out[2 + pass * kCh + threadIdx.x] = u;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = m;
        out[1] = se;
    }
}

// Heads, second pass (one block per sample): merges the pooling blocks into
// g, f and the softmax constants (max M, normaliser Z), then the MLP.
__global__ __launch_bounds__(256) void k_heads_mlp(const float* __restrict__ part, const float* __restrict__ params,
                                                   const float* __restrict__ desc, float* __restrict__ logits,
                                                   float* __restrict__ usave, float* __restrict__ hsave,
                                                   float* __restrict__ msave)
{
    __shared__ float u[kHeadIn];
    __shared__ float hid[kHidden];
    __shared__ float mz[2];
    const int b = blockIdx.x;
    const float* pp = part + (long long)b * kPoolBlocks * kPoolPart;
    if (threadIdx.x == 0) {
        float M = pp[0];
        for (int k = 1; k < kPoolBlocks; ++k) M = fmaxf(M, pp[k * kPoolPart]);
        float Z = 0.f;
        for (int k = 0; k < kPoolBlocks; ++k) Z += pp[k * kPoolPart + 1] * expf(pp[k * kPoolPart] - M);
        mz[0] = M;
        mz[1] = Z;
        msave[2 * b] = M;
        msave[2 * b + 1] = Z;
    }
    __syncthreads();
    if (threadIdx.x < kCh) {
        const int c = threadIdx.x;
        float gsum = 0.f, fsum = 0.f;
        for (int k = 0; k < kPoolBlocks; ++k) {
            gsum += pp[k * kPoolPart + 2 + kCh + c];
            fsum += pp[k * kPoolPart + 2 + c] * expf(pp[k * kPoolPart] - mz[0]);
        }
        u[c] = gsum * (1.0f / kV2);
        u[kCh + c] = fsum / mz[1];
    }
    else if (threadIdx.x >= 2 * kCh && threadIdx.x < kHeadIn)
        u[threadIdx.x] = desc[b * kDesc + threadIdx.x - 2 * kCh];
    __syncthreads();
    if (threadIdx.x < 2 * kCh) usave[b * 2 * kCh + threadIdx.x] = u[threadIdx.x];
    if (threadIdx.x < kHidden) {
        float t = params[P_BA + threadIdx.x];
        for (int i = 0; i < kHeadIn; ++i) t += params[P_WA + threadIdx.x * kHeadIn + i] * u[i];
        t = t > 0.f ? t : 0.f;
        hid[threadIdx.x] = t;
        hsave[b * kHidden + threadIdx.x] = t;
    }
    __syncthreads();
    if (threadIdx.x < kOrient) {
        float t = params[P_BO + threadIdx.x];
        for (int j = 0; j < kHidden; ++j) t += params[P_WO + threadIdx.x * kHidden + j] * hid[j];
        logits[b * kOrient + threadIdx.x] = t;
    }
}

// Losses and head gradients, one block per sample.  Orientation loss: cross
// entropy of the softmax over the sample's valid orientations against a target
// distribution; region loss: cross entropy of the heat softmax (the one the
// attention pooling uses, constants M and Z from the forward) against the
// region index (none when < 0), weighted by heat_weight.  Both are averaged
// over the batch.  Writes the per-sample head-parameter gradients (layout
// wh .. bo) and dz of layer 4 (BF16), including the gradient through the
// attention weights.
__global__ __launch_bounds__(256) void k_heads_bwd(const u16* __restrict__ a4, const float* __restrict__ params,
                                                   const float* __restrict__ desc, const float* __restrict__ heat,
                                                   const float* __restrict__ logits, const float* __restrict__ usave,
                                                   const float* __restrict__ hsave, const float* __restrict__ msave,
                                                   const float* __restrict__ target, const u8* __restrict__ omask,
                                                   const int* __restrict__ region, int batch, float heat_weight,
                                                   float* __restrict__ losses, float* __restrict__ head_partial,
                                                   u16* __restrict__ dz4)
{
    __shared__ float dheat[kV2];
    __shared__ float pis[kV2];
    __shared__ float dlog[kOrient];
    __shared__ float dpre[kHidden];
    __shared__ float u[kHeadIn];
    __shared__ float dg[kCh], df[kCh];
    __shared__ float sh[8];
    __shared__ float part[8][kCh];
    const int b = blockIdx.x;
    const float inv_b = 1.0f / float(batch);
    const u16* a = a4 + (long long)b * kV2 * kCh;
    float* hp = head_partial + (long long)b * kHeadParams;
    // orientation softmax
    const int o = threadIdx.x;
    const bool valid = o < kOrient && omask[b * kOrient + o];
    const float lo = valid ? logits[b * kOrient + o] : -INFINITY;
    const float mo = block_max(lo, sh);
    const float eo = valid ? expf(lo - mo) : 0.f;
    const float so = block_sum(eo, sh);
    float lossp = 0.f;
    if (o < kOrient) {
        const float p = eo / so;
        const float tg = valid ? target[b * kOrient + o] : 0.f;
        if (valid && tg > 0.f) lossp = -tg * (lo - mo - logf(so));
        dlog[o] = valid ? (p - tg) * inv_b : 0.f;
    }
    const float loss_o = block_sum(lossp, sh);
    // heat softmax from the forward's constants
    const int reg = region[b];
    const float M = msave[2 * b], Z = msave[2 * b + 1];
    for (int v = threadIdx.x; v < kV2; v += 256) pis[v] = expf(heat[b * kV2 + v] - M) / Z;
    if (threadIdx.x == 0) {
        losses[2 * b] = loss_o;
        losses[2 * b + 1] = reg >= 0 ? -(heat[b * kV2 + reg] - M - logf(Z)) : 0.f;
    }
    if (threadIdx.x < 2 * kCh) u[threadIdx.x] = usave[b * 2 * kCh + threadIdx.x];
    else if (threadIdx.x < kHeadIn) u[threadIdx.x] = desc[b * kDesc + threadIdx.x - 2 * kCh];
    __syncthreads();
    // orientation head
    constexpr int oWO = P_WO - P_WH, oBO = P_BO - P_WH, oWA = P_WA - P_WH, oBA = P_BA - P_WH;
    for (int e = threadIdx.x; e < kOrient * kHidden; e += 256)
        hp[oWO + e] = dlog[e / kHidden] * hsave[b * kHidden + e % kHidden];
    if (threadIdx.x < kOrient) hp[oBO + threadIdx.x] = dlog[threadIdx.x];
    if (threadIdx.x < kHidden) {
        const int j = threadIdx.x;
        float d = 0.f;
        for (int q = 0; q < kOrient; ++q) d += params[P_WO + q * kHidden + j] * dlog[q];
        dpre[j] = hsave[b * kHidden + j] > 0.f ? d : 0.f;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kHidden * kHeadIn; e += 256) hp[oWA + e] = dpre[e / kHeadIn] * u[e % kHeadIn];
    if (threadIdx.x < kHidden) hp[oBA + threadIdx.x] = dpre[threadIdx.x];
    if (threadIdx.x < 2 * kCh) {
        float d = 0.f;
        for (int j = 0; j < kHidden; ++j) d += params[P_WA + j * kHeadIn + threadIdx.x] * dpre[j];
        if (threadIdx.x < kCh) dg[threadIdx.x] = d * (1.0f / kV2);
        else df[threadIdx.x - kCh] = d;
    }
    __syncthreads();
    // attention: s_v = a_v . df; dheat_v = hw (pi_v - [v = reg]) + pi_v (s_v - sum_u pi_u s_u)
    float ps = 0.f;
    for (int v = threadIdx.x; v < kV2; v += 256) {
        float sv = 0.f;
        for (int c = 0; c < kCh; ++c) sv += bf2f(a[v * kCh + c]) * df[c];
        dheat[v] = sv;
        ps += pis[v] * sv;
    }
    const float sbar = block_sum(ps, sh);
    const float hw = reg >= 0 ? heat_weight * inv_b : 0.f;
    for (int v = threadIdx.x; v < kV2; v += 256)
        dheat[v] = hw * (pis[v] - (v == reg ? 1.f : 0.f)) + pis[v] * (dheat[v] - sbar);
    __syncthreads();
    // heat head: dwh[c] = sum_v dheat[v] a4[v][c]; dbh = sum_v dheat[v]
    const int c = threadIdx.x & 31, g = threadIdx.x >> 5;
    float s = 0.f;
    for (int v = g; v < kV2; v += 8) s += dheat[v] * bf2f(a[v * kCh + c]);
    part[g][c] = s;
    float sb = 0.f;
    for (int v = threadIdx.x; v < kV2; v += 256) sb += dheat[v];
    sb = block_sum(sb, sh);  // includes a barrier after part[] is written
    if (threadIdx.x < kCh) {
        float t = 0.f;
        for (int i = 0; i < 8; ++i) t += part[i][threadIdx.x];
        hp[threadIdx.x] = t;  // wh
    }
    if (threadIdx.x == 0) hp[P_BH - P_WH] = sb;
    __syncthreads();
    // dz4 = (a4 > 0) * (dheat[v] wh[c] + dg[c] + pi[v] df[c])
    u16* out = dz4 + (long long)b * kV2 * kCh;
    for (int e = threadIdx.x; e < kV2 * kCh; e += 256) {
        const int v = e / kCh, cc = e % kCh;
        const float d = dheat[v] * params[P_WH + cc] + dg[cc] + pis[v] * df[cc];
        out[e] = f2bf(bf2f(a[e]) > 0.f ? d : 0.f);
    }
}

// Sums per-sample parts into the flat gradient: biases from channel sums, the
// descriptor weights, and the head parameters.
__global__ void k_finish_grads(const float* __restrict__ sums1, const float* __restrict__ sums2,
                               const float* __restrict__ sums3, const float* __restrict__ sums4,
                               const float* __restrict__ desc, const float* __restrict__ head_partial, int batch,
                               float* __restrict__ grad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 4 * kCh) {
        const int l = i / kCh, c = i % kCh;
        const float* sums = l == 0 ? sums1 : l == 1 ? sums2 : l == 2 ? sums3 : sums4;
        float s = 0.f;
        for (int b = 0; b < batch; ++b)
            for (int g = 0; g < kSumSegments; ++g) s += sums[(b * kSumSegments + g) * kCh + c];
        const int off = l == 0 ? P_B1 : l == 1 ? P_B2 : l == 2 ? P_B3 : P_B4;
        grad[off + c] = s;
        return;
    }
    int j = i - 4 * kCh;
    if (j < kCh * kDesc) {
        const int co = j / kDesc, d = j % kDesc;
        float s = 0.f;
        for (int b = 0; b < batch; ++b) {
            float t = 0.f;
            for (int g = 0; g < kSumSegments; ++g) t += sums2[(b * kSumSegments + g) * kCh + co];
            s += t * desc[b * kDesc + d];
        }
        grad[P_WF + j] = s;
        return;
    }
    j -= kCh * kDesc;
    if (j < kHeadParams) {
        float s = 0.f;
        for (int b = 0; b < batch; ++b) s += head_partial[(long long)b * kHeadParams + j];
        grad[P_WH + j] = s;
    }
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 copies of the conv weights: W [co][K] (layer 1 padded to 32) and the
// transposed Wt [ci][t][co] for the input gradients of layers 2-4.
__global__ void k_cast_weights(const float* __restrict__ p, u16* __restrict__ w1, u16* __restrict__ w2,
                               u16* __restrict__ w3, u16* __restrict__ w4, u16* __restrict__ t2, u16* __restrict__ t3,
                               u16* __restrict__ t4)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kCh * kK1) {
        const int co = i / kK1, t = i % kK1;
        w1[i] = t < kTaps ? f2bf(p[P_W1 + co * kTaps + t]) : u16(0);
    }
    if (i < kCh * kK) {
        const int co = i / kK, k = i % kK, t = k / kCh, ci = k % kCh;
        const int ti = ci * kK + t * kCh + co;  // transposed position
        w2[i] = f2bf(p[P_W2 + i]);
        w3[i] = f2bf(p[P_W3 + i]);
        w4[i] = f2bf(p[P_W4 + i]);
        t2[ti] = w2[i];
        t3[ti] = w3[i];
        t4[ti] = w4[i];
    }
}

// Proposal selection for one sample: the top_o valid orientations by logit
// (ties: lower index) into the active list, and the top_r regions by heat
// logit (ties: lower index) into the region mask (16 per axis).  Regions are
// ordered by a bitonic sort of 64-bit keys (descending heat, then index).
__device__ __forceinline__ unsigned descending_float_key(float f)
{
    const unsigned u = __float_as_uint(f);
    const unsigned ascending = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return ~ascending;
}

__global__ __launch_bounds__(1024) void k_select(const float* __restrict__ logits, const float* __restrict__ prior,
                                                 const float* __restrict__ heat,
                                                 const OrientDesc* __restrict__ descs, int orientations,
                                                 const int* __restrict__ cfg, int* __restrict__ act,
                                                 u8* __restrict__ region)
{
    __shared__ unsigned long long keys[kV2];
    __shared__ int nvalid;
    const int top_o = cfg[0], top_r = cfg[1];
    const bool all_regions = top_r >= kV2;  // uniform: no region mask, no sort
    if (!all_regions) {
        for (int v = threadIdx.x; v < kV2; v += blockDim.x)
            keys[v] = ((unsigned long long)descending_float_key(heat[v]) << 32) | unsigned(v);
    }
    if (threadIdx.x == 0) nvalid = 0;
    __syncthreads();
    const int o = threadIdx.x;
    const bool valid = o < orientations && o < kOrient && descs[o].lx > 0;
    if (valid) atomicAdd(&nvalid, 1);
    if (valid) {
        // orientation score: the logit plus the scene's prior term
        const float lo = logits[o] + prior[o];
        int rank = 0;
        for (int q = 0; q < orientations && q < kOrient; ++q) {
            if (q == o || descs[q].lx <= 0) continue;
            const float lq = logits[q] + prior[q];
            rank += (lq > lo || (lq == lo && q < o)) ? 1 : 0;
        }
        if (rank < top_o) act[kActHeader + rank] = o;
    }
    if (all_regions) {
        __syncthreads();
        if (threadIdx.x == 0) {
            act[0] = nvalid < top_o ? nvalid : top_o;
            act[1] = 0;
        }
        return;
    }
    // bitonic sort, ascending
    for (int size = 2; size <= kV2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < kV2 / 2; i += blockDim.x) {
                const int lo_i = 2 * i - (i & (stride - 1));
                const int hi_i = lo_i + stride;
                const bool up = (lo_i & size) == 0;
                const unsigned long long a = keys[lo_i], b = keys[hi_i];
                if ((a > b) == up) {
                    keys[lo_i] = b;
                    keys[hi_i] = a;
                }
            }
            __syncthreads();
        }
    }
    for (int r = threadIdx.x; r < kV2; r += blockDim.x) region[unsigned(keys[r] & 0xffffffffu)] = r < top_r ? 1 : 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        act[0] = nvalid < top_o ? nvalid : top_o;
        act[1] = 1;
        act[2] = 16;
        act[3] = 16;
        act[4] = 16;
    }
}

// Lazy exact minima, the exact part of a lazy learned step.  For every
// orientation the proposal keeps the last lowest feasible key found for it
// (key) and whether that key is still its current minimum (valid).
// Occupancy only grows, so an orientation's feasible set only shrinks: a
// minimum that is still free is still the minimum, and one that the new
// occupancy blocks is a lower bound of the new minimum (keys order as the
// plain rule, (z, y, x, orientation)).

// After a full search every orientation's minimum is current.
__global__ void k_lazy_init(const unsigned long long* __restrict__ orient_best, int orientations,
                            unsigned long long* __restrict__ key, unsigned* __restrict__ valid)
{
    const int o = threadIdx.x;
    if (o >= kOrient) return;
    key[o] = o < orientations ? orient_best[o] : ~0ULL;
    valid[o] = 1;
}

// A current key whose kernel now overlaps the occupancy (or leaves the
// window) is no longer the minimum.  Blocks (orientation, slice): each wave
// takes kernel rows (z, y) in turn, its lanes the row's x cells, so no cell
// index is divided; every block that finds an overlap clears the flag.
constexpr int kLazySlices = 32;

__global__ __launch_bounds__(256) void k_lazy_check(const u8* __restrict__ env, int ex, int ey, int ez,
                                                    const u8* __restrict__ kcells,
                                                    const OrientDesc* __restrict__ descs,
                                                    const unsigned long long* __restrict__ key,
                                                    unsigned* __restrict__ valid)
{
    const int o = blockIdx.x;
    const unsigned long long k = key[o];
    if (!valid[o] || k == ~0ULL) return;  // uniform
    const int ux = int((k >> 16) & 0xffffULL), uy = int((k >> 32) & 0xffffULL), uz = int(k >> 48);
    const OrientDesc d = descs[o];
    const int lane = threadIdx.x & 63, waves = blockDim.x >> 6;
    const int rows = d.ky * d.kz;
    unsigned hit = 0;
    for (int r = blockIdx.y * waves + (threadIdx.x >> 6); r < rows; r += waves * gridDim.y) {
        const int z = r / d.ky, y = r - z * d.ky;
        const u8* krow = kcells + d.cells_offset + (long long)r * d.kx;
        const bool inside = uz + z < ez && uy + y < ey;
        const long long erow = (long long)ex * ((uy + y) + (long long)ey * (uz + z)) + ux;
        for (int x = lane; x < d.kx; x += 64)
            if (krow[x] && (!inside || ux + x >= ex || env[erow + x])) hit = 1;
    }
    if (__any(hit) && lane == 0) valid[o] = 0;
}

// The search set of a lazy learned step: the orientations whose key is not
// current and whose lower bound lies below the lowest current key (only
// they can beat it); the top_o of them by score (logit + prior; ties: lower
// index) go into the active list.  No region mask.
__global__ __launch_bounds__(128) void k_lazy_select(const float* __restrict__ logits, const float* __restrict__ prior,
                                                     const OrientDesc* __restrict__ descs, int orientations,
                                                     const int* __restrict__ cfg,
                                                     const unsigned long long* __restrict__ key,
                                                     const unsigned* __restrict__ valid, int* __restrict__ act)
{
    __shared__ unsigned long long best;
    __shared__ int ncand;
    __shared__ float score[kOrient];
    __shared__ unsigned char cand[kOrient];
    const int top_o = cfg[0];
    const int o = threadIdx.x;
    const int no = orientations < kOrient ? orientations : kOrient;
    if (o == 0) {
        best = ~0ULL;
        ncand = 0;
    }
    __syncthreads();
    const bool fits = o < no && descs[o].lx > 0;
    const bool cur = fits && valid[o] != 0;
    const unsigned long long ko = o < no ? key[o] : ~0ULL;
    if (cur) atomicMin(&best, ko);
    __syncthreads();
    const bool is_cand = fits && !cur && ko < best;
    if (o < kOrient) {
        cand[o] = is_cand ? 1 : 0;
        score[o] = o < no ? logits[o] + prior[o] : 0.f;
    }
    if (is_cand) atomicAdd(&ncand, 1);
    __syncthreads();
    if (is_cand) {
        const float so = score[o];
        int rank = 0;
        for (int q = 0; q < no; ++q) {
            if (q == o || !cand[q]) continue;
            const float sq = score[q];
            rank += (sq > so || (sq == so && q < o)) ? 1 : 0;
        }
        if (rank < top_o) act[kActHeader + rank] = o;
    }
    if (o == 0) {
        act[0] = ncand < top_o ? ncand : top_o;
        act[1] = 0;
    }
}

// After the search: the searched orientations' minima become current, and
// the step's choice is the lowest current key (the recheck and the stamp
// read it from the result).
__global__ __launch_bounds__(128) void k_lazy_merge(const unsigned long long* __restrict__ orient_best,
                                                    const int* __restrict__ act, int orientations,
                                                    unsigned long long* __restrict__ key,
                                                    unsigned* __restrict__ valid, StepDeviceResult* __restrict__ r)
{
    __shared__ unsigned long long best;
    const int i = threadIdx.x;
    const int no = orientations < kOrient ? orientations : kOrient;
    if (i == 0) best = ~0ULL;
    if (i < act[0]) {
        const int o = act[kActHeader + i];
        key[o] = orient_best[o];
        valid[o] = 1;
    }
    __syncthreads();
    if (i < no && valid[i]) atomicMin(&best, key[i]);
    __syncthreads();
    if (i == 0) r->key = best;
}

// Lane-layout self-test of the MFMA tile: C = A (32x16) * B (16x32).
__global__ void k_mfma_probe(const u16* __restrict__ a, const u16* __restrict__ b, float* __restrict__ c)
{
    const int lane = threadIdx.x, hi = lane >> 5, c32 = lane & 31;
    u16 ha[8], hb[8];
    for (int j = 0; j < 8; ++j) {
        ha[j] = a[c32 * 16 + hi * 8 + j];
        hb[j] = b[(hi * 8 + j) * 32 + c32];
    }
    f32x16 acc;
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pack8(ha), pack8(hb), acc, 0, 0, 0);
    for (int i = 0; i < 16; ++i) c[acc_row(i, lane) * 32 + c32] = acc[i];
}

}  // namespace

class Proposal {
public:
    Proposal(Context& ctx, int max_batch) : ctx_(ctx), max_batch_(max_batch)
    {
        use_device(ctx_);
        SP_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        params_.allocate(sizeof(float) * P_TOTAL);
        grads_.allocate(sizeof(float) * P_TOTAL);
        m_.allocate(sizeof(float) * P_TOTAL);
        v_.allocate(sizeof(float) * P_TOTAL);
        w1_.allocate(sizeof(u16) * kCh * kK1);
        for (int l = 0; l != 3; ++l) {
            w_[l].allocate(sizeof(u16) * kCh * kK);
            wt_[l].allocate(sizeof(u16) * kCh * kK);
        }
        const std::size_t B = std::size_t(max_batch);
        slots_.allocate(sizeof(unsigned) * B);
        desc_.allocate(sizeof(float) * B * kDesc);
        film_.allocate(sizeof(float) * B * kCh);
        a1_.allocate(sizeof(u16) * B * kV1 * kCh);
        a2_.allocate(sizeof(u16) * B * kV2 * kCh);
        a3_.allocate(sizeof(u16) * B * kV2 * kCh);
        a4_.allocate(sizeof(u16) * B * kV2 * kCh);
        heat_.allocate(sizeof(float) * B * kV2);
        logits_.allocate(sizeof(float) * B * kOrient);
        usave_.allocate(sizeof(float) * B * 2 * kCh);
        hsave_.allocate(sizeof(float) * B * kHidden);
        msave_.allocate(sizeof(float) * B * 2);
        pool_.allocate(sizeof(float) * B * kPoolBlocks * kPoolPart);
        infer_x_.allocate(kV0);
        infer_slot_.allocate(sizeof(unsigned));
        infer_desc_.allocate(sizeof(float) * kDesc);
        cfg_.allocate(sizeof(int) * 2);
        prior_.allocate(sizeof(float) * kOrient);
        lazy_key_.allocate(sizeof(unsigned long long) * kOrient);
        lazy_valid_.allocate(sizeof(unsigned) * kOrient);
        SP_HIP_CHECK(hipMemsetAsync(prior_.get(), 0, prior_.bytes(), stream_));
        SP_HIP_CHECK(hipMemsetAsync(infer_slot_.get(), 0, sizeof(unsigned), stream_));
        SP_HIP_CHECK(hipMemsetAsync(m_.get(), 0, m_.bytes(), stream_));
        SP_HIP_CHECK(hipMemsetAsync(v_.get(), 0, v_.bytes(), stream_));
        SP_HIP_CHECK(hipMemsetAsync(params_.get(), 0, params_.bytes(), stream_));
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
    }
    ~Proposal()
    {
        if (stream_) (void)hipStreamDestroy(stream_);
    }

    void set_parameters(const float* p, std::uint64_t n)
    {
        if (n != std::uint64_t(P_TOTAL)) throw std::invalid_argument("parameter count");
        use_device(ctx_);
        SP_HIP_CHECK(hipMemcpyAsync(params_.get(), p, sizeof(float) * P_TOTAL, hipMemcpyHostToDevice, stream_));
        SP_HIP_CHECK(hipMemsetAsync(m_.get(), 0, m_.bytes(), stream_));
        SP_HIP_CHECK(hipMemsetAsync(v_.get(), 0, v_.bytes(), stream_));
        adam_t_ = 0;
        cast(stream_);
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    // The scene's orientation prior: selection ranks logit[o] + weight * prior[o]
    // (zero beyond count, and everywhere until set).
    void set_prior(const float* prior, std::uint32_t count, float weight)
    {
        if (count > std::uint32_t(kOrient)) throw std::invalid_argument("the prior has at most 96 entries");
        if (!std::isfinite(weight)) throw std::invalid_argument("prior weight");
        float h[kOrient] {};
        for (std::uint32_t o = 0; o != count; ++o) {
            if (!std::isfinite(prior[o])) throw std::invalid_argument("prior values must be finite");
            h[o] = weight * prior[o];
        }
        use_device(ctx_);
        SP_HIP_CHECK(hipMemcpyAsync(prior_.get(), h, sizeof h, hipMemcpyHostToDevice, stream_));
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    void get_parameters(float* p, std::uint64_t n)
    {
        if (n < std::uint64_t(P_TOTAL)) throw std::length_error("capacity");
        use_device(ctx_);
        SP_HIP_CHECK(hipMemcpyAsync(p, params_.get(), sizeof(float) * P_TOTAL, hipMemcpyDeviceToHost, stream_));
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    void reserve(std::uint32_t samples)
    {
        use_device(ctx_);
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
        bank_.allocate(std::size_t(samples) * kV0);
        bank_count_ = samples;
    }

    void load_samples(const std::uint8_t* host, std::uint32_t first, std::uint32_t count)
    {
        if (std::uint64_t(first) + count > bank_count_) throw std::invalid_argument("sample slot");
        use_device(ctx_);
        SP_HIP_CHECK(hipMemcpyAsync(bank_.as<u8>() + std::size_t(first) * kV0, host, std::size_t(count) * kV0,
                                    hipMemcpyHostToDevice, stream_));
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    void read_sample(std::uint32_t slot, std::uint8_t* host)
    {
        if (slot >= bank_count_) throw std::invalid_argument("sample slot");
        use_device(ctx_);
        SP_HIP_CHECK(hipMemcpyAsync(host, bank_.as<u8>() + std::size_t(slot) * kV0, kV0, hipMemcpyDeviceToHost, stream_));
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    // The session's current occupancy, downsampled, into bank slot `slot`.
    void record(sp_hip_session* s, std::uint32_t slot)
    {
        if (slot >= bank_count_) throw std::invalid_argument("sample slot");
        const SessionView v = session_view(s);
        check_extent(v);
        use_device(ctx_);
        hipLaunchKernelGGL(k_downsample, dim3(kV0 / 256), dim3(256), 0, v.stream, v.env, v.ext[0], v.ext[1], v.ext[2],
                           bank_.as<u8>() + std::size_t(slot) * kV0);
        SP_HIP_CHECK(hipGetLastError());
        SP_HIP_CHECK(hipStreamSynchronize(v.stream));
    }

    // Forward (and, with targets, losses + gradients, optionally an Adam step)
    // over bank samples `slots`.
    void run(const std::uint32_t* slots, std::uint32_t batch, const float* desc, float* logits, float* heat,
             const float* target, const std::uint8_t* omask, const std::int32_t* region, float heat_weight, float lr,
             int update, float* losses, float* grads)
    {
        if (batch == 0 || batch > std::uint32_t(max_batch_)) throw std::invalid_argument("batch size");
        for (std::uint32_t i = 0; i != batch; ++i)
            if (slots[i] >= bank_count_) throw std::invalid_argument("sample slot");
        use_device(ctx_);
        SP_HIP_CHECK(hipMemcpyAsync(slots_.get(), slots, sizeof(unsigned) * batch, hipMemcpyHostToDevice, stream_));
        SP_HIP_CHECK(hipMemcpyAsync(desc_.get(), desc, sizeof(float) * kDesc * batch, hipMemcpyHostToDevice, stream_));
        forward(stream_, bank_.as<u8>(), slots_.as<unsigned>(), int(batch), desc_.as<float>());
        if (target) {
            for (std::uint32_t i = 0; i != batch; ++i)
                if (region[i] >= kV2) throw std::invalid_argument("region index");
            ensure_training(int(batch));
            SP_HIP_CHECK(hipMemcpyAsync(target_.get(), target, sizeof(float) * kOrient * batch, hipMemcpyHostToDevice,
                                        stream_));
            SP_HIP_CHECK(hipMemcpyAsync(omask_.get(), omask, kOrient * batch, hipMemcpyHostToDevice, stream_));
            SP_HIP_CHECK(hipMemcpyAsync(region_.get(), region, sizeof(int) * batch, hipMemcpyHostToDevice, stream_));
            backward(stream_, int(batch), heat_weight);
            if (update) {
                ++adam_t_;
                const float b1 = 0.9f, b2 = 0.999f;
                const float c1 = 1.f - std::pow(b1, float(adam_t_)), c2 = 1.f - std::pow(b2, float(adam_t_));
                hipLaunchKernelGGL(k_adam, dim3((P_TOTAL + 255) / 256), dim3(256), 0, stream_, params_.as<float>(),
                                   grads_.as<float>(), m_.as<float>(), v_.as<float>(), P_TOTAL, lr, b1, b2, 1e-8f, c1,
                                   c2);
                cast(stream_);
            }
        }
        SP_HIP_CHECK(hipGetLastError());
        if (logits)
            SP_HIP_CHECK(hipMemcpyAsync(logits, logits_.get(), sizeof(float) * kOrient * batch, hipMemcpyDeviceToHost,
                                        stream_));
        if (heat)
            SP_HIP_CHECK(hipMemcpyAsync(heat, heat_.get(), sizeof(float) * kV2 * batch, hipMemcpyDeviceToHost, stream_));
        if (target && losses) {
            std::vector<float> lp(2 * batch);
            SP_HIP_CHECK(hipMemcpyAsync(lp.data(), losses_.get(), sizeof(float) * 2 * batch, hipMemcpyDeviceToHost,
                                        stream_));
            SP_HIP_CHECK(hipStreamSynchronize(stream_));
            losses[0] = losses[1] = 0.f;
            for (std::uint32_t b = 0; b != batch; ++b) {
                losses[0] += lp[2 * b] / float(batch);
                losses[1] += lp[2 * b + 1] / float(batch);
            }
        }
        if (target && grads)
            SP_HIP_CHECK(hipMemcpyAsync(grads, grads_.get(), sizeof(float) * P_TOTAL, hipMemcpyDeviceToHost, stream_));
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    // One learned placement step on the session: downsample, forward, select
    // the top orientations and regions, search them exactly (one graph).
    // Falls back to the full search when nothing in the proposal is feasible.
    void step(sp_hip_session* s, const float* desc, std::uint32_t top_o, std::uint32_t top_r, bool stamp,
              bool fallback, sp_step_result& out, int& used_fallback)
    {
        const SessionView v = session_view(s);
        check_extent(v);
        if (v.orientations > kOrient) throw std::invalid_argument("the proposal ranks at most 96 orientations");
        if (top_o == 0 || top_o > std::uint32_t(v.orientations)) throw std::invalid_argument("top orientations");
        if (top_r == 0 || top_r > std::uint32_t(kV2)) throw std::invalid_argument("top regions");
        use_device(ctx_);
        if (std::memcmp(desc, last_desc_, sizeof last_desc_) != 0 || !desc_set_) {
            std::memcpy(last_desc_, desc, sizeof last_desc_);
            SP_HIP_CHECK(hipMemcpyAsync(infer_desc_.get(), last_desc_, sizeof last_desc_, hipMemcpyHostToDevice,
                                        v.stream));
            desc_set_ = true;
        }
        if (int(top_o) != cfg_host_[0] || int(top_r) != cfg_host_[1]) {
            cfg_host_[0] = int(top_o);
            cfg_host_[1] = int(top_r);
            SP_HIP_CHECK(hipMemcpyAsync(cfg_.get(), cfg_host_, sizeof cfg_host_, hipMemcpyHostToDevice, v.stream));
        }
        auto propose = [&, v](hipStream_t st) {
            hipLaunchKernelGGL(k_downsample, dim3(kV0 / 256), dim3(256), 0, st, v.env, v.ext[0], v.ext[1], v.ext[2],
                               infer_x_.as<u8>());
            forward(st, infer_x_.as<u8>(), infer_slot_.as<unsigned>(), 1, infer_desc_.as<float>());
            hipLaunchKernelGGL(k_select, dim3(1), dim3(1024), 0, st, logits_.as<float>(), prior_.as<float>(),
                               heat_.as<float>(), v.descs, v.orientations, cfg_.as<int>(), v.act, v.region);
        };
        session_launch_proposed(s, this, propose, int(top_o), stamp);
        session_read_result(s, out);
        used_fallback = 0;
        if (out.orientation < 0 && fallback) {
            session_launch_full(s, stamp);
            session_read_result(s, out);
            used_fallback = 1;
        }
    }

    // One lazy learned step (see k_lazy_*): the orientations whose last
    // minimum the new occupancy blocked and that could beat the lowest
    // current minimum are ranked by the network (plus the prior), the top_o
    // of them are searched exactly, and the choice is the lowest current
    // minimum.  Exact (the plain rule's choice) whenever no more than top_o
    // orientations need a search.  The first step on a session after a
    // reset, and a step that finds nothing, run the full search, which makes
    // every minimum current.
    void lazy_step(sp_hip_session* s, const float* desc, std::uint32_t top_o, bool stamp, sp_step_result& out,
                   int& used_full, std::uint32_t& searched)
    {
        const SessionView v = session_view(s);
        check_extent(v);
        if (v.orientations > kOrient) throw std::invalid_argument("the proposal ranks at most 96 orientations");
        if (top_o == 0 || top_o > std::uint32_t(v.orientations)) throw std::invalid_argument("top orientations");
        use_device(ctx_);
        used_full = 0;
        searched = 0;
        auto full = [&] {
            session_launch_full(s, stamp);
            hipLaunchKernelGGL(k_lazy_init, dim3(1), dim3(128), 0, v.stream, v.orient_best, v.orientations,
                               lazy_key_.as<unsigned long long>(), lazy_valid_.as<unsigned>());
            session_read_result(s, out);
            lazy_serial_ = v.serial;
            lazy_rev_ = v.reset_rev;
            used_full = 1;
            searched = std::uint32_t(v.fitting);
        };
        if (lazy_serial_ != v.serial || lazy_rev_ != v.reset_rev) {
            full();
            return;
        }
        if (std::memcmp(desc, last_desc_, sizeof last_desc_) != 0 || !desc_set_) {
            std::memcpy(last_desc_, desc, sizeof last_desc_);
            SP_HIP_CHECK(hipMemcpyAsync(infer_desc_.get(), last_desc_, sizeof last_desc_, hipMemcpyHostToDevice,
                                        v.stream));
            desc_set_ = true;
        }
        if (int(top_o) != cfg_host_[0]) {
            cfg_host_[0] = int(top_o);
            SP_HIP_CHECK(hipMemcpyAsync(cfg_.get(), cfg_host_, sizeof cfg_host_, hipMemcpyHostToDevice, v.stream));
        }
        auto propose = [&, v](hipStream_t st) {
            hipLaunchKernelGGL(k_lazy_check, dim3(v.orientations, kLazySlices), dim3(256), 0, st, v.env, v.ext[0], v.ext[1],
                               v.ext[2], v.kcells, v.descs, lazy_key_.as<unsigned long long>(),
                               lazy_valid_.as<unsigned>());
            hipLaunchKernelGGL(k_downsample, dim3(kV0 / 256), dim3(256), 0, st, v.env, v.ext[0], v.ext[1], v.ext[2],
                               infer_x_.as<u8>());
            forward(st, infer_x_.as<u8>(), infer_slot_.as<unsigned>(), 1, infer_desc_.as<float>());
            hipLaunchKernelGGL(k_lazy_select, dim3(1), dim3(128), 0, st, logits_.as<float>(), prior_.as<float>(),
                               v.descs, v.orientations, cfg_.as<int>(), lazy_key_.as<unsigned long long>(),
                               lazy_valid_.as<unsigned>(), v.act);
        };
        const std::function<void(hipStream_t)> merge = [&, v](hipStream_t st) {
            hipLaunchKernelGGL(k_lazy_merge, dim3(1), dim3(128), 0, st, v.orient_best, v.act, v.orientations,
                               lazy_key_.as<unsigned long long>(), lazy_valid_.as<unsigned>(), v.result);
        };
        session_launch_proposed(s, this, propose, int(top_o), stamp, &merge);
        int n = 0;
        SP_HIP_CHECK(hipMemcpyAsync(&n, v.act, sizeof n, hipMemcpyDeviceToHost, v.stream));
        session_read_result(s, out);
        searched = std::uint32_t(n);
        // nothing current and feasible: orientations left unsearched may
        // still hold a placement
        if (out.orientation < 0) full();
    }

    void lazy_reset() { lazy_serial_ = 0; }

    void mfma_probe(const std::uint16_t* a, const std::uint16_t* b, float* c)
    {
        use_device(ctx_);
        DeviceBuffer da(32 * 16 * 2), db(16 * 32 * 2), dc(32 * 32 * 4);
        SP_HIP_CHECK(hipMemcpyAsync(da.get(), a, 32 * 16 * 2, hipMemcpyHostToDevice, stream_));
        SP_HIP_CHECK(hipMemcpyAsync(db.get(), b, 16 * 32 * 2, hipMemcpyHostToDevice, stream_));
        hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, stream_, da.as<u16>(), db.as<u16>(), dc.as<float>());
        SP_HIP_CHECK(hipGetLastError());
        SP_HIP_CHECK(hipMemcpyAsync(c, dc.get(), 32 * 32 * 4, hipMemcpyDeviceToHost, stream_));
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
    }

private:
    static void check_extent(const SessionView& v)
    {
        for (int a = 0; a != 3; ++a)
            if (v.ext[a] < kIn) throw std::invalid_argument("the proposal needs an environment of at least 64 cells per axis");
    }

    void cast(hipStream_t st)
    {
        hipLaunchKernelGGL(k_cast_weights, dim3((kCh * kK + 255) / 256), dim3(256), 0, st, params_.as<float>(),
                           w1_.as<u16>(), w_[0].as<u16>(), w_[1].as<u16>(), w_[2].as<u16>(), wt_[0].as<u16>(),
                           wt_[1].as<u16>(), wt_[2].as<u16>());
    }

    void forward(hipStream_t st, const u8* bank, const unsigned* slots, int batch, const float* desc)
    {
        const float* p = params_.as<float>();
        hipLaunchKernelGGL(k_film, dim3((batch * kCh + 255) / 256), dim3(256), 0, st, p, desc, batch, film_.as<float>());
        hipLaunchKernelGGL((k_conv<0, 1, 2, 32>), dim3(kV1 / 128, batch), dim3(256), 0, st, bank, slots, w1_.as<u16>(),
                           p + P_B1, nullptr, nullptr, a1_.as<u16>());
        hipLaunchKernelGGL((k_conv<0, 32, 2, 16>), dim3(kV2 / 32, batch), dim3(256), 0, st, a1_.get(), nullptr,
                           w_[0].as<u16>(), p + P_B2, film_.as<float>(), nullptr, a2_.as<u16>());
        hipLaunchKernelGGL((k_conv<0, 32, 1, 16>), dim3(kV2 / 32, batch), dim3(256), 0, st, a2_.get(), nullptr,
                           w_[1].as<u16>(), p + P_B3, nullptr, nullptr, a3_.as<u16>());
        hipLaunchKernelGGL((k_conv<0, 32, 1, 16>), dim3(kV2 / 32, batch), dim3(256), 0, st, a3_.get(), nullptr,
                           w_[2].as<u16>(), p + P_B4, nullptr, nullptr, a4_.as<u16>());
        hipLaunchKernelGGL(k_heads_pool, dim3(kPoolBlocks, batch), dim3(256), 0, st, a4_.as<u16>(), p,
                           heat_.as<float>(), pool_.as<float>());
        hipLaunchKernelGGL(k_heads_mlp, dim3(batch), dim3(256), 0, st, pool_.as<float>(), p, desc, logits_.as<float>(),
                           usave_.as<float>(), hsave_.as<float>(), msave_.as<float>());
    }

    void ensure_training(int batch)
    {
        if (trained_batch_ >= batch) return;
        const std::size_t B = std::size_t(max_batch_);
        dz4_.allocate(sizeof(u16) * B * kV2 * kCh);
        dz3_.allocate(sizeof(u16) * B * kV2 * kCh);
        dz2_.allocate(sizeof(u16) * B * kV2 * kCh);
        dz1_.allocate(sizeof(u16) * B * kV1 * kCh);
        const std::size_t chunks_l1 = B * kV1 / kRowsPerChunk, chunks = B * kV2 / kRowsPerChunk;
        partial_.allocate(sizeof(float) * std::max(chunks_l1, chunks * kTaps) * kCh * kCh);
        for (auto& s : sums_) s.allocate(sizeof(float) * B * kSumSegments * kCh);
        head_partial_.allocate(sizeof(float) * B * kHeadParams);
        target_.allocate(sizeof(float) * B * kOrient);
        omask_.allocate(B * kOrient);
        region_.allocate(sizeof(int) * B);
        losses_.allocate(sizeof(float) * 2 * B);
        trained_batch_ = max_batch_;
    }

    void backward(hipStream_t st, int batch, float heat_weight)
    {
        const float* p = params_.as<float>();
        float* g = grads_.as<float>();
        hipLaunchKernelGGL(k_heads_bwd, dim3(batch), dim3(256), 0, st, a4_.as<u16>(), p, desc_.as<float>(),
                           heat_.as<float>(), logits_.as<float>(), usave_.as<float>(), hsave_.as<float>(),
                           msave_.as<float>(), target_.as<float>(), omask_.as<u8>(), region_.as<int>(), batch,
                           heat_weight, losses_.as<float>(), head_partial_.as<float>(), dz4_.as<u16>());
        const int chunks2 = batch * kV2 / kRowsPerChunk, chunks1 = batch * kV1 / kRowsPerChunk;
        const int red2 = (kTaps * kCh * kCh + 255) / 256;
        // layer 4: weights (input a3), input gradient gated by a3
        hipLaunchKernelGGL((k_conv_wgrad<32, 1, 16>), dim3(kTaps, chunks2), dim3(256), 0, st, a3_.get(), nullptr,
                           dz4_.as<u16>(), partial_.as<float>());
        hipLaunchKernelGGL(k_reduce_wgrad, dim3(red2), dim3(256), 0, st, partial_.as<float>(), chunks2, kTaps, 0,
                           g + P_W4);
        hipLaunchKernelGGL(k_channel_sums, dim3(batch, kSumSegments), dim3(256), 0, st, dz4_.as<u16>(), kV2,
                           sums_[3].as<float>());
        hipLaunchKernelGGL((k_conv<1, 32, 1, 16>), dim3(kV2 / 32, batch), dim3(256), 0, st, dz4_.get(), nullptr,
                           wt_[2].as<u16>(), nullptr, nullptr, a3_.as<u16>(), dz3_.as<u16>());
        // layer 3
        hipLaunchKernelGGL((k_conv_wgrad<32, 1, 16>), dim3(kTaps, chunks2), dim3(256), 0, st, a2_.get(), nullptr,
                           dz3_.as<u16>(), partial_.as<float>());
        hipLaunchKernelGGL(k_reduce_wgrad, dim3(red2), dim3(256), 0, st, partial_.as<float>(), chunks2, kTaps, 0,
                           g + P_W3);
        hipLaunchKernelGGL(k_channel_sums, dim3(batch, kSumSegments), dim3(256), 0, st, dz3_.as<u16>(), kV2,
                           sums_[2].as<float>());
        hipLaunchKernelGGL((k_conv<1, 32, 1, 16>), dim3(kV2 / 32, batch), dim3(256), 0, st, dz3_.get(), nullptr,
                           wt_[1].as<u16>(), nullptr, nullptr, a2_.as<u16>(), dz2_.as<u16>());
        // layer 2 (stride 2): weights from a1, input gradient at 32^3 gated by a1
        hipLaunchKernelGGL((k_conv_wgrad<32, 2, 16>), dim3(kTaps, chunks2), dim3(256), 0, st, a1_.get(), nullptr,
                           dz2_.as<u16>(), partial_.as<float>());
        hipLaunchKernelGGL(k_reduce_wgrad, dim3(red2), dim3(256), 0, st, partial_.as<float>(), chunks2, kTaps, 0,
                           g + P_W2);
        hipLaunchKernelGGL(k_channel_sums, dim3(batch, kSumSegments), dim3(256), 0, st, dz2_.as<u16>(), kV2,
                           sums_[1].as<float>());
        hipLaunchKernelGGL((k_conv<1, 32, 2, 32>), dim3(kV1 / 32, batch), dim3(256), 0, st, dz2_.get(), nullptr,
                           wt_[0].as<u16>(), nullptr, nullptr, a1_.as<u16>(), dz1_.as<u16>());
        // layer 1: weights only
        hipLaunchKernelGGL((k_conv_wgrad<1, 2, 32>), dim3(1, chunks1), dim3(256), 0, st, bank_.get(), slots_.as<unsigned>(),
                           dz1_.as<u16>(), partial_.as<float>());
        hipLaunchKernelGGL(k_reduce_wgrad, dim3((kCh * kCh + 255) / 256), dim3(256), 0, st, partial_.as<float>(), chunks1,
                           1, 1, g + P_W1);
        hipLaunchKernelGGL(k_channel_sums, dim3(batch, kSumSegments), dim3(256), 0, st, dz1_.as<u16>(), kV1,
                           sums_[0].as<float>());
        const int fin = 4 * kCh + kCh * kDesc + kHeadParams;
        hipLaunchKernelGGL(k_finish_grads, dim3((fin + 255) / 256), dim3(256), 0, st, sums_[0].as<float>(),
                           sums_[1].as<float>(), sums_[2].as<float>(), sums_[3].as<float>(), desc_.as<float>(),
                           head_partial_.as<float>(), batch, g);
    }

    Context& ctx_;
    int max_batch_;
    hipStream_t stream_ {};
    DeviceBuffer params_, grads_, m_, v_, w1_, w_[3], wt_[3];
    DeviceBuffer slots_, desc_, film_, a1_, a2_, a3_, a4_, heat_, logits_, usave_, hsave_, msave_, pool_;
    DeviceBuffer infer_x_, infer_slot_, infer_desc_, cfg_, prior_;
    DeviceBuffer lazy_key_, lazy_valid_;   // per orientation: last minimum, and whether it is current
    // the session lazy_key_ belongs to (serial 0: none) and its feasibility
    // revision then: a step on another session, or after any change other
    // than stamping, searches every orientation again
    std::uint64_t lazy_serial_ {}, lazy_rev_ {};
    DeviceBuffer dz4_, dz3_, dz2_, dz1_, partial_, sums_[4], head_partial_, target_, omask_, region_, losses_;
    DeviceBuffer bank_;
    std::uint32_t bank_count_ {};
    int trained_batch_ {};
    int adam_t_ {};
    float last_desc_[kDesc] {};
    bool desc_set_ {};
    int cfg_host_[2] { -1, -1 };
};

}  // namespace sphip

struct sp_hip_proposal {
    sphip::Proposal* p {};
    sp_hip_ctx* ctx {};
};

namespace {
template <typename F> int proposal_guarded(sp_hip_ctx* ctx, F&& f)
{
    auto fail = [&](int status, const std::string& msg) {
        if (ctx) ctx->c.last_error = msg;
        return status;
    };
    try {
        f();
        return SP_OK;
    }
    catch (const sphip::DeviceError& e) {
        return fail(SP_DEVICE_ERROR, e.what);
    }
    catch (const std::length_error& e) {
        return fail(SP_CAPACITY, e.what());
    }
    catch (const std::bad_alloc&) {
        return fail(SP_CORRELATION_ALLOCATION, "host allocation failed");
    }
    catch (const std::exception& e) {
        return fail(SP_INVALID_ARGUMENT, e.what());
    }
}
}  // namespace

extern "C" {

uint32_t sp_hip_proposal_parameter_count(void) { return uint32_t(sphip::P_TOTAL); }

int sp_hip_proposal_create(sp_hip_ctx* ctx, uint32_t max_batch, sp_hip_proposal** out)
{
    if (!ctx || !out || max_batch == 0 || max_batch > 1024) return SP_INVALID_ARGUMENT;
    return proposal_guarded(ctx, [&] {
        auto* p = new sp_hip_proposal;
        p->ctx = ctx;
        try {
            p->p = new sphip::Proposal(ctx->c, int(max_batch));
        }
        catch (...) {
            delete p;
            throw;
        }
        *out = p;
    });
}

void sp_hip_proposal_destroy(sp_hip_proposal* p)
{
    if (!p) return;
    delete p->p;
    delete p;
}

int sp_hip_proposal_set_parameters(sp_hip_proposal* p, const float* params, uint64_t count)
{
    if (!p || !params) return SP_INVALID_ARGUMENT;
    return proposal_guarded(p->ctx, [&] { p->p->set_parameters(params, count); });
}

int sp_hip_proposal_set_prior(sp_hip_proposal* p, const float* prior, uint32_t count, float weight)
{
    if (!p || (count && !prior)) return SP_INVALID_ARGUMENT;
    return proposal_guarded(p->ctx, [&] { p->p->set_prior(prior, count, weight); });
}

int sp_hip_proposal_get_parameters(sp_hip_proposal* p, float* params, uint64_t count)
{
    if (!p || !params) return SP_INVALID_ARGUMENT;
    return proposal_guarded(p->ctx, [&] { p->p->get_parameters(params, count); });
}

int sp_hip_proposal_reserve(sp_hip_proposal* p, uint32_t samples)
{
    if (!p) return SP_INVALID_ARGUMENT;
    return proposal_guarded(p->ctx, [&] { p->p->reserve(samples); });
}

int sp_hip_proposal_load_samples(sp_hip_proposal* p, const uint8_t* samples, uint32_t first, uint32_t count)
{
    if (!p || (count && !samples)) return SP_INVALID_ARGUMENT;
    return proposal_guarded(p->ctx, [&] { p->p->load_samples(samples, first, count); });
}

int sp_hip_proposal_read_sample(sp_hip_proposal* p, uint32_t slot, uint8_t* sample)
{
    if (!p || !sample) return SP_INVALID_ARGUMENT;
    return proposal_guarded(p->ctx, [&] { p->p->read_sample(slot, sample); });
}

int sp_hip_proposal_record(sp_hip_proposal* p, sp_hip_session* s, uint32_t slot)
{
    if (!p || !s) return SP_INVALID_ARGUMENT;
    return proposal_guarded(p->ctx, [&] { p->p->record(s, slot); });
}

int sp_hip_proposal_forward(sp_hip_proposal* p, const uint32_t* slots, uint32_t batch, const float* descriptors,
                            float* logits, float* heat)
{
    if (!p || !slots || !descriptors) return SP_INVALID_ARGUMENT;
    return proposal_guarded(p->ctx, [&] {
        p->p->run(slots, batch, descriptors, logits, heat, nullptr, nullptr, nullptr, 0.f, 0.f, 0, nullptr, nullptr);
    });
}

int sp_hip_proposal_train(sp_hip_proposal* p, const uint32_t* slots, uint32_t batch, const float* descriptors,
                          const float* orientation_target, const uint8_t* orientation_mask, const int32_t* region,
                          float heat_weight, float learning_rate, int update, float losses[2], float* gradients)
{
    if (!p || !slots || !descriptors || !orientation_target || !orientation_mask || !region) return SP_INVALID_ARGUMENT;
    return proposal_guarded(p->ctx, [&] {
        p->p->run(slots, batch, descriptors, nullptr, nullptr, orientation_target, orientation_mask, region, heat_weight,
                  learning_rate, update, losses, gradients);
    });
}

int sp_hip_proposal_step(sp_hip_proposal* p, sp_hip_session* s, const float descriptor[32], uint32_t top_orientations,
                         uint32_t top_regions, int stamp, int fallback, sp_step_result* out, int* used_fallback)
{
    if (!p || !s || !descriptor || !out) return SP_INVALID_ARGUMENT;
    int fb = 0;
    const int rc = proposal_guarded(p->ctx, [&] {
        p->p->step(s, descriptor, top_orientations, top_regions, stamp != 0, fallback != 0, *out, fb);
    });
    if (used_fallback) *used_fallback = fb;
    if (rc == SP_OK && ((out->numeric_flags | out->run_flags) & 7u)) {
        p->ctx->c.last_error = "step numeric check failed";
        return SP_CORRELATION_NUMERIC;
    }
    return rc;
}

int sp_hip_proposal_lazy_step(sp_hip_proposal* p, sp_hip_session* s, const float descriptor[32],
                              uint32_t top_orientations, int stamp, sp_step_result* out, int* used_full,
                              uint32_t* searched)
{
    if (!p || !s || !descriptor || !out) return SP_INVALID_ARGUMENT;
    int full = 0;
    std::uint32_t n = 0;
    const int rc = proposal_guarded(p->ctx, [&] {
        p->p->lazy_step(s, descriptor, top_orientations, stamp != 0, *out, full, n);
    });
    if (used_full) *used_full = full;
    if (searched) *searched = n;
    if (rc == SP_OK && ((out->numeric_flags | out->run_flags) & 7u)) {
        p->ctx->c.last_error = "step numeric check failed";
        return SP_CORRELATION_NUMERIC;
    }
    return rc;
}

int sp_hip_proposal_lazy_reset(sp_hip_proposal* p)
{
    if (!p) return SP_INVALID_ARGUMENT;
    return proposal_guarded(p->ctx, [&] { p->p->lazy_reset(); });
}

int sp_hip_proposal_mfma_probe(sp_hip_proposal* p, const uint16_t* a, const uint16_t* b, float* c)
{
    if (!p || !a || !b || !c) return SP_INVALID_ARGUMENT;
    return proposal_guarded(p->ctx, [&] { p->p->mfma_probe(a, b, c); });
}

}  // extern "C"
