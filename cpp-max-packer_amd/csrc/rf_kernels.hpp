// Register-FFT versions of the step's passes for power-of-R lengths (256 and
// 512).  Same layouts and semantics as fft_kernels.hpp; W pencils per block,
// P = N/R threads per pencil.
#pragma once

#include "fft_kernels.hpp"
#include "sp_regfft.hpp"

#include <type_traits>

namespace sphip {

template <typename T> __device__ __forceinline__ void load_tw_lds(cx<T>* dst, const cx<T>* tw, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = tw[i];
}

// An incremental update's target.  The C2R pass of a circular update window
// (period nx, ny, N; the new cells at offsets [0, a) per axis) adds each
// count of the new cells' correlation support into the main session's fields
// instead of writing a window field: window index j stands for the linear
// shift j when j < a and j - n otherwise, the support is a + k - 1 shifts
// per axis (k the kernel's extent), and F_o(u + shift) += sign * count for
// main corners inside the legal range, u the placement's corner (`origin`)
// less `shift`.  An orientation whose lowest feasible corner receives a
// count is flagged for a rescan.
struct FieldUpdate {
    int* F;
    const OrientDesc* md;
    const long long* moff;
    const unsigned long long* origin;
    const unsigned long long* best;
    int* rescan;
    int ax, ay, az;
    int shift, sign;
};

// Window index j (period n) as the linear shift it stands for, or -2^30
// (no corner of a grid of at most 2^16 cells per axis plus it is legal) when
// j lies outside the support of a new-cell extent a and a kernel extent k.
__device__ __forceinline__ int window_shift(int j, int n, int a, int k)
{
    const int aa = min(a, n), neg = min(a + k - 1, n) - aa;
    return j < aa ? j : j >= n - neg ? j - n : -(1 << 30);
}

// An update window's index j that lies in the gap between the support's
// two runs (the update's counts there are zero); a = 0: no gap.
__device__ __forceinline__ bool window_gap(int j, int n, int a, int k)
{
    return a > 0 && window_shift(j, n, a, k) == -(1 << 30);
}

// Strided pencils (consecutive pencils contiguous in memory).  Same arguments
// as k_c2c.  Lane mapping: pencil = lane % W (fastest), t = lane / W.
template <typename T, int N, int R, int W>
__global__ void __launch_bounds__(W*(N / R)) rf_cols(const cx<T>* in, long long in_o, long long in_sn, long long in_sw,
                                                      long long in_sb, cx<T>* out, long long out_o, long long out_sn,
                                                      long long out_sw, long long out_sb, int count_w, int n_in,
                                                      int n_out, const OrientDesc* descs, int mode, const int* act,
                                                      const cx<T>* __restrict__ tw, int fwd, int gap_ax = 0,
                                                      int gap_ay = 0)
{
    using T2 = cx<T>;
    using S = RegShape<N, R>;
    constexpr int P = S::P;
    extern __shared__ unsigned char smem_raw[];
    T2* tws = reinterpret_cast<T2*>(smem_raw);
    T2* lds = tws + N;
    const int o = descs ? slot_orientation(act) : int(blockIdx.z);
    if (o < 0) return;
    int cw = count_w, ni = n_in, no = n_out;
    int kx = 0, ky = 0;
    if (descs) {
        const OrientDesc d = descs[o];
        kx = d.kx;
        ky = d.ky;
        if (mode == 1) {
            if (d.lx <= 0) return;
            cw = d.kx;
            ni = d.ky;
        }
        else if (mode == 2) {
            cw = d.lx;
            no = d.ly;
        }
    }
    const int p0 = blockIdx.x * W;
    if (p0 >= cw) return;
    const int w = threadIdx.x % W, t = threadIdx.x / W;
    // mode 2 with gaps (an update window's inverse y pass whose C2R applies
    // the counts): columns and rows outside the update's support are
    // neither read nor written, since that pass reads those columns as zero
    // and skips those rows
    const bool live = p0 + w < cw && !(mode == 2 && window_gap(p0 + w, cw, gap_ax, kx));
    const cx<T>* src = in + o * in_o + blockIdx.y * in_sb + (p0 + w) * in_sw;
    T2* dst = out + o * out_o + blockIdx.y * out_sb + (p0 + w) * out_sw;
    load_tw_lds<T>(tws, tw, N);
    T2 v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int n = t + P * q;
        v[q] = (live && n < ni) ? nt_load(&src[n * in_sn]) : T2 { 0, 0 };
    }
    __syncthreads();
    reg_fft<T, N, R>(v, lds + w * S::stride, tws, t, fwd != 0);
    if (live) {
#pragma unroll
        for (int s = 0; s < R; ++s) {
            const int n = t + P * s;
            if (n < no && !(mode == 2 && window_gap(n, N, gap_ay, ky))) nt_store(&dst[n * out_sn], v[s]);
        }
    }
}

// Contiguous pencils (rows).  Lane mapping: t = lane % P (fastest), pencil = lane / P.
template <typename T, int N, int R, int W>
__global__ void __launch_bounds__(W*(N / R)) rf_rows(cx<T>* data, int rows, long long row_pitch, long long batch_pitch,
                                                      const cx<T>* __restrict__ tw, int fwd)
{
    using T2 = cx<T>;
    using S = RegShape<N, R>;
    constexpr int P = S::P;
    extern __shared__ unsigned char smem_raw[];
    T2* tws = reinterpret_cast<T2*>(smem_raw);
    T2* lds = tws + N;
    const int t = threadIdx.x % P, w = threadIdx.x / P;
    const int r = blockIdx.x * W + w;
    const bool live = r < rows;
    T2* row = data + blockIdx.y * batch_pitch + (long long)r * row_pitch;
    load_tw_lds<T>(tws, tw, N);
    T2 v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = live ? row[t + P * q] : T2 { 0, 0 };
    __syncthreads();
    reg_fft<T, N, R>(v, lds + w * S::stride, tws, t, fwd != 0);
    if (live) {
#pragma unroll
        for (int s = 0; s < R; ++s) row[t + P * s] = v[s];
    }
}

// The product pass, batched over orientations: a block takes a tile of W
// spectral rows (kz, y0..y0+W) of the environment spectrum, keeps them in
// registers, and for each of its orientations finishes the kernel's x
// transform from its partially transformed row H[kz][y][x < kx], multiplies
// by the conjugate, runs the inverse x transform and stores x < lx into
// out[o][kz][y][x].  (The fused path's column-major form is rf_mul_xt.)
// The work is (row tile, orientation) units, tile-major: nh * ceil(ny / W)
// tiles of W rows times the active orientations.  The grid is 1-D and block
// b takes units [b U / G, (b + 1) U / G), loading the environment rows again
// whenever its range enters a new tile; a grid of one block per tile gives
// each block one tile with all orientations.
template <typename T, int N, int R, int W>
__global__ void __launch_bounds__(W*(N / R))
rf_mul_x(const cx<T>* __restrict__ H, long long h_o, int h_px, const cx<T>* __restrict__ SB, int ny, int nh,
         cx<T>* __restrict__ out, long long out_o, const OrientDesc* __restrict__ descs, const int* __restrict__ act,
         const cx<T>* __restrict__ tw, int gap_ax = 0)
{
    using T2 = cx<T>;
    using S = RegShape<N, R>;
    constexpr int P = S::P;
    extern __shared__ unsigned char smem_raw[];
    T2* tws = reinterpret_cast<T2*>(smem_raw);
    T2* lds = tws + N;
    const int t = int(threadIdx.x) % P;
    const int w = int(threadIdx.x) / P;
    const int count = act[0];
    const int ytiles = (ny + W - 1) / W;
    const long long units = (long long)ytiles * nh * count;
    const long long u0 = units * blockIdx.x / gridDim.x, u1 = units * (blockIdx.x + 1) / gridDim.x;
    if (u0 >= u1) return;  // uniform across the block
    load_tw_lds<T>(tws, tw, N);
    __syncthreads();
    T2* pl = lds + w * S::stride;
    for (long long u = u0; u < u1;) {
        // the part of this block's range in row tile `tile`: orientations [i0, i1)
        const int tile = int(u / count);
        const long long base = (long long)tile * count;
        const int i0 = int(u - base), i1 = int(min((long long)count, u1 - base));
        u = base + i1;
        const int kz = tile / ytiles, y = (tile % ytiles) * W + w;
        const bool live = y < ny;
        const long long row = (long long)kz * ny + y;
        T2 sb[R];
#pragma unroll
        for (int s = 0; s < R; ++s) sb[s] = live ? SB[row * N + t + P * s] : T2 { 0, 0 };
        for (int i = i0; i < i1; ++i) {
            const int o = act[kActHeader + i];
            const int kx = descs[o].kx, lx = descs[o].lx;
            if (lx <= 0) continue;  // uniform across the block
            const T2* hrow = H + o * h_o + row * h_px;
            T2 v[R];
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int x = t + P * q;
                v[q] = (live && x < kx) ? nt_load(&hrow[x]) : T2 { 0, 0 };
            }
            // the kernel row has kx <= h_px nonzero values: at radix 16 with up to
            // three leading groups of 16, the first pass leaves the zeros out
            // (uniform over the block).  A pencil is P consecutive lanes of one
            // wave, so its exchanges are wave-local (the table was synchronised
            // before the loop).
            constexpr bool WV = 64 % P == 0;
            const int lead = (kx + P - 1) / P;
            if (R == 16 && lead == 1) reg_fft<T, N, R, 1, WV>(v, pl, tws, t, true);
            else if (R == 16 && lead == 2) reg_fft<T, N, R, 2, WV>(v, pl, tws, t, true);
            else if (R == 16 && lead == 3) reg_fft<T, N, R, 3, WV>(v, pl, tws, t, true);
            else reg_fft<T, N, R, 0, WV>(v, pl, tws, t, true);
#pragma unroll
            for (int s = 0; s < R; ++s) v[s] = cmulc(sb[s], v[s]);
            reg_fft<T, N, R, 0, WV>(v, pl, tws, t, false);
            if (live) {
                T2* orow = out + o * out_o + row * N;
#pragma unroll
                for (int s = 0; s < R; ++s) {
                    const int x = t + P * s;
                    // gap_ax: an update window's columns outside the support
                    // are not needed (see rf_cols)
                    if (x < lx && !window_gap(x, lx, gap_ax, kx)) nt_store(&orow[x], v[s]);
                }
            }
        }
    }
}

// The column-major product pass of the fused path (FP32, N = 256, radix 16).
// Same work as rf_mul_x (kernel x transform, product with conj, inverse x)
// plus the environment's own x transform: the environment spectrum SB comes
// transformed along z and y only, and each row tile transforms its rows
// along x once for all of its orientations (the step then has no separate
// environment x pass).  The output is out[o][x][plane][y] (x stride
// (nh - 1) ny), which rf_yz_argmin reads:
//  * lanes take the pencil fastest (w = lane % W rows of a tile, t = lane / W
//    the 16 threads of a pencil), so that a store covers W consecutive y;
//  * the kernel spectra are kept y-fastest, H[o][kz][x < h_lines][y], so a
//    tile's kernel rows are read as whole 128-byte lines of W consecutive y
//    per x (no line is fetched for the part of a row past its kx values);
//  * the FFT exchange is element-major, lds[W pad(i) + w]: the W pencils of
//    a tile sit side by side, so the lanes of an exchange write or read touch
//    consecutive LDS addresses;
//  * the exchange twiddles are [q][t] tables, one per direction.
// The planes whose rows are Hermitian in y, kz = 0 and kz = n/2, are packed
// into plane 0 as row(0) + i row(n/2) by rf_mul_xt_plane0; this kernel covers
// the planes kz = 1 .. n/2 - 1, as (row tile, orientation) units shared out
// over the grid as rf_mul_x does.
constexpr int kXtPad = 272;  // LDS values per pencil of the element-major exchange (pad(255) + 1 rounded up)

template <typename T> constexpr std::size_t xt_lds_bytes(int w)
{
    return sizeof(cx<T>) * (std::size_t(2 * 256) + std::size_t(w) * kXtPad);
}

// One 256-point transform (radix 16 x 16) of W pencils in the element-major
// exchange; v in = x[t + 16 q], out = X[t + 16 s].  LEAD > 0: x[t + 16 q] is
// zero for q >= LEAD.  xl = the exchange region + this lane's pencil w.
template <typename T, int W, int LEAD>
__device__ __forceinline__ void xfft256(cx<T>* v, cx<T>* xl, const cx<T>* tq, int t, bool fwd)
{
    constexpr int R = 16;
    __builtin_assume(t >= 0 && t < R);
    if constexpr (LEAD > 0) dft16_lead<T, LEAD>(v, fwd);
    else dft16<T>(v, fwd);
    __syncthreads();  // the previous exchange's reads are done
    // pad(16 t + s) = 17 t + s and pad(16 q + t) = 17 q + t for s, t < 16
#pragma unroll
    for (int s = 0; s < R; ++s) xl[W * ((R + 1) * t + s)] = v[s];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < R; ++q) {
        v[q] = xl[W * ((R + 1) * q + t)];
        if (q > 0) v[q] = cmul(v[q], tq[q * R + t]);
    }
    dft16<T>(v, fwd);
}

// Values per lane of a kernel row that rf_mul_xt loads one orientation
// ahead: x = t + 16 q for q < kXtAhead, every value of a row with kx <= 48.
constexpr int kXtAhead = 3;

// The values x = t + 16 q, q >= kXtAhead, of a kernel row (zero past kx;
// none are loaded for kx <= 48).  hcol = H[o][kz][0][y] (x stride ny).
template <typename T>
__device__ __forceinline__ void xt_kernel_rest(cx<T>* v, const cx<T>* __restrict__ hcol, int ny, int kx, bool live,
                                               int t)
{
    using T2 = cx<T>;
    constexpr int R = 16, P = 16;
    if (kx > kXtAhead * P) {  // uniform over the block
#pragma unroll
        for (int q = kXtAhead; q < R; ++q) {
            const int x = t + P * q;
            v[q] = (live && x < kx) ? nt_load(&hcol[(long long)x * ny]) : T2 { 0, 0 };
        }
        // the loads complete within this branch: the waits after it would
        // otherwise count them as possibly in flight on every path
#pragma unroll
        for (int q = kXtAhead; q < R; ++q) asm volatile("" : "+v"(v[q].x), "+v"(v[q].y));
    }
    else {
#pragma unroll
        for (int q = kXtAhead; q < R; ++q) v[q] = T2 { 0, 0 };
    }
}

// The kernel's x transform of one tile row (kx <= 48: the first pass leaves
// the zero groups out); v holds the row, zero past kx.
template <typename T, int W>
__device__ __forceinline__ void xt_kernel_fft(cx<T>* v, int kx, cx<T>* xl, const cx<T>* tqf, int t)
{
    constexpr int P = 16;
    const int lead = (kx + P - 1) / P;  // uniform over the block
    if (lead == 1) xfft256<T, W, 1>(v, xl, tqf, t, true);
    else if (lead == 2) xfft256<T, W, 2>(v, xl, tqf, t, true);
    else if (lead == 3) xfft256<T, W, 3>(v, xl, tqf, t, true);
    else xfft256<T, W, 0>(v, xl, tqf, t, true);
}

template <typename T>
__device__ __forceinline__ void xt_tables(cx<T>* tqf, cx<T>* tqi, const cx<T>* __restrict__ tw)
{
    constexpr int N = 256, R = 16;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const int q = i / R, tt = i % R;
        const cx<T> w = tw[(tt * q) & (N - 1)];
        tqf[i] = w;
        tqi[i] = cx<T> { w.x, -w.y };
    }
}

template <typename T, int W>
__global__ void __launch_bounds__(W * 16)
rf_mul_xt(const cx<T>* __restrict__ H, long long h_o, int h_lines, const cx<T>* __restrict__ SB, int ny, int nh,
          cx<T>* __restrict__ out, long long out_o, const OrientDesc* __restrict__ descs, const int* __restrict__ act,
          const cx<T>* __restrict__ tw, T sb_scale)
{
    using T2 = cx<T>;
    constexpr int N = 256, R = 16, P = 16;
    extern __shared__ unsigned char smem_raw[];
    T2* tqf = reinterpret_cast<T2*>(smem_raw);
    T2* tqi = tqf + N;
    const int t = int(threadIdx.x) / W, w = int(threadIdx.x) % W;
    T2* xl = tqi + N + w;
    const int count = act[0];
    const int ytiles = (ny + W - 1) / W;
    const int planes = nh - 1;  // output planes (plane 0 packed)
    const long long units = (long long)ytiles * (nh - 2) * count;
    const long long u0 = units * blockIdx.x / gridDim.x, u1 = units * (blockIdx.x + 1) / gridDim.x;
    if (u0 >= u1) return;  // uniform across the block
    xt_tables<T>(tqf, tqi, tw);
    __syncthreads();
    for (long long u = u0; u < u1;) {
        // the part of this block's range in row tile `tile`: orientations [i0, i1)
        const int tile = int(u / count);
        const long long base = (long long)tile * count;
        const int i0 = int(u - base), i1 = int(min((long long)count, u1 - base));
        u = base + i1;
        const int kz = 1 + tile / ytiles, y = (tile % ytiles) * W + w;
        const bool live = y < ny;
        const long long row = (long long)kz * ny + y;
        // byte offset of (t, y) within an x plane group of the output (below
        // 2^31: at most 15 (nh - 1) ny + ny values)
        const unsigned int lane_off = unsigned(y + t * planes * ny) * unsigned(sizeof(T2));
        // the tile's environment rows, transformed along x here (once per
        // tile, for every orientation of it)
        T2 sb[R];
#pragma unroll
        for (int s = 0; s < R; ++s) sb[s] = live ? SB[row * N + t + P * s] : T2 { 0, 0 };
        xfft256<T, W, 0>(sb, xl, tqf, t, true);
#pragma unroll
        for (int s = 0; s < R; ++s) sb[s] = T2 { sb[s].x * sb_scale, sb[s].y * sb_scale };
        // The next orientation's descriptor and the first values of its
        // kernel row are loaded while the current one is transformed.  The
        // loads use clamped addresses (values past kx are replaced by zero
        // when taken) and are taken after the inverse transform, before the
        // stores: the wait for them then covers no store of the same
        // iteration, only memory operations an iteration old.
        int on = 0, kxn = 0, lxn = 0;
        T2 ahead[kXtAhead], first[kXtAhead];
        const int yc = live ? y : ny - 1;
        auto fetch = [&](int i) {
            on = act[kActHeader + i];
            kxn = descs[on].kx;
            lxn = descs[on].lx;
            const int last = min(max(kxn, 1), h_lines) - 1;
            const T2* hcol = H + on * h_o + (long long)kz * h_lines * ny + yc;
#pragma unroll
            for (int q = 0; q < kXtAhead; ++q) ahead[q] = nt_load(&hcol[(long long)min(t + P * q, last) * ny]);
        };
        auto take = [&] {
#pragma unroll
            for (int q = 0; q < kXtAhead; ++q) {
                first[q] = (live && t + P * q < kxn) ? ahead[q] : T2 { 0, 0 };
                // taken here, not moved past the stores that follow (the wait
                // for the loads would then include those stores)
                asm volatile("" : "+v"(first[q].x), "+v"(first[q].y));
            }
        };
        fetch(i0);
        take();
        for (int i = i0; i < i1; ++i) {
            const int o = on, kx = kxn, lx = lxn;
            T2 v[R];
#pragma unroll
            for (int q = 0; q < kXtAhead; ++q) v[q] = first[q];
            xt_kernel_rest<T>(v, H + o * h_o + (long long)kz * h_lines * ny + y, ny, kx, live, t);
            if (i + 1 < i1) fetch(i + 1);  // uniform across the block
            if (lx <= 0) {  // uniform across the block
                take();
                continue;
            }
            xt_kernel_fft<T, W>(v, kx, xl, tqf, t);
#pragma unroll
            for (int s = 0; s < R; ++s) v[s] = cmulc(sb[s], v[s]);
            xfft256<T, W, 0>(v, xl, tqi, t, false);
            take();
            if (live) {
                // value x = t + 16 s at out[o][x][kz][y]: a base uniform over
                // the block for each s (x = 16 s) plus this lane's byte offset
                // of (t, y), so that no store address is formed in vector
                // registers
                const long long xs = (long long)planes * ny;
                T2* const base = out + o * out_o + (long long)kz * ny;
#pragma unroll
                for (int s = 0; s < R; ++s) {
                    const int x = t + P * s;
                    char* const bs = reinterpret_cast<char*>(base + (long long)(P * s) * xs);
                    if (x < lx) nt_store(reinterpret_cast<T2*>(bs + lane_off), v[s]);
                }
            }
        }
    }
}

// Plane 0 of rf_mul_xt's output: row(kz 0) + i row(kz n/2) per (row,
// orientation).  The inverse x transform is linear, so the two rows'
// products are combined before it, P0 + i P_half, and a packed row takes
// three transforms.  One block per (tile of W rows, active orientation slot).
template <typename T, int W>
__global__ void __launch_bounds__(W * 16)
rf_mul_xt_plane0(const cx<T>* __restrict__ H, long long h_o, int h_lines, const cx<T>* __restrict__ SB, int ny, int nh,
                 cx<T>* __restrict__ out, long long out_o, const OrientDesc* __restrict__ descs,
                 const int* __restrict__ act, const cx<T>* __restrict__ tw, T sb_scale)
{
    using T2 = cx<T>;
    constexpr int N = 256, R = 16, P = 16;
    extern __shared__ unsigned char smem_raw[];
    T2* tqf = reinterpret_cast<T2*>(smem_raw);
    T2* tqi = tqf + N;
    if (int(blockIdx.y) >= act[0]) return;  // uniform
    const int o = act[kActHeader + blockIdx.y];
    const int kx = descs[o].kx, lx = descs[o].lx;
    if (lx <= 0) return;  // uniform
    const int t = int(threadIdx.x) / W, w = int(threadIdx.x) % W;
    T2* xl = tqi + N + w;
    const int y = int(blockIdx.x) * W + w;
    const bool live = y < ny;
    xt_tables<T>(tqf, tqi, tw);
    __syncthreads();
    T2 p[R];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int kz = h == 0 ? 0 : nh - 1;
        const long long row = (long long)kz * ny + y;
        // the environment row, transformed along x here
        T2 e[R];
#pragma unroll
        for (int s = 0; s < R; ++s) e[s] = live ? SB[row * N + t + P * s] : T2 { 0, 0 };
        xfft256<T, W, 0>(e, xl, tqf, t, true);
        T2 v[R];
        const T2* hcol = H + o * h_o + (long long)kz * h_lines * ny + y;
#pragma unroll
        for (int q = 0; q < kXtAhead; ++q) {
            const int x = t + P * q;
            v[q] = (live && x < kx) ? nt_load(&hcol[(long long)x * ny]) : T2 { 0, 0 };
        }
        xt_kernel_rest<T>(v, hcol, ny, kx, live, t);
        xt_kernel_fft<T, W>(v, kx, xl, tqf, t);
#pragma unroll
        for (int s = 0; s < R; ++s) {
            const T2 q = cmulc(T2 { e[s].x * sb_scale, e[s].y * sb_scale }, v[s]);
            p[s] = h == 0 ? q : T2 { p[s].x - q.y, p[s].y + q.x };
        }
        // the second row's loads are not issued while the first is live
        __builtin_amdgcn_sched_barrier(0);
    }
    xfft256<T, W, 0>(p, xl, tqi, t, false);
    if (live) {
        T2* ocol = out + o * out_o + y;
        const long long xs = (long long)(nh - 1) * ny;
#pragma unroll
        for (int s = 0; s < R; ++s) {
            const int x = t + P * s;
            if (x < lx) nt_store(&ocol[x * xs], p[s]);
        }
    }
}

// R2C along z with column pairs: W real columns (x) per block at row y, two
// columns per complex pencil.  Input uint8 or real; descs selects kernel mode.
template <typename T, int N, int R, int W, typename In>
__global__ void __launch_bounds__((W / 2) * (N / R)) rf_r2c_z(const In* __restrict__ in, int in_px, int in_py, int ex,
                                                               int ey, int ez, cx<T>* __restrict__ out, long long out_bs,
                                                               int out_px, int out_py, int cx_, int cy_,
                                                               const OrientDesc* descs, const int* act,
                                                               const cx<T>* __restrict__ tw)
{
    using T2 = cx<T>;
    using S = RegShape<N, R>;
    constexpr int P = S::P;
    constexpr int PW = W / 2;
    extern __shared__ unsigned char smem_raw[];
    T2* tws = reinterpret_cast<T2*>(smem_raw);
    T2* lds = tws + N;
    const int b = descs ? slot_orientation(act) : int(blockIdx.z);
    if (b < 0) return;
    int ex_ = ex, ey_ = ey, ez_ = ez, px = in_px, py = in_py, cxx = cx_, cyy = cy_;
    const In* src = in;
    if (descs) {
        const OrientDesc d = descs[b];
        if (d.lx <= 0) return;
        ex_ = px = cxx = d.kx;
        ey_ = py = cyy = d.ky;
        ez_ = d.kz;
        src = in + d.cells_offset;
    }
    const int y = blockIdx.y;
    const int x0 = blockIdx.x * W;
    if (y >= cyy || x0 >= cxx) return;
    const int p = threadIdx.x % PW, t = threadIdx.x / PW;
    load_tw_lds<T>(tws, tw, N);
    const int xa = x0 + p, xb = x0 + p + PW;
    T2 v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int z = t + P * q;
        const bool zin = y < ey_ && z < ez_;
        const T a = (zin && xa < ex_) ? T(to_real(src[xa + (long long)px * (y + (long long)py * z)])) : T(0);
        const T bb = (zin && xb < ex_) ? T(to_real(src[xb + (long long)px * (y + (long long)py * z)])) : T(0);
        v[q] = T2 { a, bb };
    }
    __syncthreads();
    T2* pl = lds + p * S::stride;
    reg_fft<T, N, R>(v, pl, tws, t, true);
    // unpack needs X[N-k]: exchange the spectrum through LDS
    __syncthreads();
#pragma unroll
    for (int s = 0; s < R; ++s) pl[S::pad(t + P * s)] = v[s];
    __syncthreads();
    T2* dst = out + b * out_bs;
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const int k = t + P * s;
        if (k > N / 2) continue;
        const T2 xk = v[s];
        const T2 xn = pl[S::pad((N - k) & (N - 1))];
        const T2 A = { T(0.5) * (xk.x + xn.x), T(0.5) * (xk.y - xn.y) };
        const T2 B = { T(0.5) * (xk.y + xn.y), T(0.5) * (xn.x - xk.x) };
        const long long base = (long long)out_px * (y + (long long)out_py * k);
        if (xa < cxx) dst[base + xa] = A;
        if (xb < cxx) dst[base + xb] = B;
    }
}

// C2R along z of column pairs + rounding, the reference's per-value checks,
// the feasible mask and the lowest (z, y, x) key (see k_c2r_z_argmin; OUT and
// the rounding as there).  APPLY: the counts go into another session's
// fields by `fu` (no key, no field output); rows outside the support are
// skipped whole.
template <typename T, int N, int R, int W, bool OUT, bool APPLY = false>
__global__ void __launch_bounds__((W / 2) * (N / R)) rf_c2r_z_argmin(const cx<T>* __restrict__ in, long long in_o,
                                                                      int nx, int ny, const OrientDesc* __restrict__ descs,
                                                                      T inv_total, StepDeviceResult* __restrict__ result,
                                                                      int* __restrict__ fields,
                                                                      const long long* __restrict__ field_offsets,
                                                                      const unsigned long long* __restrict__ excl,
                                                                      double* __restrict__ raw, StepFilter filt,
                                                                      const cx<T>* __restrict__ tw, FieldUpdate fu = {})
{
    using T2 = cx<T>;
    using S = RegShape<N, R>;
    constexpr int P = S::P;
    constexpr int PW = W / 2;
    extern __shared__ unsigned char smem_raw[];
    T2* tws = reinterpret_cast<T2*>(smem_raw);
    T2* lds = tws + N;
    __shared__ unsigned long long best;
    __shared__ unsigned int res_bits, flags;
    const int o = slot_orientation(filt.act);
    if (o < 0) return;
    const OrientDesc d = descs[o];
    const int y = blockIdx.y;
    const int x0 = blockIdx.x * W;
    if (d.lx <= 0 || d.ly <= 0 || d.lz <= 0 || y >= d.ly || x0 >= d.lx) return;
    // APPLY: the main orientation, the target row and the placement's corner
    OrientDesc m {};
    int ty = 0, ox = 0, oz = 0;
    int* fm = nullptr;
    if constexpr (APPLY) {
        m = fu.md[o];
        if (m.lx <= 0) return;
        const unsigned long long key = *fu.origin;
        ox = int((key >> 16) & 0xffffULL) - fu.shift;
        oz = int(key >> 48) - fu.shift;
        ty = int((key >> 32) & 0xffffULL) - fu.shift + window_shift(y, ny, fu.ay, m.ky);
        if (ty < 0 || ty >= m.ly) return;  // outside the support or the legal range: no count lands
        fm = fu.F + fu.moff[o];
    }
    if (threadIdx.x == 0) {
        best = ~0ULL;
        res_bits = 0;
        flags = 0;
    }
    const int p = threadIdx.x % PW, t = threadIdx.x / PW;
    load_tw_lds<T>(tws, tw, N);
    const int xa = x0 + p, xb = x0 + p + PW;
    const cx<T>* src = in + o * in_o + (long long)nx * y + x0;
    const long long plane = (long long)nx * ny;
    // Stage the half spectrum once: rows k in [0, N/2] of the W columns, then
    // every thread builds its Hermitian-extended inputs from LDS.
    constexpr int NH = N / 2 + 1;
    constexpr int SROW = W + 1;
    T2* stage = lds;  // NH * SROW elements, reused by the FFT exchange afterwards
    const int lim = d.lx - x0;
    {
        // all loads of this thread first (memory-level parallelism), then LDS;
        // the block is a whole number of rows, so each thread keeps one column
        // and steps KSTEP rows per load
        constexpr int BLOCK = (W / 2) * (N / R);
        static_assert(BLOCK % W == 0, "block must cover whole rows");
        constexpr int KSTEP = BLOCK / W;
        constexpr int ITER = (NH + KSTEP - 1) / KSTEP;
        const int c = threadIdx.x % W, k0 = threadIdx.x / W;
        // APPLY: columns in the gap between the support's runs read as zero
        const bool col_in = c < lim && !(APPLY && window_gap(x0 + c, nx, fu.ax, m.kx));
        const T2* p = src + k0 * plane + c;
        const long long step = KSTEP * plane;
        T2 tmp[ITER];
#pragma unroll
        for (int j = 0; j < ITER; ++j)
            tmp[j] = (k0 + j * KSTEP < NH && col_in) ? p[j * step] : T2 { 0, 0 };
#pragma unroll
        for (int j = 0; j < ITER; ++j)
            if (k0 + j * KSTEP < NH) stage[(k0 + j * KSTEP) * SROW + c] = tmp[j];
    }
    __syncthreads();
    T2 v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int k = t + P * q;
        const bool mirror = k > N / 2;
        const int kk = mirror ? N - k : k;
        const T2 a = stage[kk * SROW + p];
        const T2 b = stage[kk * SROW + p + PW];
        v[q] = mirror ? T2 { a.x + b.y, b.x - a.y } : T2 { a.x - b.y, a.y + b.x };
    }
    __syncthreads();
    reg_fft<T, N, R>(v, lds + p * S::stride, tws, t, false);
    const T max_out = T(d.mass < (long long)result->env_mass ? d.mass : (long long)result->env_mass);
    unsigned long long mine = ~0ULL;
    float my_res = 0.f;
    unsigned int my_flags = 0;
    int* fo = OUT && fields ? fields + field_offsets[o] : nullptr;
    const bool any_excluded = excl[0] != 0;
    const bool regions_on = filt.act && filt.act[1];
    // APPLY: each value's count and target offset in the orientation's field
    // (its legal box holds fewer than 2^31 corners), applied after the loop
    int cnt[2][R] = {}, toff[2][R] = {};
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const int z = t + P * s;
        if (z >= d.lz) continue;
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int x = half ? xb : xa;
            if (x >= d.lx) continue;
            const T val = (half ? v[s].y : v[s].x) * inv_total;
            const T rv = rint(val);
            const T res = fabs(val - rv);
            my_res = fmaxf(my_res, float(res));
            if (!(res < T(0.25))) my_flags |= 1u;
            if (rv < T(0) || rv > max_out) my_flags |= 2u;
            if (OUT) {
                if (fo) fo[x + (long long)d.lx * (y + (long long)d.ly * z)] = int(rv);
                if (raw) raw[field_offsets[o] + x + (long long)d.lx * (y + (long long)d.ly * z)] = double(val);
            }
            if constexpr (APPLY) {
                const int tz = oz + window_shift(z, N, fu.az, m.kz);
                const int tx = ox + window_shift(x, nx, fu.ax, m.kx);
                const int c = int(rv);
                const bool in = c != 0 && tx >= 0 && tx < m.lx && tz >= 0 && tz < m.lz;
                cnt[half][s] = in ? c : 0;
                toff[half][s] = in ? tx + m.lx * (ty + m.ly * tz) : 0;
                continue;
            }
            if (rv == T(0)) {
                const unsigned long long key = ((unsigned long long)z << 48) | ((unsigned long long)y << 32) |
                                               ((unsigned long long)x << 16) | (unsigned long long)o;
                if ((!any_excluded || !excluded(excl, key)) && (!regions_on || region_allowed(filt, x, y, z)))
                    mine = key < mine ? key : mine;
            }
        }
    }
    if constexpr (APPLY) {
        // the orientation's current minimum as an offset (-1: none)
        const unsigned long long b = fu.best[o];
        const int boff = b == ~0ULL ? -1
                                    : int((b >> 16) & 0xffffULL) +
                                          m.lx * (int((b >> 32) & 0xffffULL) + m.ly * int(b >> 48));
        bool hit = false;
        // a half's sixteen loads are issued together, before any of its
        // stores (each target corner receives from exactly one window value)
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            int old[R];
#pragma unroll
            for (int s = 0; s < R; ++s) old[s] = cnt[half][s] != 0 ? fm[toff[half][s]] : 0;
#pragma unroll
            for (int s = 0; s < R; ++s)
                if (cnt[half][s] != 0) {
                    fm[toff[half][s]] = old[s] + fu.sign * cnt[half][s];
                    hit = hit || toff[half][s] == boff;
                }
        }
        if (hit) fu.rescan[o] = 1;
    }
    // wave-level reduction, then one LDS atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long om = __shfl_down(mine, off, 64);
        mine = om < mine ? om : mine;
        my_res = fmaxf(my_res, __shfl_down(my_res, off, 64));
        my_flags |= __shfl_down(my_flags, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (mine != ~0ULL) atomicMin(&best, mine);
        if (my_res > 0.f) atomicMax(&res_bits, __float_as_uint(my_res));
        if (my_flags) atomicOr(&flags, my_flags);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (best != ~0ULL) {
            atomicMin(&result->key, best);
            if (filt.orient_best) atomicMin(&filt.orient_best[o], best);
        }
        if (res_bits) atomicMax(&result->max_residual_bits, res_bits);
        if (flags) atomicOr(&result->flags, flags);
    }
}


// Fused inverse y + C2R along z + checks + argmin for N = ny = nz = 256 at
// radix 16.  A block of 1024 threads owns one x column of one orientation at
// a time: the column's half spectrum in[o][x][kz][y] (128 x 256 values,
// written by rf_mul_xt with plane 0 holding kz 0 + i kz 128, and
// already scaled by 1/(nx ny nz)) is read once, transformed along y in
// registers, and transformed along z in two chunks of up to 128 legal rows
// through LDS, so the inverse y output never goes back to memory.
//  * y pass: 64 groups of 16 threads, two pencils per group (planes g and
//    g + 64).  Plane 0 is Hermitian in y in both parts, so its transform
//    gives the kz = 0 values as real parts and the kz = 128 values as
//    imaginary parts.  Thread t keeps Y[kz][y = t + 16 s] for s < 16.
//  * z pass: the rows of a chunk are staged as stage[kz][row] (line 0 keeps
//    the packed kz0 + i kz128 values); group g takes rows (2g, 2g + 1) of the
//    chunk as one complex pencil through the Hermitian extension (as
//    rf_c2r_z_argmin does for column pairs), and the real and imaginary
//    outputs are the two rows' values at z = t + 16 s.  A wave whose eight
//    rows of the chunk all lie past the legal range skips them.
// The whole column is in registers between the two passes (64 VGPRs of the
// 128 a thread has at 1024 threads), so the transforms use reg_fft_lean
// (exchange twiddles from a [q][t] table, reads four at a time).
// The grid is one resident block per CU; block b takes the columns
// c = b, b + G, ... of the (x, active slot) list, c = x * count + slot, and
// skips columns outside its orientation's legal x range.  The first pencil
// of the block's next column is loaded while the last chunk of the current
// one is transformed (its registers are free once that chunk is staged), so
// a column's first loads overlap the previous column's work.
// The checks are those of rf_c2r_z_argmin kept as running extremes: the
// largest residual (NaN-propagating, so a NaN raises the residual flag as a
// per-value test does), the smallest and largest rounded count, and a bit
// per zero count; the lowest feasible key of a column is taken from the
// zero bits after its transforms.  Keys, residuals and flags are reduced
// once per block, the per-orientation minima (when asked for) once per
// column.
constexpr int kYzThreads = 1024;
constexpr int kYzRow = 18;                             // LDS values per exchange row (16-byte stores)
constexpr int kYzPencil = 16 * kYzRow;                 // LDS values per exchange pencil
constexpr int kYzStageRow = 132;                       // LDS values per staged kz line (128 rows + pad)

template <typename T> constexpr std::size_t yz_lds_bytes()
{
    constexpr std::size_t fft = std::size_t(64) * kYzPencil;
    constexpr std::size_t stage = std::size_t(128) * kYzStageRow;
    // twiddle table, exchange / staging region, one parked word per thread
    return sizeof(cx<T>) * (256 + (fft > stage ? fft : stage)) + sizeof(unsigned int) * kYzThreads;
}

// Smallest / largest over a wave (every lane gets the result).
__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v)
{
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}
__device__ __forceinline__ float wave_max_f32(float v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

template <typename T>
__global__ void __launch_bounds__(kYzThreads) rf_yz_argmin(const cx<T>* __restrict__ in, long long in_o,
                                                            const OrientDesc* __restrict__ descs,
                                                            StepDeviceResult* __restrict__ result,
                                                            const unsigned long long* __restrict__ excl, StepFilter filt,
                                                            const cx<T>* __restrict__ tw, int slots, int lmax)
{
    using T2 = cx<T>;
    constexpr int N = 256, R = 16, P = 16, NP = N / 2, G = kYzThreads / P;
    using S = RegShape<N, R>;
    static_assert(S::P == P && G == 64 && 2 * G == NP, "256-point pencils of 16 threads, two planes per group");
    static_assert(std::is_same_v<T, float>, "the fused pass runs in FP32 (staged pairs are read as float4)");
    static_assert(64 % P == 0, "a pencil's threads are lanes of one wave");
    constexpr int kRegion = 64 * kYzPencil > 128 * kYzStageRow ? 64 * kYzPencil : 128 * kYzStageRow;
    extern __shared__ __attribute__((aligned(16))) unsigned char yz_smem[];  // 16-byte stores and loads
    T2* tws = reinterpret_cast<T2*>(yz_smem);
    T2* lds = tws + N;
    unsigned int* parked = reinterpret_cast<unsigned int*>(lds + kRegion);
    __shared__ unsigned long long best;
    __shared__ unsigned int res_bits, flags;
    const int count = filt.act ? min(filt.act[0], slots) : slots;
    if (count <= 0 || lmax <= 0) return;  // uniform
    const int total = lmax * count;
    // the first legal column at or after c in this block's sequence, or total
    auto next_col = [&](int c, int& o, int& x) {
        for (; c < total; c += int(gridDim.x)) {
            const int sl = c % count, xx = c / count;
            const int oo = filt.act ? filt.act[kActHeader + sl] : sl;
            const OrientDesc& dd = descs[oo];
            if (xx < dd.lx && dd.ly > 0 && dd.lz > 0) {
                o = oo;
                x = xx;
                return c;
            }
        }
        return total;
    };
    int o = 0, x = 0;
    int c = next_col(int(blockIdx.x), o, x);
    if (c >= total) return;  // uniform
    if (threadIdx.x == 0) {
        best = ~0ULL;
        res_bits = 0;
        flags = 0;
    }
    if (threadIdx.x < N) {
        // inverse exchange twiddles as [q][t]: conj(w^(t q))
        const int q = threadIdx.x / R, tt = threadIdx.x % R;
        const T2 w = tw[(tt * q) & (N - 1)];
        tws[threadIdx.x] = T2 { w.x, -w.y };
    }
    // lane indices, formed afresh where they are used (values derived from
    // them, such as LDS and load addresses, are then not hoisted out of the
    // column loop and held across it)
    auto lane_t = [] {
        int v = int(threadIdx.x) % P;
        asm volatile("" : "+v"(v));
        __builtin_assume(v >= 0 && v < P);
        return v;
    };
    auto lane_g = [] {
        int v = int(threadIdx.x) / P;
        asm volatile("" : "+v"(v));
        __builtin_assume(v >= 0 && v < G);
        return v;
    };
    constexpr long long kColumn = (long long)NP * N;  // values per x column
    T2 a[R], b[R];
    {
        const T2* p0 = in + o * in_o + x * kColumn + (long long)lane_g() * N + lane_t();
#pragma unroll
        for (int q = 0; q < R; ++q) a[q] = nt_load(&p0[P * q]);
    }
    __syncthreads();
    const bool any_excluded = excl[0] != 0;
    const bool regions_on = filt.act && filt.act[1];
    // per lane: largest residual and smallest rounded count over every
    // column; per wave (uniform): the lowest feasible key and the range flag
    T r1 = T(0), lo = T(0);
    unsigned long long wave_key = ~0ULL;
    unsigned int wave_range = 0;
    T2* stage = lds;
    while (c < total) {
        const OrientDesc d = descs[o];
        {
            const int t = lane_t(), g = lane_g();
            T2* pen = lds + g * kYzPencil;
            const T2* p1 = in + o * in_o + x * kColumn + (long long)(g + G) * N + t;
            // y pass.  Its exchanges touch only each group's own region.  The
            // second pencil's loads are issued inside the first transform,
            // after its exchange writes, when the first pencil's registers are
            // free.
            auto load_b = [&] {
#pragma unroll
                for (int q = 0; q < R; ++q) b[q] = nt_load(&p1[P * q]);
            };
            reg_fft_lean<T, N, R, false, decltype(load_b), kYzRow, false>(a, pen, tws, t, false, load_b);
            // (the two transforms are not interleaved by the scheduler: their
            // working sets together do not fit the registers)
            __builtin_amdgcn_sched_barrier(0);
            reg_fft_lean<T, N, R, false, NoWork, kYzRow, false>(b, pen, tws, t, false);
        }
        int o2 = o, x2 = x;
        const int c2 = next_col(c + int(gridDim.x), o2, x2);
        const T max_out = T(d.mass < (long long)result->env_mass ? d.mass : (long long)result->env_mass);
        T col_hi = T(0);
        unsigned int zero_last = 0;
        // one chunk of rows [128 C, 128 C + 128); PF: the next column's first
        // pencil is loaded once this chunk is staged (the last chunk); PARK:
        // this chunk's zero bits are parked in LDS until the column ends
        auto chunk = [&](auto cc, auto pf, auto park) {
            constexpr int C = decltype(cc)::value;
            constexpr bool PF = decltype(pf)::value, PARK = decltype(park)::value;
            const int row0 = 128 * C;
            const int t = lane_t(), g = lane_g();
            T2* pen = lds + g * kYzPencil;
            __syncthreads();
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const int r = t + P * s;  // row within the chunk: y = row0 + r
                stage[g * kYzStageRow + r] = a[8 * C + s];
                stage[(g + G) * kYzStageRow + r] = b[8 * C + s];
            }
            if (PF) {
                // (after the block's last column the current one is read
                // again, unused: the loads stay unconditional)
                const bool more = c2 < total;
                const T2* pn = in + (more ? o2 : o) * in_o + (more ? x2 : x) * kColumn + (long long)g * N + t;
#pragma unroll
                for (int q = 0; q < R; ++q) a[q] = nt_load(&pn[P * q]);
            }
            __syncthreads();
            if (row0 >= d.ly) {  // uniform: a column of at most 128 legal rows
                zero_last = 0;
                return;
            }
            const int y1 = row0 + 2 * g;
            // wave w holds rows row0 + 8 w ... row0 + 8 w + 7; a wave none of
            // whose rows is legal (y >= ly) transforms and checks nothing and
            // takes only the transform's block barrier (uniform per wave)
            const bool live = row0 + 8 * __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6) < d.ly;
            T2 v[R];
            if (live) {
                // line 0 holds (kz0 + i kz128) of both rows: k = 0 takes the
                // real parts and k = 128 the imaginary parts (thread 0, q = 0
                // and 8)
#pragma unroll
                for (int h = 0; h < 4; ++h) {
#pragma unroll
                    for (int q = 4 * h; q < 4 * h + 4; ++q) {
                        const int k = t + P * q;
                        const bool mirror = k > N / 2;
                        const int kk = (k == N / 2) ? 0 : mirror ? N - k : k;
                        const float4 ab = *reinterpret_cast<const float4*>(&stage[kk * kYzStageRow + 2 * g]);
                        const T2 u { ab.x, ab.y }, w { ab.z, ab.w };
                        T2 e = mirror ? T2 { u.x + w.y, w.x - u.y } : T2 { u.x - w.y, u.y + w.x };
                        if (q == 0 && t == 0) e = T2 { u.x, w.x };
                        if (q == 8 && t == 0) e = T2 { u.y, w.y };
                        v[q] = e;
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
                dft_reg<T, R, false>(v, false);
            }
            // the exchange regions alias the staged lines other waves read
            __syncthreads();
            unsigned int zeros = 0;
            if (live) {
                reg_fft_lean_second<T, N, R, NoWork, kYzRow, false>(v, pen, tws, t, false);
                // checks over every value of the wave's rows, kept as running
                // extremes.  Every circular value, legal or not, is in exact
                // arithmetic an integer in [0, min(masses)] and lies under the
                // same a-priori bound, so the checks need no legality test per
                // value (the legal values alone are the reference's; checking
                // more of them is stricter).
                auto check = [&](int s, T& ra, T& rb) {
                    const T va = v[s].x, vb = v[s].y;
                    ra = rint(va);
                    rb = rint(vb);
                    r1 = __builtin_elementwise_maximum(r1,
                                                       __builtin_elementwise_maximum(fabsf(va - ra), fabsf(vb - rb)));
                    lo = fminf(lo, fminf(ra, rb));
                    col_hi = fmaxf(col_hi, fmaxf(ra, rb));
                };
                // Zero counts are marked in bit 2 s + half, which orders them
                // as their keys (z, then y); the bits of values outside the
                // legal range are cleared with the lane's legal-bit mask.
#pragma unroll
                for (int s = 0; s < R; ++s) {
                    T ra, rb;
                    check(s, ra, rb);
                    zeros |= (ra == T(0) ? 1u << (2 * s) : 0u) | (rb == T(0) ? 2u << (2 * s) : 0u);
                }
                // legal bits: 2 s + half with t + 16 s < lz and y1 + half < ly
                const int ns = min(max((d.lz - t + P - 1) / P, 0), R);
                const unsigned int zmask = ns >= R ? ~0u : (1u << (2 * ns)) - 1u;
                const unsigned int ymask = (y1 < d.ly ? 0x55555555u : 0u) | (y1 + 1 < d.ly ? 0xAAAAAAAAu : 0u);
                zeros &= zmask & ymask;
            }
            if (PARK) parked[threadIdx.x] = zeros;
            else zero_last = zeros;
        };
        using C0 = std::integral_constant<int, 0>;
        using C1 = std::integral_constant<int, 1>;
        chunk(C0 {}, std::false_type {}, std::true_type {});
        chunk(C1 {}, std::true_type {}, std::false_type {});
        // the range check's upper end is per orientation
        if (wave_max_f32(col_hi) > max_out) wave_range = 2u;
        // this thread's lowest feasible key in the column: per chunk the
        // first marked value (bit order is key order within a chunk) that is
        // neither rejected nor outside the allowed regions, and the lower of
        // the two chunks' keys (z comes first in a key, so either chunk can
        // hold the lower one)
        unsigned long long col_key = ~0ULL;
        {
            const int t = lane_t(), g = lane_g();
#pragma unroll
            for (int C = 0; C < 2; ++C) {
                unsigned int zeros = C == 0 ? parked[threadIdx.x] : zero_last;
                const int y1 = 128 * C + 2 * g;
                while (zeros) {
                    const int i = __builtin_ctz(zeros);
                    const int z = t + P * (i >> 1), y = y1 + (i & 1);
                    const unsigned long long key = ((unsigned long long)z << 48) | ((unsigned long long)y << 32) |
                                                   ((unsigned long long)x << 16) | (unsigned long long)o;
                    bool ok = !any_excluded || !excluded(excl, key);
                    if (ok && regions_on) {
                        // (a copy of the filter per use: the region divisions
                        // are then not prepared outside the column loop)
                        StepFilter f = filt;
                        asm volatile("" : "+s"(f.ex), "+s"(f.ey), "+s"(f.ez));
                        ok = region_allowed(f, x, y, z);
                    }
                    if (ok) {
                        col_key = key < col_key ? key : col_key;
                        break;
                    }
                    zeros &= zeros - 1;
                }
            }
        }
        if (__builtin_amdgcn_ballot_w64(col_key != ~0ULL)) {
            // the wave's minimum, held as a scalar
            col_key = wave_min_u64(col_key);
            col_key = ((unsigned long long)__builtin_amdgcn_readfirstlane(unsigned(col_key >> 32)) << 32) |
                      (unsigned long long)__builtin_amdgcn_readfirstlane(unsigned(col_key));
            wave_key = col_key < wave_key ? col_key : wave_key;
            if (filt.orient_best && (threadIdx.x & 63) == 0) atomicMin(&filt.orient_best[o], col_key);
        }
        c = c2;
        o = o2;
        x = x2;
    }
    unsigned int my_flags = (!(r1 < T(0.25)) ? 1u : 0u) | (lo < T(0) ? 2u : 0u) | wave_range;
    float my_res = float(r1);
    for (int off = 32; off > 0; off >>= 1) {
        my_res = __builtin_elementwise_maximum(my_res, __shfl_down(my_res, off, 64));
        my_flags |= __shfl_down(my_flags, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (wave_key != ~0ULL) atomicMin(&best, wave_key);
        if (!(my_res <= 0.f)) atomicMax(&res_bits, __float_as_uint(my_res));
        if (my_flags) atomicOr(&flags, my_flags);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (best != ~0ULL) atomicMin(&result->key, best);
        if (res_bits) atomicMax(&result->max_residual_bits, res_bits);
        if (flags) atomicOr(&result->flags, flags);
    }
}

}  // namespace sphip
