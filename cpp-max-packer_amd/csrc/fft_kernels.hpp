// Global kernels of the spectral correlation on gfx950.
//
// Layout: a real grid is X-fastest, idx = x + nx*(y + ny*z).  Spectra keep the
// half (Hermitian) axis on z, the slowest axis: S[kz][y][x] with kz in
// [0, nz/2], so every spectral row along x is a full aligned row.  Forward =
// R2C along z, then C2C along y, then C2C along x; the inverse runs x, y, then
// C2R along z.  Two real z-columns share one complex FFT (pairing) in both
// directions.
#pragma once

#include "sp_fft.hpp"

namespace sphip {

// Per-orientation geometry for the batched step (device array).
struct OrientDesc {
    int kx, ky, kz;           // kernel window shape
    int lx, ly, lz;           // legal translations per axis (E - K + 1; <= 0: does not fit)
    long long cells_offset;   // into the concatenated kernel cells
    long long mass;           // kernel ones
};

// Result of a step (device).  key packs (uz, uy, ux, orientation) in 16 bits each.
struct StepDeviceResult {
    unsigned long long key;
    unsigned int flags;       // bit0 residual >= 0.25, bit1 range, bit2 the chosen corner's direct recheck
    unsigned int max_residual_bits;
    unsigned long long env_mass;
    unsigned int run_flags;   // union of the flags of earlier steps since the host last read the result
    unsigned int pad;
};

// Active orientations of a step: act[0] = count, act[1] = regions on,
// act[2..4] = regions per axis, act[5..] = orientation per grid z slot.
// Null: every slot is its own orientation.
constexpr int kActHeader = 5;
__device__ __forceinline__ int slot_orientation(const int* act)
{
    if (!act) return int(blockIdx.z);
    if (int(blockIdx.z) >= act[0]) return -1;
    return act[kActHeader + blockIdx.z];
}

// What the argmin pass keeps: the active list, an optional mask over
// regions of the environment (a kernel corner u lies in region
// floor(u * regions / extent) per axis) and optional per-orientation minima.
struct StepFilter {
    const int* act;
    const unsigned char* region_mask;
    int ex, ey, ez;
    unsigned long long* orient_best;
};

__device__ __forceinline__ bool region_allowed(const StepFilter& f, int ux, int uy, int uz)
{
    if (!f.act || !f.act[1]) return true;
    const int* a = f.act;
    const int rx = ux * a[2] / f.ex, ry = uy * a[3] / f.ey, rz = uz * a[4] / f.ez;
    return f.region_mask[rx + a[2] * (ry + a[3] * rz)] != 0;
}

// Rejected translations (gap_fill.cpp:517-520): excl[0] = count, then sorted
// packed keys.  A rejected key is never chosen again.
__device__ __forceinline__ bool excluded(const unsigned long long* excl, unsigned long long key)
{
    const unsigned long long n = excl[0];
    if (n == 0) return false;
    unsigned long long lo = 0, hi = n;
    while (lo < hi) {
        const unsigned long long mid = (lo + hi) / 2;
        if (excl[1 + mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && excl[1 + lo] == key;
}

template <typename T> __device__ __forceinline__ void load_twiddles(cx<T>* lds_tw, const cx<T>* tw, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) lds_tw[i] = tw[i];
}


// Builds P full-length complex pencils X_p[k] = A_k + i B_k (and the Hermitian
// mirror X_p[n-k] = conj(A_k) + i conj(B_k)) from the half spectra of the two
// real columns p and p + P at src[k * plane + column], k in [0, nh).  Every
// slot k in [0, n) is written exactly once.  valid = columns present.
template <typename T>
__device__ __forceinline__ void hermitian_pairs(cx<T>* X, int sw, const cx<T>* src, long long plane, int n, int nh,
                                                int P, int valid)
{
    using T2 = cx<T>;
    for (int i = threadIdx.x; i < nh * P; i += blockDim.x) {
        const int k = i / P, p = i % P;
        const T2 a = p < valid ? src[k * plane + p] : T2 { 0, 0 };
        const T2 b = p + P < valid ? src[k * plane + p + P] : T2 { 0, 0 };
        X[k * sw + p] = T2 { a.x - b.y, a.y + b.x };
        const int km = (n - k) % n;
        if (km != k) X[km * sw + p] = T2 { a.x + b.y, b.x - a.y };
    }
}

template <typename In> __device__ __forceinline__ double to_real(In v) { return double(v); }

// ---------------------------------------------------------------------------
// R2C along z.  Grid: x = column tiles (W columns of x), y = row y, z = batch.
// in: real grid of shape (ex, ey, ez) with pitches (in_px, in_py) per batch at
// in + batch*in_bs; columns x < cx, y < cy are transformed (zeros beyond the
// input extents), written to out[kz][y][x] with pitches (out_px, out_py), for
// kz in [0, nz/2].  Per-batch extents come from descs when non-null (kernel
// cells of orientation b).
template <typename T, typename In>
__global__ void __launch_bounds__(256) k_r2c_z(const In* __restrict__ in, long long in_bs, int in_px, int in_py,
                                               int ex, int ey, int ez, cx<T>* __restrict__ out, long long out_bs,
                                               int out_px, int out_py, int cx_, int cy_, const OrientDesc* descs,
                                               const int* act, const cx<T>* __restrict__ tw, FftPlan plan, int W)
{
    using T2 = cx<T>;
    extern __shared__ unsigned char smem_raw[];
    const int nz = plan.n;
    const int P = W / 2;        // complex pencils (pairs)
    const int sw = P + 1;
    T2* tws = reinterpret_cast<T2*>(smem_raw);
    T2* A = tws + nz;
    T2* B = A + nz * sw;
    const int b = descs ? slot_orientation(act) : int(blockIdx.z);
    if (b < 0) return;
    int ex_ = ex, ey_ = ey, ez_ = ez, px = in_px, py = in_py, cxx = cx_, cyy = cy_;
    const In* src = in + b * in_bs;
    if (descs) {
        const OrientDesc d = descs[b];
        if (d.lx <= 0) return;  // kernel does not fit the window: no spectrum needed
        ex_ = px = cxx = d.kx;
        ey_ = py = cyy = d.ky;
        ez_ = d.kz;
        src = in + d.cells_offset;
    }
    const int y = blockIdx.y;
    const int x0 = blockIdx.x * W;
    if (y >= cyy || x0 >= cxx) return;  // uniform per block
    load_twiddles<T>(tws, tw, nz);
    for (int i = threadIdx.x; i < nz * W; i += blockDim.x) {
        const int z = i / W, w = i % W;
        const int x = x0 + w;
        T v = 0;
        if (x < ex_ && y < ey_ && z < ez_) v = T(to_real(src[x + (long long)px * (y + (long long)py * z)]));
        T2* slot = &A[z * sw + (w % P)];
        if (w < P) slot->x = v;
        else slot->y = v;
    }
    __syncthreads();
    T2* R = lds_fft<T>(A, B, tws, plan, P, sw, true);
    const int nh = nz / 2 + 1;
    T2* dst = out + b * out_bs;
    for (int i = threadIdx.x; i < nh * W; i += blockDim.x) {
        const int k = i / W, w = i % W;
        const int p = w % P;
        const int x = x0 + w;
        if (x >= cxx) continue;
        const T2 xk = R[k * sw + p];
        const T2 xn = R[((nz - k) % nz) * sw + p];
        T2 v;
        if (w < P) v = { T(0.5) * (xk.x + xn.x), T(0.5) * (xk.y - xn.y) };   // (Xk + conj Xn)/2
        else v = { T(0.5) * (xk.y + xn.y), T(0.5) * (xn.x - xk.x) };        // (Xk - conj Xn)/(2i)
        dst[x + (long long)out_px * (y + (long long)out_py * k)] = v;
    }
}

// ---------------------------------------------------------------------------
// Generic strided C2C.  Pencil p (of count_w, stride s_w) in batch b (stride
// s_b), point n at stride s_n, separate in/out strides; only the first n_in
// input points are read (zeros after) and the first n_out outputs written.
// Grid: x = pencil tiles of W, y = batch, z = orientation (offset o_in/o_out).
// When descs is set, mode selects per-orientation extents:
//   mode 1: count_w = kx, n_in = ky (kernel y-pass);  mode 2: count_w = lx, n_out = ly (inverse y-pass).
template <typename T>
__global__ void __launch_bounds__(256) k_c2c(const cx<T>* in, long long in_o, long long in_sn, long long in_sw,
                                             long long in_sb, cx<T>* out, long long out_o, long long out_sn,
                                             long long out_sw, long long out_sb, int count_w, int n_in, int n_out,
                                             const OrientDesc* descs, int mode, const int* act,
                                             const cx<T>* __restrict__ tw, FftPlan plan, int W, int fwd)
{
    using T2 = cx<T>;
    extern __shared__ unsigned char smem_raw[];
    const int n = plan.n;
    const int sw = W + 1;
    T2* tws = reinterpret_cast<T2*>(smem_raw);
    T2* A = tws + n;
    T2* B = A + n * sw;
    const int o = descs ? slot_orientation(act) : int(blockIdx.z);
    if (o < 0) return;
    int cw = count_w, ni = n_in, no = n_out;
    if (descs) {
        const OrientDesc d = descs[o];
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
    const int wv = min(W, cw - p0);
    const cx<T>* src = in + o * in_o + blockIdx.y * in_sb + p0 * in_sw;
    T2* dst = out + o * out_o + blockIdx.y * out_sb + p0 * out_sw;
    load_twiddles<T>(tws, tw, n);
    if (in_sn == 1) {
        for (int i = threadIdx.x; i < n * W; i += blockDim.x) {
            const int w = i / n, k = i % n;
            A[k * sw + w] = (w < wv && k < ni) ? src[k + w * in_sw] : T2 { 0, 0 };
        }
    }
    else {
        for (int i = threadIdx.x; i < n * W; i += blockDim.x) {
            const int k = i / W, w = i % W;
            A[k * sw + w] = (w < wv && k < ni) ? src[k * in_sn + w * in_sw] : T2 { 0, 0 };
        }
    }
    __syncthreads();
    T2* R = lds_fft<T>(A, B, tws, plan, W, sw, fwd != 0);
    if (out_sn == 1) {
        for (int i = threadIdx.x; i < n * W; i += blockDim.x) {
            const int w = i / n, k = i % n;
            if (w < wv && k < no) dst[k + w * out_sw] = R[k * sw + w];
        }
    }
    else {
        for (int i = threadIdx.x; i < n * W; i += blockDim.x) {
            const int k = i / W, w = i % W;
            if (w < wv && k < no) dst[k * out_sn + w * out_sw] = R[k * sw + w];
        }
    }
}

// ---------------------------------------------------------------------------
// Row pass with the product: for each spectral row (kz, y) take the kernel's
// partially transformed row H[kz][y][x < kx] (zeros beyond), finish its x
// transform, form G = S_B * conj(S_K) and run the inverse x transform; store
// x < lx into out[kz][y][x].  With k_spectral set, H rows are already full
// x spectra (no forward pass).  Grid: x = row tiles of W rows (y), y = kz,
// z = orientation.
template <typename T>
__global__ void __launch_bounds__(256) k_mul_x(const cx<T>* __restrict__ H, long long h_o, int h_px, int h_py,
                                               const cx<T>* __restrict__ SB, int nx, int ny, cx<T>* __restrict__ out,
                                               long long out_o, const OrientDesc* descs, int k_spectral,
                                               int lx_all, const int* act, const cx<T>* __restrict__ tw, FftPlan plan,
                                               int W)
{
    using T2 = cx<T>;
    extern __shared__ unsigned char smem_raw[];
    const int n = plan.n;  // == nx
    const int sw = W + 1;
    T2* tws = reinterpret_cast<T2*>(smem_raw);
    T2* K0 = tws + n;
    T2* K1 = K0 + n * sw;
    T2* B0 = K1 + n * sw;
    T2* B1 = B0 + n * sw;
    const int o = descs ? slot_orientation(act) : int(blockIdx.z);
    if (o < 0) return;
    const int kz = blockIdx.y;
    const int y0 = blockIdx.x * W;
    int kx = nx, lx = lx_all;
    if (descs) {
        kx = k_spectral ? nx : descs[o].kx;
        lx = descs[o].lx;
    }
    if (lx <= 0) return;
    const int wv = min(W, ny - y0);
    load_twiddles<T>(tws, tw, n);
    const T2* hrow = H + o * h_o + (long long)h_px * (y0 + (long long)h_py * kz);
    const T2* brow = SB + (long long)nx * (y0 + (long long)ny * kz);
    for (int i = threadIdx.x; i < n * W; i += blockDim.x) {
        const int w = i / n, x = i % n;
        const bool ok = w < wv;
        K0[x * sw + w] = (ok && x < kx) ? hrow[x + (long long)h_px * w] : T2 { 0, 0 };
        B0[x * sw + w] = ok ? brow[x + (long long)nx * w] : T2 { 0, 0 };
    }
    __syncthreads();
    const T2* KS = K0;
    if (!k_spectral) KS = lds_fft<T>(K0, K1, tws, plan, W, sw, true);
    for (int i = threadIdx.x; i < n * W; i += blockDim.x) {
        const int x = i / W, w = i % W;
        B0[x * sw + w] = cmulc(B0[x * sw + w], KS[x * sw + w]);
    }
    __syncthreads();
    T2* R = lds_fft<T>(B0, B1, tws, plan, W, sw, false);
    T2* orow = out + o * out_o + (long long)nx * (y0 + (long long)ny * kz);
    for (int i = threadIdx.x; i < n * W; i += blockDim.x) {
        const int w = i / n, x = i % n;
        if (w < wv && x < lx) orow[x + (long long)nx * w] = R[x * sw + w];
    }
}

// ---------------------------------------------------------------------------
// C2R along z for the legal columns (x < lx, y < ly), fused with rounding, the
// reference's per-value checks (residual < 0.25, 0 <= v <= min(mass_B,
// mass_K)), the feasible mask (round(v) == 0) and the lowest (z, y, x) choice.
// Grid: x = column tiles (W of x), y = row y, z = orientation.  OUT: the
// overlap fields and/or raw values are written (else only the choice).
// Rounding is to nearest-even: it differs from round-half-away only at an
// exact .5, whose residual 0.5 already fails the step's check.
template <typename T, bool OUT>
__global__ void __launch_bounds__(256) k_c2r_z_argmin(const cx<T>* __restrict__ in, long long in_o, int nx, int ny,
                                                      const OrientDesc* __restrict__ descs, T inv_total,
                                                      StepDeviceResult* __restrict__ result, int* __restrict__ fields,
                                                      const long long* __restrict__ field_offsets,
                                                      const unsigned long long* __restrict__ excl,
                                                      double* __restrict__ raw, StepFilter filt,
                                                      const cx<T>* __restrict__ tw, FftPlan plan, int W)
{
    using T2 = cx<T>;
    extern __shared__ unsigned char smem_raw[];
    const int nz = plan.n;
    const int P = W / 2;
    const int sw = P + 1;
    T2* tws = reinterpret_cast<T2*>(smem_raw);
    T2* A = tws + nz;
    T2* B = A + nz * sw;
    __shared__ unsigned long long best;
    __shared__ unsigned int res_bits, flags;
    const int o = slot_orientation(filt.act);
    if (o < 0) return;
    const OrientDesc d = descs[o];
    const int y = blockIdx.y;
    const int x0 = blockIdx.x * W;
    if (d.lx <= 0 || d.ly <= 0 || d.lz <= 0 || y >= d.ly || x0 >= d.lx) return;
    if (threadIdx.x == 0) {
        best = ~0ULL;
        res_bits = 0;
        flags = 0;
    }
    load_twiddles<T>(tws, tw, nz);
    const int nh = nz / 2 + 1;
    const cx<T>* src = in + o * in_o + (long long)nx * y + x0;
    const long long plane = (long long)nx * ny;
    // columns x >= lx were never written by the earlier passes: read as zero
    hermitian_pairs<T>(A, sw, src, plane, nz, nh, P, d.lx - x0);
    __syncthreads();
    T2* R = lds_fft<T>(A, B, tws, plan, P, sw, false);
    const T max_out = T(d.mass < (long long)result->env_mass ? d.mass : (long long)result->env_mass);
    unsigned long long mine = ~0ULL;
    float my_res = 0.f;
    unsigned int my_flags = 0;
    int* fo = OUT && fields ? fields + field_offsets[o] : nullptr;
    const bool any_excluded = excl[0] != 0;
    const bool regions_on = filt.act && filt.act[1];
    // blockDim.x is a multiple of W: each thread keeps one column w
    const int w = threadIdx.x % W;
    const int x = x0 + w;
    if (x < d.lx) {
        for (int z = threadIdx.x / W; z < d.lz; z += blockDim.x / W) {
            const T2 r = R[z * sw + (w % P)];
            const T v = (w < P ? r.x : r.y) * inv_total;
            const T rv = rint(v);
            const T res = fabs(v - rv);
            my_res = fmaxf(my_res, float(res));
            if (!(res < T(0.25))) my_flags |= 1u;
            if (rv < T(0) || rv > max_out) my_flags |= 2u;
            if (OUT) {
                if (fo) fo[x + (long long)d.lx * (y + (long long)d.ly * z)] = int(rv);
                if (raw) raw[field_offsets[o] + x + (long long)d.lx * (y + (long long)d.ly * z)] = double(v);
            }
            if (rv == T(0)) {
                const unsigned long long key = ((unsigned long long)z << 48) | ((unsigned long long)y << 32) |
                                               ((unsigned long long)x << 16) | (unsigned long long)o;
                if ((!any_excluded || !excluded(excl, key)) && (!regions_on || region_allowed(filt, x, y, z)))
                    mine = key < mine ? key : mine;
            }
        }
    }
    if (mine != ~0ULL) atomicMin(&best, mine);
    if (my_res > 0.f) atomicMax(&res_bits, __float_as_uint(my_res));
    if (my_flags) atomicOr(&flags, my_flags);
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

// C2R along z storing the real field (times scale) for every column x < cx_,
// y < cy_ and z < cz_ into out[z][y][x] (pitches nx, ny).  Used by the
// full-linear correlation.
template <typename T>
__global__ void __launch_bounds__(256) k_c2r_z_store(const cx<T>* __restrict__ in, int nx, int ny, T scale,
                                                     T* __restrict__ out, const cx<T>* __restrict__ tw, FftPlan plan,
                                                     int W)
{
    using T2 = cx<T>;
    extern __shared__ unsigned char smem_raw[];
    const int nz = plan.n;
    const int P = W / 2;
    const int sw = P + 1;
    T2* tws = reinterpret_cast<T2*>(smem_raw);
    T2* A = tws + nz;
    T2* B = A + nz * sw;
    const int y = blockIdx.y;
    const int x0 = blockIdx.x * W;
    if (x0 >= nx) return;
    load_twiddles<T>(tws, tw, nz);
    const int nh = nz / 2 + 1;
    const long long plane = (long long)nx * ny;
    const cx<T>* src = in + (long long)nx * y + x0;
    hermitian_pairs<T>(A, sw, src, plane, nz, nh, P, nx - x0);
    __syncthreads();
    T2* R = lds_fft<T>(A, B, tws, plan, P, sw, false);
    for (int i = threadIdx.x; i < nz * W; i += blockDim.x) {
        const int z = i / W, w = i % W;
        const int x = x0 + w;
        if (x >= nx) continue;
        const T2 r = R[z * sw + (w % P)];
        out[x + (long long)nx * (y + (long long)ny * z)] = (w < P ? r.x : r.y) * scale;
    }
}

}  // namespace sphip
