// Proximity field and ranked candidate pages on the device.
//
// engine/pack_solver/src/field_proximity.hpp:32-132 (build_proximity): exact
// squared Euclidean distance from every cell to the nearest occupied cell of
// the environment, then exp(-sqrt(d2)/2), 0 where nothing is occupied.
// engine/pack_solver/src/spectral_pipeline.cpp:1770-1857: per orientation,
// every legal translation gets score .65 * top - .35 * mean proximity; the 32
// best after an exclusive cursor form a page, ordered by (score, translation
// z, y, x) -- in low-overlap pages by binary overlap first -- and ordinary
// pages keep only rounded-zero overlaps and are rechecked directly.
#pragma once

#include "fft_kernels.hpp"

namespace sphip {

// Squared distances are exact 32-bit integers; "no occupied cell on any line
// reached" is kEdtInf.  Lines of up to 1024 cells keep every finite value and
// kEdtInf + (n-1)^2 below 2^32.
constexpr unsigned kEdtInf = 0x7fffffffu;
constexpr int kMaxRankExtent = 1024;
constexpr int kPageSize = 32;
constexpr int kPageCap = 2048;        // LDS merge buffer (partial pages of one orientation)
constexpr int kPartialCap = 2048;     // LDS candidate buffer of a page block (a power of two)
constexpr int kPageBlockThreads = 256;
constexpr int kPagePerIter = 4 * kPageBlockThreads;
constexpr int kMaxPageBlocks = kPageCap / kPageSize;

// X pass: one wave per row (y, z).  Lane c keeps the ballot of chunk c (64
// cells) of the row, prefix/suffix scans over lanes give the last occupied
// cell before and the first after each chunk, and every cell takes the
// nearer of its left and right occupied neighbours.  Rows of at most 4096
// cells (64 chunks).  No block barrier: a wave past the last row exits.
__global__ __launch_bounds__(256) void k_edt_rows(const unsigned char* __restrict__ occ, int ex, long long rows,
                                                  unsigned* __restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const long long row = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= rows) return;
    const unsigned char* src = occ + row * ex;
    unsigned* dst = out + row * ex;
    const int chunks = (ex + 63) >> 6;
    unsigned long long mine = 0;
    for (int c = 0; c < chunks; ++c) {
        const int x = c * 64 + lane;
        const unsigned long long m = __ballot(x < ex && src[x] != 0);
        if (lane == c) mine = m;
    }
    int last = mine ? lane * 64 + 63 - __clzll((long long)mine) : -1;
    int first = mine ? lane * 64 + __ffsll((unsigned long long)mine) - 1 : 0x7fffffff;
    for (int off = 1; off < 64; off <<= 1) {
        const int l = __shfl_up(last, off, 64);
        const int f = __shfl_down(first, off, 64);
        if (lane >= off) last = l > last ? l : last;
        if (lane + off < 64) first = f < first ? f : first;
    }
    for (int c = 0; c < chunks; ++c) {
        const unsigned long long m = __shfl(mine, c, 64);
        const int before = c > 0 ? __shfl(last, c - 1, 64) : -1;
        const int after = c + 1 < 64 ? __shfl(first, c + 1, 64) : 0x7fffffff;
        const int x = c * 64 + lane;
        const unsigned long long lo = lane == 63 ? m : (m & ((2ull << lane) - 1ull));
        const unsigned long long hi = m & (~0ull << lane);
        const int left = lo ? c * 64 + 63 - __clzll((long long)lo) : before;
        const int right = hi ? c * 64 + __ffsll(hi) - 1 : after;
        unsigned d = kEdtInf;
        if (left >= 0) d = unsigned(x - left) * unsigned(x - left);
        if (right != 0x7fffffff) {
            const unsigned r = unsigned(right - x) * unsigned(right - x);
            d = r < d ? r : d;
        }
        if (x < ex) dst[x] = d;
    }
}

// Y or Z pass: d(q) = min_p f(p) + (q - p)^2 along lines of n cells at
// `stride`, tx consecutive x per block staged in LDS.  The search runs
// outward from q and stops once r^2 reaches the best value: exact, since
// f >= 0.  The final pass also writes exp(-sqrt(d)/2) from a table the host
// filled with the same expression (0 for kEdtInf).
template <bool FINAL>
__global__ __launch_bounds__(256) void k_edt_lines(unsigned* __restrict__ f, int ex, int n, long long stride,
                                                   long long outer_stride, int tx, const double* __restrict__ table,
                                                   double* __restrict__ values)
{
    extern __shared__ unsigned edt_tile[];
    const int x0 = blockIdx.x * tx;
    const long long base = (long long)blockIdx.y * outer_stride + x0;
    const int w = ex - x0 < tx ? ex - x0 : tx;
    const int total = n * tx;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int c = i % tx, q = i / tx;
        edt_tile[i] = c < w ? f[base + q * stride + c] : kEdtInf;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int c = i % tx, q = i / tx;
        if (c >= w) continue;
        unsigned best = edt_tile[i];
        const int reach = q > n - 1 - q ? q : n - 1 - q;
        for (int r = 1; r <= reach; ++r) {
            const unsigned rr = unsigned(r) * unsigned(r);
            if (rr >= best) break;
            if (q >= r) {
                const unsigned v = edt_tile[(q - r) * tx + c] + rr;
                best = v < best ? v : best;
            }
            if (q + r < n) {
                const unsigned v = edt_tile[(q + r) * tx + c] + rr;
                best = v < best ? v : best;
            }
        }
        const long long at = base + q * stride + c;
        f[at] = best;
        if (FINAL) values[at] = best >= kEdtInf ? 0.0 : table[best];
    }
}

// A page entry: order (overlap, score, position); overlap is 0 in ordinary
// pages, `score` holds order-preserving bits of the FP64 score, `pos` packs
// the corner (uz << 32 | uy << 16 | ux).  `direct` is the direct recheck.
struct PageKey {
    unsigned long long score;
    unsigned long long pos;
    unsigned int overlap;
    unsigned int direct;
};

__device__ __forceinline__ bool key_less(const PageKey& a, const PageKey& b)
{
    if (a.overlap != b.overlap) return a.overlap < b.overlap;
    if (a.score != b.score) return a.score < b.score;
    return a.pos < b.pos;
}

__device__ __forceinline__ PageKey max_page_key() { return PageKey { ~0ull, ~0ull, 0xffffffffu, 0u }; }

// Monotone map of doubles to unsigned integers (-0 taken as +0, so equal
// scores compare equal as in the reference's `!=`).
__host__ __device__ inline unsigned long long score_order_bits(double s)
{
    if (s == 0.0) s = 0.0;
    unsigned long long b;
    __builtin_memcpy(&b, &s, sizeof b);
    return (b >> 63) ? ~b : (b | (1ull << 63));
}

__host__ __device__ inline double score_from_order_bits(unsigned long long k)
{
    const unsigned long long b = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
    double s;
    __builtin_memcpy(&s, &b, sizeof s);
    return s;
}

// spectral_pipeline.cpp:1797-1801, evaluated in the same order with no
// fused multiply-add, so equal inputs give the reference expression's bits.
struct RankParams {
    double origin_z, pitch, container_min_z, container_height;
    long long env_first_z;
    int mode;           // 0 ordinary, 1 low overlap
    int any_cursor;
};

__device__ __forceinline__ double page_score(double origin_z, double pitch, double container_min_z,
                                          double container_height, double max_z, long long tz, double prox,
                                          double mass)
{
#pragma clang fp contract(off)
    const double world_z = origin_z + pitch * double(tz);
    const double top = (world_z + max_z - container_min_z) / container_height;
    const double mean = prox / mass;
    return .65 * top - .35 * mean;
}

// Bitonic sort of buf[0, m) (m a power of two) by key_less, whole block.
__device__ void page_sort(PageKey* buf, int m)
{
    for (int size = 2; size <= m; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = threadIdx.x; t < m / 2; t += blockDim.x) {
                const int lo = 2 * t - (t & (stride - 1));
                const int hi = lo + stride;
                const bool up = (lo & size) == 0;
                const PageKey a = buf[lo], b = buf[hi];
                if (key_less(b, a) == up) {
                    buf[lo] = b;
                    buf[hi] = a;
                }
            }
            __syncthreads();
        }
}

// Keeps the best kPageSize of buf[0, count) at its front and the threshold
// (the kPageSize-th best, or the maximum key while fewer are kept).
__device__ void page_flush(PageKey* buf, int* count, PageKey* thr)
{
    const int n = *count;
    int m = kPageSize;
    while (m < n) m <<= 1;
    for (int i = n + threadIdx.x; i < m; i += blockDim.x) buf[i] = max_page_key();
    __syncthreads();
    page_sort(buf, m);
    if (threadIdx.x == 0) {
        *count = n < kPageSize ? n : kPageSize;
        *thr = n >= kPageSize ? buf[kPageSize - 1] : max_page_key();
    }
    __syncthreads();
}

// Stage 1: block (chunk, slot) scans a contiguous chunk of one orientation's
// legal corners (z-major) and keeps its best kPageSize candidates in LDS.
// Rejected corners (incremental fields hold them at 2^30 and above, or the
// rejected-translation list) never enter a page.
__global__ __launch_bounds__(kPageBlockThreads) void k_page_partial(
    const int* __restrict__ F, const double* __restrict__ P, const long long* __restrict__ offsets,
    const OrientDesc* __restrict__ descs, const int* __restrict__ olist, RankParams rp,
    const double* __restrict__ max_z, const long long* __restrict__ kfirst_z, const PageKey* __restrict__ cursors,
    const unsigned long long* __restrict__ excl, PageKey* __restrict__ partial, int* __restrict__ partial_count)
{
    __shared__ PageKey buf[kPartialCap];
    __shared__ int count;
    __shared__ PageKey thr;
    const int slot = blockIdx.y;
    const int o = olist[slot];
    const OrientDesc d = descs[o];
    const long long plane = (long long)(d.lx > 0 ? d.lx : 0) * (d.ly > 0 ? d.ly : 0);
    const long long total = d.lx > 0 ? plane * d.lz : 0;
    const long long chunk = (total + gridDim.x - 1) / gridDim.x;
    long long begin = chunk * blockIdx.x;
    begin = begin < total ? begin : total;
    const long long end = begin + chunk < total ? begin + chunk : total;
    if (threadIdx.x == 0) {
        count = 0;
        thr = max_page_key();
    }
    __syncthreads();
    const int* fo = F + offsets[o];
    const double* po = P + offsets[o];
    const PageKey cur = cursors[slot];
    const bool has_cur = rp.any_cursor && cur.direct != 0;
    const bool any_excluded = excl[0] != 0;
    const double mz = max_z[o];
    const double mass = double(d.mass);
    // corner -> (z, y, x) by FP64 reciprocals, corrected to the exact
    // quotient (indices stay below 2^30, so the estimate is off by at most 1)
    const unsigned uplane = unsigned(plane), ulx = unsigned(d.lx > 0 ? d.lx : 1);
    const double inv_plane = 1.0 / double(uplane > 0 ? uplane : 1u), inv_lx = 1.0 / double(ulx);
    constexpr int J = kPagePerIter / kPageBlockThreads;
    // the overlaps of the next iteration load while this one is scored, and
    // the proximity loads of its eligible corners are issued together
    int ov[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const long long i = begin + j * kPageBlockThreads + threadIdx.x;
        ov[j] = i < end ? fo[i] : (1 << 30);
    }
    for (long long i0 = begin; i0 < end; i0 += kPagePerIter) {
        bool ok[J];
        double pv[J];
#pragma unroll
        for (int j = 0; j < J; ++j) {
            ok[j] = ov[j] < (1 << 30) && (rp.mode != 0 || ov[j] == 0);
            pv[j] = ok[j] ? po[i0 + j * kPageBlockThreads + threadIdx.x] : 0.0;
        }
        // issued after the proximity loads: memory counters complete in
        // order, so scoring waits for those while these stay in flight
        int nv[J];
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const long long i = i0 + kPagePerIter + j * kPageBlockThreads + threadIdx.x;
            nv[j] = i < end ? fo[i] : (1 << 30);
        }
#pragma unroll
        for (int j = 0; j < J; ++j) {
            if (!ok[j]) continue;
            const unsigned ui = unsigned(i0 + j * kPageBlockThreads + threadIdx.x);
            unsigned z = unsigned(double(ui) * inv_plane);
            if (z * uplane > ui) --z;
            else if ((z + 1u) * uplane <= ui) ++z;
            const unsigned rem = ui - z * uplane;
            unsigned y = unsigned(double(rem) * inv_lx);
            if (y * ulx > rem) --y;
            else if ((y + 1u) * ulx <= rem) ++y;
            const unsigned x = rem - y * ulx;
            if (any_excluded) {
                const unsigned long long skey = ((unsigned long long)z << 48) | ((unsigned long long)y << 32) |
                                                ((unsigned long long)x << 16) | (unsigned long long)o;
                if (excluded(excl, skey)) continue;
            }
            const double s = page_score(rp.origin_z, rp.pitch, rp.container_min_z, rp.container_height, mz,
                                        rp.env_first_z + z - kfirst_z[o], pv[j], mass);
            const PageKey k { score_order_bits(s),
                              ((unsigned long long)z << 32) | ((unsigned long long)y << 16) | (unsigned long long)x,
                              rp.mode ? unsigned(ov[j]) : 0u, 0u };
            if (has_cur && !key_less(cur, k)) continue;
            if (!key_less(k, thr)) continue;
            const int at = atomicAdd(&count, 1);
            buf[at] = k;
        }
        // every thread reads the count before any thread adds again, so the
        // flush (which holds barriers) is taken by the whole block or by none
        __syncthreads();
        const int filled = count;
        __syncthreads();
        if (filled > kPartialCap - kPagePerIter) page_flush(buf, &count, &thr);
#pragma unroll
        for (int j = 0; j < J; ++j) ov[j] = nv[j];
    }
    __syncthreads();
    if (count > 0) page_flush(buf, &count, &thr);
    const int n = count;
    PageKey* dst = partial + ((long long)slot * gridDim.x + blockIdx.x) * kPageSize;
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = buf[i];
    if (threadIdx.x == 0) partial_count[slot * gridDim.x + blockIdx.x] = n;
}

// Stage 2: one block per slot merges the partial pages (at most
// kMaxPageBlocks of them) into the page.
__global__ __launch_bounds__(kPageBlockThreads) void k_page_merge(const PageKey* __restrict__ partial,
                                                                  const int* __restrict__ partial_count, int blocks,
                                                                  PageKey* __restrict__ pages,
                                                                  int* __restrict__ page_count)
{
    __shared__ PageKey buf[kPageCap];
    __shared__ int base[kMaxPageBlocks + 1];
    const int slot = blockIdx.x;
    if (threadIdx.x == 0) {
        int s = 0;
        for (int b = 0; b < blocks; ++b) {
            base[b] = s;
            s += partial_count[slot * blocks + b];
        }
        base[blocks] = s;
    }
    __syncthreads();
    const int n = base[blocks];
    for (int b = 0; b < blocks; ++b) {
        const int c = base[b + 1] - base[b];
        for (int i = threadIdx.x; i < c; i += blockDim.x)
            buf[base[b] + i] = partial[((long long)slot * blocks + b) * kPageSize + i];
    }
    int m = kPageSize;
    while (m < n) m <<= 1;
    for (int i = n + threadIdx.x; i < m; i += blockDim.x) buf[i] = max_page_key();
    __syncthreads();
    page_sort(buf, m);
    const int keep = n < kPageSize ? n : kPageSize;
    for (int i = threadIdx.x; i < keep; i += blockDim.x) pages[(long long)slot * kPageSize + i] = buf[i];
    if (threadIdx.x == 0) page_count[slot] = keep;
}

// Stage 3, ordinary pages: block (entry, slot) rechecks one page entry
// directly against the occupancy (spectral_pipeline.cpp:1825-1855); each wave
// walks kernel rows (y, z), its lanes the row's cells.
__global__ __launch_bounds__(kPageBlockThreads) void k_page_recheck(const int* __restrict__ olist,
                                                                    const OrientDesc* __restrict__ descs,
                                                                    const unsigned char* __restrict__ kcells,
                                                                    const unsigned char* __restrict__ env, int ex, int ey,
                                                                    PageKey* __restrict__ pages,
                                                                    const int* __restrict__ page_count)
{
    const int slot = blockIdx.y, e = blockIdx.x;
    if (e >= page_count[slot]) return;
    PageKey* k = pages + (long long)slot * kPageSize + e;
    const unsigned long long p = k->pos;
    const int ux = int(p & 0xffff), uy = int((p >> 16) & 0xffff), uz = int(p >> 32);
    const OrientDesc d = descs[olist[slot]];
    const unsigned char* kc = kcells + d.cells_offset;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, waves = blockDim.x >> 6;
    int hit = 0;
    for (int q = wave; q < d.ky * d.kz && !hit; q += waves) {
        const int cz = q / d.ky, cy = q - cz * d.ky;
        const unsigned char* krow = kc + (long long)q * d.kx;
        const unsigned char* erow = env + ux + (long long)ex * ((uy + cy) + (long long)ey * (uz + cz));
        for (int cx = lane; cx < d.kx; cx += 64)
            if (krow[cx] && erow[cx]) hit = 1;
        hit = __any(hit);
    }
    hit = __syncthreads_or(hit);
    if (threadIdx.x == 0) k->direct = hit ? 0u : 1u;
}

}  // namespace sphip
