// Batched placement step (gap_fill.cpp:363-509 semantics) on gfx950.
//
// One step = environment forward FFT, every orientation's pruned kernel
// spectrum, the product and inverse passes, rounding + checks + feasible mask
// + lowest (z, y, x) choice, and stamping the chosen kernel.  Every kernel of a
// step runs on one stream; with capture_graph the sequence is recorded once as
// a hipGraph and replayed per step.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <numeric>
#include <set>
#include <type_traits>

#include "common.hpp"
#include "fft_kernels.hpp"
#include "rank_kernels.hpp"
#include "rf_kernels.hpp"
#include "session.hpp"

namespace sphip {

namespace {

constexpr int kThreads = 256;
// Blocks of the single-pose kernels (recheck, stamp): one wave per kernel row
// or so at typical kernel sizes, so that a row's dependent loads are not
// waited for one after another
constexpr int kSmallGrid = 256;
constexpr std::size_t kMaxExcluded = 65536;
constexpr unsigned int kCopyLogCap = 1u << 20;  // copies whose footprint a session keeps (removable copies)
constexpr std::size_t kLdsBudget = 64 * 1024;

// Largest tile width (power of two, <= cap) whose LDS footprint fits.
int choose_w(int n, std::size_t elem, int buffers, bool pairs, int cap, int min_w)
{
    int best = min_w;
    for (int w = min_w; w <= cap; w *= 2) {
        const int pencils = pairs ? w / 2 : w;
        const std::size_t bytes = elem * (std::size_t(n) + std::size_t(buffers) * n * (pencils + 1));
        if (bytes <= kLdsBudget) best = w;
    }
    return best;
}

template <typename T> std::size_t lds_bytes(int n, int buffers, int pencils)
{
    return sizeof(cx<T>) * (std::size_t(n) + std::size_t(buffers) * n * (pencils + 1));
}

__global__ void k_reset_result(StepDeviceResult* r, unsigned long long* orient_best, int orientations)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        r->key = ~0ULL;
        r->run_flags |= r->flags;
        r->flags = 0;
        r->max_residual_bits = 0;
    }
    for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < orientations; o += gridDim.x * blockDim.x)
        orient_best[o] = ~0ULL;
}

__global__ void k_set_key(StepDeviceResult* r, unsigned long long key)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) r->key = key;
}

// Direct recheck of the step's choice, the one translation the step acts on
// (the reference's deterministic direct probes, correlation.cpp:549-577, ask
// the same of eight outputs): every cell of the chosen kernel at the chosen
// corner must be free in the occupancy the choice was made on.  Any overlap
// sets flag bit 2, and the stamp that follows is skipped.
__global__ void k_verify_choice(const unsigned char* __restrict__ env, int ex, int ey, int ez,
                                const unsigned char* __restrict__ kcells, const OrientDesc* __restrict__ descs,
                                StepDeviceResult* r)
{
    const unsigned long long key = r->key;
    if (key == ~0ULL) return;
    const int o = int(key & 0xffffULL);
    const int ux = int((key >> 16) & 0xffffULL), uy = int((key >> 32) & 0xffffULL), uz = int(key >> 48);
    const OrientDesc d = descs[o];
    // each wave takes kernel rows (z, y) in turn and its lanes the row's x
    // cells (no per-cell index division)
    const int lane = threadIdx.x & 63, waves = blockDim.x >> 6;
    const int rows = d.ky * d.kz;
    unsigned int hit = 0;
    for (int rw = blockIdx.x * waves + (threadIdx.x >> 6); rw < rows; rw += waves * gridDim.x) {
        const int z = rw / d.ky, y = rw - z * d.ky;
        const unsigned char* krow = kcells + d.cells_offset + (long long)rw * d.kx;
        // a corner outside the window, or a kernel cell beyond it, is an overlap
        const bool inside = uz + z < ez && uy + y < ey;
        const long long erow = (long long)ex * ((uy + y) + (long long)ey * (uz + z)) + ux;
        for (int x = lane; x < d.kx; x += 64)
            if (krow[x] && (!inside || ux + x >= ex || env[erow + x])) hit = 1;
    }
    if (__any(hit) && lane == 0) atomicOr(&r->flags, 4u);
}

// Placed copies (conservative_fields.cpp:1666-1887, BlockedField): every
// stamped copy is appended to a log of its keys (log[0] = copies stamped,
// log[1 + id] = key of copy id while id < capacity) and adds one to the placed
// count of every cell of its stamp footprint.  A cell is blocked while the
// environment mask blocks it or its placed count is nonzero.
struct CopyTrack {
    unsigned int* counts;     // placed count per environment cell (null: not tracked)
    unsigned long long* log;
    unsigned int log_cap;
};

// ORs the chosen kernel into the occupancy (gap_fill.cpp:98-121 for the chosen
// pose) and adds the newly set cells to the environment mass.  A step whose
// checks failed (any flag) places nothing, as a failed correlation places
// nothing in the reference.
__global__ void k_stamp(unsigned char* env, int ex, int ey, int ez, const unsigned char* __restrict__ kcells,
                        const OrientDesc* __restrict__ descs, StepDeviceResult* r, CopyTrack track,
                        unsigned char* collect = nullptr, int cx = 0, int cy = 0, int mx = 0, int my = 0, int mz = 0,
                        unsigned long long* collect_mass = nullptr, int shift = 0)
{
    const unsigned long long key = r->key;
    if (key == ~0ULL || r->flags != 0) return;
    if (track.log && blockIdx.x == 0 && threadIdx.x == 0) {
        const unsigned long long id = track.log[0];
        if (id < track.log_cap) track.log[1 + id] = key;
        track.log[0] = id + 1;
    }
    const int o = int(key & 0xffffULL);
    const int ux = int((key >> 16) & 0xffffULL), uy = int((key >> 32) & 0xffffULL), uz = int(key >> 48);
    const OrientDesc d = descs[o];
    unsigned long long added = 0;
    // each wave takes kernel rows (z, y) in turn and its lanes the row's x
    // cells (distinct kernel cells land on distinct environment cells)
    const int lane = threadIdx.x & 63, waves = blockDim.x >> 6;
    const int rows = d.ky * d.kz;
    for (int rw = blockIdx.x * waves + (threadIdx.x >> 6); rw < rows; rw += waves * gridDim.x) {
        const int z = rw / d.ky, y = rw - z * d.ky;
        const int gy = uy - shift + y, gz = uz - shift + z;
        if (gy < 0 || gz < 0 || gy >= ey || gz >= ez) continue;
        const unsigned char* krow = kcells + d.cells_offset + (long long)rw * d.kx;
        const long long erow = (long long)ex * (gy + (long long)ey * gz);
        for (int x = lane; x < d.kx; x += 64) {
            const int gx = ux - shift + x;
            if (!krow[x] || gx < 0 || gx >= ex) continue;
            const long long idx = erow + gx;
            if (track.counts) atomicAdd(track.counts + idx, 1u);
            unsigned char* c = env + idx;
            if (*c == 0) {
                *c = 1;
                ++added;
                // newly set cell, at kernel cell + margin in the update window
                if (collect) collect[(x + mx) + (long long)cx * ((y + my) + (long long)cy * (z + mz))] = 1;
            }
        }
    }
    // one atomic per wave (a wave's sum is at most the kernel's cell count)
    unsigned int wadded = unsigned(added);
    for (int off = 32; off > 0; off >>= 1) wadded += __shfl_xor(wadded, off, 64);
    if (lane == 0 && wadded) {
        atomicAdd(&r->env_mass, (unsigned long long)wadded);
        if (collect_mass) atomicAdd(collect_mass, (unsigned long long)wadded);
    }
}

// conservative_fields.cpp:1848-1887 BlockedField::remove: subtracts one placed
// count over the footprint of the copy with `key` (the stamp kernel at the
// corner minus `shift`); a cell whose count reaches zero and that the mask
// leaves free becomes free, is collected at stamp-kernel coordinates into
// the update window (when given) and leaves the environment mass.  A count
// that was already zero sets bit 0 of *fault (the count invariant).
__global__ void k_unstamp(unsigned char* env, const unsigned char* __restrict__ mask, int ex, int ey, int ez,
                          const unsigned char* __restrict__ kcells, const OrientDesc* __restrict__ descs,
                          unsigned long long key, unsigned int* counts, StepDeviceResult* r, unsigned int* fault,
                          unsigned char* collect, int cx, int cy, unsigned long long* collect_mass, int shift)
{
    const int o = int(key & 0xffffULL);
    const int ux = int((key >> 16) & 0xffffULL), uy = int((key >> 32) & 0xffffULL), uz = int(key >> 48);
    const OrientDesc d = descs[o];
    const long long n = (long long)d.kx * d.ky * d.kz;
    unsigned long long freed = 0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        if (!kcells[d.cells_offset + i]) continue;
        const int x = int(i % d.kx), y = int((i / d.kx) % d.ky), z = int(i / ((long long)d.kx * d.ky));
        const int gx = ux - shift + x, gy = uy - shift + y, gz = uz - shift + z;
        if (gx < 0 || gy < 0 || gz < 0 || gx >= ex || gy >= ey || gz >= ez) continue;
        const long long idx = gx + (long long)ex * (gy + (long long)ey * gz);
        const unsigned int old = atomicSub(counts + idx, 1u);
        if (old == 0u) {
            atomicAdd(counts + idx, 1u);
            atomicOr(fault, 1u);
            continue;
        }
        if (old == 1u && mask[idx] == 0 && env[idx] != 0) {
            env[idx] = 0;
            ++freed;
            if (collect) collect[x + (long long)cx * (y + (long long)cy * z)] = 1;
        }
    }
    if (freed) {
        atomicAdd(&r->env_mass, 0ULL - freed);
        if (collect_mass) atomicAdd(collect_mass, freed);
    }
}

// After a removal every orientation's feasible set may grow: each minimum is
// searched again from the first corner.
__global__ void k_rescan_all(unsigned long long* __restrict__ best, int* __restrict__ rescan, int orientations)
{
    for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < orientations; o += gridDim.x * blockDim.x) {
        best[o] = 0ULL;
        rescan[o] = 1;
    }
}


// ---- pair clearance (conservative_fields.cpp:189-204, 398-434, 1015-1122) --
// A placed copy blocks every cell within the clearance stencil of its cells.
// Each orientation's stamp kernel is its kernel dilated by the stencil, in a
// window grown by the halo on every side; stamping places it at the corner
// minus the halo.

__global__ void k_dilate(const unsigned char* __restrict__ kcells, const OrientDesc* __restrict__ kd,
                         const OrientDesc* __restrict__ sd, unsigned char* __restrict__ scells, const int* __restrict__ st,
                         int count, int halo)
{
    const int o = blockIdx.y;
    const OrientDesc k = kd[o], d = sd[o];
    const long long n = (long long)k.kx * k.ky * k.kz;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        if (!kcells[k.cells_offset + i]) continue;
        const int x = int(i % k.kx) + halo, y = int((i / k.kx) % k.ky) + halo, z = int(i / ((long long)k.kx * k.ky)) + halo;
        for (int j = 0; j < count; ++j) {
            const int sx = x + st[3 * j], sy = y + st[3 * j + 1], sz = z + st[3 * j + 2];
            scells[d.cells_offset + sx + (long long)d.kx * (sy + (long long)d.ky * sz)] = 1;
        }
    }
}

// ---- incremental overlap fields (gap_fill.cpp:426-451,524-584 restated) ----
// The occupancy only grows, so after new cells N are set every orientation's
// overlap field grows by the correlation of N with its kernel, and its set of
// feasible corners only shrinks.  F (int32, exact) is updated by adding that
// small correlation, computed on an update window around N; an orientation
// whose lowest feasible key became blocked is rescanned forward from it.

__device__ __forceinline__ unsigned long long pack_key(int ux, int uy, int uz, int o)
{
    return ((unsigned long long)uz << 48) | ((unsigned long long)uy << 32) | ((unsigned long long)ux << 16) |
           (unsigned long long)o;
}

// U_o is the circular correlation (period n' per axis) of the new cells,
// which sit at offsets [0, Kp) of the update window, with kernel o; n' >=
// Kp + K_o - 1, so its shift s stands for the linear shift s when
// s < n' - (K_o - 1) and s - n' otherwise.  F_o(u_p + shift) += U_o(s) for
// corners inside the main legal range, u_p the pending placement's corner;
// flags orientations whose current lowest feasible corner received overlap.
// Only the shifts where U_o can be nonzero are visited (a >= Kp bounds the
// new cells' extent per axis).

__global__ void k_apply_update(int* __restrict__ F, const OrientDesc* __restrict__ md, const long long* __restrict__ moff,
                               const int* __restrict__ U, const OrientDesc* __restrict__ ud,
                               const long long* __restrict__ uoff, const unsigned long long* __restrict__ origin,
                               int nxu, int nyu, int nzu, int ax, int ay, int az, const unsigned long long* __restrict__ best,
                               int* __restrict__ rescan, int shift, int sign)
{
    const int o = blockIdx.y;
    const OrientDesc d = md[o], e = ud[o];
    if (d.lx <= 0 || e.lx <= 0) return;
    const unsigned long long key = *origin;
    const int ox = int((key >> 16) & 0xffffULL) - shift, oy = int((key >> 32) & 0xffffULL) - shift,
              oz = int(key >> 48) - shift;
    // The update's support: the changed cells lie in [0, a) of the window per
    // axis, so their correlation with this kernel is nonzero only at shifts
    // [-(k - 1), a - 1], i.e. at circular indices [0, a) and [n - (k - 1), n).
    // Support index j < a is shift j; j >= a is shift j - s (s = a + k - 1).
    const int sx = min(ax + d.kx - 1, nxu), sy = min(ay + d.ky - 1, nyu), sz = min(az + d.kz - 1, nzu);
    const int axx = min(ax, nxu), ayy = min(ay, nyu), azz = min(az, nzu);
    const int* u = U + uoff[o];
    int* f = F + moff[o];
    const unsigned long long b = best[o];
    // each wave takes support rows (z, y) in turn and its lanes the row's x
    const int lane = threadIdx.x & 63, waves = blockDim.x >> 6;
    const int rows = sy * sz;
    for (int rw = blockIdx.x * waves + (threadIdx.x >> 6); rw < rows; rw += waves * gridDim.x) {
        const int jz = rw / sy, jy = rw - jz * sy;
        const int dz = jz < azz ? jz : jz - sz, dy = jy < ayy ? jy : jy - sy;
        const int tz = oz + dz, ty = oy + dy;
        if (ty < 0 || tz < 0 || ty >= d.ly || tz >= d.lz) continue;
        const int cz = dz < 0 ? nzu + dz : dz, cy = dy < 0 ? nyu + dy : dy;
        const int* urow = u + (long long)e.lx * (cy + (long long)e.ly * cz);
        int* frow = f + (long long)d.lx * (ty + (long long)d.ly * tz);
        for (int jx = lane; jx < sx; jx += 64) {
            const int dx = jx < axx ? jx : jx - sx;
            const int tx = ox + dx;
            if (tx < 0 || tx >= d.lx) continue;
            const int v = urow[dx < 0 ? nxu + dx : dx];
            if (v == 0) continue;
            frow[tx] += sign * v;
            if (pack_key(tx, ty, tz, o) == b) rescan[o] = 1;
        }
    }
}

// Marks a rejected corner blocked for good and flags a rescan if it was the minimum.
__global__ void k_block_corner(int* __restrict__ F, const OrientDesc* __restrict__ md, const long long* __restrict__ moff,
                               int o, int ux, int uy, int uz, const unsigned long long* __restrict__ best,
                               int* __restrict__ rescan)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const OrientDesc d = md[o];
    if (d.lx <= 0 || ux >= d.lx || uy >= d.ly || uz >= d.lz) return;
    F[moff[o] + ux + (long long)d.lx * (uy + (long long)d.ly * uz)] += 1 << 30;
    if (pack_key(ux, uy, uz, o) == best[o]) rescan[o] = 1;
}

// When flagged, the lowest (z, y, x) corner of an orientation with F == 0 at
// or after its old minimum (the feasible set only shrinks).  Keys follow the
// field's linear order (z-major, then y, then x), so the scan walks chunks of
// kRescanChunk corners from the old minimum's index: kRescanBlocks blocks
// share an orientation, block j taking chunks j, j + B, ... in order, and
// stops before a chunk past a zero already found (found[o] only decreases
// and each chunk has one owner, so every chunk below the final result is
// scanned in full).  k_rescan_finish then publishes found[o] as the minimum
// and resets it to ~0 for the next rescan; with one block per orientation
// (gridDim.y == 1, for small fields) the block does that itself.
constexpr int kRescanBlocks = 8;
constexpr int kRescanChunk = 8192;
constexpr long long kRescanWide = 1LL << 22;  // corners per orientation from which kRescanBlocks share one
__global__ __launch_bounds__(1024) void k_rescan(const int* __restrict__ F, const OrientDesc* __restrict__ md,
                                                 const long long* __restrict__ moff,
                                                 unsigned long long* __restrict__ best, int* __restrict__ rescan,
                                                 unsigned long long* __restrict__ found)
{
    __shared__ int stop;
    const int o = blockIdx.x;
    if (!rescan[o]) return;
    const OrientDesc d = md[o];
    const long long plane = d.lx > 0 ? (long long)d.lx * d.ly : 0, total = plane * (d.lx > 0 ? d.lz : 0);
    auto index_of = [&](unsigned long long k) {
        return (long long)(k >> 48) * plane + (long long)((k >> 32) & 0xffffULL) * d.lx + (long long)((k >> 16) & 0xffffULL);
    };
    const unsigned long long b0 = best[o];
    const int* f = F + moff[o];
    // (no feasible corner was left when b0 is ~0: nothing to scan)
    for (long long c0 = b0 == ~0ULL ? total : index_of(b0) + (long long)blockIdx.y * kRescanChunk; c0 < total;
         c0 += (long long)gridDim.y * kRescanChunk) {
        if (threadIdx.x == 0) {
            const unsigned long long k = __hip_atomic_load(&found[o], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            stop = k != ~0ULL && index_of(k) < c0;
        }
        __syncthreads();
        if (stop) break;
        // a thread's first zero is its smallest
        const long long c1 = c0 + kRescanChunk < total ? c0 + kRescanChunk : total;
        long long first = -1;
        for (long long i = c0 + threadIdx.x; i < c1; i += blockDim.x)
            if (f[i] == 0) {
                first = i;
                break;
            }
        unsigned long long mine = ~0ULL;
        if (first >= 0) {
            const int z = int(first / plane);
            const long long rem = first - z * plane;
            const int y = int(rem / d.lx);
            mine = pack_key(int(rem - (long long)y * d.lx), y, z, o);
        }
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long om = __shfl_down(mine, off, 64);
            mine = om < mine ? om : mine;
        }
        if ((threadIdx.x & 63) == 0 && mine != ~0ULL) atomicMin(&found[o], mine);
        __syncthreads();  // every wave has read `stop` before it is written again
    }
    if (gridDim.y == 1) {
        __syncthreads();  // every wave's atomicMin is done
        if (threadIdx.x == 0) {
            best[o] = found[o];
            found[o] = ~0ULL;
            rescan[o] = 0;
        }
    }
}

__global__ void k_rescan_finish(unsigned long long* __restrict__ best, int* __restrict__ rescan,
                                unsigned long long* __restrict__ found, int orientations)
{
    for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < orientations; o += gridDim.x * blockDim.x) {
        if (!rescan[o]) continue;
        best[o] = found[o];
        found[o] = ~0ULL;
        rescan[o] = 0;
    }
}

// The step's choice from the maintained minima; numeric flags of the update
// step carried over.
__global__ void k_choose(StepDeviceResult* __restrict__ res, const unsigned long long* __restrict__ best, int orientations,
                         const StepDeviceResult* __restrict__ upd, int take_flags)
{
    __shared__ unsigned long long m;
    if (threadIdx.x == 0) m = ~0ULL;
    __syncthreads();
    unsigned long long mine = ~0ULL;
    for (int o = threadIdx.x; o < orientations; o += blockDim.x) mine = best[o] < mine ? best[o] : mine;
    atomicMin(&m, mine);
    __syncthreads();
    if (threadIdx.x == 0) {
        res->key = m;
        if (take_flags) {
            res->run_flags |= res->flags;
            res->flags = upd->flags;
            res->max_residual_bits = upd->max_residual_bits;
        }
    }
}

// ---- register-FFT dispatch for power-of-R lengths -------------------------
template <typename T, int N, int R> constexpr std::size_t rf_lds(int pencils)
{
    return sizeof(cx<T>) * (std::size_t(N) + std::size_t(pencils) * RegShape<N, R>::stride);
}

inline bool rf_supported(int n) { return n == 128 || n == 256 || n == 512; }

// ---- tile candidates (radix R, tile width W) for the register-FFT passes ---
// Every candidate of a list is compiled; the autotuner times the ones whose
// LDS fits the device and keeps the fastest per shape.
struct Tile {
    int r = 0, w = 0;
};
template <int N, int R, int W> struct TileT {
    static constexpr int n = N, r = R, w = W;
};
template <typename... Ts> struct TileList {};

template <typename F, typename... Ts> bool dispatch_tile(TileList<Ts...>, int n, Tile t, F&& f)
{
    return ((n == Ts::n && t.r == Ts::r && t.w == Ts::w ? (f(Ts {}), true) : false) || ...);
}

template <typename... Ts> std::vector<Tile> tiles_for(TileList<Ts...>, int n)
{
    std::vector<Tile> v;
    ((n == Ts::n ? v.push_back(Tile { Ts::r, Ts::w }) : void()), ...);
    return v;
}

// P1 (kernel x pass + product + inverse x), P2 (inverse y), P3 (C2R z + argmin)
// (128 runs at R = 16 as the mixed 16 x 8 transform, eight threads a pencil)
using MulTiles = TileList<TileT<128, 16, 16>, TileT<128, 16, 32>, TileT<128, 16, 64>, TileT<256, 16, 8>,
                          TileT<256, 16, 16>, TileT<256, 16, 32>, TileT<256, 4, 4>, TileT<256, 4, 8>, TileT<256, 4, 16>,
                          TileT<512, 8, 2>, TileT<512, 8, 4>, TileT<512, 8, 8>, TileT<512, 8, 16>>;
using ColTiles = TileList<TileT<128, 16, 32>, TileT<128, 16, 64>, TileT<128, 16, 128>, TileT<256, 16, 16>,
                          TileT<256, 16, 32>, TileT<256, 16, 64>, TileT<256, 4, 8>, TileT<256, 4, 16>, TileT<512, 8, 4>,
                          TileT<512, 8, 8>, TileT<512, 8, 16>>;
using C2rTiles = TileList<TileT<128, 16, 64>, TileT<128, 16, 128>, TileT<128, 16, 256>, TileT<256, 16, 32>,
                          TileT<256, 16, 64>, TileT<256, 16, 128>, TileT<256, 4, 16>, TileT<256, 4, 32>, TileT<512, 8, 8>,
                          TileT<512, 8, 16>, TileT<512, 8, 32>>;

enum { kPassMul = 0, kPassInvY = 1, kPassC2r = 2, kPasses = 3 };

inline Tile default_tile(int pass, int n)
{
    if (pass == kPassMul) return n == 128 ? Tile { 16, 32 } : n == 256 ? Tile { 16, 16 } : Tile { 8, 4 };
    if (pass == kPassInvY) return n == 128 ? Tile { 16, 64 } : n == 256 ? Tile { 16, 32 } : Tile { 8, 16 };
    return n == 128 ? Tile { 16, 128 } : n == 256 ? Tile { 16, 64 } : Tile { 8, 16 };
}

inline std::vector<Tile> tile_candidates(int pass, int n)
{
    if (pass == kPassMul) return tiles_for(MulTiles {}, n);
    if (pass == kPassInvY) return tiles_for(ColTiles {}, n);
    return tiles_for(C2rTiles {}, n);
}

// Dynamic LDS bytes of a pass at a tile (elem = bytes per complex value).
inline std::size_t tile_lds(int pass, int n, Tile t, std::size_t elem)
{
    const std::size_t stride = std::size_t(n) + n / 16 + 1;
    if (pass == kPassC2r) {
        const std::size_t fft = elem * (std::size_t(n) + std::size_t(t.w / 2) * stride);
        const std::size_t stage = elem * (std::size_t(n) + std::size_t(n / 2 + 1) * (t.w + 1));
        return std::max(fft, stage);
    }
    return elem * (std::size_t(n) + std::size_t(t.w) * stride);
}

template <typename T>
void rf_cols_launch(int n, Tile tile, hipStream_t st, int count_w_max, int batches, int orients, const cx<T>* in,
                    long long in_o, long long in_sn, long long in_sw, long long in_sb, cx<T>* out, long long out_o,
                    long long out_sn, long long out_sw, long long out_sb, int count_w, int n_in, int n_out,
                    const OrientDesc* descs, int mode, const int* act, const cx<T>* tw, int fwd, int gap_ax = 0,
                    int gap_ay = 0)
{
    const bool ok = dispatch_tile(ColTiles {}, n, tile, [&](auto tt) {
        using TT = decltype(tt);
        constexpr int N = TT::n, R = TT::r, W = TT::w;
        hipLaunchKernelGGL((rf_cols<T, N, R, W>), dim3((count_w_max + W - 1) / W, batches, orients), dim3(W * (N / R)),
                           (rf_lds<T, N, R>(W)), st, in, in_o, in_sn, in_sw, in_sb, out, out_o, out_sn, out_sw, out_sb,
                           count_w, n_in, n_out, descs, mode, act, tw, fwd, gap_ax, gap_ay);
    });
    if (!ok) throw DeviceError { "no such FFT tile" };
}

template <typename T>
void rf_rows_launch(int n, hipStream_t st, cx<T>* data, int rows, long long row_pitch, int batches, long long batch_pitch,
                    const cx<T>* tw, int fwd)
{
    if (!rf_supported(n)) throw DeviceError { "register FFT length" };
    if (n == 128) {
        constexpr int N = 128, R = 16, W = 32;
        hipLaunchKernelGGL((rf_rows<T, N, R, W>), dim3((rows + W - 1) / W, batches), dim3(W * (N / R)), (rf_lds<T, N, R>(W)),
                           st, data, rows, row_pitch, batch_pitch, tw, fwd);
    }
    else if (n == 256) {
        constexpr int N = 256, R = 16, W = 16;
        hipLaunchKernelGGL((rf_rows<T, N, R, W>), dim3((rows + W - 1) / W, batches), dim3(W * (N / R)), (rf_lds<T, N, R>(W)),
                           st, data, rows, row_pitch, batch_pitch, tw, fwd);
    }
    else {
        constexpr int N = 512, R = 8, W = 4;
        hipLaunchKernelGGL((rf_rows<T, N, R, W>), dim3((rows + W - 1) / W, batches), dim3(W * (N / R)), (rf_lds<T, N, R>(W)),
                           st, data, rows, row_pitch, batch_pitch, tw, fwd);
    }
}

// Grid of the column-major product pass: one block per row tile (all
// orientations each) when the tiles fill whole rounds of resident blocks to
// within 2%, which keeps the blocks in flight on neighbouring rows; otherwise
// as many blocks as the device holds at once (its CUs times the blocks per CU
// that registers and LDS allow), each with an even share of the tiles times
// orientations.  The row-major form always takes one block per tile: its row
// stores of neighbouring blocks then meet, which measured faster than evenly
// shared blocks.
inline int mul_grid(const void* fn, int threads, std::size_t lds, int cus, long long tiles, int slots)
{
    int per = 0;
    SP_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, threads, lds));
    const long long cap = (long long)std::max(per, 1) * std::max(cus, 1);
    const long long rounds = (tiles + cap - 1) / cap;
    if ((rounds * cap - tiles) * 50 <= rounds * cap) return int(tiles);
    return int(std::max(1LL, std::min(cap, tiles * slots)));
}

// P1 over the active orientations.  One block per row tile when the tiles
// fill at least one round of resident blocks; fewer tiles (the 128-point
// window: 520 tiles of 128 threads) share their (tile, orientation) units
// evenly over one round, which would otherwise leave most of each CU idle.
template <typename T>
void rf_mul_launch(int n, Tile tile, hipStream_t st, int cus, int slots, int ny, int nh, const cx<T>* H, long long h_o,
                   int h_px, const cx<T>* SB, cx<T>* out, long long out_o, const OrientDesc* descs, const int* act,
                   const cx<T>* tw, int gap_ax = 0)
{
    const bool ok = dispatch_tile(MulTiles {}, n, tile, [&](auto tt) {
        using TT = decltype(tt);
        constexpr int N = TT::n, R = TT::r, W = TT::w;
        const long long tiles = (long long)(ny + W - 1) / W * nh;
        const int threads = W * (N / R);
        const std::size_t lds = rf_lds<T, N, R>(W);
        int per = 0;
        SP_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(&rf_mul_x<T, N, R, W>),
                                                                  threads, lds));
        const long long cap = (long long)std::max(per, 1) * std::max(cus, 1);
        const long long grid = tiles >= cap ? tiles : std::min(cap, tiles * std::max(slots, 1));
        hipLaunchKernelGGL((rf_mul_x<T, N, R, W>), dim3(unsigned(grid)), dim3(threads), lds, st, H, h_o, h_px, SB, ny, nh,
                           out, out_o, descs, act, tw, gap_ax);
    });
    if (!ok) throw DeviceError { "no such FFT tile" };
}

// P1 with the column-major output of the fused inverse y and z pass (FP32,
// N = 256, radix 16; the tile's width is kept when it is one of 8, 16, 32).
// H: the kernel spectra y-fastest, H[o][kz][x < h_lines][y] (h_o values per
// orientation).  Returns the grid (blocks) of the main launch and sets *tiles
// to its row tiles.
template <typename T>
int rf_mul_xt_launch(Tile tile, hipStream_t st, int cus, int slots, int ny, int nh, const cx<T>* H, long long h_o,
                     int h_lines, const cx<T>* SB, cx<T>* out, long long out_o, const OrientDesc* descs, const int* act,
                     const cx<T>* tw, T sb_scale, long long* tiles)
{
    constexpr int R = 16;
    int used = 0;
    auto go = [&](auto ww) {
        constexpr int W = decltype(ww)::value;
        const std::size_t lds = xt_lds_bytes<T>(W);
        // the output packs kz = 0 and kz = n/2 into plane 0 (its own launch, one
        // block per row tile and orientation); the main pass has the planes
        // kz = 1 .. n/2 - 1
        const int ytiles = (ny + W - 1) / W;
        hipLaunchKernelGGL((rf_mul_xt_plane0<T, W>), dim3(ytiles, slots), dim3(W * R), lds, st, H, h_o, h_lines, SB, ny,
                           nh, out, out_o, descs, act, tw, sb_scale);
        *tiles = (long long)ytiles * (nh - 2);
        const int grid = mul_grid(reinterpret_cast<const void*>(&rf_mul_xt<T, W>), W * R, lds, cus, *tiles, slots);
        used = grid;
        hipLaunchKernelGGL((rf_mul_xt<T, W>), dim3(grid), dim3(W * R), lds, st, H, h_o, h_lines, SB, ny, nh, out, out_o,
                           descs, act, tw, sb_scale);
    };
    if (tile.r == R && tile.w == 8) go(std::integral_constant<int, 8> {});
    else if (tile.r == R && tile.w == 32) go(std::integral_constant<int, 32> {});
    else go(std::integral_constant<int, 16> {});
    return used;
}

template <typename T, typename In>
void rf_r2c_launch(int n, hipStream_t st, int cols_max, int rows_max, int batches, const In* in, int in_px, int in_py,
                   int ex, int ey, int ez, cx<T>* out, long long out_bs, int out_px, int out_py, int cx_, int cy_,
                   const OrientDesc* descs, const int* act, const cx<T>* tw)
{
    if (!rf_supported(n)) throw DeviceError { "register FFT length" };
    if (n == 128) {
        constexpr int N = 128, R = 16, W = 64;
        hipLaunchKernelGGL((rf_r2c_z<T, N, R, W, In>), dim3((cols_max + W - 1) / W, rows_max, batches),
                           dim3((W / 2) * (N / R)), (rf_lds<T, N, R>(W / 2)), st, in, in_px, in_py, ex, ey, ez, out,
                           out_bs, out_px, out_py, cx_, cy_, descs, act, tw);
    }
    else if (n == 256) {
        constexpr int N = 256, R = 16, W = 32;
        hipLaunchKernelGGL((rf_r2c_z<T, N, R, W, In>), dim3((cols_max + W - 1) / W, rows_max, batches),
                           dim3((W / 2) * (N / R)), (rf_lds<T, N, R>(W / 2)), st, in, in_px, in_py, ex, ey, ez, out,
                           out_bs, out_px, out_py, cx_, cy_, descs, act, tw);
    }
    else {
        constexpr int N = 512, R = 8, W = 16;
        hipLaunchKernelGGL((rf_r2c_z<T, N, R, W, In>), dim3((cols_max + W - 1) / W, rows_max, batches),
                           dim3((W / 2) * (N / R)), (rf_lds<T, N, R>(W / 2)), st, in, in_px, in_py, ex, ey, ez, out,
                           out_bs, out_px, out_py, cx_, cy_, descs, act, tw);
    }
}

// Update windows whose z axis is this long apply their counts in the C2R
// pass (the 128-point windows that lengths 90 .. 128 round up to); others
// write a window field that k_apply_update adds.
constexpr int kApplyN = 128;

template <typename T>
void rf_c2r_launch(int n, Tile tile, hipStream_t st, int lx_max, int ly_max, int orients, const cx<T>* in, long long in_o,
                   int nx, int ny, const OrientDesc* descs, T inv_total, StepDeviceResult* res, int* fields,
                   const long long* offsets, const unsigned long long* excl, double* raw, StepFilter filt,
                   const cx<T>* tw, const FieldUpdate* fu = nullptr)
{
    const bool ok = dispatch_tile(C2rTiles {}, n, tile, [&](auto tt) {
        using TT = decltype(tt);
        constexpr int N = TT::n, R = TT::r, W = TT::w;
        const std::size_t lds = std::max(rf_lds<T, N, R>(W / 2), sizeof(cx<T>) * (N + (N / 2 + 1) * (W + 1)));
        if (fu) {
            if constexpr (N == kApplyN)
                hipLaunchKernelGGL((rf_c2r_z_argmin<T, N, R, W, false, true>), dim3((lx_max + W - 1) / W, ly_max, orients),
                                   dim3((W / 2) * (N / R)), lds, st, in, in_o, nx, ny, descs, inv_total, res, nullptr,
                                   offsets, excl, nullptr, filt, tw, *fu);
            else
                throw DeviceError { "update window length" };
        }
        else if (fields || raw)
            hipLaunchKernelGGL((rf_c2r_z_argmin<T, N, R, W, true>), dim3((lx_max + W - 1) / W, ly_max, orients),
                               dim3((W / 2) * (N / R)), lds, st, in, in_o, nx, ny, descs, inv_total, res, fields,
                               offsets, excl, raw, filt, tw);
        else
            hipLaunchKernelGGL((rf_c2r_z_argmin<T, N, R, W, false>), dim3((lx_max + W - 1) / W, ly_max, orients),
                               dim3((W / 2) * (N / R)), lds, st, in, in_o, nx, ny, descs, inv_total, res, fields,
                               offsets, excl, raw, filt, tw);
    });
    if (!ok) throw DeviceError { "no such FFT tile" };
}

}  // namespace

class Session {
public:
    Session(Context& ctx, const sp_session_desc& d, hipStream_t shared_stream = nullptr) : ctx_(ctx), desc_(d)
    {
        for (int a = 0; a != 3; ++a) {
            ext_[a] = int(d.environment_shape[a]);
            efirst_[a] = d.environment_first[a];
            const int g = good_length(ext_[a]);
            if (g == 0) throw DeviceError { "FFT_LENGTH" };
            n_[a] = g;
        }
        O_ = int(d.orientations);
        kernels_.resize(O_);
        descs_.resize(O_);
        use_device(ctx_);
        if (shared_stream) stream_ = shared_stream;
        else {
            SP_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
            owns_stream_ = true;
        }
        int lds = 0;
        SP_HIP_CHECK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx_.device));
        max_lds_ = std::size_t(lds);
        SP_HIP_CHECK(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, ctx_.device));
        env_.allocate(std::size_t(ext_[0]) * ext_[1] * ext_[2]);
        SP_HIP_CHECK(hipMemsetAsync(env_.get(), 0, env_.bytes(), stream_));
        result_.allocate(sizeof(StepDeviceResult));
        StepDeviceResult r0 {};
        r0.key = ~0ULL;
        SP_HIP_CHECK(hipMemcpyAsync(result_.get(), &r0, sizeof r0, hipMemcpyHostToDevice, stream_));
        SP_HIP_CHECK(hipHostMalloc(&host_result_, sizeof(StepDeviceResult), hipHostMallocDefault));
        excl_.allocate(sizeof(unsigned long long) * (1 + kMaxExcluded));
        SP_HIP_CHECK(hipMemsetAsync(excl_.get(), 0, sizeof(unsigned long long), stream_));
        // sessions of the C ABI track their placed copies; the update window
        // and the proximity session (which share a stream) never stamp
        tracks_copies_ = shared_stream == nullptr;
        if (tracks_copies_) {
            mask_.allocate(env_.bytes());
            SP_HIP_CHECK(hipMemsetAsync(mask_.get(), 0, mask_.bytes(), stream_));
            counts_.allocate(env_.bytes() * sizeof(unsigned int));
            SP_HIP_CHECK(hipMemsetAsync(counts_.get(), 0, counts_.bytes(), stream_));
            copylog_.allocate(sizeof(unsigned long long) * (1 + kCopyLogCap));
            SP_HIP_CHECK(hipMemsetAsync(copylog_.get(), 0, sizeof(unsigned long long), stream_));
            fault_.allocate(sizeof(unsigned int));
        }
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    CopyTrack track() const
    {
        if (!tracks_copies_) return CopyTrack { nullptr, nullptr, 0u };
        return CopyTrack { counts_.as<unsigned int>(), copylog_.as<unsigned long long>(), kCopyLogCap };
    }
    ~Session()
    {
        drop_graphs();
        if (host_result_) (void)hipHostFree(host_result_);
        upd_.reset();
        if (stream_ && owns_stream_) (void)hipStreamDestroy(stream_);
    }

    void set_environment(const std::uint8_t* cells, std::uint64_t count)
    {
        if (count != std::uint64_t(ext_[0]) * ext_[1] * ext_[2]) throw std::invalid_argument("environment size");
        for (std::uint64_t i = 0; i != count; ++i)
            if (cells[i] > 1) throw std::domain_error("CORRELATION_INVALID_ENVIRONMENT");
        use_device(ctx_);
        SP_HIP_CHECK(hipMemcpyAsync(env_.get(), cells, count, hipMemcpyHostToDevice, stream_));
        if (tracks_copies_) {
            // the given occupancy is the fixed mask; placed copies start over
            SP_HIP_CHECK(hipMemcpyAsync(mask_.get(), env_.get(), count, hipMemcpyDeviceToDevice, stream_));
            SP_HIP_CHECK(hipMemsetAsync(counts_.get(), 0, counts_.bytes(), stream_));
            SP_HIP_CHECK(hipMemsetAsync(copylog_.get(), 0, sizeof(unsigned long long), stream_));
            removed_.clear();
            frozen_copies_ = 0;
        }
        std::uint64_t mass = 0;
        for (std::uint64_t i = 0; i != count; ++i) mass += cells[i];
        StepDeviceResult r0 {};
        r0.key = ~0ULL;
        r0.env_mass = mass;
        SP_HIP_CHECK(hipMemcpyAsync(result_.get(), &r0, sizeof r0, hipMemcpyHostToDevice, stream_));
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
        inc_ready_ = pending_ = false;
        ++env_rev_;
        ++reset_rev_;
    }

    void get_environment(std::uint8_t* cells, std::uint64_t count)
    {
        if (count != std::uint64_t(ext_[0]) * ext_[1] * ext_[2]) throw std::invalid_argument("environment size");
        use_device(ctx_);
        SP_HIP_CHECK(hipMemcpyAsync(cells, env_.get(), count, hipMemcpyDeviceToHost, stream_));
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    void set_kernel(std::uint32_t o, const std::uint8_t* cells, const std::uint32_t shape[3], const std::int64_t first[3])
    {
        if (int(o) >= O_) throw std::invalid_argument("orientation index");
        Kernel& k = kernels_[o];
        std::uint64_t n = 1;
        for (int a = 0; a != 3; ++a) {
            if (shape[a] == 0) throw std::invalid_argument("kernel shape");
            k.shape[a] = int(shape[a]);
            k.first[a] = first[a];
            n *= shape[a];
        }
        k.cells.assign(cells, cells + n);
        k.mass = 0;
        for (auto c : k.cells) {
            if (c > 1) throw std::domain_error("CORRELATION_INVALID_KERNEL");
            k.mass += c;
        }
        k.set = true;
        prepared_ = false;
    }

    // Selects the register-FFT passes for power-of-R lengths (default) or the
    // generic LDS Stockham passes everywhere (for cross-checking the two).
    void set_fast_path(bool on)
    {
        // Switching the passes of a set-up incremental update window took the
        // GPU to an illegal memory access whose cause is not found yet; it is
        // refused until it is.
        if (upd_ && on != use_rf_)
            throw std::invalid_argument("set_fast_path must precede the first incremental step");
        if (on != use_rf_) {
            use_rf_ = on;
            spectra_ready_ = false;
            tuned_ = false;
            drop_graphs();
        }
        if (upd_) {
            upd_->set_fast_path(on);  // the incremental window's passes too
            upd_->keep_fields(!upd_->can_apply());  // its own fields only where k_apply_update reads them
        }
    }

    // Restricts the next steps to a list of orientations and, optionally, to
    // a mask over regions of the environment.  No list and no mask: every
    // orientation everywhere (the plain search).  Orientations whose kernel
    // does not fit are dropped from the list.
    void set_active(const std::uint32_t* orients, std::uint32_t count, const std::uint8_t* mask,
                    const std::uint32_t regions[3])
    {
        prepare();
        if (count == 0 && !mask) {
            subset_ = false;
            return;
        }
        if (incremental_) throw std::invalid_argument("incremental steps search every orientation");
        std::vector<int> list;
        std::vector<char> seen(O_, 0);
        if (count) {
            for (std::uint32_t i = 0; i != count; ++i) {
                if (orients[i] >= std::uint32_t(O_)) throw std::invalid_argument("orientation index");
                if (seen[orients[i]]) throw std::invalid_argument("orientation listed twice");
                seen[orients[i]] = 1;
                if (descs_[orients[i]].lx > 0) list.push_back(int(orients[i]));
            }
        }
        else
            for (int o = 0; o != O_; ++o)
                if (descs_[o].lx > 0) list.push_back(o);
        int r[3] = { 1, 1, 1 };
        if (mask) {
            for (int a = 0; a != 3; ++a) {
                if (!regions || regions[a] == 0 || regions[a] > std::uint32_t(kMaxRegionsPerAxis))
                    throw std::invalid_argument("regions per axis");
                r[a] = int(regions[a]);
            }
        }
        std::vector<int> head(kActHeader + list.size());
        head[0] = int(list.size());
        head[1] = mask ? 1 : 0;
        head[2] = r[0];
        head[3] = r[1];
        head[4] = r[2];
        std::copy(list.begin(), list.end(), head.begin() + kActHeader);
        use_device(ctx_);
        SP_HIP_CHECK(hipMemcpyAsync(act_sub_.get(), head.data(), head.size() * sizeof(int), hipMemcpyHostToDevice, stream_));
        if (mask)
            SP_HIP_CHECK(hipMemcpyAsync(region_.get(), mask, std::size_t(r[0]) * r[1] * r[2], hipMemcpyHostToDevice,
                                        stream_));
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
        subset_ = true;
        sub_slots_ = 8;
        while (sub_slots_ < int(list.size())) sub_slots_ *= 2;
        sub_slots_ = std::max(1, std::min(sub_slots_, std::max(nfit_, 1)));
    }

    // Lowest feasible key of each orientation in the last step (~0: none).
    void orientation_minima(std::uint64_t* out, std::uint32_t count)
    {
        if (!prepared_) throw std::invalid_argument("no step run");
        if (count < std::uint32_t(O_)) throw std::length_error("capacity");
        use_device(ctx_);
        SP_HIP_CHECK(hipMemcpyAsync(out, orient_best_.get(), sizeof(std::uint64_t) * O_, hipMemcpyDeviceToHost, stream_));
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    // gap_fill.cpp:517-520: a rejected proposal's translation stays blocked
    // for that orientation.
    void exclude(std::uint32_t o, const std::int64_t t[3])
    {
        if (int(o) >= O_ || !kernels_[o].set) throw std::invalid_argument("orientation index");
        unsigned long long key = o;
        const int shift[3] = { 16, 32, 48 };
        for (int a = 0; a != 3; ++a) {
            const std::int64_t u = t[a] + kernels_[o].first[a] - efirst_[a];
            if (u < 0 || u > 0xffff) throw std::invalid_argument("translation outside the window");
            key |= (unsigned long long)u << shift[a];
        }
        auto it = std::lower_bound(excluded_.begin(), excluded_.end(), key);
        if (it != excluded_.end() && *it == key) return;
        if (excluded_.size() >= kMaxExcluded) throw std::length_error("too many rejected translations");
        excluded_.insert(it, key);
        ++reset_rev_;
        std::vector<unsigned long long> buf(1 + excluded_.size());
        buf[0] = excluded_.size();
        std::copy(excluded_.begin(), excluded_.end(), buf.begin() + 1);
        use_device(ctx_);
        SP_HIP_CHECK(hipMemcpyAsync(excl_.get(), buf.data(), buf.size() * sizeof(unsigned long long),
                                    hipMemcpyHostToDevice, stream_));
        if (incremental_ && inc_ready_ && inc_effective_) {
            if (pending_) apply_pending();
            hipLaunchKernelGGL(k_block_corner, dim3(1), dim3(64), 0, stream_, fields_.as<int>(), d_descs_.as<OrientDesc>(),
                               d_offsets_.as<long long>(), int(o), int((key >> 16) & 0xffff), int((key >> 32) & 0xffff),
                               int(key >> 48), orient_best_.as<unsigned long long>(), rescan_.as<int>());
            launch_rescan();
            SP_HIP_CHECK(hipGetLastError());
        }
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    // Stamp kernel o at translation t (outside the step graph).
    void stamp_at(std::uint32_t o, const std::int64_t t[3])
    {
        prepare();
        if (int(o) >= O_ || descs_[o].lx <= 0) throw std::invalid_argument("orientation index");
        unsigned long long key = o;
        const int shift[3] = { 16, 32, 48 };
        for (int a = 0; a != 3; ++a) {
            const std::int64_t u = t[a] + kernels_[o].first[a] - efirst_[a];
            if (u < 0 || u > 0xffff) throw std::invalid_argument("translation outside the window");
            key |= (unsigned long long)u << shift[a];
        }
        use_device(ctx_);
        StepDeviceResult* res = result_.as<StepDeviceResult>();
        const bool inc = incremental_ && inc_ready_ && inc_effective_;
        if (inc && pending_) apply_pending();
        ++env_rev_;
        hipLaunchKernelGGL(k_set_key, dim3(1), dim3(64), 0, stream_, res, key);
        if (inc) enqueue_collecting_stamp();
        else enqueue_stamp(nullptr, 0, 0, nullptr);
        SP_HIP_CHECK(hipGetLastError());
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    // The direct recheck of the chosen corner (not for the circular update
    // window or a real-valued proximity session, whose result key is unused).
    void enqueue_verify()
    {
        if (circular_ || real_env_) return;
        hipLaunchKernelGGL(k_verify_choice, dim3(kSmallGrid), dim3(kThreads), 0, stream_, env_.as<unsigned char>(), ext_[0],
                           ext_[1], ext_[2], kcells_.as<unsigned char>(), d_descs_.as<OrientDesc>(),
                           result_.as<StepDeviceResult>());
        mark("verify");
    }

    // ---- placed copies (conservative_fields.cpp:1666-1887) -----------------
    std::uint64_t copy_count()
    {
        if (!tracks_copies_) return 0;
        unsigned long long n = 0;
        use_device(ctx_);
        SP_HIP_CHECK(hipMemcpyAsync(&n, copylog_.get(), sizeof n, hipMemcpyDeviceToHost, stream_));
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
        return n;
    }

    struct CopyInfo {
        unsigned long long key;
        bool removed;
    };

    // Every copy stamped since the environment was set, in stamp order
    // (copy id = position); copies beyond the log's capacity are counted
    // but not listed.
    std::vector<CopyInfo> copies()
    {
        const std::uint64_t n = copy_count();
        const std::uint64_t kept = std::min<std::uint64_t>(n, kCopyLogCap);
        std::vector<unsigned long long> keys(kept);
        if (kept) {
            SP_HIP_CHECK(hipMemcpyAsync(keys.data(), copylog_.as<unsigned long long>() + 1, kept * sizeof(unsigned long long),
                                        hipMemcpyDeviceToHost, stream_));
            SP_HIP_CHECK(hipStreamSynchronize(stream_));
        }
        std::vector<CopyInfo> out(kept);
        for (std::uint64_t i = 0; i != kept; ++i) out[i] = CopyInfo { keys[i], removed_.count(i) != 0 };
        return out;
    }

    void key_pose(unsigned long long key, std::uint32_t& o, std::int64_t t[3]) const
    {
        o = std::uint32_t(key & 0xffffULL);
        const std::int64_t u[3] = { std::int64_t((key >> 16) & 0xffffULL), std::int64_t((key >> 32) & 0xffffULL),
                                     std::int64_t(key >> 48) };
        for (int a = 0; a != 3; ++a) t[a] = efirst_[a] + u[a] - kernels_[o].first[a];
    }

    // BlockedField::remove (conservative_fields.cpp:1848-1887): the copy's
    // placed counts are subtracted; cells that no copy and no mask block any
    // longer become free.  Incremental fields subtract the freed cells'
    // correlation, and every orientation's minimum is searched again.
    void remove_copy(std::uint64_t id)
    {
        if (!tracks_copies_) throw std::invalid_argument("FIELD_COPY_ID");
        const std::uint64_t n = copy_count();
        if (id >= n || id >= kCopyLogCap || removed_.count(id)) throw std::invalid_argument("FIELD_COPY_ID");
        prepare();
        if (id < frozen_copies_) throw std::invalid_argument("FIELD_COPY_ID");  // stamped under other stamp kernels
        use_device(ctx_);
        unsigned long long key = 0;
        SP_HIP_CHECK(hipMemcpyAsync(&key, copylog_.as<unsigned long long>() + 1 + id, sizeof key, hipMemcpyDeviceToHost,
                                    stream_));
        SP_HIP_CHECK(hipMemsetAsync(fault_.get(), 0, sizeof(unsigned int), stream_));
        const bool inc = incremental_ && inc_ready_ && inc_effective_;
        if (inc && pending_) apply_pending();
        const bool dil = halo_ > 0;
        const unsigned char* kc = dil ? scells_.as<unsigned char>() : kcells_.as<unsigned char>();
        const OrientDesc* kd = dil ? sdescs_.as<OrientDesc>() : d_descs_.as<OrientDesc>();
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
        if (inc) {
            Session& u = *upd_;
            SP_HIP_CHECK(hipMemsetAsync(u.env_.get(), 0, u.env_.bytes(), stream_));
            StepDeviceResult* ur = u.result_.as<StepDeviceResult>();
            SP_HIP_CHECK(hipMemsetAsync(&ur->env_mass, 0, sizeof(ur->env_mass), stream_));
            hipLaunchKernelGGL(k_unstamp, dim3(64), dim3(kThreads), 0, stream_, env_.as<unsigned char>(),
                               mask_.as<unsigned char>(), ext_[0], ext_[1], ext_[2], kc, kd, key,
                               counts_.as<unsigned int>(), result_.as<StepDeviceResult>(), fault_.as<unsigned int>(),
                               u.env_.as<unsigned char>(), u.ext_[0], u.ext_[1], &ur->env_mass, halo_);
            // the removed copy's key as the update window's origin
            rm_key_.ensure(sizeof(unsigned long long));
            SP_HIP_CHECK(hipMemcpyAsync(rm_key_.get(), &key, sizeof key, hipMemcpyHostToDevice, stream_));
            apply_window(rm_key_.as<unsigned long long>(), -1);
            hipLaunchKernelGGL(k_rescan_all, dim3(1), dim3(256), 0, stream_, orient_best_.as<unsigned long long>(),
                               rescan_.as<int>(), O_);
            launch_rescan();
        }
        else
            hipLaunchKernelGGL(k_unstamp, dim3(64), dim3(kThreads), 0, stream_, env_.as<unsigned char>(),
                               mask_.as<unsigned char>(), ext_[0], ext_[1], ext_[2], kc, kd, key,
                               counts_.as<unsigned int>(), result_.as<StepDeviceResult>(), fault_.as<unsigned int>(),
                               nullptr, 0, 0, nullptr, halo_);
        SP_HIP_CHECK(hipGetLastError());
        unsigned int fault = 0;
        SP_HIP_CHECK(hipMemcpyAsync(&fault, fault_.get(), sizeof fault, hipMemcpyDeviceToHost, stream_));
        StepDeviceResult ur {};
        if (inc) SP_HIP_CHECK(hipMemcpyAsync(&ur, upd_->result_.get(), sizeof ur, hipMemcpyDeviceToHost, stream_));
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
        removed_.insert(id);
        ++env_rev_;
        ++reset_rev_;
        if (fault) throw std::domain_error("FIELD_COUNT_UNDERFLOW");
        if (ur.flags) throw std::domain_error("CORRELATION_NUMERIC");
    }

    void placed_counts(std::uint32_t* out, std::uint64_t count)
    {
        if (count != std::uint64_t(ext_[0]) * ext_[1] * ext_[2]) throw std::invalid_argument("environment size");
        if (!tracks_copies_) throw std::invalid_argument("no placed copies");
        use_device(ctx_);
        SP_HIP_CHECK(hipMemcpyAsync(out, counts_.get(), count * sizeof(std::uint32_t), hipMemcpyDeviceToHost, stream_));
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    // Stamps the result's key: the kernel, or with a pair clearance its
    // dilated stamp kernel at the corner minus the halo; optionally collects
    // the newly set cells (at stamp-kernel coordinates) and counts them.
    void enqueue_stamp(unsigned char* collect, int cx, int cy, unsigned long long* cmass)
    {
        const bool dil = halo_ > 0;
        hipLaunchKernelGGL(k_stamp, dim3(kSmallGrid), dim3(kThreads), 0, stream_, env_.as<unsigned char>(), ext_[0], ext_[1],
                           ext_[2], dil ? scells_.as<unsigned char>() : kcells_.as<unsigned char>(),
                           dil ? sdescs_.as<OrientDesc>() : d_descs_.as<OrientDesc>(), result_.as<StepDeviceResult>(),
                           track(), collect, cx, cy, 0, 0, 0, cmass, halo_);
    }

    // Pair clearance between placed copies (0: none).
    void set_pair_clearance(double clearance, double pitch)
    {
        if (!std::isfinite(clearance) || clearance < 0.0) throw std::invalid_argument("clearance must be finite and >= 0");
        if (!std::isfinite(pitch) || !(pitch > 0.0)) throw std::invalid_argument("pitch must be finite and > 0");
        pair_clearance_ = clearance;
        pitch_ = pitch;
        prepared_ = false;
    }

    // Dilated stamp kernels for the pair clearance (after the kernels are uploaded).
    void prepare_stamp_kernels()
    {
        halo_ = 0;
        if (pair_clearance_ == 0.0) return;
        const long double r = std::ceil(static_cast<long double>(pair_clearance_) / pitch_) + 1.0L;
        if (!(r < 4096.0L)) throw std::invalid_argument("clearance halo too large");
        halo_ = int(r);
        std::vector<int> st;
        for (int z = -halo_; z <= halo_; ++z)
            for (int y = -halo_; y <= halo_; ++y)
                for (int x = -halo_; x <= halo_; ++x) {
                    const int d[3] = { x, y, z };
                    if (offset_may_reach(d, pitch_, pair_clearance_)) {
                        st.push_back(x);
                        st.push_back(y);
                        st.push_back(z);
                    }
                }
        std::vector<OrientDesc> sd(descs_);
        long long off = 0;
        for (int o = 0; o != O_; ++o) {
            sd[o].kx = descs_[o].kx + 2 * halo_;
            sd[o].ky = descs_[o].ky + 2 * halo_;
            sd[o].kz = descs_[o].kz + 2 * halo_;
            sd[o].cells_offset = off;
            off += (long long)sd[o].kx * sd[o].ky * sd[o].kz;
        }
        scells_.allocate(std::size_t(std::max(off, 1LL)));
        sdescs_.allocate(sizeof(OrientDesc) * O_);
        stencil_.allocate(sizeof(int) * std::max<std::size_t>(st.size(), 3));
        SP_HIP_CHECK(hipMemsetAsync(scells_.get(), 0, scells_.bytes(), stream_));
        SP_HIP_CHECK(hipMemcpyAsync(sdescs_.get(), sd.data(), sizeof(OrientDesc) * O_, hipMemcpyHostToDevice, stream_));
        SP_HIP_CHECK(hipMemcpyAsync(stencil_.get(), st.data(), sizeof(int) * st.size(), hipMemcpyHostToDevice, stream_));
        hipLaunchKernelGGL(k_dilate, dim3(64, O_), dim3(kThreads), 0, stream_, kcells_.as<unsigned char>(),
                           d_descs_.as<OrientDesc>(), sdescs_.as<OrientDesc>(), scells_.as<unsigned char>(),
                           stencil_.as<int>(), int(st.size() / 3), halo_);
        SP_HIP_CHECK(hipGetLastError());
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    // ---- incremental mode ------------------------------------------------
    // Incremental steps keep the fields unless the update window turned out
    // to have no benefit (then steps stay full and need none).
    bool fields_needed() const { return keep_fields_ || (incremental_ && (!inc_setup_done_ || inc_effective_)); }

    void set_incremental(bool on)
    {
        if (on == incremental_) return;
        if (on && subset_) throw std::invalid_argument("incremental steps search every orientation");
        incremental_ = on;
        prepared_ = false;  // fields buffer and incremental state
    }

    // Update session: a circular window of n' = good length >= 2 Kmax - 1
    // cells per axis holding the newly set cells at offsets [0, Kp), whose
    // fields are every circular shift; used when n'^3 is at most half the
    // main FFT grid.
    void inc_setup()
    {
        inc_setup_done_ = true;
        inc_effective_ = false;
        if (nfit_ == 0) return;
        sp_session_desc d {};
        double vol_u = 1.0, vol_m = 1.0;
        for (int a = 0; a != 3; ++a) {
            margin_[a] = 0;
            const int need = 2 * kmax_[a] + 2 * halo_ - 1;
            int g = good_length(need);
            if (g == 0) return;
            // from 90 points (at most 2.9x the points) the 128-point register
            // transforms beat the LDS Stockham passes at the shorter length:
            // about 3x their speed per point, measured at 96 points
            if (use_rf_ && g >= 90 && need <= 128) g = 128;
            vol_u *= g;
            vol_m *= n_[a];
            d.environment_shape[a] = std::uint32_t(g);
            d.environment_first[a] = 0;
        }
        if (2.0 * vol_u > vol_m) return;
        d.orientations = std::uint32_t(O_);
        d.precision = precision_;
        d.capture_graph = desc_.capture_graph;
        upd_ = std::make_unique<Session>(ctx_, d, stream_);
        upd_->circular_ = true;
        upd_->set_fast_path(use_rf_);
        for (int o = 0; o != O_; ++o) {
            const Kernel& k = kernels_[o];
            const std::uint32_t sh[3] = { std::uint32_t(k.shape[0]), std::uint32_t(k.shape[1]), std::uint32_t(k.shape[2]) };
            const std::int64_t zero[3] = { 0, 0, 0 };
            upd_->set_kernel(std::uint32_t(o), k.cells.data(), sh, zero);
        }
        // a window whose C2R pass applies its counts writes no field of its own
        upd_->keep_fields(!upd_->can_apply());
        upd_->prepare();
        rescan_.allocate(sizeof(int) * O_);
        SP_HIP_CHECK(hipMemsetAsync(rescan_.get(), 0, sizeof(int) * O_, stream_));
        rfound_.allocate(sizeof(unsigned long long) * O_);
        SP_HIP_CHECK(hipMemsetAsync(rfound_.get(), 0xff, sizeof(unsigned long long) * O_, stream_));
        inc_effective_ = true;
    }

    // Stamps the result's key and collects the newly set cells into the
    // update session's (cleared) occupancy, counting them as its mass.
    void enqueue_collecting_stamp()
    {
        Session& u = *upd_;
        SP_HIP_CHECK(hipMemsetAsync(u.env_.get(), 0, u.env_.bytes(), stream_));
        StepDeviceResult* ur = u.result_.as<StepDeviceResult>();
        SP_HIP_CHECK(hipMemsetAsync(&ur->env_mass, 0, sizeof(ur->env_mass), stream_));
        enqueue_stamp(u.env_.as<unsigned char>(), u.ext_[0], u.ext_[1], &ur->env_mass);
        pending_ = true;
    }

    // Adds the correlation of the pending new cells with every kernel into
    // the fields and rescans orientations whose minimum was blocked.
    void apply_pending()
    {
        apply_window(&result_.as<StepDeviceResult>()->key, 1);
        launch_rescan();
        SP_HIP_CHECK(hipGetLastError());
        pending_ = false;
    }

    // Adds sign times the update window's correlation (its occupancy holds
    // the changed cells, placed at `origin`) into the fields: inside the
    // window's C2R pass where that pass is the register one, else by
    // k_apply_update from the window's own fields.
    void apply_window(const unsigned long long* origin, int sign)
    {
        Session& u = *upd_;
        const int ax = kmax_[0] + 2 * halo_, ay = kmax_[1] + 2 * halo_, az = kmax_[2] + 2 * halo_;
        if (u.can_apply()) {
            const FieldUpdate fu { fields_.as<int>(), d_descs_.as<OrientDesc>(), d_offsets_.as<long long>(), origin,
                                   orient_best_.as<unsigned long long>(), rescan_.as<int>(), ax, ay, az, halo_, sign };
            u.launch_apply(fu, sign > 0 ? 0 : 1);
            return;
        }
        u.launch_mode(false, false);
        hipLaunchKernelGGL(k_apply_update, dim3(64, O_), dim3(kThreads), 0, stream_, fields_.as<int>(),
                           d_descs_.as<OrientDesc>(), d_offsets_.as<long long>(), u.fields_.as<int>(),
                           u.d_descs_.as<OrientDesc>(), u.d_offsets_.as<long long>(), origin, u.n_[0], u.n_[1], u.n_[2],
                           ax, ay, az, orient_best_.as<unsigned long long>(), rescan_.as<int>(), halo_, sign);
    }

    bool can_apply() const { return pass_uses_rf(kPassC2r) && n_[2] == kApplyN; }

    // Window sessions: one step whose C2R pass adds its counts into another
    // session's fields by `fu` (no stamp, no field output).  `variant` keys
    // the cached graph: the pointers in `fu` are fixed per variant until the
    // owner prepares again, which drops this session.
    void launch_apply(const FieldUpdate& fu, int variant)
    {
        prepare();
        use_device(ctx_);
        ensure_kernel_spectra();
        autotune();
        apply_ = fu;
        apply_on_ = true;
        try {
            if (!desc_.capture_graph) {
                enqueue(false, false);
                SP_HIP_CHECK(hipGetLastError());
            }
            else
                run_graph(4 + variant, false, nullptr, nfit_, [&] { enqueue(false, false); });
        }
        catch (...) {
            apply_on_ = false;
            throw;
        }
        apply_on_ = false;
    }

    // Searches again the minima of the orientations flagged in rescan_: over
    // kRescanBlocks blocks per orientation and a finishing kernel for fields
    // of kRescanWide corners or more per orientation, else one block each
    // (the extra blocks and launch cost more than they save on small fields).
    void launch_rescan()
    {
        const bool wide = fields_total_ >= (long long)std::max(O_, 1) * kRescanWide;
        hipLaunchKernelGGL(k_rescan, dim3(O_, wide ? kRescanBlocks : 1), dim3(1024), 0, stream_, fields_.as<int>(),
                           d_descs_.as<OrientDesc>(), d_offsets_.as<long long>(), orient_best_.as<unsigned long long>(),
                           rescan_.as<int>(), rfound_.as<unsigned long long>());
        if (wide)
            hipLaunchKernelGGL(k_rescan_finish, dim3((O_ + 255) / 256), dim3(256), 0, stream_,
                               orient_best_.as<unsigned long long>(), rescan_.as<int>(), rfound_.as<unsigned long long>(),
                               O_);
    }

    void launch_incremental(bool stamp)
    {
        if (stamp) ++env_rev_;
        prepare();
        use_device(ctx_);
        if (!inc_setup_done_) inc_setup();
        if (!inc_effective_) {
            launch_mode(stamp, false);
            return;
        }
        bool fresh = false;
        if (!inc_ready_) {
            // one full step gives the exact fields and every orientation's minimum
            launch_mode(false, false);
            fresh = true;
            for (unsigned long long k : excluded_) {
                const int o = int(k & 0xffff);
                hipLaunchKernelGGL(k_block_corner, dim3(1), dim3(64), 0, stream_, fields_.as<int>(),
                                   d_descs_.as<OrientDesc>(), d_offsets_.as<long long>(), o, int((k >> 16) & 0xffff),
                                   int((k >> 32) & 0xffff), int(k >> 48), orient_best_.as<unsigned long long>(),
                                   rescan_.as<int>());
            }
            inc_ready_ = true;
            pending_ = false;
        }
        else if (pending_)
            apply_pending();
        hipLaunchKernelGGL(k_choose, dim3(1), dim3(256), 0, stream_, result_.as<StepDeviceResult>(),
                           orient_best_.as<unsigned long long>(), O_, upd_->result_.as<StepDeviceResult>(),
                           fresh ? 0 : 1);
        if (!fresh) enqueue_verify();
        if (stamp) enqueue_collecting_stamp();
        SP_HIP_CHECK(hipGetLastError());
    }

    void keep_fields(bool keep)
    {
        if (keep != keep_fields_) {
            keep_fields_ = keep;
            prepared_ = false;
        }
    }

    // Unrounded legal values (as double) of the next steps, for measuring the
    // FP32 deviation from FP64 on the same occupancy.
    void keep_raw(bool keep)
    {
        if (keep != keep_raw_) {
            keep_raw_ = keep;
            prepared_ = false;
        }
    }

    void raw_values(std::uint32_t o, double* out, std::uint64_t count)
    {
        if (!keep_raw_ || !prepared_) throw std::invalid_argument("raw values not kept");
        if (int(o) >= O_) throw std::invalid_argument("orientation index");
        const OrientDesc& d = descs_[o];
        const std::uint64_t n = std::uint64_t(std::max(d.lx, 0)) * std::max(d.ly, 0) * std::max(d.lz, 0);
        if (count < n) throw std::length_error("capacity");
        if (n == 0) return;
        use_device(ctx_);
        SP_HIP_CHECK(hipMemcpyAsync(out, raw_.as<double>() + field_offsets_[o], n * sizeof(double), hipMemcpyDeviceToHost,
                                    stream_));
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    // Precision chosen for the session: FP32 unless its a-priori bound, with
    // the environment mass at its worst case (every cell), exceeds 0.4.
    int precision() const { return precision_; }
    double bound() const { return bound_; }

    void prepare()
    {
        if (prepared_) return;
        // copies stamped before the stamp kernels are rebuilt keep footprints
        // the new kernels may not reproduce: they can no longer be removed
        if (tracks_copies_) frozen_copies_ = copy_count();
        for (const auto& k : kernels_)
            if (!k.set) throw std::invalid_argument("kernel not set");
        std::uint64_t max_mass = 0;
        for (const auto& k : kernels_) max_mass = std::max(max_mass, k.mass);
        const std::uint64_t cells = std::uint64_t(n_[0]) * n_[1] * n_[2];
        bound_ = fp32_error_bound(cells, cells, max_mass);
        precision_ = desc_.precision == SP_PRECISION_FP64 ? 64 : desc_.precision == SP_PRECISION_FP32 ? 32
                     : (bound_ > 0.4 ? 64 : 32);
        // descriptors and concatenated kernels
        std::vector<std::uint8_t> all;
        kmax_ = { 1, 1, 1 };
        lmax_ = { 0, 0, 0 };
        field_offsets_.assign(O_, 0);
        long long foff = 0;
        for (int o = 0; o != O_; ++o) {
            const auto& k = kernels_[o];
            OrientDesc& d = descs_[o];
            d.kx = k.shape[0];
            d.ky = k.shape[1];
            d.kz = k.shape[2];
            d.lx = ext_[0] - k.shape[0] + 1;
            d.ly = ext_[1] - k.shape[1] + 1;
            d.lz = ext_[2] - k.shape[2] + 1;
            if (d.lx <= 0 || d.ly <= 0 || d.lz <= 0) d.lx = d.ly = d.lz = 0;  // does not fit (gap_fill.cpp:382-390)
            else if (circular_) {
                // every circular shift of the FFT grid (window = grid)
                d.lx = n_[0];
                d.ly = n_[1];
                d.lz = n_[2];
            }
            d.cells_offset = (long long)all.size();
            d.mass = (long long)k.mass;
            all.insert(all.end(), k.cells.begin(), k.cells.end());
            if (d.lx > 0)
                for (int a = 0; a != 3; ++a) kmax_[a] = std::max(kmax_[a], k.shape[a]);
            lmax_[0] = std::max(lmax_[0], d.lx);
            lmax_[1] = std::max(lmax_[1], d.ly);
            lmax_[2] = std::max(lmax_[2], d.lz);
            field_offsets_[o] = foff;
            foff += (long long)d.lx * d.ly * d.lz;
        }
        if (kmax_[0] > n_[0] || kmax_[1] > n_[1] || kmax_[2] > n_[2]) {
            // pruned kernel buffers assume the kernel fits the FFT grid
            for (int a = 0; a != 3; ++a) kmax_[a] = std::min(kmax_[a], n_[a]);
        }
        use_device(ctx_);
        kcells_.allocate(std::max<std::size_t>(all.size(), 1));
        SP_HIP_CHECK(hipMemcpyAsync(kcells_.get(), all.data(), all.size(), hipMemcpyHostToDevice, stream_));
        d_descs_.allocate(sizeof(OrientDesc) * O_);
        SP_HIP_CHECK(hipMemcpyAsync(d_descs_.get(), descs_.data(), sizeof(OrientDesc) * O_, hipMemcpyHostToDevice, stream_));
        prepare_stamp_kernels();
        // active lists: [count, regions on, regions x, y, z, orientation per slot]
        std::vector<int> fit { 0, 0, 1, 1, 1 };
        for (int o = 0; o != O_; ++o)
            if (descs_[o].lx > 0) fit.push_back(o);
        nfit_ = int(fit.size()) - kActHeader;
        fit[0] = nfit_;
        act_full_.allocate(sizeof(int) * fit.size());
        SP_HIP_CHECK(hipMemcpyAsync(act_full_.get(), fit.data(), sizeof(int) * fit.size(), hipMemcpyHostToDevice, stream_));
        act_sub_.allocate(sizeof(int) * (kActHeader + O_));
        SP_HIP_CHECK(hipMemcpyAsync(act_sub_.get(), fit.data(), sizeof(int) * fit.size(), hipMemcpyHostToDevice, stream_));
        subset_ = false;
        sub_slots_ = std::max(nfit_, 1);
        if (!region_.get()) region_.allocate(std::size_t(kMaxRegionsPerAxis) * kMaxRegionsPerAxis * kMaxRegionsPerAxis);
        orient_best_.allocate(sizeof(unsigned long long) * O_);
        d_offsets_.allocate(sizeof(long long) * O_);
        SP_HIP_CHECK(hipMemcpyAsync(d_offsets_.get(), field_offsets_.data(), sizeof(long long) * O_,
                                    hipMemcpyHostToDevice, stream_));
        if (fields_needed()) fields_.allocate(std::max<long long>(foff, 1) * sizeof(int));
        else fields_.release();
        if (keep_raw_) raw_.allocate(std::max<long long>(foff, 1) * sizeof(double));
        else raw_.release();
        fields_total_ = foff;
        const std::size_t elem = precision_ == 64 ? sizeof(double2) : sizeof(float2);
        const std::size_t nh = std::size_t(n_[2] / 2 + 1);
        sb_.allocate(elem * nh * n_[1] * n_[0]);
        h1_.allocate(elem * O_ * nh * kmax_[1] * kmax_[0]);
        // H2 rows start on 128-byte lines (the product pass reads one per
        // orientation and spectral row)
        const int per_line = int(128 / elem);
        h2p_ = (kmax_[0] + per_line - 1) / per_line * per_line;
        h2_.allocate(elem * O_ * nh * n_[1] * h2p_);
        t_.allocate(elem * O_ * nh * n_[1] * n_[0]);
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
        for (int a = 0; a != 3; ++a) {
            if (!make_plan(n_[a], plans_[a])) throw DeviceError { "FFT_LENGTH" };
            // twiddle tables are allocated here, never inside a graph capture
            if (precision_ == 64) (void)twiddles<double>(n_[a]);
            else (void)twiddles<float>(n_[a]);
        }
        drop_graphs();
        spectra_ready_ = false;
        tuned_ = false;
        upd_.reset();
        prox_.reset();
        inc_setup_done_ = inc_ready_ = pending_ = false;
        for (int p = 0; p != kPasses; ++p)
            if (!pinned_[p]) {
                tile_[p] = Tile {};
                tile_ms_[p] = 0.f;
            }
        prepared_ = true;
        ++reset_rev_;
    }

    // Optional per-kernel timing: when prof_ is set, an event is recorded after
    // every launch (outside graph capture only).
    void mark(const char* name)
    {
        if (!prof_) return;
        hipEvent_t e;
        SP_HIP_CHECK(hipEventCreate(&e));
        SP_HIP_CHECK(hipEventRecord(e, stream_));
        prof_events_.push_back(e);
        prof_names_.push_back(name);
    }

    // G1 + G2: each listed orientation's kernel spectrum up to its y pass (H2).
    // It does not depend on the occupancy; with resident kernel spectra it
    // runs once per preparation instead of in every step.
    template <typename T> void enqueue_kernel_spectra(const int* act, int slots)
    {
        using T2 = cx<T>;
        const int ny = n_[1], nz = n_[2];
        const int nh = nz / 2 + 1;
        const T2* twy = twiddles<T>(ny);
        const T2* twz = twiddles<T>(nz);
        T2* H1 = h1_.as<T2>();
        T2* H2 = h2_.as<T2>();
        const OrientDesc* dd = d_descs_.as<OrientDesc>();
        const std::size_t e = sizeof(T2);
        const long long h1o = (long long)nh * kmax_[1] * kmax_[0];
        const long long h2o = (long long)nh * ny * h2p_;
        // G1: kernel R2C along z over its own columns (x < kx, y < ky)
        if (rf_supported(nz) && use_rf_) {
            rf_r2c_launch<T, unsigned char>(nz, stream_, kmax_[0], kmax_[1], slots, kcells_.as<unsigned char>(), 0, 0, 0,
                                            0, 0, H1, h1o, kmax_[0], kmax_[1], 0, 0, dd, act, twz);
        }
        else {
            const int W = choose_w(nz, e, 2, true, 32, 2);
            hipLaunchKernelGGL((k_r2c_z<T, unsigned char>), dim3((kmax_[0] + W - 1) / W, kmax_[1], slots),
                               dim3(kThreads), lds_bytes<T>(nz, 2, W / 2), stream_, kcells_.as<unsigned char>(), 0LL, 0,
                               0, 0, 0, 0, H1, h1o, kmax_[0], kmax_[1], 0, 0, dd, act, twz, plans_[2], W);
        }
        mark("kernel_r2c_z");
        // G2: kernel C2C along y for x < kx, reading its ky nonzero rows.  The
        // fused path's product pass reads the result y-fastest,
        // H2[o][kz][x < h2p][y] (the same size per orientation).
        const bool yfast = fused_yz();
        if (rf_supported(ny) && use_rf_) {
            rf_cols_launch<T>(ny, default_tile(kPassInvY, ny), stream_, kmax_[0], nh, slots, H1, h1o, (long long)kmax_[0], 1LL,
                              (long long)kmax_[1] * kmax_[0], H2, h2o, yfast ? 1LL : (long long)h2p_,
                              yfast ? (long long)ny : 1LL, (long long)ny * h2p_, 0, 0, ny, dd, 1, act, twy, 1);
        }
        else {
            const int W = choose_w(ny, e, 2, false, 16, 1);
            hipLaunchKernelGGL((k_c2c<T>), dim3((kmax_[0] + W - 1) / W, nh, slots), dim3(kThreads),
                               lds_bytes<T>(ny, 2, W), stream_, H1, h1o, (long long)kmax_[0], 1LL,
                               (long long)kmax_[1] * kmax_[0], H2, h2o, (long long)h2p_, 1LL,
                               (long long)ny * h2p_, 0, 0, ny, dd, 1, act, twy, plans_[1], W, 1);
        }
        mark("kernel_c2c_y");
        spectra_yfast_ = yfast;
    }

    // One register-FFT pass of the step (P1, P2 or P3) at a given tile.
    template <typename T> void rf_pass(int pass, Tile tile, const int* act, int slots)
    {
        using T2 = cx<T>;
        const int nx = n_[0], ny = n_[1], nz = n_[2];
        const int nh = nz / 2 + 1;
        T2* SB = sb_.as<T2>();
        T2* H2 = h2_.as<T2>();
        T2* TT = t_.as<T2>();
        const OrientDesc* dd = d_descs_.as<OrientDesc>();
        const long long h2o = (long long)nh * ny * h2p_;
        const long long to = (long long)nh * ny * nx;
        if (pass == kPassMul) {
            if constexpr (std::is_same_v<T, float>) {
                if (fused_yz()) {
                    // the fused pass's 1/(nx ny nz) = 2^-24, applied exactly to the environment rows
                    mul_grid_ = rf_mul_xt_launch<T>(tile, stream_, cus_, slots, ny, nh, H2, h2o, h2p_, SB, TT, to, dd,
                                                    act, twiddles<T>(nx), T(1) / (T(nx) * T(ny) * T(nz)), &mul_tiles_);
                    return;
                }
            }
            // an update window whose C2R pass applies its counts needs no
            // column or row outside the update's support
            rf_mul_launch<T>(nx, tile, stream_, cus_, slots, ny, nh, H2, h2o, h2p_, SB, TT, to, dd, act, twiddles<T>(nx),
                             apply_on_ ? apply_.ax : 0);
        }
        else if (pass == kPassInvY) {
            rf_cols_launch<T>(ny, tile, stream_, lmax_[0], nh, slots, TT, to, (long long)nx, 1LL, (long long)nx * ny, TT, to,
                              (long long)nx, 1LL, (long long)nx * ny, 0, ny, 0, dd, 2, act, twiddles<T>(ny), 0,
                              apply_on_ ? apply_.ax : 0, apply_on_ ? apply_.ay : 0);
        }
        else {
            const T inv_total = T(1) / (T(nx) * T(ny) * T(nz));
            const StepFilter filt { act, region_.as<unsigned char>(), ext_[0], ext_[1], ext_[2],
                                    orient_best_.as<unsigned long long>() };
            rf_c2r_launch<T>(nz, tile, stream_, lmax_[0], lmax_[1], slots, TT, to, nx, ny, dd, inv_total,
                             result_.as<StepDeviceResult>(), fields_needed() ? fields_.as<int>() : nullptr,
                             d_offsets_.as<long long>(), excl_.as<unsigned long long>(),
                             keep_raw_ ? raw_.as<double>() : nullptr, filt, twiddles<T>(nz),
                             apply_on_ ? &apply_ : nullptr);
        }
    }

    // The fused inverse y + inverse z pass (rf_yz_argmin) replaces P2 and P3
    // on FP32 256^3 grids when the step keeps no per-value output (fields or
    // raw values) and neither pass has a pinned tile.
    bool fused_yz() const
    {
        return fuse_yz_ && use_rf_ && precision_ == 32 && n_[0] == 256 && n_[1] == 256 && n_[2] == 256 &&
               !fields_needed() && !keep_raw_ && !pinned_[kPassInvY] && !pinned_[kPassC2r] &&
               yz_lds_bytes<float>() <= max_lds_;
    }

    bool fused_active()
    {
        prepare();
        return fused_yz();
    }

    void set_fused_passes(bool on)
    {
        if (on != fuse_yz_) {
            fuse_yz_ = on;
            tuned_ = false;
            drop_graphs();
        }
    }

    template <typename T> void enqueue_yz(const int* act, int slots)
    {
        if constexpr (std::is_same_v<T, float>) {
            const int nx = n_[0], ny = n_[1], nz = n_[2];
            const long long to = (long long)(nz / 2 + 1) * ny * nx;
            const StepFilter filt { act, region_.as<unsigned char>(), ext_[0], ext_[1], ext_[2],
                                    orient_best_.as<unsigned long long>() };
            // one resident block per CU walks its share of the legal columns
            const int grid = std::max(1, std::min(cus_, lmax_[0] * slots));
            hipLaunchKernelGGL((rf_yz_argmin<T>), dim3(grid), dim3(kYzThreads), yz_lds_bytes<T>(), stream_,
                               t_.as<cx<T>>(), to, d_descs_.as<OrientDesc>(), result_.as<StepDeviceResult>(),
                               excl_.as<unsigned long long>(), filt, twiddles<T>(nz), slots, lmax_[0]);
        }
        else
            throw std::logic_error("the fused inverse y and z pass runs in FP32");
    }

    // One generic (LDS Stockham) pass at tile width W (pencils per block).
    template <typename T> void gen_pass(int pass, int W, const int* act, int slots)
    {
        using T2 = cx<T>;
        const int nx = n_[0], ny = n_[1], nz = n_[2];
        const int nh = nz / 2 + 1;
        T2* SB = sb_.as<T2>();
        T2* H2 = h2_.as<T2>();
        T2* TT = t_.as<T2>();
        const OrientDesc* dd = d_descs_.as<OrientDesc>();
        const long long h2o = (long long)nh * ny * h2p_;
        const long long to = (long long)nh * ny * nx;
        if (pass == kPassMul) {
            hipLaunchKernelGGL((k_mul_x<T>), dim3((ny + W - 1) / W, nh, slots), dim3(kThreads), lds_bytes<T>(nx, 4, W),
                               stream_, H2, h2o, h2p_, ny, SB, nx, ny, TT, to, dd, 0, 0, act, twiddles<T>(nx),
                               plans_[0], W);
        }
        else if (pass == kPassInvY) {
            hipLaunchKernelGGL((k_c2c<T>), dim3((lmax_[0] + W - 1) / W, nh, slots), dim3(kThreads), lds_bytes<T>(ny, 2, W),
                               stream_, TT, to, (long long)nx, 1LL, (long long)nx * ny, TT, to, (long long)nx, 1LL,
                               (long long)nx * ny, 0, ny, 0, dd, 2, act, twiddles<T>(ny), plans_[1], W, 0);
        }
        else {
            const T inv_total = T(1) / (T(nx) * T(ny) * T(nz));
            const StepFilter filt { act, region_.as<unsigned char>(), ext_[0], ext_[1], ext_[2],
                                    orient_best_.as<unsigned long long>() };
            if (W < 2 || kThreads % W != 0) throw std::invalid_argument("inverse z tile width must divide the block");
            int* fo = fields_needed() ? fields_.as<int>() : nullptr;
            double* ro = keep_raw_ ? raw_.as<double>() : nullptr;
            if (fo || ro)
                hipLaunchKernelGGL((k_c2r_z_argmin<T, true>), dim3((lmax_[0] + W - 1) / W, lmax_[1], slots),
                                   dim3(kThreads), lds_bytes<T>(nz, 2, W / 2), stream_, TT, to, nx, ny, dd, inv_total,
                                   result_.as<StepDeviceResult>(), fo, d_offsets_.as<long long>(),
                                   excl_.as<unsigned long long>(), ro, filt, twiddles<T>(nz), plans_[2], W);
            else
                hipLaunchKernelGGL((k_c2r_z_argmin<T, false>), dim3((lmax_[0] + W - 1) / W, lmax_[1], slots),
                                   dim3(kThreads), lds_bytes<T>(nz, 2, W / 2), stream_, TT, to, nx, ny, dd, inv_total,
                                   result_.as<StepDeviceResult>(), fo, d_offsets_.as<long long>(),
                                   excl_.as<unsigned long long>(), ro, filt, twiddles<T>(nz), plans_[2], W);
        }
    }

    bool pass_uses_rf(int pass) const { return use_rf_ && rf_supported(pass_length(pass)); }

    // Default generic tile width (the LDS heuristic) of a pass.
    int generic_default_w(int pass) const
    {
        const std::size_t e = precision_ == 64 ? sizeof(double2) : sizeof(float2);
        if (pass == kPassMul) return choose_w(n_[0], e, 4, false, 16, 1);
        if (pass == kPassInvY) return choose_w(n_[1], e, 2, false, 16, 1);
        return choose_w(n_[2], e, 2, true, 32, 2);
    }

    std::size_t generic_lds(int pass, int W) const
    {
        const std::size_t e = precision_ == 64 ? sizeof(double2) : sizeof(float2);
        const int n = pass_length(pass);
        if (pass == kPassMul) return e * (std::size_t(n) + 4 * std::size_t(n) * (W + 1));
        if (pass == kPassInvY) return e * (std::size_t(n) + 2 * std::size_t(n) * (W + 1));
        return e * (std::size_t(n) + 2 * std::size_t(n) * (W / 2 + 1));
    }

    std::vector<Tile> generic_candidates(int pass) const
    {
        std::vector<Tile> v;
        for (int w : { 1, 2, 4, 8, 16, 32, 64 }) {
            if (pass == kPassC2r && w < 2) continue;
            if (generic_lds(pass, w) <= max_lds_) v.push_back(Tile { 0, w });
        }
        return v;
    }

    template <typename T> void run_pass(int pass, Tile tile, const int* act, int slots)
    {
        if (pass_uses_rf(pass)) rf_pass<T>(pass, tile.r ? tile : default_tile(pass, pass_length(pass)), act, slots);
        else gen_pass<T>(pass, tile.r == 0 && tile.w > 0 ? tile.w : generic_default_w(pass), act, slots);
    }

    int pass_length(int pass) const { return pass == kPassMul ? n_[0] : pass == kPassInvY ? n_[1] : n_[2]; }

    bool tile_fits(int pass, Tile t) const
    {
        const std::size_t elem = precision_ == 64 ? sizeof(double2) : sizeof(float2);
        return tile_lds(pass, pass_length(pass), t, elem) <= max_lds_;
    }

    // Project-written autotuner: times every fitting tile of each register-FFT
    // pass on this session's buffers (one warm-up and three timed launches with
    // HIP events) and keeps the fastest; choices are cached per shape in the
    // context.  Pinned passes are left alone.  Runs outside graph capture.
    template <typename T> void autotune_passes()
    {
        const int nx = n_[0], ny = n_[1], nz = n_[2];
        bool any = false;
        for (int pass = 0; pass != kPasses; ++pass) any = any || !pinned_[pass];
        if (!any) return;
        // Tune on real data: one full step without stamping leaves the
        // environment spectrum and the inverse-y output of the current
        // occupancy in place.  P3 (whose cost depends on how many counts are
        // zero) is timed first on that output and only reads it; P1 rewrites
        // the buffer from the spectra each launch; P2 runs in place last (its
        // cost does not depend on the values).  The next step recomputes all.
        enqueue_step<T>(false, act_full_.as<int>(), nfit_);
        SP_HIP_CHECK(hipGetLastError());
        hipEvent_t e0, e1;
        SP_HIP_CHECK(hipEventCreate(&e0));
        SP_HIP_CHECK(hipEventCreate(&e1));
        try {
            for (int pass : { int(kPassC2r), int(kPassMul), int(kPassInvY) }) {
                const int n = pass_length(pass);
                if (pinned_[pass]) continue;
                // the inverse z pass pairs columns into complex transforms by
                // their place in the tile, so its width sets the arithmetic too
                if (fixed_radix_ && pass == kPassC2r) continue;
                // the fused pass replaces P2 and P3 and has no tile; P1 then
                // writes the column-major layout at radix 16
                const bool fused = fused_yz();
                if (fused && pass != kPassMul) continue;
                const bool rf = pass_uses_rf(pass);
                char key[160];
                std::snprintf(key, sizeof key, "p%d%s%s%s n%d,%d,%d f%d k%d,%d l%d,%d s%d", pass, rf ? "r" : "g",
                              fixed_radix_ ? "d" : "", fused ? "y" : "", nx, ny, nz, precision_, kmax_[0], kmax_[1],
                              lmax_[0], lmax_[1], nfit_);
                auto it = ctx_.tiles.find(key);
                if (it != ctx_.tiles.end()) {
                    tile_[pass] = Tile { it->second.first, it->second.second };
                    tile_ms_[pass] = -1.f;
                    continue;
                }
                float best = -1.f;
                Tile best_tile = rf ? default_tile(pass, n) : Tile { 0, generic_default_w(pass) };
                for (const Tile t : rf ? tile_candidates(pass, n) : generic_candidates(pass)) {
                    if (rf && !tile_fits(pass, t)) continue;
                    // the radix sets the arithmetic; the tile width does not
                    if (rf && fixed_radix_ && t.r != default_tile(pass, n).r) continue;
                    if (fused && t.r != 16) continue;
                    run_pass<T>(pass, t, act_full_.as<int>(), nfit_);
                    if (hipGetLastError() != hipSuccess) continue;  // launch refused: not a candidate here
                    SP_HIP_CHECK(hipEventRecord(e0, stream_));
                    for (int i = 0; i != 3; ++i) run_pass<T>(pass, t, act_full_.as<int>(), nfit_);
                    SP_HIP_CHECK(hipEventRecord(e1, stream_));
                    SP_HIP_CHECK(hipEventSynchronize(e1));
                    SP_HIP_CHECK(hipGetLastError());
                    float ms = 0.f;
                    SP_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                    if (best < 0.f || ms < best) {
                        best = ms;
                        best_tile = t;
                    }
                }
                tile_[pass] = best_tile;
                tile_ms_[pass] = best / 3.f;
                ctx_.tiles[key] = { best_tile.r, best_tile.w };
            }
        }
        catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }

    void autotune()
    {
        if (!autotune_ || tuned_ || nfit_ == 0 || lmax_[0] <= 0) return;
        const bool prof = prof_;
        prof_ = false;
        if (precision_ == 64) autotune_passes<double>();
        else autotune_passes<float>();
        prof_ = prof;
        tuned_ = true;
    }

    void set_autotune(bool on)
    {
        autotune_ = on;
        tuned_ = false;
        if (!on)
            for (int p = 0; p != kPasses; ++p)
                if (!pinned_[p]) tile_[p] = Tile {};
        drop_graphs();
    }

    Tile effective_tile(int pass) const
    {
        const Tile t = tile_[pass];
        if (pass_uses_rf(pass)) return t.r ? t : default_tile(pass, pass_length(pass));
        return Tile { 0, t.r == 0 && t.w > 0 ? t.w : generic_default_w(pass) };
    }

    void set_tile(int pass, int r, int w)
    {
        prepare();
        if (pass < 0 || pass >= kPasses) throw std::invalid_argument("pass");
        const int n = pass_length(pass);
        if (r == 0 && w == 0) {
            pinned_[pass] = false;
            tile_[pass] = default_tile(pass, n);
            tuned_ = false;
            drop_graphs();
            return;
        }
        bool known = false;
        if (pass_uses_rf(pass)) {
            for (const Tile t : tile_candidates(pass, n)) known = known || (t.r == r && t.w == w);
            if (!known) throw std::invalid_argument("no such tile");
            if (!tile_fits(pass, Tile { r, w })) throw std::invalid_argument("tile exceeds the LDS of the device");
        }
        else {
            for (const Tile t : generic_candidates(pass)) known = known || (t.r == r && t.w == w);
            if (!known) throw std::invalid_argument("no such tile (generic passes take radix 0)");
        }
        pinned_[pass] = true;
        tile_[pass] = Tile { r, w };
        tile_ms_[pass] = 0.f;
        drop_graphs();
    }

    void get_tile(int pass, int& r, int& w, float& ms)
    {
        prepare();
        if (pass < 0 || pass >= kPasses) throw std::invalid_argument("pass");
        const Tile t = effective_tile(pass);
        r = t.r;
        w = t.w;
        ms = tile_ms_[pass];
    }

    std::vector<Tile> fitting_tiles(int pass)
    {
        prepare();
        if (pass < 0 || pass >= kPasses) throw std::invalid_argument("pass");
        std::vector<Tile> out;
        const int n = pass_length(pass);
        if (!pass_uses_rf(pass)) return generic_candidates(pass);
        for (const Tile t : tile_candidates(pass, n))
            if (tile_fits(pass, t)) out.push_back(t);
        return out;
    }

    template <typename T, typename In> void enqueue_env_r2c(const In* src, cx<T>* SB, const cx<T>* twz)
    {
        const int nx = n_[0], ny = n_[1], nz = n_[2];
        if (rf_supported(nz) && use_rf_) {
            rf_r2c_launch<T, In>(nz, stream_, nx, ny, 1, src, ext_[0], ext_[1], ext_[0], ext_[1], ext_[2], SB, 0LL, nx,
                                 ny, nx, ny, nullptr, nullptr, twz);
        }
        else {
            const int W = choose_w(nz, sizeof(cx<T>), 2, true, 32, 2);
            hipLaunchKernelGGL((k_r2c_z<T, In>), dim3((nx + W - 1) / W, ny, 1), dim3(kThreads),
                               lds_bytes<T>(nz, 2, W / 2), stream_, src, 0LL, ext_[0], ext_[1], ext_[0], ext_[1], ext_[2],
                               SB, 0LL, nx, ny, nx, ny, nullptr, nullptr, twz, plans_[2], W);
        }
    }

    template <typename T>
    void enqueue_step(bool stamp, const int* act, int slots, const std::function<void(hipStream_t)>* merge = nullptr)
    {
        using T2 = cx<T>;
        const int nx = n_[0], ny = n_[1], nz = n_[2];
        const int nh = nz / 2 + 1;
        const T2* twx = twiddles<T>(nx);
        const T2* twy = twiddles<T>(ny);
        const T2* twz = twiddles<T>(nz);
        T2* SB = sb_.as<T2>();
        T2* H2 = h2_.as<T2>();
        T2* TT = t_.as<T2>();
        const OrientDesc* dd = d_descs_.as<OrientDesc>();
        StepDeviceResult* res = result_.as<StepDeviceResult>();
        const std::size_t e = sizeof(T2);
        mark("start");
        hipLaunchKernelGGL(k_reset_result, dim3(1), dim3(256), 0, stream_, res, orient_best_.as<unsigned long long>(), O_);
        mark("reset");
        // E1: environment R2C along z (uint8 occupancy, or the real-valued
        // field of a proximity session -> spectrum planes)
        if (real_env_) {
            if constexpr (std::is_same_v<T, double>) enqueue_env_r2c<T, double>(envd_.as<double>(), SB, twz);
            else throw std::logic_error("real-valued environments run in FP64");
        }
        else
            enqueue_env_r2c<T, unsigned char>(env_.as<unsigned char>(), SB, twz);
        mark("env_r2c_z");
        // E2: C2C along y (columns x, batch kz)
        if (rf_supported(ny) && use_rf_) {
            rf_cols_launch<T>(ny, default_tile(kPassInvY, ny), stream_, nx, nh, 1, SB, 0LL, (long long)nx, 1LL, (long long)nx * ny, SB, 0LL,
                              (long long)nx, 1LL, (long long)nx * ny, nx, ny, ny, nullptr, 0, nullptr, twy, 1);
            mark("env_c2c_y");
        }
        else {
            const int W = choose_w(ny, e, 2, false, 16, 1);
            hipLaunchKernelGGL((k_c2c<T>), dim3((nx + W - 1) / W, nh, 1), dim3(kThreads), lds_bytes<T>(ny, 2, W),
                               stream_, SB, 0LL, (long long)nx, 1LL, (long long)nx * ny, SB, 0LL, (long long)nx, 1LL,
                               (long long)nx * ny, nx, ny, ny, nullptr, 0, nullptr, twy, plans_[1], W, 1);
            mark("env_c2c_y");
        }
        // E3: C2C along x (rows, batch kz); the fused path's product pass
        // transforms each row tile along x itself
        if (!fused_yz()) {
            if (rf_supported(nx) && use_rf_) {
                rf_rows_launch<T>(nx, stream_, SB, ny, (long long)nx, nh, (long long)nx * ny, twx, 1);
            }
            else {
                const int W = choose_w(nx, e, 2, false, 16, 1);
                hipLaunchKernelGGL((k_c2c<T>), dim3((ny + W - 1) / W, nh, 1), dim3(kThreads), lds_bytes<T>(nx, 2, W),
                                   stream_, SB, 0LL, 1LL, (long long)nx, (long long)nx * ny, SB, 0LL, 1LL,
                                   (long long)nx, (long long)nx * ny, ny, nx, nx, nullptr, 0, nullptr, twx, plans_[0],
                                   W, 1);
            }
            mark("env_c2c_x");
        }
        const long long h2o = (long long)nh * ny * h2p_;
        const long long to = (long long)nh * ny * nx;
        if (slots > 0 && !resident_) enqueue_kernel_spectra<T>(act, slots);
        if (slots > 0) {
            // P1: finish kernel x transform, product with conj, inverse x
            run_pass<T>(kPassMul, tile_[kPassMul], act, slots);
            mark("mul_x");
        }
        if (slots > 0 && lmax_[0] > 0 && fused_yz()) {
            // P2 + P3 in one pass over each legal x column
            enqueue_yz<T>(act, slots);
            mark("yz_argmin");
        }
        else if (slots > 0 && lmax_[0] > 0) {
            // P2: inverse C2C along y for the legal columns, rows y < ly written
            run_pass<T>(kPassInvY, tile_[kPassInvY], act, slots);
            mark("inv_c2c_y");
            // P3: C2R along z + round + checks + feasible + lowest (z, y, x)
            run_pass<T>(kPassC2r, tile_[kPassC2r], act, slots);
            mark("c2r_z_argmin");
        }
        if (merge) (*merge)(stream_);
        enqueue_verify();
        if (stamp) {
            enqueue_stamp(nullptr, 0, 0, nullptr);
            mark("stamp");
        }
    }

    // Resident kernel spectra: H2 of every fitting orientation, computed once
    // after preparation (outside any graph) and reused by every step.
    void ensure_kernel_spectra()
    {
        // (computed again when the fused pass is switched on or off: its
        // product pass reads them in the other layout)
        if (!resident_ || (spectra_ready_ && spectra_yfast_ == fused_yz()) || nfit_ == 0) return;
        const bool prof = prof_;
        prof_ = false;
        if (precision_ == 64) enqueue_kernel_spectra<double>(act_full_.as<int>(), nfit_);
        else enqueue_kernel_spectra<float>(act_full_.as<int>(), nfit_);
        prof_ = prof;
        SP_HIP_CHECK(hipGetLastError());
        spectra_ready_ = true;
    }

    void set_resident_kernels(bool on)
    {
        if (on != resident_) {
            resident_ = on;
            spectra_ready_ = false;
            drop_graphs();
        }
    }

    void enqueue(bool stamp, bool subset)
    {
        const int* act = subset ? act_sub_.as<int>() : act_full_.as<int>();
        const int slots = subset ? sub_slots_ : nfit_;
        if (precision_ == 64) enqueue_step<double>(stamp, act, slots);
        else enqueue_step<float>(stamp, act, slots);
    }

    void drop_graphs()
    {
        for (auto& g : graphs_) {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            if (g.graph) (void)hipGraphDestroy(g.graph);
        }
        graphs_.clear();
    }

    // Replays the graph recorded for (what, stamp, owner, slots), recording it
    // on first use.  Device pointers inside stay valid until prepare() runs
    // again, which drops every graph.
    void run_graph(int what, bool stamp, const void* owner, int slots, const std::function<void()>& record)
    {
        for (auto& g : graphs_)
            if (g.what == what && g.stamp == stamp && g.owner == owner && g.slots == slots) {
                SP_HIP_CHECK(hipGraphLaunch(g.exec, stream_));
                return;
            }
        if (graphs_.size() >= 8) drop_graphs();
        CachedGraph g { what, stamp, owner, slots, nullptr, nullptr };
        SP_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        try {
            record();
        }
        catch (...) {
            hipGraph_t broken = nullptr;
            (void)hipStreamEndCapture(stream_, &broken);
            if (broken) (void)hipGraphDestroy(broken);
            throw;
        }
        SP_HIP_CHECK(hipStreamEndCapture(stream_, &g.graph));
        SP_HIP_CHECK(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
        graphs_.push_back(g);
        SP_HIP_CHECK(hipGraphLaunch(g.exec, stream_));
    }

    void launch(bool stamp)
    {
        if (incremental_) launch_incremental(stamp);
        else launch_mode(stamp, subset_);
    }

    void launch_mode(bool stamp, bool subset)
    {
        if (stamp) ++env_rev_;
        prepare();
        use_device(ctx_);
        ensure_kernel_spectra();
        autotune();
        if (!desc_.capture_graph) {
            enqueue(stamp, subset);
            SP_HIP_CHECK(hipGetLastError());
            return;
        }
        run_graph(subset ? 1 : 0, stamp, nullptr, subset ? sub_slots_ : nfit_, [&] { enqueue(stamp, subset); });
    }

    // A proposal step: `propose` enqueues work on the session stream that
    // writes the subset list (at most `slots` orientations) and the region
    // mask on the device, then the subset search runs; both in one graph.
    void launch_proposed(const void* owner, const std::function<void(hipStream_t)>& propose, int slots, bool stamp,
                         const std::function<void(hipStream_t)>* merge = nullptr)
    {
        if (incremental_) throw std::invalid_argument("proposal steps need a session without incremental steps");
        prepare();
        use_device(ctx_);
        ensure_kernel_spectra();
        autotune();
        if (slots <= 0 || slots > O_) throw std::invalid_argument("proposal slots");
        if (stamp) ++env_rev_;
        auto record = [&] {
            propose(stream_);
            if (precision_ == 64) enqueue_step<double>(stamp, act_sub_.as<int>(), slots, merge);
            else enqueue_step<float>(stamp, act_sub_.as<int>(), slots, merge);
        };
        if (!desc_.capture_graph) {
            record();
            SP_HIP_CHECK(hipGetLastError());
            return;
        }
        run_graph(merge ? 3 : 2, stamp, owner, slots, record);
    }

    SessionView view()
    {
        prepare();
        SessionView v {};
        v.env = env_.as<unsigned char>();
        for (int a = 0; a != 3; ++a) v.ext[a] = ext_[a];
        v.orientations = O_;
        v.fitting = nfit_;
        v.act = act_sub_.as<int>();
        v.region = region_.as<unsigned char>();
        v.max_regions_per_axis = kMaxRegionsPerAxis;
        v.descs = d_descs_.as<OrientDesc>();
        v.kcells = kcells_.as<unsigned char>();
        v.orient_best = orient_best_.as<unsigned long long>();
        v.result = result_.as<StepDeviceResult>();
        v.stream = stream_;
        v.serial = serial_;
        v.reset_rev = reset_rev_;
        return v;
    }

    void product_grid(std::uint32_t& blocks, std::uint64_t& tiles) const
    {
        blocks = std::uint32_t(mul_grid_);
        tiles = std::uint64_t(mul_tiles_);
    }

    // Back-to-back launches of one pass of the step (as a rocprof kernel trace
    // sees them): `launches` launches recorded in one graph and replayed
    // between two events on the session stream, after a step has filled its
    // inputs.  Every pass rewrites the same output from the same inputs, so
    // the session's state is unchanged.  Returns ms per launch.
    float time_pass(const std::string& name, int launches)
    {
        if (launches < 1 || launches > 1000) throw std::invalid_argument("launches");
        prepare();
        use_device(ctx_);
        ensure_kernel_spectra();
        autotune();
        const int* act = act_full_.as<int>();
        const int slots = nfit_;
        if (slots == 0 || lmax_[0] == 0) throw std::invalid_argument("no orientation fits");
        std::function<void()> one;
        if (name == "mul_x")
            one = [&] {
                if (precision_ == 64) run_pass<double>(kPassMul, tile_[kPassMul], act, slots);
                else run_pass<float>(kPassMul, tile_[kPassMul], act, slots);
            };
        else if (name == "yz_argmin") {
            if (!fused_yz()) throw std::invalid_argument("the fused pass does not run on this session");
            one = [&] { enqueue_yz<float>(act, slots); };
        }
        else if (name == "inv_c2c_y" || name == "c2r_z_argmin") {
            if (fused_yz()) throw std::invalid_argument("the separate inverse passes do not run on this session");
            const int pass = name == "inv_c2c_y" ? kPassInvY : kPassC2r;
            one = [&, pass] {
                if (precision_ == 64) run_pass<double>(pass, tile_[pass], act, slots);
                else run_pass<float>(pass, tile_[pass], act, slots);
            };
        }
        else
            throw std::invalid_argument("unknown pass");
        // inputs of every pass: one step without stamping
        launch_mode(false, false);
        one();
        SP_HIP_CHECK(hipGetLastError());
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        SP_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        try {
            for (int i = 0; i != launches; ++i) one();
        }
        catch (...) {
            (void)hipStreamEndCapture(stream_, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        SP_HIP_CHECK(hipStreamEndCapture(stream_, &graph));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        float ms = 0.f;
        try {
            SP_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            SP_HIP_CHECK(hipGraphLaunch(exec, stream_));  // warm-up replay
            SP_HIP_CHECK(hipEventCreate(&e0));
            SP_HIP_CHECK(hipEventCreate(&e1));
            SP_HIP_CHECK(hipEventRecord(e0, stream_));
            SP_HIP_CHECK(hipGraphLaunch(exec, stream_));
            SP_HIP_CHECK(hipEventRecord(e1, stream_));
            SP_HIP_CHECK(hipEventSynchronize(e1));
            SP_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        }
        catch (...) {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (exec) (void)hipGraphExecDestroy(exec);
            (void)hipGraphDestroy(graph);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        return ms / float(launches);
    }

    // One step without the graph, timing each kernel between events.
    std::vector<std::pair<std::string, float>> profile_step(bool stamp)
    {
        prepare();
        use_device(ctx_);
        ensure_kernel_spectra();
        autotune();
        prof_ = true;
        prof_events_.clear();
        prof_names_.clear();
        try {
            enqueue(stamp, subset_);
        }
        catch (...) {
            prof_ = false;
            throw;
        }
        prof_ = false;
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
        std::vector<std::pair<std::string, float>> out;
        for (std::size_t i = 1; i < prof_events_.size(); ++i) {
            float ms = 0;
            SP_HIP_CHECK(hipEventElapsedTime(&ms, prof_events_[i - 1], prof_events_[i]));
            out.emplace_back(prof_names_[i], ms);
        }
        for (auto e : prof_events_) (void)hipEventDestroy(e);
        prof_events_.clear();
        return out;
    }

    void read_result(sp_step_result& out)
    {
        use_device(ctx_);
        SP_HIP_CHECK(hipMemcpyAsync(host_result_, result_.get(), sizeof(StepDeviceResult), hipMemcpyDeviceToHost, stream_));
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
        const StepDeviceResult r = *host_result_;
        if (r.run_flags) {
            // reported once: later reads start a new run
            SP_HIP_CHECK(hipMemsetAsync(&result_.as<StepDeviceResult>()->run_flags, 0, sizeof(unsigned int), stream_));
            SP_HIP_CHECK(hipStreamSynchronize(stream_));
        }
        out = sp_step_result {};
        out.precision = precision_;
        out.error_bound = bound_;
        out.numeric_flags = r.flags;
        out.run_flags = r.run_flags;
        float res;
        std::memcpy(&res, &r.max_residual_bits, sizeof res);
        out.max_residual = res;
        if (r.key == ~0ULL) {
            out.orientation = -1;
            return;
        }
        const int o = int(r.key & 0xffffULL);
        const std::int64_t u[3] = { std::int64_t((r.key >> 16) & 0xffffULL), std::int64_t((r.key >> 32) & 0xffffULL),
                                     std::int64_t(r.key >> 48) };
        out.orientation = o;
        for (int a = 0; a != 3; ++a) {
            // key = t + kernel_first = environment_first + u; t = key - kernel_first
            out.key[a] = efirst_[a] + u[a];
            out.translation[a] = out.key[a] - kernels_[o].first[a];
        }
    }

    void overlap_field(std::uint32_t o, std::int32_t* out, std::uint64_t count, std::uint32_t legal[3])
    {
        if (!fields_needed() || !prepared_) throw std::invalid_argument("fields not kept");
        if (int(o) >= O_) throw std::invalid_argument("orientation index");
        const OrientDesc& d = descs_[o];
        legal[0] = std::uint32_t(std::max(d.lx, 0));
        legal[1] = std::uint32_t(std::max(d.ly, 0));
        legal[2] = std::uint32_t(std::max(d.lz, 0));
        const std::uint64_t n = std::uint64_t(legal[0]) * legal[1] * legal[2];
        if (count < n) throw std::length_error("capacity");
        if (n == 0) return;
        use_device(ctx_);
        SP_HIP_CHECK(hipMemcpyAsync(out, fields_.as<int>() + field_offsets_[o], n * sizeof(int), hipMemcpyDeviceToHost,
                                    stream_));
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    // ---- ranked pages (spectral_pipeline.cpp:1343,1563,1740-1857) ---------
    // Score constants and each orientation's oriented-bounds max z.  Ranking
    // keeps the binary overlap of every legal corner (fields) and adds an FP64
    // proximity session sharing this stream.
    void set_ranking(const sp_ranking_desc& r, const double* max_z, std::uint32_t count)
    {
        if (count != std::uint32_t(O_)) throw std::invalid_argument("one oriented max z per orientation");
        const double height = r.container_max_z_mm - r.container_min_z_mm;
        if (!std::isfinite(r.lattice_origin_z_mm) || !std::isfinite(r.pitch_mm) || !(r.pitch_mm > 0.0) ||
            !std::isfinite(r.container_min_z_mm) || !std::isfinite(height) || !(height > 0.0))
            throw std::domain_error("SPECTRAL_PIPELINE_CONTAINER_HEIGHT");
        for (std::uint32_t o = 0; o != count; ++o)
            if (!std::isfinite(max_z[o])) throw std::invalid_argument("oriented max z must be finite");
        for (int a = 0; a != 3; ++a)
            if (ext_[a] > kMaxRankExtent) throw std::invalid_argument("ranking needs at most 1024 cells per axis");
        rank_desc_ = r;
        rank_max_z_.assign(max_z, max_z + count);
        rank_on_ = true;
        keep_fields(true);
        prox_.reset();
    }

    // Binary overlaps and the proximity correlation of every legal corner for
    // the current occupancy: one binary step (full or incremental, no stamp),
    // the proximity field, and one FP64 step of the proximity session.
    void rank()
    {
        if (!rank_on_) throw std::invalid_argument("ranking not set");
        if (subset_) throw std::invalid_argument("ranking searches every orientation");
        prepare();
        // spectral_pipeline.cpp:1735-1740: the mean proximity divides by the kernel's ones
        for (int o = 0; o != O_; ++o)
            if (descs_[o].lx > 0 && descs_[o].mass == 0) throw std::domain_error("SPECTRAL_PIPELINE_EMPTY_KERNEL");
        use_device(ctx_);
        if (incremental_) launch_incremental(false);
        else launch_mode(false, false);
        ensure_prox();
        enqueue_edt();
        mark("proximity_field");
        prox_->launch_mode(false, false);
        SP_HIP_CHECK(hipGetLastError());
        rank_rev_ = env_rev_;
    }

    void ensure_prox()
    {
        if (prox_) return;
        sp_session_desc d = desc_;
        d.precision = SP_PRECISION_FP64;
        auto p = std::make_unique<Session>(ctx_, d, stream_);
        const std::size_t cells = std::size_t(ext_[0]) * ext_[1] * ext_[2];
        p->real_env_ = true;
        p->envd_.allocate(sizeof(double) * cells);
        for (int o = 0; o != O_; ++o) {
            const Kernel& k = kernels_[o];
            const std::uint32_t sh[3] = { std::uint32_t(k.shape[0]), std::uint32_t(k.shape[1]), std::uint32_t(k.shape[2]) };
            p->set_kernel(std::uint32_t(o), k.cells.data(), sh, k.first);
        }
        p->keep_raw(true);
        // unrounded values: every process must compute the same bits, so only
        // the tile widths of the product and inverse y passes are tuned (the
        // radix and the inverse z tile set the arithmetic; integer fields are
        // indifferent to both)
        p->fixed_radix_ = true;
        p->set_fast_path(use_rf_);
        p->set_autotune(autotune_);
        p->prepare();
        edt_.allocate(sizeof(unsigned) * cells);
        // exp(-sqrt(d2)/2) for every squared distance the grid can hold, by the
        // reference's own expression (field_proximity.hpp:129)
        std::uint64_t dmax = 0;
        for (int a = 0; a != 3; ++a) dmax += std::uint64_t(ext_[a] - 1) * std::uint64_t(ext_[a] - 1);
        std::vector<double> table(dmax + 1);
        for (std::uint64_t v = 0; v <= dmax; ++v) table[v] = std::exp(-std::sqrt(static_cast<double>(v)) / 2.);
        edt_table_.allocate(sizeof(double) * table.size());
        SP_HIP_CHECK(hipMemcpyAsync(edt_table_.get(), table.data(), sizeof(double) * table.size(), hipMemcpyHostToDevice,
                                    stream_));
        std::vector<long long> kfz(O_);
        for (int o = 0; o != O_; ++o) kfz[o] = (long long)kernels_[o].first[2];
        rank_max_z_d_.allocate(sizeof(double) * O_);
        rank_kfz_d_.allocate(sizeof(long long) * O_);
        SP_HIP_CHECK(hipMemcpyAsync(rank_max_z_d_.get(), rank_max_z_.data(), sizeof(double) * O_, hipMemcpyHostToDevice,
                                    stream_));
        SP_HIP_CHECK(hipMemcpyAsync(rank_kfz_d_.get(), kfz.data(), sizeof(long long) * O_, hipMemcpyHostToDevice, stream_));
        rank_olist_.allocate(sizeof(int) * O_);
        rank_cursor_.allocate(sizeof(PageKey) * O_);
        rank_partial_.allocate(sizeof(PageKey) * std::size_t(O_) * kMaxPageBlocks * kPageSize);
        rank_pcount_.allocate(sizeof(int) * std::size_t(O_) * kMaxPageBlocks);
        rank_pages_.allocate(sizeof(PageKey) * std::size_t(O_) * kPageSize);
        rank_count_.allocate(sizeof(int) * O_);
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
        prox_ = std::move(p);
    }

    static int edt_tile_width(int n)
    {
        int tx = 64;
        while (tx > 1 && std::size_t(n) * tx * sizeof(unsigned) > 65536) tx >>= 1;
        return tx;
    }

    // build_proximity on the device: x rows, then y and z lines (the last
    // writing the FP64 field into the proximity session's environment).
    void enqueue_edt()
    {
        const long long rows = (long long)ext_[1] * ext_[2];
        hipLaunchKernelGGL(k_edt_rows, dim3(unsigned((rows + 3) / 4)), dim3(256), 0, stream_, env_.as<unsigned char>(),
                           ext_[0], rows, edt_.as<unsigned>());
        int tx = edt_tile_width(ext_[1]);
        hipLaunchKernelGGL(k_edt_lines<false>, dim3((ext_[0] + tx - 1) / tx, ext_[2]), dim3(256),
                           std::size_t(ext_[1]) * tx * sizeof(unsigned), stream_, edt_.as<unsigned>(), ext_[0], ext_[1],
                           (long long)ext_[0], (long long)ext_[0] * ext_[1], tx, nullptr, nullptr);
        tx = edt_tile_width(ext_[2]);
        hipLaunchKernelGGL(k_edt_lines<true>, dim3((ext_[0] + tx - 1) / tx, ext_[1]), dim3(256),
                           std::size_t(ext_[2]) * tx * sizeof(unsigned), stream_, edt_.as<unsigned>(), ext_[0], ext_[2],
                           (long long)ext_[0] * ext_[1], (long long)ext_[0], tx, edt_table_.as<double>(),
                           prox_->envd_.as<double>());
        SP_HIP_CHECK(hipGetLastError());
    }

    // One page per listed orientation (mode 0 ordinary, 1 low overlap), each
    // after its cursor when has_cursor[i] is set.  Recomputes the fields when
    // the occupancy changed since the last rank().
    void ranked_pages(int mode, const std::uint32_t* orients, std::uint32_t count, const sp_ranked_candidate* cursors,
                      const std::uint8_t* has_cursor, sp_ranked_candidate* out, std::uint32_t* counts)
    {
        if (mode != 0 && mode != 1) throw std::invalid_argument("page mode");
        if (!rank_on_) throw std::invalid_argument("ranking not set");
        prepare();
        if (!prox_ || rank_rev_ != env_rev_) rank();
        use_device(ctx_);
        std::vector<int> olist;
        std::vector<int> slot_of(count, -1);
        std::vector<PageKey> keys;
        bool any_cursor = false;
        std::vector<char> listed(O_, 0);
        for (std::uint32_t i = 0; i != count; ++i) {
            const std::uint32_t o = orients[i];
            if (o >= std::uint32_t(O_)) throw std::invalid_argument("orientation index");
            if (listed[o]) throw std::invalid_argument("orientation listed twice");
            listed[o] = 1;
            counts[i] = 0;
            if (descs_[o].lx <= 0) continue;  // a kernel wider than the window has no legal translation
            PageKey k { ~0ull, ~0ull, 0xffffffffu, 0u };
            if (cursors && has_cursor && has_cursor[i]) {
                const sp_ranked_candidate& c = cursors[i];
                unsigned long long pos = 0;
                const int shift[3] = { 0, 16, 32 };
                for (int a = 0; a != 3; ++a) {
                    const std::int64_t u = c.translation[a] + kernels_[o].first[a] - efirst_[a];
                    if (u < 0 || u > 0xffff) throw std::invalid_argument("cursor translation outside the window");
                    pos |= (unsigned long long)u << shift[a];
                }
                if (!std::isfinite(c.score)) throw std::invalid_argument("cursor score must be finite");
                double ov = mode ? c.binary_overlap : 0.0;
                if (!(ov >= 0.0) || ov > 4294967295.0 || ov != std::floor(ov))
                    throw std::invalid_argument("cursor overlap must be a non-negative integer");
                k = PageKey { score_order_bits(c.score), pos, unsigned(ov), 1u };
                any_cursor = true;
            }
            slot_of[i] = int(olist.size());
            olist.push_back(int(o));
            keys.push_back(k);
        }
        const int slots = int(olist.size());
        if (slots == 0) return;
        long long most = 0;
        for (int o : olist) most = std::max(most, (long long)descs_[o].lx * descs_[o].ly * descs_[o].lz);
        long long blocks = std::max<long long>(1, (4096 + slots - 1) / slots);
        blocks = std::min<long long>(blocks, kMaxPageBlocks);
        blocks = std::max<long long>(1, std::min<long long>(blocks, (most + kPagePerIter - 1) / kPagePerIter));
        SP_HIP_CHECK(hipMemcpyAsync(rank_olist_.get(), olist.data(), sizeof(int) * slots, hipMemcpyHostToDevice, stream_));
        SP_HIP_CHECK(hipMemcpyAsync(rank_cursor_.get(), keys.data(), sizeof(PageKey) * slots, hipMemcpyHostToDevice,
                                    stream_));
        RankParams rp {};
        rp.origin_z = rank_desc_.lattice_origin_z_mm;
        rp.pitch = rank_desc_.pitch_mm;
        rp.container_min_z = rank_desc_.container_min_z_mm;
        rp.container_height = rank_desc_.container_max_z_mm - rank_desc_.container_min_z_mm;
        rp.env_first_z = (long long)efirst_[2];
        rp.mode = mode;
        rp.any_cursor = any_cursor ? 1 : 0;
        hipLaunchKernelGGL(k_page_partial, dim3(unsigned(blocks), unsigned(slots)), dim3(kPageBlockThreads), 0, stream_,
                           fields_.as<int>(), prox_->raw_.as<double>(), d_offsets_.as<long long>(),
                           d_descs_.as<OrientDesc>(), rank_olist_.as<int>(), rp, rank_max_z_d_.as<double>(),
                           rank_kfz_d_.as<long long>(), rank_cursor_.as<PageKey>(), excl_.as<unsigned long long>(),
                           rank_partial_.as<PageKey>(), rank_pcount_.as<int>());
        hipLaunchKernelGGL(k_page_merge, dim3(unsigned(slots)), dim3(kPageBlockThreads), 0, stream_,
                           rank_partial_.as<PageKey>(), rank_pcount_.as<int>(), int(blocks), rank_pages_.as<PageKey>(),
                           rank_count_.as<int>());
        if (mode == 0)
            hipLaunchKernelGGL(k_page_recheck, dim3(unsigned(kPageSize), unsigned(slots)), dim3(kPageBlockThreads), 0,
                               stream_, rank_olist_.as<int>(), d_descs_.as<OrientDesc>(), kcells_.as<unsigned char>(),
                               env_.as<unsigned char>(), ext_[0], ext_[1], rank_pages_.as<PageKey>(),
                               rank_count_.as<int>());
        SP_HIP_CHECK(hipGetLastError());
        std::vector<PageKey> pages(std::size_t(slots) * kPageSize);
        std::vector<int> pc(slots);
        SP_HIP_CHECK(hipMemcpyAsync(pages.data(), rank_pages_.get(), sizeof(PageKey) * pages.size(),
                                    hipMemcpyDeviceToHost, stream_));
        SP_HIP_CHECK(hipMemcpyAsync(pc.data(), rank_count_.get(), sizeof(int) * slots, hipMemcpyDeviceToHost, stream_));
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
        for (std::uint32_t i = 0; i != count; ++i) {
            const int sl = slot_of[i];
            if (sl < 0) continue;
            const int o = olist[sl];
            counts[i] = std::uint32_t(pc[sl]);
            for (int j = 0; j != pc[sl]; ++j) {
                const PageKey& k = pages[std::size_t(sl) * kPageSize + j];
                sp_ranked_candidate& c = out[std::size_t(i) * kPageSize + j];
                const std::int64_t u[3] = { std::int64_t(k.pos & 0xffff), std::int64_t((k.pos >> 16) & 0xffff),
                                            std::int64_t(k.pos >> 32) };
                for (int a = 0; a != 3; ++a) c.translation[a] = efirst_[a] + u[a] - kernels_[o].first[a];
                c.score = score_from_order_bits(k.score);
                c.binary_overlap = double(k.overlap);
                c.orientation = std::uint32_t(o);
                c.direct_zero_overlap = k.direct;
            }
        }
    }

    // spectral.cpp:1485-1590, the ordinary part of the greedy trial loop:
    // fresh pages of every orientation for the current occupancy, the best
    // page head by rank_less (score, orientation, translation z, y, x) taken
    // in turn, a consumed full page refilled after its cursor, and the first
    // head that passed the direct recheck and `validate` (the caller's exact
    // check; all accepted when null) stamped -- then everything again for the
    // new occupancy, until no page has a candidate or max_copies are placed.
    std::vector<sp_placement> ranked_fill(std::uint32_t max_copies, sp_validate_fn validate, void* user)
    {
        if (!rank_on_) throw std::invalid_argument("ranking not set");
        prepare();
        const auto rank_less = [](const sp_ranked_candidate& a, const sp_ranked_candidate& b) {
            if (a.score != b.score) return a.score < b.score;
            if (a.orientation != b.orientation) return a.orientation < b.orientation;
            return std::tie(a.translation[2], a.translation[1], a.translation[0]) <
                   std::tie(b.translation[2], b.translation[1], b.translation[0]);
        };
        std::vector<std::uint32_t> all(O_);
        std::iota(all.begin(), all.end(), 0u);
        std::vector<sp_ranked_candidate> buf(std::size_t(O_) * kPageSize);
        std::vector<std::uint32_t> cnt(O_);
        std::vector<sp_placement> placed;
        while (placed.size() < max_copies) {
            ranked_pages(0, all.data(), std::uint32_t(O_), nullptr, nullptr, buf.data(), cnt.data());
            std::vector<std::vector<sp_ranked_candidate>> pages(O_);
            std::vector<std::size_t> index(O_, 0);
            std::vector<char> exhausted(O_);
            std::vector<sp_ranked_candidate> cursor(O_);
            for (int o = 0; o != O_; ++o) {
                pages[o].assign(buf.begin() + std::ptrdiff_t(o) * kPageSize,
                                buf.begin() + std::ptrdiff_t(o) * kPageSize + cnt[o]);
                exhausted[o] = pages[o].empty();
            }
            bool admitted = false;
            for (;;) {
                std::vector<std::uint32_t> refill;
                for (int o = 0; o != O_; ++o) {
                    if (exhausted[o] || index[o] != pages[o].size()) continue;
                    // a page shorter than a full one was the last of its orientation
                    if (pages[o].size() < std::size_t(kPageSize)) exhausted[o] = 1;
                    else refill.push_back(std::uint32_t(o));
                }
                if (!refill.empty()) {
                    std::vector<sp_ranked_candidate> cur(refill.size());
                    std::vector<std::uint8_t> has(refill.size(), 1);
                    for (std::size_t i = 0; i != refill.size(); ++i) cur[i] = cursor[refill[i]];
                    ranked_pages(0, refill.data(), std::uint32_t(refill.size()), cur.data(), has.data(), buf.data(),
                                 cnt.data());
                    for (std::size_t i = 0; i != refill.size(); ++i) {
                        const std::uint32_t o = refill[i];
                        pages[o].assign(buf.begin() + std::ptrdiff_t(i) * kPageSize,
                                        buf.begin() + std::ptrdiff_t(i) * kPageSize + cnt[i]);
                        index[o] = 0;
                        exhausted[o] = cnt[i] == 0;
                    }
                }
                int best = -1;
                for (int o = 0; o != O_; ++o)
                    if (index[o] < pages[o].size() &&
                        (best < 0 || rank_less(pages[o][index[o]], pages[best][index[best]])))
                        best = o;
                if (best < 0) break;
                const sp_ranked_candidate c = pages[best][index[best]++];
                cursor[best] = c;
                if (!c.direct_zero_overlap) continue;
                if (validate && !validate(user, c.orientation, c.translation)) continue;
                stamp_at(c.orientation, c.translation);
                sp_placement pl {};
                for (int a = 0; a != 3; ++a) pl.translation[a] = c.translation[a];
                pl.orientation = c.orientation;
                pl.score = c.score;
                placed.push_back(pl);
                admitted = true;
                break;
            }
            if (!admitted) break;
        }
        return placed;
    }

    // The proximity field of the last rank(): exp(-sqrt(d2)/2) and d2
    // (0x7fffffff where nothing is occupied), X-fastest over the window.
    void proximity_field(double* values, std::uint32_t* squared, std::uint64_t count)
    {
        if (!prox_) throw std::invalid_argument("no ranking run");
        const std::uint64_t cells = std::uint64_t(ext_[0]) * ext_[1] * ext_[2];
        if (count < cells) throw std::length_error("capacity");
        use_device(ctx_);
        if (values)
            SP_HIP_CHECK(hipMemcpyAsync(values, prox_->envd_.get(), cells * sizeof(double), hipMemcpyDeviceToHost, stream_));
        if (squared)
            SP_HIP_CHECK(hipMemcpyAsync(squared, edt_.get(), cells * sizeof(unsigned), hipMemcpyDeviceToHost, stream_));
        SP_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    // The proximity correlation of the last rank() at orientation o's legal corners.
    void proximity_values(std::uint32_t o, double* out, std::uint64_t count)
    {
        if (!prox_) throw std::invalid_argument("no ranking run");
        prox_->raw_values(o, out, count);
    }

    hipStream_t stream() const { return stream_; }
    std::uint64_t device_bytes() const
    {
        return env_.bytes() + kcells_.bytes() + sb_.bytes() + h1_.bytes() + h2_.bytes() + t_.bytes() + fields_.bytes();
    }

private:
    struct Kernel {
        int shape[3] {};
        std::int64_t first[3] {};
        std::vector<std::uint8_t> cells;
        std::uint64_t mass {};
        bool set {};
    };
    Context& ctx_;
    sp_session_desc desc_;
    int ext_[3] {}, n_[3] {};
    std::int64_t efirst_[3] {};
    int O_ {};
    std::vector<Kernel> kernels_;
    std::vector<OrientDesc> descs_;
    std::vector<long long> field_offsets_;
    std::array<int, 3> kmax_ {}, lmax_ {};
    FftPlan plans_[3] {};
    hipStream_t stream_ {};
    DeviceBuffer env_, result_, kcells_, d_descs_, d_offsets_, fields_, sb_, h1_, h2_, t_, excl_;
    DeviceBuffer act_full_, act_sub_, region_, orient_best_;
    std::vector<unsigned long long> excluded_;
    int nfit_ {};
    bool subset_ {};
    int sub_slots_ { 1 };
    StepDeviceResult* host_result_ {};
    long long fields_total_ {};
    bool keep_fields_ {}, keep_raw_ {}, prepared_ {};
    bool fuse_yz_ { true };
    DeviceBuffer raw_;
    int precision_ { 32 };
    double bound_ {};
    struct CachedGraph {
        int what;
        bool stamp;
        const void* owner;
        int slots;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<CachedGraph> graphs_;
    bool prof_ {};
    bool use_rf_ { true };
    bool resident_ { true };
    bool spectra_ready_ {};
    bool spectra_yfast_ {};    // H2 is y-fastest (the fused path's layout)
    bool owns_stream_ {};
    bool incremental_ {}, inc_setup_done_ {}, inc_effective_ {}, inc_ready_ {}, pending_ {};
    bool circular_ {};
    double pair_clearance_ {}, pitch_ { 1.0 };
    int halo_ {};
    DeviceBuffer scells_, sdescs_, stencil_;
    int margin_[3] {};
    std::unique_ptr<Session> upd_;
    DeviceBuffer rescan_;
    DeviceBuffer rfound_;   // per orientation: lowest zero key found by the running rescan (~0 between rescans)
    FieldUpdate apply_ {};  // set only while launch_apply enqueues
    bool apply_on_ {};
    bool autotune_ { true }, tuned_ {};
    Tile tile_[kPasses] {};
    float tile_ms_[kPasses] {};
    bool pinned_[kPasses] {};
    std::size_t max_lds_ { 65536 };
    int cus_ { 1 };  // compute units of the session's device
    int h2p_ { 0 };  // H2 row pitch (values)
    int mul_grid_ {};          // blocks of the last column-major product pass
    long long mul_tiles_ {};   // its row tiles (grid < tiles: shares of tiles x orientations)
    bool tracks_copies_ {};
    DeviceBuffer mask_, counts_, copylog_, fault_, rm_key_;
    std::set<std::uint64_t> removed_;
    std::uint64_t frozen_copies_ {};
    std::vector<hipEvent_t> prof_events_;
    std::vector<const char*> prof_names_;
    // occupancy revision (set_environment and every stamp) and ranking state
    std::uint64_t env_rev_ {}, rank_rev_ { ~0ull };
    // the session's process-wide serial, and the revision of every change
    // other than a stamp that alters which positions are feasible (a new
    // environment or kernel, an excluded translation, a removed copy, a new
    // clearance): between two such changes occupancy only grows
    const std::uint64_t serial_ { next_serial() };
    std::uint64_t reset_rev_ {};
    static std::uint64_t next_serial()
    {
        static std::atomic<std::uint64_t> n { 0 };
        return ++n;
    }
    bool real_env_ {}, rank_on_ {}, fixed_radix_ {};
    DeviceBuffer envd_;
    sp_ranking_desc rank_desc_ {};
    std::vector<double> rank_max_z_;
    std::unique_ptr<Session> prox_;
    DeviceBuffer edt_, edt_table_, rank_max_z_d_, rank_kfz_d_, rank_olist_, rank_cursor_, rank_partial_, rank_pcount_,
        rank_pages_, rank_count_;
};

}  // namespace sphip

struct sp_hip_session {
    sphip::Session* s {};
    sp_hip_ctx* ctx {};
};

using sphip::DeviceError;

namespace sphip {
SessionView session_view(sp_hip_session* s) { return s->s->view(); }

void session_launch_proposed(sp_hip_session* s, const void* owner, const std::function<void(hipStream_t)>& propose,
                             int slots, bool stamp, const std::function<void(hipStream_t)>* merge)
{
    s->s->launch_proposed(owner, propose, slots, stamp, merge);
}

void session_launch_full(sp_hip_session* s, bool stamp) { s->s->launch_mode(stamp, false); }

void session_read_result(sp_hip_session* s, sp_step_result& out) { s->s->read_result(out); }
}  // namespace sphip

namespace {
int fail(sp_hip_ctx* ctx, int status, const std::string& msg)
{
    if (ctx) ctx->c.last_error = msg;
    return status;
}

template <typename F> int guarded(sp_hip_ctx* ctx, F&& f)
{
    try {
        f();
        return SP_OK;
    }
    catch (const DeviceError& e) {
        if (e.what == "FFT_LENGTH") return fail(ctx, SP_FFT_LENGTH_UNSUPPORTED, "padded FFT length unsupported");
        return fail(ctx, SP_DEVICE_ERROR, e.what);
    }
    catch (const std::domain_error& e) {
        const std::string w = e.what();
        if (w == "CORRELATION_INVALID_ENVIRONMENT") return fail(ctx, SP_CORRELATION_INVALID_ENVIRONMENT, w);
        if (w == "CORRELATION_INVALID_KERNEL") return fail(ctx, SP_CORRELATION_INVALID_KERNEL, w);
        if (w == "FIELD_COUNT_UNDERFLOW") return fail(ctx, SP_FIELD_COUNT_UNDERFLOW, w);
        if (w == "CORRELATION_NUMERIC") return fail(ctx, SP_CORRELATION_NUMERIC, w);
        return fail(ctx, SP_INVALID_ARGUMENT, w);
    }
    catch (const std::length_error& e) {
        return fail(ctx, SP_CAPACITY, e.what());
    }
    catch (const std::bad_alloc&) {
        return fail(ctx, SP_CORRELATION_ALLOCATION, "host allocation failed");
    }
    catch (const std::invalid_argument& e) {
        if (std::string(e.what()) == "FIELD_COPY_ID") return fail(ctx, SP_FIELD_COPY_ID, "copy id is not present");
        return fail(ctx, SP_INVALID_ARGUMENT, e.what());
    }
    catch (const std::exception& e) {
        return fail(ctx, SP_INVALID_ARGUMENT, e.what());
    }
}
}  // namespace

extern "C" {

int sp_hip_session_create(sp_hip_ctx* ctx, const sp_session_desc* desc, sp_hip_session** out)
{
    if (!ctx || !desc || !out || desc->orientations == 0 || desc->orientations > 65535) return SP_INVALID_ARGUMENT;
    for (int a = 0; a != 3; ++a)
        if (desc->environment_shape[a] == 0 || desc->environment_shape[a] > 65535) return SP_INVALID_ARGUMENT;
    return guarded(ctx, [&] {
        auto* s = new sp_hip_session;
        s->ctx = ctx;
        try {
            s->s = new sphip::Session(ctx->c, *desc);
        }
        catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

void sp_hip_session_destroy(sp_hip_session* s)
{
    if (!s) return;
    delete s->s;
    delete s;
}

int sp_hip_session_set_environment(sp_hip_session* s, const uint8_t* cells, uint64_t count)
{
    if (!s || !cells) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->set_environment(cells, count); });
}

int sp_hip_session_get_environment(sp_hip_session* s, uint8_t* cells, uint64_t count)
{
    if (!s || !cells) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->get_environment(cells, count); });
}

int sp_hip_session_set_kernel(sp_hip_session* s, uint32_t o, const uint8_t* cells, uint64_t count,
                              const uint32_t shape[3], const int64_t first[3])
{
    if (!s || !cells || !shape || !first) return SP_INVALID_ARGUMENT;
    std::uint64_t n = 1;
    for (int a = 0; a != 3; ++a) {
        if (shape[a] == 0 || shape[a] > 65535) return fail(s->ctx, SP_CORRELATION_SHAPE_OVERFLOW, "kernel shape");
        n *= shape[a];
    }
    if (count != n) return fail(s->ctx, SP_CORRELATION_INVALID_FIELD, "kernel cell count does not match its shape");
    return guarded(s->ctx, [&] { s->s->set_kernel(o, cells, shape, first); });
}

int sp_hip_session_set_fast_path(sp_hip_session* s, int on)
{
    if (!s) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->set_fast_path(on != 0); });
}

int sp_hip_session_set_fused_passes(sp_hip_session* s, int on)
{
    if (!s) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->set_fused_passes(on != 0); });
}

int sp_hip_session_get_fused_passes(sp_hip_session* s, int32_t* active)
{
    if (!s || !active) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { *active = s->s->fused_active() ? 1 : 0; });
}

int sp_hip_session_set_active(sp_hip_session* s, const uint32_t* orientations, uint32_t count,
                              const uint8_t* region_mask, const uint32_t regions[3])
{
    if (!s || (count && !orientations)) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->set_active(orientations, count, region_mask, regions); });
}

int sp_hip_session_set_autotune(sp_hip_session* s, int on)
{
    if (!s) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->set_autotune(on != 0); });
}

int sp_hip_session_set_tile(sp_hip_session* s, int pass, int radix, int width)
{
    if (!s) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->set_tile(pass, radix, width); });
}

int sp_hip_session_get_tile(sp_hip_session* s, int pass, int32_t* radix, int32_t* width, float* ms)
{
    if (!s || !radix || !width || !ms) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] {
        int r = 0, w = 0;
        float t = 0.f;
        s->s->get_tile(pass, r, w, t);
        *radix = r;
        *width = w;
        *ms = t;
    });
}

int sp_hip_session_tile_candidates(sp_hip_session* s, int pass, int32_t* pairs, uint32_t capacity, uint32_t* count)
{
    if (!s || !count) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] {
        const auto tiles = s->s->fitting_tiles(pass);
        *count = uint32_t(tiles.size());
        for (uint32_t i = 0; i < tiles.size() && i < capacity && pairs; ++i) {
            pairs[2 * i] = tiles[i].r;
            pairs[2 * i + 1] = tiles[i].w;
        }
    });
}

int sp_hip_session_set_pair_clearance(sp_hip_session* s, double clearance_mm, double pitch_mm)
{
    if (!s) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->set_pair_clearance(clearance_mm, pitch_mm); });
}

int sp_hip_session_set_incremental(sp_hip_session* s, int on)
{
    if (!s) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->set_incremental(on != 0); });
}

int sp_hip_session_set_resident_kernels(sp_hip_session* s, int on)
{
    if (!s) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->set_resident_kernels(on != 0); });
}

int sp_hip_session_orientation_minima(sp_hip_session* s, uint64_t* keys, uint32_t count)
{
    if (!s || !keys) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->orientation_minima(keys, count); });
}

int sp_hip_session_exclude(sp_hip_session* s, uint32_t orientation, const int64_t translation[3])
{
    if (!s || !translation) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->exclude(orientation, translation); });
}

int sp_hip_session_stamp(sp_hip_session* s, uint32_t orientation, const int64_t translation[3])
{
    if (!s || !translation) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->stamp_at(orientation, translation); });
}

int sp_hip_session_copies(sp_hip_session* s, sp_placed_copy* out, uint64_t capacity, uint64_t* count)
{
    if (!s || !count || (capacity && !out)) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] {
        const auto all = s->s->copies();
        *count = all.size();
        if (all.size() > capacity) throw std::length_error("capacity");
        for (std::size_t i = 0; i != all.size(); ++i) {
            sp_placed_copy c {};
            c.copy_id = i;
            s->s->key_pose(all[i].key, c.orientation, c.translation);
            c.removed = all[i].removed ? 1u : 0u;
            out[i] = c;
        }
    });
}

int sp_hip_session_remove_copy(sp_hip_session* s, uint64_t copy_id)
{
    if (!s) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->remove_copy(copy_id); });
}

int sp_hip_session_placed_counts(sp_hip_session* s, uint32_t* counts, uint64_t count)
{
    if (!s || !counts) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->placed_counts(counts, count); });
}

int sp_hip_session_keep_raw(sp_hip_session* s, int keep)
{
    if (!s) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->keep_raw(keep != 0); });
}

int sp_hip_session_raw_values(sp_hip_session* s, uint32_t orientation, double* out, uint64_t count)
{
    if (!s || !out) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->raw_values(orientation, out, count); });
}

int sp_hip_session_keep_fields(sp_hip_session* s, int keep)
{
    if (!s) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->keep_fields(keep != 0); });
}

int sp_hip_session_step(sp_hip_session* s, int stamp, sp_step_result* out)
{
    if (!s || !out) return SP_INVALID_ARGUMENT;
    int rc = guarded(s->ctx, [&] {
        s->s->launch(stamp != 0);
        s->s->read_result(*out);
    });
    if (rc == SP_OK && ((out->numeric_flags | out->run_flags) & 7u))
        return fail(s->ctx, SP_CORRELATION_NUMERIC, "step numeric check failed");
    return rc;
}

int sp_hip_session_run(sp_hip_session* s, uint32_t steps, int stamp)
{
    if (!s) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] {
        for (uint32_t i = 0; i != steps; ++i) s->s->launch(stamp != 0);
    });
}

int sp_hip_session_last_result(sp_hip_session* s, sp_step_result* out)
{
    if (!s || !out) return SP_INVALID_ARGUMENT;
    const int rc = guarded(s->ctx, [&] { s->s->read_result(*out); });
    if (rc == SP_OK && ((out->numeric_flags | out->run_flags) & 7u))
        return fail(s->ctx, SP_CORRELATION_NUMERIC, "a step's numeric or direct check failed");
    return rc;
}

int sp_hip_session_overlap_field(sp_hip_session* s, uint32_t o, int32_t* out, uint64_t count, uint32_t legal[3])
{
    if (!s || !legal) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->overlap_field(o, out, count, legal); });
}

int sp_hip_session_profile_step(sp_hip_session* s, int stamp, float* kernel_ms, char* names, uint32_t capacity,
                                uint32_t* count)
{
    if (!s || !count) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] {
        const auto prof = s->s->profile_step(stamp != 0);
        *count = uint32_t(prof.size());
        for (uint32_t i = 0; i < prof.size() && i < capacity; ++i) {
            if (kernel_ms) kernel_ms[i] = prof[i].second;
            if (names) {
                std::strncpy(names + 32 * i, prof[i].first.c_str(), 31);
                names[32 * i + 31] = 0;
            }
        }
    });
}

int sp_hip_session_product_grid(sp_hip_session* s, uint32_t* blocks, uint64_t* tiles)
{
    if (!s || !blocks || !tiles) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->product_grid(*blocks, *tiles); });
}

int sp_hip_session_time_pass(sp_hip_session* s, const char* pass, uint32_t launches, float* ms)
{
    if (!s || !pass || !ms) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { *ms = s->s->time_pass(pass, int(launches)); });
}

void* sp_hip_session_stream(sp_hip_session* s) { return s ? static_cast<void*>(s->s->stream()) : nullptr; }

uint64_t sp_hip_session_device_bytes(const sp_hip_session* s) { return s ? s->s->device_bytes() : 0; }


int sp_hip_session_set_ranking(sp_hip_session* s, const sp_ranking_desc* desc, const double* object_max_z_mm,
                               uint32_t count)
{
    if (!s || !desc || (!object_max_z_mm && count)) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->set_ranking(*desc, object_max_z_mm, count); });
}

int sp_hip_session_rank(sp_hip_session* s)
{
    if (!s) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->rank(); });
}

int sp_hip_session_ranked_pages(sp_hip_session* s, int mode, const uint32_t* orientations, uint32_t count,
                                const sp_ranked_candidate* cursors, const uint8_t* has_cursor,
                                sp_ranked_candidate* pages, uint32_t* page_counts)
{
    if (!s || (count && (!orientations || !pages || !page_counts))) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx,
                   [&] { s->s->ranked_pages(mode, orientations, count, cursors, has_cursor, pages, page_counts); });
}

int sp_hip_session_proximity_field(sp_hip_session* s, double* values, uint32_t* squared, uint64_t count)
{
    if (!s) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->proximity_field(values, squared, count); });
}

int sp_hip_session_proximity_values(sp_hip_session* s, uint32_t o, double* values, uint64_t count)
{
    if (!s || !values) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] { s->s->proximity_values(o, values, count); });
}

int sp_hip_session_ranked_fill(sp_hip_session* s, uint32_t max_copies, sp_validate_fn validate, void* user,
                               sp_placement* placements, uint32_t capacity, uint32_t* placed)
{
    if (!s || !placed || (capacity && !placements)) return SP_INVALID_ARGUMENT;
    return guarded(s->ctx, [&] {
        const auto v = s->s->ranked_fill(std::min(max_copies, capacity), validate, user);
        for (std::size_t i = 0; i != v.size(); ++i) placements[i] = v[i];
        *placed = std::uint32_t(v.size());
    });
}

}  // extern "C"
